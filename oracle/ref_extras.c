/*
 * Test-infrastructure glue compiled together with the reference flam3
 * sources (see oracle/Makefile).  It only exposes, as plain functions,
 * things the reference offers as macros or needs a little set-up for, so the
 * Python tests can drive the reference through ctypes.  Nothing here is part
 * of the product library.
 */
#include "private.h"
#include "isaac.h"
#include "filters.h"
#include "variations.h"
#include "palettes.h"

/* irand is a macro (isaac.h:49-52). */
unsigned long ref_irand(randctx *rc) { return irand(rc); }

/* flam3_init_frame seeds from the isaac_seed env var (flam3.c:2494-2516);
 * this seeds from an explicit string the same way. */
void ref_seed_rc(randctx *rc, const char *seed) {
   memset(rc->randrsl, 0, RANDSIZ * sizeof(ub4));
   strncpy((char *)&rc->randrsl, seed, RANDSIZ * sizeof(ub4));
   irandinit(rc, 1);
}

/* flam3_create_de_filters returns a struct by value (filters.c:272). */
int ref_de_filters(double max_rad, double min_rad, double curve, int ss,
                   int *max_filtered_counts, int *max_filter_index, int *kernel_size,
                   double **widths, double **coefs) {
   flam3_de_helper de = flam3_create_de_filters(max_rad, min_rad, curve, ss);
   *max_filtered_counts = de.max_filtered_counts;
   *max_filter_index = de.max_filter_index;
   *kernel_size = de.kernel_size;
   *widths = de.filter_widths;
   *coefs = de.filter_coefs;
   return de.kernel_size < 0;
}

void ref_free(void *p) { free(p); }

/* Free an array of genomes returned by flam3_parse_xml2. */
void ref_free_genomes(flam3_genome *g, int n) {
   int i;
   for (i = 0; i < n; i++) clear_cp(&g[i], 0);
   free(g);
}

/* Interpolate a genome at time t into freshly malloc'd storage. */
flam3_genome *ref_interpolate(flam3_genome *cps, int ncps, double t) {
   flam3_genome *r = (flam3_genome *)calloc(1, sizeof(flam3_genome));
   flam3_interpolate(cps, ncps, t, 0, r);
   return r;
}

flam3_genome *ref_interpolate_stagger(flam3_genome *cps, int ncps, double t, double stagger) {
   flam3_genome *r = (flam3_genome *)calloc(1, sizeof(flam3_genome));
   flam3_interpolate(cps, ncps, t, stagger, r);
   return r;
}

size_t ref_sizeof_genome(void) { return sizeof(flam3_genome); }
size_t ref_sizeof_xform(void) { return sizeof(flam3_xform); }
size_t ref_sizeof_frame(void) { return sizeof(flam3_frame); }
