/*
 * fl3_host.h - host-side preparation used by flam3_render before anything
 * is sent to the GPU: ISAAC streams, per-xform precalculation, the xform
 * selection tables, and the filter banks.  All FP64, all plain C.
 */
#ifndef FL3_HOST_H
#define FL3_HOST_H

#include <stdint.h>
#include "flam3_abi.h"
#include "fl3h.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- ISAAC (isaac.c, isaac.h; RANDSIZL = 4) ---------------- */
void fl3_isaac_refill(randctx *rc);
void fl3_isaac_init(randctx *rc, int use_seed);
static inline ub4 fl3_irand(randctx *rc) {
   if (rc->randcnt-- == 0) {
      fl3_isaac_refill(rc);
      rc->randcnt = FLAM3_RANDSIZ - 1;
   }
   return rc->randrsl[rc->randcnt];
}
double fl3_isaac_01(randctx *rc);
double fl3_isaac_11(randctx *rc);

/* ---------------- genome helpers ---------------- */
#define FL3_CHOOSE_GRAIN 16384
#define FL3_EPS 1e-10

/* Deep copy / release of a genome (xform array, chaos rows, motion lists). */
int fl3_genome_copy(flam3_genome *dst, const flam3_genome *src);
void fl3_genome_free(flam3_genome *g);
/* Default field values of a genome / of freshly added xforms. */
void fl3_genome_defaults(flam3_genome *g, int defaults_on);
void fl3_xform_defaults(flam3_xform *xf, int index);
int fl3_genome_add_xforms(flam3_genome *g, int n, int padding, int final_flag);

/* Genome at time t from a time-sorted list (flam3.c:717-802).  Returns 0 on
 * success.  dst must be zeroed or previously produced by this library. */
int fl3_interpolate(const flam3_genome *cps, int ncps, double t, double stagger, flam3_genome *dst);

/* Active-variation lists and precalc flags for every xform (variations.c:2016). */
int fl3_prepare_precalc_flags(flam3_genome *cp);
/* Derived per-variation constants for one xform (variations.c:2004). */
void fl3_xform_precalc(flam3_xform *xf);
double fl3_adjust_percentage(double in);

/* Selection table as the reference lays it out: rows x 16384 uint16, row 0
 * from the plain weights, rows 1..n_std from weights x xaos row
 * (flam3.c:76-111, 165-224).  Sets cp->chaos_enable.  Returns NULL on error
 * (negative or all-zero weights).  *nrows receives the row count. */
unsigned short *fl3_create_xform_distrib(flam3_genome *cp, int *nrows);
/* The same table as step breakpoints: for row r and xform j,
 * starts[r*nstd + j] is the first slot whose entry is >= j (so entry(slot) =
 * number of j >= 1 with starts[j] <= slot).  Exact for any table built by
 * fl3_create_xform_distrib. */
void fl3_distrib_breakpoints(const unsigned short *distrib, int nrows, int nstd, int *starts);
/* The breakpoints of the rows the walkers use (row 0 without xaos, rows
 * 0..n_std with it; *nrows receives the count) without building the table:
 * what fl3_distrib_breakpoints gives for fl3_create_xform_distrib's table.
 * starts holds (n_std + 1) * n_std ints.  Sets cp->chaos_enable.  Returns 0,
 * or 1 on negative or all-zero weights.  Thread-safe (a small cache of
 * recent inputs and results). */
int fl3_xform_starts(flam3_genome *cp, int *starts, int *nrows);

/* ---------------- filters (filters.c) ---------------- */
typedef struct {
   int max_filtered_counts;
   int max_filter_index;
   int kernel_size;
   double *filter_widths;  /* max_filter_index+1 used entries */
   double *filter_coefs;   /* rows of kernel_size, 1/8-symmetric octant */
   int nfilters;           /* allocated rows */
} fl3_de_filters;

double fl3_spatial_filter_value(int knum, double x);
extern const double fl3_spatial_support[14];
/* Returns the filter width (>0) and a malloc'd fwidth^2 kernel, or -1. */
int fl3_create_spatial_filter(const flam3_genome *g0, double pixel_aspect, int field, double **filter);
/* Returns 0 on success. */
int fl3_create_de_filters(double max_rad, double min_rad, double curve, int ss, fl3_de_filters *de);
void fl3_free_de_filters(fl3_de_filters *de);
double fl3_create_temporal_filter(int numsteps, int filter_type, double filter_exp, double filter_width,
                                  double **temporal_filter, double **temporal_deltas);

/* ---------------- job packing (fl3_job.c) ---------------- */
/* Pack a prepared genome (precalc flags and per-variation constants done)
 * with its selection table (NULL: built here) and camera/palette (NULL:
 * none) into a malloc'd job blob; *bytes receives its size. */
void *fl3_pack_job(const flam3_genome *cp, const unsigned short *distrib, const fl3h_camera *cam,
                   size_t *bytes);
/* The same from the table's breakpoints (fl3_xform_starts, which also set
 * cp->chaos_enable). */
void *fl3_pack_job_starts(const flam3_genome *cp, const int *starts, const fl3h_camera *cam, size_t *bytes);
void *fl3_pack_camera_job(const fl3h_camera *cam, size_t *bytes);

/* Several renders in two phases (fl3h_render_many = prepare + execute):
 * host work of every render, then the device work.  The requests must stay
 * valid until fl3_many_execute returns. */
typedef struct fl3_many fl3_many;
fl3_many *fl3_many_prepare(fl3h_request *reqs, int n, const fl3h_render_opts *opts);
int fl3_many_execute(fl3_many *B);
void fl3_many_free(fl3_many *B);

/* Downloaded images of a group of shared renders, not yet copied into the
 * callers' buffers.  By default fl3_many_execute copies them itself; with a
 * delivery set it downloads each group into a pinned slot it acquires and
 * hands the group over, so that the copies (and the renders' completion)
 * run on another thread while the device goes on with the next group. */
typedef struct fl3_images fl3_images;
typedef struct {
   void *ctx;
   int (*acquire)(void *ctx);                     /* a free slot (0 or 1); may block */
   void (*hand_over)(void *ctx, fl3_images *im);  /* later: fl3_images_deliver(im) */
} fl3_delivery;
void fl3_many_set_delivery(fl3_many *B, const fl3_delivery *d);
/* Copy the group's images into the callers' buffers, complete their stats
 * and set their rc to 0 (rc stays 1 for a group whose download failed); frees
 * im and returns the slot it held, now free again.  B must not be freed
 * before every group handed over from it has been delivered. */
int fl3_images_deliver(fl3_images *im);

#ifdef __cplusplus
}
#endif
#endif
