/*
 * Genome at a given time (flam3_interpolate, flam3.c:717-802).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "fl3_host.h"

int fl3_interpolate_multi(const flam3_genome *cps, int ncps, double t, double stagger, flam3_genome *dst);

int fl3_interpolate(const flam3_genome *cps, int ncps, double t, double stagger, flam3_genome *dst) {
   if (ncps < 1) {
      fprintf(stderr, "flam3_amd: no genomes to render\n");
      return 1;
   }
   fl3_genome_free(dst);
   if (ncps == 1) return fl3_genome_copy(dst, &cps[0]);
   return fl3_interpolate_multi(cps, ncps, t, stagger, dst);
}

int fl3h_interpolate(const flam3_genome *genomes, int ngenomes, double time, double stagger,
                     flam3_genome *result) {
   return fl3_interpolate(genomes, ngenomes, time, stagger, result);
}

void fl3h_genome_clear(flam3_genome *g) { fl3_genome_free(g); }
