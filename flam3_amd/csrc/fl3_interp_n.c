/*
 * Interpolation between two or more genomes (flam3_interpolate with
 * ncps > 1, flam3.c:717-802 and interpolation.c).
 */
#include <stdio.h>
#include "fl3_host.h"

int fl3_interpolate_multi(const flam3_genome *cps, int ncps, double t, double stagger, flam3_genome *dst) {
   (void)cps; (void)ncps; (void)t; (void)stagger; (void)dst;
   fprintf(stderr, "flam3_amd: interpolation between genomes is not available in this build\n");
   return 1;
}
