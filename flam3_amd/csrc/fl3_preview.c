/*
 * Clients of the batched device path in the genome-search preview loop
 * (flam3-genome.c:1020-1116): the bounding-box estimate of a new genome and
 * the colour trials of flam3_improve_colors.
 *
 * fl3h_estimate_bounding_box keeps flam3_estimate_bounding_box's result and
 * random-number consumption: one FP64 chaos-game trajectory (the GPU's
 * replay iteration, fl3h_iterate) and the same quantiles.
 *
 * fl3h_improve_colors draws every candidate colouring first -- each trial's
 * colours depend only on the random sequence, not on the earlier trials'
 * scores -- and renders all trial images in one fl3h_render_many call (small
 * renders share walker launches) instead of one flam3_render per trial.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "fl3_host.h"

static int cmp_double(const void *a, const void *b) {
   const double x = *(const double *)a, y = *(const double *)b;
   return x < y ? -1 : (x > y ? 1 : 0);
}

/* flam3_estimate_bounding_box (flam3.c:3574-3639): start point from rc, a
 * 20-step fuse, nsamples points, then the eps and 1 - eps quantiles in x and
 * y (the extremes when eps * nsamples truncates to 0); eps grows to 3x the
 * bad-value rate and is capped at 0.3.  Returns the bad-value count or -1. */
int fl3h_estimate_bounding_box(flam3_genome *cp, double eps, int nsamples, double *bmin, double *bmax,
                               randctx *rc) {
   double *points, *v, mn[2], mx[2];
   unsigned short *distrib;
   int i, bv = 0, low, high, nrows = 0;
   if (!cp || !bmin || !bmax || !rc) return -1;
   if (nsamples <= 0) nsamples = 10000;
   points = (double *)malloc(sizeof(double) * 4 * (size_t)nsamples);
   if (!points) return -1;
   points[0] = fl3_isaac_11(rc);
   points[1] = fl3_isaac_11(rc);
   points[2] = 0.0;
   points[3] = 0.0;
   if (fl3_prepare_precalc_flags(cp)) {
      free(points);
      return -1;
   }
   distrib = fl3_create_xform_distrib(cp, &nrows);
   if (!distrib) {
      free(points);
      return -1;
   }
   i = fl3h_iterate(cp, nsamples, 20, points, distrib, rc, 1, &bv);
   free(distrib);
   if (i) {
      free(points);
      return -1;
   }
   if (bv / (double)nsamples > eps) eps = 3 * bv / (double)nsamples;
   if (eps > 0.3) eps = 0.3;
   low = (int)(nsamples * eps);
   high = nsamples - low;
   mn[0] = mn[1] = 1e10;
   mx[0] = mx[1] = -1e10;
   for (i = 0; i < nsamples; i++) {
      const double *p = points + 4 * (size_t)i;
      if (p[0] < mn[0]) mn[0] = p[0];
      if (p[1] < mn[1]) mn[1] = p[1];
      if (p[0] > mx[0]) mx[0] = p[0];
      if (p[1] > mx[1]) mx[1] = p[1];
   }
   if (low == 0) {
      bmin[0] = mn[0];
      bmin[1] = mn[1];
      bmax[0] = mx[0];
      bmax[1] = mx[1];
      free(points);
      return bv;
   }
   /* the k-th smallest coordinate does not depend on how a sort orders
    * points with equal keys, so sorting the coordinates alone gives the
    * reference's picks after its sort of whole points */
   v = (double *)malloc(sizeof(double) * (size_t)nsamples);
   if (!v) {
      free(points);
      return -1;
   }
   for (int k = 0; k < 2; k++) {
      for (i = 0; i < nsamples; i++) v[i] = points[4 * (size_t)i + k];
      qsort(v, (size_t)nsamples, sizeof(double), cmp_double);
      bmin[k] = v[low];
      bmax[k] = v[high];
   }
   free(v);
   free(points);
   return bv;
}

/* ---------------------------------------------------------------- colours */

/* flam3_random01 (flam3.c:2485): the C library's random(), 28 bits. */
static double random01(void) { return (random() & 0xfffffff) / (double)0xfffffff; }

/* random_xform (palettes.c:351-359): an xform with positive weight other
 * than `excluded`, at most 100 draws. */
static int random_xform(const flam3_genome *g, int excluded) {
   int ntries = 0;
   while (ntries++ < 100) {
      const int i = (int)(random() % g->num_xforms);
      if (g->xform[i].density > 0.0 && i != excluded) return i;
   }
   return -1;
}

/* change_colors (palettes.c:458-474) without a palette change. */
static void change_colors(flam3_genome *g) {
   int i, x0, x1;
   for (i = 0; i < g->num_xforms; i++) g->xform[i].color = random01();
   x0 = random_xform(g, -1);
   x1 = random_xform(g, x0);
   if (x0 >= 0 && (random() & 1)) g->xform[x0].color = 0.0;
   if (x1 >= 0 && (random() & 1)) g->xform[x1].color = 1.0;
}

/* try_colors' render settings (palettes.c:362-407): about 10000 pixels at
 * quality 1, no supersampling, no density estimation, one batch and one
 * temporal sample. */
static void trial_geometry(flam3_genome *g) {
   const double scalar = sqrt(10000.0 / (double)(g->width * g->height));
   g->sample_density = 1;
   g->spatial_oversample = 1;
   g->estimator = 0.0;
   g->width = (int)(g->width * scalar);
   g->height = (int)(g->height * scalar);
   g->pixels_per_unit *= scalar;
   g->nbatches = 1;
   g->ntemporal_samples = 1;
}

/* try_colors' score (palettes.c:409-456): the occupied cells of a res^3
 * colour histogram of the image.  The reference returns (double)(hits /
 * res3), an integer division, which is 0 unless every cell is hit; kept. */
static double color_score(const unsigned char *img, int npix, int res) {
   const int res3 = res * res * res;
   int *hist = (int *)calloc((size_t)res3, sizeof(int));
   int i, hits = 0;
   if (!hist) return -1.0;
   for (i = 0; i < npix; i++) {
      const unsigned char *p = img + 3 * (size_t)i;
      hist[(p[0] * res / 256) + (p[1] * res / 256) * res + (p[2] * res / 256) * res * res]++;
   }
   for (i = 0; i < res3; i++)
      if (hist[i]) hits++;
   free(hist);
   return (double)(hits / res3);
}

/* flam3_improve_colors (palettes.c:476-504) with change_palette = 0: the
 * original colouring and ntries random recolourings (cumulative on g, in the
 * reference's order of random() calls) are scored by their trial images,
 * all rendered in one fl3h_render_many call; g ends with the first colouring
 * of the highest score.  Palette changes need the palette library
 * (flam3_get_palette), which is not part of this library: change_palette !=
 * 0 returns 1 without touching g.  Returns 0, or 1 on failure (g unchanged). */
int fl3h_improve_colors(flam3_genome *g, int ntries, int change_palette, int color_resolution) {
   const int n = ntries + 1;
   flam3_genome *cands, *trials;
   flam3_frame *frames;
   fl3h_request *reqs;
   stat_struct *stats;
   unsigned char **imgs;
   int i, best_i = 0, rc = 1;
   double best = -1.0;
   if (!g || ntries < 0 || color_resolution <= 0) return 1;
   if (change_palette) {
      fprintf(stderr, "flam3_amd: fl3h_improve_colors: palette changes need the palette library\n");
      return 1;
   }
   cands = (flam3_genome *)calloc((size_t)n, sizeof(flam3_genome));
   trials = (flam3_genome *)calloc((size_t)n, sizeof(flam3_genome));
   frames = (flam3_frame *)calloc((size_t)n, sizeof(flam3_frame));
   reqs = (fl3h_request *)calloc((size_t)n, sizeof(fl3h_request));
   stats = (stat_struct *)calloc((size_t)n, sizeof(stat_struct));
   imgs = (unsigned char **)calloc((size_t)n, sizeof(unsigned char *));
   if (!cands || !trials || !frames || !reqs || !stats || !imgs) goto done;
   for (i = 0; i < n; i++) {
      if (fl3_genome_copy(&cands[i], g)) goto done;
      if (i > 0) {
         /* the recolouring applies to the previous candidate's colours */
         int k;
         for (k = 0; k < g->num_xforms; k++) cands[i].xform[k].color = cands[i - 1].xform[k].color;
         change_colors(&cands[i]);
      }
      if (fl3_genome_copy(&trials[i], &cands[i])) goto done;
      trial_geometry(&trials[i]);
      imgs[i] = (unsigned char *)calloc((size_t)trials[i].width * trials[i].height + 1, 3);
      if (!imgs[i]) goto done;
      fl3h_init_frame(&frames[i], getenv("isaac_seed"));
      frames[i].bits = 33;
      frames[i].bytes_per_channel = 1;
      frames[i].verbose = 0;
      frames[i].genomes = &trials[i];
      frames[i].ngenomes = 1;
      frames[i].earlyclip = 1;
      frames[i].pixel_aspect_ratio = 1.0;
      frames[i].progress = NULL;
      frames[i].nthreads = 1;
      frames[i].sub_batch_size = 10000;
      reqs[i].frame = &frames[i];
      reqs[i].out = imgs[i];
      reqs[i].field = flam3_field_both;
      reqs[i].nchan = 3;
      reqs[i].transp = 0;
      reqs[i].stats = &stats[i];
   }
   if (fl3h_render_many(reqs, n, NULL) != 0) {
      fprintf(stderr, "flam3_amd: error rendering test images for the colour trials\n");
      goto done;
   }
   for (i = 0; i < n; i++) {
      const double b = color_score(imgs[i], trials[i].width * trials[i].height, color_resolution);
      if (b < 0) goto done;
      if (b > best) {
         best = b;
         best_i = i;
      }
   }
   for (i = 0; i < g->num_xforms; i++) g->xform[i].color = cands[best_i].xform[i].color;
   rc = 0;
done:
   for (i = 0; i < n; i++) {
      if (cands) fl3_genome_free(&cands[i]);
      if (trials) fl3_genome_free(&trials[i]);
      if (imgs) free(imgs[i]);
   }
   free(cands);
   free(trials);
   free(frames);
   free(reqs);
   free(stats);
   free(imgs);
   return rc;
}
