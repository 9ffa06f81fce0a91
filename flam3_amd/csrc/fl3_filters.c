/*
 * Filter banks built on the host and uploaded once per render (or per
 * batch): the separable spatial filter used by the downsample, the
 * variable-width density-estimation kernels, and the temporal (motion blur)
 * weights.  They follow filters.c step for step so that the tables are
 * bit-identical to the reference's; the reference's kernel numbering is kept
 * as-is, including its mismatch with the names in flam3.h (id 6, named
 * lanczos3 there, evaluates the Mitchell kernel with support 2.0).
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "fl3_host.h"

#define DE_THRESH 100

/* Support of each spatial kernel, indexed like the evaluator below
 * (filters.c:29-45). */
const double fl3_spatial_support[14] = {
   1.5, 1.0, 0.5, 1.0, 1.5, 2.0, 2.0, 1.0, 2.0, 1.0, 1.0, 3.0, 2.0, 1.5
};

static double sinc(double x) {
   x *= M_PI;
   if (x != 0) return sin(x) / x;
   return 1.0;
}

static double k_gaussian(double x) { return exp(-2.0 * x * x) * sqrt(2.0 / M_PI); }

static double k_hermite(double t) {
   if (t < 0.0) t = -t;
   if (t < 1.0) return (2.0 * t - 3.0) * t * t + 1.0;
   return 0.0;
}

static double k_box(double t) { return (t > -0.5 && t <= 0.5) ? 1.0 : 0.0; }

static double k_triangle(double t) {
   if (t < 0.0) t = -t;
   return t < 1.0 ? 1.0 - t : 0.0;
}

static double k_bell(double t) {
   if (t < 0) t = -t;
   if (t < .5) return .75 - t * t;
   if (t < 1.5) {
      t = t - 1.5;
      return .5 * (t * t);
   }
   return 0.0;
}

static double k_bspline(double t) {
   double tt;
   if (t < 0) t = -t;
   if (t < 1) {
      tt = t * t;
      return (.5 * tt * t) - tt + (2.0 / 3.0);
   } else if (t < 2) {
      t = 2 - t;
      return (1.0 / 6.0) * (t * t * t);
   }
   return 0.0;
}

static double k_mitchell(double t) {
   const double B = 1.0 / 3.0, C = 1.0 / 3.0;
   double tt = t * t;
   if (t < 0) t = -t;
   if (t < 1.0) {
      t = ((12.0 - 9.0 * B - 6.0 * C) * (t * tt)) + ((-18.0 + 12.0 * B + 6.0 * C) * tt) + (6.0 - 2 * B);
      return t / 6.0;
   } else if (t < 2.0) {
      t = ((-1.0 * B - 6.0 * C) * (t * tt)) + ((6.0 * B + 30.0 * C) * tt) +
          ((-12.0 * B - 48.0 * C) * t) + (8.0 * B + 24 * C);
      return t / 6.0;
   }
   return 0.0;
}

static double k_blackman(double x) { return 0.42 + 0.5 * cos(M_PI * x) + 0.08 * cos(2 * M_PI * x); }

static double k_catrom(double x) {
   if (x < -2.0) return 0.0;
   if (x < -1.0) return 0.5 * (4.0 + x * (8.0 + x * (5.0 + x)));
   if (x < 0.0) return 0.5 * (2.0 + x * x * (-5.0 - 3.0 * x));
   if (x < 1.0) return 0.5 * (2.0 + x * x * (-5.0 + 3.0 * x));
   if (x < 2.0) return 0.5 * (4.0 + x * (-8.0 + x * (5.0 - x)));
   return 0.0;
}

static double k_hanning(double x) { return 0.5 + 0.5 * cos(M_PI * x); }
static double k_hamming(double x) { return 0.54 + 0.46 * cos(M_PI * x); }

static double k_lanczos3(double t) {
   if (t < 0) t = -t;
   return t < 3.0 ? sinc(t) * sinc(t / 3.0) : 0.0;
}

static double k_lanczos2(double t) {
   if (t < 0) t = -t;
   return t < 2.0 ? sinc(t) * sinc(t / 2.0) : 0.0;
}

static double k_quadratic(double x) {
   if (x < -1.5) return 0.0;
   if (x < -0.5) return 0.5 * (x + 1.5) * (x + 1.5);
   if (x < 0.5) return 0.75 - x * x;
   if (x < 1.5) return 0.5 * (x - 1.5) * (x - 1.5);
   return 0.0;
}

/* Kernel evaluator by number (filters.c:172-202). */
double fl3_spatial_filter_value(int knum, double x) {
   switch (knum) {
      case 0: return k_gaussian(x);
      case 1: return k_hermite(x);
      case 2: return k_box(x);
      case 3: return k_triangle(x);
      case 4: return k_bell(x);
      case 5: return k_bspline(x);
      case 6: return k_mitchell(x);
      case 7: return sinc(x) * k_blackman(x);
      case 8: return k_catrom(x);
      case 9: return sinc(x) * k_hanning(x);
      case 10: return sinc(x) * k_hamming(x);
      case 11: return k_lanczos3(x) * sinc(x / 3.0);
      case 12: return k_lanczos2(x) * sinc(x / 2.0);
      default: return k_quadratic(x);
   }
}

/* Separable downsample filter, fwidth x fwidth, normalised to sum 1
 * (filters.c:217-270). */
int fl3_create_spatial_filter(const flam3_genome *g0, double aspect, int field, double **filter) {
   int knum = g0->spatial_filter_select;
   int ss = g0->spatial_oversample;
   double supp = fl3_spatial_support[knum];
   double fw = 2.0 * supp * ss * g0->spatial_filter_radius / aspect;
   double adjust, sum = 0.0, inv;
   int fwidth = ((int)fw) + 1, i, j, n;
   double *f;
   if ((fwidth ^ ss) & 1) fwidth++;
   adjust = (fw > 0.0) ? supp * fwidth / fw : 1.0;
   f = (double *)calloc((size_t)fwidth * fwidth, sizeof(double));
   if (!f) return -1;
   for (i = 0; i < fwidth; i++) {
      for (j = 0; j < fwidth; j++) {
         double ii = ((2.0 * i + 1.0) / (double)fwidth - 1.0) * adjust;
         double jj = ((2.0 * j + 1.0) / (double)fwidth - 1.0) * adjust;
         if (field) jj *= 2.0;
         jj /= aspect;
         f[i + j * fwidth] = fl3_spatial_filter_value(knum, ii) * fl3_spatial_filter_value(knum, jj);
      }
   }
   n = fwidth * fwidth;
   for (i = 0; i < n; i++) sum += f[i];
   if (sum == 0.0) {
      fprintf(stderr, "Spatial filter value is too small: %g.  Terminating.\n", g0->spatial_filter_radius);
      free(f);
      return -1;
   }
   inv = 1.0 / sum;
   for (i = 0; i < n; i++) f[i] *= inv;
   *filter = f;
   return fwidth;
}

/* Density-estimation kernel bank (filters.c:272-407).  Kernel k has radius
 * h_k that shrinks with the hit count; only the octant j >= k >= 0 of each
 * (2r-1)^2 Gaussian is stored. */
static int build_de_filters(double max_rad, double min_rad, double curve, int ss, fl3_de_filters *de) {
   double cmax, cmin, nf_d;
   int nf, max_ind, row, half, loop;
   const int keep = 100;

   de->kernel_size = -1;
   de->filter_coefs = NULL;
   de->filter_widths = NULL;
   if (curve <= 0.0) {
      fprintf(stderr, "estimator curve must be > 0\n");
      return 1;
   }
   if (max_rad < min_rad) {
      fprintf(stderr, "estimator must be larger than estimator_minimum.\n");
      return 1;
   }
   cmax = max_rad * ss + 1;
   cmin = min_rad * ss + 1;
   nf_d = pow(cmax / cmin, 1.0 / curve);
   if (nf_d > 1e7) {
      fprintf(stderr, "too many filters required in this configuration (%g)\n", nf_d);
      return 1;
   }
   nf = (int)ceil(nf_d);
   if (nf > keep) {
      max_ind = (int)ceil(DE_THRESH + pow(nf - DE_THRESH, curve)) + 1;
      de->max_filtered_counts = (int)pow((double)(max_ind - DE_THRESH), 1.0 / curve) + DE_THRESH;
   } else {
      max_ind = nf;
      de->max_filtered_counts = max_ind;
   }
   row = (int)(2 * ceil(cmax) - 1);
   half = (row - 1) / 2;
   de->kernel_size = (half + 1) * (2 + half) / 2;
   de->nfilters = max_ind;
   de->filter_coefs = (double *)calloc((size_t)max_ind * de->kernel_size, sizeof(double));
   de->filter_widths = (double *)calloc((size_t)max_ind, sizeof(double));
   if (!de->filter_coefs || !de->filter_widths) return 1;

   de->max_filter_index = 0;
   for (loop = 0; loop < max_ind; loop++) {
      double sum = 0.0, h, d;
      int a, b, idx;
      if (loop < keep)
         h = cmax / pow(loop + 1, curve);
      else
         h = cmax / pow(pow(loop - keep, 1.0 / curve) + keep + 1, curve);
      if (h <= cmin) {
         h = cmin;
         de->max_filter_index = loop;
      }
      de->filter_widths[loop] = h;
      for (a = -half; a <= half; a++)
         for (b = -half; b <= half; b++) {
            d = sqrt((double)(a * a + b * b)) / h;
            if (d <= 1.0) sum += k_gaussian(fl3_spatial_support[0] * d);
         }
      idx = loop * de->kernel_size;
      for (a = 0; a <= half; a++)
         for (b = 0; b <= a; b++) {
            d = sqrt((double)(a * a + b * b)) / h;
            de->filter_coefs[idx++] = (d > 1.0) ? 0.0 : k_gaussian(fl3_spatial_support[0] * d) / sum;
         }
      if (de->max_filter_index > 0) break;
   }
   if (de->max_filter_index == 0) de->max_filter_index = max_ind - 1;
   return 0;
}

/* The bank is a pure function of (max_rad, min_rad, curve, ss) and takes
 * most of a millisecond to build at radius 9 and oversample 2, once per
 * frame: the last few banks are kept and handed out as copies. */
#define DE_CACHE 8
static struct {
   double key[3];
   int ss, used;
   fl3_de_filters de;
} de_cache[DE_CACHE];
static int de_cache_next;
static pthread_mutex_t de_cache_mu = PTHREAD_MUTEX_INITIALIZER;

static int copy_de_filters(const fl3_de_filters *src, fl3_de_filters *dst) {
   const size_t nc = (size_t)src->nfilters * (size_t)src->kernel_size;
   *dst = *src;
   dst->filter_coefs = (double *)malloc(sizeof(double) * (nc > 0 ? nc : 1));
   dst->filter_widths = (double *)malloc(sizeof(double) * (size_t)(src->nfilters > 0 ? src->nfilters : 1));
   if (!dst->filter_coefs || !dst->filter_widths) {
      free(dst->filter_coefs);
      free(dst->filter_widths);
      dst->filter_coefs = dst->filter_widths = NULL;
      return 1;
   }
   memcpy(dst->filter_coefs, src->filter_coefs, sizeof(double) * nc);
   memcpy(dst->filter_widths, src->filter_widths, sizeof(double) * (size_t)src->nfilters);
   return 0;
}

int fl3_create_de_filters(double max_rad, double min_rad, double curve, int ss, fl3_de_filters *de) {
   const double key[3] = {max_rad, min_rad, curve};
   int i, rc;
   pthread_mutex_lock(&de_cache_mu);
   for (i = 0; i < DE_CACHE; i++)
      if (de_cache[i].used && de_cache[i].ss == ss && !memcmp(de_cache[i].key, key, sizeof(key))) {
         rc = copy_de_filters(&de_cache[i].de, de);
         pthread_mutex_unlock(&de_cache_mu);
         return rc;
      }
   pthread_mutex_unlock(&de_cache_mu);
   rc = build_de_filters(max_rad, min_rad, curve, ss, de);
   if (rc) return rc;
   pthread_mutex_lock(&de_cache_mu);
   i = de_cache_next;
   de_cache_next = (de_cache_next + 1) % DE_CACHE;
   if (de_cache[i].used) fl3_free_de_filters(&de_cache[i].de);
   de_cache[i].used = 0;
   if (copy_de_filters(de, &de_cache[i].de) == 0) {
      memcpy(de_cache[i].key, key, sizeof(key));
      de_cache[i].ss = ss;
      de_cache[i].used = 1;
   }
   pthread_mutex_unlock(&de_cache_mu);
   return 0;
}

void fl3_free_de_filters(fl3_de_filters *de) {
   free(de->filter_coefs);
   free(de->filter_widths);
   de->filter_coefs = NULL;
   de->filter_widths = NULL;
}

/* Motion-blur sample offsets and weights (filters.c:409-489); returns the
 * mean weight used in the density normalisation. */
double fl3_create_temporal_filter(int n, int type, double fexp, double width,
                                  double **filter_out, double **deltas_out) {
   double maxf = 0.0, sum = 0.0, slpx, half;
   double *deltas = (double *)malloc(sizeof(double) * n);
   double *f = (double *)malloc(sizeof(double) * n);
   int i;
   if (n == 1) {
      deltas[0] = 0;
      f[0] = 1.0;
      *deltas_out = deltas;
      *filter_out = f;
      return 1.0;
   }
   for (i = 0; i < n; i++) deltas[i] = ((double)i / (double)(n - 1) - 0.5) * width;
   if (type == flam3_temporal_exp) {
      for (i = 0; i < n; i++) {
         slpx = (fexp >= 0) ? ((double)i + 1.0) / n : (double)(n - i) / n;
         f[i] = pow(slpx, fabs(fexp));
         if (f[i] > maxf) maxf = f[i];
      }
   } else if (type == flam3_temporal_gaussian) {
      half = n / 2.0;
      for (i = 0; i < n; i++) {
         f[i] = k_gaussian(fl3_spatial_support[0] * fabs(i - half) / half);
         if (f[i] > maxf) maxf = f[i];
      }
   } else {
      for (i = 0; i < n; i++) f[i] = 1.0;
      maxf = 1.0;
   }
   for (i = 0; i < n; i++) {
      f[i] /= maxf;
      sum += f[i];
   }
   sum /= n;
   *deltas_out = deltas;
   *filter_out = f;
   return sum;
}
