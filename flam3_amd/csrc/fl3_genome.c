/*
 * Genome bookkeeping on the host: deep copies, defaults, the per-xform
 * precalculation that decides which variations run, and the xform selection
 * tables.  The selection table must be identical slot for slot to the
 * reference's (flam3.c:76-111, 165-224) because the device picks xforms with
 * the same 14 random bits the reference uses; the GPU consumes it as
 * per-row step breakpoints instead of 32 KB tables.
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "fl3_host.h"

/* Variation indices that need a precalculated quantity (flam3.h:87-185). */
enum {
   V_POLAR = 5, V_HANDKERCHIEF = 6, V_HEART = 7, V_DISC = 8, V_SPIRAL = 9,
   V_HYPERBOLIC = 10, V_DIAMOND = 11, V_EX = 12, V_JULIA = 13, V_POWER = 19,
   V_RINGS = 21, V_FAN = 22, V_BLOB = 23, V_FAN2 = 25, V_RINGS2 = 26,
   V_JULIAN = 32, V_JULIASCOPE = 33, V_RADIAL_BLUR = 36, V_NGON = 38,
   V_DISC2 = 49, V_SUPER_SHAPE = 50, V_FLOWER = 51, V_CONIC = 52, V_CPOW = 59,
   V_ESCHER = 63, V_PRE_BLUR = 67, V_POLAR2 = 70, V_WEDGE = 77,
   V_WEDGE_JULIA = 78, V_WEDGE_SPH = 79, V_WHORL = 80, V_LOG = 83
};

static int num_std(const flam3_genome *g) {
   return g->num_xforms - (g->final_xform_index >= 0);
}

/* ---------------------------------------------------------------- copies */

static int copy_xform(flam3_xform *dst, const flam3_xform *src) {
   *dst = *src;
   dst->motion = NULL;
   dst->num_motion = 0;
   if (src->num_motion > 0 && src->motion) {
      dst->motion = (flam3_xform *)malloc(sizeof(flam3_xform) * src->num_motion);
      if (!dst->motion) return 1;
      memcpy(dst->motion, src->motion, sizeof(flam3_xform) * src->num_motion);
      dst->num_motion = src->num_motion;
   }
   return 0;
}

void fl3_genome_free(flam3_genome *g) {
   int i, ns;
   if (!g) return;
   if (g->xform) {
      ns = num_std(g);
      if (g->chaos) {
         for (i = 0; i < ns; i++) free(g->chaos[i]);
         free(g->chaos);
      }
      for (i = 0; i < g->num_xforms; i++) free(g->xform[i].motion);
      free(g->xform);
   }
   g->xform = NULL;
   g->chaos = NULL;
   g->num_xforms = 0;
   g->final_xform_index = -1;
   g->final_xform_enable = 0;
}

/* Copy with the standard xforms first and the final xform last, the order
 * flam3_copy produces (flam3.c:1171-1209). */
int fl3_genome_copy(flam3_genome *dst, const flam3_genome *src) {
   int i, j, ns = num_std(src), k = 0;
   *dst = *src;
   dst->xform = (flam3_xform *)calloc(src->num_xforms > 0 ? src->num_xforms : 1, sizeof(flam3_xform));
   dst->chaos = (double **)calloc(ns > 0 ? ns : 1, sizeof(double *));
   if (!dst->xform || !dst->chaos) return 1;
   for (i = 0; i < src->num_xforms; i++) {
      if (i == src->final_xform_index) continue;
      if (copy_xform(&dst->xform[k++], &src->xform[i])) return 1;
   }
   if (src->final_xform_index >= 0) {
      if (copy_xform(&dst->xform[k], &src->xform[src->final_xform_index])) return 1;
      dst->final_xform_index = k;
   }
   for (i = 0; i < ns; i++) {
      dst->chaos[i] = (double *)malloc(sizeof(double) * ns);
      if (!dst->chaos[i]) return 1;
      for (j = 0; j < ns; j++)
         dst->chaos[i][j] = (src->chaos && src->chaos[i]) ? src->chaos[i][j] : 1.0;
   }
   return 0;
}

/* ---------------------------------------------------------------- defaults */

/* Field defaults of a genome (flam3.c:1270-1369, clear_cp). */
void fl3_genome_defaults(flam3_genome *cp, int defaults_on) {
   cp->palette_index = -1;
   cp->center[0] = cp->center[1] = 0.0;
   cp->rot_center[0] = cp->rot_center[1] = 0.0;
   cp->gamma = 4.0;
   cp->vibrancy = 1.0;
   cp->contrast = 1.0;
   cp->brightness = 4.0;
   cp->symmetry = 0;
   cp->hue_rotation = 0.0;
   cp->rotate = 0.0;
   cp->pixels_per_unit = 50;
   cp->interpolation = flam3_interpolation_linear;
   cp->palette_interpolation = flam3_palette_interpolation_hsv_circular;
   cp->hsv_rgb_palette_blend = 0.0;
   cp->genome_index = 0;
   memset(cp->parent_fname, 0, FLAM3_PARENT_FN_LEN);
   cp->highlight_power = -1.0;
   if (defaults_on) {
      cp->background[0] = cp->background[1] = cp->background[2] = 0.0;
      cp->width = 100;
      cp->height = 100;
      cp->spatial_oversample = 1;
      cp->spatial_filter_radius = 0.5;
      cp->zoom = 0.0;
      cp->sample_density = 1;
      cp->estimator = 9.0;
      cp->estimator_minimum = 0.0;
      cp->estimator_curve = 0.4;
      cp->gam_lin_thresh = 0.01;
      cp->nbatches = 1;
      cp->ntemporal_samples = 1000;
      cp->spatial_filter_select = flam3_gaussian_kernel;
      cp->interpolation_type = flam3_inttype_log;
      cp->temporal_filter_type = flam3_temporal_box;
      cp->temporal_filter_width = 1.0;
      cp->temporal_filter_exp = 0.0;
      cp->palette_mode = flam3_palette_mode_step;
   } else {
      cp->background[0] = cp->background[1] = cp->background[2] = -1.0;
      cp->zoom = 999999999;
      cp->spatial_oversample = -1;
      cp->spatial_filter_radius = -1;
      cp->width = -1;
      cp->height = -1;
      cp->sample_density = -1;
      cp->estimator = -1;
      cp->estimator_minimum = -1;
      cp->estimator_curve = -1;
      cp->gam_lin_thresh = -1;
      cp->nbatches = 0;
      cp->ntemporal_samples = 0;
      cp->spatial_filter_select = -1;
      cp->interpolation_type = -1;
      cp->temporal_filter_type = -1;
      cp->temporal_filter_width = -1;
      cp->temporal_filter_exp = -999;
      cp->palette_mode = -1;
   }
   fl3_genome_free(cp);
}

/* Parameter defaults of a new xform (variations.c:2401-2563). */
void fl3_xform_defaults(flam3_xform *x, int index) {
   int j;
   memset(x, 0, sizeof(*x));
   x->color_speed = 0.5;
   x->animate = 1.0;
   x->color = index & 1;
   x->opacity = 1.0;
   x->var[0] = 1.0;
   for (j = 1; j < FLAM3_NVARIATIONS; j++) x->var[j] = 0.0;
   x->c[0][0] = 1.0; x->c[1][1] = 1.0;
   x->post[0][0] = 1.0; x->post[1][1] = 1.0;
   x->blob_high = 1.0; x->blob_waves = 1.0;
   x->bent2_x = 1.0; x->bent2_y = 1.0;
   x->cell_size = 1.0;
   x->cpow_r = 1.0; x->cpow_power = 1.0;
   x->curve_xlength = 1.0; x->curve_ylength = 1.0;
   x->oscope_separation = 1.0; x->oscope_frequency = M_PI; x->oscope_amplitude = 1.0;
   x->wedge_count = 1.0;
   x->wedge_sph_count = 1.0;
   x->wedge_julia_power = 1.0; x->wedge_julia_count = 1.0;
   x->wedgeJulia_cn = 0.5; x->wedgeJulia_rN = 1.0;
   x->auger_freq = 1.0; x->auger_weight = 0.5; x->auger_scale = 1.0;
   x->julian_power = 1.0; x->julian_dist = 1.0; x->julian_rN = 1.0; x->julian_cn = 0.5;
   x->juliascope_power = 1.0; x->juliascope_dist = 1.0; x->juliascope_rN = 1.0; x->juliascope_cn = 0.5;
   x->radialBlur_zoomvar = 1.0;
   x->pie_slices = 6.0; x->pie_thickness = 0.5;
   x->ngon_sides = 5; x->ngon_power = 3; x->ngon_circle = 1; x->ngon_corners = 2;
   x->curl_c1 = 1.0;
   x->rectangles_x = 1.0; x->rectangles_y = 1.0;
   x->amw_amp = 1.0;
   x->super_shape_n1 = 1.0; x->super_shape_n2 = 1.0; x->super_shape_n3 = 1.0;
   x->conic_eccentricity = 1.0;
}

/* Append n xforms (flam3.c:1033-1093): new xforms get defaults, a final
 * xform stays last, xaos rows/columns of new standard xforms are 1. */
int fl3_genome_add_xforms(flam3_genome *g, int n, int padding, int final_flag) {
   int old_num = g->num_xforms, oldstd = num_std(g), i, j, numstd;
   flam3_xform *nx = (flam3_xform *)realloc(g->xform, sizeof(flam3_xform) * (old_num + n));
   if (!nx) return 1;
   g->xform = nx;
   g->num_xforms += n;
   for (i = old_num; i < g->num_xforms; i++) {
      fl3_xform_defaults(&g->xform[i], i);
      g->xform[i].padding = padding ? 1 : 0;
   }
   if (g->final_xform_index >= 0 && g->final_xform_index != g->num_xforms - 1) {
      flam3_xform tmp = g->xform[g->final_xform_index];
      for (i = g->final_xform_index; i < g->num_xforms - 1; i++) g->xform[i] = g->xform[i + 1];
      g->final_xform_index = g->num_xforms - 1;
      g->xform[g->final_xform_index] = tmp;
   }
   if (final_flag) {
      g->final_xform_enable = 1;
      g->final_xform_index = g->num_xforms - 1;
   } else {
      double **nc;
      numstd = num_std(g);
      for (i = 0; i < oldstd; i++) {
         double *row = (double *)realloc(g->chaos[i], sizeof(double) * numstd);
         if (!row) return 1;
         g->chaos[i] = row;
         for (j = oldstd; j < numstd; j++) row[j] = 1.0;
      }
      nc = (double **)realloc(g->chaos, sizeof(double *) * (numstd > 0 ? numstd : 1));
      if (!nc) return 1;
      g->chaos = nc;
      for (i = oldstd; i < numstd; i++) {
         g->chaos[i] = (double *)malloc(sizeof(double) * numstd);
         if (!g->chaos[i]) return 1;
         for (j = 0; j < numstd; j++) g->chaos[i][j] = 1.0;
      }
   }
   return 0;
}

/* ---------------------------------------------------------------- precalc */

/* Opacity remap 10^(log2(in)) used for vis_adjusted (interpolation.c:22-29). */
double fl3_adjust_percentage(double in) {
   if (in == 0.0) return 0.0;
   return pow(10.0, -log(1.0 / in) / log(2));
}

static int is_identity(double s[3][2]) {
   return s[0][0] == 1.0 && s[0][1] == 0.0 && s[1][0] == 0.0 &&
          s[1][1] == 1.0 && s[2][0] == 0.0 && s[2][1] == 0.0;
}

int fl3_prepare_precalc_flags(flam3_genome *cp) {
   int i, j;
   for (i = 0; i < cp->num_xforms; i++) {
      flam3_xform *x = &cp->xform[i];
      int n = 0;
      if (x->density < 0.0) {
         fprintf(stderr, "xform %d weight must be non-negative, not %g.\n", i, x->density);
         return 1;
      }
      /* Xforms that can never be chosen keep whatever they had. */
      if (i != cp->final_xform_index && x->density == 0.0) continue;

      x->vis_adjusted = fl3_adjust_percentage(x->opacity);
      x->precalc_angles_flag = 0;
      x->precalc_atan_xy_flag = 0;
      x->precalc_atan_yx_flag = 0;
      x->has_preblur = 0;
      x->has_post = !is_identity(x->post);
      for (j = 0; j < FLAM3_NVARIATIONS; j++) {
         if (x->var[j] == 0) continue;
         x->varFunc[n] = j;
         x->active_var_weights[n] = x->var[j];
         n++;
         switch (j) {
            case V_POLAR: case V_HANDKERCHIEF: case V_HEART: case V_DISC:
            case V_EX: case V_JULIA: case V_FAN: case V_FAN2: case V_DISC2:
            case V_POLAR2:
               x->precalc_atan_xy_flag = 1; break;
            case V_SPIRAL: case V_HYPERBOLIC: case V_DIAMOND: case V_POWER:
            case V_RINGS: case V_RINGS2:
               x->precalc_angles_flag = 1; break;
            case V_BLOB:
               x->precalc_atan_xy_flag = 1; x->precalc_angles_flag = 1; break;
            case V_JULIAN: case V_JULIASCOPE: case V_RADIAL_BLUR: case V_NGON:
            case V_SUPER_SHAPE: case V_FLOWER: case V_CONIC: case V_CPOW:
            case V_ESCHER: case V_WEDGE: case V_WEDGE_JULIA: case V_WEDGE_SPH:
            case V_WHORL: case V_LOG:
               x->precalc_atan_yx_flag = 1; break;
            case V_PRE_BLUR:
               x->has_preblur = x->var[j]; break;
            default: break;
         }
      }
      x->num_active_vars = n;
   }
   return 0;
}

/* Derived constants (variations.c:1942-2014). */
void fl3_xform_precalc(flam3_xform *x) {
   double ang, s, c, add, k;
   ang = x->perspective_angle * M_PI / 2.0;
   x->persp_vsin = sin(ang);
   x->persp_vfcos = x->perspective_dist * cos(ang);

   x->julian_rN = fabs(x->julian_power);
   x->julian_cn = x->julian_dist / (double)x->julian_power / 2.0;

   x->juliascope_rN = fabs(x->juliascope_power);
   x->juliascope_cn = x->juliascope_dist / (double)x->juliascope_power / 2.0;

   ang = x->radial_blur_angle * M_PI / 2.0;
   x->radialBlur_spinvar = sin(ang);
   x->radialBlur_zoomvar = cos(ang);

   x->waves_dx2 = 1.0 / (x->c[2][0] * x->c[2][0] + FL3_EPS);
   x->waves_dy2 = 1.0 / (x->c[2][1] * x->c[2][1] + FL3_EPS);

   add = x->disc2_twist;
   x->disc2_timespi = x->disc2_rot * M_PI;
   s = sin(add);
   c = cos(add);
   x->disc2_sinadd = s;
   x->disc2_cosadd = c - 1;
   if (add > 2 * M_PI) {
      k = 1 + add - 2 * M_PI;
      x->disc2_cosadd *= k;
      x->disc2_sinadd *= k;
   }
   if (add < -2 * M_PI) {
      k = 1 + add + 2 * M_PI;
      x->disc2_cosadd *= k;
      x->disc2_sinadd *= k;
   }

   x->super_shape_pm_4 = x->super_shape_m / 4.0;
   x->super_shape_pneg1_n1 = -1.0 / x->super_shape_n1;

   x->wedgeJulia_cf = 1.0 - x->wedge_julia_angle * x->wedge_julia_count * M_1_PI * 0.5;
   x->wedgeJulia_rN = fabs(x->wedge_julia_power);
   x->wedgeJulia_cn = x->wedge_julia_dist / x->wedge_julia_power / 2.0;
}

/* ---------------------------------------------------------------- selection */

/* One row of the table: weights optionally scaled by xaos row `from`. */
static int fill_row(const flam3_genome *cp, int from, unsigned short *row) {
   int i, j, ns = num_std(cp);
   double total = 0.0, step, acc, edge;
   for (i = 0; i < ns; i++) {
      double d = cp->xform[i].density;
      if (from >= 0) d *= cp->chaos[from][i];
      if (d < 0.0) {
         fprintf(stderr, "xform weight must be non-negative, not %g.\n", d);
         return 1;
      }
      total += d;
   }
   if (total == 0.0) {
      fprintf(stderr, "cannot iterate empty flame.\n");
      return 1;
   }
   step = total / FL3_CHOOSE_GRAIN;
   /* Slot i gets the first xform whose cumulative weight exceeds i*step,
    * walking the cumulative sum in the same order as the reference. */
   j = 0;
   edge = cp->xform[0].density * (from >= 0 ? cp->chaos[from][0] : 1.0);
   acc = 0.0;
   for (i = 0; i < FL3_CHOOSE_GRAIN; i++) {
      while (acc >= edge) {
         j++;
         if (from >= 0)
            edge += cp->xform[j].density * cp->chaos[from][j];
         else
            edge += cp->xform[j].density;
      }
      row[i] = (unsigned short)j;
      acc += step;
   }
   return 0;
}

unsigned short *fl3_create_xform_distrib(flam3_genome *cp, int *nrows) {
   int ns = num_std(cp), rows = ns + 1, i, j, r, unity = 1;
   unsigned short *t = (unsigned short *)calloc((size_t)rows * FL3_CHOOSE_GRAIN, sizeof(unsigned short));
   if (!t) return NULL;
   if (fill_row(cp, -1, t)) { free(t); return NULL; }
   for (i = 0; i < ns; i++)
      for (j = 0; j < ns; j++)
         if (fabs(cp->chaos[i][j] - 1.0) > FL3_EPS) unity = 0;
   cp->chaos_enable = !unity;
   if (cp->chaos_enable) {
      r = 0;
      for (i = 0; i < cp->num_xforms; i++) {
         if (i == cp->final_xform_index) continue;
         r++;
         if (fill_row(cp, i, t + (size_t)r * FL3_CHOOSE_GRAIN)) { free(t); return NULL; }
      }
   }
   if (nrows) *nrows = rows;
   return t;
}

/* Breakpoints of the selection rows the walkers use, without the table:
 * each row filled as fl3_create_xform_distrib fills it, into one row buffer,
 * and reduced to its breakpoints.  The last STARTS_CACHE inputs (weights
 * and, with xaos, the xaos rows) and their breakpoints are kept, since the
 * temporal samples of a frame, and the frames of an animation, usually share
 * their weights; the lookup compares the inputs bit for bit. */
#define STARTS_CACHE 8
typedef struct {
   double *key;
   size_t nkey;
   int *starts;
   int nstarts, rows;
} starts_entry;
static starts_entry starts_cache[STARTS_CACHE];
static int starts_next;
static pthread_mutex_t starts_mu = PTHREAD_MUTEX_INITIALIZER;

int fl3_xform_starts(flam3_genome *cp, int *starts, int *nrows) {
   int ns = num_std(cp), i, j, r, unity = 1, rows, e;
   size_t nkey, k = 0;
   double *key;
   unsigned short *row;
   for (i = 0; i < ns; i++)
      for (j = 0; j < ns; j++)
         if (fabs(cp->chaos[i][j] - 1.0) > FL3_EPS) unity = 0;
   cp->chaos_enable = !unity;
   rows = unity ? 1 : ns + 1;
   nkey = 2 + (size_t)ns + (unity ? 0 : (size_t)ns * ns);
   key = (double *)malloc(sizeof(double) * nkey);
   if (!key) return 1;
   key[k++] = ns;
   key[k++] = cp->final_xform_index;
   for (i = 0; i < ns; i++) key[k++] = cp->xform[i].density;
   if (!unity)
      for (i = 0; i < ns; i++)
         for (j = 0; j < ns; j++) key[k++] = cp->chaos[i][j];
   pthread_mutex_lock(&starts_mu);
   for (e = 0; e < STARTS_CACHE; e++) {
      const starts_entry *c = &starts_cache[e];
      if (c->key && c->nkey == nkey && memcmp(c->key, key, sizeof(double) * nkey) == 0) {
         memcpy(starts, c->starts, sizeof(int) * (size_t)c->nstarts);
         pthread_mutex_unlock(&starts_mu);
         free(key);
         if (nrows) *nrows = rows;
         return 0;
      }
   }
   pthread_mutex_unlock(&starts_mu);
   row = (unsigned short *)malloc(sizeof(unsigned short) * FL3_CHOOSE_GRAIN);
   if (!row) {
      free(key);
      return 1;
   }
   if (fill_row(cp, -1, row)) goto fail;
   fl3_distrib_breakpoints(row, 1, ns, starts);
   if (!unity) {
      r = 0;
      for (i = 0; i < cp->num_xforms; i++) {
         if (i == cp->final_xform_index) continue;
         r++;
         if (fill_row(cp, i, row)) goto fail;
         fl3_distrib_breakpoints(row, 1, ns, starts + (size_t)r * ns);
      }
   }
   free(row);
   {
      int *copy = (int *)malloc(sizeof(int) * (size_t)rows * ns);
      if (copy) {
         memcpy(copy, starts, sizeof(int) * (size_t)rows * ns);
         pthread_mutex_lock(&starts_mu);
         starts_entry *c = &starts_cache[starts_next];
         starts_next = (starts_next + 1) % STARTS_CACHE;
         free(c->key);
         free(c->starts);
         c->key = key;
         c->nkey = nkey;
         c->starts = copy;
         c->nstarts = rows * ns;
         c->rows = rows;
         pthread_mutex_unlock(&starts_mu);
         key = NULL;
      }
   }
   free(key);
   if (nrows) *nrows = rows;
   return 0;
fail:
   free(row);
   free(key);
   return 1;
}

void fl3_distrib_breakpoints(const unsigned short *distrib, int nrows, int nstd, int *starts) {
   int r, j, i;
   for (r = 0; r < nrows; r++) {
      const unsigned short *row = distrib + (size_t)r * FL3_CHOOSE_GRAIN;
      i = 0;
      for (j = 0; j < nstd; j++) {
         while (i < FL3_CHOOSE_GRAIN && row[i] < j) i++;
         starts[r * nstd + j] = i;
      }
   }
}
