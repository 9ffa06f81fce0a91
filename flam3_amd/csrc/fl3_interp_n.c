/*
 * Genome at time t between control points: flam3_interpolate with ncps > 1
 * (flam3.c:717-802), flam3_align (interpolation.c:768-1031),
 * flam3_interpolate_n (interpolation.c:373-708) and its helpers.  Host C,
 * FP64, same operation order as the reference so the interpolated genome
 * (and with it the render tables) is bit-identical.
 */
#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "fl3_host.h"
#include "fl3_tonemath.h"

#define IEPS 1e-10

enum {
   VI_LINEAR = 0, VI_SPHERICAL = 2, VI_POLAR = 5, VI_RINGS = 21, VI_FAN = 22, VI_BLOB = 23, VI_PDJ = 24,
   VI_FAN2 = 25, VI_RINGS2 = 26, VI_PERSPECTIVE = 30, VI_JULIAN = 32, VI_JULIASCOPE = 33, VI_RADIAL_BLUR = 36,
   VI_PIE = 37, VI_NGON = 38, VI_CURL = 39, VI_RECTANGLES = 40, VI_DISC2 = 49, VI_SUPER_SHAPE = 50,
   VI_FLOWER = 51, VI_CONIC = 52, VI_PARABOLA = 53, VI_BENT2 = 54, VI_BIPOLAR = 55, VI_CELL = 58, VI_CPOW = 59,
   VI_CURVE = 60, VI_ESCHER = 63, VI_LAZYSUSAN = 65, VI_MODULUS = 68, VI_OSCILLOSCOPE = 69, VI_POPCORN2 = 71,
   VI_SEPARATION = 73, VI_SPLIT = 74, VI_SPLITS = 75, VI_STRIPES = 76, VI_WEDGE = 77, VI_WEDGE_JULIA = 78,
   VI_WEDGE_SPH = 79, VI_WHORL = 80, VI_WAVES2 = 81, VI_AUGER = 96, VI_FLUX = 97, VI_MOBIUS = 98
};

#define XO(f) offsetof(flam3_xform, f)
#define XD(x, off) (*(double *)((char *)(x) + (off)))

/* Parameter fields that travel with a parametric variation when alignment
 * borrows them from a neighbouring control point (flam3.c:804-1003).  The
 * reference copies flower_petals twice and never flower_holes; kept. */
static const struct {
   int var;
   int n;
   size_t off[8];
} PARAM_BLOCKS[] = {
   {VI_BLOB, 3, {XO(blob_low), XO(blob_high), XO(blob_waves)}},
   {VI_PDJ, 4, {XO(pdj_a), XO(pdj_b), XO(pdj_c), XO(pdj_d)}},
   {VI_FAN2, 2, {XO(fan2_x), XO(fan2_y)}},
   {VI_RINGS2, 1, {XO(rings2_val)}},
   {VI_PERSPECTIVE, 4, {XO(perspective_angle), XO(perspective_dist), XO(persp_vsin), XO(persp_vfcos)}},
   {VI_JULIAN, 4, {XO(julian_power), XO(julian_dist), XO(julian_rN), XO(julian_cn)}},
   {VI_JULIASCOPE, 4, {XO(juliascope_power), XO(juliascope_dist), XO(juliascope_rN), XO(juliascope_cn)}},
   {VI_RADIAL_BLUR, 1, {XO(radial_blur_angle)}},
   {VI_PIE, 3, {XO(pie_slices), XO(pie_rotation), XO(pie_thickness)}},
   {VI_NGON, 4, {XO(ngon_sides), XO(ngon_power), XO(ngon_corners), XO(ngon_circle)}},
   {VI_CURL, 2, {XO(curl_c1), XO(curl_c2)}},
   {VI_RECTANGLES, 2, {XO(rectangles_x), XO(rectangles_y)}},
   {VI_DISC2, 2, {XO(disc2_rot), XO(disc2_twist)}},
   {VI_SUPER_SHAPE, 6, {XO(super_shape_rnd), XO(super_shape_m), XO(super_shape_n1), XO(super_shape_n2),
                        XO(super_shape_n3), XO(super_shape_holes)}},
   {VI_FLOWER, 1, {XO(flower_petals)}},
   {VI_CONIC, 2, {XO(conic_eccentricity), XO(conic_holes)}},
   {VI_PARABOLA, 2, {XO(parabola_height), XO(parabola_width)}},
   {VI_BENT2, 2, {XO(bent2_x), XO(bent2_y)}},
   {VI_BIPOLAR, 1, {XO(bipolar_shift)}},
   {VI_CELL, 1, {XO(cell_size)}},
   {VI_CPOW, 3, {XO(cpow_i), XO(cpow_r), XO(cpow_power)}},
   {VI_CURVE, 4, {XO(curve_xamp), XO(curve_yamp), XO(curve_xlength), XO(curve_ylength)}},
   {VI_ESCHER, 1, {XO(escher_beta)}},
   {VI_LAZYSUSAN, 5, {XO(lazysusan_x), XO(lazysusan_y), XO(lazysusan_spin), XO(lazysusan_space),
                      XO(lazysusan_twist)}},
   {VI_MODULUS, 2, {XO(modulus_x), XO(modulus_y)}},
   {VI_OSCILLOSCOPE, 4, {XO(oscope_separation), XO(oscope_frequency), XO(oscope_amplitude),
                         XO(oscope_damping)}},
   {VI_POPCORN2, 3, {XO(popcorn2_x), XO(popcorn2_y), XO(popcorn2_c)}},
   {VI_SEPARATION, 4, {XO(separation_x), XO(separation_y), XO(separation_xinside), XO(separation_yinside)}},
   {VI_SPLIT, 2, {XO(split_xsize), XO(split_ysize)}},
   {VI_SPLITS, 2, {XO(splits_x), XO(splits_y)}},
   {VI_STRIPES, 2, {XO(stripes_space), XO(stripes_warp)}},
   {VI_WEDGE, 4, {XO(wedge_angle), XO(wedge_hole), XO(wedge_count), XO(wedge_swirl)}},
   {VI_WEDGE_JULIA, 7, {XO(wedge_julia_angle), XO(wedge_julia_count), XO(wedge_julia_power),
                        XO(wedge_julia_dist), XO(wedgeJulia_cf), XO(wedgeJulia_cn), XO(wedgeJulia_rN)}},
   {VI_WEDGE_SPH, 4, {XO(wedge_sph_angle), XO(wedge_sph_hole), XO(wedge_sph_count), XO(wedge_sph_swirl)}},
   {VI_WHORL, 2, {XO(whorl_inside), XO(whorl_outside)}},
   {VI_WAVES2, 4, {XO(waves2_scalex), XO(waves2_scaley), XO(waves2_freqx), XO(waves2_freqy)}},
   {VI_AUGER, 4, {XO(auger_sym), XO(auger_weight), XO(auger_freq), XO(auger_scale)}},
   {VI_FLUX, 1, {XO(flux_spread)}},
   {VI_MOBIUS, 8, {XO(mobius_re_a), XO(mobius_re_b), XO(mobius_re_c), XO(mobius_re_d), XO(mobius_im_a),
                   XO(mobius_im_b), XO(mobius_im_c), XO(mobius_im_d)}},
};

/* Xform fields blended linearly (interpolation.c:546-638). */
static const size_t XF_BLEND[] = {
   XO(opacity), XO(animate), XO(blob_low), XO(blob_high), XO(blob_waves), XO(pdj_a), XO(pdj_b), XO(pdj_c),
   XO(pdj_d), XO(fan2_x), XO(fan2_y), XO(rings2_val), XO(perspective_angle), XO(perspective_dist),
   XO(julian_power), XO(julian_dist), XO(juliascope_power), XO(juliascope_dist), XO(radial_blur_angle),
   XO(pie_slices), XO(pie_rotation), XO(pie_thickness), XO(ngon_sides), XO(ngon_power), XO(ngon_circle),
   XO(ngon_corners), XO(curl_c1), XO(curl_c2), XO(rectangles_x), XO(rectangles_y), XO(amw_amp), XO(disc2_rot),
   XO(disc2_twist), XO(super_shape_rnd), XO(super_shape_m), XO(super_shape_n1), XO(super_shape_n2),
   XO(super_shape_n3), XO(super_shape_holes), XO(flower_petals), XO(flower_holes), XO(conic_eccentricity),
   XO(conic_holes), XO(parabola_height), XO(parabola_width), XO(bent2_x), XO(bent2_y), XO(bipolar_shift),
   XO(cell_size), XO(cpow_r), XO(cpow_i), XO(cpow_power), XO(curve_xamp), XO(curve_yamp), XO(curve_xlength),
   XO(curve_ylength), XO(escher_beta), XO(lazysusan_x), XO(lazysusan_y), XO(lazysusan_twist),
   XO(lazysusan_space), XO(lazysusan_spin), XO(modulus_x), XO(modulus_y), XO(oscope_separation),
   XO(oscope_frequency), XO(oscope_amplitude), XO(oscope_damping), XO(popcorn2_x), XO(popcorn2_y),
   XO(popcorn2_c), XO(separation_x), XO(separation_xinside), XO(separation_y), XO(separation_yinside),
   XO(split_xsize), XO(split_ysize), XO(splits_x), XO(splits_y), XO(stripes_space), XO(stripes_warp),
   XO(wedge_angle), XO(wedge_hole), XO(wedge_count), XO(wedge_swirl), XO(wedge_julia_angle),
   XO(wedge_julia_count), XO(wedge_julia_power), XO(wedge_julia_dist), XO(wedge_sph_angle), XO(wedge_sph_hole),
   XO(wedge_sph_count), XO(wedge_sph_swirl), XO(whorl_inside), XO(whorl_outside), XO(waves2_scalex),
   XO(waves2_scaley), XO(waves2_freqx), XO(waves2_freqy), XO(auger_sym), XO(auger_weight), XO(auger_freq),
   XO(auger_scale), XO(flux_spread), XO(mobius_re_a), XO(mobius_im_a), XO(mobius_re_b), XO(mobius_im_b),
   XO(mobius_re_c), XO(mobius_im_c), XO(mobius_re_d), XO(mobius_im_d),
};

#define GO(f) offsetof(flam3_genome, f)
/* Genome fields blended linearly (interpolation.c:477-505). */
static const size_t G_BLEND[] = {
   GO(brightness), GO(contrast), GO(highlight_power), GO(gamma), GO(vibrancy), GO(hue_rotation),
   GO(center[0]), GO(center[1]), GO(rot_center[0]), GO(rot_center[1]), GO(background[0]), GO(background[1]),
   GO(background[2]), GO(pixels_per_unit), GO(spatial_filter_radius), GO(temporal_filter_exp),
   GO(temporal_filter_width), GO(sample_density), GO(zoom), GO(rotate), GO(estimator), GO(estimator_minimum),
   GO(estimator_curve), GO(gam_lin_thresh),
};
#define GD(g, off) (*(double *)((char *)(g) + (off)))

static int nstd_of(const flam3_genome *g) { return g->num_xforms - (g->final_xform_index >= 0); }

static void borrow_params(flam3_xform *dst, const flam3_xform *src, int var) {
   size_t b;
   int k;
   for (b = 0; b < sizeof(PARAM_BLOCKS) / sizeof(PARAM_BLOCKS[0]); b++)
      if (PARAM_BLOCKS[b].var == var) {
         for (k = 0; k < PARAM_BLOCKS[b].n; k++) XD(dst, PARAM_BLOCKS[b].off[k]) = XD(src, PARAM_BLOCKS[b].off[k]);
         return;
      }
}

static int copy_xform_deep(flam3_xform *dst, const flam3_xform *src) {
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

/* src's fields with its xform list padded to nstd standard xforms (padding
 * flag set on the added ones) plus a final xform when want_final
 * (flam3_copyx, flam3.c:1211-1268). */
static int copy_padded(flam3_genome *dst, const flam3_genome *src, int nstd, int want_final) {
   int i, srcstd = nstd_of(src);
   *dst = *src;
   dst->num_xforms = 0;
   dst->xform = NULL;
   dst->chaos = NULL;
   dst->final_xform_index = -1;
   if (fl3_genome_add_xforms(dst, nstd, 1, 0)) return 1;
   for (i = 0; i < srcstd; i++) {
      if (copy_xform_deep(&dst->xform[i], &src->xform[i])) return 1;
      memcpy(dst->chaos[i], src->chaos[i], sizeof(double) * srcstd);
   }
   if (want_final) {
      if (fl3_genome_add_xforms(dst, 1, 1, 1)) return 1;
      if (src->final_xform_enable > 0) {
         if (copy_xform_deep(&dst->xform[dst->num_xforms - 1], &src->xform[src->final_xform_index])) return 1;
      } else {
         /* a final xform interpolated against nothing stays still */
         flam3_xform *x = &dst->xform[dst->num_xforms - 1];
         x->num_motion = 0;
         x->motion = NULL;
         x->animate = 0.0;
         x->color_speed = 0.0;
      }
   } else {
      dst->final_xform_index = -1;
      dst->final_xform_enable = 0;
   }
   return 0;
}

static void set_affine(flam3_xform *x, double a, double b, double c, double d) {
   x->c[0][0] = a;
   x->c[0][1] = b;
   x->c[1][0] = c;
   x->c[1][1] = d;
   x->c[2][0] = 0.0;
   x->c[2][1] = 0.0;
}

/* Give every control point the same xform count and final-xform slot, borrow
 * missing parametric-variation parameters from a neighbour, and turn padding
 * xforms into a good "identity" for their neighbours (flam3_align,
 * interpolation.c:768-1031). */
static int align(flam3_genome *dst, const flam3_genome *src, int n) {
   int i, xf, j, ii, fnd, max_nx, max_fx, tnx, aligned = 1;
   const int ref = n / 2 - 1;
   max_nx = nstd_of(&src[0]);
   max_fx = src[0].final_xform_enable;
   for (i = 1; i < n; i++) {
      tnx = nstd_of(&src[i]);
      if (tnx != max_nx) {
         aligned = 0;
         if (tnx > max_nx) max_nx = tnx;
      }
      if (src[i].final_xform_enable != max_fx) {
         aligned = 0;
         max_fx |= src[i].final_xform_enable;
      }
   }
   for (i = 0; i < n; i++)
      if (copy_padded(&dst[i], &src[i], max_nx, max_fx)) return 1;
   if (dst[ref].interpolation_type == flam3_inttype_compat || dst[ref].interpolation_type == flam3_inttype_older)
      return 0;

   for (i = 0; i < n; i++) {
      for (xf = 0; xf < max_nx + max_fx; xf++) {
         flam3_xform *x = &dst[i].xform[xf];
         /* parametric variations start at blob (23) */
         for (j = 23; j < FLAM3_NVARIATIONS; j++) {
            if (x->var[j] != 0) continue;
            if (i > 0) {
               if (dst[i - 1].xform[xf].var[j] != 0) borrow_params(x, &dst[i - 1].xform[xf], j);
            } else if (i < n - 1) {
               if (dst[i + 1].xform[xf].var[j] != 0) borrow_params(x, &dst[i + 1].xform[xf], j);
            }
         }
         if (x->padding != 1 || aligned) continue;

         x->var[VI_LINEAR] = 0.0;
         fnd = 0;
         if ((i == 0 && dst[i].interpolation_type == flam3_inttype_log) ||
             (i > 0 && dst[i - 1].interpolation_type == flam3_inttype_log)) {
            for (ii = -1; ii <= 1; ii += 2) {
               const flam3_xform *nb;
               if (i + ii < 0 || i + ii >= n) continue;
               nb = &dst[i + ii].xform[xf];
               if (nb->padding == 1) continue;
               /* holes: blend against a half-turned identity of weight -1 */
               if (nb->var[VI_SPHERICAL] > 0 || nb->var[VI_NGON] > 0 || nb->var[VI_JULIAN] > 0 ||
                   nb->var[VI_JULIASCOPE] > 0 || nb->var[VI_POLAR] > 0 || nb->var[VI_WEDGE_SPH] > 0 ||
                   nb->var[VI_WEDGE_JULIA] > 0) {
                  x->var[VI_LINEAR] = -1.0;
                  set_affine(x, -1.0, 0.0, 0.0, -1.0);
                  fnd = -1;
               }
            }
         }
         if (fnd == 0) {
            for (ii = -1; ii <= 1; ii += 2) {
               const flam3_xform *nb;
               if (i + ii < 0 || i + ii >= n) continue;
               nb = &dst[i + ii].xform[xf];
               if (nb->padding == 1) continue;
               if (nb->var[VI_RECTANGLES] > 0) {
                  x->var[VI_RECTANGLES] = 1.0;
                  x->rectangles_x = 0.0;
                  x->rectangles_y = 0.0;
                  fnd++;
               }
               if (nb->var[VI_RINGS2] > 0) {
                  x->var[VI_RINGS2] = 1.0;
                  x->rings2_val = 0.0;
                  fnd++;
               }
               if (nb->var[VI_FAN2] > 0) {
                  x->var[VI_FAN2] = 1.0;
                  x->fan2_x = 0.0;
                  x->fan2_y = 0.0;
                  fnd++;
               }
               if (nb->var[VI_BLOB] > 0) {
                  x->var[VI_BLOB] = 1.0;
                  x->blob_low = 1.0;
                  x->blob_high = 1.0;
                  x->blob_waves = 1.0;
                  fnd++;
               }
               if (nb->var[VI_PERSPECTIVE] > 0) {
                  x->var[VI_PERSPECTIVE] = 1.0;
                  x->perspective_angle = 0.0;
                  fnd++;
               }
               if (nb->var[VI_CURL] > 0) {
                  x->var[VI_CURL] = 1.0;
                  x->curl_c1 = 0.0;
                  x->curl_c2 = 0.0;
                  fnd++;
               }
               if (nb->var[VI_SUPER_SHAPE] > 0) {
                  x->var[VI_SUPER_SHAPE] = 1.0;
                  x->super_shape_n1 = 2.0;
                  x->super_shape_n2 = 2.0;
                  x->super_shape_n3 = 2.0;
                  x->super_shape_rnd = 0.0;
                  x->super_shape_holes = 0.0;
                  fnd++;
               }
            }
         }
         if (fnd == 0) {
            for (ii = -1; ii <= 1; ii += 2) {
               const flam3_xform *nb;
               if (i + ii < 0 || i + ii >= n) continue;
               nb = &dst[i + ii].xform[xf];
               if (nb->padding == 1) continue;
               if (nb->var[VI_FAN] > 0) {
                  x->var[VI_FAN] = 1.0;
                  fnd++;
               }
               if (nb->var[VI_RINGS] > 0) {
                  x->var[VI_RINGS] = 1.0;
                  fnd++;
               }
            }
            if (fnd > 0) set_affine(x, 0.0, 1.0, 1.0, 0.0);
         }
         if (fnd == 0)
            x->var[VI_LINEAR] = 1.0;
         else if (fnd > 0) {
            double sum = 0.0;
            for (j = 0; j < FLAM3_NVARIATIONS; j++) sum += x->var[j];
            for (j = 0; j < FLAM3_NVARIATIONS; j++) x->var[j] /= sum;
         }
      }
   }
   return 0;
}

/* ---------------------------------------------------------------- affines */

static int is_identity(double m[3][2]) {
   return m[0][0] == 1.0 && m[0][1] == 0.0 && m[1][0] == 0.0 && m[1][1] == 1.0 && m[2][0] == 0.0 &&
          m[2][1] == 0.0;
}

/* Columns of xform xfi's affine (post when use_post) as angle, magnitude
 * and translation per control point, angles unwrapped along the control
 * points (interpolation.c:247-324). */
static void to_polar(const flam3_genome *cp, int n, int xfi, int use_post, double ang[4][2], double mag[4][2],
                     double trn[4][2]) {
   int k, col;
   for (k = 0; k < n; k++) {
      int zero[2] = {0, 0};
      const flam3_xform *x = &cp[k].xform[xfi];
      for (col = 0; col < 2; col++) {
         const double a0 = use_post ? x->post[col][0] : x->c[col][0];
         const double a1 = use_post ? x->post[col][1] : x->c[col][1];
         ang[k][col] = atan2(a1, a0);
         mag[k][col] = sqrt(a0 * a0 + a1 * a1);
         if (mag[k][col] == 0.0) zero[col] = 1;
         trn[k][col] = use_post ? x->post[2][col] : x->c[2][col];
      }
      if (zero[0] && !zero[1])
         ang[k][0] = ang[k][1];
      else if (!zero[0] && zero[1])
         ang[k][1] = ang[k][0];
   }
   for (col = 0; col < 2; col++) {
      for (k = 1; k < n; k++) {
         const double wind = cp[k].xform[xfi].wind[col];
         if (wind > 0 && !use_post) {
            const double lo = wind - 2 * M_PI;
            while (ang[k - 1][col] < lo) ang[k - 1][col] += 2 * M_PI;
            while (ang[k - 1][col] > lo + 2 * M_PI) ang[k - 1][col] -= 2 * M_PI;
            while (ang[k][col] < lo) ang[k][col] += 2 * M_PI;
            while (ang[k][col] > lo + 2 * M_PI) ang[k][col] -= 2 * M_PI;
         } else {
            const double d = ang[k][col] - ang[k - 1][col];
            if (d > M_PI + IEPS)
               ang[k][col] -= 2 * M_PI;
            else if (d < -(M_PI - IEPS))
               ang[k][col] += 2 * M_PI;
         }
      }
   }
}

/* Blend angles linearly and magnitudes logarithmically (linearly when one
 * is below e^-10), translations linearly, and convert back
 * (interpolation.c:194-245).  out must be cleared. */
static void from_polar(const double *c, int n, double ang[4][2], double mag[4][2], double trn[4][2],
                       double out[3][2]) {
   double acc_ang[2] = {0.0, 0.0}, acc_mag[2] = {0.0, 0.0};
   int linear_mag[2] = {0, 0}, k, col;
   for (col = 0; col < 2; col++)
      for (k = 0; k < n; k++)
         if (log(mag[k][col]) < -10) linear_mag[col] = 1;
   for (k = 0; k < n; k++)
      for (col = 0; col < 2; col++) {
         acc_ang[col] += c[k] * ang[k][col];
         if (linear_mag[col])
            acc_mag[col] += c[k] * mag[k][col];
         else
            acc_mag[col] += c[k] * log(mag[k][col]);
         out[2][col] += c[k] * trn[k][col];
      }
   for (col = 0; col < 2; col++) {
      const double m = linear_mag[col] ? acc_mag[col] : exp(acc_mag[col]);
      out[col][0] = m * cos(acc_ang[col]);
      out[col][1] = m * sin(acc_ang[col]);
   }
}

static void clear3x2(double m[3][2]) { memset(m, 0, sizeof(double) * 6); }

static void add_scaled3x2(double s, double src[3][2], double dst[3][2]) {
   int a, b;
   for (a = 0; a < 3; a++)
      for (b = 0; b < 2; b++) dst[a][b] += s * src[a][b];
}

/* Weight of control point 0 for xform i when the xforms start their
 * transitions one after another (interpolation.c:339-369). */
static double stagger_weight(double t, double stagger, int nx, int i) {
   const double spread = stagger * ((double)(nx - 1) / nx);
   const double st = spread * (nx - 1 - i) / (nx - 1);
   const double et = st + (1 - spread);
   double u;
   if (t <= st) return 0;
   if (t >= et) return 1;
   u = (t - st) / (1 - spread);
   return 3 * u * u - 2 * u * u * u;
}

/* ---------------------------------------------------------------- blend */

static void blend_palette(flam3_genome *res, const flam3_genome *cpi, int n, const double *c) {
   int i, k, l;
   const int mode = cpi[0].palette_interpolation;
   if (mode == flam3_palette_interpolation_sweep) {
      for (i = 0; i < 256; i++) res->palette[i] = cpi[(i < 256 * c[0]) ? 0 : 1].palette[i];
      return;
   }
   double rgb_part = 0.0;
   if (mode == flam3_palette_interpolation_rgb)
      rgb_part = 1.0;
   else if (mode == flam3_palette_interpolation_hsv_circular)
      rgb_part = cpi[0].hsv_rgb_palette_blend;
   for (i = 0; i < 256; i++) {
      double rgb[3] = {0, 0, 0}, hsv[3] = {0, 0, 0}, back[3], count = 0, index = 0;
      int opaque = 1;
      for (k = 0; k < n; k++) {
         const flam3_palette_entry *e = &cpi[k].palette[i];
         double h[3];
         fl3_rgb2hsv(e->color, h);
         if (n == 2 && k == 0 && mode == flam3_palette_interpolation_hsv_circular) {
            /* hue of the first point moved so the blend takes the short way round */
            double h2[3];
            fl3_rgb2hsv(cpi[1].palette[i].color, h2);
            if (h2[0] - h[0] > 3.0)
               h[0] += 6.0;
            else if (h2[0] - h[0] < -3.0)
               h[0] -= 6.0;
         }
         for (l = 0; l < 3; l++) {
            rgb[l] += c[k] * e->color[l];
            hsv[l] += c[k] * h[l];
         }
         count += c[k] * e->color[3];
         if (e->color[3] != 1.0) opaque = 0;
         index += c[k] * e->index;
      }
      if (opaque) count = 1.0;
      fl3_hsv2rgb(hsv, back);
      flam3_palette_entry *o = &res->palette[i];
      for (l = 0; l < 3; l++) o->color[l] = rgb_part * rgb[l] + (1.0 - rgb_part) * back[l];
      o->color[3] = count;
      o->index = index;
      for (l = 0; l < 4; l++) {
         if (o->color[l] < 0.0) o->color[l] = 0.0;
         if (o->color[l] > 1.0) o->color[l] = 1.0;
      }
      if (o->index < 0.0) o->index = 0.0;
      if (o->index > 255.0) o->index = 255.0;
   }
}

static double lin(const double *c, int n, const flam3_genome *cpi, size_t off) {
   double v = 0.0;
   int k;
   for (k = 0; k < n; k++) v += c[k] * GD(&cpi[k], off);
   return v;
}

#define BLEND_INT(field)                                              \
   do {                                                               \
      double tt_ = 0.0;                                               \
      for (k = 0; k < n; k++) tt_ += c[k] * cpi[k].field;             \
      res->field = (int)rint(tt_);                                    \
   } while (0)

#define BLEND_X(expr_dst, field)                                      \
   do {                                                               \
      double v_ = 0.0;                                                \
      for (k = 0; k < n; k++) v_ += c[k] * cpi[k].xform[i].field;     \
      expr_dst = v_;                                                  \
   } while (0)

/* res: cleared genome with the aligned xform layout.  c is modified and
 * restored when staggering (interpolation.c:373-708). */
static void blend(flam3_genome *res, int n, const flam3_genome *cpi, double *c, double stagger) {
   int i, j, k, numstd;
   size_t f;
   blend_palette(res, cpi, n, c);
   res->palette_index = -1;
   res->symmetry = 0;
   res->spatial_filter_select = cpi[0].spatial_filter_select;
   res->temporal_filter_type = cpi[0].temporal_filter_type;
   res->palette_mode = cpi[0].palette_mode;
   res->interpolation_type = cpi[0].interpolation_type;
   res->palette_interpolation = cpi[0].palette_interpolation;
   res->hsv_rgb_palette_blend = cpi[0].hsv_rgb_palette_blend;
   for (f = 0; f < sizeof(G_BLEND) / sizeof(G_BLEND[0]); f++) GD(res, G_BLEND[f]) = lin(c, n, cpi, G_BLEND[f]);
   BLEND_INT(width);
   BLEND_INT(height);
   BLEND_INT(spatial_oversample);
   BLEND_INT(nbatches);
   BLEND_INT(ntemporal_samples);

   numstd = nstd_of(&cpi[0]);
   for (i = 0; i < numstd; i++)
      for (j = 0; j < numstd; j++) {
         double v = 0.0;
         for (k = 0; k < n; k++) v += c[k] * cpi[k].chaos[i][j];
         res->chaos[i][j] = v < 0 ? 0 : v;
      }

   for (i = 0; i < cpi[0].num_xforms; i++) {
      flam3_xform *x = &res->xform[i];
      const int staggered = n == 2 && stagger > 0 && i != cpi[0].final_xform_index;
      double save[2] = {c[0], c[1]};
      double v;
      if (staggered) {
         c[0] = stagger_weight(save[0], stagger, numstd, i);
         c[1] = 1.0 - c[0];
      }
      BLEND_X(v, density);
      x->density = v < 0.0 ? 0.0 : v;
      BLEND_X(x->color, color);
      if (x->color < 0) x->color = 0;
      if (x->color > 1) x->color = 1;
      BLEND_X(x->color_speed, color_speed);
      if (x->color_speed < 0) x->color_speed = 0;
      if (x->color_speed > 1) x->color_speed = 1;
      for (f = 0; f < sizeof(XF_BLEND) / sizeof(XF_BLEND[0]); f++) {
         double s = 0.0;
         for (k = 0; k < n; k++) s += c[k] * XD(&cpi[k].xform[i], XF_BLEND[f]);
         XD(x, XF_BLEND[f]) = s;
      }
      for (j = 0; j < FLAM3_NVARIATIONS; j++) BLEND_X(x->var[j], var[j]);

      {
         int all_id = 1;
         for (k = 0; k < n; k++) all_id &= is_identity(((flam3_genome *)cpi)[k].xform[i].post);
         if (cpi[0].interpolation_type == flam3_inttype_log) {
            double ang[4][2], mag[4][2], trn[4][2];
            clear3x2(x->c);
            to_polar(cpi, n, i, 0, ang, mag, trn);
            from_polar(c, n, ang, mag, trn, x->c);
            clear3x2(x->post);
            if (all_id) {
               x->post[0][0] = 1.0;
               x->post[1][1] = 1.0;
            } else {
               to_polar(cpi, n, i, 1, ang, mag, trn);
               from_polar(c, n, ang, mag, trn, x->post);
            }
         } else {
            clear3x2(x->c);
            clear3x2(x->post);
            for (k = 0; k < n; k++) {
               add_scaled3x2(c[k], ((flam3_genome *)cpi)[k].xform[i].c, x->c);
               add_scaled3x2(c[k], ((flam3_genome *)cpi)[k].xform[i].post, x->post);
            }
            if (all_id) {
               clear3x2(x->post);
               x->post[0][0] = 1.0;
               x->post[1][1] = 1.0;
            }
         }
      }
      if (staggered) {
         c[0] = save[0];
         c[1] = save[1];
      }
   }
}

static void free_all(flam3_genome *g, int n) {
   int i;
   for (i = 0; i < n; i++) fl3_genome_free(&g[i]);
}

int fl3_interpolate_multi(const flam3_genome *cps, int ncps, double t, double stagger, flam3_genome *dst) {
   int i1, i2, smooth = 0, nalign, rc = 1;
   double c[4];
   flam3_genome cpi[4];
   memset(cpi, 0, sizeof(cpi));
   if (cps[0].time >= t) {
      i1 = 0;
      i2 = 1;
   } else if (cps[ncps - 1].time <= t) {
      i1 = ncps - 2;
      i2 = ncps - 1;
   } else {
      i1 = 0;
      while (cps[i1].time < t) i1++;
      i1--;
      i2 = i1 + 1;
   }
   c[0] = (cps[i2].time - t) / (cps[i2].time - cps[i1].time);
   c[1] = 1.0 - c[0];

   if (cps[i1].interpolation != flam3_interpolation_linear) {
      /* Catmull-Rom over four control points needs one on each side; the
       * reference reports the first and last segments and then reads past
       * the array; here they fall back to linear. */
      if (i1 == 0 || i2 == ncps - 1)
         fprintf(stderr, "error: cannot use smooth interpolation on %s segment.\nreverting to linear interpolation.\n",
                 i1 == 0 ? "first" : "last");
      else
         smooth = 1;
   }
   nalign = smooth ? 4 : 2;
   if (align(cpi, smooth ? &cps[i1 - 1] : &cps[i1], nalign)) goto out;

   fl3_genome_free(dst);
   fl3_genome_defaults(dst, 1);
   dst->num_xforms = 0;
   dst->xform = NULL;
   dst->chaos = NULL;
   dst->final_xform_index = -1;
   dst->final_xform_enable = 0;
   if (cpi[0].final_xform_index >= 0) {
      if (fl3_genome_add_xforms(dst, cpi[0].num_xforms - 1, 0, 0)) goto out;
      if (fl3_genome_add_xforms(dst, 1, 0, 1)) goto out;
   } else if (fl3_genome_add_xforms(dst, cpi[0].num_xforms, 0, 0))
      goto out;
   dst->time = t;
   dst->interpolation = flam3_interpolation_linear;
   dst->interpolation_type = cpi[0].interpolation_type;
   dst->palette_interpolation = flam3_palette_interpolation_hsv_circular;
   dst->hsv_rgb_palette_blend = 0.0;

   if (!smooth) {
      blend(dst, 2, cpi, c, stagger);
   } else {
      /* Catmull-Rom weights at u = c[1] (interpolation.c:326-337) */
      const double u = c[1], u2 = u * u, u3 = u2 * u;
      double w[4];
      w[0] = (2 * u2 - u - u3) / 2;
      w[1] = (3 * u3 - 5 * u2 + 2) / 2;
      w[2] = (4 * u2 - 3 * u3 + u) / 2;
      w[3] = (u3 - u2) / 2;
      blend(dst, 4, cpi, w, 0);
   }
   rc = 0;
out:
   free_all(cpi, nalign);
   return rc;
}
