/*
 * flam3_abi.h - the data types that cross the flam3_render boundary.
 *
 * These declarations are layout-compatible with the reference's public
 * header (flam3.h:185-191 stat_struct, flam3.h:212-464 flam3_xform,
 * flam3.h:466-541 flam3_genome, flam3.h:602-615 flam3_frame, isaac.h:26-35
 * randctx, isaacs.h:17 ub4 = unsigned long).  A program built against the
 * reference flam3.h can hand its flam3_frame to this library unchanged: the
 * field order, types and padding are identical (checked by the CPU tests,
 * which compare sizeof/offsetof against the reference build).  Field names
 * are kept so code reads like the reference; comments are this project's.
 *
 * The library never includes libxml2: the one libxml2 type in a genome
 * (the edit history document) is carried as an opaque pointer.
 */
#ifndef FLAM3_ABI_H
#define FLAM3_ABI_H

#ifdef __cplusplus
extern "C" {
#endif

#define FLAM3_NVARIATIONS 99
#define FLAM3_NAME_LEN 64
#define FLAM3_PARENT_FN_LEN 30
#define FLAM3_RANDSIZL 4
#define FLAM3_RANDSIZ (1 << FLAM3_RANDSIZL)

/* ISAAC context.  On LP64 every word is 64 bits wide; only the seeding mix
 * sees the upper halves (isaac.c:48-58). */
typedef unsigned long ub4;
typedef struct randctx {
   ub4 randcnt;
   ub4 randrsl[FLAM3_RANDSIZ];
   ub4 randmem[FLAM3_RANDSIZ];
   ub4 randa;
   ub4 randb;
   ub4 randc;
} randctx;

typedef struct {
   double badvals;
   long int num_iters;
   int render_seconds;
} stat_struct;

typedef struct {
   double index;
   double color[4]; /* r, g, b premultiplied by a, and a; each 0..1 */
} flam3_palette_entry;
typedef flam3_palette_entry flam3_palette[256];

/* One affine + variations function of the IFS. */
typedef struct xform {
   double var[FLAM3_NVARIATIONS]; /* blend weight of each variation */
   double c[3][2];                /* affine: x' = c00 x + c10 y + c20 */
   double post[3][2];             /* post affine, same convention */
   double density;                /* selection weight */
   double color;                  /* colour coordinate this map pulls toward */
   double color_speed;
   double animate;
   double opacity;
   double vis_adjusted;           /* opacity remapped by adjust_percentage */
   int padding;
   double wind[2];
   int precalc_angles_flag;
   int precalc_atan_xy_flag;
   int precalc_atan_yx_flag;
   double has_preblur;            /* pre_blur weight, 0 when absent */
   int has_post;

   /* variation parameters */
   double blob_low, blob_high, blob_waves;
   double pdj_a, pdj_b, pdj_c, pdj_d;
   double fan2_x, fan2_y;
   double rings2_val;
   double perspective_angle, perspective_dist;
   double julian_power, julian_dist;
   double juliascope_power, juliascope_dist;
   double radial_blur_angle;
   double pie_slices, pie_rotation, pie_thickness;
   double ngon_sides, ngon_power, ngon_circle, ngon_corners;
   double curl_c1, curl_c2;
   double rectangles_x, rectangles_y;
   double amw_amp;
   double disc2_rot, disc2_twist;
   double super_shape_rnd, super_shape_m, super_shape_n1, super_shape_n2,
          super_shape_n3, super_shape_holes;
   double flower_petals, flower_holes;
   double conic_eccentricity, conic_holes;
   double parabola_height, parabola_width;
   double bent2_x, bent2_y;
   double bipolar_shift;
   double cell_size;
   double cpow_r, cpow_i, cpow_power;
   double curve_xamp, curve_yamp;
   double curve_xlength, curve_ylength;
   double escher_beta;
   double lazysusan_spin, lazysusan_space, lazysusan_twist;
   double lazysusan_x, lazysusan_y;
   double modulus_x, modulus_y;
   double oscope_separation, oscope_frequency, oscope_amplitude, oscope_damping;
   double popcorn2_x, popcorn2_y, popcorn2_c;
   double separation_x, separation_xinside;
   double separation_y, separation_yinside;
   double split_xsize, split_ysize;
   double splits_x, splits_y;
   double stripes_space, stripes_warp;
   double wedge_angle, wedge_hole;
   double wedge_count, wedge_swirl;
   double wedge_julia_angle, wedge_julia_count, wedge_julia_power, wedge_julia_dist;
   double wedge_sph_angle, wedge_sph_count;
   double wedge_sph_hole, wedge_sph_swirl;
   double whorl_inside, whorl_outside;
   double waves2_freqx, waves2_scalex;
   double waves2_freqy, waves2_scaley;
   double auger_sym, auger_weight;
   double auger_freq, auger_scale;
   double flux_spread;
   double mobius_re_a, mobius_im_a;
   double mobius_re_b, mobius_im_b;
   double mobius_re_c, mobius_im_c;
   double mobius_re_d, mobius_im_d;

   /* values derived from the parameters before iterating */
   double persp_vsin, persp_vfcos;
   double julian_rN, julian_cn;
   double juliascope_rN, juliascope_cn;
   double wedgeJulia_rN, wedgeJulia_cn, wedgeJulia_cf;
   double radialBlur_spinvar, radialBlur_zoomvar;
   double waves_dx2, waves_dy2;
   double disc2_sinadd, disc2_cosadd, disc2_timespi;
   double super_shape_pm_4, super_shape_pneg1_n1;

   int num_active_vars;
   double active_var_weights[FLAM3_NVARIATIONS];
   int varFunc[FLAM3_NVARIATIONS];

   int motion_freq;
   int motion_func;
   struct xform *motion;
   int num_motion;
} flam3_xform;

typedef struct {
   char flame_name[FLAM3_NAME_LEN + 1];
   double time;
   int interpolation;
   int interpolation_type;
   int palette_interpolation;
   double hsv_rgb_palette_blend;
   int num_xforms;
   int final_xform_index;
   int final_xform_enable;
   flam3_xform *xform;
   double **chaos;                /* xaos matrix, num_std x num_std */
   int chaos_enable;
   int genome_index;
   char parent_fname[FLAM3_PARENT_FN_LEN];
   int symmetry;
   flam3_palette palette;
   char *input_image;
   int palette_index;
   double brightness;
   double contrast;
   double gamma;
   double highlight_power;
   int width, height;             /* final image size in pixels */
   int spatial_oversample;
   double center[2];
   double rot_center[2];
   double rotate;                 /* degrees */
   double vibrancy;
   double hue_rotation;
   double background[3];
   double zoom;
   double pixels_per_unit;
   double spatial_filter_radius;
   int spatial_filter_select;
   double sample_density;         /* "quality": samples per output pixel */
   int nbatches;
   int ntemporal_samples;
   double estimator;              /* density-estimation max radius */
   double estimator_curve;
   double estimator_minimum;
   void *edits;                   /* libxml2 xmlDocPtr in the reference */
   double gam_lin_thresh;
   int palette_index0;
   double hue_rotation0;
   int palette_index1;
   double hue_rotation1;
   double palette_blend;
   int temporal_filter_type;
   double temporal_filter_width, temporal_filter_exp;
   int palette_mode;
} flam3_genome;

typedef struct {
   double pixel_aspect_ratio;
   flam3_genome *genomes;
   int ngenomes;
   int verbose;
   int bits;
   int bytes_per_channel;
   int earlyclip;
   double time;
   int (*progress)(void *, double, int, double);
   void *progress_parameter;
   randctx rc;
   int nthreads;
   int sub_batch_size;
} flam3_frame;

/* Constants with the reference's values (flam3.h). */
#define flam3_field_both 0
#define flam3_field_even 1
#define flam3_field_odd 2

#define flam3_palette_mode_step 0
#define flam3_palette_mode_linear 1

#define flam3_interpolation_linear 0
#define flam3_interpolation_smooth 1

#define flam3_inttype_linear 0
#define flam3_inttype_log 1
#define flam3_inttype_compat 2
#define flam3_inttype_older 3

#define flam3_palette_interpolation_hsv 0
#define flam3_palette_interpolation_sweep 1
#define flam3_palette_interpolation_rgb 2
#define flam3_palette_interpolation_hsv_circular 3

#define flam3_temporal_box 0
#define flam3_temporal_gaussian 1
#define flam3_temporal_exp 2

#define flam3_gaussian_kernel 0

#ifdef __cplusplus
}
#endif
#endif /* FLAM3_ABI_H */
