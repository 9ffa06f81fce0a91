/*
 * Packing of one prepared genome (plus camera and palette) into the job
 * blob the kernels load into LDS (layout in fl3_pack.h).
 *
 * Variations are re-indexed by the set used anywhere in the genome ("union
 * slots", ascending variation number): a workgroup loops over the union with
 * a wave-uniform variation id and each walker multiplies by its own xform's
 * weight for that slot, skipping zeros.  Per walker this adds the same terms
 * in the same order as the reference's active-variation list
 * (variations.c:2043-2048, 2177-2381).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "fl3_host.h"
#include "fl3_pack.h"

#define VAR_PRE_BLUR 67
#define VAR_JULIA 13

static int usable(const flam3_genome *cp, int i) {
   if (i == cp->final_xform_index) return 1;
   return cp->xform[i].density > 0.0;
}

static void pack_xform(const flam3_xform *x, fl3_xform_pack *d) {
   int k;
   double *p = d->p;
   memset(d, 0, sizeof(*d));
   for (k = 0; k < 3; k++) {
      d->c[2 * k] = x->c[k][0];
      d->c[2 * k + 1] = x->c[k][1];
      d->post[2 * k] = x->post[k][0];
      d->post[2 * k + 1] = x->post[k][1];
   }
   d->color = x->color;
   d->color_speed = x->color_speed;
   d->vis = x->vis_adjusted;
   d->opacity = x->opacity;
   d->preblur = x->has_preblur;
   d->flags = (x->precalc_atan_xy_flag > 0 ? FL3_PRE_ATAN_XY : 0) |
              (x->precalc_angles_flag > 0 ? FL3_PRE_ANGLES : 0) |
              (x->precalc_atan_yx_flag > 0 ? FL3_PRE_ATAN_YX : 0) |
              (x->has_post ? FL3_HAS_POST : 0);
   p[P_BLOB_LOW] = x->blob_low; p[P_BLOB_HIGH] = x->blob_high; p[P_BLOB_WAVES] = x->blob_waves;
   p[P_PDJ_A] = x->pdj_a; p[P_PDJ_B] = x->pdj_b; p[P_PDJ_C] = x->pdj_c; p[P_PDJ_D] = x->pdj_d;
   p[P_FAN2_X] = x->fan2_x; p[P_FAN2_Y] = x->fan2_y; p[P_RINGS2_VAL] = x->rings2_val;
   p[P_PERSP_DIST] = x->perspective_dist; p[P_PERSP_VSIN] = x->persp_vsin; p[P_PERSP_VFCOS] = x->persp_vfcos;
   p[P_JULIAN_POWER] = x->julian_power; p[P_JULIAN_RN] = x->julian_rN; p[P_JULIAN_CN] = x->julian_cn;
   p[P_JSCOPE_POWER] = x->juliascope_power; p[P_JSCOPE_RN] = x->juliascope_rN; p[P_JSCOPE_CN] = x->juliascope_cn;
   p[P_RB_SPINVAR] = x->radialBlur_spinvar; p[P_RB_ZOOMVAR] = x->radialBlur_zoomvar;
   p[P_PIE_SLICES] = x->pie_slices; p[P_PIE_ROTATION] = x->pie_rotation; p[P_PIE_THICKNESS] = x->pie_thickness;
   p[P_NGON_SIDES] = x->ngon_sides; p[P_NGON_POWER] = x->ngon_power;
   p[P_NGON_CIRCLE] = x->ngon_circle; p[P_NGON_CORNERS] = x->ngon_corners;
   p[P_CURL_C1] = x->curl_c1; p[P_CURL_C2] = x->curl_c2;
   p[P_RECT_X] = x->rectangles_x; p[P_RECT_Y] = x->rectangles_y;
   p[P_DISC2_TIMESPI] = x->disc2_timespi; p[P_DISC2_SINADD] = x->disc2_sinadd; p[P_DISC2_COSADD] = x->disc2_cosadd;
   p[P_SS_RND] = x->super_shape_rnd; p[P_SS_N2] = x->super_shape_n2; p[P_SS_N3] = x->super_shape_n3;
   p[P_SS_HOLES] = x->super_shape_holes; p[P_SS_PM4] = x->super_shape_pm_4; p[P_SS_PNEG1_N1] = x->super_shape_pneg1_n1;
   p[P_FLOWER_PETALS] = x->flower_petals; p[P_FLOWER_HOLES] = x->flower_holes;
   p[P_CONIC_ECC] = x->conic_eccentricity; p[P_CONIC_HOLES] = x->conic_holes;
   p[P_PARA_H] = x->parabola_height; p[P_PARA_W] = x->parabola_width;
   p[P_BENT2_X] = x->bent2_x; p[P_BENT2_Y] = x->bent2_y;
   p[P_BIPOLAR_SHIFT] = x->bipolar_shift; p[P_CELL_SIZE] = x->cell_size;
   p[P_CPOW_R] = x->cpow_r; p[P_CPOW_I] = x->cpow_i; p[P_CPOW_POWER] = x->cpow_power;
   p[P_CURVE_XAMP] = x->curve_xamp; p[P_CURVE_YAMP] = x->curve_yamp;
   p[P_CURVE_XLEN] = x->curve_xlength; p[P_CURVE_YLEN] = x->curve_ylength;
   p[P_ESCHER_BETA] = x->escher_beta;
   p[P_LS_SPIN] = x->lazysusan_spin; p[P_LS_SPACE] = x->lazysusan_space; p[P_LS_TWIST] = x->lazysusan_twist;
   p[P_LS_X] = x->lazysusan_x; p[P_LS_Y] = x->lazysusan_y;
   p[P_MOD_X] = x->modulus_x; p[P_MOD_Y] = x->modulus_y;
   p[P_OSC_SEP] = x->oscope_separation; p[P_OSC_FREQ] = x->oscope_frequency;
   p[P_OSC_AMP] = x->oscope_amplitude; p[P_OSC_DAMP] = x->oscope_damping;
   p[P_POP2_X] = x->popcorn2_x; p[P_POP2_Y] = x->popcorn2_y; p[P_POP2_C] = x->popcorn2_c;
   p[P_SEP_X] = x->separation_x; p[P_SEP_XIN] = x->separation_xinside;
   p[P_SEP_Y] = x->separation_y; p[P_SEP_YIN] = x->separation_yinside;
   p[P_SPLIT_X] = x->split_xsize; p[P_SPLIT_Y] = x->split_ysize;
   p[P_SPLITS_X] = x->splits_x; p[P_SPLITS_Y] = x->splits_y;
   p[P_STRIPES_SPACE] = x->stripes_space; p[P_STRIPES_WARP] = x->stripes_warp;
   p[P_WEDGE_ANGLE] = x->wedge_angle; p[P_WEDGE_HOLE] = x->wedge_hole;
   p[P_WEDGE_COUNT] = x->wedge_count; p[P_WEDGE_SWIRL] = x->wedge_swirl;
   p[P_WJ_ANGLE] = x->wedge_julia_angle; p[P_WJ_COUNT] = x->wedge_julia_count;
   p[P_WJ_POWER] = x->wedge_julia_power; p[P_WJ_RN] = x->wedgeJulia_rN;
   p[P_WJ_CN] = x->wedgeJulia_cn; p[P_WJ_CF] = x->wedgeJulia_cf;
   p[P_WS_ANGLE] = x->wedge_sph_angle; p[P_WS_COUNT] = x->wedge_sph_count;
   p[P_WS_HOLE] = x->wedge_sph_hole; p[P_WS_SWIRL] = x->wedge_sph_swirl;
   p[P_WHORL_IN] = x->whorl_inside; p[P_WHORL_OUT] = x->whorl_outside;
   p[P_W2_FREQX] = x->waves2_freqx; p[P_W2_SCALEX] = x->waves2_scalex;
   p[P_W2_FREQY] = x->waves2_freqy; p[P_W2_SCALEY] = x->waves2_scaley;
   p[P_AUGER_SYM] = x->auger_sym; p[P_AUGER_WEIGHT] = x->auger_weight;
   p[P_AUGER_FREQ] = x->auger_freq; p[P_AUGER_SCALE] = x->auger_scale;
   p[P_FLUX_SPREAD] = x->flux_spread;
   p[P_MOB_RE_A] = x->mobius_re_a; p[P_MOB_IM_A] = x->mobius_im_a;
   p[P_MOB_RE_B] = x->mobius_re_b; p[P_MOB_IM_B] = x->mobius_im_b;
   p[P_MOB_RE_C] = x->mobius_re_c; p[P_MOB_IM_C] = x->mobius_im_c;
   p[P_MOB_RE_D] = x->mobius_re_d; p[P_MOB_IM_D] = x->mobius_im_d;
   p[P_WAVES_DX2] = x->waves_dx2; p[P_WAVES_DY2] = x->waves_dy2;
}

static size_t align8(size_t n) { return (n + 7) & ~(size_t)7; }

/* distrib: the selection table, or starts: its breakpoints
 * (fl3_xform_starts), or neither (built here). */
static void *pack_job(const flam3_genome *cp, const unsigned short *distrib, const int *starts,
                      const fl3h_camera *cam, size_t *bytes) {
   int nxf = cp->num_xforms, nstd = nxf - (cp->final_xform_index >= 0);
   int used[FL3_NVAR], nunion = 0, i, j, u, nrows;
   unsigned short *own = NULL;
   size_t off, total;
   unsigned char *blob;
   fl3_job_hdr *h;

   if (nxf < 1 || nxf > FL3_MAX_XFORMS || nstd < 1) {
      fprintf(stderr, "flam3_amd: genomes need 1..%d xforms (got %d)\n", FL3_MAX_XFORMS, nxf);
      return NULL;
   }
   memset(used, 0, sizeof(used));
   for (i = 0; i < nxf; i++) {
      const flam3_xform *x = &cp->xform[i];
      if (!usable(cp, i)) continue;
      if (x->num_active_vars < 0 || x->num_active_vars > FL3_NVAR) {
         fprintf(stderr, "flam3_amd: xform %d has not been prepared\n", i);
         return NULL;
      }
      for (j = 0; j < x->num_active_vars; j++) {
         int v = x->varFunc[j];
         if (v >= 0 && v < FL3_NVAR && v != VAR_PRE_BLUR) used[v] = 1;
      }
   }
   for (j = 0; j < FL3_NVAR; j++)
      if (used[j]) nunion++;

   if (!distrib && !starts) {
      own = fl3_create_xform_distrib((flam3_genome *)cp, &nrows);
      if (!own) return NULL;
      distrib = own;
   }
   nrows = cp->chaos_enable ? nstd + 1 : 1;

   off = align8(sizeof(fl3_job_hdr));
   {
      size_t o_xf = off, o_w, o_st, o_dm, o_vl;
      off += sizeof(fl3_xform_pack) * nxf;
      o_w = off;
      off = align8(off + sizeof(double) * nxf * (nunion > 0 ? nunion : 1));
      o_st = off;
      off = align8(off + sizeof(int32_t) * nrows * nstd);
      o_dm = off;
      off += sizeof(double) * 1024;
      o_vl = off;
      /* (and four pairs of padding: a walker reads an xform's first four
       * pairs at once, whatever its count) */
      off = align8(off + sizeof(int32_t) * nxf * (2 + 2 * (size_t)nunion) + 8 * sizeof(int32_t));
      total = off;
      blob = (unsigned char *)calloc(1, total);
      if (!blob) { free(own); return NULL; }
      h = (fl3_job_hdr *)blob;
      h->off_xf = (int64_t)o_xf;
      h->off_w = (int64_t)o_w;
      h->off_starts = (int64_t)o_st;
      h->off_dmap = (int64_t)o_dm;
      h->off_vl = (int64_t)o_vl;
   }
   h->bytes = (int32_t)total;
   h->nxf = nxf;
   h->nstd = nstd;
   h->final_index = (cp->final_xform_enable == 1 && cp->final_xform_index >= 0) ? cp->final_xform_index : -1;
   h->chaos_enable = cp->chaos_enable;
   h->nunion = nunion;
   h->nrows = nrows;
   h->final_opacity = h->final_index >= 0 ? cp->xform[h->final_index].opacity : 1.0;
   u = 0;
   for (j = 0; j < FL3_NVAR; j++)
      if (used[j]) h->union_ids[u++] = j;
   for (; u < FL3_NVAR; u++) h->union_ids[u] = -1;

   {
      fl3_xform_pack *px = (fl3_xform_pack *)(blob + h->off_xf);
      double *w = (double *)(blob + h->off_w);
      int32_t *st = (int32_t *)(blob + h->off_starts);
      int slot[FL3_NVAR];
      u = 0;
      for (j = 0; j < FL3_NVAR; j++) slot[j] = used[j] ? u++ : -1;
      for (i = 0; i < nxf; i++) {
         const flam3_xform *x = &cp->xform[i];
         pack_xform(x, &px[i]);
         if (!usable(cp, i)) {
            px[i].flags = 0;
            px[i].preblur = 0;
            continue;
         }
         for (j = 0; j < x->num_active_vars; j++) {
            int v = x->varFunc[j];
            if (v >= 0 && v < FL3_NVAR && v != VAR_PRE_BLUR) w[i * nunion + slot[v]] = x->active_var_weights[j];
         }
      }
      if (starts)
         memcpy(st, starts, sizeof(int32_t) * (size_t)nrows * nstd);
      else
         fl3_distrib_breakpoints(distrib, nrows, nstd, st);
      /* per xform, its nonzero weights (as the walkers' floats) and their
       * variation ids, in union order: the list a wavefront that applies
       * one xform reads with scalar loads */
      for (i = 0; i < nxf; i++) {
         int32_t *vl = (int32_t *)(blob + h->off_vl) + (size_t)i * (2 + 2 * (size_t)nunion);
         int k = 0;
         for (u = 0; u < nunion; u++) {
            const float wf = (float)w[i * nunion + u];
            if (wf != 0.0f) {
               vl[2 + 2 * k] = h->union_ids[u];
               memcpy(&vl[3 + 2 * k], &wf, sizeof(wf));
               if (h->union_ids[u] == VAR_JULIA) px[i].flags |= FL3_DRAWS_JULIA;
               k++;
            }
         }
         vl[0] = k;
         vl[1] = 0;
      }
   }

   if (cam) {
      double *dm = (double *)(blob + h->off_dmap);
      h->cmap_size = cam->cmap_size;
      h->palette_mode = cam->palette_mode;
      h->rotate = cam->rotate;
      h->width = cam->width;
      h->height = cam->height;
      for (j = 0; j < 4; j++) h->bounds[j] = cam->bounds[j];
      h->rot[0] = cam->rot[0][0]; h->rot[1] = cam->rot[0][1];
      h->rot[2] = cam->rot[1][0]; h->rot[3] = cam->rot[1][1];
      h->rot_center[0] = cam->rot_center[0];
      h->rot_center[1] = cam->rot_center[1];
      h->ws0 = cam->ws0; h->wb0s0 = cam->wb0s0; h->hs1 = cam->hs1; h->hb1s1 = cam->hb1s1;
      memcpy(dm, cam->dmap, sizeof(double) * 1024);
   } else {
      h->cmap_size = 256;
   }
   free(own);
   *bytes = total;
   return blob;
}

void *fl3_pack_job(const flam3_genome *cp, const unsigned short *distrib, const fl3h_camera *cam,
                   size_t *bytes) {
   return pack_job(cp, distrib, NULL, cam, bytes);
}

void *fl3_pack_job_starts(const flam3_genome *cp, const int *starts, const fl3h_camera *cam, size_t *bytes) {
   return pack_job(cp, NULL, starts, cam, bytes);
}

/* A job with one identity xform and only the camera/palette set: what the
 * deposit stage needs. */
void *fl3_pack_camera_job(const fl3h_camera *cam, size_t *bytes) {
   flam3_genome g;
   flam3_xform x;
   unsigned short table[FL3_CHOOSE_GRAIN];
   void *r;
   memset(&g, 0, sizeof(g));
   fl3_xform_defaults(&x, 0);
   x.density = 1.0;
   g.num_xforms = 1;
   g.final_xform_index = -1;
   g.xform = &x;
   if (fl3_prepare_precalc_flags(&g)) return NULL;
   memset(table, 0, sizeof(table));
   r = fl3_pack_job(&g, table, cam, bytes);
   return r;
}
