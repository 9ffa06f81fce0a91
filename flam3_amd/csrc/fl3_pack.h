/*
 * fl3_pack.h - the packed genome ("job") the host hands to the kernels.
 *
 * One job is one interpolated genome at one temporal sample plus the camera
 * and palette needed to deposit its points.  It is a variable-size blob of
 * 8-byte aligned sections; a workgroup copies it into LDS (converted to the
 * iteration precision) before iterating.  Shared by the C host code and the
 * HIP device code, so plain C only.
 */
#ifndef FL3_PACK_H
#define FL3_PACK_H

#include <stdint.h>

#define FL3_NVAR 99
#define FL3_MAX_XFORMS 64   /* standard + final xforms per genome on device */

/* Variation parameter slots (per xform).  Names follow flam3_xform fields. */
enum fl3_param {
   P_BLOB_LOW, P_BLOB_HIGH, P_BLOB_WAVES,
   P_PDJ_A, P_PDJ_B, P_PDJ_C, P_PDJ_D,
   P_FAN2_X, P_FAN2_Y, P_RINGS2_VAL,
   P_PERSP_DIST, P_PERSP_VSIN, P_PERSP_VFCOS,
   P_JULIAN_POWER, P_JULIAN_RN, P_JULIAN_CN,
   P_JSCOPE_POWER, P_JSCOPE_RN, P_JSCOPE_CN,
   P_RB_SPINVAR, P_RB_ZOOMVAR,
   P_PIE_SLICES, P_PIE_ROTATION, P_PIE_THICKNESS,
   P_NGON_SIDES, P_NGON_POWER, P_NGON_CIRCLE, P_NGON_CORNERS,
   P_CURL_C1, P_CURL_C2, P_RECT_X, P_RECT_Y,
   P_DISC2_TIMESPI, P_DISC2_SINADD, P_DISC2_COSADD,
   P_SS_RND, P_SS_N2, P_SS_N3, P_SS_HOLES, P_SS_PM4, P_SS_PNEG1_N1,
   P_FLOWER_PETALS, P_FLOWER_HOLES, P_CONIC_ECC, P_CONIC_HOLES,
   P_PARA_H, P_PARA_W, P_BENT2_X, P_BENT2_Y, P_BIPOLAR_SHIFT, P_CELL_SIZE,
   P_CPOW_R, P_CPOW_I, P_CPOW_POWER,
   P_CURVE_XAMP, P_CURVE_YAMP, P_CURVE_XLEN, P_CURVE_YLEN,
   P_ESCHER_BETA,
   P_LS_SPIN, P_LS_SPACE, P_LS_TWIST, P_LS_X, P_LS_Y,
   P_MOD_X, P_MOD_Y,
   P_OSC_SEP, P_OSC_FREQ, P_OSC_AMP, P_OSC_DAMP,
   P_POP2_X, P_POP2_Y, P_POP2_C,
   P_SEP_X, P_SEP_XIN, P_SEP_Y, P_SEP_YIN,
   P_SPLIT_X, P_SPLIT_Y, P_SPLITS_X, P_SPLITS_Y,
   P_STRIPES_SPACE, P_STRIPES_WARP,
   P_WEDGE_ANGLE, P_WEDGE_HOLE, P_WEDGE_COUNT, P_WEDGE_SWIRL,
   P_WJ_ANGLE, P_WJ_COUNT, P_WJ_POWER, P_WJ_RN, P_WJ_CN, P_WJ_CF,
   P_WS_ANGLE, P_WS_COUNT, P_WS_HOLE, P_WS_SWIRL,
   P_WHORL_IN, P_WHORL_OUT,
   P_W2_FREQX, P_W2_SCALEX, P_W2_FREQY, P_W2_SCALEY,
   P_AUGER_SYM, P_AUGER_WEIGHT, P_AUGER_FREQ, P_AUGER_SCALE,
   P_FLUX_SPREAD,
   P_MOB_RE_A, P_MOB_IM_A, P_MOB_RE_B, P_MOB_IM_B,
   P_MOB_RE_C, P_MOB_IM_C, P_MOB_RE_D, P_MOB_IM_D,
   P_WAVES_DX2, P_WAVES_DY2,
   FL3_NPARAM
};

/* precalc flags of an xform */
#define FL3_PRE_ATAN_XY 1
#define FL3_PRE_ANGLES 2
#define FL3_PRE_ATAN_YX 4
#define FL3_HAS_POST 8
#define FL3_DRAWS_JULIA 16   /* julia (variation 13) is among the xform's nonzero float weights */

typedef struct {
   double c[6];        /* c00 c01 c10 c11 c20 c21 (x' = c00 x + c10 y + c20) */
   double post[6];
   double color, color_speed, vis, opacity, preblur;
   int32_t flags;
   int32_t pad_;
   double p[FL3_NPARAM];
} fl3_xform_pack;

/* Fixed header of a job; sections follow at the given byte offsets. */
typedef struct {
   int32_t bytes;          /* total blob size */
   int32_t nxf;            /* xforms in the xf section (std then final) */
   int32_t nstd;
   int32_t final_index;    /* -1: no final xform applied */
   int32_t chaos_enable;
   int32_t nunion;         /* variations used by any xform, ascending */
   int32_t nrows;          /* selection rows (1, or nstd+1 with xaos) */
   int32_t cmap_size;      /* 256 - apo_palette */
   int32_t palette_mode;
   int32_t rotate;         /* camera rotation active */
   int32_t width, height;  /* histogram cells (with gutter) */
   int32_t union_ids[FL3_NVAR];
   int32_t pad0_;
   double bounds[4];
   double rot[4];          /* rot[0][0] rot[0][1] rot[1][0] rot[1][1] */
   double rot_center[2];
   double ws0, wb0s0, hs1, hb1s1;
   double final_opacity;
   int64_t off_xf;         /* fl3_xform_pack[nxf] */
   int64_t off_w;          /* double[nxf * nunion] weights by union slot */
   int64_t off_starts;     /* int32[nrows * nstd] selection breakpoints */
   int64_t off_dmap;       /* double[256 * 4] palette x 255 x temporal weight */
   int64_t off_vl;         /* int32[nxf][2 + 2 nunion] + 8 (padding): per xform the
                              number of its weights that are nonzero as floats, 0,
                              then (variation id, the weight's float bits) for each,
                              in union order */
} fl3_job_hdr;

#endif
