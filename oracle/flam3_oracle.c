/*
 * flam3_oracle.c - TEST INFRASTRUCTURE: CPU restatement of the flam3 render
 * hot path, the parity oracle of the GPU library (see flam3_oracle.h).
 * Built with -O2, no -ffast-math, no FP contraction, the same flags as the
 * strict reference build it is pinned against.  Each function cites the
 * reference lines it restates.
 */
#define _GNU_SOURCE
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "flam3_oracle.h"

/* ---- reference functions used for tables and interpolation (libflam3ref_strict.so) ---- */
typedef struct {
   int max_filtered_counts;
   int max_filter_index;
   int kernel_size;
   double *filter_widths;
   double *filter_coefs;
} ref_de_helper;                                   /* filters.h:26-32 */
extern void flam3_interpolate(flam3_genome *cps, int ncps, double t, double stagger, flam3_genome *result);
extern int prepare_precalc_flags(flam3_genome *cp);
extern void xform_precalc(flam3_genome *cp, int xi);
extern unsigned short *flam3_create_xform_distrib(flam3_genome *cp);
extern ref_de_helper flam3_create_de_filters(double max_rad, double min_rad, double curve, int ss);
extern int flam3_create_spatial_filter(flam3_frame *spec, int field, double **filter);
extern double flam3_create_temporal_filter(int numsteps, int filter_type, double filter_exp, double filter_width,
                                           double **temporal_filter, double **temporal_deltas);
extern void clear_cp(flam3_genome *cp, int default_flag);
extern double flam3_calc_alpha(double density, double gamma, double linrange);
extern void flam3_calc_newrgb(double *cbuf, double ls, double highpow, double *newrgb);
extern void sincos(double x, double *s, double *c);

#define EPS (1e-10)
#define badvalue(x) (((x) != (x)) || ((x) > 1e10) || ((x) < -1e10))

/* ================================================================ ISAAC */

#define M32 0xffffffffUL

/* isaac.c:25-45 */
void or_isaac(randctx *ctx) {
   ub4 a = ctx->randa, b = (ctx->randb + (++ctx->randc)) & M32, x, y;
   ub4 *mm = ctx->randmem, *r = ctx->randrsl;
   int i;
   for (i = 0; i < 16; i++) {
      ub4 sh;
      x = mm[i];
      switch (i % 4) {
         case 0: sh = a << 13; break;
         case 1: sh = a >> 6; break;
         case 2: sh = a << 2; break;
         default: sh = a >> 16; break;
      }
      a = ((a ^ sh) + mm[(i + 8) % 16]) & M32;
      mm[i] = y = (mm[(x >> 2) & 15] + a + b) & M32;
      r[i] = b = (mm[((y >> 4) >> 2) & 15] + x) & M32;
   }
   ctx->randb = b;
   ctx->randa = a;
}

#define OR_MIX(a, b, c, d, e, f, g, h) \
   {                                   \
      a ^= b << 11; d += a; b += c;    \
      b ^= c >> 2; e += b; c += d;     \
      c ^= d << 8; f += c; d += e;     \
      d ^= e >> 16; g += d; e += f;    \
      e ^= f << 10; h += e; f += g;    \
      f ^= g >> 4; a += f; g += h;     \
      g ^= h << 8; b += g; h += a;     \
      h ^= a >> 9; c += h; a += b;     \
   }

/* isaac.c:61-110 */
void or_irandinit(randctx *ctx, int flag) {
   ub4 a, b, c, d, e, f, g, h, *m = ctx->randmem, *r = ctx->randrsl;
   int i, pass;
   ctx->randa = ctx->randb = ctx->randc = 0;
   a = b = c = d = e = f = g = h = 0x9e3779b9;
   for (i = 0; i < 4; ++i) OR_MIX(a, b, c, d, e, f, g, h);
   for (pass = 0; pass < (flag ? 2 : 1); pass++) {
      const ub4 *src = pass ? m : r;
      for (i = 0; i < 16; i += 8) {
         if (flag) {
            a += src[i]; b += src[i + 1]; c += src[i + 2]; d += src[i + 3];
            e += src[i + 4]; f += src[i + 5]; g += src[i + 6]; h += src[i + 7];
         }
         OR_MIX(a, b, c, d, e, f, g, h);
         m[i] = a; m[i + 1] = b; m[i + 2] = c; m[i + 3] = d;
         m[i + 4] = e; m[i + 5] = f; m[i + 6] = g; m[i + 7] = h;
      }
   }
   or_isaac(ctx);
   ctx->randcnt = 16;
}

/* isaac.h:49-52 */
unsigned long or_irand(randctx *r) {
   if (!r->randcnt--) {
      or_isaac(r);
      r->randcnt = 15;
   }
   return r->randrsl[r->randcnt];
}

/* flam3.c:2519-2525 */
double or_rand01(randctx *ct) { return ((int)or_irand(ct) & 0xfffffff) / (double)0xfffffff; }
double or_rand11(randctx *ct) { return (((int)or_irand(ct) & 0xfffffff) - 0x7ffffff) / (double)0x7ffffff; }
static int or_randbit(randctx *ct) { return (int)or_irand(ct) & 1; }

/* ================================================================ variations */

typedef struct {
   double tx, ty, sumsq, sq, atan, sina, cosa, atanyx, p0, p1;
   flam3_xform *x;
   randctx *rc;
} vstate;

#define R01 or_rand01(f->rc)

/* variations.c:142-1937, one case per variation */
static void or_var(int v, vstate *f, double w) {
   flam3_xform *X = f->x;
   switch (v) {
   case 0: f->p0 += w * f->tx; f->p1 += w * f->ty; break;                       /* :142 */
   case 1: f->p0 += w * sin(f->tx); f->p1 += w * sin(f->ty); break;             /* :153 */
   case 2: { double r2 = w / (f->sumsq + EPS); f->p0 += r2 * f->tx; f->p1 += r2 * f->ty; } break;
   case 3: { double c1, c2, nx, ny; sincos(f->sumsq, &c1, &c2);                  /* :178 */
      nx = c1 * f->tx - c2 * f->ty; ny = c2 * f->tx + c1 * f->ty;
      f->p0 += w * nx; f->p1 += w * ny; } break;
   case 4: { double r = w / (f->sq + EPS);                                      /* :202 */
      f->p0 += (f->tx - f->ty) * (f->tx + f->ty) * r; f->p1 += 2.0 * f->tx * f->ty * r; } break;
   case 5: { double nx = f->atan * M_1_PI, ny = f->sq - 1.0;                    /* :218 */
      f->p0 += w * nx; f->p1 += w * ny; } break;
   case 6: { double a = f->atan, r = f->sq;                                     /* :232 */
      f->p0 += w * r * sin(a + r); f->p1 += w * r * cos(a - r); } break;
   case 7: { double a = f->sq * f->atan, ca, sa, r = w * f->sq;                  /* :246 */
      sincos(a, &sa, &ca); f->p0 += r * sa; f->p1 += (-r) * ca; } break;
   case 8: { double a = f->atan * M_1_PI, r = M_PI * f->sq, sr, cr;              /* :264 */
      sincos(r, &sr, &cr); f->p0 += w * sr * a; f->p1 += w * cr * a; } break;
   case 9: { double r = f->sq + EPS, r1 = w / r, sr, cr;                        /* :282 */
      sincos(r, &sr, &cr); f->p0 += r1 * (f->cosa + sr); f->p1 += r1 * (f->sina - cr); } break;
   case 10: { double r = f->sq + EPS;                                           /* :298 */
      f->p0 += w * f->sina / r; f->p1 += w * f->cosa * r; } break;
   case 11: { double r = f->sq, sr, cr; sincos(r, &sr, &cr);                     /* :311 */
      f->p0 += w * f->sina * cr; f->p1 += w * f->cosa * sr; } break;
   case 12: { double a = f->atan, r = f->sq, n0 = sin(a + r), n1 = cos(a - r);   /* :326 */
      double m0 = n0 * n0 * n0 * r, m1 = n1 * n1 * n1 * r;
      f->p0 += w * (m0 + m1); f->p1 += w * (m0 - m1); } break;
   case 13: { double r, a = 0.5 * f->atan, sa, ca;                              /* :350 */
      if (or_randbit(f->rc)) a += M_PI;
      r = w * sqrt(f->sq); sincos(a, &sa, &ca);
      f->p0 += r * ca; f->p1 += r * sa; } break;
   case 14: { double nx = f->tx, ny = f->ty;                                    /* :375 */
      if (nx < 0.0) nx = nx * 2.0;
      if (ny < 0.0) ny = ny / 2.0;
      f->p0 += w * nx; f->p1 += w * ny; } break;
   case 15: { double c10 = X->c[1][0], c11 = X->c[1][1];                        /* :396 */
      double nx = f->tx + c10 * sin(f->ty * X->waves_dx2), ny = f->ty + c11 * sin(f->tx * X->waves_dy2);
      f->p0 += w * nx; f->p1 += w * ny; } break;
   case 16: { double r = f->sq; r = 2 * w / (r + 1);                            /* :415 */
      f->p0 += r * f->ty; f->p1 += r * f->tx; } break;
   case 17: { double dx = tan(3 * f->ty), dy = tan(3 * f->tx);                   /* :433 */
      double nx = f->tx + X->c[2][0] * sin(dx), ny = f->ty + X->c[2][1] * sin(dy);
      f->p0 += w * nx; f->p1 += w * ny; } break;
   case 18: { double dx = w * exp(f->tx - 1.0), dy = M_PI * f->ty, sdy, cdy;      /* :452 */
      sincos(dy, &sdy, &cdy); f->p0 += dx * cdy; f->p1 += dx * sdy; } break;
   case 19: { double r = w * pow(f->sq, f->sina);                               /* :472 */
      f->p0 += r * f->cosa; f->p1 += r * f->sina; } break;
   case 20: { double a = f->tx * M_PI, sa, ca, nx, ny; sincos(a, &sa, &ca);      /* :489 */
      nx = ca * cosh(f->ty); ny = -sa * sinh(f->ty);
      f->p0 += w * nx; f->p1 += w * ny; } break;
   case 21: { double dx = X->c[2][0] * X->c[2][0] + EPS, r = f->sq;             /* :508 */
      r = w * (fmod(r + dx, 2 * dx) - dx + r * (1 - dx));
      f->p0 += r * f->cosa; f->p1 += r * f->sina; } break;
   case 22: { double dx = M_PI * (X->c[2][0] * X->c[2][0] + EPS), dy = X->c[2][1], dx2 = 0.5 * dx;  /* :528 */
      double a = f->atan, r = w * f->sq, sa, ca;
      a += (fmod(a + dy, dx) > dx2) ? -dx2 : dx2;
      sincos(a, &sa, &ca); f->p0 += r * ca; f->p1 += r * sa; } break;
   case 23: { double r = f->sq, a = f->atan, bd = X->blob_high - X->blob_low;   /* :557 */
      r = r * (X->blob_low + bd * (0.5 + 0.5 * sin(X->blob_waves * a)));
      f->p0 += w * f->sina * r; f->p1 += w * f->cosa * r; } break;
   case 24: { double nx1 = cos(X->pdj_b * f->tx), nx2 = sin(X->pdj_c * f->tx);  /* :579 */
      double ny1 = sin(X->pdj_a * f->ty), ny2 = cos(X->pdj_d * f->ty);
      f->p0 += w * (ny1 - nx1); f->p1 += w * (nx2 - ny2); } break;
   case 25: { double dy = X->fan2_y, dx = M_PI * (X->fan2_x * X->fan2_x + EPS), dx2 = 0.5 * dx;  /* :598 */
      double a = f->atan, sa, ca, r = w * f->sq, t = a + dy - dx * (int)((a + dy) / dx);
      if (t > dx2) a = a - dx2; else a = a + dx2;
      sincos(a, &sa, &ca); f->p0 += r * sa; f->p1 += r * ca; } break;
   case 26: { double r = f->sq, dx = X->rings2_val * X->rings2_val + EPS;      /* :640 */
      r += -2.0 * dx * (int)((r + dx) / (2.0 * dx)) + r * (1.0 - dx);
      f->p0 += w * f->sina * r; f->p1 += w * f->cosa * r; } break;
   case 27: { double r = (w * 2.0) / (f->sq + 1.0); f->p0 += r * f->tx; f->p1 += r * f->ty; } break;
   case 28: { double r = w / (0.25 * (f->sumsq) + 1); f->p0 += r * f->tx; f->p1 += r * f->ty; } break;
   case 29: f->p0 += w * sin(f->tx); f->p1 += w * f->ty; break;                  /* :680 */
   case 30: { double t = 1.0 / (X->perspective_dist - f->ty * X->persp_vsin);   /* :687 */
      f->p0 += w * X->perspective_dist * f->tx * t; f->p1 += w * X->persp_vfcos * f->ty * t; } break;
   case 31: { double tmpr, sinr, cosr, r; tmpr = R01 * 2 * M_PI;               /* :696 */
      sincos(tmpr, &sinr, &cosr); r = w * R01;
      f->p0 += f->tx * r * cosr; f->p1 += f->ty * r * sinr; } break;
   case 32: { int t_rnd = (int)trunc((X->julian_rN) * R01);                    /* :710 */
      double tmpr = (f->atanyx + 2 * M_PI * t_rnd) / X->julian_power;
      double r = w * pow(f->sumsq, X->julian_cn), sina, cosa;
      sincos(tmpr, &sina, &cosa); f->p0 += r * cosa; f->p1 += r * sina; } break;
   case 33: { int t_rnd = (int)trunc((X->juliascope_rN) * R01); double tmpr, r, sina, cosa;  /* :725 */
      if ((t_rnd & 1) == 0) tmpr = (2 * M_PI * t_rnd + f->atanyx) / X->juliascope_power;
      else tmpr = (2 * M_PI * t_rnd - f->atanyx) / X->juliascope_power;
      sincos(tmpr, &sina, &cosa); r = w * pow(f->sumsq, X->juliascope_cn);
      f->p0 += r * cosa; f->p1 += r * sina; } break;
   case 34: { double tmpr, sinr, cosr, r; tmpr = R01 * 2 * M_PI;               /* :746 */
      sincos(tmpr, &sinr, &cosr); r = w * R01;
      f->p0 += r * cosr; f->p1 += r * sinr; } break;
   case 35: { double ang, r, sina, cosa; ang = R01 * 2 * M_PI;                  /* :760 */
      sincos(ang, &sina, &cosa);
      r = w * (R01 + R01 + R01 + R01 - 2.0);
      f->p0 += r * cosa; f->p1 += r * sina; } break;
   case 36: { double rndG, ra, rz, tmpa, sa, ca;                                /* :775 */
      rndG = w * (R01 + R01 + R01 + R01 - 2.0);
      ra = f->sq; tmpa = f->atanyx + X->radialBlur_spinvar * rndG;
      sincos(tmpa, &sa, &ca); rz = X->radialBlur_zoomvar * rndG - 1;
      f->p0 += ra * ca + rz * f->tx; f->p1 += ra * sa + rz * f->ty; } break;
   case 37: { double a, r, sa, ca; int sl;                                      /* :795 */
      sl = (int)(R01 * X->pie_slices + 0.5);
      a = X->pie_rotation + 2.0 * M_PI * (sl + R01 * X->pie_thickness) / X->pie_slices;
      r = w * R01; sincos(a, &sa, &ca);
      f->p0 += r * ca; f->p1 += r * sa; } break;
   case 38: { double r_factor, theta, phi, b, amp;                              /* :811 */
      r_factor = pow(f->sumsq, X->ngon_power / 2.0);
      theta = f->atanyx; b = 2 * M_PI / X->ngon_sides;
      phi = theta - (b * floor(theta / b));
      if (phi > b / 2) phi -= b;
      amp = X->ngon_corners * (1.0 / (cos(phi) + EPS) - 1.0) + X->ngon_circle;
      amp /= (r_factor + EPS);
      f->p0 += w * f->tx * amp; f->p1 += w * f->ty * amp; } break;
   case 39: { double re = 1.0 + X->curl_c1 * f->tx + X->curl_c2 * (f->tx * f->tx - f->ty * f->ty);  /* :832 */
      double im = X->curl_c1 * f->ty + 2.0 * X->curl_c2 * f->tx * f->ty, r = w / (re * re + im * im);
      f->p0 += (f->tx * re + f->ty * im) * r; f->p1 += (f->ty * re - f->tx * im) * r; } break;
   case 40:                                                                     /* :843 */
      if (X->rectangles_x == 0) f->p0 += w * f->tx;
      else f->p0 += w * ((2 * floor(f->tx / X->rectangles_x) + 1) * X->rectangles_x - f->tx);
      if (X->rectangles_y == 0) f->p1 += w * f->ty;
      else f->p1 += w * ((2 * floor(f->ty / X->rectangles_y) + 1) * X->rectangles_y - f->ty);
      break;
   case 41: { double ang = R01 * w * M_PI, sinr, cosr; sincos(ang, &sinr, &cosr);  /* :857 */
      f->p0 += w * sinr; f->p1 += w * (sinr * sinr) / cosr; } break;
   case 42: f->p0 += w * sin(f->tx) / cos(f->ty); f->p1 += w * tan(f->ty); break;  /* :885 */
   case 43: f->p0 += w * (R01 - 0.5); f->p1 += w * (R01 - 0.5); break;            /* :900 */
   case 44: { double ang = w * R01 * M_PI, r = w / (f->sumsq + EPS), tanr = w * tan(ang) * r;  /* :915 */
      f->p0 += tanr * cos(f->tx); f->p1 += tanr * sin(f->ty); } break;
   case 45: { double r = R01 * w * f->sq, sinr, cosr; sincos(r, &sinr, &cosr);     /* :946 */
      f->p0 += w * f->tx * (cosr + sinr); f->p1 += w * f->tx * (cosr - sinr); } break;
   case 46: { double r = w * f->sq, cr = cos(r), icr = 1.0 / cr;                /* :976 */
      f->p0 += w * f->tx;
      if (cr < 0) f->p1 += w * (icr + 1); else f->p1 += w * (icr - 1); } break;
   case 47: { double r = R01 * w * f->sq, sinr, cosr, diff;                      /* :998 */
      sincos(r, &sinr, &cosr); diff = log10(sinr * sinr) + cosr;
      if (badvalue(diff)) diff = -30.0;
      f->p0 += w * f->tx * diff; f->p1 += w * f->tx * (diff - sinr * M_PI); } break;
   case 48: { double s = f->tx * f->tx - f->ty * f->ty, r = w * sqrt(1.0 / (s * s + EPS));  /* :1033 */
      f->p0 += f->tx * r; f->p1 += f->ty * r; } break;
   case 49: { double r, t, sinr, cosr; t = X->disc2_timespi * (f->tx + f->ty);   /* :1054 */
      sincos(t, &sinr, &cosr); r = w * f->atan / M_PI;
      f->p0 += (sinr + X->disc2_cosadd) * r; f->p1 += (cosr + X->disc2_sinadd) * r; } break;
   case 50: { double theta, t1, t2, r, st, ct, myrnd;                            /* :1092 */
      theta = X->super_shape_pm_4 * f->atanyx + M_PI_4; sincos(theta, &st, &ct);
      t1 = fabs(ct); t1 = pow(t1, X->super_shape_n2);
      t2 = fabs(st); t2 = pow(t2, X->super_shape_n3);
      myrnd = X->super_shape_rnd;
      r = w * ((myrnd * R01 + (1.0 - myrnd) * f->sq) - X->super_shape_holes) *
          pow(t1 + t2, X->super_shape_pneg1_n1) / f->sq;
      f->p0 += r * f->tx; f->p1 += r * f->ty; } break;
   case 51: { double theta = f->atanyx;                                         /* :1118 */
      double r = w * (R01 - X->flower_holes) * cos(X->flower_petals * theta) / f->sq;
      f->p0 += r * f->tx; f->p1 += r * f->ty; } break;
   case 52: { double ct = f->tx / f->sq;                                        /* :1133 */
      double r = w * (R01 - X->conic_holes) * X->conic_eccentricity / (1 + X->conic_eccentricity * ct) / f->sq;
      f->p0 += r * f->tx; f->p1 += r * f->ty; } break;
   case 53: { double r = f->sq, sr, cr; sincos(r, &sr, &cr);                     /* :1148 */
      f->p0 += X->parabola_height * w * sr * sr * R01;
      f->p1 += X->parabola_width * w * cr * R01; } break;
   case 54: { double nx = f->tx, ny = f->ty;                                    /* :1164 */
      if (nx < 0.0) nx = nx * X->bent2_x;
      if (ny < 0.0) ny = ny * X->bent2_y;
      f->p0 += w * nx; f->p1 += w * ny; } break;
   case 55: { double x2y2 = f->sumsq, t = x2y2 + 1, x2 = 2 * f->tx;             /* :1180 */
      double ps = -M_PI_2 * X->bipolar_shift, y = 0.5 * atan2(2.0 * f->ty, x2y2 - 1.0) + ps;
      if (y > M_PI_2) y = -M_PI_2 + fmod(y + M_PI_2, M_PI);
      else if (y < -M_PI_2) y = M_PI_2 - fmod(M_PI_2 - y, M_PI);
      f->p0 += w * 0.25 * M_2_PI * log((t + x2) / (t - x2)); f->p1 += w * M_2_PI * y; } break;
   case 56: { double roundX = rint(f->tx), roundY = rint(f->ty);               /* :1199 */
      double offsetX = f->tx - roundX, offsetY = f->ty - roundY;
      if (R01 >= 0.75) {
         f->p0 += w * (offsetX * 0.5 + roundX); f->p1 += w * (offsetY * 0.5 + roundY);
      } else if (fabs(offsetX) >= fabs(offsetY)) {
         if (offsetX >= 0.0) {
            f->p0 += w * (offsetX * 0.5 + roundX + 0.25);
            f->p1 += w * (offsetY * 0.5 + roundY + 0.25 * offsetY / offsetX);
         } else {
            f->p0 += w * (offsetX * 0.5 + roundX - 0.25);
            f->p1 += w * (offsetY * 0.5 + roundY - 0.25 * offsetY / offsetX);
         }
      } else {
         if (offsetY >= 0.0) {
            f->p1 += w * (offsetY * 0.5 + roundY + 0.25);
            f->p0 += w * (offsetX * 0.5 + roundX + offsetX / offsetY * 0.25);
         } else {
            f->p1 += w * (offsetY * 0.5 + roundY - 0.25);
            f->p0 += w * (offsetX * 0.5 + roundX - offsetX / offsetY * 0.25);
         }
      } } break;
   case 57: { double wx = w * 1.3029400317411197908970256609023, y2 = f->ty * 2.0;  /* :1238 */
      double r = wx * sqrt(fabs(f->ty * f->tx) / (EPS + f->tx * f->tx + y2 * y2));
      f->p0 += r * f->tx; f->p1 += r * y2; } break;
   case 58: { double inv = 1.0 / X->cell_size;                                  /* :1253 */
      int x = floor(f->tx * inv), y = floor(f->ty * inv);
      double dx = f->tx - x * X->cell_size, dy = f->ty - y * X->cell_size;
      if (y >= 0) { if (x >= 0) { y *= 2; x *= 2; } else { y *= 2; x = -(2 * x + 1); } }
      else { if (x >= 0) { y = -(2 * y + 1); x *= 2; } else { y = -(2 * y + 1); x = -(2 * x + 1); } }
      f->p0 += w * (dx + x * X->cell_size); f->p1 -= w * (dy + y * X->cell_size); } break;
   case 59: { double a = f->atanyx, lnr = 0.5 * log(f->sumsq);                  /* :1291 */
      double va = 2.0 * M_PI / X->cpow_power, vc = X->cpow_r / X->cpow_power, vd = X->cpow_i / X->cpow_power;
      double ang = vc * a + vd * lnr + va * floor(X->cpow_power * R01), sa, ca;
      double m = w * exp(vc * lnr - vd * a);
      sincos(ang, &sa, &ca); f->p0 += m * ca; f->p1 += m * sa; } break;
   case 60: { double xl = X->curve_xlength * X->curve_xlength, yl = X->curve_ylength * X->curve_ylength;  /* :1312 */
      if (xl < 1E-20) xl = 1E-20;
      if (yl < 1E-20) yl = 1E-20;
      f->p0 += w * (f->tx + X->curve_xamp * exp(-f->ty * f->ty / xl));
      f->p1 += w * (f->ty + X->curve_yamp * exp(-f->tx * f->tx / yl)); } break;
   case 61: { double tmp = f->sumsq + 1.0, tmp2 = 2.0 * f->tx;                  /* :1328 */
      double r1 = sqrt(tmp + tmp2), r2 = sqrt(tmp - tmp2), xmax = (r1 + r2) * 0.5;
      double a1 = log(xmax + sqrt(xmax - 1.0)), a2 = -acos(f->tx / xmax), ww = w / 11.57034632;
      double snv, csv, snhu, cshu;
      sincos(a1, &snv, &csv); snhu = sinh(a2); cshu = cosh(a2);
      if (f->ty > 0.0) snv = -snv;
      f->p0 += ww * cshu * csv; f->p1 += ww * snhu * snv; } break;
   case 62: { double tmp = f->sumsq + 1.0, x2 = 2.0 * f->tx;                    /* :1354 */
      double xmax = 0.5 * (sqrt(tmp + x2) + sqrt(tmp - x2)), a = f->tx / xmax, b = 1.0 - a * a;
      double ssx = xmax - 1.0, ww = w / M_PI_2;
      if (b < 0) b = 0; else b = sqrt(b);
      if (ssx < 0) ssx = 0; else ssx = sqrt(ssx);
      f->p0 += ww * atan2(a, b);
      if (f->ty > 0) f->p1 += ww * log(xmax + ssx); else f->p1 -= ww * log(xmax + ssx); } break;
   case 63: { double seb, ceb, vc, vd, m, n, sn, cn, a = f->atanyx, lnr = 0.5 * log(f->sumsq);  /* :1385 */
      sincos(X->escher_beta, &seb, &ceb);
      vc = 0.5 * (1.0 + ceb); vd = 0.5 * seb;
      m = w * exp(vc * lnr - vd * a); n = vc * a + vd * lnr;
      sincos(n, &sn, &cn); f->p0 += m * cn; f->p1 += m * sn; } break;
   case 64: { double expx = exp(f->tx) * 0.5, expnx = 0.25 / expx, sn, cn, tmp;  /* :1412 */
      sincos(f->ty, &sn, &cn); tmp = w / (expx + expnx - cn);
      f->p0 += tmp * (expx - expnx); f->p1 += tmp * sn; } break;
   case 65: { double x = f->tx - X->lazysusan_x, y = f->ty + X->lazysusan_y, r = sqrt(x * x + y * y);  /* :1428 */
      double sina, cosa;
      if (r < w) {
         double a = atan2(y, x) + X->lazysusan_spin + X->lazysusan_twist * (w - r);
         sincos(a, &sina, &cosa); r = w * r;
         f->p0 += r * cosa + X->lazysusan_x; f->p1 += r * sina - X->lazysusan_y;
      } else {
         r = w * (1.0 + X->lazysusan_space / r);
         f->p0 += r * x + X->lazysusan_x; f->p1 += r * y - X->lazysusan_y;
      } } break;
   case 66: { double r2 = f->sumsq, w2 = w * w;                                 /* :1456 */
      if (r2 < w2) { double r = w * sqrt(w2 / r2 - 1.0); f->p0 += r * f->tx; f->p1 += r * f->ty; }
      else { f->p0 += w * f->tx; f->p1 += w * f->ty; } } break;
   case 68: { double xr = 2 * X->modulus_x, yr = 2 * X->modulus_y;             /* :1498 */
      if (f->tx > X->modulus_x) f->p0 += w * (-X->modulus_x + fmod(f->tx + X->modulus_x, xr));
      else if (f->tx < -X->modulus_x) f->p0 += w * (X->modulus_x - fmod(X->modulus_x - f->tx, xr));
      else f->p0 += w * f->tx;
      if (f->ty > X->modulus_y) f->p1 += w * (-X->modulus_y + fmod(f->ty + X->modulus_y, yr));
      else if (f->ty < -X->modulus_y) f->p1 += w * (X->modulus_y - fmod(X->modulus_y - f->ty, yr));
      else f->p1 += w * f->ty; } break;
   case 69: { double tpf = 2 * M_PI * X->oscope_frequency, t;                   /* :1521 */
      if (X->oscope_damping == 0.0) t = X->oscope_amplitude * cos(tpf * f->tx) + X->oscope_separation;
      else t = X->oscope_amplitude * exp(-fabs(f->tx) * X->oscope_damping) * cos(tpf * f->tx) + X->oscope_separation;
      if (fabs(f->ty) <= t) { f->p0 += w * f->tx; f->p1 -= w * f->ty; }
      else { f->p0 += w * f->tx; f->p1 += w * f->ty; } } break;
   case 70: { double p2v = w / M_PI; f->p0 += p2v * f->atan; f->p1 += p2v / 2.0 * log(f->sumsq); } break;
   case 71:                                                                     /* :1554 */
      f->p0 += w * (f->tx + X->popcorn2_x * sin(tan(f->ty * X->popcorn2_c)));
      f->p1 += w * (f->ty + X->popcorn2_y * sin(tan(f->tx * X->popcorn2_c)));
      break;
   case 72: { double t = f->sumsq, r = 1.0 / (f->sq * (t + 1.0 / (w + EPS)));   /* :1563 */
      f->p0 += f->tx * r; f->p1 += f->ty * r; } break;
   case 73: { double sx2 = X->separation_x * X->separation_x, sy2 = X->separation_y * X->separation_y;  /* :1584 */
      if (f->tx > 0.0) f->p0 += w * (sqrt(f->tx * f->tx + sx2) - f->tx * X->separation_xinside);
      else f->p0 -= w * (sqrt(f->tx * f->tx + sx2) + f->tx * X->separation_xinside);
      if (f->ty > 0.0) f->p1 += w * (sqrt(f->ty * f->ty + sy2) - f->ty * X->separation_yinside);
      else f->p1 -= w * (sqrt(f->ty * f->ty + sy2) + f->ty * X->separation_yinside); } break;
   case 74:                                                                     /* :1603 */
      if (cos(f->tx * X->split_xsize * M_PI) >= 0) f->p1 += w * f->ty; else f->p1 -= w * f->ty;
      if (cos(f->ty * X->split_ysize * M_PI) >= 0) f->p0 += w * f->tx; else f->p0 -= w * f->tx;
      break;
   case 75:                                                                     /* :1619 */
      if (f->tx >= 0) f->p0 += w * (f->tx + X->splits_x); else f->p0 += w * (f->tx - X->splits_x);
      if (f->ty >= 0) f->p1 += w * (f->ty + X->splits_y); else f->p1 += w * (f->ty - X->splits_y);
      break;
   case 76: { double roundx = floor(f->tx + 0.5), offsetx = f->tx - roundx;     /* :1635 */
      f->p0 += w * (offsetx * (1.0 - X->stripes_space) + roundx);
      f->p1 += w * (f->ty + offsetx * offsetx * X->stripes_warp); } break;
   case 77: { double r = f->sq, a = f->atanyx + X->wedge_swirl * r;             /* :1649 */
      double c = floor((X->wedge_count * a + M_PI) * M_1_PI * 0.5);
      double comp_fac = 1 - X->wedge_angle * X->wedge_count * M_1_PI * 0.5, sa, ca;
      a = a * comp_fac + c * X->wedge_angle; sincos(a, &sa, &ca);
      r = w * (r + X->wedge_hole); f->p0 += r * ca; f->p1 += r * sa; } break;
   case 78: { double r = w * pow(f->sumsq, X->wedgeJulia_cn);                   /* :1671 */
      int t_rnd = (int)((X->wedgeJulia_rN) * R01);
      double a = (f->atanyx + 2 * M_PI * t_rnd) / X->wedge_julia_power;
      double c = floor((X->wedge_julia_count * a + M_PI) * M_1_PI * 0.5), sa, ca;
      a = a * X->wedgeJulia_cf + c * X->wedge_julia_angle; sincos(a, &sa, &ca);
      f->p0 += r * ca; f->p1 += r * sa; } break;
   case 79: { double r = 1.0 / (f->sq + EPS), a = f->atanyx + X->wedge_sph_swirl * r;  /* :1689 */
      double c = floor((X->wedge_sph_count * a + M_PI) * M_1_PI * 0.5);
      double comp_fac = 1 - X->wedge_sph_angle * X->wedge_sph_count * M_1_PI * 0.5, sa, ca;
      a = a * comp_fac + c * X->wedge_sph_angle; sincos(a, &sa, &ca);
      r = w * (r + X->wedge_sph_hole); f->p0 += r * ca; f->p1 += r * sa; } break;
   case 80: { double r = f->sq, a, sa, ca;                                      /* :1710 */
      if (r < w) a = f->atanyx + X->whorl_inside / (w - r);
      else a = f->atanyx + X->whorl_outside / (w - r);
      sincos(a, &sa, &ca); f->p0 += w * r * ca; f->p1 += w * r * sa; } break;
   case 81:                                                                     /* :1735 */
      f->p0 += w * (f->tx + X->waves2_scalex * sin(f->ty * X->waves2_freqx));
      f->p1 += w * (f->ty + X->waves2_scaley * sin(f->tx * X->waves2_freqy));
      break;
   case 82: { double e = exp(f->tx), c, s; sincos(f->ty, &s, &c);               /* :1747 */
      f->p0 += w * e * c; f->p1 += w * e * s; } break;
   case 83: f->p0 += w * 0.5 * log(f->sumsq); f->p1 += w * f->atanyx; break;    /* :1756 */
   case 84: { double s, c, sh, ch; sincos(f->tx, &s, &c); sh = sinh(f->ty); ch = cosh(f->ty);  /* :1763 */
      f->p0 += w * s * ch; f->p1 += w * c * sh; } break;
   case 85: { double s, c, sh, ch; sincos(f->tx, &s, &c); sh = sinh(f->ty); ch = cosh(f->ty);  /* :1773 */
      f->p0 += w * c * ch; f->p1 -= w * s * sh; } break;
   case 86: { double s, c, sh, ch, den; sincos(2 * f->tx, &s, &c);              /* :1783 */
      sh = sinh(2.0 * f->ty); ch = cosh(2.0 * f->ty); den = 1.0 / (c + ch);
      f->p0 += w * den * s; f->p1 += w * den * sh; } break;
   case 87: { double s, c, sh, ch, den; sincos(f->tx, &s, &c);                  /* :1795 */
      sh = sinh(f->ty); ch = cosh(f->ty); den = 2.0 / (cos(2 * f->tx) + cosh(2 * f->ty));
      f->p0 += w * den * c * ch; f->p1 += w * den * s * sh; } break;
   case 88: { double s, c, sh, ch, den; sincos(f->tx, &s, &c);                  /* :1807 */
      sh = sinh(f->ty); ch = cosh(f->ty); den = 2.0 / (cosh(2.0 * f->ty) - cos(2.0 * f->tx));
      f->p0 += w * den * s * ch; f->p1 -= w * den * c * sh; } break;
   case 89: { double s, c, sh, ch, den; sincos(2.0 * f->tx, &s, &c);            /* :1819 */
      sh = sinh(2.0 * f->ty); ch = cosh(2.0 * f->ty); den = 1.0 / (ch - c);
      f->p0 += w * den * s; f->p1 += w * den * -1 * sh; } break;
   case 90: { double s, c, sh, ch; sincos(f->ty, &s, &c); sh = sinh(f->tx); ch = cosh(f->tx);  /* :1831 */
      f->p0 += w * sh * c; f->p1 += w * ch * s; } break;
   case 91: { double s, c, sh, ch; sincos(f->ty, &s, &c); sh = sinh(f->tx); ch = cosh(f->tx);  /* :1841 */
      f->p0 += w * ch * c; f->p1 += w * sh * s; } break;
   case 92: { double s, c, sh, ch, den; sincos(2.0 * f->ty, &s, &c);            /* :1851 */
      sh = sinh(2.0 * f->tx); ch = cosh(2.0 * f->tx); den = 1.0 / (c + ch);
      f->p0 += w * den * sh; f->p1 += w * den * s; } break;
   case 93: { double s, c, sh, ch, den; sincos(f->ty, &s, &c);                  /* :1863 */
      sh = sinh(f->tx); ch = cosh(f->tx); den = 2.0 / (cos(2.0 * f->ty) + cosh(2.0 * f->tx));
      f->p0 += w * den * c * ch; f->p1 -= w * den * s * sh; } break;
   case 94: { double s, c, sh, ch, den; sincos(f->ty, &s, &c);                  /* :1875 */
      sh = sinh(f->tx); ch = cosh(f->tx); den = 2.0 / (cosh(2.0 * f->tx) - cos(2.0 * f->ty));
      f->p0 += w * den * sh * c; f->p1 -= w * den * ch * s; } break;
   case 95: { double s, c, sh, ch, den; sincos(2.0 * f->ty, &s, &c);            /* :1887 */
      sh = sinh(2.0 * f->tx); ch = cosh(2.0 * f->tx); den = 1.0 / (ch - c);
      f->p0 += w * den * sh; f->p1 += w * den * s; } break;
   case 96: { double s = sin(X->auger_freq * f->tx), t = sin(X->auger_freq * f->ty);  /* :1899 */
      double dy = f->ty + X->auger_weight * (X->auger_scale * s / 2.0 + fabs(f->ty) * s);
      double dx = f->tx + X->auger_weight * (X->auger_scale * t / 2.0 + fabs(f->tx) * t);
      f->p0 += w * (f->tx + X->auger_sym * (dx - f->tx)); f->p1 += w * dy; } break;
   case 97: { double xpw = f->tx + w, xmw = f->tx - w;                          /* :1911 */
      double avgr = w * (2 + X->flux_spread) * sqrt(sqrt(f->ty * f->ty + xpw * xpw) / sqrt(f->ty * f->ty + xmw * xmw));
      double avga = (atan2(f->ty, xmw) - atan2(f->ty, xpw)) * 0.5;
      f->p0 += avgr * cos(avga); f->p1 += avgr * sin(avga); } break;
   case 98: {                                                                   /* :1923 */
      double re_u = X->mobius_re_a * f->tx - X->mobius_im_a * f->ty + X->mobius_re_b;
      double im_u = X->mobius_re_a * f->ty + X->mobius_im_a * f->tx + X->mobius_im_b;
      double re_v = X->mobius_re_c * f->tx - X->mobius_im_c * f->ty + X->mobius_re_d;
      double im_v = X->mobius_re_c * f->ty + X->mobius_im_c * f->tx + X->mobius_im_d;
      double rad_v = w / (re_v * re_v + im_v * im_v);
      f->p0 += rad_v * (re_u * re_v + im_u * im_v); f->p1 += rad_v * (im_u * re_v - re_u * im_v); } break;
   default: break;
   }
}

/* variations.c:2129-2399 */
int or_apply_xform(flam3_genome *cp, int fn, const double *p, double *q, randctx *rc) {
   flam3_xform *X = &cp->xform[fn];
   vstate f;
   double s1 = X->color_speed, p2 = p[2];
   int n;
   f.rc = rc;
   f.x = X;
   q[2] = s1 * X->color + (1.0 - s1) * p2;
   q[3] = X->vis_adjusted;
   f.tx = X->c[0][0] * p[0] + X->c[1][0] * p[1] + X->c[2][0];
   f.ty = X->c[0][1] * p[0] + X->c[1][1] * p[1] + X->c[2][1];
   if (X->has_preblur != 0.0) {                                                /* :1480 */
      double rndG = X->has_preblur * (or_rand01(rc) + or_rand01(rc) + or_rand01(rc) + or_rand01(rc) - 2.0);
      double rndA = or_rand01(rc) * 2.0 * M_PI, sinA, cosA;
      sincos(rndA, &sinA, &cosA);
      f.tx += rndG * cosA;
      f.ty += rndG * sinA;
   }
   f.sumsq = f.tx * f.tx + f.ty * f.ty;
   f.sq = sqrt(f.sumsq);
   f.atan = f.sina = f.cosa = f.atanyx = 0.0;
   if (X->precalc_atan_xy_flag > 0) f.atan = atan2(f.tx, f.ty);
   if (X->precalc_angles_flag > 0) {
      f.sina = f.tx / f.sq;
      f.cosa = f.ty / f.sq;
   }
   if (X->precalc_atan_yx_flag > 0) f.atanyx = atan2(f.ty, f.tx);
   f.p0 = 0.0;
   f.p1 = 0.0;
   for (n = 0; n < X->num_active_vars; n++) or_var(X->varFunc[n], &f, X->active_var_weights[n]);
   if (X->has_post) {
      q[0] = X->post[0][0] * f.p0 + X->post[1][0] * f.p1 + X->post[2][0];
      q[1] = X->post[0][1] * f.p0 + X->post[1][1] * f.p1 + X->post[2][1];
   } else {
      q[0] = f.p0;
      q[1] = f.p1;
   }
   if (badvalue(q[0]) || badvalue(q[1])) {
      q[0] = or_rand11(rc);
      q[1] = or_rand11(rc);
      return 1;
   }
   return 0;
}

/* flam3.c:232-299 */
int or_iterate(flam3_genome *cp, int n, int fuse, double *samples, const unsigned short *distrib, randctx *rc) {
   double p[4], q[4];
   int i, consec = 0, badvals = 0, lastxf = 0, fn;
   const int fx = cp->final_xform_index;
   p[0] = samples[0]; p[1] = samples[1]; p[2] = samples[2]; p[3] = samples[3];
   for (i = 0; i < cp->num_xforms; i++) xform_precalc(cp, i);
   for (i = -4 * fuse; i < 4 * n; i += 4) {
      if (cp->chaos_enable)
         fn = distrib[lastxf * 16384 + (((unsigned)or_irand(rc)) & 16383)];
      else
         fn = distrib[((unsigned)or_irand(rc)) & 16383];
      if (or_apply_xform(cp, fn, p, q, rc) > 0) {
         consec++;
         badvals++;
         if (consec < 5) {
            memcpy(p, q, sizeof(p));
            i -= 4;
            continue;
         } else
            consec = 0;
      } else
         consec = 0;
      lastxf = fn + 1;
      memcpy(p, q, sizeof(p));
      if (cp->final_xform_enable == 1) {
         if (cp->xform[fx].opacity == 1 || or_rand01(rc) < cp->xform[fx].opacity) {
            or_apply_xform(cp, fx, p, q, rc);
            q[3] = p[3];
         }
      }
      if (i >= 0) memcpy(samples + i, q, sizeof(q));
   }
   return badvals;
}

/* ================================================================ deposit */

/* One saturating add into a uint32 channel (flam3.c:3838-3840 / 3690-3706). */
static void bump_u32(unsigned int *dest, double delta, int mt) {
   if (mt) {
      unsigned int d = (unsigned int)delta;
      if (UINT_MAX - *dest > d) *dest += d;
      else *dest = UINT_MAX;
   } else {
      if (UINT_MAX - *dest > delta) *dest += delta;
      else *dest = UINT_MAX;
   }
}

/* rect.c:411-516 */
void or_deposit(const fl3h_camera *cam, const double *pts, long n, int bits, int mt, void *buckets) {
   long j;
   const int cmap_size = cam->cmap_size, cmap_size_m1 = cam->cmap_size - 1;
   for (j = 0; j < n; j++) {
      const double *p = pts + 4 * j;
      double p0, p1, p00, p11, dbl_index0, dbl_frac, interpcolor[4], logvis;
      int ci, color_index0;
      long idx;
      if (cam->rotate) {
         p00 = p[0] - cam->rot_center[0];
         p11 = p[1] - cam->rot_center[1];
         p0 = p00 * cam->rot[0][0] + p11 * cam->rot[0][1] + cam->rot_center[0];
         p1 = p00 * cam->rot[1][0] + p11 * cam->rot[1][1] + cam->rot_center[1];
      } else {
         p0 = p[0];
         p1 = p[1];
      }
      if (!(p0 >= cam->bounds[0] && p1 >= cam->bounds[1] && p0 <= cam->bounds[2] && p1 <= cam->bounds[3]))
         continue;
      if (p[3] == 0) continue;
      logvis = p[3];
      idx = (int)(cam->ws0 * p0 - cam->wb0s0) + cam->width * (int)(cam->hs1 * p1 - cam->hb1s1);
      if (idx < 0 || idx >= (long)cam->width * cam->height) continue;  /* the reference would overrun */
      dbl_index0 = p[2] * cmap_size;
      color_index0 = (int)(dbl_index0);
      if (cam->palette_mode == flam3_palette_mode_linear) {
         if (color_index0 < 0) {
            color_index0 = 0;
            dbl_frac = 0;
         } else if (color_index0 >= cmap_size_m1) {
            color_index0 = cmap_size_m1 - 1;
            dbl_frac = 1.0;
         } else
            dbl_frac = dbl_index0 - (double)color_index0;
         for (ci = 0; ci < 4; ci++)
            interpcolor[ci] = cam->dmap[color_index0][ci] * (1.0 - dbl_frac) + cam->dmap[color_index0 + 1][ci] * dbl_frac;
      } else {
         if (color_index0 < 0) color_index0 = 0;
         else if (color_index0 >= cmap_size_m1) color_index0 = cmap_size_m1;
         for (ci = 0; ci < 4; ci++) interpcolor[ci] = cam->dmap[color_index0][ci];
      }
      if (bits == 64) {
         double *b = (double *)buckets + 5 * idx;
         if (p[3] == 1.0) {
            for (ci = 0; ci < 4; ci++) b[ci] += interpcolor[ci];
            b[4] += 255.0;
         } else {
            for (ci = 0; ci < 4; ci++) b[ci] += logvis * interpcolor[ci];
            b[4] += logvis * 255.0;
         }
      } else {
         unsigned int *b = (unsigned int *)buckets + 5 * idx;
         if (p[3] == 1.0) {
            for (ci = 0; ci < 4; ci++) bump_u32(b + ci, interpcolor[ci], mt);
            bump_u32(b + 4, 255.0, mt);
         } else {
            for (ci = 0; ci < 4; ci++) bump_u32(b + ci, logvis * interpcolor[ci], mt);
            bump_u32(b + 4, logvis * 255.0, mt);
         }
      }
   }
}

/* ================================================================ density estimation */

static double bucket_ch(const void *b, int bits, long cell, int ch) {
   if (bits == 64) return ((const double *)b)[5 * cell + ch];
   return (double)((const unsigned int *)b)[5 * cell + ch];
}

/* add_c_to_accum (flam3.c:3777-3785 bits 32, 3827-3835 bits 33, 3727-3735 bits 64) */
static void add_c(void *acc, int bits, int x, int y, int W, int H, const double *c) {
   long cell;
   int k;
   if (!(y >= 0 && y < H && x >= 0 && x < W)) return;
   cell = x + (long)y * W;
   for (k = 0; k < 4; k++) {
      if (bits == 64) ((double *)acc)[4 * cell + k] += c[k];
      else if (bits == 33) ((float *)acc)[4 * cell + k] += c[k];
      else {
         unsigned int *d = (unsigned int *)acc + 4 * cell + k;
         if (UINT_MAX - *d > c[k]) *d += c[k];
         else *d = UINT_MAX;
      }
   }
}

/* rect.c:57-175 for all rows (one thread).  Cells in the row past the end
 * come from `past` (the reference reads the accumulator there) or are 0. */
void or_density(const void *buckets, int width, int height, int bits, int oversample, double k1, double k2,
                double curve, int max_filtered_counts, int max_filter_index, int kernel_size,
                const double *filter_widths, const double *filter_coefs, void *accum, int phantom) {
   const int ss = (int)floor(oversample / 2.0);
   const int scf = !(oversample & 1);
   const double scfact = pow(oversample / (oversample + 1.0), 2.0);
   const int wid = width, hig = height;
   const int myspan = height - 2 * (oversample - 1) + 1;
   const int str = oversample - 1, enr = (oversample - 1) + myspan;
   const size_t bsz = bits == 64 ? 40 : 20;
   int i, j;
   /* the reference's view: buckets immediately followed by the accumulator */
   unsigned char *view = (unsigned char *)malloc(bsz * ((size_t)width * height + width));
   memcpy(view, buckets, bsz * (size_t)width * height);
   if (phantom)
      memcpy(view + bsz * (size_t)width * height, accum, bsz * (size_t)width);
   else
      memset(view + bsz * (size_t)width * height, 0, bsz * (size_t)width);
   for (j = str; j < enr; j++) {
      for (i = oversample - 1; i < wid - (oversample - 1); i++) {
         int ii, jj, f_select_int, f_coef_idx, arr_filt_width;
         double f_select = 0.0, c[4], ls;
         long cell = i + (long)j * wid;
         if (phantom && j == hig) /* refresh: the accumulator may have changed */
            memcpy(view + bsz * (size_t)width * height, accum, bsz * (size_t)width);
         if (bucket_ch(view, bits, cell, 4) == 0 || bucket_ch(view, bits, cell, 3) == 0) continue;
         for (ii = -ss; ii <= ss; ii++)
            for (jj = -ss; jj <= ss; jj++)
               f_select += bucket_ch(view, bits, (i + ii) + (long)(j + jj) * wid, 4) / 255.0;
         if (scf) f_select *= scfact;
         if (f_select > max_filtered_counts)
            f_select_int = max_filter_index;
         else if (f_select <= 100)
            f_select_int = (int)ceil(f_select) - 1;
         else
            f_select_int = 100 + (int)floor(pow(f_select - 100, curve));
         if (f_select_int > max_filter_index) f_select_int = max_filter_index;
         f_coef_idx = f_select_int * kernel_size;
         arr_filt_width = (int)ceil(filter_widths[f_select_int]) - 1;
         for (jj = 0; jj <= arr_filt_width; jj++) {
            for (ii = 0; ii <= jj; ii++) {
               if (filter_coefs[f_coef_idx] == 0.0) {
                  f_coef_idx++;
                  continue;
               }
               c[0] = bucket_ch(view, bits, cell, 0);
               c[1] = bucket_ch(view, bits, cell, 1);
               c[2] = bucket_ch(view, bits, cell, 2);
               c[3] = bucket_ch(view, bits, cell, 3);
               ls = filter_coefs[f_coef_idx] * (k1 * log(1.0 + c[3] * k2)) / c[3];
               c[0] *= ls; c[1] *= ls; c[2] *= ls; c[3] *= ls;
               if (jj == 0 && ii == 0) {
                  add_c(accum, bits, i + ii, j + jj, wid, hig, c);
               } else if (ii == 0) {
                  add_c(accum, bits, i + jj, j, wid, hig, c);
                  add_c(accum, bits, i - jj, j, wid, hig, c);
                  add_c(accum, bits, i, j + jj, wid, hig, c);
                  add_c(accum, bits, i, j - jj, wid, hig, c);
               } else if (jj == ii) {
                  add_c(accum, bits, i + ii, j + jj, wid, hig, c);
                  add_c(accum, bits, i - ii, j + jj, wid, hig, c);
                  add_c(accum, bits, i + ii, j - jj, wid, hig, c);
                  add_c(accum, bits, i - ii, j - jj, wid, hig, c);
               } else {
                  add_c(accum, bits, i + ii, j + jj, wid, hig, c);
                  add_c(accum, bits, i - ii, j + jj, wid, hig, c);
                  add_c(accum, bits, i + ii, j - jj, wid, hig, c);
                  add_c(accum, bits, i - ii, j - jj, wid, hig, c);
                  add_c(accum, bits, i + jj, j + ii, wid, hig, c);
                  add_c(accum, bits, i - jj, j + ii, wid, hig, c);
                  add_c(accum, bits, i + jj, j - ii, wid, hig, c);
                  add_c(accum, bits, i - jj, j - ii, wid, hig, c);
               }
               f_coef_idx++;
            }
         }
      }
   }
   free(view);
}

/* rect.c:947-974 */
void or_logscale(const void *buckets, int width, int height, int bits, double k1, double k2, void *accum) {
   long cell, n = (long)width * height;
   for (cell = 0; cell < n; cell++) {
      double c[4], ls;
      int k;
      for (k = 0; k < 4; k++) c[k] = bucket_ch(buckets, bits, cell, k);
      if (0.0 == c[3]) continue;
      ls = (k1 * log(1.0 + c[3] * k2)) / c[3];
      for (k = 0; k < 4; k++) c[k] *= ls;
      add_c(accum, bits, (int)(cell % width), (int)(cell / width), width, height, c);
   }
}

/* ================================================================ tone map */

static double acc_get(const void *acc, int bits, long cell, int k) {
   if (bits == 64) return ((const double *)acc)[4 * cell + k];
   if (bits == 33) return ((const float *)acc)[4 * cell + k];
   return ((const unsigned int *)acc)[4 * cell + k];
}

static void acc_set(void *acc, int bits, long cell, int k, double v) {
   if (bits == 64) ((double *)acc)[4 * cell + k] = v;
   else if (bits == 33) ((float *)acc)[4 * cell + k] = v;
   else ((unsigned int *)acc)[4 * cell + k] = v;
}

/* Colour of one pixel or cell from its filtered accumulator (rect.c:1093-1131,
 * 1167-1203). */
static void tone(double *t, const fl3h_tonemap_params *tp, double *alpha_out) {
   double alpha, ls, tmp = t[3] / 255, newrgb[3], a;
   int rgbi;
   if (t[3] <= 0) {
      alpha = 0.0;
      ls = 0.0;
   } else {
      alpha = flam3_calc_alpha(tmp, tp->g, tp->linrange);
      ls = tp->vibrancy * 256.0 * alpha / tmp;
      if (alpha < 0.0) alpha = 0.0;
      if (alpha > 1.0) alpha = 1.0;
   }
   flam3_calc_newrgb(t, ls, tp->highpow, newrgb);
   for (rgbi = 0; rgbi < 3; rgbi++) {
      a = newrgb[rgbi];
      a += (1.0 - tp->vibrancy) * 256.0 * pow(t[rgbi] / 255, tp->g);
      if (tp->nchan <= 3 || tp->transp == 0)
         a += ((1.0 - alpha) * tp->background[rgbi]);
      else {
         if (alpha > 0) a /= alpha;
         else a = 0;
      }
      if (a > 255) a = 255;
      if (a < 0) a = 0;
      t[rgbi] = a;
   }
   *alpha_out = alpha;
}

/* rect.c:1067-1247 */
void or_tonemap(void *accum, int width, int height, int bits, const fl3h_tonemap_params *tp,
                const double *filter, void *out, int out_stride_pixels) {
   int i, j, x, y, ii, jj, rgbi;
   const int fw = tp->filter_width, os = tp->oversample, nchan = tp->nchan, bpc = tp->bytes_per_channel;
   if (tp->earlyclip) {
      long cell, n = (long)width * height;
      for (cell = 0; cell < n; cell++) {
         double t[4], alpha;
         for (rgbi = 0; rgbi < 4; rgbi++) t[rgbi] = acc_get(accum, bits, cell, rgbi);
         tone(t, tp, &alpha);
         for (rgbi = 0; rgbi < 3; rgbi++) acc_set(accum, bits, cell, rgbi, t[rgbi]);
         acc_set(accum, bits, cell, 3, alpha);
      }
   }
   y = tp->de_offset;
   for (j = 0; j < tp->image_height; j++) {
      x = tp->de_offset;
      for (i = 0; i < tp->image_width; i++) {
         double t[4] = {0.0, 0.0, 0.0, 0.0}, a;
         unsigned char *p8 = (unsigned char *)out + (size_t)nchan * bpc * (i + (size_t)j * out_stride_pixels);
         unsigned short *p16 = (unsigned short *)p8;
         for (ii = 0; ii < fw; ii++)
            for (jj = 0; jj < fw; jj++) {
               double k = filter[ii + jj * fw];
               long cell = x + ii + (long)(y + jj) * width;
               t[0] += k * acc_get(accum, bits, cell, 0);
               t[1] += k * acc_get(accum, bits, cell, 1);
               t[2] += k * acc_get(accum, bits, cell, 2);
               t[3] += k * acc_get(accum, bits, cell, 3);
            }
         if (!tp->earlyclip) {
            double alpha;
            tone(t, tp, &alpha);
            t[3] = alpha;
         }
         for (rgbi = 0; rgbi < 3; rgbi++) {
            a = t[rgbi];
            if (a > 255) a = 255;
            if (a < 0) a = 0;
            if (2 == bpc) {
               a *= 256.0;
               p16[rgbi] = (unsigned short)a;
            } else
               p8[rgbi] = (unsigned char)a;
         }
         if (t[3] > 1) t[3] = 1;
         if (t[3] < 0) t[3] = 0;
         if (nchan > 3) {
            if (tp->transp == 1) {
               if (2 == bpc) p16[3] = (unsigned short)(t[3] * 65535);
               else p8[3] = (unsigned char)(t[3] * 255);
            } else {
               if (2 == bpc) p16[3] = 65535;
               else p8[3] = 255;
            }
         }
         x += os;
      }
      y += os;
   }
}

/* ================================================================ render */

/* render_rectangle (rect.c:529-1286) with one iteration thread. */
int or_render(flam3_frame *spec, void *out, int field, int nchan, int transp, stat_struct *stats, int phantom) {
   flam3_genome cp;
   int i, j, k, batch_num, tsn, image_width, image_height, out_width, filter_width, oversample;
   int nbatches, nts, gutter_width, max_gnm_de_fw, de_offset, width, height, vib_gam_n = 0, bits = spec->bits;
   double *filter, *temporal_filter, *temporal_deltas, sumfilt, vibrancy = 0.0, gamma = 0.0;
   double background[3] = {0.0, 0.0, 0.0}, ppux = 0, ppuy = 0, badvals = 0.0, highpow;
   size_t nbuckets, bsz, asz;
   void *buckets, *accumulate;
   double *points;
   fl3h_camera cam;
   randctx trc;
   const int SBS = spec->sub_batch_size;
   const int fuse = spec->earlyclip ? 100 : 15;
   const char *ai = getenv("apo_palette");
   const int cmap_size = 256 - (ai ? atoi(ai) : 0);

   stats->num_iters = 0;
   memset(&cp, 0, sizeof(cp));
   flam3_interpolate(spec->genomes, spec->ngenomes, spec->time, 0, &cp);
   oversample = cp.spatial_oversample;
   highpow = cp.highlight_power;
   nbatches = cp.nbatches;
   nts = cp.ntemporal_samples;
   if (nbatches < 1 || oversample < 1) return 1;
   image_width = cp.width;
   out_width = image_width;
   if (field) {
      image_height = cp.height / 2;
      if (field == flam3_field_odd) out = (unsigned char *)out + nchan * spec->bytes_per_channel * out_width;
      out_width *= 2;
   } else
      image_height = cp.height;
   filter_width = flam3_create_spatial_filter(spec, field, &filter);
   if (filter_width < 0) return 1;
   sumfilt = flam3_create_temporal_filter(nbatches * nts, cp.temporal_filter_type, cp.temporal_filter_exp,
                                          cp.temporal_filter_width, &temporal_filter, &temporal_deltas);
   gutter_width = (filter_width - oversample) / 2;
   max_gnm_de_fw = 0;
   for (i = 0; i < spec->ngenomes; i++) {
      int w = (int)ceil(spec->genomes[i].estimator) * oversample;
      if (w > max_gnm_de_fw) max_gnm_de_fw = w;
   }
   if (max_gnm_de_fw > 0) max_gnm_de_fw += (oversample - 1);
   if (max_gnm_de_fw > gutter_width) {
      de_offset = max_gnm_de_fw - gutter_width;
      gutter_width = max_gnm_de_fw;
   } else
      de_offset = 0;
   height = oversample * image_height + 2 * gutter_width;
   width = oversample * image_width + 2 * gutter_width;
   nbuckets = (size_t)width * height;
   bsz = bits == 64 ? 40 : 20;
   asz = bits == 64 ? 32 : 16;
   buckets = calloc(nbuckets, bsz);
   accumulate = calloc(nbuckets, asz);
   points = (double *)malloc(sizeof(double) * 4 * (size_t)SBS);
   memset(&cam, 0, sizeof(cam));

   for (batch_num = 0; batch_num < nbatches; batch_num++) {
      double sample_density = 0.0, k1, area, k2;
      ref_de_helper de;
      memset(buckets, 0, nbuckets * bsz);
      flam3_interpolate(spec->genomes, spec->ngenomes, spec->time + temporal_deltas[batch_num * nts], 0, &cp);
      if (cp.estimator < 0.0 || cp.estimator_minimum < 0.0) return 1;
      if (bits <= 32 && cp.estimator > 0.0) cp.estimator = 0.0;
      if (cp.estimator > 0.0) {
         de = flam3_create_de_filters(cp.estimator, cp.estimator_minimum, cp.estimator_curve, oversample);
         if (de.kernel_size < 0) return 1;
      } else
         de.max_filter_index = 0;
      for (tsn = 0; tsn < nts; tsn++) {
         const double color_scalar = temporal_filter[batch_num * nts + tsn];
         unsigned short *distrib;
         double batch_size, thread_bs, sub_batch, scale, t0, t1, corner0, corner1, shift = 0.0;
         flam3_interpolate(spec->genomes, spec->ngenomes, spec->time + temporal_deltas[batch_num * nts + tsn], 0, &cp);
         if (prepare_precalc_flags(&cp)) return 1;
         distrib = flam3_create_xform_distrib(&cp);
         if (!distrib) return 1;
         for (j = 0; j < 256; j++)
            for (k = 0; k < 4; k++) cam.dmap[j][k] = (cp.palette[(j * 256) / 256].color[k] * 255) * color_scalar;
         cam.cmap_size = cmap_size;
         cam.palette_mode = cp.palette_mode;
         if (cp.sample_density <= 0.0) return 1;
         scale = pow(2.0, cp.zoom);
         sample_density = cp.sample_density * scale * scale;
         ppux = cp.pixels_per_unit * scale;
         ppuy = field ? (ppux / 2.0) : ppux;
         ppux /= spec->pixel_aspect_ratio;
         if (field == flam3_field_even) shift = -0.5;
         if (field == flam3_field_odd) shift = 0.5;
         shift = shift / ppux;
         t0 = (double)gutter_width / (oversample * ppux);
         t1 = (double)gutter_width / (oversample * ppuy);
         corner0 = cp.center[0] - image_width / ppux / 2.0;
         corner1 = cp.center[1] - image_height / ppuy / 2.0;
         cam.bounds[0] = corner0 - t0;
         cam.bounds[1] = corner1 - t1 + shift;
         cam.bounds[2] = corner0 + image_width / ppux + t0;
         cam.bounds[3] = corner1 + image_height / ppuy + t1 + shift;
         cam.rot[0][0] = cos(cp.rotate * 2 * M_PI / 360.0);
         cam.rot[0][1] = -sin(cp.rotate * 2 * M_PI / 360.0);
         cam.rot[1][0] = -cam.rot[0][1];
         cam.rot[1][1] = cam.rot[0][0];
         cam.rot_center[0] = cp.rot_center[0];
         cam.rot_center[1] = cp.rot_center[1];
         cam.rotate = cp.rotate != 0.0;
         cam.width = width;
         cam.height = height;
         cam.ws0 = width * (1.0 / (cam.bounds[2] - cam.bounds[0]));
         cam.wb0s0 = cam.ws0 * cam.bounds[0];
         cam.hs1 = height * (1.0 / (cam.bounds[3] - cam.bounds[1]));
         cam.hb1s1 = cam.hs1 * cam.bounds[1];
         batch_size = sample_density * image_width * image_height / (nbatches * nts);
         stats->num_iters += batch_size;
         /* one thread: reseed from the frame stream (rect.c:858-865) */
         memset(&trc, 0, sizeof(trc));
         for (k = 0; k < 16; k++) trc.randrsl[k] = or_irand(&spec->rc);
         or_irandinit(&trc, 1);
         thread_bs = batch_size / 1.0;
         for (sub_batch = 0; sub_batch < thread_bs; sub_batch += SBS) {
            int sbs = (sub_batch + SBS > thread_bs) ? (int)(thread_bs - sub_batch) : SBS;
            points[0] = or_rand11(&trc);
            points[1] = or_rand11(&trc);
            points[2] = or_rand01(&trc);
            points[3] = or_rand01(&trc);
            badvals += or_iterate(&cp, sbs, fuse, points, distrib, &trc);
            or_deposit(&cam, points, sbs, bits, 0, buckets);
         }
         free(distrib);
         vibrancy += cp.vibrancy;
         gamma += cp.gamma;
         background[0] += cp.background[0];
         background[1] += cp.background[1];
         background[2] += cp.background[2];
         vib_gam_n++;
      }
      k1 = (cp.contrast * cp.brightness * 255 * 268.0 * (1.0 / (double)nbatches)) / 256;
      area = image_width * image_height / (ppux * ppuy);
      k2 = (oversample * oversample * nbatches) / (cp.contrast * area * 255 * sample_density * sumfilt);
      if (de.max_filter_index == 0)
         or_logscale(buckets, width, height, bits, k1, k2, accumulate);
      else
         or_density(buckets, width, height, bits, oversample, k1, k2, cp.estimator_curve, de.max_filtered_counts,
                    de.max_filter_index, de.kernel_size, de.filter_widths, de.filter_coefs, accumulate, phantom);
      if (de.max_filter_index > 0) {
         free(de.filter_coefs);
         free(de.filter_widths);
      }
   }
   {
      fl3h_tonemap_params tp;
      memset(&tp, 0, sizeof(tp));
      tp.image_width = image_width;
      tp.image_height = image_height;
      tp.oversample = oversample;
      tp.de_offset = de_offset;
      tp.filter_width = filter_width;
      tp.nchan = nchan;
      tp.transp = transp;
      tp.bytes_per_channel = spec->bytes_per_channel;
      tp.earlyclip = spec->earlyclip;
      tp.g = 1.0 / (gamma / vib_gam_n);
      tp.linrange = cp.gam_lin_thresh;
      tp.vibrancy = vibrancy / vib_gam_n;
      tp.highpow = highpow;
      for (k = 0; k < 3; k++) tp.background[k] = background[k] / (vib_gam_n / 256.0);
      or_tonemap(accumulate, width, height, bits, &tp, filter, out, out_width);
   }
   stats->badvals = badvals;
   free(filter);
   free(temporal_filter);
   free(temporal_deltas);
   free(buckets);
   free(accumulate);
   free(points);
   clear_cp(&cp, 0);
   return 0;
}
