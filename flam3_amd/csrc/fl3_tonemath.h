/*
 * fl3_tonemath.h - per-pixel colour math of the final pass, FP64, callable
 * from host and device code.  Same formulas and branch structure as
 * palettes.c:199-349 (rgb2hsv, hsv2rgb, flam3_calc_alpha,
 * flam3_calc_newrgb) so that 8-bit output agrees with the reference to the
 * last bit except where a value sits on a rounding edge.
 */
#ifndef FL3_TONEMATH_H
#define FL3_TONEMATH_H

#include <math.h>

#if defined(__HIPCC__)
#define FL3_HD __host__ __device__ static inline
#else
#define FL3_HD static inline
#endif

#define FL3_PREFILTER_WHITE 255.0

FL3_HD void fl3_rgb2hsv(const double *rgb, double *hsv) {
   double rd = rgb[0], gd = rgb[1], bd = rgb[2], h, s, v, mx, mn, del;
   if (rd >= gd) mx = (rd >= bd) ? rd : bd;
   else mx = (gd >= bd) ? gd : bd;
   if (rd <= gd) mn = (rd <= bd) ? rd : bd;
   else mn = (gd <= bd) ? gd : bd;
   del = mx - mn;
   v = mx;
   s = (mx != 0.0) ? del / mx : 0.0;
   h = 0;
   if (s != 0.0) {
      double rc = (mx - rd) / del, gc = (mx - gd) / del, bc = (mx - bd) / del;
      if (rd == mx) h = bc - gc;
      else if (gd == mx) h = 2 + rc - bc;
      else if (bd == mx) h = 4 + gc - rc;
      if (h < 0) h += 6;
   }
   hsv[0] = h;
   hsv[1] = s;
   hsv[2] = v;
}

FL3_HD void fl3_hsv2rgb(const double *hsv, double *rgb) {
   double h = hsv[0], s = hsv[1], v = hsv[2], f, p, q, t, rd, gd, bd;
   int j;
   while (h >= 6.0) h = h - 6.0;
   while (h < 0.0) h = h + 6.0;
   j = (int)floor(h);
   f = h - j;
   p = v * (1 - s);
   q = v * (1 - (s * f));
   t = v * (1 - (s * (1 - f)));
   switch (j) {
      case 1: rd = q; gd = v; bd = p; break;
      case 2: rd = p; gd = v; bd = t; break;
      case 3: rd = p; gd = q; bd = v; break;
      case 4: rd = t; gd = p; bd = v; break;
      case 5: rd = v; gd = p; bd = q; break;
      default: rd = v; gd = t; bd = p; break;
   }
   rgb[0] = rd;
   rgb[1] = gd;
   rgb[2] = bd;
}

/* Alpha from normalised density, with a linear ramp below linrange. */
FL3_HD double fl3_calc_alpha(double dnorm, double gamma, double linrange) {
   double funcval = pow(linrange, gamma);
   if (dnorm > 0) {
      if (dnorm < linrange) {
         double frac = dnorm / linrange;
         return (1.0 - frac) * dnorm * (funcval / linrange) + frac * pow(dnorm, gamma);
      }
      return pow(dnorm, gamma);
   }
   return 0;
}

/* Scaled colour with highlight control: when a channel would exceed 255
 * and highpow >= 0, saturation is reduced in HSV instead of clipping. */
FL3_HD void fl3_calc_newrgb(const double *cbuf, double ls, double highpow, double *newrgb) {
   double maxa = -1.0, maxc = 0, newls, a;
   int k;
   if (ls == 0.0 || (cbuf[0] == 0.0 && cbuf[1] == 0.0 && cbuf[2] == 0.0)) {
      newrgb[0] = newrgb[1] = newrgb[2] = 0.0;
      return;
   }
   for (k = 0; k < 3; k++) {
      a = ls * (cbuf[k] / FL3_PREFILTER_WHITE);
      if (a > maxa) {
         maxa = a;
         maxc = cbuf[k] / FL3_PREFILTER_WHITE;
      }
   }
   if (maxa > 255 && highpow >= 0.0) {
      double hsv[3], lsratio;
      newls = 255.0 / maxc;
      lsratio = pow(newls / ls, highpow);
      for (k = 0; k < 3; k++) newrgb[k] = newls * (cbuf[k] / FL3_PREFILTER_WHITE) / 255.0;
      fl3_rgb2hsv(newrgb, hsv);
      hsv[1] *= lsratio;
      fl3_hsv2rgb(hsv, newrgb);
      for (k = 0; k < 3; k++) newrgb[k] *= 255.0;
   } else {
      double adjhlp = -highpow;
      newls = 255.0 / maxc;
      if (adjhlp > 1) adjhlp = 1;
      if (maxa <= 255) adjhlp = 1.0;
      for (k = 0; k < 3; k++)
         newrgb[k] = ((1.0 - adjhlp) * newls + adjhlp * ls) * (cbuf[k] / FL3_PREFILTER_WHITE);
   }
}

#endif
