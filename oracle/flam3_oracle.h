/*
 * flam3_oracle.h - TEST INFRASTRUCTURE.  A CPU restatement of the flam3
 * render hot path (iteration, deposit, density estimation, tone map) used
 * as the parity oracle for the GPU library.  Never linked into the product.
 *
 * Table construction (precalc flags, selection tables, filter banks) and
 * genome interpolation are taken from the reference library built in place
 * (oracle/_ref/libflam3ref_strict.so), which this library links against;
 * everything on the hot path is restated here.  The restatement is pinned
 * against the reference by the CPU tests (bit-identical iterate output and
 * byte-identical full renders at nthreads = 1).
 */
#ifndef FLAM3_ORACLE_H
#define FLAM3_ORACLE_H

#include "flam3_abi.h"
#include "fl3h.h"

/* ISAAC (isaac.c:25-110, isaac.h:49-52) with the LP64 word type. */
void or_isaac(randctx *ctx);
void or_irandinit(randctx *ctx, int flag);
unsigned long or_irand(randctx *ctx);
double or_rand01(randctx *ctx);
double or_rand11(randctx *ctx);

/* apply_xform (variations.c:2129-2399); cp prepared by the caller. */
int or_apply_xform(flam3_genome *cp, int fn, const double *p, double *q, randctx *rc);
/* flam3_iterate (flam3.c:232-299). */
int or_iterate(flam3_genome *cp, int n, int fuse, double *samples, const unsigned short *distrib, randctx *rc);

/* Deposit loop of rect.c:411-516.  mt = 0: single-thread adds
 * (dest = (uint)(dest + delta), flam3.c:3838-3840); mt = 1: atomic variant
 * (dest + (uint)delta, flam3.c:3690-3706). */
void or_deposit(const fl3h_camera *cam, const double *pts, long n, int bits, int mt, void *buckets);

/* de_thread (rect.c:57-251) over the whole histogram, single threaded.
 * phantom = 1 reads the row past the histogram from `phantom_row` as the
 * reference's memory layout does (the accumulator follows the buckets);
 * phantom = 0 treats it as empty. */
void or_density(const void *buckets, int width, int height, int bits, int oversample, double k1, double k2,
                double curve, int max_filtered_counts, int max_filter_index, int kernel_size,
                const double *filter_widths, const double *filter_coefs, void *accum, int phantom);
void or_logscale(const void *buckets, int width, int height, int bits, double k1, double k2, void *accum);
/* rect.c:1067-1247 */
void or_tonemap(void *accum, int width, int height, int bits, const fl3h_tonemap_params *tp,
                const double *filter, void *out, int out_stride_pixels);

/* render_rectangle (rect.c:529-1286) at nthreads = 1 (phantom as above). */
int or_render(flam3_frame *f, void *out, int field, int nchan, int transp, stat_struct *stats, int phantom);

#endif
