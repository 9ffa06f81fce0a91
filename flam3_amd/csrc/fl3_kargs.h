// fl3_kargs.h - argument blocks of the HIP kernels, shared by the kernel
// translation units and the host launch code.
#ifndef FL3_KARGS_H
#define FL3_KARGS_H
#include <stdint.h>

struct IterArgs {
   const unsigned char *job;
   void *buckets;
   long long nsamples;      // fast: total plotted samples in the launch
   int fuse;
   unsigned long long key0, key1;
   unsigned long long *badvals;
   uint32_t *sat;
   // replay
   uint32_t *isaac;         // 36 words per walker, updated in place
   double batch_size;       // replay: samples per walker (double, rect.c:840)
   int sbs;                 // replay: sub-batch size
   int nwalkers;
};

struct PointsArgs {
   const unsigned char *job;
   const double *start;
   double *out;
   uint32_t *isaac;
   int *badvals;
   int n, fuse, nwalkers;
};

struct DEArgs {
   const void *buckets;
   void *accum;
   const double *coefs;
   const double *widths;
   int width, height, oversample;
   int max_filtered_counts, max_filter_index, kernel_size;
   double k1, k2, curve;
};

struct TMArgs {
   const void *accum;
   const double *filter;
   unsigned char *out;      // device image, image_width x image_height
   int W;                   // accumulator width
   int iw, ih, os, de_offset, fw;
   int nchan, transp, bpc, earlyclip;
   double g, linrange, vibrancy, highpow;
   double bg[3];
};

#endif
