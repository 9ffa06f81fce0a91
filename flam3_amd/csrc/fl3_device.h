/*
 * fl3_device.h - internal C interface between the host render driver
 * (fl3_render.c) and the HIP kernels (fl3_device.hip).  Device buffers live
 * in a per-device context owned by fl3_device.hip and are reused across
 * renders; every call returns 0 on success and sets fl3d_error() otherwise.
 */
#ifndef FL3_DEVICE_H
#define FL3_DEVICE_H

#include <stddef.h>
#include <stdint.h>
#include "fl3_host.h"

#ifdef __cplusplus
extern "C" {
#endif

const char *fl3d_error(void);

/* Histogram + accumulator for one render, cleared. */
int fl3d_begin(int width, int height, int bits);
/* Device bytes a render may use (the device's free memory plus what the
 * library's buffers already hold) and the device's total memory. */
int fl3d_device_memory(double *avail, double *total);
/* Zero the histogram (start of each batch, rect.c:721). */
int fl3d_clear_buckets(void);
/* Histograms of n renders in one allocation (cleared); fl3d_select(r) makes
 * render r the one that fl3d_clear_buckets, fl3d_iterate*, fl3d_density,
 * fl3d_logscale and fl3d_tonemap* work on.  fl3d_begin(w, h, bits) is the
 * one-render case. */
int fl3d_begin_many(const int *width, const int *height, int n, int bits);
int fl3d_select(int r);

typedef struct {
   int fuse;            /* 15, or 100 with earlyclip (rect.c:36-37, 266) */
   int64_t nsamples;    /* plotted samples in this launch */
   uint64_t key[2];     /* walker stream key drawn from the frame's ISAAC */
   int exact;           /* saturating adds on every channel */
   int nwalkers;        /* 0: chosen by the library */
   int deposit;         /* 0 auto, 1 atomics, 2 tile-binned */
   int64_t chunk;       /* binned: samples per chunk, 0 = library choice */
   const int64_t *job_nsamples;   /* fl3d_iterate_batch: job j plots job_nsamples[j] (at most
                                     nsamples); NULL: nsamples each */
} fl3d_iter_params;

/* Iterate one job (one genome at one temporal sample) and deposit into the
 * histogram.  *saturated is set when a hit counter reached UINT_MAX in the
 * fast mode (the caller then repeats the batch with exact = 1). */
int fl3d_iterate(const void *job, size_t job_bytes, const fl3d_iter_params *p,
                 double *badvals, int *saturated);
/* Jobs of several begun renders in one launch (per-sample atomic deposits
 * into each job's render); walkers: a multiple of 256, e.g. from
 * fl3d_job_walkers(nsamples, n, 0) for a run of n jobs of nsamples each,
 * which reproduces fl3d_iterate_batch's walkers.  badvals / saturated: one
 * entry per render.  exact: every job deposits with saturating adds. */
typedef struct {
   const void *job;
   size_t bytes;
   int64_t nsamples;
   uint64_t key[2];
   int render;
   int walkers;
   int exact;       /* saturating deposits (a render that could reach 2^32 in a cell) */
   int run;         /* jobs of a render with equal run share their kernel tier, as
                       one fl3d_iterate_batch launch of that render alone would */
} fl3d_multi_job;
long long fl3d_job_walkers(int64_t nsamples, int njobs, int requested);
/* Size the histogram, accumulator and density-estimation buffers for renders
 * of hist_bytes histogram and accumulator bytes in all (no later growth). */
int fl3d_reserve(double hist_bytes);
int fl3d_iterate_multi(const fl3d_multi_job *jobs, int njobs, int fuse, int exact, int nrenders,
                       double *badvals, int *saturated);
/* The same for njobs jobs in one launch (the temporal samples of a batch):
 * p->nsamples samples each (or p->job_nsamples[j]), job j's walkers keyed by
 * keys[2j], keys[2j+1]. */
int fl3d_iterate_batch(const void *const *jobs, const size_t *job_bytes, int njobs, const fl3d_iter_params *p,
                       const uint64_t *keys, double *badvals, int *saturated);

typedef struct {
   int oversample;
   double k1, k2, curve;
   int exact;           /* contributions formed and summed as the reference does */
} fl3d_de_params;

/* Density estimation into the accumulator (rect.c:57-251). */
int fl3d_density(const fl3d_de_params *p, const fl3_de_filters *de);
/* Log scaling without DE (rect.c:947-974). */
int fl3d_logscale(double k1, double k2);
/* Density estimation (or log scaling where de[i] is NULL or has no filters)
 * of renders 0..n-1 of fl3d_begin_many, batched into one launch pair where
 * possible. */
int fl3d_density_many(const fl3d_de_params *p, const fl3_de_filters *const *de, int n);

typedef struct {
   int image_width, image_height;  /* output pixels (height halved for fields) */
   int oversample, de_offset, filter_width;
   int nchan, transp, bytes_per_channel, earlyclip;
   double g, linrange, vibrancy, highpow;
   double background[3];
} fl3d_tm_params;

/* Spatial filter + tone map + quantise (rect.c:1067-1247), written to a
 * host buffer with a row stride of out_stride bytes. */
int fl3d_tonemap(const fl3d_tm_params *p, const double *filter, void *out_host, size_t out_stride);
/* The same in three steps for several renders: device buffers for all
 * images and filters, one tone map per selected render at byte offsets,
 * then one download into the callers' rows. */
typedef struct {
   size_t img_off;
   void *out;
   size_t out_stride, row_bytes;
   int rows;
} fl3d_image_dst;
int fl3d_images_begin(size_t image_bytes, size_t filter_bytes);
int fl3d_tonemap_async(const fl3d_tm_params *p, const double *filter, size_t img_off, size_t filt_off);
int fl3d_images_fetch(const fl3d_image_dst *d, int n);
/* The first `bytes` of the image arena into pinned host buffer `slot` (0 or
 * 1, kept until the next download into the same slot); *host points at it. */
int fl3d_images_download(int slot, size_t bytes, const unsigned char **host);
/* 1 if the two jobs may share 4-byte records in one launch: job0 has a
 * stepped palette and visibility 1 for every xform, and job has the same
 * palette and visibilities. */
int fl3d_jobs_share_palette(const void *job0, const void *job);
/* Tone maps of renders 0..n-1 of fl3d_begin_many in one launch (after
 * fl3d_images_begin; render i's image at img_off[i]). */
int fl3d_tonemap_many(const fl3d_tm_params *p, const double *const *filters, const size_t *img_off, int n);

/* Replay iteration: nwalkers FP64 walkers, walker w being reference thread w
 * with ISAAC state isaac[36w..36w+35] (count, 16 results, 16 memory words,
 * a, b, c; updated in place), batch_size samples each in sub-batches of sbs. */
int fl3d_replay(const void *job, size_t job_bytes, uint32_t *isaac, int nwalkers,
                double batch_size, int sbs, int fuse, double *badvals);

/* Device-resident copies for tests. */
int fl3d_download_buckets(void *host, size_t bytes);
int fl3d_download_accum(void *host, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif
