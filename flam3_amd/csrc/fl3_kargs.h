// fl3_kargs.h - argument blocks of the HIP kernels, shared by the kernel
// translation units and the host launch code.
#ifndef FL3_KARGS_H
#define FL3_KARGS_H
#include <stdint.h>

// Independent renders in one launch (per-sample atomic deposits): the
// workgroups of a job are contiguous, workgroup b runs descs[blk_job[b]].
struct JobDesc {
   long long off;              // byte offset of the job in IterArgs::jobs
   long long nsamples;         // plotted samples of the job
   unsigned long long key0, key1;
   long long cell0;            // first histogram cell of the job's render
   int block0, nblocks;        // the job's workgroups
   int render;                 // counters of the render: badvals[2 render], saturation [2 render + 1]
   int pad;
};

struct IterArgs {
   const unsigned char *job;
   const JobDesc *descs;       // non-null: independent renders (direct deposits or job records)
   const int *blk_job;         // job of each workgroup of the launch
   int blk0;                   // the launch's first workgroup in the record / walker-state columns
   long long wtot;             // walkers of all launches sharing the records (0: this launch's)
   long long row0;             // independent renders with records: first row of this chunk
   // several jobs (the temporal samples of a batch) in one launch: workgroup
   // b iterates job b % njobs with walker key keys[2 * job .. +1]; nsamples
   // is then per job
   const unsigned char *jobs;
   const long long *job_off;
   const unsigned long long *keys;
   int njobs;
   // several jobs with their own sample counts (at most nsamples' share):
   // job j's plotted samples in all, and the rows of earlier chunks
   const long long *job_ns;
   long long ns_row0;
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
   // binned deposit: one 8-byte record per plotted sample, at i*W + walker;
   // rows_out rows are written (walkers with fewer samples pad with no-ops)
   int rec_mode;              // 0 deposit with atomics, 1 8-byte records, 2 4-byte records,
                              // 3 8-byte records carrying the job instead of the palette fraction
   unsigned long long *rec;   // 8-byte records, or 4-byte ones in the same buffer
   long long rows_out;
   unsigned *blk_hist;        // [workgroup][super-tile] record counts
   unsigned *bin_total;       // [tile] record counts over all workgroups (added atomically)
   int bin_shift, nbins;
   unsigned hist_off;         // LDS byte offset of the workgroup's tile counts
   // job records carrying the colour: [job][256] palette entries as bytes
   // R | G << 8 | B << 16 | A << 24 (null: palette index, xform and job),
   // held in LDS at pal_off
   const uint32_t *pal8;
   unsigned pal_off;
   // one xform per wavefront and step, points permuted between steps
   // through 6 KB of LDS at shuf_off (genomes without xaos); 0: per lane
   int wave_xf;
   unsigned shuf_off;
   // walker state kept between launches of one batch (10 words per walker,
   // structure of arrays): resume = continue from it without fuse steps
   uint32_t *wstate;
   int resume;
};

// Binned deposit limits: tiles per histogram (the sort keeps a counter per
// tile in LDS), samples per chunk of records, records per accumulation work
// item and the batch size from which the library prefers binning to atomics.
// A work item is what one workgroup accumulates alone: 2^18 records keeps a
// tile cell's partial sum below 2^18 * 255 < 2^32 and splits the densest
// tiles of a frame (2.2e6 of a chunk's 2.6e8 records at 1080p, against a mean
// of 1.3e5) over several workgroups, so that the last items do not trail.
#define FL3_MAX_BINS 8192
#define FL3_SUPER_SHIFT 5           // tiles per super-tile = 32

#define FL3_BIN_CHUNK_SAMPLES (1LL << 28)
#define FL3_CHUNK_BUFS 2            // record buffers in rotation: the walkers run one chunk ahead of the sort
#define FL3_BIN_ITEM_RECORDS (1 << 18)
#define FL3_BIN_MIN_SAMPLES (1LL << 20)

// Record formats of fl3k_launch_bin beyond 0 (8-byte records), 1 (4-byte
// records) and 2 (job records): the same records into double buckets.
#define FL3_BIN_F64_WIDE 3
#define FL3_BIN_F64_NARROW 4

// Binned deposit (walker records -> tile-sorted records -> LDS tiles).
struct BinArgs {
   const unsigned char *job;       // for palette / visibility tables
   void *buckets;
   const void *rec;                // records in walker order, one block of rows x 256 per workgroup
   void *sorted;                   // records grouped by tile
   unsigned *blk_hist;             // [walker workgroup][super-tile] counts
   int nblk;                       // walker workgroups (256 walkers each)
   long long walkers;              // row stride of rec
   long long rows;                 // record rows in this chunk
   unsigned *bin_total;            // [tile] record counts (summed by the walkers)
   unsigned *bin_start, *item_start, *item_counter;
   // two-pass sort: walker order -> super-tiles of FL3_SUPER tiles (sorted)
   // -> tiles (final, the record buffer reused)
   unsigned *soff;                 // [walker workgroup][super-tile] first slot in sorted
   unsigned *cursor;               // [tile] next free slot of the tile in final
   void *final;
   int nsuper;
   int shift;                      // cells per tile = 1 << shift
   int nbins;                      // tiles
   int chunk;                      // records per accumulation work item
   int round;                      // 4-byte records: records per LDS round (0: not packed)
   int width, height;
   // job records: palette (floor and float) and xform visibility per job
   const uint32_t *dmapu_all;      // [job][1024]
   const float *dmapf_all;         // [job][1024]
   const float *vis_all;           // [job][vis_stride]
   int vis_stride;
   // banded histograms: the records' cells are relative to cell_base; the
   // input of the first sort pass (or count pass) holds rec_count records
   // (0: every slot of the blocks is a record or FL3_REC_NONE)
   unsigned cell_base;
   long long rec_count;
   int rec_rel;                    // records hold cells relative to cell_base (4-byte band records)
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
   uint16_t *findex;        // per cell filter index, 0xffff = no contribution
   void *srcv;              // per cell AT[4]: colour x k1 log(1 + c3 k2) / c3
   const void *coefs_t;     // the kernel coefficients in the accumulator type
   double *srcl;            // per cell k1 log(1 + c3 k2) (the reference's per-tap numerator)
   int exact;               // sum each contribution as the reference does (double, in source order)
   int width, height, oversample;
   int max_filtered_counts, max_filter_index, kernel_size;
   int half;                // largest kernel radius (cells)
   double k1, k2, curve;
   // per 16x16 block: 1 + the largest kernel radius of its sources (0:
   // none), for each gather region's halo; nullptr: the widest kernel's halo
   // everywhere
   int *blkmax;
   int nbx;                 // blocks per row
   unsigned long long *taps;   // if set: 64 counters, summed: the gather footprint (2r + 1)^2 of all sources
   // float accumulator, ncoef > 0: the rows cut to what a tap reaches (row i
   // from ccoef[coff[i]], (r + 1)(r + 2) / 2 entries for radius r), which the
   // gather stages in ncoef floats of dynamic LDS; 0: taps read coefs_t
   const float *ccoef;
   const int *coff;
   int ncoef;
};
#define FL3_DE_LDS_COEFS 8192   // most floats of a cut filter bank the gather stages in LDS

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
