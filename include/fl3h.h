/*
 * fl3h.h - C ABI of the MI355X flam3 render path (libflam3_amd.so).
 *
 * The drop-in entry points are flam3_render and flam3_render_memory_required
 * with the reference's signatures and semantics.  The fl3h_* functions expose
 * the individual stages on host buffers so that each can be checked against
 * the reference in isolation; the reference has these stages only as static
 * functions or inline loops, cited per function below.  Every function
 * returns 0 on success, or non-zero with a message in fl3h_last_error().
 * Plain C types only; no device pointers cross this interface.
 */
#ifndef FL3H_H
#define FL3H_H

#include <stddef.h>
#include <stdint.h>
#include "flam3_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces flam3_render (flam3.h:623, flam3.c:3894-3932; body rect.c:529-1286).
 * out: W*H*nchan*bytes_per_channel bytes, rows of W pixels (every other row
 * with field = even/odd).  Returns 1 on invalid parameters (message on
 * stderr), 0 otherwise, including after an abort through f->progress. */
int flam3_render(flam3_frame *f, void *out, int field, int nchan, int transp, stat_struct *stats);

/* Replaces flam3_render_memory_required (flam3.h:625, flam3.c:3874-3887):
 * the reference's estimate, ss^2 * W * H * bytes * 9. */
double flam3_render_memory_required(flam3_frame *f);

/* Options of the extended render call. */
typedef struct {
   int replay;     /* 0: FP32 production walkers.  1: FP64 walkers that each
                      replay one reference iteration thread (f->nthreads of
                      them), with the reference's ISAAC streams and sub-batches */
   int nwalkers;   /* production walkers per launch, 0 = library choice */
   int deposit;    /* production histogram deposit: 0 library choice (tile-binned
                      for large batches), 1 per-sample atomics, 2 tile-binned */
   long long chunk_samples; /* binned deposit: samples per chunk (rounded up to
                      a multiple of the walker count), 0 = library choice */
   int temporal_batch; /* production: temporal samples per walker launch, 0 = all
                      of a batch (1 = one launch per temporal sample, the
                      reference's rect.c:754-931 loop) */
} fl3h_render_opts;

/* One render of fl3h_render_many: flam3_render's arguments and its return
 * value (rc). */
typedef struct {
   flam3_frame *frame;
   void *out;
   int field, nchan, transp;
   stat_struct *stats;
   int rc;
} fl3h_request;

/* Renders n frames, each exactly as fl3h_render(frame, out, field, nchan,
 * transp, stats, opts) would (same image, stats and f->rc), with small
 * renders sharing the GPU: the walkers of all temporal samples of the
 * renders with a stepped palette at bits 32/33 (up to 32e6 histogram cells
 * per launch), and of those with a linear palette below 2^24 plotted
 * samples (one batch, production mode, the bits of the first such render,
 * no progress callback) run in one launch, then their density estimation
 * and tone maps, and their images come back in one copy.  Other renders run
 * one after another.  Returns the number of
 * requests whose rc is not 0.  No counterpart in the reference: the
 * preview loop of flam3-genome (flam3-genome.c:1020-1116) and
 * flam3-animate's frame loop (flam3-animate.c:225-327) call flam3_render
 * once per image. */
int fl3h_render_many(fl3h_request *reqs, int n, const fl3h_render_opts *opts);

/* Render server (no counterpart in the reference; for callers that render
 * many independent images: a genome-search preview loop like
 * flam3-genome.c:1020-1116, or flam3-animate's frames): continuous batching
 * on one GPU.  A planner thread admits pending requests in arrival order
 * (up to max_batch renders whose histogram + accumulator bytes fit
 * hist_budget_bytes; the first always) and does their host work while a
 * runner thread has the previous batch on the GPU (fl3h_render_many's
 * shared launches).  Each request's frame, out and stats must stay valid
 * until fl3h_server_wait returns for it; each comes out as fl3h_render(...,
 * &opts) makes it alone. */
typedef struct {
   double hist_budget_bytes;   /* per batch; 0 = 1.15e9 (the 32e6 cells one shared launch bins) */
   int max_batch;              /* renders per batch; 0 = 512 */
   fl3h_render_opts opts;      /* for every request */
} fl3h_server_opts;
typedef struct fl3h_server fl3h_server;
fl3h_server *fl3h_server_create(const fl3h_server_opts *o);   /* NULL o: defaults */
/* Queues a flam3_render(f, out, field, nchan, transp, stats); returns its
 * ticket (>= 0) or -1. */
long long fl3h_server_submit(fl3h_server *s, flam3_frame *f, void *out, int field, int nchan, int transp,
                             stat_struct *stats);
/* Blocks until the request is done; returns what flam3_render would return
 * and, optionally, the seconds from submit to done. */
int fl3h_server_wait(fl3h_server *s, long long ticket, double *latency_seconds);
int fl3h_server_done(fl3h_server *s, long long ticket);   /* 1 once finished */
/* Batches run, renders done, and the seconds the planner spent on host
 * work and the runner on device work (each pointer may be NULL). */
void fl3h_server_counts(fl3h_server *s, long long *batches, long long *renders, double *prep_seconds,
                        double *gpu_seconds);
/* Finishes every submitted request, then stops the threads. */
void fl3h_server_destroy(fl3h_server *s);

/* flam3_render with options; flam3_render(f, ...) == fl3h_render(f, ..., NULL). */
int fl3h_render(flam3_frame *f, void *out, int field, int nchan, int transp, stat_struct *stats,
                const fl3h_render_opts *opts);

/* Batched flam3_iterate (flam3.h:555, flam3.c:232-299) in FP64 on the GPU:
 * walker w runs the reference loop from samples[w*4n .. w*4n+3] with ISAAC
 * context rc[w] and writes its n samples to samples[w*4n ...]; each rc[w] is
 * advanced as the reference advances it.  Like the reference, the caller has
 * run the precalc flags and built xform_distrib (flam3_create_xform_distrib
 * layout); the per-variation constants are recomputed on cp in place.
 * badvals (may be NULL) receives the per-walker bad-value counts. */
int fl3h_iterate(flam3_genome *cp, int n, int fuse, double *samples,
                 const unsigned short *xform_distrib, randctx *rc, int nwalkers, int *badvals);

/* flam3_xform_preview (flam3.h:543, flam3.c:301-349) on the GPU in FP64:
 * apply xform xi `depth` times to each point of the (2*numvals+1)^2 grid over
 * [-range, range]^2; result receives x,y pairs.  Prepares cp like the
 * reference does (precalc flags with xi's weight set to 1 temporarily). */
int fl3h_xform_preview(flam3_genome *cp, int xi, double range, int numvals, int depth,
                       double *result, randctx *rc);

/* One FP32 application of xform xi (apply_xform, variations.c:2129-2399) to
 * each point (x, y, colour, unused) in the production walkers' arithmetic and
 * variation tier (tier -1: the tier the genome's variations select; 0-5 a
 * restricted tier; 6 all variations).  out[4i..4i+3] = x, y, colour, 1.0 if
 * the result was a bad value.  No counterpart in the reference: it exposes
 * the FP32 walkers' xform step to tests. */
int fl3h_apply_xform_fp32(flam3_genome *cp, int xi, const double *points, long long npoints, double *out,
                          int tier);

/* Camera and palette of one temporal sample, as rect.c:778-827 sets them up. */
typedef struct {
   double bounds[4];       /* x0 y0 x1 y1, inclusive, gutter included */
   double rot[2][2];
   double rot_center[2];
   int rotate;             /* genome rotate != 0 */
   double ws0, wb0s0, hs1, hb1s1;
   int width, height;      /* histogram cells */
   int cmap_size;          /* 256 - apo_palette */
   int palette_mode;       /* 0 step, 1 linear */
   double dmap[256][4];    /* palette colour x 255 x temporal weight */
} fl3h_camera;

/* Deposit loop of rect.c:411-516 for a given point stream (x, y, colour,
 * visibility per point) into a histogram in the reference layout: 5 channels
 * per cell, uint32 (bits 32/33; truncated, saturating adds as
 * flam3.c:3690-3706) or double (bits 64).  buckets is read and updated. */
int fl3h_deposit_points(const fl3h_camera *cam, const double *points, long long npoints,
                        int bits, void *buckets);

/* Density estimation (de_thread, rect.c:57-251, with add_c_to_accum
 * flam3.c:3827-3835) or, when max_filter_index == 0, the plain log scaling
 * of rect.c:947-974.  buckets: width*height cells of 5 channels (uint32 for
 * bits 33, double for 64); accum: 4 channels (float / double), read and
 * updated.  The filter bank is flam3_create_de_filters' (filters.c:272). */
int fl3h_density_estimate(const void *buckets, int width, int height, int bits, int oversample,
                          double k1, double k2, double curve, int max_filtered_counts,
                          int max_filter_index, int kernel_size, int nfilters,
                          const double *filter_widths, const double *filter_coefs, void *accum);

/* The same stage with the production renders' arithmetic: contributions
 * formed in the accumulator type (the exact mode above forms each in double
 * as the reference does).  No counterpart in the reference: it exposes the
 * production path's density estimation to tests. */
int fl3h_density_estimate_production(const void *buckets, int width, int height, int bits, int oversample,
                                     double k1, double k2, double curve, int max_filtered_counts,
                                     int max_filter_index, int kernel_size, int nfilters,
                                     const double *filter_widths, const double *filter_coefs, void *accum);

/* fl3h_density_estimate on device buffers without allocating anything: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory during the call (code under HIP graph capture, tight serving
 * loops, harnesses that forbid allocation).  buckets and accum are DEVICE
 * pointers (the caller's memory, same layouts as fl3h_density_estimate);
 * the filter bank is host memory.  The intermediate per-cell data and the
 * device copy of the filter bank live in a fixed 64 MiB device array of the
 * library (one per GPU, held until the process ends): histograms above
 * fl3h_density_estimate_noalloc_max_cells(bits) cells (2,409,678 at bits 33,
 * 1,491,705 at bits 64) are rejected.  Runs on the null stream and returns
 * when the accumulator is complete.  Calls that overlap on one GPU share
 * that array and race.  Same arithmetic (the exact mode) and result as
 * fl3h_density_estimate. */
int fl3h_density_estimate_noalloc(const void *buckets, int width, int height, int bits, int oversample,
                                  double k1, double k2, double curve, int max_filtered_counts,
                                  int max_filter_index, int kernel_size, int nfilters,
                                  const double *filter_widths, const double *filter_coefs, void *accum);
long long fl3h_density_estimate_noalloc_max_cells(int bits);

typedef struct {
   int image_width, image_height, oversample, de_offset, filter_width;
   int nchan, transp, bytes_per_channel, earlyclip;
   double g;               /* 1 / mean gamma */
   double linrange;        /* gam_lin_thresh */
   double vibrancy, highpow;
   double background[3];   /* already scaled by 256 / nsamples as rect.c:1079 */
} fl3h_tonemap_params;

/* Spatial filter + tone map + quantisation of rect.c:1067-1247 (including
 * the earlyclip pre-pass) from an accumulator (4 channels, float for bits 33,
 * double for 64, uint32 for 32) into out (image_width*image_height*nchan*bpc). */
int fl3h_tonemap(const void *accum, int width, int height, int bits, const fl3h_tonemap_params *tp,
                 const double *filter, void *out);

/* Replaces flam3_estimate_bounding_box (flam3.h:588, flam3.c:3574-3639):
 * one FP64 chaos-game trajectory of nsamples points (fl3h_iterate, start
 * drawn from rc, 20 fused steps), then the eps / 1 - eps quantiles of x and
 * y (eps raised to 3x the bad-value rate, at most 0.3; the extremes when
 * eps * nsamples < 1).  rc advances as the reference advances it.  Returns
 * the number of bad values, or -1.  The genome-search preview loop
 * (flam3-genome.c:1033) calls it once per candidate genome. */
int fl3h_estimate_bounding_box(flam3_genome *cp, double eps, int nsamples, double *bmin, double *bmax,
                               randctx *rc);

/* Replaces flam3_improve_colors (flam3.h:586, palettes.c:476-504) for
 * change_palette = 0: the genome's colouring and ntries random recolourings
 * (the reference's sequence of random() calls) are scored by try_colors'
 * ~10000-pixel trial images (palettes.c:362-456, including its integer
 * score hits / res^3), all rendered in one fl3h_render_many call; g keeps
 * the first best colouring.  Palette changes (change_palette != 0) need the
 * palette library and return 1.  Returns 0, or 1 on failure. */
int fl3h_improve_colors(flam3_genome *g, int ntries, int change_palette, int color_resolution);

/* Genome ingestion on a host thread pool (no counterpart in the reference;
 * for a render server's or preview loop's input): XML documents parsed by
 * the caller's parser, with flam3_parse_xml2's signature (flam3.h:581,
 * parser.c:220-1256), on nthreads threads (0: 8).  The first document is
 * parsed alone (the parser initialises process state lazily), the rest in
 * parallel.  Each genome is checked for what rendering relies on (final
 * xform last, parser.c:802-821; positive sizes, oversampling, batches,
 * temporal samples and quality; xaos rows).  fl3h_ingest_wait hands over
 * the parser's genome array (the caller frees it as after
 * flam3_parse_xml2) and returns 0, or 1 when parsing or the checks failed. */
typedef flam3_genome *(*fl3h_parse_fn)(char *xml, char *filename, int default_flag, int *ncps);
typedef struct fl3h_ingest fl3h_ingest;
fl3h_ingest *fl3h_ingest_create(fl3h_parse_fn parse, int nthreads, int default_flag);
long long fl3h_ingest_submit(fl3h_ingest *q, const char *xml, const char *filename);
int fl3h_ingest_wait(fl3h_ingest *q, long long ticket, flam3_genome **genomes, int *ncps);
void fl3h_ingest_destroy(fl3h_ingest *q);

/* Seed f->rc from an explicit string the way flam3_init_frame seeds it from
 * the isaac_seed environment variable (flam3.h:639, flam3.c:2494-2516);
 * NULL seeds from the current time. */
void fl3h_init_frame(flam3_frame *f, const char *isaac_seed);

/* Genome construction for callers without the XML parser: defaults as
 * clear_cp (flam3.c:1270) and parse_flame_element leave them, nstd standard
 * xforms (+ a final xform) with all variation weights 0.  Free with
 * fl3h_genome_free. */
flam3_genome *fl3h_genome_new(int nstd, int has_final);
void fl3h_genome_free(flam3_genome *g);

/* Replaces flam3_interpolate (flam3.h:562, flam3.c:717-802 with
 * interpolation.c:373-1031): the genome at `time` between the control
 * points genomes[0..ngenomes), sorted by time, written into *result (whose
 * xforms, if any, are released first; zero-initialise a new result).
 * Returns 0, or 1 when ngenomes < 1 or memory runs out. */
int fl3h_interpolate(const flam3_genome *genomes, int ngenomes, double time, double stagger,
                     flam3_genome *result);
/* Release the xforms and xaos rows of a genome filled by fl3h_interpolate
 * (the struct itself stays the caller's). */
void fl3h_genome_clear(flam3_genome *g);

/* Launch tuning (no counterpart in the reference): the parameters the
 * autotuner (flam3_amd/autotune.py) chooses, loaded once per process from
 * the file named by the FL3H_TUNING environment variable, else from
 * tuning.conf beside the library ("key value" lines), else defaults.
 * Keys: wg_per_cu (walker workgroups of 256 per CU for launches of several
 * chunks, default 8), wg_per_cu_single (the same for launches of one chunk,
 * 8), min_samples
 * (samples per walker at least, 256), min_samples_small (the same for
 * launches of fewer than small_samples samples, 128; small_samples 2^23),
 * chunk_samples (binned deposit chunk,
 * 2^28), de_halo (density estimation: each gather region scans the halo its
 * neighbourhood's kernels reach 1, the widest kernel's halo everywhere 0).  With value non-NULL the key is set for the rest of the process; out
 * (may be NULL) receives the current value.  Returns 0, or 1 for an unknown
 * key or a value out of range.  The walker parameters shape the sample
 * streams: renders made under different settings are different samples of
 * the same image. */
int fl3h_tuning(const char *key, const double *value, double *out);

/* Stage timing (no counterpart in the reference): with profiling enabled
 * (reset(1); reset(2) also counts the DE footprint) every kernel launch of
 * the render path is bracketed by HIP events on the stream it runs on;
 * totals accumulate until the next reset. */
typedef struct {
   double iterate_ms, density_ms, tonemap_ms;
   long long iterate_launches, density_launches, tonemap_launches;
   long long samples;      /* plotted samples launched */
   long long cells;        /* histogram cells of the last density pass */
   double bin_ms;          /* tile-binned deposit: sort + LDS accumulation */
   long long bin_launches;
   long long de_taps;      /* density estimation: gather footprint, sum of (2r + 1)^2 over sources
                              (counted with fl3h_profile_reset(2) only: the count costs time) */
} fl3h_stage_times;
void fl3h_profile_reset(int enable);
int fl3h_profile_get(fl3h_stage_times *out);

/* Hardware-counter calibration (no counterpart in the reference): copies
 * `bytes` (a multiple of 16) between two fresh device buffers with one plain
 * copy kernel, k_copy_probe, so that a profiler's byte counters can be
 * scaled by a kernel of known traffic (`bytes` read, `bytes` written). */
int fl3h_copy_probe(long long bytes);

const char *fl3h_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
