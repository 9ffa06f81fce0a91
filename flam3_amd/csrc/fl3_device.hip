// fl3_device.hip - host side of the MI355X (gfx950) render path: the
// per-device context (buffers reused across renders, one stream), kernel
// launches, and the C interfaces over them.  The kernels themselves live in
// fl3_kernels_fast.hip and fl3_kernels_exact.hip:
//
//  * walker kernels (chaos game + histogram deposit, rect.c:253-527 with
//    flam3.c:232-299 and variations.c:2129-2399 inlined): one thread is one
//    walker; the genome lives in LDS in the iteration precision.
//      - k_iter_fast: FP32 walkers with their own counter streams (keyed from
//        the frame's ISAAC stream), for production renders;
//      - k_iter_replay: FP64 walkers that each run exactly one reference
//        thread's ISAAC stream and sub-batch loop (replay / verification);
//      - k_iterate_points: FP64 batched flam3_iterate.
//  * k_deposit_points: deposit a given point stream (stage test).
//  * k_de / k_logscale: density estimation or plain log scaling into the
//    accumulator (rect.c:57-251, 947-974).
//  * k_earlyclip + k_tonemap: spatial filter, log/gamma/vibrancy/highlight
//    tone map and quantisation fused into one pass (rect.c:1067-1247).
#include <hip/hip_runtime.h>
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <atomic>
#include <mutex>
#include <type_traits>
#include <vector>

#include "fl3_pack.h"
#include "fl3_core.h"
#include "fl3_launch.h"
#include "fl3_tonemath.h"
#include "fl3_device.h"
#include "fl3h.h"

// ------------------------------------------------------------------ errors

static thread_local char g_err[512];

static void set_err(const char *fmt, const char *a = "", long b = 0) {
   snprintf(g_err, sizeof(g_err), fmt, a, b);
}

#define HIPCHK(call)                                                          \
   do {                                                                       \
      hipError_t e_ = (call);                                                 \
      if (e_ != hipSuccess) {                                                 \
         snprintf(g_err, sizeof(g_err), "%s failed: %s (line %d)", #call,    \
                  hipGetErrorString(e_), __LINE__);                           \
         return 1;                                                            \
      }                                                                       \
   } while (0)

extern "C" const char *fl3d_error(void) { return g_err; }
extern "C" const char *fl3h_last_error(void) { return g_err; }

// ------------------------------------------------------------------ device context

// A single image of at least FL3_STRIP_MIN_BYTES comes back in this many
// strips of rows, the host copying each one out while the next arrives.
#define FL3_IMAGE_STRIPS 4
#define FL3_STRIP_MIN_BYTES (2u << 20)

struct DevBuf {
   void *p = nullptr;
   size_t cap = 0;
   int ensure(size_t bytes) {
      if (bytes <= cap) return 0;
      const size_t cap0 = cap;
      if (p) (void)hipFree(p);
      p = nullptr;
      cap = 0;
      // (a buffer that has to grow again at least doubles, up to 2 GiB: each
      // growth frees and allocates, which stalls the device for milliseconds)
      size_t want = bytes + bytes / 8;
      if (cap0 && cap0 < ((size_t)1 << 31) && want < 2 * cap0) want = 2 * cap0;
      if (hipMalloc(&p, want) != hipSuccess) {
         snprintf(g_err, sizeof(g_err), "hipMalloc of %zu bytes failed", want);
         (void)hipGetLastError();   // (reported here; later launches must not see it)
         p = nullptr;
         return 1;
      }
      cap = want;
      return 0;
   }
};

struct Ctx {
   int ready = 0;
   int device = 0;
   int ncu = 256;
   size_t lds_max = 65536;   // LDS bytes one workgroup may allocate
   hipStream_t stream = nullptr;
   // The tile-binned deposit sorts and accumulates chunk k on bin_stream
   // while the walkers produce the next chunks on stream: chunk k's records
   // and tile counts are set k % FL3_CHUNK_BUFS of crec/cblk/ctot.
   hipStream_t bin_stream = nullptr;
   hipEvent_t ev_walk[FL3_CHUNK_BUFS] = {}, ev_bin[FL3_CHUNK_BUFS] = {}, ev_join = nullptr;
   DevBuf buckets, accum, job, table, image, isaac, points, misc, findex;
   // what the DE table buffer holds: the coefficients and widths of the last
   // upload, valid while the buffer is the same allocation (address, size)
   std::vector<double> de_up;
   void *de_up_p = nullptr;
   size_t de_up_cap = 0;
   int de_up_nf = 0, de_up_ks = 0;   // its row count in use (max_filter_index + 1) and row length
   DevBuf crec[FL3_CHUNK_BUFS], cblk[FL3_CHUNK_BUFS], ctot[FL3_CHUNK_BUFS];   // tile-binned deposit
   DevBuf sorted, wstate, bins, jobtab, soff;
   DevBuf srcv, srcl;                  // density estimation per-source values
   DevBuf detab, deargs;               // batched density estimation: tables, per-render arguments
   DevBuf deblk;                       // DE: widest source kernel per 16x16 block
   DevBuf detaps;                      // DE gather footprint counters (stage profile)
   DevBuf blk2, soff2, bins2;          // banded histograms: second-level block counts, offsets, counters
   unsigned *band_host = nullptr;      // pinned: first-level tile starts of the last chunk
   hipEvent_t ev_band = nullptr;
   hipEvent_t ev_strip[FL3_IMAGE_STRIPS] = {};   // fl3d_images_fetch: strip k of the image is in staging
   unsigned long long *counters = nullptr;   // [0] badvals, [1] saturation flag
   void *staging = nullptr;                   // pinned host buffer for the image download
   size_t staging_cap = 0;
   void *slot[2] = {nullptr, nullptr};        // pinned buffers of fl3d_images_download
   size_t slot_cap[2] = {0, 0};
   int width = 0, height = 0, bits = 33;
   // Several renders may share the histogram buffers (fl3d_begin_many):
   // render r's cells start at rcell0[r]; width/height/cell0 are those of
   // the selected render, which the per-render calls work on.
   std::vector<long long> rcell0;
   std::vector<int> rwidth, rheight;
   long long cell0 = 0;
   DevBuf rcount, descs;   // per-render counters, job descriptors + workgroup table
};

static Ctx g_ctx;
static std::recursive_mutex g_mu;
static std::atomic<int> g_ready{0};   // g_ctx.ready, for readers without g_mu

static int ctx_init() {
   if (g_ctx.ready) return 0;
   HIPCHK(hipGetDevice(&g_ctx.device));
   hipDeviceProp_t prop;
   HIPCHK(hipGetDeviceProperties(&prop, g_ctx.device));
   g_ctx.ncu = prop.multiProcessorCount;
   g_ctx.lds_max = prop.sharedMemPerBlock;
   HIPCHK(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
   HIPCHK(hipStreamCreateWithFlags(&g_ctx.bin_stream, hipStreamNonBlocking));
   for (int i = 0; i < FL3_CHUNK_BUFS; i++) {
      HIPCHK(hipEventCreateWithFlags(&g_ctx.ev_walk[i], hipEventDisableTiming));
      HIPCHK(hipEventCreateWithFlags(&g_ctx.ev_bin[i], hipEventDisableTiming));
   }
   HIPCHK(hipEventCreateWithFlags(&g_ctx.ev_join, hipEventDisableTiming));
   HIPCHK(hipEventCreateWithFlags(&g_ctx.ev_band, hipEventDisableTiming));
   for (int i = 0; i < FL3_IMAGE_STRIPS; i++)
      HIPCHK(hipEventCreateWithFlags(&g_ctx.ev_strip[i], hipEventDisableTiming));
   HIPCHK(hipHostMalloc((void **)&g_ctx.band_host, sizeof(unsigned) * (FL3_MAX_BINS + 1), hipHostMallocDefault));
   HIPCHK(hipMalloc((void **)&g_ctx.counters, 64));
   g_ctx.ready = 1;
   g_ready.store(1, std::memory_order_release);
   return 0;
}

// Device memory a render may use: the device's free bytes plus what this
// context already holds (its buffers are reused for the next renders).
extern "C" int fl3d_device_memory(double *avail, double *total) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   if (ctx_init()) return 1;
   size_t fr = 0, tot = 0;
   HIPCHK(hipMemGetInfo(&fr, &tot));
   const DevBuf *bufs[] = {&g_ctx.buckets, &g_ctx.accum, &g_ctx.job, &g_ctx.table, &g_ctx.image, &g_ctx.isaac,
                           &g_ctx.points, &g_ctx.misc, &g_ctx.findex, &g_ctx.sorted, &g_ctx.wstate, &g_ctx.bins,
                           &g_ctx.jobtab, &g_ctx.soff, &g_ctx.srcv, &g_ctx.srcl, &g_ctx.detab, &g_ctx.deargs,
                           &g_ctx.rcount, &g_ctx.descs, &g_ctx.deblk};
   double held = 0.0;
   for (const DevBuf *b : bufs) held += (double)b->cap;
   for (int i = 0; i < FL3_CHUNK_BUFS; i++)
      held += (double)(g_ctx.crec[i].cap + g_ctx.cblk[i].cap + g_ctx.ctot[i].cap);
   if (avail) *avail = (double)fr + held;
   if (total) *total = (double)tot;
   return 0;
}

// ------------------------------------------------------------------ stage timing
//
// Optional HIP-event timing of each stage on the context's stream (the
// stream every kernel runs on).  Pending intervals are folded in at the next
// synchronisation point of the render.
struct StageTimer {
   int enabled = 0;
   hipEvent_t ev[64][2];
   int stage[64];
   int n = 0, created = 0;
   double ms[5] = {0, 0, 0, 0, 0};
   long long launches[5] = {0, 0, 0, 0, 0};
   long long samples = 0, cells = 0;
};
static StageTimer g_prof;
enum { ST_ITER = 0, ST_DE = 1, ST_TONE = 2, ST_BIN = 3, ST_OTHER = 4 };

static void prof_flush() {
   for (int i = 0; i < g_prof.n; i++) {
      float ms = 0.0f;
      if (hipEventSynchronize(g_prof.ev[i][1]) == hipSuccess &&
          hipEventElapsedTime(&ms, g_prof.ev[i][0], g_prof.ev[i][1]) == hipSuccess) {
         g_prof.ms[g_prof.stage[i]] += ms;
         g_prof.launches[g_prof.stage[i]]++;
      }
   }
   g_prof.n = 0;
}

static int prof_begin(int stage, hipStream_t st = nullptr) {
   if (!g_prof.enabled) return -1;
   if (!g_prof.created) {
      for (int i = 0; i < 64; i++) {
         if (hipEventCreate(&g_prof.ev[i][0]) != hipSuccess || hipEventCreate(&g_prof.ev[i][1]) != hipSuccess)
            return -1;
      }
      g_prof.created = 1;
   }
   if (g_prof.n == 64) prof_flush();
   const int k = g_prof.n++;
   g_prof.stage[k] = stage;
   (void)hipEventRecord(g_prof.ev[k][0], st ? st : g_ctx.stream);
   return k;
}

static void prof_end(int k, hipStream_t st = nullptr) {
   if (k >= 0) (void)hipEventRecord(g_prof.ev[k][1], st ? st : g_ctx.stream);
}

extern "C" void fl3h_profile_reset(int enable) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   prof_flush();
   g_prof.enabled = enable;
   for (int i = 0; i < 5; i++) {
      g_prof.ms[i] = 0.0;
      g_prof.launches[i] = 0;
   }
   g_prof.samples = 0;
   g_prof.cells = 0;
   if (g_ctx.ready && g_ctx.detaps.ensure(64 * sizeof(unsigned long long)) == 0) {
      (void)hipMemsetAsync(g_ctx.detaps.p, 0, 64 * sizeof(unsigned long long), g_ctx.stream);
      (void)hipStreamSynchronize(g_ctx.stream);
   }
}

// (counted only with profiling level 2: the count adds a reduction and
// atomics to the preparation pass)
static unsigned long long *de_taps() {
   return g_prof.enabled >= 2 && g_ctx.detaps.p ? (unsigned long long *)g_ctx.detaps.p : nullptr;
}

extern "C" int fl3h_profile_get(fl3h_stage_times *out) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   prof_flush();
   out->iterate_ms = g_prof.ms[ST_ITER];
   out->iterate_launches = g_prof.launches[ST_ITER];
   out->density_ms = g_prof.ms[ST_DE];
   out->density_launches = g_prof.launches[ST_DE];
   out->tonemap_ms = g_prof.ms[ST_TONE];
   out->tonemap_launches = g_prof.launches[ST_TONE];
   out->samples = g_prof.samples;
   out->cells = g_prof.cells;
   out->bin_ms = g_prof.ms[ST_BIN];
   out->bin_launches = g_prof.launches[ST_BIN];
   out->de_taps = 0;
   if (g_ctx.ready && g_ctx.detaps.p) {
      unsigned long long t[64];
      if (hipMemcpy(t, g_ctx.detaps.p, sizeof(t), hipMemcpyDeviceToHost) == hipSuccess)
         for (int i = 0; i < 64; i++) out->de_taps += (long long)t[i];
   }
   return 0;
}

extern "C" int fl3h_copy_probe(long long bytes) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   if (ctx_init()) return 1;
   if (bytes <= 0 || bytes % 16) {
      set_err("fl3h_copy_probe: bytes must be a positive multiple of 16, not %s%ld", "", (long)bytes);
      return 1;
   }
   void *src = nullptr, *dst = nullptr;
   HIPCHK(hipMalloc(&src, (size_t)bytes));
   if (hipMalloc(&dst, (size_t)bytes) != hipSuccess) {
      (void)hipFree(src);
      set_err("fl3h_copy_probe: out of device memory");
      return 1;
   }
   int rc = fl3k_launch_copy_probe(src, dst, bytes / 16, g_ctx.stream);
   if (hipStreamSynchronize(g_ctx.stream) != hipSuccess) rc = 1;
   (void)hipFree(src);
   (void)hipFree(dst);
   if (rc) set_err("fl3h_copy_probe: launch failed");
   return rc;
}

static size_t bucket_bytes(int bits) { return bits == 64 ? 40 : 20; }
#define DG_REGION_HOST 16   // cells per side of a density-estimation workgroup (DG_REGION)
static size_t accum_bytes(int bits);
// the selected render's histogram and accumulator
static void *cur_buckets() { return (unsigned char *)g_ctx.buckets.p + g_ctx.cell0 * bucket_bytes(g_ctx.bits); }
static void *cur_accum() { return (unsigned char *)g_ctx.accum.p + g_ctx.cell0 * accum_bytes(g_ctx.bits); }
static size_t accum_bytes(int bits) { return bits == 64 ? 32 : 16; }

extern "C" int fl3d_begin(int width, int height, int bits) {
   return fl3d_begin_many(&width, &height, 1, bits);
}

extern "C" int fl3d_begin_many(const int *width, const int *height, int n, int bits) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   if (ctx_init()) return 1;
   if (n < 1) {
      set_err("fl3d_begin_many: no renders");
      return 1;
   }
   g_ctx.rcell0.resize(n);
   g_ctx.rwidth.assign(width, width + n);
   g_ctx.rheight.assign(height, height + n);
   size_t cells = 0;
   for (int r = 0; r < n; r++) {
      g_ctx.rcell0[r] = (long long)cells;
      cells += (size_t)width[r] * height[r];
   }
   if (g_ctx.buckets.ensure(cells * bucket_bytes(bits) + 16)) return 1;
   if (g_ctx.accum.ensure(cells * accum_bytes(bits) + 16)) return 1;
   if (g_ctx.rcount.ensure(16 * (size_t)n)) return 1;
   g_ctx.bits = bits;
   HIPCHK(hipMemsetAsync(g_ctx.accum.p, 0, cells * accum_bytes(bits), g_ctx.stream));
   HIPCHK(hipMemsetAsync(g_ctx.buckets.p, 0, cells * bucket_bytes(bits), g_ctx.stream));
   return fl3d_select(0);
}

// Device buffers sized up front for renders whose histograms and
// accumulators take hist_bytes in all (a server's admission budget), so
// that no batch within it stalls on growing them: histogram, accumulator
// and the density estimation's per-cell buffers.
extern "C" int fl3d_reserve(double hist_bytes) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   if (ctx_init()) return 1;
   if (!(hist_bytes > 0)) return 0;
   // per cell 20 + 16 B (bits 32/33) or 40 + 32 B (bits 64): the buckets
   // take 5/9 of either, the accumulator 4/9; at most hist_bytes / 36 cells.
   // A batch's groups share the buffers one after another, each of at most
   // the shared launch's 32e6 cells (a larger lone render grows them).
   if (hist_bytes > 32e6 * 36.0) hist_bytes = 32e6 * 36.0;
   const size_t cells = (size_t)(hist_bytes / 36.0) + 1;
   if (g_ctx.buckets.ensure((size_t)(hist_bytes * 5.0 / 9.0) + 16) ||
       g_ctx.accum.ensure((size_t)(hist_bytes * 4.0 / 9.0) + 16) || g_ctx.findex.ensure(sizeof(uint16_t) * cells) ||
       g_ctx.srcv.ensure(16 * cells) || g_ctx.srcl.ensure(sizeof(double) * cells) ||
       g_ctx.deblk.ensure(sizeof(int) * (cells / (DG_REGION_HOST * DG_REGION_HOST) + 4096)))
      return 1;
   return 0;
}

extern "C" int fl3d_select(int r) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   if (r < 0 || r >= (int)g_ctx.rcell0.size()) {
      set_err("fl3d_select: no render %s%ld", "", (long)r);
      return 1;
   }
   g_ctx.cell0 = g_ctx.rcell0[r];
   g_ctx.width = g_ctx.rwidth[r];
   g_ctx.height = g_ctx.rheight[r];
   return 0;
}

extern "C" int fl3d_clear_buckets(void) {
   size_t n = (size_t)g_ctx.width * g_ctx.height;
   HIPCHK(hipMemsetAsync(cur_buckets(), 0, n * bucket_bytes(g_ctx.bits), g_ctx.stream));
   return 0;
}

static int upload_job(const void *job, size_t bytes) {
   if (g_ctx.job.ensure(bytes)) return 1;
   HIPCHK(hipMemcpyAsync(g_ctx.job.p, job, bytes, hipMemcpyHostToDevice, g_ctx.stream));
   return 0;
}

static int read_counters(double *badvals, int *saturated) {
   unsigned long long c[2];
   HIPCHK(hipMemcpyAsync(c, g_ctx.counters, sizeof(c), hipMemcpyDeviceToHost, g_ctx.stream));
   HIPCHK(hipStreamSynchronize(g_ctx.stream));
   if (badvals) *badvals = (double)c[0];
   if (saturated) *saturated = c[1] != 0;
   return 0;
}

// Walkers per launch: 8 workgroups of 256 per CU, never fewer than ~256
// samples per walker (each walker pays its fuse steps once).
// ------------------------------------------------------------------ tuning
//
// Launch parameters chosen by the autotuner (flam3_amd/autotune.py), read at
// set-up from the file FL3H_TUNING names, else from tuning.conf beside the
// library ("key value" lines), else the defaults below.  They shape the
// walkers (and so the sample streams) the same way for every render of the
// process, alone or shared.
struct Tuning {
   int wg_per_cu = 8;               // walker workgroups of 256 per CU (launches of several chunks)
   int wg_per_cu_single = 8;        // the same for launches of one chunk (no sort to overlap)
   // samples per walker at least, for launches of at least small_samples
   // samples; smaller ones (previews) take at least min_samples_small each:
   // more walkers with fewer samples, so fewer steps in a row.  Every walker
   // also pays its fuse steps, so fewer samples each cost throughput
   // (preview serving with 64 / 128 / 256 for every launch: p50 1.27 / 1.63 /
   // 2.48 ms at 1000 requests/s, 4700 / 5900 / 6900 requests/s all at once)
   long long min_samples = 256;
   long long min_samples_small = 128;
   long long small_samples = 1LL << 23;
   long long chunk_samples = FL3_BIN_CHUNK_SAMPLES;   // binned deposit: samples per chunk
   int de_halo = 1;                 // DE gather halo: per region from its sources (1) or the widest kernel's (0)
};
static Tuning g_tune;
static std::once_flag g_tune_once;

static int tune_set(const char *key, double v) {
   if (!strcmp(key, "wg_per_cu") && v >= 1 && v <= 16) g_tune.wg_per_cu = (int)v;
   else if (!strcmp(key, "wg_per_cu_single") && v >= 1 && v <= 16) g_tune.wg_per_cu_single = (int)v;
   else if (!strcmp(key, "min_samples") && v >= 1 && v <= 65536) g_tune.min_samples = (long long)v;
   else if (!strcmp(key, "min_samples_small") && v >= 1 && v <= 65536) g_tune.min_samples_small = (long long)v;
   else if (!strcmp(key, "small_samples") && v >= 0 && v <= (double)(1LL << 40)) g_tune.small_samples = (long long)v;
   else if (!strcmp(key, "chunk_samples") && v >= (1 << 20) && v <= (double)(1LL << 30))
      g_tune.chunk_samples = (long long)v;
   else if (!strcmp(key, "de_halo") && (v == 0 || v == 1)) g_tune.de_halo = (int)v;
   else return 1;
   return 0;
}

static void tune_load() {
   const char *path = getenv("FL3H_TUNING");
   char buf[1024];
   if (!path) {
      Dl_info di;
      if (!dladdr((void *)&tune_set, &di) || !di.dli_fname) return;
      snprintf(buf, sizeof(buf), "%s", di.dli_fname);
      char *slash = strrchr(buf, '/');
      if (!slash) return;
      snprintf(slash + 1, sizeof(buf) - (size_t)(slash + 1 - buf), "tuning.conf");
      path = buf;
   }
   FILE *fp = fopen(path, "r");
   if (!fp) return;
   char line[256], key[64];
   double v;
   while (fgets(line, sizeof(line), fp))
      if (line[0] != '#' && sscanf(line, "%63s %lf", key, &v) == 2 && tune_set(key, v))
         fprintf(stderr, "flam3_amd: %s: ignoring tuning entry %s %g\n", path, key, v);
   fclose(fp);
}

static void tune_init() { std::call_once(g_tune_once, tune_load); }

// Set (or, with value NULL, read) one tuning parameter for the rest of the
// process; returns 0, or 1 for an unknown key or a value out of range.
extern "C" int fl3h_tuning(const char *key, const double *value, double *out) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   tune_init();
   if (!key) return 1;
   if (value && tune_set(key, *value)) return 1;
   double v;
   if (!strcmp(key, "wg_per_cu")) v = g_tune.wg_per_cu;
   else if (!strcmp(key, "wg_per_cu_single")) v = g_tune.wg_per_cu_single;
   else if (!strcmp(key, "min_samples")) v = (double)g_tune.min_samples;
   else if (!strcmp(key, "min_samples_small")) v = (double)g_tune.min_samples_small;
   else if (!strcmp(key, "small_samples")) v = (double)g_tune.small_samples;
   else if (!strcmp(key, "chunk_samples")) v = (double)g_tune.chunk_samples;
   else if (!strcmp(key, "de_halo")) v = g_tune.de_halo;
   else return 1;
   if (out) *out = v;
   return 0;
}

// Walkers of a launch: enough for every CU to hold its tuned number of
// workgroups (fewer when the launch runs in several chunks, whose sort
// passes share the CUs with the next chunk's walkers), at least min_samples
// samples each (min_samples_small for launches below small_samples), a
// multiple of 256.
static long long choose_walkers(long long nsamples, int requested) {
   if (requested > 0) return requested;
   tune_init();
   const int wg = nsamples > g_tune.chunk_samples ? g_tune.wg_per_cu : g_tune.wg_per_cu_single;
   long long maxw = (long long)g_ctx.ncu * wg * 256;
   long long w = nsamples / (nsamples < g_tune.small_samples ? g_tune.min_samples_small : g_tune.min_samples);
   if (w > maxw) w = maxw;
   if (w < 256) w = 256;
   return (w + 255) / 256 * 256;
}

// Walkers of one job among njobs that share nsamples, as fl3d_iterate_batch
// splits them: a multiple of 256.
static long long job_walkers(long long nsamples, int njobs, int requested) {
   long long W = choose_walkers(nsamples * njobs, requested);
   long long Wj = (W / njobs) / 256 * 256;
   return Wj < 256 ? 256 : Wj;
}


// Smallest walker-kernel tier holding every variation the genome uses.
static int pick_tier(const fl3_job_hdr *h) {
   int lo = FL3_NVAR, hi = -1;
   for (int u = 0; u < h->nunion; u++) {
      const int v = h->union_ids[u];
      if (v >= FL3_CLASSIC) {
         lo = v < lo ? v : lo;
         hi = v > hi ? v : hi;
      }
   }
   if (hi < 0) return 0;
   for (int t = 1; t <= FL3_NBLOCKS; t++)
      if (lo >= fl3_tier_lo(t) && hi < fl3_tier_hi(t)) return t;
   return FL3_TIER_FULL;
}

// Record, per-workgroup count and tile-total buffers of the chunks of a
// binned deposit (as many sets as chunks, at most FL3_CHUNK_BUFS).
static int ensure_chunk_bufs(long long nchunks, size_t rec_bytes, size_t blk_bytes, size_t tot_bytes) {
   const int nb = nchunks < FL3_CHUNK_BUFS ? (nchunks > 0 ? (int)nchunks : 1) : FL3_CHUNK_BUFS;
   for (int i = 0; i < nb; i++)
      if (g_ctx.crec[i].ensure(rec_bytes) || g_ctx.cblk[i].ensure(blk_bytes) || g_ctx.ctot[i].ensure(tot_bytes))
         return 1;
   return 0;
}

// Tile shift of the binned deposit for the current histogram (cells per
// tile = 1 << shift), or 0 when the histogram is too large for it.
static int bin_shift_cells(long long ncells);
static int bin_shift_cells(long long ncells) {
   // 4096-cell tiles when LDS allows (fewer, larger scatter runs), else 2048
   for (int sh = 12; sh >= 11; sh--)
      if (((ncells + (1LL << sh) - 1) >> sh) <= FL3_MAX_BINS && fl3k_bin_lds_bytes(sh) <= g_ctx.lds_max) return sh;
   return 0;
}

// Tile shift of the accumulation into double buckets (bits 64: five doubles
// per cell in LDS), 0 if LDS is too small.
static int f64_tile_shift() {
   for (int sh = 11; sh >= 10; sh--)
      if (fl3k_bin_f64_lds_bytes(sh) <= g_ctx.lds_max) return sh;
   return 0;
}

// Accumulation tile shift for a histogram of ncells cells and the bucket
// type; *band: 0 if its tiles fit the tile limit, else the shift of a band
// (FL3_MAX_BINS tiles), the unit the records are first grouped by.
static int acc_tile_shift(long long ncells, int bits, int *band) {
   int sh = 0;
   if (bits == 64)
      sh = f64_tile_shift();
   else
      for (int t = 12; t >= 11 && !sh; t--)
         if (fl3k_bin_lds_bytes(t) <= g_ctx.lds_max) sh = t;
   *band = 0;
   if (!sh) return 0;
   if (((ncells + (1LL << sh) - 1) >> sh) > FL3_MAX_BINS) {
      const int bs = sh + 13;   // FL3_MAX_BINS = 2^13 tiles per band
      // bands are the first level's super-tiles: at most 256 of them, and
      // 32-bit record cells
      if (((ncells + (1LL << bs) - 1) >> bs) > 256 || ncells >= (1LL << 31)) return 0;
      *band = bs;
   }
   return sh;
}

// 4-byte records (cell | palette index << 24) suffice when every plotted
// point has visibility 1 and the palette is stepped: the deposit is then a
// function of the cell and the palette index alone.
static bool narrow_records_ok(const fl3_job_hdr *h, long long ncells) {
   if (h->palette_mode != 0 || (ncells >= 0 && ncells >= (1LL << 24) - 1)) return false;
   const fl3_xform_pack *px = (const fl3_xform_pack *)((const unsigned char *)h + h->off_xf);
   for (int e = 0; e < h->nstd; e++)
      if (px[e].vis != 1.0) return false;
   return true;
}

// Job records take the packed accumulation when every job has a stepped
// palette and full visibility (a record then adds its job's palette entry
// and 255 hits, whatever its xform).
static bool job_records_packed(const fl3_job_hdr *const *h, int n) {
   for (int j = 0; j < n; j++) {
      if (h[j]->palette_mode != 0) return false;
      const fl3_xform_pack *px = (const fl3_xform_pack *)((const unsigned char *)h[j] + h[j]->off_xf);
      for (int e = 0; e < h[j]->nxf; e++)
         if (px[e].vis != 1.0) return false;
   }
   return true;
}

// The jobs of one launch (one per temporal sample of a batch) as the
// device sees them.
struct JobSet {
   int n = 1;
   const fl3_job_hdr *const *h = nullptr;  // host headers
   const unsigned char *dev_one = nullptr; // n == 1: the job
   const unsigned char *dev_all = nullptr; // n > 1: concatenated jobs
   const long long *dev_off = nullptr;     // n > 1: byte offsets
   const unsigned long long *dev_keys = nullptr;
   const long long *dev_ns = nullptr;      // n > 1 with counts of their own: samples per job
   size_t lds = 0;                         // largest genome LDS image
   size_t lds_rec = 0;                     // the same without the palette (record modes)
   int tier = 0, nxf_max = 0;
   bool same_tables = true;                // one palette and visibility set for all jobs
   bool linear = false;
};

static bool same_tables(const fl3_job_hdr *a, const fl3_job_hdr *b) {
   if (a->nxf != b->nxf || a->nstd != b->nstd || a->palette_mode != b->palette_mode) return false;
   if (memcmp((const unsigned char *)a + a->off_dmap, (const unsigned char *)b + b->off_dmap, sizeof(double) * 1024))
      return false;
   const fl3_xform_pack *pa = (const fl3_xform_pack *)((const unsigned char *)a + a->off_xf);
   const fl3_xform_pack *pb = (const fl3_xform_pack *)((const unsigned char *)b + b->off_xf);
   for (int e = 0; e < a->nxf; e++)
      if (pa[e].vis != pb[e].vis) return false;
   return true;
}

// Two jobs whose 4-byte records deposit alike: stepped palettes and
// visibility 1 for every xform that plots, and palette entries that
// truncate to the same integers (a 4-byte record adds its entry's integers
// and 255 hits; interpolated palettes of one animation differ only in bits
// the truncation drops).
static bool same_narrow_tables(const fl3_job_hdr *a, const fl3_job_hdr *b) {
   if (!narrow_records_ok(a, 0) || !narrow_records_ok(b, 0)) return false;
   const double *da = (const double *)((const unsigned char *)a + a->off_dmap);
   const double *db = (const double *)((const unsigned char *)b + b->off_dmap);
   for (int e = 0; e < 1024; e++)
      if ((uint32_t)da[e] != (uint32_t)db[e]) return false;
   return true;
}

extern "C" int fl3d_jobs_share_palette(const void *job0, const void *job) {
   return same_narrow_tables((const fl3_job_hdr *)job0, (const fl3_job_hdr *)job);
}

// Palette (as integers and floats) and visibilities of every job of a
// launch, for the accumulation of job records; *pal8 (if not null) gets
// the palettes as bytes R | G << 8 | B << 16 | A << 24 per entry, for
// records that carry the colour, when every entry is below 256 (else null).
static int upload_job_tables(const fl3_job_hdr *const *h, int n, int nxf_max, BinArgs *b,
                             const uint32_t **pal8 = nullptr) {
   const int vs = nxf_max > 0 ? nxf_max : 1;
   std::vector<uint32_t> du((size_t)n * 1024), pb((size_t)n * 256);
   std::vector<float> df((size_t)n * 1024), vis((size_t)n * vs, 1.0f);
   bool bytes = true;
   for (int j = 0; j < n; j++) {
      const unsigned char *hb = (const unsigned char *)h[j];
      const double *dm = (const double *)(hb + h[j]->off_dmap);
      const fl3_xform_pack *px = (const fl3_xform_pack *)(hb + h[j]->off_xf);
      for (int e = 0; e < 1024; e++) {
         du[(size_t)j * 1024 + e] = (uint32_t)dm[e];
         df[(size_t)j * 1024 + e] = (float)dm[e];
         bytes = bytes && dm[e] < 256.0;
      }
      for (int e = 0; e < 256; e++) {
         const uint32_t *c = &du[(size_t)j * 1024 + 4 * e];
         pb[(size_t)j * 256 + e] = (c[0] & 0xffu) | (c[1] & 0xffu) << 8 | (c[2] & 0xffu) << 16 | (c[3] & 0xffu) << 24;
      }
      for (int e = 0; e < h[j]->nxf; e++) vis[(size_t)j * vs + e] = (float)px[e].vis;
   }
   const size_t nb = du.size() * 4, fb = df.size() * 4, vb = vis.size() * 4, pbb = pb.size() * 4;
   if (g_ctx.jobtab.ensure(nb + fb + vb + pbb)) return 1;
   unsigned char *t = (unsigned char *)g_ctx.jobtab.p;
   HIPCHK(hipMemcpyAsync(t, du.data(), nb, hipMemcpyHostToDevice, g_ctx.stream));
   HIPCHK(hipMemcpyAsync(t + nb, df.data(), fb, hipMemcpyHostToDevice, g_ctx.stream));
   HIPCHK(hipMemcpyAsync(t + nb + fb, vis.data(), vb, hipMemcpyHostToDevice, g_ctx.stream));
   HIPCHK(hipMemcpyAsync(t + nb + fb + vb, pb.data(), pbb, hipMemcpyHostToDevice, g_ctx.stream));
   HIPCHK(hipStreamSynchronize(g_ctx.stream));  // the host vectors go out of scope
   b->dmapu_all = (const uint32_t *)t;
   b->dmapf_all = (const float *)(t + nb);
   b->vis_all = (const float *)(t + nb + fb);
   b->vis_stride = vs;
   if (pal8) *pal8 = bytes ? (const uint32_t *)(t + nb + fb + vb) : nullptr;
   return 0;
}

// Job records carry the colour (and take the packed accumulation) when
// every job that writes them has full visibility and a stepped palette of
// entries below 256, the tiles hold 2048 or 4096 cells and the walker's LDS
// has room for the palette after its tile counts.
static bool colour_records(const fl3_job_hdr *const *h, int n, const uint32_t *pal8, int shift, size_t lds_walker) {
   return pal8 && (shift == 11 || shift == 12) && job_records_packed(h, n) &&
          ((lds_walker + 15) & ~(size_t)15) + 1024 <= g_ctx.lds_max;
}

// Records per LDS round of the packed 4-byte-record accumulation: the
// count (16 bits) and the R and G sums (24 bits) must not overflow; a
// multiple of the records one pass of the workgroup loads; 0 if the palette
// is too bright for a useful round.
static int narrow_round(const fl3_job_hdr *const *h, int n) {
   double maxd = 255.0;
   for (int j = 0; j < n; j++) {
      const double *dm = (const double *)((const unsigned char *)h[j] + h[j]->off_dmap);
      for (int e = 0; e < 1024; e++) maxd = dm[e] > maxd ? dm[e] : maxd;
   }
   const long long pass = fl3k_bin_pass_records();
   long long lim = (long long)(16777215.0 / (maxd + 1.0));
   if (lim > 65535) lim = 65535;
   return (int)(lim / pass * pass);
}

// Largest records per accumulation work item: a cell's sums must fit 32
// bits (and the packed pairs must not carry): at most 2^18 records, fewer
// if a palette entry exceeds 255.  Double buckets (fixed64) sum multiples of
// 2^-32 in signed 64-bit integers: records x largest |entry| < 2^31.
static int item_records(const fl3_job_hdr *const *h, int n, int fixed64 = 0) {
   double maxd = 255.0;
   for (int j = 0; j < n; j++) {
      const double *dm = (const double *)((const unsigned char *)h[j] + h[j]->off_dmap);
      for (int e = 0; e < 1024; e++) {
         const double v = fixed64 ? fabs(dm[e]) : dm[e];
         maxd = v > maxd ? v : maxd;
      }
   }
   const double lim = (fixed64 ? 2147483647.0 : 4294967295.0) / (maxd + 1.0);
   const int c = lim < (double)FL3_BIN_ITEM_RECORDS ? (int)lim : FL3_BIN_ITEM_RECORDS;
   return c < 1 ? 1 : c;
}

// Genomes whose walkers draw the xform per wavefront: no xaos, and room in
// LDS for the exchange buffers beside the largest genome image, tile counts
// and palette bytes any launch of them can need (so that the choice does
// not depend on what else shares a launch).
static bool wave_ok(const fl3_job_hdr *h) {
   return !h->chaos_enable && ((lds_bytes_for<float>(h, true) + 15) & ~(size_t)15) + 4 * (size_t)FL3_MAX_BINS + 16 +
                                      1024 + 16 + (2 * 3 * 256 * 4 + 256 * 2) <=
                                  g_ctx.lds_max;
}

static bool jobs_wave_ok(const fl3_job_hdr *const *h, int n) {
   for (int j = 0; j < n; j++)
      if (!wave_ok(h[j])) return false;
   return true;
}

// Walker LDS with the point-exchange buffers of the one-xform-per-wavefront
// mode appended (two 3 x 256 floats and the 256-entry xform-selection
// table, at a 16-byte boundary); the mode stays off when they would not
// fit.
#define FL3_SHUF_BYTES (2 * 3 * 256 * 4 + 256 * 2)
static size_t with_shuffle(IterArgs &a, size_t lds, bool want) {
   const size_t off = (lds + 15) & ~(size_t)15;
   if (want && off + FL3_SHUF_BYTES <= g_ctx.lds_max) {
      a.wave_xf = 1;
      a.shuf_off = (unsigned)off;
      return off + FL3_SHUF_BYTES;
   }
   a.wave_xf = 0;
   a.shuf_off = 0;
   return lds;
}

// Binned deposit: the walkers write records in chunks of `per` rows (the
// walker state carries over between chunks, so the trajectories are those of
// a single launch); each chunk is tile-sorted and accumulated.
// One chunk of a banded histogram on stream bs: the first level (scans of
// the walkers' counts, records grouped by band into b1.sorted), then per
// band with records: a counting pass, the two sort passes (through the
// walker buffer) back into the band's range and the accumulation.  The
// host reads the band boundaries, so it waits for this chunk's walkers.
static int bin_banded(const BinArgs &b1, int format, int band, int shift, hipStream_t bs) {
   // double buckets of 2^24-cell bands, visibility 1, stepped palette: the
   // band pass writes 4-byte records (cell within the band | palette index)
   const int narrow = format == FL3_BIN_F64_NARROW;
   if (fl3k_launch_bin_split(&b1, narrow, bs)) return 1;
   HIPCHK(hipMemcpyAsync(g_ctx.band_host, b1.bin_start, sizeof(unsigned) * ((size_t)b1.nbins + 1),
                         hipMemcpyDeviceToHost, bs));
   HIPCHK(hipEventRecord(g_ctx.ev_band, bs));
   HIPCHK(hipEventSynchronize(g_ctx.ev_band));
   const long long ncells = (long long)g_ctx.width * g_ctx.height;
   const int rows2 = 64;   // second-level blocks: 64 x 256 records
   const size_t rsz = narrow ? 4 : 8;
   const unsigned *st = g_ctx.band_host;
   for (int j = 0; j < b1.nsuper; j++) {
      const int t0 = j << FL3_SUPER_SHIFT, t1 = std::min(b1.nbins, t0 + (1 << FL3_SUPER_SHIFT));
      const long long n = (long long)st[t1] - st[t0];
      if (n <= 0) continue;
      BinArgs b = b1;
      b.cell_base = (unsigned)((long long)j << band);
      b.rec_rel = narrow;
      b.rec = (const unsigned char *)b1.sorted + rsz * st[t0];
      b.rec_count = n;
      b.rows = rows2;
      b.nblk = (int)((n + 256LL * rows2 - 1) / (256LL * rows2));
      b.walkers = 256LL * b.nblk;
      b.shift = shift;
      const long long cells = std::min(ncells - (long long)b.cell_base, 1LL << band);
      b.nbins = (int)((cells + (1LL << shift) - 1) >> shift);
      b.nsuper = (b.nbins + (1 << FL3_SUPER_SHIFT) - 1) >> FL3_SUPER_SHIFT;
      if (g_ctx.blk2.ensure(sizeof(unsigned) * (size_t)b.nblk * b.nsuper) ||
          g_ctx.soff2.ensure(sizeof(unsigned) * (size_t)b.nblk * b.nsuper))
         return 1;
      b.blk_hist = (unsigned *)g_ctx.blk2.p;
      b.soff = (unsigned *)g_ctx.soff2.p;
      // the bands' counters, at fixed places for any tile count (the tile
      // counts start at zero and each scan leaves them there)
      unsigned *bw = (unsigned *)g_ctx.bins2.p;
      b.bin_total = bw;
      b.bin_start = bw + FL3_MAX_BINS;
      b.item_start = bw + 2 * FL3_MAX_BINS + 1;
      b.item_counter = bw + 3 * FL3_MAX_BINS + 2;
      b.cursor = bw + 3 * FL3_MAX_BINS + 4;
      b.sorted = (void *)b1.rec;                                // the walker buffer, free after the band pass
      b.final = (unsigned char *)b1.sorted + rsz * st[t0];      // the band's own range
      if (fl3k_launch_bin_count(&b, narrow, bs) || fl3k_launch_bin(&b, g_ctx.ncu, format, bs)) return 1;
   }
   return 0;
}

// shift: the accumulation tiles' (cells per tile = 1 << shift); band: 0, or
// the band shift of a histogram of more than FL3_MAX_BINS tiles: the
// walkers then count their records per 2^(band - FL3_SUPER_SHIFT)-cell tile,
// the first sort pass groups them by band, and each band is counted, sorted
// and accumulated on its own (iterate_banded_chunk).
static int iterate_binned(IterArgs a, long long W, long long Wj, const JobSet &js, const fl3d_iter_params *p,
                          int shift, int band) {
   const long long ncells = (long long)g_ctx.width * g_ctx.height;
   const int wshift = band ? band - FL3_SUPER_SHIFT : shift;   // the walkers' counting tiles
   const int nbins = (int)((ncells + (1LL << wshift) - 1) >> wshift);
   bool narrow = !band && narrow_records_ok(js.h[0], ncells);
   for (int j = 1; narrow && j < js.n; j++) narrow = same_narrow_tables(js.h[0], js.h[j]);
   // double buckets of a banded histogram: the walkers write 8-byte records,
   // the band pass 4-byte ones (cells within a 2^24-cell band) when the
   // tables allow
   bool band_narrow = band == 24 && g_ctx.bits == 64 && narrow_records_ok(js.h[0], -1);
   for (int j = 1; band_narrow && j < js.n; j++) band_narrow = same_narrow_tables(js.h[0], js.h[j]);
   // one palette for all jobs: narrow or wide records; else the job index
   // rides in the record (step palettes only: it takes the fraction's bits);
   // double buckets take narrow or wide records
   const int format = g_ctx.bits == 64 ? (narrow || band_narrow ? FL3_BIN_F64_NARROW : FL3_BIN_F64_WIDE)
                                       : (narrow ? 1 : (js.same_tables ? 0 : 2));
   long long per = p->chunk > 0 ? (p->chunk + W - 1) / W : (g_tune.chunk_samples / W);
   // banded: chunks of at least 4 samples per cell (up to 2^31 records), as
   // every chunk adds nearly every tile of the histogram once
   if (band && p->chunk <= 0) {
      const long long want = 4 * ncells < (1LL << 28) ? (1LL << 28) : (4 * ncells > (1LL << 31) ? (1LL << 31) : 4 * ncells);
      per = (want + W - 1) / W;
   }
   if (per < 1) per = 1;
   const long long rows_total = (p->nsamples + Wj - 1) / Wj;
   if (per > rows_total) per = rows_total;
   if ((double)W * per >= 4.0e9) {
      set_err("binned deposit chunk exceeds 2^32 records");
      return 1;
   }
   const long long nblk = W / 256;
   const size_t rec_bytes = (narrow ? 4 : 8) * (size_t)W * per;
   const long long nchunks = (p->nsamples + Wj * per - 1) / (Wj * per);
   if (g_ctx.sorted.ensure(rec_bytes)) return 1;
   if (g_ctx.wstate.ensure(sizeof(uint32_t) * 10 * (size_t)W)) return 1;
   const size_t nsup = (size_t)((nbins + (1 << FL3_SUPER_SHIFT) - 1) >> FL3_SUPER_SHIFT);
   if (ensure_chunk_bufs(nchunks, rec_bytes, sizeof(unsigned) * (size_t)nblk * nsup, sizeof(unsigned) * (size_t)nbins))
      return 1;
   const int nsuper = (nbins + (1 << FL3_SUPER_SHIFT) - 1) >> FL3_SUPER_SHIFT;
   if (g_ctx.bins.ensure(sizeof(unsigned) * (4 * (size_t)nbins + 8))) return 1;
   if (band) {
      if (g_ctx.bins2.ensure(sizeof(unsigned) * (4 * (size_t)FL3_MAX_BINS + 8))) return 1;
      HIPCHK(hipMemsetAsync(g_ctx.bins2.p, 0, sizeof(unsigned) * FL3_MAX_BINS, g_ctx.stream));
   }
   if (g_ctx.soff.ensure(sizeof(unsigned) * (size_t)nblk * nsuper)) return 1;
   unsigned *bw = (unsigned *)g_ctx.bins.p;
   BinArgs b;
   memset(&b, 0, sizeof(b));
   b.job = js.n > 1 ? js.dev_all : js.dev_one;  // palette tables of job 0 (identical for all when narrow)
   b.buckets = a.buckets;
   b.rec = g_ctx.crec[0].p;
   b.sorted = g_ctx.sorted.p;
   b.blk_hist = (unsigned *)g_ctx.cblk[0].p;
   b.nblk = (int)nblk;
   b.walkers = W;
   b.bin_total = bw;
   b.bin_start = bw + nbins;
   b.item_start = bw + 2 * nbins + 1;
   b.item_counter = bw + 3 * nbins + 2;
   b.cursor = bw + 3 * nbins + 4;
   b.soff = (unsigned *)g_ctx.soff.p;
   b.nsuper = nsuper;
   b.shift = wshift;
   b.nbins = nbins;
   b.width = g_ctx.width;
   b.height = g_ctx.height;
   b.chunk = item_records(js.h, js.n, format >= FL3_BIN_F64_WIDE);
   const uint32_t *pal8 = nullptr;
   if (format == 2 && upload_job_tables(js.h, js.n, js.nxf_max, &b, &pal8)) return 1;
   a.rec_mode = narrow ? 2 : (format == 2 ? 3 : 1);
   a.wstate = (uint32_t *)g_ctx.wstate.p;
   a.bin_shift = wshift;
   a.nbins = nbins;
   a.hist_off = (unsigned)((js.lds_rec + 15) & ~(size_t)15);
   size_t lds_rec = a.hist_off + sizeof(unsigned) * (size_t)nbins;
   // job records with the colour in them: the packed accumulation
   const bool colour = format == 2 && colour_records(js.h, js.n, pal8, shift, lds_rec);
   a.pal8 = colour ? pal8 : nullptr;
   a.pal_off = (unsigned)((lds_rec + 15) & ~(size_t)15);
   if (colour) lds_rec = a.pal_off + 1024;
   lds_rec = with_shuffle(a, lds_rec, jobs_wave_ok(js.h, js.n));
   b.round = narrow || colour ? narrow_round(js.h, js.n) : 0;
   // chunk k: walkers on the render stream, then its sort and accumulation
   // on the bin stream, which chunk k + 2's walkers wait for before they
   // overwrite the same record buffer
   hipStream_t ws = g_ctx.stream, bs = g_ctx.bin_stream;
   HIPCHK(hipEventRecord(g_ctx.ev_join, ws));   // earlier work on the histogram
   HIPCHK(hipStreamWaitEvent(bs, g_ctx.ev_join, 0));
   long long k = 0;
   for (long long done = 0; done < p->nsamples; k++) {
      const int buf = (int)(k % FL3_CHUNK_BUFS);
      const long long c = p->nsamples - done < Wj * per ? p->nsamples - done : Wj * per;
      a.nsamples = c;
      a.rows_out = (c + Wj - 1) / Wj;
      a.ns_row0 = done / Wj;   // (done: whole rows of every walker)
      a.resume = done > 0;
      a.rec = (unsigned long long *)g_ctx.crec[buf].p;
      a.blk_hist = (unsigned *)g_ctx.cblk[buf].p;
      a.bin_total = (unsigned *)g_ctx.ctot[buf].p;
      if (k >= FL3_CHUNK_BUFS) HIPCHK(hipStreamWaitEvent(ws, g_ctx.ev_bin[buf], 0));
      HIPCHK(hipMemsetAsync(a.bin_total, 0, sizeof(unsigned) * (size_t)nbins, ws));
      int pk = prof_begin(ST_ITER, ws);
      int lrc = fl3k_launch_iter_fast(&a, (unsigned)nblk, lds_rec, ws, g_ctx.bits, 1, js.tier);
      prof_end(pk, ws);
      if (lrc) {
         set_err("k_iter_fast launch failed");
         return 1;
      }
      HIPCHK(hipEventRecord(g_ctx.ev_walk[buf], ws));
      HIPCHK(hipStreamWaitEvent(bs, g_ctx.ev_walk[buf], 0));
      b.rec = a.rec;
      b.final = a.rec;   // the walker buffer takes the tile-grouped records
      b.blk_hist = a.blk_hist;
      b.bin_total = a.bin_total;
      b.rows = a.rows_out;
      pk = prof_begin(ST_BIN, bs);
      if (band)
         lrc = bin_banded(b, format, band, shift, bs);
      else
         lrc = fl3k_launch_bin(&b, g_ctx.ncu, format, bs);
      prof_end(pk, bs);
      if (lrc) {
         set_err("binned deposit launch failed");
         return 1;
      }
      HIPCHK(hipEventRecord(g_ctx.ev_bin[buf], bs));
      done += c;
   }
   // the render stream continues once every chunk is accumulated
   HIPCHK(hipEventRecord(g_ctx.ev_join, bs));
   HIPCHK(hipStreamWaitEvent(ws, g_ctx.ev_join, 0));
   return 0;
}

// One launch over njobs jobs (the temporal samples of a batch), p->nsamples
// plotted samples each, job j's walkers keyed by keys[2j], keys[2j + 1].
extern "C" int fl3d_iterate_batch(const void *const *jobs, const size_t *job_bytes, int njobs,
                                  const fl3d_iter_params *p, const uint64_t *keys, double *badvals,
                                  int *saturated) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   if (njobs < 1) {
      set_err("fl3d_iterate_batch: no jobs");
      return 1;
   }
   std::vector<const fl3_job_hdr *> hs((size_t)njobs);
   JobSet js;
   js.n = njobs;
   for (int j = 0; j < njobs; j++) {
      hs[j] = (const fl3_job_hdr *)jobs[j];
      const size_t l = lds_bytes_for<float>(hs[j]), lr = lds_bytes_for<float>(hs[j], false);
      js.lds = l > js.lds ? l : js.lds;
      js.lds_rec = lr > js.lds_rec ? lr : js.lds_rec;
      const int t = pick_tier(hs[j]);
      js.tier = t > js.tier ? t : js.tier;
      js.nxf_max = hs[j]->nxf > js.nxf_max ? hs[j]->nxf : js.nxf_max;
      if (hs[j]->palette_mode != 0) js.linear = true;
      if (j > 0 && !same_tables(hs[0], hs[j])) js.same_tables = false;
   }
   js.h = hs.data();
   std::vector<long long> off((size_t)njobs);
   if (njobs == 1) {
      if (upload_job(jobs[0], job_bytes[0])) return 1;
      js.dev_one = (const unsigned char *)g_ctx.job.p;
   } else {
      size_t total = 0;
      for (int j = 0; j < njobs; j++) {
         off[j] = (long long)total;
         total += (job_bytes[j] + 255) & ~(size_t)255;
      }
      // after the jobs: byte offsets, walker keys and (if given) sample counts
      const size_t tab = sizeof(long long) * njobs + sizeof(uint64_t) * 2 * njobs +
                         (p->job_nsamples ? sizeof(long long) * njobs : 0);
      std::vector<unsigned char> host(total + tab);
      for (int j = 0; j < njobs; j++) memcpy(host.data() + off[j], jobs[j], job_bytes[j]);
      memcpy(host.data() + total, off.data(), sizeof(long long) * njobs);
      memcpy(host.data() + total + sizeof(long long) * njobs, keys, sizeof(uint64_t) * 2 * njobs);
      if (p->job_nsamples)
         for (int j = 0; j < njobs; j++) {
            if (p->job_nsamples[j] < 0 || p->job_nsamples[j] > p->nsamples) {
               set_err("fl3d_iterate_batch: a job's sample count exceeds the launch's");
               return 1;
            }
            const long long v = p->job_nsamples[j];
            memcpy(host.data() + total + (sizeof(long long) + 2 * sizeof(uint64_t)) * njobs + sizeof(long long) * j,
                   &v, sizeof(long long));
         }
      if (g_ctx.job.ensure(total + tab)) return 1;
      HIPCHK(hipMemcpyAsync(g_ctx.job.p, host.data(), total + tab, hipMemcpyHostToDevice, g_ctx.stream));
      HIPCHK(hipStreamSynchronize(g_ctx.stream));  // host is a local buffer
      js.dev_all = (const unsigned char *)g_ctx.job.p;
      js.dev_off = (const long long *)(js.dev_all + total);
      js.dev_keys = (const unsigned long long *)(js.dev_all + total + sizeof(long long) * njobs);
      if (p->job_nsamples)
         js.dev_ns = (const long long *)(js.dev_all + total + (sizeof(long long) + 2 * sizeof(uint64_t)) * njobs);
   }
   HIPCHK(hipMemsetAsync(g_ctx.counters, 0, 16, g_ctx.stream));
   if (p->nsamples > 0) {
      IterArgs a;
      memset(&a, 0, sizeof(a));
      a.job = js.dev_one;
      a.jobs = js.dev_all;
      a.job_off = js.dev_off;
      a.keys = js.dev_keys;
      a.njobs = njobs;
      a.job_ns = js.dev_ns;
      a.buckets = cur_buckets();
      a.nsamples = p->nsamples;
      a.fuse = p->fuse;
      a.key0 = keys[0];
      a.key1 = keys[1];
      a.badvals = g_ctx.counters;
      a.sat = (uint32_t *)(g_ctx.counters + 1);
      // walkers: a multiple of 256 per job
      const long long Wj = job_walkers(p->nsamples, njobs, p->nwalkers);
      const long long W = Wj * njobs;
      int band = 0;
      const int shift = acc_tile_shift((long long)g_ctx.width * g_ctx.height, g_ctx.bits, &band);
      // double buckets and banded histograms take one palette and one set
      // of visibilities for all jobs (4- or 8-byte records, no job records)
      const bool binnable = js.nxf_max <= 256 && shift > 0 && !(njobs > 1 && js.linear && !js.same_tables) &&
                            (js.same_tables || (g_ctx.bits != 64 && !band));
      if (p->deposit == 2 && !binnable) {
         set_err("tile-binned deposit needs at most 256 xforms, at most 256 bands of %s%ld tiles and, for "
                 "bits 64, banded histograms or temporal samples with linear palettes, one palette for all",
                 "", (long)FL3_MAX_BINS);
         return 1;
      }
      const bool binned =
          binnable && (p->deposit == 2 || (p->deposit == 0 && p->nsamples * njobs >= FL3_BIN_MIN_SAMPLES));
      if (p->job_nsamples && njobs > 1)
         for (int j = 0; j < njobs; j++) g_prof.samples += p->job_nsamples[j];
      else
         g_prof.samples += p->nsamples * njobs;
      if (binned) {
         if (iterate_binned(a, W, Wj, js, p, shift, band)) return 1;
      } else {
         const int pk = prof_begin(ST_ITER);
         const size_t lw = with_shuffle(a, js.lds, jobs_wave_ok(js.h, js.n));
         const int lrc =
             fl3k_launch_iter_fast(&a, (unsigned)(W / 256), lw, g_ctx.stream, g_ctx.bits, p->exact, js.tier);
         prof_end(pk);
         if (lrc) {
            set_err("k_iter_fast launch failed");
            return 1;
         }
      }
   }
   return read_counters(badvals, saturated);
}

// Callable while another thread runs device work: after the first call it
// only reads the CU count.
extern "C" long long fl3d_job_walkers(int64_t nsamples, int njobs, int requested) {
   if (!g_ready.load(std::memory_order_acquire)) {
      std::lock_guard<std::recursive_mutex> lk(g_mu);
      if (ctx_init()) return -1;
   }
   return job_walkers(nsamples, njobs > 0 ? njobs : 1, requested);
}

// Jobs of several renders (begun with fl3d_begin_many) in one launch.  A
// run of consecutive jobs of one render (equal `run`) is what
// fl3d_iterate_batch would launch for that render alone: it gets that
// launch's kernel tier, so its arithmetic is the same.  Jobs with a stepped
// palette write job records into the renders' joint cell space, which the
// tile-binned deposit sums (the same integers per-sample atomics would add);
// jobs with a linear palette, or every job when the renders' cells exceed
// the tile limits, deposit with per-sample atomics (saturating for
// jobs[j].exact).  Each tier, and each deposit kind, is a launch of its own.
// Launch groups: deposit kind (direct, direct saturating, records) x walker
// mode (one xform per wavefront, per lane) x kernel tier.
enum { MG_T = FL3_TIER_FULL + 1, MG_DIRECT = 0, MG_EXACT = 2 * MG_T, MG_REC = 4 * MG_T, MG_N = 6 * MG_T };

extern "C" int fl3d_iterate_multi(const fl3d_multi_job *jobs, int njobs, int fuse, int exact, int nrenders,
                                  double *badvals, int *saturated) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   if (njobs < 1 || nrenders < 1 || nrenders > (int)g_ctx.rcell0.size()) {
      set_err("fl3d_iterate_multi: %s%ld jobs for the begun renders", "", (long)njobs);
      return 1;
   }
   size_t total = 0, lds = 0, lds_rec = 0;
   long long samples = 0;
   int nxf_max = 0;
   std::vector<const fl3_job_hdr *> hs((size_t)njobs);
   std::vector<int> tier((size_t)njobs), grp((size_t)njobs);
   for (int j = 0; j < njobs; j++) {
      hs[j] = (const fl3_job_hdr *)jobs[j].job;
      nxf_max = hs[j]->nxf > nxf_max ? hs[j]->nxf : nxf_max;
      samples += jobs[j].nsamples;
   }
   for (int j = 0; j < njobs;) {
      int e = j, t = 0;
      while (e < njobs && jobs[e].render == jobs[j].render && jobs[e].run == jobs[j].run) {
         const int te = pick_tier(hs[e]);
         t = te > t ? te : t;
         e++;
      }
      for (int k = j; k < e; k++) tier[k] = t;
      j = e;
   }
   // per render: its walkers draw the xform per wavefront (as fl3h_render
   // would decide for it alone)
   std::vector<char> rwave((size_t)nrenders, 1);
   for (int j = 0; j < njobs; j++)
      if (jobs[j].render >= 0 && jobs[j].render < nrenders && !wave_ok(hs[j])) rwave[(size_t)jobs[j].render] = 0;
   // the joint cell space of the renders, in tiles
   long long tcells = 0;
   for (size_t r = 0; r < g_ctx.rcell0.size(); r++) tcells += (long long)g_ctx.rwidth[r] * g_ctx.rheight[r];
   const int shift = g_ctx.bits != 64 && njobs < 65536 && nxf_max <= 256 ? bin_shift_cells(tcells) : 0;
   const int nbins = shift > 0 ? (int)((tcells + (1LL << shift) - 1) >> shift) : 0;
   bool use_rec = shift > 0 && tcells < 0xffffffffLL;
   std::vector<JobDesc> d((size_t)njobs);
   long long gblocks[MG_N] = {0};
   for (int j = 0; j < njobs; j++) {
      if (jobs[j].walkers < 256 || jobs[j].walkers % 256 || jobs[j].render < 0 || jobs[j].render >= nrenders) {
         set_err("fl3d_iterate_multi: job %s%ld needs a multiple of 256 walkers and a begun render", "", (long)j);
         return 1;
      }
      const bool rec = use_rec && hs[j]->palette_mode == 0;
      const int sub = tier[j] + (rwave[jobs[j].render] ? 0 : MG_T);
      grp[j] = rec ? MG_REC + sub : (exact || jobs[j].exact ? MG_EXACT : MG_DIRECT) + sub;
      const size_t l = lds_bytes_for<float>(hs[j], !rec);
      if (rec)
         lds_rec = l > lds_rec ? l : lds_rec;
      else
         lds = l > lds ? l : lds;
      memset(&d[j], 0, sizeof(JobDesc));
      d[j].off = (long long)total;
      d[j].nsamples = jobs[j].nsamples;
      d[j].key0 = jobs[j].key[0];
      d[j].key1 = jobs[j].key[1];
      d[j].cell0 = g_ctx.rcell0[jobs[j].render];
      d[j].block0 = (int)gblocks[grp[j]];   // within its group's launch
      d[j].nblocks = jobs[j].walkers / 256;
      d[j].render = jobs[j].render;
      gblocks[grp[j]] += d[j].nblocks;
      total += (jobs[j].bytes + 255) & ~(size_t)255;
   }
   const unsigned hist_off = (unsigned)((lds_rec + 15) & ~(size_t)15);
   if (hist_off + sizeof(unsigned) * (size_t)nbins > g_ctx.lds_max) use_rec = false;
   if (!use_rec)   // every job deposits with atomics
      for (int j = 0; j < njobs; j++)
         if (grp[j] >= MG_REC) {
            gblocks[grp[j]] -= d[j].nblocks;
            grp[j] = (exact || jobs[j].exact ? MG_EXACT : MG_DIRECT) + tier[j] + (rwave[jobs[j].render] ? 0 : MG_T);
            d[j].block0 = (int)gblocks[grp[j]];
            gblocks[grp[j]] += d[j].nblocks;
            const size_t l = lds_bytes_for<float>(hs[j]);
            lds = l > lds ? l : lds;
         }
   long long gstart[MG_N + 1] = {0};
   for (int g = 0; g < MG_N; g++) {
      if (gblocks[g] > 0x7fffffffLL) {
         set_err("fl3d_iterate_multi: too many workgroups");
         return 1;
      }
      gstart[g + 1] = gstart[g] + gblocks[g];
   }
   const long long blocks = gstart[MG_N];
   std::vector<int> blk((size_t)blocks);
   for (int j = 0; j < njobs; j++)
      for (int b = 0; b < d[j].nblocks; b++) blk[(size_t)(gstart[grp[j]] + d[j].block0 + b)] = j;
   const size_t dbytes = sizeof(JobDesc) * njobs, bbytes = sizeof(int) * (size_t)blocks;
   std::vector<unsigned char> host(total + dbytes + bbytes);
   for (int j = 0; j < njobs; j++) memcpy(host.data() + d[j].off, jobs[j].job, jobs[j].bytes);
   memcpy(host.data() + total, d.data(), dbytes);
   memcpy(host.data() + total + dbytes, blk.data(), bbytes);
   if (g_ctx.job.ensure(host.size())) return 1;
   HIPCHK(hipMemcpyAsync(g_ctx.job.p, host.data(), host.size(), hipMemcpyHostToDevice, g_ctx.stream));
   HIPCHK(hipMemsetAsync(g_ctx.rcount.p, 0, 16 * (size_t)nrenders, g_ctx.stream));
   IterArgs a;
   memset(&a, 0, sizeof(a));
   a.jobs = (const unsigned char *)g_ctx.job.p;
   a.descs = (const JobDesc *)(a.jobs + total);
   a.buckets = g_ctx.buckets.p;
   a.badvals = (unsigned long long *)g_ctx.rcount.p;
   a.fuse = fuse;
   const int *blk_dev = (const int *)(a.jobs + total + dbytes);
   g_prof.samples += samples;
   const size_t lds_lane = lds;
   const size_t lds_wave = with_shuffle(a, lds, true);
   const bool wave_fits = a.wave_xf != 0;
   for (int g = 0; g < MG_REC; g++) {
      if (!gblocks[g]) continue;
      a.blk_job = blk_dev + gstart[g];
      a.wave_xf = wave_fits && (g / MG_T) % 2 == 0;
      const int pk = prof_begin(ST_ITER);
      const int lrc = fl3k_launch_iter_fast(&a, (unsigned)gblocks[g], a.wave_xf ? lds_wave : lds_lane, g_ctx.stream,
                                            g_ctx.bits, g >= MG_EXACT, g % MG_T);
      prof_end(pk);
      if (lrc) {
         set_err("k_iter_fast launch failed");
         return 1;
      }
   }
   const long long Wrec = 256 * (gstart[MG_N] - gstart[MG_REC]);
   if (Wrec > 0) {
      // job records in chunks of `per` rows, sorted and accumulated on the
      // bin stream while the walkers produce the next chunk
      long long rows_total = 0;
      int jrec = -1;   // the first job that writes records
      for (int j = 0; j < njobs; j++)
         if (grp[j] >= MG_REC) {
            const long long wj = 256LL * d[j].nblocks, r = (d[j].nsamples + wj - 1) / wj;
            rows_total = r > rows_total ? r : rows_total;
            if (jrec < 0) jrec = j;
         }
      // 4-byte records (cell | palette index) when every job that writes
      // records has the first one's palette and visibilities (all 1) and the
      // joint cell space fits 24 bits; else job records
      bool narrow = narrow_records_ok(hs[jrec], tcells);
      for (int j = jrec + 1; narrow && j < njobs; j++)
         if (grp[j] >= MG_REC && !same_narrow_tables(hs[jrec], hs[j])) narrow = false;
      long long per = g_tune.chunk_samples / Wrec;
      if (per < 1) per = 1;
      if (per > rows_total) per = rows_total;
      if ((double)Wrec * per >= 4.0e9) {
         set_err("binned deposit chunk exceeds 2^32 records");
         return 1;
      }
      const long long nblk = Wrec / 256;
      const long long nchunks = (rows_total + per - 1) / per;
      const size_t rec_bytes = (narrow ? 4 : 8) * (size_t)Wrec * per;
      if (g_ctx.sorted.ensure(rec_bytes)) return 1;
      if (g_ctx.wstate.ensure(sizeof(uint32_t) * 10 * (size_t)Wrec)) return 1;
      const size_t nsup = (size_t)((nbins + (1 << FL3_SUPER_SHIFT) - 1) >> FL3_SUPER_SHIFT);
      if (ensure_chunk_bufs(nchunks, rec_bytes, sizeof(unsigned) * (size_t)nblk * nsup,
                            sizeof(unsigned) * (size_t)nbins))
         return 1;
      const int nsuper = (nbins + (1 << FL3_SUPER_SHIFT) - 1) >> FL3_SUPER_SHIFT;
      if (g_ctx.bins.ensure(sizeof(unsigned) * (4 * (size_t)nbins + 8))) return 1;
      if (g_ctx.soff.ensure(sizeof(unsigned) * (size_t)nblk * nsuper)) return 1;
      unsigned *bw = (unsigned *)g_ctx.bins.p;
      BinArgs b;
      memset(&b, 0, sizeof(b));
      b.job = a.jobs + d[jrec].off;   // the palette of 4-byte records
      b.buckets = g_ctx.buckets.p;
      b.sorted = g_ctx.sorted.p;
      b.nblk = (int)nblk;
      b.walkers = Wrec;
      b.bin_total = bw;
      b.bin_start = bw + nbins;
      b.item_start = bw + 2 * nbins + 1;
      b.item_counter = bw + 3 * nbins + 2;
      b.cursor = bw + 3 * nbins + 4;
      b.soff = (unsigned *)g_ctx.soff.p;
      b.nsuper = nsuper;
      b.shift = shift;
      b.nbins = nbins;
      b.width = (int)(tcells < 0x7fffffffLL ? tcells : 0x7fffffffLL);
      b.height = 1;
      b.chunk = item_records(hs.data(), njobs);
      const uint32_t *pal8 = nullptr;
      if (!narrow && upload_job_tables(hs.data(), njobs, nxf_max, &b, &pal8)) return 1;
      IterArgs ar = a;
      ar.rec_mode = narrow ? 2 : 3;   // 4-byte or job records
      ar.wstate = (uint32_t *)g_ctx.wstate.p;
      ar.bin_shift = shift;
      ar.nbins = nbins;
      ar.hist_off = hist_off;
      ar.wtot = Wrec;
      size_t lds_r = hist_off + sizeof(unsigned) * (size_t)nbins;
      {
         // records with the colour in them (the packed accumulation) when
         // every job that writes records allows it
         std::vector<const fl3_job_hdr *> rh;
         for (int j = 0; j < njobs; j++)
            if (grp[j] >= MG_REC) rh.push_back(hs[j]);
         const bool colour = !narrow && colour_records(rh.data(), (int)rh.size(), pal8, shift, lds_r);
         ar.pal8 = colour ? pal8 : nullptr;
         ar.pal_off = (unsigned)((lds_r + 15) & ~(size_t)15);
         if (colour) lds_r = ar.pal_off + 1024;
         b.round = narrow || colour ? narrow_round(rh.data(), (int)rh.size()) : 0;
      }
      const size_t lds_r_lane = lds_r;
      lds_r = with_shuffle(ar, lds_r, true);
      const bool wave_fits_r = ar.wave_xf != 0;
      hipStream_t ws = g_ctx.stream, bs = g_ctx.bin_stream;
      HIPCHK(hipEventRecord(g_ctx.ev_join, ws));
      HIPCHK(hipStreamWaitEvent(bs, g_ctx.ev_join, 0));
      for (long long k = 0; k < nchunks; k++) {
         const int buf = (int)(k % FL3_CHUNK_BUFS);
         ar.row0 = k * per;
         ar.rows_out = rows_total - ar.row0 < per ? rows_total - ar.row0 : per;
         ar.resume = k > 0;
         ar.rec = (unsigned long long *)g_ctx.crec[buf].p;
         ar.blk_hist = (unsigned *)g_ctx.cblk[buf].p;
         ar.bin_total = (unsigned *)g_ctx.ctot[buf].p;
         if (k >= FL3_CHUNK_BUFS) HIPCHK(hipStreamWaitEvent(ws, g_ctx.ev_bin[buf], 0));
         HIPCHK(hipMemsetAsync(ar.bin_total, 0, sizeof(unsigned) * (size_t)nbins, ws));
         for (int g = MG_REC; g < MG_N; g++) {
            if (!gblocks[g]) continue;
            ar.blk_job = blk_dev + gstart[g];
            ar.blk0 = (int)(gstart[g] - gstart[MG_REC]);
            ar.wave_xf = wave_fits_r && ((g - MG_REC) / MG_T) % 2 == 0;
            const int pk = prof_begin(ST_ITER, ws);
            const int lrc = fl3k_launch_iter_fast(&ar, (unsigned)gblocks[g], ar.wave_xf ? lds_r : lds_r_lane, ws,
                                                  g_ctx.bits, 1, (g - MG_REC) % MG_T);
            prof_end(pk, ws);
            if (lrc) {
               set_err("k_iter_fast launch failed");
               return 1;
            }
         }
         HIPCHK(hipEventRecord(g_ctx.ev_walk[buf], ws));
         HIPCHK(hipStreamWaitEvent(bs, g_ctx.ev_walk[buf], 0));
            b.rec = ar.rec;
         b.final = ar.rec;
         b.blk_hist = ar.blk_hist;
         b.bin_total = ar.bin_total;
            b.rows = ar.rows_out;
         const int pk = prof_begin(ST_BIN, bs);
         const int lrc = fl3k_launch_bin(&b, g_ctx.ncu, narrow ? 1 : 2, bs);
         prof_end(pk, bs);
         if (lrc) {
            set_err("binned deposit launch failed");
            return 1;
         }
         HIPCHK(hipEventRecord(g_ctx.ev_bin[buf], bs));
      }
      HIPCHK(hipEventRecord(g_ctx.ev_join, bs));
      HIPCHK(hipStreamWaitEvent(ws, g_ctx.ev_join, 0));
   }
   std::vector<unsigned long long> c(2 * (size_t)nrenders);
   HIPCHK(hipMemcpyAsync(c.data(), g_ctx.rcount.p, 16 * (size_t)nrenders, hipMemcpyDeviceToHost, g_ctx.stream));
   HIPCHK(hipStreamSynchronize(g_ctx.stream));   // also keeps host alive for the upload
   for (int r = 0; r < nrenders; r++) {
      if (badvals) badvals[r] = (double)c[2 * r];
      if (saturated) saturated[r] = (c[2 * r + 1] & 0xffffffffull) != 0;
   }
   return 0;
}

extern "C" int fl3d_iterate(const void *job, size_t job_bytes, const fl3d_iter_params *p, double *badvals,
                            int *saturated) {
   const void *jobs[1] = {job};
   const size_t bytes[1] = {job_bytes};
   const uint64_t keys[2] = {p->key[0], p->key[1]};
   return fl3d_iterate_batch(jobs, bytes, 1, p, keys, badvals, saturated);
}

// Replay launch: isaac holds 36 words per walker (host-seeded); updated.
static int iterate_replay(const void *job, size_t job_bytes, uint32_t *isaac_host, int nwalkers,
                          double batch_size, int sbs, int fuse, double *badvals) {
   const fl3_job_hdr *h = (const fl3_job_hdr *)job;
   if (upload_job(job, job_bytes)) return 1;
   size_t ib = sizeof(uint32_t) * 36 * (size_t)nwalkers;
   if (g_ctx.isaac.ensure(ib)) return 1;
   HIPCHK(hipMemcpyAsync(g_ctx.isaac.p, isaac_host, ib, hipMemcpyHostToDevice, g_ctx.stream));
   HIPCHK(hipMemsetAsync(g_ctx.counters, 0, 16, g_ctx.stream));
   IterArgs a;
   memset(&a, 0, sizeof(a));
   a.job = (const unsigned char *)g_ctx.job.p;
   a.buckets = cur_buckets();
   a.fuse = fuse;
   a.badvals = g_ctx.counters;
   a.sat = (uint32_t *)(g_ctx.counters + 1);
   a.isaac = (uint32_t *)g_ctx.isaac.p;
   a.batch_size = batch_size;
   a.sbs = sbs;
   a.nwalkers = nwalkers;
   const dim3 grid((nwalkers + 63) / 64), block(64);
   const size_t lds = lds_bytes_for<double>(h);
   if (fl3k_launch_replay(&a, grid.x, lds, g_ctx.stream, g_ctx.bits)) {
      set_err("k_iter_replay launch failed");
      return 1;
   }
   HIPCHK(hipMemcpyAsync(isaac_host, g_ctx.isaac.p, ib, hipMemcpyDeviceToHost, g_ctx.stream));
   return read_counters(badvals, nullptr);
}

// The rows of a filter bank cut to the coefficients a tap can reach (row i,
// radius r = ceil(width) - 1 as the gather computes it: its first
// (r + 1)(r + 2) / 2 entries), back to back as floats (rounded to nearest,
// as the device's conversion) into cf, each row's start into off (both
// skipped if cf is null: the count alone).  Returns the float count, or 0
// (the gather then reads the full table) if a row would need more entries
// than the table's rows hold or the cut bank exceeds FL3_DE_LDS_COEFS floats.
static int de_cut_bank(const fl3_de_filters *de, int *off, float *cf) {
   const int nf = de->max_filter_index + 1;
   long n = 0;
   for (int i = 0; i < nf; i++) {
      const long r = (long)ceil(de->filter_widths[i]) - 1;
      const long len = (r + 1) * (r + 2) / 2;
      if (r < 0 || len > de->kernel_size || n + len > FL3_DE_LDS_COEFS) return 0;
      if (cf) {
         off[i] = (int)n;
         for (long t = 0; t < len; t++) cf[n + t] = (float)de->filter_coefs[(size_t)i * de->kernel_size + (size_t)t];
      }
      n += len;
   }
   return (int)n;
}

// A cut bank as the batched tables hold it: row starts, then the floats,
// each part padded to 16 bytes.  Bytes of the row starts, and of the whole
// (0 without a cut bank).
static size_t de_cut_rows_bytes(const fl3_de_filters *de) {
   return (sizeof(int) * (size_t)(de->max_filter_index + 1) + 15) & ~(size_t)15;
}
static size_t de_cut_bytes(const fl3_de_filters *de, int ncut) {
   if (ncut <= 0) return 0;
   return de_cut_rows_bytes(de) + ((sizeof(float) * (size_t)ncut + 15) & ~(size_t)15);
}

// DE tables on the device: coefficients (double), widths (double) and, for
// the float accumulator, a float copy of the coefficients and the cut bank
// (de_cut_bank; *ncoef 0 if there is none).
static int upload_de(const fl3_de_filters *de, double **coefs, double **widths, float **coefs_f,
                     const int **coff = nullptr, const float **ccoef = nullptr, int *ncoef = nullptr) {
   size_t nc = (size_t)de->nfilters * de->kernel_size, nw = de->nfilters;
   const int ncut = de_cut_bank(de, nullptr, nullptr);
   std::vector<int> off(ncut ? (size_t)de->max_filter_index + 1 : 0);
   std::vector<float> cf((size_t)ncut);
   if (ncut) de_cut_bank(de, off.data(), cf.data());
   const size_t cut0 = (sizeof(double) * (nc + nw) + sizeof(float) * nc + 15) & ~(size_t)15;
   const size_t cut1 = cut0 + sizeof(int) * off.size();
   const void *p0 = g_ctx.table.p;
   const size_t cap0 = g_ctx.table.cap;
   if (g_ctx.table.ensure(cut1 + sizeof(float) * (size_t)ncut)) {
      g_ctx.de_up_p = nullptr;
      return 1;
   }
   if (g_ctx.table.p != p0 || g_ctx.table.cap != cap0) g_ctx.de_up_p = nullptr;   // (a new allocation)
   *coefs = (double *)g_ctx.table.p;
   *widths = *coefs + nc;
   *coefs_f = (float *)(*widths + nw);
   if (ncoef) {
      *coff = (const int *)((unsigned char *)g_ctx.table.p + cut0);
      *ccoef = (const float *)((unsigned char *)g_ctx.table.p + cut1);
      *ncoef = ncut;
   }
   // the same tables as the last upload (equal estimator parameters, as in
   // every frame of a still or an animation): they are on the device
   // already, so no copies, conversion or wait
   if (g_ctx.de_up_p == g_ctx.table.p && g_ctx.de_up_cap == g_ctx.table.cap && g_ctx.de_up.size() == nc + nw &&
       g_ctx.de_up_nf == de->max_filter_index + 1 && g_ctx.de_up_ks == de->kernel_size &&
       memcmp(g_ctx.de_up.data(), de->filter_coefs, sizeof(double) * nc) == 0 &&
       memcmp(g_ctx.de_up.data() + nc, de->filter_widths, sizeof(double) * nw) == 0)
      return 0;
   g_ctx.de_up_p = nullptr;
   HIPCHK(hipMemcpyAsync(*coefs, de->filter_coefs, sizeof(double) * nc, hipMemcpyHostToDevice, g_ctx.stream));
   HIPCHK(hipMemcpyAsync(*widths, de->filter_widths, sizeof(double) * nw, hipMemcpyHostToDevice, g_ctx.stream));
   if (ncut) {
      HIPCHK(hipMemcpyAsync((unsigned char *)g_ctx.table.p + cut0, off.data(), sizeof(int) * off.size(),
                            hipMemcpyHostToDevice, g_ctx.stream));
      HIPCHK(hipMemcpyAsync((unsigned char *)g_ctx.table.p + cut1, cf.data(), sizeof(float) * (size_t)ncut,
                            hipMemcpyHostToDevice, g_ctx.stream));
   }
   if (fl3k_launch_f64_to_f32(*coefs, *coefs_f, (long)nc, g_ctx.stream)) {
      set_err("k_f64_to_f32 launch failed");
      return 1;
   }
   // the caller may free its host tables once this returns
   HIPCHK(hipStreamSynchronize(g_ctx.stream));
   g_ctx.de_up.assign(de->filter_coefs, de->filter_coefs + nc);
   g_ctx.de_up.insert(g_ctx.de_up.end(), de->filter_widths, de->filter_widths + nw);
   g_ctx.de_up_p = g_ctx.table.p;
   g_ctx.de_up_cap = g_ctx.table.cap;
   g_ctx.de_up_nf = de->max_filter_index + 1;
   g_ctx.de_up_ks = de->kernel_size;
   return 0;
}

extern "C" int fl3d_density(const fl3d_de_params *p, const fl3_de_filters *de) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   DEArgs a;
   memset(&a, 0, sizeof(a));
   double *coefs, *widths;
   float *coefs_f;
   const int *coff;
   const float *ccoef;
   int ncoef;
   if (upload_de(de, &coefs, &widths, &coefs_f, &coff, &ccoef, &ncoef)) return 1;
   a.buckets = cur_buckets();
   a.accum = cur_accum();
   a.coefs = coefs;
   a.widths = widths;
   a.coefs_t = g_ctx.bits == 64 ? (const void *)coefs : (const void *)coefs_f;
   if (g_ctx.bits != 64) {
      a.coff = coff;
      a.ccoef = ccoef;
      a.ncoef = ncoef;
   }
   a.width = g_ctx.width;
   a.height = g_ctx.height;
   a.oversample = p->oversample;
   a.max_filtered_counts = de->max_filtered_counts;
   a.max_filter_index = de->max_filter_index;
   a.kernel_size = de->kernel_size;
   a.k1 = p->k1;
   a.k2 = p->k2;
   a.curve = p->curve;
   int half = 0;
   while ((half + 1) * (half + 2) / 2 < de->kernel_size) half++;
   a.half = half;
   if (de->max_filter_index >= 1024) {
      set_err("density estimation: more than 1024 kernels");
      return 1;
   }
   if (g_ctx.width >= 65536 || g_ctx.height >= 65536) {
      set_err("density estimation: histogram side above 65535 cells");
      return 1;
   }
   if (g_ctx.findex.ensure(sizeof(uint16_t) * (size_t)g_ctx.width * g_ctx.height)) return 1;
   a.findex = (uint16_t *)g_ctx.findex.p;
   if (g_ctx.srcv.ensure(accum_bytes(g_ctx.bits) * (size_t)g_ctx.width * g_ctx.height)) return 1;
   if (g_ctx.srcl.ensure(sizeof(double) * (size_t)g_ctx.width * g_ctx.height)) return 1;
   a.srcv = g_ctx.srcv.p;
   a.srcl = (double *)g_ctx.srcl.p;
   a.exact = p->exact || g_ctx.bits == 64;   // (the FP64 gather is the exact one)
   a.taps = de_taps();   // gather footprint, for the stage profile
   tune_init();
   if (g_tune.de_halo) {
      a.nbx = (g_ctx.width + DG_REGION_HOST - 1) / DG_REGION_HOST;
      const size_t nblk = (size_t)a.nbx * ((g_ctx.height + DG_REGION_HOST - 1) / DG_REGION_HOST);
      if (g_ctx.deblk.ensure(sizeof(int) * nblk)) return 1;
      a.blkmax = (int *)g_ctx.deblk.p;
   }
   const int pk = prof_begin(ST_DE);
   const int lrc = fl3k_launch_de_tiled(&a, g_ctx.stream, g_ctx.bits);
   prof_end(pk);
   g_prof.cells = (long long)g_ctx.width * g_ctx.height;
   if (lrc) {
      set_err("density estimation launch failed");
      return 1;
   }
   return 0;
}

static int de_many_batch(const fl3d_de_params *p, const fl3_de_filters *const *de, const std::vector<int> &dej,
                         size_t tab, long long tcells);

// Density estimation of renders 0..n-1 of the begun histograms in one
// launch pair (FP32 accumulators; renders without DE filters, or with bits
// 64 / exact sums, take the per-render calls).  de[i] NULL or
// max_filter_index 0: log scaling with p[i].k1, p[i].k2.
extern "C" int fl3d_density_many(const fl3d_de_params *p, const fl3_de_filters *const *de, int n) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   if (n <= 0) return 0;
   if (n > (int)g_ctx.rcell0.size()) {
      set_err("fl3d_density_many: more renders than begun");
      return 1;
   }
   const bool batched = g_ctx.bits != 64;
   std::vector<int> dej;
   size_t tab = 0;
   long long tcells = 0;
   for (int i = 0; i < n; i++) {
      tcells += (long long)g_ctx.rwidth[i] * g_ctx.rheight[i];
      const bool has = de[i] && de[i]->max_filter_index > 0;
      if (has && batched && !p[i].exact && de[i]->max_filter_index < 1024 && g_ctx.rwidth[i] < 65536 &&
          g_ctx.rheight[i] < 65536) {
         dej.push_back(i);
         const size_t nc = (size_t)de[i]->nfilters * de[i]->kernel_size, nw = de[i]->nfilters;
         tab += (nc + nw) * sizeof(double) + ((nc * sizeof(float) + 15) & ~(size_t)15) +
                de_cut_bytes(de[i], de_cut_bank(de[i], nullptr, nullptr));
      }
   }
   const int nd = (int)dej.size();
   if (nd > 0 && de_many_batch(p, de, dej, tab, tcells)) return 1;
   // then the renders outside the batch, one by one
   for (int i = 0; i < n; i++) {
      const bool has = de[i] && de[i]->max_filter_index > 0;
      if (std::find(dej.begin(), dej.end(), i) != dej.end()) continue;
      if (fl3d_select(i)) return 1;
      if (has) {
         if (fl3d_density(&p[i], de[i])) return 1;
      } else if (fl3d_logscale(p[i].k1, p[i].k2))
         return 1;
   }
   return 0;
}

// The batched part of fl3d_density_many: renders dej of the begun
// histograms, tab bytes of coefficient tables, tcells cells in all.
static int de_many_batch(const fl3d_de_params *p, const fl3_de_filters *const *de, const std::vector<int> &dej,
                         size_t tab, long long tcells) {
   const int nd = (int)dej.size();
   // coefficient tables (double, then the float copy) of every render
   if (g_ctx.detab.ensure(tab)) return 1;
   if (g_ctx.findex.ensure(sizeof(uint16_t) * (size_t)tcells) || g_ctx.srcv.ensure(16 * (size_t)tcells) ||
       g_ctx.srcl.ensure(sizeof(double) * (size_t)tcells))
      return 1;
   std::vector<DEArgs> as((size_t)nd);
   std::vector<unsigned> first(2 * (size_t)(nd + 1));
   // per-block source radii for every render's gather halos
   long long nblk = 0;
   for (int k = 0; k < nd; k++)
      nblk += (long long)((g_ctx.rwidth[dej[k]] + DG_REGION_HOST - 1) / DG_REGION_HOST) *
              ((g_ctx.rheight[dej[k]] + DG_REGION_HOST - 1) / DG_REGION_HOST);
   tune_init();
   const bool halo = g_tune.de_halo != 0;
   if (halo && g_ctx.deblk.ensure(sizeof(int) * (size_t)nblk)) return 1;
   long long blk0 = 0;
   unsigned *pf = first.data(), *gf = first.data() + nd + 1;
   unsigned long long pb = 0, gb = 0;
   std::vector<unsigned char> hostt(tab);
   size_t off = 0;
   // renders with equal filter banks (the same estimator parameters and
   // oversampling) share one uploaded table
   // (with its cut bank after the float copy, if it has one)
   struct Uniq {
      const fl3_de_filters *d;
      size_t off;
      int ncut;
   };
   std::vector<Uniq> uniq;
   int ncut_max = 0;
   for (int k = 0; k < nd; k++) {
      const int i = dej[k];
      const fl3_de_filters *d = de[i];
      const size_t nc = (size_t)d->nfilters * d->kernel_size, nw = d->nfilters;
      const size_t base = (nc + nw) * sizeof(double) + ((nc * sizeof(float) + 15) & ~(size_t)15);
      size_t toff = off;
      int ncut = 0;
      bool found = false;
      for (const auto &u : uniq) {
         const fl3_de_filters *e = u.d;
         if (e->nfilters == d->nfilters && e->kernel_size == d->kernel_size &&
             e->max_filter_index == d->max_filter_index &&
             !memcmp(e->filter_coefs, d->filter_coefs, nc * sizeof(double)) &&
             !memcmp(e->filter_widths, d->filter_widths, nw * sizeof(double))) {
            toff = u.off;
            ncut = u.ncut;
            found = true;
            break;
         }
      }
      if (!found) {
         memcpy(hostt.data() + off, d->filter_coefs, nc * sizeof(double));
         memcpy(hostt.data() + off + nc * sizeof(double), d->filter_widths, nw * sizeof(double));
         // the float copy, rounded to nearest as a device conversion would
         float *ft = (float *)(hostt.data() + off + (nc + nw) * sizeof(double));
         for (size_t e = 0; e < nc; e++) ft[e] = (float)d->filter_coefs[e];
         ncut = de_cut_bank(d, nullptr, nullptr);
         if (ncut)
            de_cut_bank(d, (int *)(hostt.data() + off + base),
                        (float *)(hostt.data() + off + base + de_cut_rows_bytes(d)));
         uniq.push_back({d, off, ncut});
         off += base + de_cut_bytes(d, ncut);
      }
      ncut_max = ncut > ncut_max ? ncut : ncut_max;
      unsigned char *tb = (unsigned char *)g_ctx.detab.p + toff;
      DEArgs &a = as[k];
      memset(&a, 0, sizeof(a));
      const long long c0 = g_ctx.rcell0[i];
      a.buckets = (unsigned char *)g_ctx.buckets.p + c0 * bucket_bytes(g_ctx.bits);
      a.accum = (unsigned char *)g_ctx.accum.p + c0 * accum_bytes(g_ctx.bits);
      a.coefs = (const double *)tb;
      a.widths = (const double *)(tb + nc * sizeof(double));
      a.coefs_t = tb + (nc + nw) * sizeof(double);
      if (ncut) {
         a.coff = (const int *)(tb + base);
         a.ccoef = (const float *)(tb + base + de_cut_rows_bytes(d));
         a.ncoef = ncut;
      }
      a.findex = (uint16_t *)g_ctx.findex.p + c0;
      a.srcv = (unsigned char *)g_ctx.srcv.p + c0 * 16;
      a.srcl = (double *)g_ctx.srcl.p + c0;
      a.exact = 0;
      a.taps = de_taps();
      a.width = g_ctx.rwidth[i];
      a.height = g_ctx.rheight[i];
      a.oversample = p[i].oversample;
      a.max_filtered_counts = d->max_filtered_counts;
      a.max_filter_index = d->max_filter_index;
      a.kernel_size = d->kernel_size;
      a.k1 = p[i].k1;
      a.k2 = p[i].k2;
      a.curve = p[i].curve;
      int half = 0;
      while ((half + 1) * (half + 2) / 2 < d->kernel_size) half++;
      a.half = half;
      if (halo) {
         a.nbx = (a.width + DG_REGION_HOST - 1) / DG_REGION_HOST;
         a.blkmax = (int *)g_ctx.deblk.p + blk0;
         blk0 += (long long)a.nbx * ((a.height + DG_REGION_HOST - 1) / DG_REGION_HOST);
      }
      pf[k] = (unsigned)pb;
      gf[k] = (unsigned)gb;
      pb += (unsigned long long)((a.width + 255) / 256) * a.height;
      gb += (unsigned long long)((a.width + DG_REGION_HOST - 1) / DG_REGION_HOST) *
            ((a.height + DG_REGION_HOST - 1) / DG_REGION_HOST);
   }
   if (pb > 0x7fffffffULL || gb > 0x7fffffffULL) {
      set_err("fl3d_density_many: too many workgroups");
      return 1;
   }
   pf[nd] = (unsigned)pb;
   gf[nd] = (unsigned)gb;
   const size_t ab = sizeof(DEArgs) * nd, fb = sizeof(unsigned) * first.size();
   if (g_ctx.deargs.ensure(ab + fb + 16)) return 1;
   HIPCHK(hipMemcpyAsync(g_ctx.detab.p, hostt.data(), off, hipMemcpyHostToDevice, g_ctx.stream));
   HIPCHK(hipMemcpyAsync(g_ctx.deargs.p, as.data(), ab, hipMemcpyHostToDevice, g_ctx.stream));
   HIPCHK(hipMemcpyAsync((unsigned char *)g_ctx.deargs.p + ab, first.data(), fb, hipMemcpyHostToDevice, g_ctx.stream));
   const int pk = prof_begin(ST_DE);
   const unsigned *dfirst = (const unsigned *)((unsigned char *)g_ctx.deargs.p + ab);
   const int lrc = fl3k_launch_de_many((const DEArgs *)g_ctx.deargs.p, dfirst, (unsigned)pb, dfirst + nd + 1,
                                       (unsigned)gb, nd, g_ctx.stream, halo ? (int *)g_ctx.deblk.p : nullptr,
                                       (long)nblk, sizeof(float) * (size_t)ncut_max);
   prof_end(pk);
   if (lrc) {
      set_err("density estimation launch failed");
      return 1;
   }
   // the host vectors go out of scope
   HIPCHK(hipStreamSynchronize(g_ctx.stream));
   return 0;
}

extern "C" int fl3d_logscale(double k1, double k2) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   const long n = (long)g_ctx.width * g_ctx.height;
   const int pk = prof_begin(ST_DE);
   const int lrc = fl3k_launch_logscale(cur_buckets(), cur_accum(), n, k1, k2, g_ctx.stream, g_ctx.bits);
   prof_end(pk);
   if (lrc) {
      set_err("k_logscale launch failed");
      return 1;
   }
   return 0;
}

// Images of one or more renders: each tone map writes its image at an offset
// of the device image buffer; one copy brings them all back.
extern "C" int fl3d_images_begin(size_t image_bytes, size_t filter_bytes) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   if (g_ctx.image.ensure(image_bytes > 0 ? image_bytes : 1)) return 1;
   if (g_ctx.misc.ensure(filter_bytes > 0 ? filter_bytes : 1)) return 1;
   return 0;
}

extern "C" int fl3d_tonemap_async(const fl3d_tm_params *p, const double *filter, size_t img_off, size_t filt_off) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   const size_t row = (size_t)p->image_width * p->nchan * p->bytes_per_channel;
   const size_t img = row * p->image_height;
   const size_t fb = sizeof(double) * p->filter_width * p->filter_width;
   if (img_off + img > g_ctx.image.cap || filt_off + fb > g_ctx.misc.cap) {
      set_err("fl3d_tonemap_async: image or filter outside the buffers of fl3d_images_begin");
      return 1;
   }
   double *dfilt = (double *)((unsigned char *)g_ctx.misc.p + filt_off);
   HIPCHK(hipMemcpyAsync(dfilt, filter, fb, hipMemcpyHostToDevice, g_ctx.stream));
   TMArgs a;
   memset(&a, 0, sizeof(a));
   a.accum = cur_accum();
   a.filter = dfilt;
   a.out = (unsigned char *)g_ctx.image.p + img_off;
   a.W = g_ctx.width;
   a.iw = p->image_width;
   a.ih = p->image_height;
   a.os = p->oversample;
   a.de_offset = p->de_offset;
   a.fw = p->filter_width;
   a.nchan = p->nchan;
   a.transp = p->transp;
   a.bpc = p->bytes_per_channel;
   a.earlyclip = p->earlyclip;
   a.g = p->g;
   a.linrange = p->linrange;
   a.vibrancy = p->vibrancy;
   a.highpow = p->highpow;
   for (int k = 0; k < 3; k++) a.bg[k] = p->background[k];
   const long n = (long)g_ctx.width * g_ctx.height;
   if (p->earlyclip) {
      if (fl3k_launch_earlyclip(cur_accum(), n, &a, g_ctx.stream, g_ctx.bits)) {
         set_err("k_earlyclip launch failed");
         return 1;
      }
   }
   if (p->image_width > 0 && p->image_height > 0) {
      const int pk = prof_begin(ST_TONE);
      const int lrc = fl3k_launch_tonemap(&a, g_ctx.stream, g_ctx.bits);
      prof_end(pk);
      if (lrc) {
         set_err("k_tonemap launch failed");
         return 1;
      }
   }
   return 0;
}

// Tone maps of renders 0..n-1 of the begun histograms into the image buffer
// of fl3d_images_begin (render i at img_off[i]): earlyclip pre-passes per
// render, then one launch for all.
extern "C" int fl3d_tonemap_many(const fl3d_tm_params *p, const double *const *filters, const size_t *img_off,
                                 int n) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   if (n <= 0) return 0;
   if (n > (int)g_ctx.rcell0.size()) {
      set_err("fl3d_tonemap_many: more renders than begun");
      return 1;
   }
   size_t fb = 0;
   for (int i = 0; i < n; i++) fb += sizeof(double) * p[i].filter_width * p[i].filter_width;
   std::vector<TMArgs> as((size_t)n);
   std::vector<unsigned> first((size_t)n + 1);
   std::vector<double> hf(fb / sizeof(double) + 1);
   if (g_ctx.detab.ensure(fb + 16) || g_ctx.deargs.ensure(sizeof(TMArgs) * n + sizeof(unsigned) * (n + 1) + 16))
      return 1;
   size_t fo = 0;
   unsigned long long blocks = 0;
   for (int i = 0; i < n; i++) {
      const size_t row = (size_t)p[i].image_width * p[i].nchan * p[i].bytes_per_channel;
      if (img_off[i] + row * p[i].image_height > g_ctx.image.cap) {
         set_err("fl3d_tonemap_many: image outside the buffer of fl3d_images_begin");
         return 1;
      }
      const size_t nf = (size_t)p[i].filter_width * p[i].filter_width;
      memcpy(hf.data() + fo, filters[i], sizeof(double) * nf);
      TMArgs &a = as[i];
      memset(&a, 0, sizeof(a));
      a.accum = (unsigned char *)g_ctx.accum.p + g_ctx.rcell0[i] * accum_bytes(g_ctx.bits);
      a.filter = (const double *)g_ctx.detab.p + fo;
      a.out = (unsigned char *)g_ctx.image.p + img_off[i];
      a.W = g_ctx.rwidth[i];
      a.iw = p[i].image_width;
      a.ih = p[i].image_height;
      a.os = p[i].oversample;
      a.de_offset = p[i].de_offset;
      a.fw = p[i].filter_width;
      a.nchan = p[i].nchan;
      a.transp = p[i].transp;
      a.bpc = p[i].bytes_per_channel;
      a.earlyclip = p[i].earlyclip;
      a.g = p[i].g;
      a.linrange = p[i].linrange;
      a.vibrancy = p[i].vibrancy;
      a.highpow = p[i].highpow;
      for (int k = 0; k < 3; k++) a.bg[k] = p[i].background[k];
      first[i] = (unsigned)blocks;
      if (a.iw > 0 && a.ih > 0) blocks += (unsigned long long)((a.iw + 255) / 256) * a.ih;
      fo += nf;
   }
   first[n] = (unsigned)blocks;
   if (blocks > 0x7fffffffULL) {
      set_err("fl3d_tonemap_many: too many workgroups");
      return 1;
   }
   HIPCHK(hipMemcpyAsync(g_ctx.detab.p, hf.data(), fb, hipMemcpyHostToDevice, g_ctx.stream));
   HIPCHK(hipMemcpyAsync(g_ctx.deargs.p, as.data(), sizeof(TMArgs) * n, hipMemcpyHostToDevice, g_ctx.stream));
   unsigned *dfirst = (unsigned *)((unsigned char *)g_ctx.deargs.p + sizeof(TMArgs) * n);
   HIPCHK(hipMemcpyAsync(dfirst, first.data(), sizeof(unsigned) * (n + 1), hipMemcpyHostToDevice, g_ctx.stream));
   for (int i = 0; i < n; i++)
      if (as[i].earlyclip &&
          fl3k_launch_earlyclip(const_cast<void *>(as[i].accum), (long)g_ctx.rwidth[i] * g_ctx.rheight[i], &as[i],
                                g_ctx.stream, g_ctx.bits)) {
         set_err("k_earlyclip launch failed");
         return 1;
      }
   const int pk = prof_begin(ST_TONE);
   const int lrc = fl3k_launch_tonemap_many((const TMArgs *)g_ctx.deargs.p, dfirst, (unsigned)blocks, n,
                                            g_ctx.stream, g_ctx.bits);
   prof_end(pk);
   if (lrc) {
      set_err("k_tonemap launch failed");
      return 1;
   }
   HIPCHK(hipStreamSynchronize(g_ctx.stream));   // the host vectors go out of scope
   return 0;
}

extern "C" int fl3d_images_fetch(const fl3d_image_dst *d, int n) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   size_t end = 0;
   for (int i = 0; i < n; i++) {
      const size_t e = d[i].img_off + d[i].row_bytes * (size_t)d[i].rows;
      end = e > end ? e : end;
   }
   if (end == 0) {
      HIPCHK(hipStreamSynchronize(g_ctx.stream));
      return 0;
   }
   // one copy into pinned staging, then rows into the callers' buffers
   // (a strided copy straight to pageable memory goes row by row)
   if (g_ctx.staging_cap < end) {
      // (grown at least twofold: pinned allocations are slow)
      const size_t want = end > 2 * g_ctx.staging_cap ? end : 2 * g_ctx.staging_cap;
      if (g_ctx.staging) (void)hipHostFree(g_ctx.staging);
      g_ctx.staging = nullptr;
      g_ctx.staging_cap = 0;
      HIPCHK(hipHostMalloc(&g_ctx.staging, want, hipHostMallocDefault));
      g_ctx.staging_cap = want;
   }
   if (n == 1 && end >= FL3_STRIP_MIN_BYTES && d[0].rows >= FL3_IMAGE_STRIPS) {
      // One image in strips of rows: the host copies strip k out of staging
      // while the device copies strip k + 1 in.
      const size_t rb = d[0].row_bytes;
      int r0[FL3_IMAGE_STRIPS + 1];
      for (int k = 0; k <= FL3_IMAGE_STRIPS; k++) r0[k] = (int)((long)d[0].rows * k / FL3_IMAGE_STRIPS);
      for (int k = 0; k < FL3_IMAGE_STRIPS; k++) {
         const size_t o = d[0].img_off + rb * (size_t)r0[k];
         HIPCHK(hipMemcpyAsync((unsigned char *)g_ctx.staging + o, (const unsigned char *)g_ctx.image.p + o,
                               rb * (size_t)(r0[k + 1] - r0[k]), hipMemcpyDeviceToHost, g_ctx.stream));
         HIPCHK(hipEventRecord(g_ctx.ev_strip[k], g_ctx.stream));
      }
      const unsigned char *src = (const unsigned char *)g_ctx.staging + d[0].img_off;
      unsigned char *dst = (unsigned char *)d[0].out;
      for (int k = 0; k < FL3_IMAGE_STRIPS; k++) {
         HIPCHK(hipEventSynchronize(g_ctx.ev_strip[k]));
         if (d[0].out_stride == rb)
            memcpy(dst + rb * (size_t)r0[k], src + rb * (size_t)r0[k], rb * (size_t)(r0[k + 1] - r0[k]));
         else
            for (int j = r0[k]; j < r0[k + 1]; j++)
               memcpy(dst + (size_t)j * d[0].out_stride, src + (size_t)j * rb, rb);
      }
      return 0;
   }
   HIPCHK(hipMemcpyAsync(g_ctx.staging, g_ctx.image.p, end, hipMemcpyDeviceToHost, g_ctx.stream));
   HIPCHK(hipStreamSynchronize(g_ctx.stream));
   for (int i = 0; i < n; i++) {
      const unsigned char *src = (const unsigned char *)g_ctx.staging + d[i].img_off;
      unsigned char *dst = (unsigned char *)d[i].out;
      if (d[i].out_stride == d[i].row_bytes)
         memcpy(dst, src, d[i].row_bytes * (size_t)d[i].rows);
      else
         for (int j = 0; j < d[i].rows; j++)
            memcpy(dst + (size_t)j * d[i].out_stride, src + (size_t)j * d[i].row_bytes, d[i].row_bytes);
   }
   return 0;
}

extern "C" int fl3d_images_download(int slot, size_t bytes, const unsigned char **host) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   if (slot < 0 || slot > 1 || !host) {
      set_err("fl3d_images_download: bad slot");
      return 1;
   }
   *host = nullptr;
   if (bytes > g_ctx.image.cap) {
      set_err("fl3d_images_download: %s%ld bytes beyond the image arena", "", (long)bytes);
      return 1;
   }
   if (g_ctx.slot_cap[slot] < bytes) {
      // (grown at least twofold, within the image arena: pinned allocations are slow)
      size_t want = bytes > 2 * g_ctx.slot_cap[slot] ? bytes : 2 * g_ctx.slot_cap[slot];
      if (want > g_ctx.image.cap) want = g_ctx.image.cap;
      if (g_ctx.slot[slot]) (void)hipHostFree(g_ctx.slot[slot]);
      g_ctx.slot[slot] = nullptr;
      g_ctx.slot_cap[slot] = 0;
      HIPCHK(hipHostMalloc(&g_ctx.slot[slot], want, hipHostMallocDefault));
      g_ctx.slot_cap[slot] = want;
   }
   if (bytes) HIPCHK(hipMemcpyAsync(g_ctx.slot[slot], g_ctx.image.p, bytes, hipMemcpyDeviceToHost, g_ctx.stream));
   HIPCHK(hipStreamSynchronize(g_ctx.stream));
   *host = (const unsigned char *)g_ctx.slot[slot];
   return 0;
}

extern "C" int fl3d_tonemap(const fl3d_tm_params *p, const double *filter, void *out_host, size_t out_stride) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   fl3d_image_dst d;
   d.img_off = 0;
   d.out = out_host;
   d.out_stride = out_stride;
   d.row_bytes = (size_t)p->image_width * p->nchan * p->bytes_per_channel;
   d.rows = p->image_width > 0 ? p->image_height : 0;
   if (fl3d_images_begin(d.row_bytes * (size_t)p->image_height,
                         sizeof(double) * p->filter_width * p->filter_width))
      return 1;
   if (fl3d_tonemap_async(p, filter, 0, 0)) return 1;
   return fl3d_images_fetch(&d, 1);
}

extern "C" int fl3d_download_buckets(void *host, size_t bytes) {
   HIPCHK(hipMemcpyAsync(host, cur_buckets(), bytes, hipMemcpyDeviceToHost, g_ctx.stream));
   HIPCHK(hipStreamSynchronize(g_ctx.stream));
   return 0;
}

extern "C" int fl3d_download_accum(void *host, size_t bytes) {
   HIPCHK(hipMemcpyAsync(host, cur_accum(), bytes, hipMemcpyDeviceToHost, g_ctx.stream));
   HIPCHK(hipStreamSynchronize(g_ctx.stream));
   return 0;
}

extern "C" int fl3d_replay(const void *job, size_t job_bytes, uint32_t *isaac_host, int nwalkers,
                           double batch_size, int sbs, int fuse, double *badvals) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   return iterate_replay(job, job_bytes, isaac_host, nwalkers, batch_size, sbs, fuse, badvals);
}

// ------------------------------------------------------------------ stage entry points (fl3h.h)

// Pack a randctx into the 36-word device form.
static void pack_isaac(const randctx *rc, uint32_t *s) {
   s[0] = (uint32_t)rc->randcnt;
   for (int i = 0; i < 16; i++) s[1 + i] = (uint32_t)rc->randrsl[i];
   for (int i = 0; i < 16; i++) s[17 + i] = (uint32_t)rc->randmem[i];
   s[33] = (uint32_t)rc->randa;
   s[34] = (uint32_t)rc->randb;
   s[35] = (uint32_t)rc->randc;
}

// Back to the host layout.  randmem words the device rewrote are 32-bit;
// words it never touched keep their host (possibly 64-bit) value.
static void unpack_isaac(const uint32_t *s, randctx *rc) {
   const int refilled = (uint32_t)rc->randc != s[35];
   rc->randcnt = s[0];
   for (int i = 0; i < 16; i++) rc->randrsl[i] = s[1 + i];
   if (refilled)
      for (int i = 0; i < 16; i++) rc->randmem[i] = s[17 + i];
   rc->randa = s[33];
   rc->randb = s[34];
   rc->randc = s[35];
}

extern "C" int fl3h_iterate(flam3_genome *cp, int n, int fuse, double *samples,
                            const unsigned short *xform_distrib, randctx *rc, int nwalkers,
                            int *badvals) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   if (ctx_init()) return 1;
   if (nwalkers < 1 || n < 0) { set_err("fl3h_iterate: bad sizes"); return 1; }
   // flam3_iterate recomputes the per-variation constants of every xform
   // on the caller's genome (flam3.c:246-247)
   for (int i = 0; i < cp->num_xforms; i++) fl3_xform_precalc(&cp->xform[i]);
   size_t jb = 0;
   void *job = fl3_pack_job(cp, xform_distrib, NULL, &jb);
   if (!job) return 1;
   const fl3_job_hdr *h = (const fl3_job_hdr *)job;
   uint32_t *st = (uint32_t *)malloc(sizeof(uint32_t) * 36 * nwalkers);
   for (int w = 0; w < nwalkers; w++) pack_isaac(&rc[w], st + 36 * w);
   size_t sb = sizeof(double) * 4 * (size_t)nwalkers;
   size_t ob = sizeof(double) * 4 * (size_t)n * nwalkers;
   int rcode = 1;
   do {
      if (upload_job(job, jb)) break;
      if (g_ctx.isaac.ensure(sizeof(uint32_t) * 36 * nwalkers + sizeof(int) * nwalkers)) break;
      if (g_ctx.points.ensure(sb + ob + 16)) break;
      double *dstart = (double *)g_ctx.points.p;
      double *dout = dstart + 4 * (size_t)nwalkers;
      int *dbad = (int *)((uint32_t *)g_ctx.isaac.p + 36 * nwalkers);
      if (hipMemcpyAsync(g_ctx.isaac.p, st, sizeof(uint32_t) * 36 * nwalkers, hipMemcpyHostToDevice, g_ctx.stream) != hipSuccess) break;
      double *starts = (double *)malloc(sb);
      for (int w = 0; w < nwalkers; w++)
         for (int k = 0; k < 4; k++) starts[4 * w + k] = samples[(size_t)w * 4 * n + k];
      if (hipMemcpyAsync(dstart, starts, sb, hipMemcpyHostToDevice, g_ctx.stream) != hipSuccess) { set_err("device copy failed"); break; }
      PointsArgs a;
      a.job = (const unsigned char *)g_ctx.job.p;
      a.start = dstart;
      a.out = dout;
      a.isaac = (uint32_t *)g_ctx.isaac.p;
      a.badvals = dbad;
      a.n = n;
      a.fuse = fuse;
      a.nwalkers = nwalkers;
      if (fl3k_launch_points(&a, (nwalkers + 63) / 64, lds_bytes_for<double>(h), g_ctx.stream)) {
         set_err("k_iterate_points launch failed");
         free(starts);
         break;
      }
      int *bad = (int *)malloc(sizeof(int) * nwalkers);
      if (hipMemcpyAsync(samples, dout, ob, hipMemcpyDeviceToHost, g_ctx.stream) != hipSuccess) { set_err("device copy failed"); break; }
      if (hipMemcpyAsync(st, g_ctx.isaac.p, sizeof(uint32_t) * 36 * nwalkers, hipMemcpyDeviceToHost, g_ctx.stream) != hipSuccess) { set_err("device copy failed"); break; }
      if (hipMemcpyAsync(bad, dbad, sizeof(int) * nwalkers, hipMemcpyDeviceToHost, g_ctx.stream) != hipSuccess) { set_err("device copy failed"); break; }
      if (hipStreamSynchronize(g_ctx.stream) != hipSuccess) { set_err("k_iterate_points failed"); free(starts); free(bad); break; }
      for (int w = 0; w < nwalkers; w++) {
         unpack_isaac(st + 36 * w, &rc[w]);
         if (badvals) badvals[w] = bad[w];
      }
      free(starts);
      free(bad);
      rcode = 0;
   } while (0);
   free(st);
   free(job);
   return rcode;
}

extern "C" int fl3h_xform_preview(flam3_genome *cp, int xi, double range, int numvals, int depth,
                                  double *result, randctx *rc) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   if (ctx_init()) return 1;
   if (xi < 0 || xi >= cp->num_xforms || numvals < 0) { set_err("fl3h_xform_preview: bad arguments"); return 1; }
   // as flam3.c:313-326: weight 1 for xi while preparing
   const double oldweight = cp->xform[xi].density;
   cp->xform[xi].density = 1.0;
   if (fl3_prepare_precalc_flags(cp)) {
      cp->xform[xi].density = oldweight;
      set_err("fl3h_xform_preview: precalc failed");
      return 1;
   }
   fl3_xform_precalc(&cp->xform[xi]);
   size_t jb = 0;
   void *job = fl3_pack_job(cp, NULL, NULL, &jb);
   cp->xform[xi].density = oldweight;
   if (!job) return 1;
   const fl3_job_hdr *h = (const fl3_job_hdr *)job;
   const size_t nres = 2 * (size_t)(2 * numvals + 1) * (2 * numvals + 1);
   uint32_t st[36];
   pack_isaac(rc, st);
   int rcode = 1;
   do {
      if (upload_job(job, jb)) break;
      if (g_ctx.isaac.ensure(sizeof(st))) break;
      if (g_ctx.points.ensure(sizeof(double) * nres)) break;
      if (hipMemcpyAsync(g_ctx.isaac.p, st, sizeof(st), hipMemcpyHostToDevice, g_ctx.stream) != hipSuccess) { set_err("device copy failed"); break; }
      if (fl3k_launch_preview((const unsigned char *)g_ctx.job.p, xi, range, numvals, depth,
                              (double *)g_ctx.points.p, (uint32_t *)g_ctx.isaac.p, lds_bytes_for<double>(h),
                              g_ctx.stream)) {
         set_err("k_xform_preview launch failed");
         break;
      }
      if (hipMemcpyAsync(result, g_ctx.points.p, sizeof(double) * nres, hipMemcpyDeviceToHost, g_ctx.stream) != hipSuccess) { set_err("device copy failed"); break; }
      if (hipMemcpyAsync(st, g_ctx.isaac.p, sizeof(st), hipMemcpyDeviceToHost, g_ctx.stream) != hipSuccess) { set_err("device copy failed"); break; }
      if (hipStreamSynchronize(g_ctx.stream) != hipSuccess) { set_err("k_xform_preview failed"); break; }
      unpack_isaac(st, rc);
      rcode = 0;
   } while (0);
   free(job);
   return rcode;
}

extern "C" int fl3h_apply_xform_fp32(flam3_genome *cp, int xi, const double *points, long long npoints,
                                     double *out, int tier) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   if (ctx_init()) return 1;
   if (xi < 0 || xi >= cp->num_xforms || npoints < 0 || tier < -1 || tier > FL3_TIER_FULL) {
      set_err("fl3h_apply_xform_fp32: bad arguments");
      return 1;
   }
   const double oldweight = cp->xform[xi].density;
   cp->xform[xi].density = 1.0;
   if (fl3_prepare_precalc_flags(cp)) {
      cp->xform[xi].density = oldweight;
      set_err("fl3h_apply_xform_fp32: precalc failed");
      return 1;
   }
   fl3_xform_precalc(&cp->xform[xi]);
   size_t jb = 0;
   void *job = fl3_pack_job(cp, NULL, NULL, &jb);
   cp->xform[xi].density = oldweight;
   if (!job) return 1;
   const fl3_job_hdr *h = (const fl3_job_hdr *)job;
   if (tier < 0) tier = pick_tier(h);
   const size_t pb = sizeof(double) * 4 * (size_t)npoints;
   int rcode = 1;
   do {
      if (npoints == 0) { rcode = 0; break; }
      if (upload_job(job, jb)) break;
      if (g_ctx.points.ensure(2 * pb)) break;
      double *din = (double *)g_ctx.points.p, *dout = din + 4 * npoints;
      if (hipMemcpyAsync(din, points, pb, hipMemcpyHostToDevice, g_ctx.stream) != hipSuccess) { set_err("device copy failed"); break; }
      if (fl3k_launch_apply_fp32((const unsigned char *)g_ctx.job.p, xi, din, npoints, dout, 0x5eedULL,
                                 lds_bytes_for<float>(h), g_ctx.stream, tier)) {
         set_err("k_apply_fp32 launch failed");
         break;
      }
      if (hipMemcpyAsync(out, dout, pb, hipMemcpyDeviceToHost, g_ctx.stream) != hipSuccess) { set_err("device copy failed"); break; }
      if (hipStreamSynchronize(g_ctx.stream) != hipSuccess) { set_err("k_apply_fp32 failed"); break; }
      rcode = 0;
   } while (0);
   free(job);
   return rcode;
}

extern "C" int fl3h_deposit_points(const fl3h_camera *cam, const double *points, long long npoints,
                                   int bits, void *buckets) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   if (ctx_init()) return 1;
   if (bits != 32 && bits != 33 && bits != 64) { set_err("fl3h_deposit_points: bits must be 32, 33 or 64"); return 1; }
   size_t jb = 0;
   void *job = fl3_pack_camera_job(cam, &jb);
   if (!job) return 1;
   const fl3_job_hdr *h = (const fl3_job_hdr *)job;
   const size_t cells = (size_t)cam->width * cam->height;
   const size_t bb = cells * bucket_bytes(bits);
   const size_t pb = sizeof(double) * 4 * (size_t)npoints;
   int rcode = 1;
   do {
      if (upload_job(job, jb)) break;
      if (g_ctx.points.ensure(pb + 16)) break;
      if (g_ctx.misc.ensure(bb)) break;
      if (hipMemcpyAsync(g_ctx.points.p, points, pb, hipMemcpyHostToDevice, g_ctx.stream) != hipSuccess) { set_err("device copy failed"); break; }
      if (hipMemcpyAsync(g_ctx.misc.p, buckets, bb, hipMemcpyHostToDevice, g_ctx.stream) != hipSuccess) { set_err("device copy failed"); break; }
      if (hipMemsetAsync(g_ctx.counters, 0, 16, g_ctx.stream) != hipSuccess) { set_err("device copy failed"); break; }
      const size_t lds = lds_bytes_for<double>(h);
      const dim3 grid(npoints > 0 ? (unsigned)((npoints + 255) / 256 < 4096 ? (npoints + 255) / 256 : 4096) : 1);
      if (fl3k_launch_deposit((const unsigned char *)g_ctx.job.p, (const double *)g_ctx.points.p, npoints,
                              g_ctx.misc.p, (uint32_t *)(g_ctx.counters + 1), grid.x, lds, g_ctx.stream, bits)) {
         set_err("k_deposit_points launch failed");
         break;
      }
      if (hipMemcpyAsync(buckets, g_ctx.misc.p, bb, hipMemcpyDeviceToHost, g_ctx.stream) != hipSuccess) { set_err("device copy failed"); break; }
      if (hipStreamSynchronize(g_ctx.stream) != hipSuccess) { set_err("k_deposit_points failed"); break; }
      rcode = 0;
   } while (0);
   free(job);
   return rcode;
}

static int density_estimate_host(const void *buckets, int width, int height, int bits, int oversample, double k1,
                                 double k2, double curve, int max_filtered_counts, int max_filter_index,
                                 int kernel_size, int nfilters, const double *filter_widths,
                                 const double *filter_coefs, void *accum, int exact);

extern "C" int fl3h_density_estimate(const void *buckets, int width, int height, int bits, int oversample,
                                     double k1, double k2, double curve, int max_filtered_counts,
                                     int max_filter_index, int kernel_size, int nfilters,
                                     const double *filter_widths, const double *filter_coefs, void *accum) {
   return density_estimate_host(buckets, width, height, bits, oversample, k1, k2, curve, max_filtered_counts,
                                max_filter_index, kernel_size, nfilters, filter_widths, filter_coefs, accum, 1);
}

extern "C" int fl3h_density_estimate_production(const void *buckets, int width, int height, int bits,
                                                int oversample, double k1, double k2, double curve,
                                                int max_filtered_counts, int max_filter_index, int kernel_size,
                                                int nfilters, const double *filter_widths,
                                                const double *filter_coefs, void *accum) {
   return density_estimate_host(buckets, width, height, bits, oversample, k1, k2, curve, max_filtered_counts,
                                max_filter_index, kernel_size, nfilters, filter_widths, filter_coefs, accum, 0);
}

static int density_estimate_host(const void *buckets, int width, int height, int bits, int oversample, double k1,
                                 double k2, double curve, int max_filtered_counts, int max_filter_index,
                                 int kernel_size, int nfilters, const double *filter_widths,
                                 const double *filter_coefs, void *accum, int exact) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   if (ctx_init()) return 1;
   if (bits != 33 && bits != 64) { set_err("fl3h_density_estimate: bits must be 33 or 64"); return 1; }
   const size_t cells = (size_t)width * height;
   if (g_ctx.buckets.ensure(cells * bucket_bytes(bits)) || g_ctx.accum.ensure(cells * accum_bytes(bits))) return 1;
   g_ctx.width = width;
   g_ctx.height = height;
   g_ctx.bits = bits;
   g_ctx.cell0 = 0;
   HIPCHK(hipMemcpyAsync(g_ctx.buckets.p, buckets, cells * bucket_bytes(bits), hipMemcpyHostToDevice, g_ctx.stream));
   HIPCHK(hipMemcpyAsync(g_ctx.accum.p, accum, cells * accum_bytes(bits), hipMemcpyHostToDevice, g_ctx.stream));
   fl3_de_filters de;
   de.max_filtered_counts = max_filtered_counts;
   de.max_filter_index = max_filter_index;
   de.kernel_size = kernel_size;
   de.nfilters = nfilters;
   de.filter_widths = (double *)filter_widths;
   de.filter_coefs = (double *)filter_coefs;
   fl3d_de_params p;
   p.oversample = oversample;
   p.k1 = k1;
   p.k2 = k2;
   p.curve = curve;
   p.exact = exact;
   if (max_filter_index > 0) {
      if (fl3d_density(&p, &de)) return 1;
   } else {
      if (fl3d_logscale(k1, k2)) return 1;
   }
   HIPCHK(hipMemcpyAsync(accum, g_ctx.accum.p, cells * accum_bytes(bits), hipMemcpyDeviceToHost, g_ctx.stream));
   HIPCHK(hipStreamSynchronize(g_ctx.stream));
   return 0;
}

// ------------------------------------------------------------------ density estimation without allocation
//
// fl3h_density_estimate_noalloc works on the caller's device buffers and
// keeps its own intermediate data -- the per-cell kernel index and
// per-source values of the two DE passes, and the device copy of the filter
// bank -- in `scratch`, a fixed device array, so it allocates nothing.  Every
// GPU the library runs on holds its own copy of the array until the process
// ends.  The array's address on a device is looked up once, with that device
// current, and kept in g_scratch; host code passes that address to the
// kernels, never the array's name.  Calls on one device share the array and
// must not overlap.
#define FL3_SCRATCH_BYTES (64u << 20)
#define FL3_SCRATCH_TABLES (4u << 20)   // filter bank: coefficients (double + float) and widths
#define FL3_SCRATCH_BLOCKS (256u << 10)  // widest kernel per 16x16 block (65536 blocks)
#define FL3_MAX_DEVICES 64
static __device__ __align__(256) unsigned char scratch[FL3_SCRATCH_BYTES];
static unsigned char *g_scratch[FL3_MAX_DEVICES];
static std::mutex g_scratch_mu;

static unsigned char *scratch_on(int dev) {
   std::lock_guard<std::mutex> lk(g_scratch_mu);
   if (dev < 0 || dev >= FL3_MAX_DEVICES) {
      set_err("fl3h_density_estimate_noalloc: device index out of range");
      return nullptr;
   }
   if (!g_scratch[dev]) {
      unsigned char *ptr = nullptr;
      const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !ptr) {
         snprintf(g_err, sizeof(g_err), "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
         return nullptr;
      }
      g_scratch[dev] = ptr;
   }
   return g_scratch[dev];
}

static size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

extern "C" long long fl3h_density_estimate_noalloc_max_cells(int bits) {
   const size_t per = sizeof(uint16_t) + accum_bytes(bits == 64 ? 64 : 33) + sizeof(double);
   return (long long)((FL3_SCRATCH_BYTES - FL3_SCRATCH_TABLES - FL3_SCRATCH_BLOCKS - 3 * 256) / per);
}

extern "C" int fl3h_density_estimate_noalloc(const void *buckets, int width, int height, int bits, int oversample,
                                             double k1, double k2, double curve, int max_filtered_counts,
                                             int max_filter_index, int kernel_size, int nfilters,
                                             const double *filter_widths, const double *filter_coefs,
                                             void *accum) {
   if (bits != 33 && bits != 64) { set_err("fl3h_density_estimate_noalloc: bits must be 33 or 64"); return 1; }
   if (!buckets || !accum || width <= 0 || height <= 0 || width >= 65536 || height >= 65536) {
      set_err("fl3h_density_estimate_noalloc: bad buffers or histogram size");
      return 1;
   }
   const size_t cells = (size_t)width * height;
   if ((long long)cells > fl3h_density_estimate_noalloc_max_cells(bits)) {
      snprintf(g_err, sizeof(g_err),
               "fl3h_density_estimate_noalloc: %zu cells exceed the %lld cells its %u-byte scratch holds at bits %d",
               cells, fl3h_density_estimate_noalloc_max_cells(bits), FL3_SCRATCH_BYTES, bits);
      return 1;
   }
   if (max_filter_index == 0) {   // plain log scaling: no intermediate data
      if (fl3k_launch_logscale(buckets, accum, (long)cells, k1, k2, 0, bits)) {
         set_err("k_logscale launch failed");
         return 1;
      }
      HIPCHK(hipStreamSynchronize(0));
      return 0;
   }
   const size_t nc = (size_t)nfilters * kernel_size, nw = (size_t)nfilters;
   if (!filter_widths || !filter_coefs || nfilters <= 0 || kernel_size <= 0 || max_filter_index >= 1024 ||
       max_filter_index >= nfilters ||
       align256(sizeof(double) * (nc + nw)) + sizeof(float) * nc > FL3_SCRATCH_TABLES) {
      set_err("fl3h_density_estimate_noalloc: filter bank missing, or larger than the scratch's table space");
      return 1;
   }
   int dev = 0;
   HIPCHK(hipGetDevice(&dev));
   unsigned char *base = scratch_on(dev);
   if (!base) return 1;
   // layout: tables, the per-block kernel radii, then per cell the kernel
   // index, the log terms and the values
   double *coefs = (double *)base;
   double *widths = coefs + nc;
   float *coefs_f = (float *)(base + align256(sizeof(double) * (nc + nw)));
   int *blkmax = (int *)(base + FL3_SCRATCH_TABLES);
   unsigned char *cellp = base + FL3_SCRATCH_TABLES + FL3_SCRATCH_BLOCKS;
   uint16_t *findex = (uint16_t *)cellp;
   double *srcl = (double *)(cellp + align256(sizeof(uint16_t) * cells));
   void *srcv = (void *)((unsigned char *)srcl + align256(sizeof(double) * cells));
   HIPCHK(hipMemcpyAsync(coefs, filter_coefs, sizeof(double) * nc, hipMemcpyHostToDevice, 0));
   HIPCHK(hipMemcpyAsync(widths, filter_widths, sizeof(double) * nw, hipMemcpyHostToDevice, 0));
   if (fl3k_launch_f64_to_f32(coefs, coefs_f, (long)nc, 0)) {
      set_err("k_f64_to_f32 launch failed");
      return 1;
   }
   DEArgs a;
   memset(&a, 0, sizeof(a));
   a.buckets = buckets;
   a.accum = accum;
   a.coefs = coefs;
   a.widths = widths;
   a.coefs_t = bits == 64 ? (const void *)coefs : (const void *)coefs_f;
   a.width = width;
   a.height = height;
   a.oversample = oversample;
   a.max_filtered_counts = max_filtered_counts;
   a.max_filter_index = max_filter_index;
   a.kernel_size = kernel_size;
   a.k1 = k1;
   a.k2 = k2;
   a.curve = curve;
   int half = 0;
   while ((half + 1) * (half + 2) / 2 < kernel_size) half++;
   a.half = half;
   a.findex = findex;
   a.srcv = srcv;
   a.srcl = srcl;
   a.exact = 1;
   a.nbx = (width + DG_REGION_HOST - 1) / DG_REGION_HOST;
   if ((size_t)a.nbx * ((height + DG_REGION_HOST - 1) / DG_REGION_HOST) * sizeof(int) > FL3_SCRATCH_BLOCKS) {
      set_err("fl3h_density_estimate_noalloc: too many 16x16 blocks for the scratch");   // (not reached: sides < 65536)
      return 1;
   }
   a.blkmax = blkmax;
   if (fl3k_launch_de_tiled(&a, 0, bits)) {
      set_err("density estimation launch failed");
      return 1;
   }
   HIPCHK(hipStreamSynchronize(0));
   return 0;
}

extern "C" int fl3h_tonemap(const void *accum, int width, int height, int bits, const fl3h_tonemap_params *tp,
                            const double *filter, void *out) {
   std::lock_guard<std::recursive_mutex> lk(g_mu);
   if (ctx_init()) return 1;
   const size_t cells = (size_t)width * height;
   if (g_ctx.accum.ensure(cells * accum_bytes(bits))) return 1;
   g_ctx.width = width;
   g_ctx.height = height;
   g_ctx.bits = bits;
   g_ctx.cell0 = 0;
   HIPCHK(hipMemcpyAsync(g_ctx.accum.p, accum, cells * accum_bytes(bits), hipMemcpyHostToDevice, g_ctx.stream));
   fl3d_tm_params p;
   p.image_width = tp->image_width;
   p.image_height = tp->image_height;
   p.oversample = tp->oversample;
   p.de_offset = tp->de_offset;
   p.filter_width = tp->filter_width;
   p.nchan = tp->nchan;
   p.transp = tp->transp;
   p.bytes_per_channel = tp->bytes_per_channel;
   p.earlyclip = tp->earlyclip;
   p.g = tp->g;
   p.linrange = tp->linrange;
   p.vibrancy = tp->vibrancy;
   p.highpow = tp->highpow;
   for (int k = 0; k < 3; k++) p.background[k] = tp->background[k];
   return fl3d_tonemap(&p, filter, out, (size_t)tp->image_width * tp->nchan * tp->bytes_per_channel);
}

