// fl3_kernels_exact.hip - FP64 kernels that must follow the reference's
// arithmetic operation by operation: replay walkers, batched iterate, xform
// preview, point deposit, density estimation, log scaling and the tone map.
// Compiled with FP contraction off (the reference's x86-64 build has no FMA).
#include <type_traits>
#include "fl3_core.h"
#include "fl3_tonemath.h"
#include "fl3_launch.h"

// Replay walkers: walker w is reference iteration thread w (rect.c:253-527)
// with its ISAAC state seeded on the host as rect.c:858-865 does, running
// the same sub-batch loop in FP64 (exact-mode deposits).
template <int BITS>
__global__ __launch_bounds__(64) void k_iter_replay(IterArgs a) {
   unsigned char *lds = fl3_lds;
   GenL<double> G;
   load_genome<double>(a.job, lds, G);
   __syncthreads();
   const int wid = blockIdx.x * blockDim.x + threadIdx.x;
   if (wid >= a.nwalkers) return;
   IsaacRng rng;
   rng.load(a.isaac + 36 * wid);
   uint32_t bad = 0;
   const double bs = a.batch_size;
   const bool use_final = G.final_index >= 0;
   for (double sub = 0; sub < bs; sub += a.sbs) {
      int sbs = (sub + a.sbs > bs) ? (int)(bs - sub) : a.sbs;
      double p[4], q[4];
      p[0] = u11<double>(rng);
      p[1] = u11<double>(rng);
      p[2] = u01<double>(rng);
      p[3] = u01<double>(rng);
      int lastxf = 0, consec = 0;
      for (int i = -a.fuse; i < sbs;) {
         const uint32_t r = rng.next();
         const int fn = select_xform(G, G.chaos_enable ? lastxf : 0, r & 16383u);
         if (apply_xform<double>(G, fn, p, q, rng)) {
            consec++;
            bad++;
            if (consec < 5) {
               p[0] = q[0]; p[1] = q[1]; p[2] = q[2]; p[3] = q[3];
               continue;
            }
            consec = 0;
         } else
            consec = 0;
         lastxf = fn + 1;
         p[0] = q[0]; p[1] = q[1]; p[2] = q[2]; p[3] = q[3];
         if (use_final) {
            if (G.final_opacity == 1.0 || u01<double>(rng) < G.final_opacity) {
               apply_xform<double>(G, G.final_index, p, q, rng);
               q[3] = p[3];
            }
         }
         if (i >= 0) deposit<double, BITS, true>(G, q, a.buckets, a.sat);
         i++;
      }
   }
   rng.store(a.isaac + 36 * wid);
   if (bad) atomicAdd(a.badvals, (unsigned long long)bad);
}

// Batched flam3_iterate (flam3.c:232-299) in FP64: walker w starts from
// start[4w..4w+3] with ISAAC state w and writes n samples.

__global__ __launch_bounds__(64) void k_iterate_points(PointsArgs a) {
   unsigned char *lds = fl3_lds;
   GenL<double> G;
   load_genome<double>(a.job, lds, G);
   __syncthreads();
   const int wid = blockIdx.x * blockDim.x + threadIdx.x;
   if (wid >= a.nwalkers) return;
   IsaacRng rng;
   rng.load(a.isaac + 36 * wid);
   double p[4], q[4];
   for (int k = 0; k < 4; k++) p[k] = a.start[4 * wid + k];
   int lastxf = 0, consec = 0, bad = 0;
   const bool use_final = G.final_index >= 0;
   double *out = a.out + (size_t)wid * a.n * 4;
   for (int i = -a.fuse; i < a.n;) {
      const uint32_t r = rng.next();
      const int fn = select_xform(G, G.chaos_enable ? lastxf : 0, r & 16383u);
      if (apply_xform<double>(G, fn, p, q, rng)) {
         consec++;
         bad++;
         if (consec < 5) {
            p[0] = q[0]; p[1] = q[1]; p[2] = q[2]; p[3] = q[3];
            continue;
         }
         consec = 0;
      } else
         consec = 0;
      lastxf = fn + 1;
      p[0] = q[0]; p[1] = q[1]; p[2] = q[2]; p[3] = q[3];
      if (use_final) {
         if (G.final_opacity == 1.0 || u01<double>(rng) < G.final_opacity) {
            apply_xform<double>(G, G.final_index, p, q, rng);
            q[3] = p[3];
         }
      }
      if (i >= 0) {
         out[4 * i + 0] = q[0]; out[4 * i + 1] = q[1];
         out[4 * i + 2] = q[2]; out[4 * i + 3] = q[3];
      }
      i++;
   }
   rng.store(a.isaac + 36 * wid);
   a.badvals[wid] = bad;
}

// Apply xform xi `depth` times to each point of a grid (flam3.c:301-349).
__global__ __launch_bounds__(64) void k_xform_preview(const unsigned char *job, int xi, double range,
                                                      int numvals, int depth, double *result,
                                                      uint32_t *isaac) {
   unsigned char *lds = fl3_lds;
   GenL<double> G;
   load_genome<double>(job, lds, G);
   __syncthreads();
   if (blockIdx.x != 0 || threadIdx.x != 0) return;
   IsaacRng rng;
   rng.load(isaac);
   const double incr = range / (double)numvals;
   int outi = 0;
   for (int xx = -numvals; xx <= numvals; xx++)
      for (int yy = -numvals; yy <= numvals; yy++) {
         double p[4] = {(double)xx * incr, (double)yy * incr, 0.0, 0.0}, q[4];
         for (int d = 0; d < depth; d++) {
            apply_xform<double>(G, xi, p, q, rng);
            p[0] = q[0]; p[1] = q[1]; p[2] = q[2]; p[3] = q[3];
         }
         result[outi] = p[0];
         result[outi + 1] = p[1];
         outi += 2;
      }
   rng.store(isaac);
}

// Deposit a given point stream (stage test of rect.c:411-516).
template <int BITS>
__global__ __launch_bounds__(256) void k_deposit_points(const unsigned char *job, const double *pts, long long n,
                                                        void *buckets, uint32_t *sat) {
   unsigned char *lds = fl3_lds;
   GenL<double> G;
   load_genome<double>(job, lds, G);
   __syncthreads();
   for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
        i += (long long)gridDim.x * blockDim.x) {
      double q[4] = {pts[4 * i], pts[4 * i + 1], pts[4 * i + 2], pts[4 * i + 3]};
      deposit<double, BITS, true>(G, q, buckets, sat);
   }
}

// ------------------------------------------------------------------ density estimation


// ---- tiled density estimation ------------------------------------------
//
// The reference scatters each source cell's kernel into the accumulator
// (rect.c:126-173).  Here pass 1 picks every source's kernel (the hit-count
// neighbourhood sum and index selection of rect.c:92-117) and pass 2 gives
// each workgroup one output tile held in LDS: it visits every source within
// the largest kernel radius of the tile, adds the part of that source's
// kernel that falls inside the tile with LDS atomics, and writes the tile
// back once.  No global atomics; each accumulator cell is read and written
// once per batch.
#define DE_SKIP 0xffffu
#define DE_TW 64
#define DE_TH 64
#define DE_PLANE (DE_TW * DE_TH)  // the LDS tile holds one plane per channel

// Channel i of the histogram as a double: uint32 cells with a float
// accumulator (bits 33), double cells with a double one (bits 64).
__device__ __forceinline__ double bucket_value(const void *b, long i, float) { return (double)((const uint32_t *)b)[i]; }
__device__ __forceinline__ double bucket_value(const void *b, long i, double) { return ((const double *)b)[i]; }

template <typename BT, typename AT>
__global__ __launch_bounds__(256) void k_de_prep(DEArgs a) {
   const int os = a.oversample;
   const int ss = os / 2;
   const bool scf = !(os & 1);
   const double scfact = pow(os / (os + 1.0), 2.0);
   const int W = a.width, H = a.height;
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   const int j = blockIdx.y;
   if (i >= W) return;
   const long cell = i + (long)j * W;
   uint16_t out = DE_SKIP;
   const BT *b = (const BT *)a.buckets;
   // rows (os - 1) .. H - os + 1 (rect.c:979-1000 spans H - 2 (os - 1) + 1
   // rows from os - 1); for os 1 that includes row H, past the histogram,
   // which counts as empty here
   if (i >= os - 1 && i < W - (os - 1) && j >= os - 1 && j <= H - os + 1) {
      const double c4 = (double)b[cell * 5 + 4], c3 = (double)b[cell * 5 + 3];
      if (c4 != 0 && c3 != 0) {
         double fsel = 0.0;
         for (int ii = -ss; ii <= ss; ii++)
            for (int jj = -ss; jj <= ss; jj++) {
               const int y = j + jj;
               if (y < H) fsel += (double)b[((i + ii) + (long)y * W) * 5 + 4] / 255.0;
            }
         if (scf) fsel *= scfact;
         int fi;
         if (fsel > a.max_filtered_counts)
            fi = a.max_filter_index;
         else if (fsel <= 100)
            fi = (int)ceil(fsel) - 1;
         else
            fi = 100 + (int)floor(pow(fsel - 100, a.curve));
         if (fi > a.max_filter_index) fi = a.max_filter_index;
         out = (uint16_t)fi;
         // the per-source factor of rect.c:137, k1 log(1 + c3 k2) / c3,
         // applied to the colour once; the kernel coefficient multiplies
         // it per tap
         const double lnum = a.k1 * log(1.0 + c3 * a.k2);
         a.srcl[cell] = lnum;
         const double sc = lnum / c3;
         AT *v = (AT *)a.srcv + cell * 4;
         v[0] = (AT)((double)b[cell * 5 + 0] * sc);
         v[1] = (AT)((double)b[cell * 5 + 1] * sc);
         v[2] = (AT)((double)b[cell * 5 + 2] * sc);
         v[3] = (AT)(c3 * sc);
      }
   }
   a.findex[cell] = out;
}

// Add the part of source (sx, sy)'s kernel that falls in the rectangle
// [x0, x1] x [y0, y1] of the tile, positions p0, p0 + step, ...  The kernel
// is radial: the coefficient of offset (dx, dy) is entry
// hi (hi + 1) / 2 + lo of the source's row, hi = max(|dx|, |dy|),
// lo = min(|dx|, |dy|) (the triangle walk of rect.c:123-170).
template <typename AT>
__device__ __forceinline__ void de_splat(AT *tile, const AT *crow, const AT *v, int sx, int sy, int x0, int x1,
                                         int y0, int y1, int tx0, int ty0, int p0, int step) {
   const int w = x1 - x0 + 1, n = w * (y1 - y0 + 1);
   const AT v0 = v[0], v1 = v[1], v2 = v[2], v3 = v[3];
   for (int p = p0; p < n; p += step) {
      const int y = y0 + p / w, x = x0 + p % w;
      const int dx = abs(x - sx), dy = abs(y - sy);
      const int hi = max(dx, dy), lo = min(dx, dy);
      const AT coef = crow[hi * (hi + 1) / 2 + lo];
      if (coef == (AT)0) continue;
      AT *t = tile + (y - ty0) * DE_TW + (x - tx0);
      atomicAdd(t, coef * v0);
      atomicAdd(t + DE_PLANE, coef * v1);
      atomicAdd(t + 2 * DE_PLANE, coef * v2);
      atomicAdd(t + 3 * DE_PLANE, coef * v3);
   }
}

// One workgroup per 64x64 output tile, accumulated in LDS.  Each wave takes
// 64 consecutive sources of the tile's halo; a source whose kernel covers at
// most DE_LANE_TAPS cells of the tile is splatted by its own lane, larger
// ones are splatted by the whole wave, one after another, so lanes do not
// idle behind the widest kernel of their wave.
#define DE_LANE_TAPS 24
template <typename AT>
__global__ __launch_bounds__(256) void k_de_tile(DEArgs a) {
   __shared__ AT tile[DE_TH * DE_TW * 4];
   __shared__ int radius[1024];
   const int W = a.width, H = a.height, R = a.half;
   const int tx0 = blockIdx.x * DE_TW, ty0 = blockIdx.y * DE_TH;
   const int txe = min(tx0 + DE_TW, W) - 1, tye = min(ty0 + DE_TH, H) - 1;
   for (int e = threadIdx.x; e < DE_TH * DE_TW * 4; e += blockDim.x) tile[e] = (AT)0;
   for (int e = threadIdx.x; e <= a.max_filter_index && e < 1024; e += blockDim.x)
      radius[e] = (int)ceil(a.widths[e]) - 1;
   __syncthreads();
   const AT *coefs = (const AT *)a.coefs_t;
   const AT *srcv = (const AT *)a.srcv;
   const int sx0 = max(0, tx0 - R), sx1 = min(W, tx0 + DE_TW + R);
   const int sy0 = max(0, ty0 - R), sy1 = min(H, ty0 + DE_TH + R);
   const int sw = sx1 - sx0, nsrc = sw * (sy1 - sy0);
   const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
   for (int base = wave * 64; base < nsrc; base += blockDim.x) {
      const int s = base + lane;
      int sx = 0, sy = 0, fi = 0, x0 = 1, x1 = 0, y0 = 1, y1 = 0;
      long cell = 0;
      if (s < nsrc) {
         sx = sx0 + s % sw;
         sy = sy0 + s / sw;
         cell = sx + (long)sy * W;
         const unsigned f = a.findex[cell];
         if (f != DE_SKIP) {
            fi = (int)f;
            const int r = radius[fi];
            x0 = max(sx - r, tx0); x1 = min(sx + r, txe);
            y0 = max(sy - r, ty0); y1 = min(sy + r, tye);
         }
      }
      const int taps = (x0 <= x1 && y0 <= y1) ? (x1 - x0 + 1) * (y1 - y0 + 1) : 0;
      if (taps > 0 && taps <= DE_LANE_TAPS)
         de_splat<AT>(tile, coefs + (long)fi * a.kernel_size, srcv + cell * 4, sx, sy, x0, x1, y0, y1, tx0, ty0, 0,
                      1);
      unsigned long long big = __ballot(taps > DE_LANE_TAPS);
      while (big) {
         const int l = __builtin_ctzll(big);
         big &= big - 1;
         const int bsx = __builtin_amdgcn_readlane(sx, l), bsy = __builtin_amdgcn_readlane(sy, l);
         const int bfi = __builtin_amdgcn_readlane(fi, l);
         const int bx0 = __builtin_amdgcn_readlane(x0, l), bx1 = __builtin_amdgcn_readlane(x1, l);
         const int by0 = __builtin_amdgcn_readlane(y0, l), by1 = __builtin_amdgcn_readlane(y1, l);
         const long bcell = bsx + (long)bsy * W;
         de_splat<AT>(tile, coefs + (long)bfi * a.kernel_size, srcv + bcell * 4, bsx, bsy, bx0, bx1, by0, by1, tx0,
                      ty0, lane, 64);
      }
   }
   __syncthreads();
   AT *acc = (AT *)a.accum;
   for (int e = threadIdx.x; e < DE_TH * DE_TW; e += blockDim.x) {
      const int x = tx0 + e % DE_TW, y = ty0 + e / DE_TW;
      if (x >= W || y >= H) continue;
      AT *d = acc + (x + (long)y * W) * 4;
      for (int k = 0; k < 4; k++) d[k] += tile[k * DE_PLANE + e];
   }
}

// Load-balanced variant: per batch of 512 halo cells, the workgroup lists
// the sources whose kernel reaches the tile and the number of tile cells
// each covers (a block scan), then every thread takes an equal contiguous
// share of the batch's taps and walks them source by source.  No thread
// waits for another's wide kernel, and a tap's cell comes from stepping, not
// dividing.
#define DF_THREADS 512
template <typename AT>
__global__ __launch_bounds__(DF_THREADS) void k_de_flat(DEArgs a) {
   __shared__ AT tile[DE_TH * DE_TW * 4];
   __shared__ int radius[1024];
   __shared__ unsigned pre[DF_THREADS + 1];
   __shared__ unsigned sinfo[DF_THREADS][4];
   __shared__ unsigned wave_tot[DF_THREADS / 64];
   const int W = a.width, H = a.height, R = a.half;
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const int tx0 = blockIdx.x * DE_TW, ty0 = blockIdx.y * DE_TH;
   const int txe = min(tx0 + DE_TW, W) - 1, tye = min(ty0 + DE_TH, H) - 1;
   for (int e = tid; e < DE_TH * DE_TW * 4; e += DF_THREADS) tile[e] = (AT)0;
   for (int e = tid; e <= a.max_filter_index && e < 1024; e += DF_THREADS) radius[e] = (int)ceil(a.widths[e]) - 1;
   if (tid == 0) pre[0] = 0;
   __syncthreads();
   const AT *coefs = (const AT *)a.coefs_t;
   const AT *srcv = (const AT *)a.srcv;
   const int sx0 = max(0, tx0 - R), sx1 = min(W, tx0 + DE_TW + R);
   const int sy0 = max(0, ty0 - R), sy1 = min(H, ty0 + DE_TH + R);
   const int sw = sx1 - sx0, nsrc = sw * (sy1 - sy0);
   for (int base = 0; base < nsrc; base += DF_THREADS) {
      // this thread's halo cell: its kernel's rectangle inside the tile
      const int s = base + tid;
      unsigned taps = 0;
      int sx = 0, sy = 0, fi = 0, x0 = 0, x1 = 0, y0 = 0, y1 = 0;
      if (s < nsrc) {
         sx = sx0 + s % sw;
         sy = sy0 + s / sw;
         const unsigned f = a.findex[sx + (long)sy * W];
         if (f != DE_SKIP) {
            fi = (int)f;
            const int r = radius[fi];
            x0 = max(sx - r, tx0); x1 = min(sx + r, txe);
            y0 = max(sy - r, ty0); y1 = min(sy + r, tye);
            if (x0 <= x1 && y0 <= y1) taps = (unsigned)((x1 - x0 + 1) * (y1 - y0 + 1));
         }
      }
      // inclusive scan of the tap counts over the workgroup
      unsigned v = taps;
      for (int off = 1; off < 64; off <<= 1) {
         const unsigned n = __shfl_up(v, off, 64);
         if (lane >= off) v += n;
      }
      if (lane == 63) wave_tot[wave] = v;
      // coordinates < 65536 (the host checks the histogram size)
      sinfo[tid][0] = (unsigned)sx | ((unsigned)sy << 16);
      sinfo[tid][1] = (unsigned)fi;
      sinfo[tid][2] = (unsigned)x0 | ((unsigned)x1 << 16);
      sinfo[tid][3] = (unsigned)y0 | ((unsigned)y1 << 16);
      __syncthreads();
      for (int w = 0; w < wave; w++) v += wave_tot[w];
      pre[tid + 1] = v;
      __syncthreads();
      const unsigned S = pre[DF_THREADS];
      const unsigned t0 = (unsigned)((unsigned long long)S * tid / DF_THREADS);
      const unsigned t1 = (unsigned)((unsigned long long)S * (tid + 1) / DF_THREADS);
      if (t0 < t1) {
         int lo = 0, hi = DF_THREADS - 1;
         while (lo < hi) {  // last source whose first tap is <= t0
            const int mid = (lo + hi + 1) >> 1;
            if (pre[mid] <= t0) lo = mid;
            else hi = mid - 1;
         }
         int j = lo;
         int csx, csy, cx0, cx1, cy0, cy1;
         const AT *crow;
         AT v0, v1, v2, v3;
         auto load = [&](int jj) {
            const unsigned p0 = sinfo[jj][0], p2 = sinfo[jj][2], p3 = sinfo[jj][3];
            csx = (int)(p0 & 0xffffu); csy = (int)(p0 >> 16);
            cx0 = (int)(p2 & 0xffffu); cx1 = (int)(p2 >> 16);
            cy0 = (int)(p3 & 0xffffu); cy1 = (int)(p3 >> 16);
            crow = coefs + (long)sinfo[jj][1] * a.kernel_size;
            const AT *vv = srcv + (csx + (long)csy * W) * 4;
            v0 = vv[0]; v1 = vv[1]; v2 = vv[2]; v3 = vv[3];
         };
         load(j);
         const int wdt = cx1 - cx0 + 1;
         const unsigned k = t0 - pre[j];
         int y = cy0 + (int)(k / wdt), x = cx0 + (int)(k % wdt);
         for (unsigned t = t0; t < t1; t++) {
            if (t == pre[j + 1]) {
               j++;
               while (pre[j + 1] == pre[j]) j++;
               load(j);
               x = cx0;
               y = cy0;
            }
            const int dx = abs(x - csx), dy = abs(y - csy);
            const int hi2 = max(dx, dy), lo2 = min(dx, dy);
            const AT coef = crow[hi2 * (hi2 + 1) / 2 + lo2];
            if (coef != (AT)0) {
               AT *tp = tile + (y - ty0) * DE_TW + (x - tx0);
               atomicAdd(tp, coef * v0);
               atomicAdd(tp + DE_PLANE, coef * v1);
               atomicAdd(tp + 2 * DE_PLANE, coef * v2);
               atomicAdd(tp + 3 * DE_PLANE, coef * v3);
            }
            if (++x > cx1) {
               x = cx0;
               y++;
            }
         }
      }
      __syncthreads();
   }
   AT *acc = (AT *)a.accum;
   for (int e = tid; e < DE_TH * DE_TW; e += DF_THREADS) {
      const int x = tx0 + e % DE_TW, y = ty0 + e / DE_TW;
      if (x >= W || y >= H) continue;
      AT *d = acc + (x + (long)y * W) * 4;
      for (int c = 0; c < 4; c++) d[c] += tile[c * DE_PLANE + e];
   }
}

// Gather form: each thread owns one output cell and adds, in registers,
// the contribution of every source whose kernel reaches it, sources taken
// in row-major order -- the order in which the reference's single DE thread
// reaches a cell (rect.c:80-173) -- so the sum is deterministic and, in the
// exact mode, formed operation by operation as the reference forms it:
// ls = coef * k1 log(1 + c3 k2) / c3, c *= ls, accumulator += c (float
// accumulator: rounded once per contribution).  A workgroup covers 16x16
// cells (four 8x8 blocks, one per wavefront); it walks its halo in chunks of
// 256 cells, compacts the sources that reach it into LDS in order, and each
// wavefront skips the ones that miss its block.  No atomics.
#define DG_REGION 16
#define DG_CHUNK 256
template <typename AT, bool EXACT>
__global__ __launch_bounds__(DG_CHUNK) void k_de_gather(DEArgs a) {
   __shared__ int radius[1024];
   __shared__ unsigned s_pos[DG_CHUNK];
   __shared__ int s_fi[DG_CHUNK];
   __shared__ double s_val[DG_CHUNK][5];
   __shared__ unsigned wave_cnt[DG_CHUNK / 64];
   const int W = a.width, H = a.height, R = a.half;
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const int rx0 = blockIdx.x * DG_REGION, ry0 = blockIdx.y * DG_REGION;
   const int rx1 = min(rx0 + DG_REGION, W) - 1, ry1 = min(ry0 + DG_REGION, H) - 1;
   const int mx0 = rx0 + 8 * (wave & 1), my0 = ry0 + 8 * (wave >> 1);
   const int x = mx0 + (lane & 7), y = my0 + (lane >> 3);
   const bool live = x < W && y < H;
   AT *accp = (AT *)a.accum + (live ? (x + (long)y * W) * 4 : 0);
   AT acc[4] = {(AT)0, (AT)0, (AT)0, (AT)0};
   if (live)
      for (int c = 0; c < 4; c++) acc[c] = accp[c];
   for (int e = tid; e <= a.max_filter_index && e < 1024; e += DG_CHUNK) radius[e] = (int)ceil(a.widths[e]) - 1;
   __syncthreads();
   const int hx0 = max(0, rx0 - R), hx1 = min(W, rx1 + 1 + R);
   const int hy0 = max(0, ry0 - R), hy1 = min(H, ry1 + 1 + R);
   const int hw = hx1 - hx0, nh = hw * (hy1 - hy0);
   const AT *srcv = (const AT *)a.srcv;
   for (int base = 0; base < nh; base += DG_CHUNK) {
      const int s = base + tid;
      int flag = 0, sx = 0, sy = 0, fi = 0;
      if (s < nh) {
         sx = hx0 + s % hw;
         sy = hy0 + s / hw;
         const unsigned f = a.findex[sx + (long)sy * W];
         if (f != DE_SKIP) {
            fi = (int)f;
            const int r = radius[fi];
            const int gx = max(max(rx0 - sx, sx - rx1), 0), gy = max(max(ry0 - sy, sy - ry1), 0);
            flag = gx <= r && gy <= r;
         }
      }
      const unsigned long long m = __ballot(flag);
      const unsigned rank = __popcll(m & ((1ull << lane) - 1ull));
      if (lane == 0) wave_cnt[wave] = (unsigned)__popcll(m);
      __syncthreads();
      unsigned off = 0, n = 0;
      for (int w = 0; w < DG_CHUNK / 64; w++) {
         if (w < wave) off += wave_cnt[w];
         n += wave_cnt[w];
      }
      if (flag) {
         const unsigned idx = off + rank;
         const long cell = sx + (long)sy * W;
         s_pos[idx] = (unsigned)sx | ((unsigned)sy << 16);
         s_fi[idx] = fi;
         if (EXACT) {
            for (int c = 0; c < 4; c++) s_val[idx][c] = bucket_value(a.buckets, cell * 5 + c, (AT)0);
            s_val[idx][4] = a.srcl[cell];
         } else {
            for (int c = 0; c < 4; c++) s_val[idx][c] = (double)srcv[cell * 4 + c];
         }
      }
      __syncthreads();
      for (unsigned k = 0; k < n; k++) {
         const unsigned pos = s_pos[k];
         const int csx = (int)(pos & 0xffffu), csy = (int)(pos >> 16);
         const int cfi = s_fi[k];
         const int r = radius[cfi];
         if (csx + r < mx0 || csx - r > mx0 + 7 || csy + r < my0 || csy - r > my0 + 7) continue;
         const int dx = abs(x - csx), dy = abs(y - csy);
         const int hi = max(dx, dy), lo = min(dx, dy);
         if (!live || hi > r) continue;
         const long ci = (long)cfi * a.kernel_size + hi * (hi + 1) / 2 + lo;
         if (EXACT) {
            const double coef = a.coefs[ci];
            if (coef == 0.0) continue;
            const double c3 = s_val[k][3];
            const double ls = coef * s_val[k][4] / c3;
            for (int c = 0; c < 4; c++) acc[c] = (AT)((double)acc[c] + s_val[k][c] * ls);
         } else {
            const AT coef = ((const AT *)a.coefs_t)[ci];
            if (coef == (AT)0) continue;
            for (int c = 0; c < 4; c++) acc[c] += coef * (AT)s_val[k][c];
         }
      }
      __syncthreads();
   }
   if (live)
      for (int c = 0; c < 4; c++) accp[c] = acc[c];
}

// Log scaling of every cell without DE (rect.c:947-974).
template <typename BT, typename AT>
__global__ __launch_bounds__(256) void k_logscale(const BT *b, AT *acc, long ncells, double k1, double k2) {
   const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
   if (cell >= ncells) return;
   const double c3 = (double)b[cell * 5 + 3];
   if (c3 == 0.0) return;
   const double ls = (k1 * log(1.0 + c3 * k2)) / c3;
   for (int k = 0; k < 4; k++) {
      const double c = (double)b[cell * 5 + k] * ls;
      AT &d = acc[cell * 4 + k];
      if constexpr (std::is_same<AT, uint32_t>::value) {
         // 32-bit integer accumulator: saturating (flam3.c:3774-3776)
         d = (0xffffffffu - d > c) ? (uint32_t)((double)d + c) : 0xffffffffu;
      } else {
         d = (AT)((double)d + c);
      }
   }
}

// ------------------------------------------------------------------ tone map


__device__ __forceinline__ void tone_pixel(const double *t, double *out_rgb, double &alpha, const TMArgs &a) {
   double ls, newrgb[3];
   const double tmp = t[3] / FL3_PREFILTER_WHITE;
   if (t[3] <= 0) {
      alpha = 0.0;
      ls = 0.0;
   } else {
      alpha = fl3_calc_alpha(tmp, a.g, a.linrange);
      ls = a.vibrancy * 256.0 * alpha / tmp;
      if (alpha < 0.0) alpha = 0.0;
      if (alpha > 1.0) alpha = 1.0;
   }
   fl3_calc_newrgb(t, ls, a.highpow, newrgb);
   for (int k = 0; k < 3; k++) {
      double v = newrgb[k];
      v += (1.0 - a.vibrancy) * 256.0 * pow(t[k] / FL3_PREFILTER_WHITE, a.g);
      if (a.nchan <= 3 || a.transp == 0)
         v += ((1.0 - alpha) * a.bg[k]);
      else
         v = (alpha > 0) ? v / alpha : 0;
      if (v > 255) v = 255;
      if (v < 0) v = 0;
      out_rgb[k] = v;
   }
}

// earlyclip: tone map every accumulator cell in place before filtering
// (rect.c:1086-1135).
template <typename AT>
__global__ __launch_bounds__(256) void k_earlyclip(AT *acc, long ncells, TMArgs a) {
   const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
   if (cell >= ncells) return;
   AT *ac = acc + cell * 4;
   double t[4] = {(double)ac[0], (double)ac[1], (double)ac[2], (double)ac[3]}, rgb[3], alpha;
   tone_pixel(t, rgb, alpha, a);
   for (int k = 0; k < 3; k++) ac[k] = (AT)rgb[k];
   ac[3] = (AT)alpha;
}

template <typename AT>
__global__ __launch_bounds__(256) void k_tonemap(TMArgs a) {
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   const int j = blockIdx.y;
   if (i >= a.iw || j >= a.ih) return;
   const AT *acc = (const AT *)a.accum;
   const int x = a.de_offset + i * a.os, y = a.de_offset + j * a.os;
   double t[4] = {0.0, 0.0, 0.0, 0.0};
   for (int ii = 0; ii < a.fw; ii++)
      for (int jj = 0; jj < a.fw; jj++) {
         const double k = a.filter[ii + jj * a.fw];
         const AT *ac = acc + ((x + ii) + (long)(y + jj) * a.W) * 4;
         t[0] += k * (double)ac[0];
         t[1] += k * (double)ac[1];
         t[2] += k * (double)ac[2];
         t[3] += k * (double)ac[3];
      }
   if (!a.earlyclip) {
      double rgb[3], alpha;
      tone_pixel(t, rgb, alpha, a);
      t[0] = rgb[0]; t[1] = rgb[1]; t[2] = rgb[2];
      t[3] = alpha;
   }
   unsigned char *p = a.out + (size_t)a.nchan * a.bpc * ((size_t)i + (size_t)j * a.iw);
   for (int k = 0; k < 3; k++) {
      double v = t[k];
      if (v > 255) v = 255;
      if (v < 0) v = 0;
      if (a.bpc == 2)
         ((unsigned short *)p)[k] = (unsigned short)(v * 256.0);
      else
         p[k] = (unsigned char)v;
   }
   if (t[3] > 1) t[3] = 1;
   if (t[3] < 0) t[3] = 0;
   if (a.nchan > 3) {
      if (a.transp == 1) {
         if (a.bpc == 2) ((unsigned short *)p)[3] = (unsigned short)(t[3] * 65535);
         else p[3] = (unsigned char)(t[3] * 255);
      } else {
         if (a.bpc == 2) ((unsigned short *)p)[3] = 65535;
         else p[3] = 255;
      }
   }
}

// ------------------------------------------------------------------ launchers

extern "C" int fl3k_launch_replay(const IterArgs *a, unsigned grid, size_t lds, hipStream_t s, int bits) {
   if (bits == 64)
      hipLaunchKernelGGL((k_iter_replay<64>), dim3(grid), dim3(64), lds, s, *a);
   else
      hipLaunchKernelGGL((k_iter_replay<33>), dim3(grid), dim3(64), lds, s, *a);
   return hipGetLastError() != hipSuccess;
}

extern "C" int fl3k_launch_points(const PointsArgs *a, unsigned grid, size_t lds, hipStream_t s) {
   hipLaunchKernelGGL(k_iterate_points, dim3(grid), dim3(64), lds, s, *a);
   return hipGetLastError() != hipSuccess;
}

extern "C" int fl3k_launch_preview(const unsigned char *job, int xi, double range, int numvals, int depth,
                                   double *result, uint32_t *isaac, size_t lds, hipStream_t s) {
   hipLaunchKernelGGL(k_xform_preview, dim3(1), dim3(64), lds, s, job, xi, range, numvals, depth, result, isaac);
   return hipGetLastError() != hipSuccess;
}

extern "C" int fl3k_launch_deposit(const unsigned char *job, const double *pts, long long n, void *buckets,
                                   uint32_t *sat, unsigned grid, size_t lds, hipStream_t s, int bits) {
   if (bits == 64)
      hipLaunchKernelGGL(k_deposit_points<64>, dim3(grid), dim3(256), lds, s, job, pts, n, buckets, sat);
   else
      hipLaunchKernelGGL(k_deposit_points<33>, dim3(grid), dim3(256), lds, s, job, pts, n, buckets, sat);
   return hipGetLastError() != hipSuccess;
}

extern "C" int fl3k_launch_de_tiled(const DEArgs *a, hipStream_t s, int bits) {
   const dim3 g1((a->width + 255) / 256, a->height), b1(256);
   const dim3 g2((a->width + DE_TW - 1) / DE_TW, (a->height + DE_TH - 1) / DE_TH), b2(256);
   const dim3 g3((a->width + DG_REGION - 1) / DG_REGION, (a->height + DG_REGION - 1) / DG_REGION);
   if (bits == 64) {
      hipLaunchKernelGGL((k_de_prep<double, double>), g1, b1, 0, s, *a);
      hipLaunchKernelGGL((k_de_gather<double, true>), g3, dim3(DG_CHUNK), 0, s, *a);
   } else {
      hipLaunchKernelGGL((k_de_prep<uint32_t, float>), g1, b1, 0, s, *a);
      if (a->exact)
         hipLaunchKernelGGL((k_de_gather<float, true>), g3, dim3(DG_CHUNK), 0, s, *a);
      else
         hipLaunchKernelGGL((k_de_gather<float, false>), g3, dim3(DG_CHUNK), 0, s, *a);
   }
   return hipGetLastError() != hipSuccess;
}

extern "C" int fl3k_launch_logscale(const void *b, void *acc, long n, double k1, double k2, hipStream_t s,
                                    int bits) {
   const dim3 grid((unsigned)((n + 255) / 256)), block(256);
   if (bits == 64)
      hipLaunchKernelGGL((k_logscale<double, double>), grid, block, 0, s, (const double *)b, (double *)acc, n, k1, k2);
   else if (bits == 33)
      hipLaunchKernelGGL((k_logscale<uint32_t, float>), grid, block, 0, s, (const uint32_t *)b, (float *)acc, n, k1,
                         k2);
   else
      hipLaunchKernelGGL((k_logscale<uint32_t, uint32_t>), grid, block, 0, s, (const uint32_t *)b, (uint32_t *)acc,
                         n, k1, k2);
   return hipGetLastError() != hipSuccess;
}

extern "C" int fl3k_launch_earlyclip(void *acc, long n, const TMArgs *a, hipStream_t s, int bits) {
   const dim3 grid((unsigned)((n + 255) / 256)), block(256);
   if (bits == 64)
      hipLaunchKernelGGL((k_earlyclip<double>), grid, block, 0, s, (double *)acc, n, *a);
   else if (bits == 33)
      hipLaunchKernelGGL((k_earlyclip<float>), grid, block, 0, s, (float *)acc, n, *a);
   else
      hipLaunchKernelGGL((k_earlyclip<uint32_t>), grid, block, 0, s, (uint32_t *)acc, n, *a);
   return hipGetLastError() != hipSuccess;
}

extern "C" int fl3k_launch_tonemap(const TMArgs *a, hipStream_t s, int bits) {
   const dim3 grid((a->iw + 255) / 256, a->ih), block(256);
   if (bits == 64)
      hipLaunchKernelGGL((k_tonemap<double>), grid, block, 0, s, *a);
   else if (bits == 33)
      hipLaunchKernelGGL((k_tonemap<float>), grid, block, 0, s, *a);
   else
      hipLaunchKernelGGL((k_tonemap<uint32_t>), grid, block, 0, s, *a);
   return hipGetLastError() != hipSuccess;
}
