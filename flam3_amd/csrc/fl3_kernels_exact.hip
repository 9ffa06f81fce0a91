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


// ---- density estimation ---------------------------------------------------
//
// The reference scatters each source cell's kernel into the accumulator
// (rect.c:126-173).  Here pass 1 (k_de_prep) picks every source's kernel (the
// hit-count neighbourhood sum and index selection of rect.c:92-117) and pass
// 2 (k_de_gather) gathers, for each output cell, the contributions of the
// sources that reach it.
#define DE_SKIP 0xffffu

// Channel i of the histogram as a double: uint32 cells with a float
// accumulator (bits 33), double cells with a double one (bits 64).
__device__ __forceinline__ double bucket_value(const void *b, long i, float) { return (double)((const uint32_t *)b)[i]; }
__device__ __forceinline__ double bucket_value(const void *b, long i, double) { return ((const double *)b)[i]; }

// Per 16x16 block of cells (one DE workgroup region): 1 + the largest
// kernel radius of its sources (0: no source), so that each gather
// workgroup scans only the halo its neighbourhood's sources can reach from.
// Lanes 16k..16k+15 of a wavefront hold cells of one block (rows of 256
// cells, blocks 16 wide).
__device__ __forceinline__ void de_mark_block(const DEArgs &a, int i, int j, int r1) {
   int m = r1;
   for (int o = 1; o < 16; o <<= 1) m = max(m, __shfl_xor(m, o));
   if ((threadIdx.x & 15) == 0 && m > 0) atomicMax(&a.blkmax[(j >> 4) * a.nbx + (i >> 4)], m);
}

// One cell of the DE preparation: its kernel index (DE_SKIP if it is no
// source), per-source values and log term; returns the index.
template <typename BT, typename AT>
__device__ __forceinline__ unsigned de_prep_cell(const DEArgs &a, int i, int j, int os, int ss, bool scf,
                                                 double scfact) {
   const int W = a.width, H = a.height;
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
         if (a.exact) a.srcl[cell] = lnum;
         const double sc = lnum / c3;
         AT *v = (AT *)a.srcv + cell * 4;
         v[0] = (AT)((double)b[cell * 5 + 0] * sc);
         v[1] = (AT)((double)b[cell * 5 + 1] * sc);
         v[2] = (AT)((double)b[cell * 5 + 2] * sc);
         v[3] = (AT)(c3 * sc);
      }
   }
   a.findex[cell] = out;
   return out;
}

template <typename BT, typename AT>
__device__ __forceinline__ void de_prep_body(const DEArgs &a, int bx, int by) {
   const int os = a.oversample;
   const int ss = os / 2;
   const bool scf = !(os & 1);
   const double scfact = pow(os / (os + 1.0), 2.0);
   const int W = a.width;
   const int i = bx * blockDim.x + threadIdx.x;
   const int j = by;
   if (!a.blkmax && !a.taps) {
      if (i < W) de_prep_cell<BT, AT>(a, i, j, os, ss, scf, scfact);
      return;
   }
   // (every lane takes part in the block and wavefront reductions)
   int r1 = 0;
   if (i < W) {
      const unsigned f = de_prep_cell<BT, AT>(a, i, j, os, ss, scf, scfact);
      if (f != DE_SKIP) r1 = (int)ceil(a.widths[f]);
   }
   if (a.blkmax) de_mark_block(a, i < W ? i : W - 1, j, r1);
   if (a.taps) {
      unsigned long long t = r1 > 0 ? (unsigned long long)(2 * r1 - 1) * (unsigned long long)(2 * r1 - 1) : 0ull;
      for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
      // (64 counters, so that the wavefronts' atomics do not all meet at one address)
      if ((threadIdx.x & 63) == 0 && t) atomicAdd(a.taps + ((blockIdx.x * 4u + (threadIdx.x >> 6)) & 63u), t);
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
// 256 cells, compacts the sources that reach it into LDS in order, and lists
// for each wavefront, in the same order, those that reach its block, so a
// wavefront loops over its own sources only.  No atomics.
#define DG_REGION 16
#define DG_CHUNK 256
template <typename AT, bool EXACT>
__device__ __forceinline__ void de_gather_body(const DEArgs &a, int bx, int by) {
   // per source: packed position, kernel index and values -- doubles
   // (4 channels and the log term) in the exact mode, the accumulator type
   // (4 pre-scaled channels) otherwise
   typedef typename std::conditional<EXACT, double, AT>::type VT;
   constexpr int NV = EXACT ? 5 : 4;
   __shared__ int radius[1024];
   __shared__ int4 s_src[DG_CHUNK];   // per source: x, y, first coefficient of its kernel, radius
   __shared__ VT s_val[DG_CHUNK][NV];
   __shared__ unsigned wave_cnt[DG_CHUNK / 64];
   __shared__ unsigned short s_list[4][DG_CHUNK];   // per 8x8 block: indices into the chunk's sources
   __shared__ unsigned list_cnt[4][DG_CHUNK / 64];  // [block][wavefront that listed them]
   const int W = a.width, H = a.height, R = a.half;
   const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
   const int rx0 = bx * DG_REGION, ry0 = by * DG_REGION;
   const int rx1 = min(rx0 + DG_REGION, W) - 1, ry1 = min(ry0 + DG_REGION, H) - 1;
   const int mx0 = rx0 + 8 * (wave & 1), my0 = ry0 + 8 * (wave >> 1);
   const int x = mx0 + (lane & 7), y = my0 + (lane >> 3);
   const bool live = x < W && y < H;
   AT *accp = (AT *)a.accum + (live ? (x + (long)y * W) * 4 : 0);
   AT acc[4] = {(AT)0, (AT)0, (AT)0, (AT)0};
   if (live)
      for (int c = 0; c < 4; c++) acc[c] = accp[c];
   for (int e = tid; e <= a.max_filter_index && e < 1024; e += DG_CHUNK) radius[e] = (int)ceil(a.widths[e]) - 1;
   // The halo this region needs: the largest radius among the 16x16 blocks
   // whose widest source can reach it (-1: none can), from the preparation
   // pass's per-block maxima; without them the widest kernel's.  A smaller
   // halo holds the same reaching sources in the same row-major order.
   __shared__ int s_ext;
   int E = R;
   if (a.blkmax) {
      if (tid == 0) s_ext = -1;
      __syncthreads();
      const int nbx = a.nbx, nby = (H + DG_REGION - 1) / DG_REGION;
      const int span = (R + DG_REGION - 1) / DG_REGION, side = 2 * span + 1;
      for (int k = tid; k < side * side; k += DG_CHUNK) {
         const int qx = bx + k % side - span, qy = by + k / side - span;
         if (qx < 0 || qy < 0 || qx >= nbx || qy >= nby) continue;
         const int bm = a.blkmax[qy * nbx + qx] - 1;
         const int qx0 = qx * DG_REGION, qy0 = qy * DG_REGION;
         const int gx = max(max(rx0 - (qx0 + DG_REGION - 1), qx0 - rx1), 0);
         const int gy = max(max(ry0 - (qy0 + DG_REGION - 1), qy0 - ry1), 0);
         if (bm >= 0 && gx <= bm && gy <= bm) atomicMax(&s_ext, min(bm, R));
      }
   }
   // the cut filter bank (float accumulator), for the taps below
   float *s_coef = (float *)fl3_lds;
   constexpr bool CUT = !EXACT && std::is_same<AT, float>::value;
   if (CUT && a.ncoef > 0)
      for (int e = tid; e < a.ncoef; e += DG_CHUNK) s_coef[e] = a.ccoef[e];
   __syncthreads();
   if (a.blkmax) E = s_ext;
   const int hx0 = max(0, rx0 - E), hx1 = min(W, rx1 + 1 + E);
   const int hy0 = max(0, ry0 - E), hy1 = min(H, ry1 + 1 + E);
   const int hw = hx1 - hx0, nh = E < 0 ? 0 : hw * (hy1 - hy0);
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
      unsigned off = 0;
      for (int w = 0; w < wave; w++) off += wave_cnt[w];
      // the blocks (of the four 8x8 ones) this source reaches
      unsigned reach = 0;
      if (flag) {
         const int r = radius[fi];
         for (int b = 0; b < 4; b++) {
            const int bx0 = rx0 + 8 * (b & 1), by0 = ry0 + 8 * (b >> 1);
            reach |= (unsigned)(sx + r >= bx0 && sx - r <= bx0 + 7 && sy + r >= by0 && sy - r <= by0 + 7) << b;
         }
      }
      unsigned brank[4];
      for (int b = 0; b < 4; b++) {
         const unsigned long long mb = __ballot((reach >> b) & 1u);
         brank[b] = (unsigned)__popcll(mb & ((1ull << lane) - 1ull));
         if (lane == 0) list_cnt[b][wave] = (unsigned)__popcll(mb);
      }
      if (flag) {
         const unsigned idx = off + rank;
         const long cell = sx + (long)sy * W;
         if constexpr (EXACT)
            s_src[idx] = make_int4(sx, sy, fi * a.kernel_size, radius[fi]);
         else   // (as floats: every value here is an integer below 2^24)
            s_src[idx] = make_int4(__float_as_int((float)sx), __float_as_int((float)sy),
                                   __float_as_int((float)(CUT && a.ncoef > 0 ? a.coff[fi] : fi * a.kernel_size)),
                                   __float_as_int((float)radius[fi]));
         if constexpr (EXACT) {
            for (int c = 0; c < 4; c++) s_val[idx][c] = bucket_value(a.buckets, cell * 5 + c, (AT)0);
            s_val[idx][4] = a.srcl[cell];
         } else {
            for (int c = 0; c < 4; c++) s_val[idx][c] = srcv[cell * 4 + c];
         }
      }
      __syncthreads();
      unsigned nb = 0;
      for (int b = 0; b < 4; b++) {
         unsigned boff = 0;
         for (int w = 0; w < DG_CHUNK / 64; w++)
            if (w < wave) boff += list_cnt[b][w];
         if ((reach >> b) & 1u) s_list[b][boff + brank[b]] = (unsigned short)(off + rank);
      }
      for (int w = 0; w < DG_CHUNK / 64; w++) nb += list_cnt[wave][w];
      __syncthreads();
      if constexpr (EXACT) {
         // (the coefficient is loaded for every listed source, from index 0
         // when the tap misses this cell, so that the loads of consecutive
         // sources overlap; a missed tap adds nothing)
#pragma unroll 4
         for (unsigned i = 0; i < nb; i++) {
            const unsigned k = s_list[wave][i];
            const int4 src = s_src[k];
            const int dx = abs(x - src.x), dy = abs(y - src.y);
            const int hi = max(dx, dy), lo = min(dx, dy);
            const bool tap = live && hi <= src.w;
            const long ci = tap ? (long)src.z + hi * (hi + 1) / 2 + lo : 0;
            const double coef = a.coefs[ci];
            if (tap && coef != 0.0) {
               const double c3 = s_val[k][3];
               const double ls = coef * s_val[k][4] / c3;
               for (int c = 0; c < 4; c++) acc[c] = (AT)((double)acc[c] + s_val[k][c] * ls);
            }
         }
      } else {
         // Straight-line taps, no branches, so that the coefficient loads of
         // consecutive sources are in flight together: a tap that misses
         // this cell loads index 0 and adds it times 0 (the per-source
         // values are finite), in 32-bit index arithmetic, fused
         // multiply-adds.
         // (distances in float, exact for these integers below 2^24, so
         // that |.| comes free with the max and min; the tap's index
         // z + hi (hi + 1) / 2 + lo by exact fused multiply-adds; a missed
         // tap's index cleared by a mask, so that the body has no branch)
         // (coefficients from the cut bank in LDS when it is there, else
         // from the full table: one loop each, no per-tap choice)
         const AT *ct = (const AT *)a.coefs_t;
         const float fx = (float)x, fy = (float)y;
         auto taps = [&](auto coef_at) {
#pragma unroll 4
            for (unsigned i = 0; i < nb; i++) {
               const unsigned k = s_list[wave][i];
               const int4 src = s_src[k];
               const float ax = fx - __int_as_float(src.x), ay = fy - __int_as_float(src.y);
               const float hi = fmaxf(fabsf(ax), fabsf(ay)), lo = fminf(fabsf(ax), fabsf(ay));
               const bool tap = live && hi <= __int_as_float(src.w);
               const float cf = fmaf(fmaf(hi, hi, hi), 0.5f, lo) + __int_as_float(src.z);
               const unsigned ci = (unsigned)cf & (0u - (unsigned)tap);
               const AT cv = coef_at(ci);
               const AT coef = tap ? cv : (AT)0;
               for (int c = 0; c < 4; c++) acc[c] = fma(coef, s_val[k][c], acc[c]);
            }
         };
         if (CUT && a.ncoef > 0)
            taps([&](unsigned c) { return (AT)s_coef[c]; });
         else
            taps([&](unsigned c) { return ct[c]; });
      }
      __syncthreads();
   }
   if (live)
      for (int c = 0; c < 4; c++) accp[c] = acc[c];
}

__global__ __launch_bounds__(256) void k_fill_i32(int *p, long n, int v) {
   const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) p[i] = v;
}

template <typename BT, typename AT>
__global__ __launch_bounds__(256) void k_de_prep(DEArgs a) {
   de_prep_body<BT, AT>(a, blockIdx.x, blockIdx.y);
}

template <typename AT, bool EXACT>
__global__ __launch_bounds__(DG_CHUNK) void k_de_gather(DEArgs a) {
   de_gather_body<AT, EXACT>(a, blockIdx.x, blockIdx.y);
}

// Several renders in one launch: workgroup b belongs to the render j with
// first[j] <= b < first[j + 1] and is that render's workgroup b - first[j]
// of a grid `cols` wide.
__device__ __forceinline__ int de_job(const unsigned *first, int n, unsigned b) {
   int lo = 0, hi = n - 1;
   while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (first[mid] <= b) lo = mid;
      else hi = mid - 1;
   }
   return lo;
}

template <typename BT, typename AT>
__global__ __launch_bounds__(256) void k_de_prep_many(const DEArgs *as, const unsigned *first, int n) {
   const int j = de_job(first, n, blockIdx.x);
   const DEArgs a = as[j];
   const unsigned local = blockIdx.x - first[j], cols = (unsigned)(a.width + 255) / 256;
   de_prep_body<BT, AT>(a, (int)(local % cols), (int)(local / cols));
}

template <typename AT, bool EXACT>
__global__ __launch_bounds__(DG_CHUNK) void k_de_gather_many(const DEArgs *as, const unsigned *first, int n) {
   const int j = de_job(first, n, blockIdx.x);
   const DEArgs a = as[j];
   const unsigned local = blockIdx.x - first[j], cols = (unsigned)(a.width + DG_REGION - 1) / DG_REGION;
   de_gather_body<AT, EXACT>(a, (int)(local % cols), (int)(local / cols));
}

// Density estimation of n renders: per-render arguments in device memory,
// workgroup prefixes prep_first[n + 1] (rows of 256 cells) and
// gather_first[n + 1] (16x16 regions).  bits 32/33 with the FP32
// accumulator (exact = 0) only.
// blkmax: a device array of nblk ints the renders' blkmax pointers point
// into (cleared here), or nullptr.
extern "C" int fl3k_launch_de_many(const DEArgs *as, const unsigned *prep_first, unsigned prep_blocks,
                                   const unsigned *gather_first, unsigned gather_blocks, int n, hipStream_t s,
                                   int *blkmax, long nblk, size_t gather_lds) {
   if (n <= 0) return 0;
   if (blkmax)
      hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((nblk + 255) / 256)), dim3(256), 0, s, blkmax, nblk, 0);
   hipLaunchKernelGGL((k_de_prep_many<uint32_t, float>), dim3(prep_blocks), dim3(256), 0, s, as, prep_first, n);
   hipLaunchKernelGGL((k_de_gather_many<float, false>), dim3(gather_blocks), dim3(DG_CHUNK), gather_lds, s, as,
                      gather_first, n);
   return hipGetLastError() != hipSuccess;
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
__device__ __forceinline__ void tonemap_body(const TMArgs &a, int bx, int by) {
   const int i = bx * blockDim.x + threadIdx.x;
   const int j = by;
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
   const dim3 g3((a->width + DG_REGION - 1) / DG_REGION, (a->height + DG_REGION - 1) / DG_REGION);
   if (a->blkmax) {   // per-block radii, cleared here and marked by the preparation pass
      const long nblk = (long)g3.x * g3.y;
      hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((nblk + 255) / 256)), dim3(256), 0, s, a->blkmax, nblk, 0);
   }
   if (bits == 64) {
      hipLaunchKernelGGL((k_de_prep<double, double>), g1, b1, 0, s, *a);
      hipLaunchKernelGGL((k_de_gather<double, true>), g3, dim3(DG_CHUNK), 0, s, *a);
   } else {
      hipLaunchKernelGGL((k_de_prep<uint32_t, float>), g1, b1, 0, s, *a);
      if (a->exact)
         hipLaunchKernelGGL((k_de_gather<float, true>), g3, dim3(DG_CHUNK), 0, s, *a);
      else
         hipLaunchKernelGGL((k_de_gather<float, false>), g3, dim3(DG_CHUNK),
                            a->ncoef > 0 ? sizeof(float) * (size_t)a->ncoef : 0, s, *a);
   }
   return hipGetLastError() != hipSuccess;
}

// Double to float copy of a table on the device (the density-estimation
// coefficients of the FP32 accumulator path).
__global__ __launch_bounds__(256) void k_f64_to_f32(const double *src, float *dst, long n) {
   const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n) dst[i] = (float)src[i];
}

extern "C" int fl3k_launch_f64_to_f32(const double *src, float *dst, long n, hipStream_t s) {
   if (n <= 0) return 0;
   hipLaunchKernelGGL(k_f64_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
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

template <typename AT>
__global__ __launch_bounds__(256) void k_tonemap(TMArgs a) {
   tonemap_body<AT>(a, blockIdx.x, blockIdx.y);
}

// Tone maps of several renders in one launch (workgroup prefixes as for
// k_de_gather_many; rows of 256 pixels).
template <typename AT>
__global__ __launch_bounds__(256) void k_tonemap_many(const TMArgs *as, const unsigned *first, int n) {
   const int j = de_job(first, n, blockIdx.x);
   const TMArgs a = as[j];
   const unsigned local = blockIdx.x - first[j], cols = (unsigned)(a.iw + 255) / 256;
   tonemap_body<AT>(a, (int)(local % cols), (int)(local / cols));
}

extern "C" int fl3k_launch_tonemap_many(const TMArgs *as, const unsigned *first, unsigned blocks, int n,
                                        hipStream_t s, int bits) {
   if (n <= 0 || blocks == 0) return 0;
   if (bits == 64)
      hipLaunchKernelGGL((k_tonemap_many<double>), dim3(blocks), dim3(256), 0, s, as, first, n);
   else if (bits == 33)
      hipLaunchKernelGGL((k_tonemap_many<float>), dim3(blocks), dim3(256), 0, s, as, first, n);
   else
      hipLaunchKernelGGL((k_tonemap_many<uint32_t>), dim3(blocks), dim3(256), 0, s, as, first, n);
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
