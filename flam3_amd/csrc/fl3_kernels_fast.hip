// fl3_kernels_fast.hip - FP32 production walkers (compiled with FP
// contraction on).
#include "fl3_walker.h"
#include "fl3_launch.h"


template <bool WAVE>
static void launch_full(const IterArgs *a, unsigned grid, size_t lds, hipStream_t s, int bits, int exact) {
   if (a->rec_mode == FL3_REC_JOB)
      hipLaunchKernelGGL((k_iter_fast<33, true, FL3_REC_JOB, FL3_TIER_FULL, WAVE>), dim3(grid), dim3(256), lds, s, *a);
   else if (a->rec_mode == FL3_REC_NARROW)
      hipLaunchKernelGGL((k_iter_fast<33, true, FL3_REC_NARROW, FL3_TIER_FULL, WAVE>), dim3(grid), dim3(256), lds, s,
                         *a);
   else if (a->rec_mode == FL3_REC_WIDE)
      hipLaunchKernelGGL((k_iter_fast<33, true, FL3_REC_WIDE, FL3_TIER_FULL, WAVE>), dim3(grid), dim3(256), lds, s, *a);
   else if (bits == 64)
      hipLaunchKernelGGL((k_iter_fast<64, true, FL3_REC_DIRECT, FL3_TIER_FULL, WAVE>), dim3(grid), dim3(256), lds, s,
                         *a);
   else if (exact)
      hipLaunchKernelGGL((k_iter_fast<33, true, FL3_REC_DIRECT, FL3_TIER_FULL, WAVE>), dim3(grid), dim3(256), lds, s,
                         *a);
   else
      hipLaunchKernelGGL((k_iter_fast<33, false, FL3_REC_DIRECT, FL3_TIER_FULL, WAVE>), dim3(grid), dim3(256), lds, s,
                         *a);
}

// a->wave_xf selects the one-xform-per-wavefront walkers (genomes without
// xaos, exchange buffers at a->shuf_off).
extern "C" int fl3k_launch_iter_fast(const IterArgs *a, unsigned grid, size_t lds, hipStream_t s, int bits,
                                     int exact, int tier) {
   if (tier >= 0 && tier < FL3_TIER_FULL && (a->rec_mode || bits != 64))
      return fl3k_launch_iter_tier(a, grid, lds, s, tier, exact);
   if (a->wave_xf)
      launch_full<true>(a, grid, lds, s, bits, exact);
   else
      launch_full<false>(a, grid, lds, s, bits, exact);
   return hipGetLastError() != hipSuccess;
}

extern "C" int fl3k_launch_apply_fp32(const unsigned char *job, int xi, const double *pts, long long n, double *out,
                                      unsigned long long key, size_t lds, hipStream_t s, int tier) {
   if (tier >= 0 && tier < FL3_TIER_FULL) return fl3k_launch_apply_fp32_tier(job, xi, pts, n, out, key, lds, s, tier);
   const unsigned grid = (unsigned)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
   hipLaunchKernelGGL(k_apply_fp32<FL3_TIER_FULL>, dim3(grid), dim3(256), lds, s, job, xi, pts, n, out, key);
   return hipGetLastError() != hipSuccess;
}

// ------------------------------------------------------------------ binned deposit
//
// The walker kernel writes one record per plotted sample instead of five
// global atomics, and counts its workgroup's records per tile of 1 << shift
// cells.  The records are then grouped by tile (a counting sort: per-tile
// sums and per-workgroup offsets, tile offsets, scatter), and each tile's
// records are accumulated in LDS and added to the histogram once.  Integer
// sums commute, so the histogram is the one the atomic deposit produces; the
// tile flush applies the reference's saturation exactly.

// Launch shape of the tile pipeline: threads per accumulation workgroup and
// records each thread loads before working on them (accumulation, scatter).
#ifndef FL3_ACC_THREADS
#define FL3_ACC_THREADS 1024
#endif
#ifndef FL3_ACC_UNROLL
#define FL3_ACC_UNROLL 16
#endif

// One workgroup: exclusive scans of the tile counts and of the work items
// (ceil(count / chunk) per tile).  It also leaves the tile counts at zero
// for their next use and resets the accumulation's item counter.
__global__ __launch_bounds__(1024) void k_bin_scan(BinArgs a) {
   __shared__ unsigned long long part[1024], ipart[1024];
   const int t = threadIdx.x, per = (a.nbins + 1023) / 1024;
   if (t == 0) *a.item_counter = 0u;
   const int b0 = t * per, b1 = min(a.nbins, b0 + per);
   unsigned long long s = 0, si = 0;
   for (int b = b0; b < b1; b++) {
      s += a.bin_total[b];
      si += (a.bin_total[b] + a.chunk - 1) / a.chunk;
   }
   part[t] = s;
   ipart[t] = si;
   __syncthreads();
   for (int off = 1; off < 1024; off <<= 1) {
      unsigned long long v = t >= off ? part[t - off] : 0, vi = t >= off ? ipart[t - off] : 0;
      __syncthreads();
      part[t] += v;
      ipart[t] += vi;
      __syncthreads();
   }
   unsigned long long run = t ? part[t - 1] : 0, irun = t ? ipart[t - 1] : 0;
   for (int b = b0; b < b1; b++) {
      a.bin_start[b] = (unsigned)run;
      a.cursor[b] = (unsigned)run;
      a.item_start[b] = (unsigned)irun;
      run += a.bin_total[b];
      irun += (a.bin_total[b] + a.chunk - 1) / a.chunk;
      a.bin_total[b] = 0u;
   }
   if (t == 1023) {
      a.bin_start[a.nbins] = (unsigned)part[1023];
      a.item_start[a.nbins] = (unsigned)ipart[1023];
   }
}

template <typename RT>
__device__ __forceinline__ unsigned rec_cell(RT v) {
   return sizeof(RT) == 4 ? (unsigned)v & 0xffffffu : (unsigned)v;
}
template <typename RT>
__device__ __forceinline__ bool rec_none(RT v) {
   return (unsigned)v == FL3_REC_NONE;
}
// the record's cell relative to the band being sorted (records of a band
// pass hold that already)
template <typename RT>
__device__ __forceinline__ unsigned rel_cell(const BinArgs &a, RT v) {
   return a.rec_rel ? rec_cell(v) : rec_cell(v) - a.cell_base;
}
// slot idx of the input blocks: past rec_count (when set) there are no records
template <typename RT>
__device__ __forceinline__ RT rec_at(const BinArgs &a, const RT *rec, long long idx) {
   return a.rec_count == 0 || idx < a.rec_count ? rec[idx] : (RT)FL3_REC_NONE;
}

// First slot of each walker workgroup's records of each super-tile in the
// super-tile-grouped array: workgroup s scans super-tile s's counts over the
// walker workgroups.
__global__ __launch_bounds__(256) void k_super_off(BinArgs a) {
   __shared__ unsigned part[256];
   const int s = blockIdx.x, t = threadIdx.x;
   const int b0 = s << FL3_SUPER_SHIFT;
   const int per = (a.nblk + 255) / 256;
   const int k0 = t * per, k1 = min(a.nblk, k0 + per);
   unsigned sum = 0;
   for (int k = k0; k < k1; k++) sum += a.blk_hist[(size_t)k * a.nsuper + s];
   part[t] = sum;
   __syncthreads();
   for (int off = 1; off < 256; off <<= 1) {
      const unsigned v = t >= off ? part[t - off] : 0;
      __syncthreads();
      part[t] += v;
      __syncthreads();
   }
   unsigned run = a.bin_start[b0] + (t ? part[t - 1] : 0);
   for (int k = k0; k < k1; k++) {
      a.soff[(size_t)k * a.nsuper + s] = run;
      run += a.blk_hist[(size_t)k * a.nsuper + s];
   }
}

// Exclusive prefix sum over a 256-thread workgroup (v from each thread);
// *total gets the sum.  wsum: 4 words of LDS.  Ends with a barrier.
__device__ __forceinline__ unsigned wg_excl_scan256(unsigned v, unsigned *wsum, unsigned *total) {
   const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
   unsigned x = v;
   for (int o = 1; o < 64; o <<= 1) {
      const unsigned y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
   }
   if (lane == 63) wsum[w] = x;
   __syncthreads();
   unsigned pre = 0;
   for (int k = 0; k < w; k++) pre += wsum[k];
   *total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
   __syncthreads();
   return pre + x - v;
}

// Pass 1: workgroup k moves the records of walkers [256k, 256k + 256) (one
// column block of the record matrix) to its range of each super-tile.  With
// at most 256 super-tiles per workgroup the partly written lines stay in L2
// until they fill.
#define FL3_SCATTER_UNROLL 16
// The record as pass 1 writes it: unchanged, or (OT 4 bytes from 8-byte
// records, the band pass of a histogram of 2^24-cell bands, visibility 1
// and a stepped palette) its cell within the super-tile | palette index << 24.
template <typename RT, typename OT>
__device__ __forceinline__ OT part1_out(RT v, int ss) {
   if (sizeof(OT) == sizeof(RT)) return (OT)v;
   const unsigned long long x = (unsigned long long)v;
   return (OT)(((unsigned)x & ((1u << ss) - 1u)) | (((unsigned)(x >> 32) & 0xffu) << 24));
}

template <typename RT, typename OT = RT>
__global__ __launch_bounds__(256) void k_bin_part1(BinArgs a) {
   __shared__ unsigned base[256], cnt[256];
   const int ss = a.shift + FL3_SUPER_SHIFT;
   for (int s = threadIdx.x; s < a.nsuper; s += blockDim.x) {
      base[s] = a.soff[(size_t)blockIdx.x * a.nsuper + s];
      cnt[s] = 0;
   }
   __syncthreads();
   const RT *rec = (const RT *)a.rec;
   OT *out = (OT *)a.sorted;
   long long i = 0;
   for (; i + FL3_SCATTER_UNROLL <= a.rows; i += FL3_SCATTER_UNROLL) {
      RT v[FL3_SCATTER_UNROLL];
      for (int k = 0; k < FL3_SCATTER_UNROLL; k++) v[k] = rec_at(a, rec, (blockIdx.x * a.rows + i + k) * 256 + threadIdx.x);
      for (int k = 0; k < FL3_SCATTER_UNROLL; k++) {
         if (rec_none(v[k])) continue;
         const unsigned s = rel_cell(a, v[k]) >> ss;
         out[base[s] + atomicAdd(&cnt[s], 1u)] = part1_out<RT, OT>(v[k], ss);
      }
   }
   for (; i < a.rows; i++) {
      const RT v = rec_at(a, rec, (blockIdx.x * a.rows + i) * 256 + threadIdx.x);
      if (rec_none(v)) continue;
      const unsigned s = rel_cell(a, v) >> ss;
      out[base[s] + atomicAdd(&cnt[s], 1u)] = part1_out<RT, OT>(v, ss);
   }
}

// The same for 4-byte records of tall blocks (many rows per chunk): the
// block is taken in batches of 16 rows, each batch ordered by super-tile in
// LDS (counting sort) and written out from there, so consecutive lanes store
// consecutive slots of one super-tile's range (a few cache lines per store
// instead of one per lane).  Faster from a few dozen rows on; for short
// blocks the batches' scans and barriers cost more than they save.
#define FL3_PART1_STAGED_ROWS 64
template <typename RT>
__global__ __launch_bounds__(256) void k_bin_part1_staged(BinArgs a) {
   constexpr int P = sizeof(RT) == 4 ? 16 : 8;   // rows per batch
   __shared__ unsigned gnext[256];   // this workgroup's next slot in each super-tile
   __shared__ unsigned lcnt[256], loff[256], wsum[4];
   __shared__ RT stage[256 * P];
   const int ss = a.shift + FL3_SUPER_SHIFT;
   const int t = threadIdx.x;
   gnext[t] = t < a.nsuper ? a.soff[(size_t)blockIdx.x * a.nsuper + t] : 0u;
   const RT *rec = (const RT *)a.rec;
   RT *out = (RT *)a.sorted;
   for (long long i0 = 0; i0 < a.rows; i0 += P) {
      lcnt[t] = 0;
      __syncthreads();
      RT v[P];
      unsigned rk[P];
      for (int k = 0; k < P; k++) {
         const long long row = i0 + k;
         v[k] = row < a.rows ? rec_at(a, rec, (blockIdx.x * a.rows + row) * 256 + t) : (RT)FL3_REC_NONE;
      }
      for (int k = 0; k < P; k++)
         if (!rec_none(v[k])) rk[k] = atomicAdd(&lcnt[rel_cell(a, v[k]) >> ss], 1u);
      __syncthreads();
      unsigned tot;
      loff[t] = wg_excl_scan256(lcnt[t], wsum, &tot);
      __syncthreads();
      for (int k = 0; k < P; k++)
         if (!rec_none(v[k])) stage[loff[rel_cell(a, v[k]) >> ss] + rk[k]] = v[k];
      __syncthreads();
      for (unsigned j = t; j < tot; j += 256) {
         const RT w = stage[j];
         const unsigned sp = rel_cell(a, w) >> ss;
         out[gnext[sp] + (j - loff[sp])] = w;
      }
      __syncthreads();
      gnext[t] += lcnt[t];
   }
}

// Pass 2: workgroup c moves the super-tile-grouped records
// [c * FL3_PART2_RECORDS, ...) to their tiles.  The records of a workgroup
// usually span fewer than 256 tiles (a few super-tiles): ranks within the
// workgroup from LDS counters of that span, each tile's slots reserved with
// one global atomic, and 4-byte records ordered by tile in LDS first and
// written out from there, consecutive lanes to consecutive slots (8-byte
// records would need twice the LDS and halve the workgroups a CU holds, so
// they are written from registers).  A wider span (records of many sparse
// tiles) reserves one slot per record.  The order within a tile is
// arbitrary: the accumulation adds integers (or doubles, whose sums then
// differ in rounding only).
#define FL3_PART2_PER 32
#define FL3_PART2_RECORDS (256 * FL3_PART2_PER)
template <typename RT>
__global__ __launch_bounds__(256) void k_bin_part2(BinArgs a) {
   RT *stage = (RT *)fl3_lds;   // 4-byte records: [FL3_PART2_RECORDS]
   __shared__ unsigned lcnt[256], lofs[256], gofs[256], wsum[4];
   const long long total = a.bin_start[a.nbins];
   const long long r0 = (long long)blockIdx.x * FL3_PART2_RECORDS;
   if (r0 >= total) return;
   const RT *in = (const RT *)a.sorted;
   RT *out = (RT *)a.final;
   RT v[FL3_PART2_PER];
   for (int k = 0; k < FL3_PART2_PER; k++) {
      const long long r = r0 + k * 256 + threadIdx.x;
      v[k] = r < total ? in[r] : (RT)FL3_REC_NONE;
   }
   const long long rl = (total < r0 + FL3_PART2_RECORDS ? total : r0 + FL3_PART2_RECORDS) - 1;
   const int t_lo = (int)((rel_cell(a, in[r0]) >> a.shift) & ~((1u << FL3_SUPER_SHIFT) - 1u));
   const int t_hi = (int)((rel_cell(a, in[rl]) >> a.shift) | ((1u << FL3_SUPER_SHIFT) - 1u));
   if (t_hi - t_lo >= 256) {
      for (int k = 0; k < FL3_PART2_PER; k++)
         if (!rec_none(v[k])) out[atomicAdd(&a.cursor[rel_cell(a, v[k]) >> a.shift], 1u)] = v[k];
      return;
   }
   lcnt[threadIdx.x] = 0;
   __syncthreads();
   unsigned rank[FL3_PART2_PER];
   for (int k = 0; k < FL3_PART2_PER; k++)
      if (!rec_none(v[k])) rank[k] = atomicAdd(&lcnt[(rel_cell(a, v[k]) >> a.shift) - t_lo], 1u);
   __syncthreads();
   const int b = t_lo + (int)threadIdx.x;
   const unsigned c = lcnt[threadIdx.x];   // (zero past t_hi and the last tile)
   unsigned tot;
   const unsigned l = wg_excl_scan256(c, wsum, &tot);
   lofs[threadIdx.x] = l;
   gofs[threadIdx.x] = c ? atomicAdd(&a.cursor[b], c) : 0u;
   __syncthreads();
   if (sizeof(RT) == 4) {
      for (int k = 0; k < FL3_PART2_PER; k++)
         if (!rec_none(v[k])) stage[lofs[(rel_cell(a, v[k]) >> a.shift) - t_lo] + rank[k]] = v[k];
      __syncthreads();
      for (unsigned j = threadIdx.x; j < tot; j += 256) {
         const RT w = stage[j];
         const unsigned bb = (rel_cell(a, w) >> a.shift) - (unsigned)t_lo;
         out[gofs[bb] + (j - lofs[bb])] = w;
      }
   } else
      for (int k = 0; k < FL3_PART2_PER; k++)
         if (!rec_none(v[k])) out[gofs[(rel_cell(a, v[k]) >> a.shift) - t_lo] + rank[k]] = v[k];
}

// Persistent workgroups take work items (tile, chunk of its records) from a
// counter.  The tile's channels live in LDS as two 64-bit pairs (R|G, B|A)
// and the hit count: a work item adds at most 2^18 deltas of at most 255 to
// a cell, so no 32-bit half carries into the other.  A tile that has one
// work item is added to the histogram without atomics.
template <typename RT, bool JOBREC>
__global__ __launch_bounds__(FL3_ACC_THREADS) void k_bin_accum(BinArgs a) {
   const fl3_job_hdr *h = (const fl3_job_hdr *)a.job;
   const int tile_cells = 1 << a.shift;
   unsigned long long *rg = (unsigned long long *)fl3_lds;   // tile_cells
   unsigned long long *ba = rg + tile_cells;                  // tile_cells
   uint32_t *hits = (uint32_t *)(ba + tile_cells);            // tile_cells
   uint32_t *dmapu = hits + tile_cells;                       // 1024
   float *dmapf = (float *)(dmapu + 1024);                    // 1024
   float *vis = dmapf + 1024;                                 // 256
   __shared__ int item_sh;
   const double *pd = (const double *)(a.job + h->off_dmap);
   const fl3_xform_pack *px = (const fl3_xform_pack *)(a.job + h->off_xf);
   for (int e = threadIdx.x; e < 1024; e += blockDim.x) {
      dmapu[e] = (uint32_t)pd[e];
      dmapf[e] = (float)pd[e];
   }
   for (int e = threadIdx.x; e < h->nxf && e < 256; e += blockDim.x) vis[e] = (float)px[e].vis;
   const int linear = h->palette_mode == 1;
   const long long total_items = a.item_start[a.nbins];
   const RT *sorted = (const RT *)a.final;
   for (;;) {
      __syncthreads();
      if (threadIdx.x == 0) item_sh = (int)atomicAdd(a.item_counter, 1u);
      for (int e = threadIdx.x; e < tile_cells; e += blockDim.x) {
         rg[e] = 0;
         ba[e] = 0;
         hits[e] = 0;
      }
      __syncthreads();
      const long long item = item_sh;
      if (item >= total_items) break;
      int lo = 0, hi = a.nbins - 1;
      while (lo < hi) {  // last tile whose first item is <= item
         const int mid = (lo + hi + 1) >> 1;
         if ((long long)a.item_start[mid] <= item) lo = mid;
         else hi = mid - 1;
      }
      const int b = lo;
      const long long k = item - a.item_start[b];
      const long long r0 = (long long)a.bin_start[b] + k * a.chunk;
      const long long r1 = min((long long)a.bin_start[b + 1], r0 + a.chunk);
      const bool single = (a.item_start[b + 1] - a.item_start[b]) == 1;
      const unsigned cell0 = a.cell_base + ((unsigned)b << a.shift);
      for (long long rb = r0; rb < r1; rb += FL3_ACC_UNROLL * blockDim.x) {
         RT v[FL3_ACC_UNROLL];
         for (int u = 0; u < FL3_ACC_UNROLL; u++) {
            const long long r = rb + u * blockDim.x + threadIdx.x;
            v[u] = r < r1 ? sorted[r] : (RT)FL3_REC_NONE;
         }
         for (int u = 0; u < FL3_ACC_UNROLL; u++) {
            if (rec_none(v[u])) continue;
            const unsigned local = rec_cell(v[u]) - cell0;
            uint32_t d[5];
            if (sizeof(RT) == 4) {
               const int ci = (int)((unsigned)v[u] >> 24);
               for (int c = 0; c < 4; c++) d[c] = dmapu[ci * 4 + c];
               d[4] = 255u;
            } else if (JOBREC) {
               // the record's job selects palette and visibility (stepped palette)
               const unsigned long long w = (unsigned long long)v[u];
               const int ci = (int)((w >> 32) & 0xff), fn = (int)((w >> 40) & 0xff), jb = (int)(w >> 48);
               const float vv = a.vis_all[jb * a.vis_stride + fn];
               const uint32_t *du = a.dmapu_all + jb * 1024 + ci * 4;
               if (vv == 1.0f) {
                  for (int c = 0; c < 4; c++) d[c] = du[c];
                  d[4] = 255u;
               } else {
                  const float *df = a.dmapf_all + jb * 1024 + ci * 4;
                  for (int c = 0; c < 4; c++) d[c] = (uint32_t)(vv * df[c]);
                  d[4] = (uint32_t)(vv * 255.0f);
               }
            } else {
               const unsigned long long w = (unsigned long long)v[u];
               const int ci = (int)((w >> 32) & 0xff), fn = (int)((w >> 40) & 0xff);
               const float vv = vis[fn];
               if (linear) {
                  const float f = (float)(w >> 48) * (1.0f / 65535.0f);
                  for (int c = 0; c < 4; c++) {
                     const float col = dmapf[ci * 4 + c] * (1.0f - f) + dmapf[ci * 4 + 4 + c] * f;
                     d[c] = (uint32_t)(vv == 1.0f ? col : vv * col);
                  }
               } else if (vv == 1.0f) {
                  for (int c = 0; c < 4; c++) d[c] = dmapu[ci * 4 + c];
               } else {
                  for (int c = 0; c < 4; c++) d[c] = (uint32_t)(vv * dmapf[ci * 4 + c]);
               }
               d[4] = vv == 1.0f ? 255u : (uint32_t)(vv * 255.0f);
            }
            atomicAdd(&rg[local], (unsigned long long)d[0] | ((unsigned long long)d[1] << 32));
            atomicAdd(&ba[local], (unsigned long long)d[2] | ((unsigned long long)d[3] << 32));
            atomicAdd(&hits[local], d[4]);
         }
      }
      __syncthreads();
      uint32_t *bk = (uint32_t *)a.buckets + (size_t)cell0 * 5;
      const long long ncells = (long long)a.width * a.height;
      const int nvalid = (int)min((long long)tile_cells, ncells - (long long)cell0);
      for (int e = threadIdx.x; e < nvalid * 5; e += blockDim.x) {
         const int cl = e / 5, ch = e % 5;
         const uint32_t v = ch == 4 ? hits[cl] : (uint32_t)((ch < 2 ? rg[cl] : ba[cl]) >> (32 * (ch & 1)));
         if (!v) continue;
         if (single) {
            const unsigned long long s2 = (unsigned long long)bk[e] + v;
            bk[e] = s2 > 0xffffffffull ? 0xffffffffu : (uint32_t)s2;
         } else
            sat_add(bk + e, v);
      }
   }
}

// Accumulation of records of visibility 1 with a stepped palette (a record
// adds its palette entry and 255 hits): 4-byte records, or job records that
// carry the entry's bytes (palettes of entries below 256).  The tile's cells
// live in LDS as two packed 64-bit words, R:24 | G:24 | count:16 and B:32 | A:32, so a record
// costs two LDS atomics (and, for 4-byte records, one 16-byte read of its
// palette entry, kept in LDS in that packed form); a round of at most `round` records keeps every
// field below its width, after which each thread folds the CPT cells it
// owns into registers and clears them.  The tile is then added to the
// histogram from the registers (non-atomic saturating add if the tile has
// one work item).
template <int CPT, bool JOB>
__global__ __launch_bounds__(FL3_ACC_THREADS) void k_bin_accum_narrow(BinArgs a) {
   typedef typename std::conditional<JOB, unsigned long long, unsigned>::type RT;
   const fl3_job_hdr *h = (const fl3_job_hdr *)a.job;
   const int tile_cells = CPT * FL3_ACC_THREADS;
   unsigned long long *p0 = (unsigned long long *)fl3_lds;   // tile_cells
   unsigned long long *p1 = p0 + tile_cells;                  // tile_cells
   ulonglong2 *pal = (ulonglong2 *)(p1 + tile_cells);          // 256
   __shared__ int item_sh;
   const double *pd = (const double *)(a.job + h->off_dmap);
   if (!JOB)
      for (int e = threadIdx.x; e < 256; e += blockDim.x) {
         const uint32_t r = (uint32_t)pd[4 * e], g = (uint32_t)pd[4 * e + 1], bb = (uint32_t)pd[4 * e + 2], al = (uint32_t)pd[4 * e + 3];
         ulonglong2 w;
         w.x = (unsigned long long)r | ((unsigned long long)g << 24) | (1ull << 48);
         w.y = (unsigned long long)bb | ((unsigned long long)al << 32);
         pal[e] = w;
      }
   for (int e = threadIdx.x; e < tile_cells; e += blockDim.x) {
      p0[e] = 0;
      p1[e] = 0;
   }
   const long long total_items = a.item_start[a.nbins];
   const RT *sorted = (const RT *)a.final;
   const long long step = (long long)FL3_ACC_UNROLL * FL3_ACC_THREADS;
   for (;;) {
      __syncthreads();
      if (threadIdx.x == 0) item_sh = (int)atomicAdd(a.item_counter, 1u);
      __syncthreads();
      const long long item = item_sh;
      if (item >= total_items) break;
      int lo = 0, hi = a.nbins - 1;
      while (lo < hi) {  // last tile whose first item is <= item
         const int mid = (lo + hi + 1) >> 1;
         if ((long long)a.item_start[mid] <= item) lo = mid;
         else hi = mid - 1;
      }
      const int b = lo;
      const long long k = item - a.item_start[b];
      const long long r0 = (long long)a.bin_start[b] + k * a.chunk;
      const long long r1 = min((long long)a.bin_start[b + 1], r0 + a.chunk);
      const bool single = (a.item_start[b + 1] - a.item_start[b]) == 1;
      const unsigned cell0 = a.cell_base + ((unsigned)b << a.shift);
      uint32_t acc[CPT][5];
      for (int j = 0; j < CPT; j++)
         for (int c = 0; c < 5; c++) acc[j][c] = 0;
      for (long long rs = r0; rs < r1; rs += a.round) {
         const long long re = min(rs + a.round, r1);
         for (long long rb = rs; rb < re; rb += step) {
            RT v[FL3_ACC_UNROLL];
            for (int u = 0; u < FL3_ACC_UNROLL; u++) {
               const long long r = rb + u * FL3_ACC_THREADS + threadIdx.x;
               v[u] = r < re ? sorted[r] : (RT)FL3_REC_NONE;
            }
            for (int u = 0; u < FL3_ACC_UNROLL; u++) {
               if ((unsigned)v[u] == FL3_REC_NONE) continue;
               // 4-byte record: cell | palette index << 24; job record: cell
               // | the palette entry's bytes R, G, B, A << 32
               if (JOB) {
                  const unsigned local = (unsigned)v[u] - cell0;
                  const unsigned c = (unsigned)(v[u] >> 32);
                  atomicAdd(&p0[local], (unsigned long long)(c & 0xffu) | ((unsigned long long)((c >> 8) & 0xffu) << 24) |
                                            (1ull << 48));
                  atomicAdd(&p1[local], (unsigned long long)((c >> 16) & 0xffu) | ((unsigned long long)(c >> 24) << 32));
               } else {
                  const unsigned local = ((unsigned)v[u] & 0xffffffu) - cell0;
                  const ulonglong2 w = pal[(unsigned)v[u] >> 24];
                  atomicAdd(&p0[local], w.x);
                  atomicAdd(&p1[local], w.y);
               }
            }
         }
         __syncthreads();
         for (int j = 0; j < CPT; j++) {
            const int e = threadIdx.x + j * FL3_ACC_THREADS;
            const unsigned long long x0 = p0[e], x1 = p1[e];
            p0[e] = 0;
            p1[e] = 0;
            acc[j][0] += (uint32_t)(x0 & 0xffffffu);
            acc[j][1] += (uint32_t)((x0 >> 24) & 0xffffffu);
            acc[j][2] += (uint32_t)x1;
            acc[j][3] += (uint32_t)(x1 >> 32);
            acc[j][4] += 255u * (uint32_t)(x0 >> 48);
         }
         __syncthreads();
      }
      uint32_t *bk = (uint32_t *)a.buckets + (size_t)cell0 * 5;
      const long long ncells = (long long)a.width * a.height;
      const int nvalid = (int)min((long long)tile_cells, ncells - (long long)cell0);
      // every load (or atomic add) of the thread's cells is issued before
      // any result is used, so that their latencies overlap
      uint32_t old[CPT][5];
      if (nvalid == tile_cells) {
         // (a tile inside the histogram: straight-line code over all its
         // cells, so that no store waits for an earlier one; a tile of one
         // work item owns its cells, and adding zero leaves a cell as it is)
         uint32_t *bt = bk + (size_t)threadIdx.x * 5;
         if (single) {
            for (int j = 0; j < CPT; j++)
               for (int c = 0; c < 5; c++) old[j][c] = bt[(size_t)j * FL3_ACC_THREADS * 5 + c];
            for (int j = 0; j < CPT; j++)
               for (int c = 0; c < 5; c++) {
                  const unsigned long long s2 = (unsigned long long)old[j][c] + acc[j][c];
                  old[j][c] = s2 > 0xffffffffull ? 0xffffffffu : (uint32_t)s2;
               }
            for (int j = 0; j < CPT; j++)
               for (int c = 0; c < 5; c++) bt[(size_t)j * FL3_ACC_THREADS * 5 + c] = old[j][c];
         } else {
            for (int j = 0; j < CPT; j++)
               for (int c = 0; c < 5; c++) old[j][c] = atomicAdd(bt + (size_t)j * FL3_ACC_THREADS * 5 + c, acc[j][c]);
            for (int j = 0; j < CPT; j++)
               for (int c = 0; c < 5; c++)
                  if (old[j][c] > 0xffffffffu - acc[j][c])
                     atomicExch(bt + (size_t)j * FL3_ACC_THREADS * 5 + c, 0xffffffffu);
         }
      } else if (single) {
         for (int j = 0; j < CPT; j++)
            for (int c = 0; c < 5; c++) {
               const int e = threadIdx.x + j * FL3_ACC_THREADS;
               old[j][c] = e < nvalid && acc[j][c] ? bk[(size_t)e * 5 + c] : 0u;
            }
         // (the sums first, each in a register of its own, so that no store
         // waits for an earlier one to release its data register)
         for (int j = 0; j < CPT; j++)
            for (int c = 0; c < 5; c++) {
               const unsigned long long s2 = (unsigned long long)old[j][c] + acc[j][c];
               old[j][c] = s2 > 0xffffffffull ? 0xffffffffu : (uint32_t)s2;
            }
         for (int j = 0; j < CPT; j++)
            for (int c = 0; c < 5; c++) {
               const int e = threadIdx.x + j * FL3_ACC_THREADS;
               if (e < nvalid && acc[j][c]) bk[(size_t)e * 5 + c] = old[j][c];
            }
      } else {
         for (int j = 0; j < CPT; j++)
            for (int c = 0; c < 5; c++) {
               const int e = threadIdx.x + j * FL3_ACC_THREADS;
               old[j][c] = e < nvalid && acc[j][c] ? atomicAdd(bk + (size_t)e * 5 + c, acc[j][c]) : 0u;
            }
         // (a saturating add: the sum stops at UINT_MAX, as sat_add)
         for (int j = 0; j < CPT; j++)
            for (int c = 0; c < 5; c++) {
               const int e = threadIdx.x + j * FL3_ACC_THREADS;
               const uint32_t v = acc[j][c];
               if (e < nvalid && v && old[j][c] > 0xffffffffu - v) atomicExch(bk + (size_t)e * 5 + c, 0xffffffffu);
            }
      }
   }
}

// Counting pass of a band (the walkers' counting for records sorted by band
// first): workgroup k counts the records of its block of rows x 256 slots
// per tile in LDS, adds the counts to bin_total and stores its counts per
// super-tile at blk_hist[k][super-tile].
template <typename RT>
__global__ __launch_bounds__(256) void k_bin_count(BinArgs a) {
   unsigned *hist = (unsigned *)fl3_lds;   // [nbins]
   for (int b = threadIdx.x; b < a.nbins; b += blockDim.x) hist[b] = 0;
   __syncthreads();
   const RT *rec = (const RT *)a.rec;
   // (the rows' loads are issued in batches before their counts, so that a
   // thread has many loads in flight)
   long long i = 0;
   for (; i + FL3_SCATTER_UNROLL <= a.rows; i += FL3_SCATTER_UNROLL) {
      RT v[FL3_SCATTER_UNROLL];
      for (int k = 0; k < FL3_SCATTER_UNROLL; k++) v[k] = rec_at(a, rec, (blockIdx.x * a.rows + i + k) * 256 + threadIdx.x);
      for (int k = 0; k < FL3_SCATTER_UNROLL; k++)
         if (!rec_none(v[k])) atomicAdd(&hist[rel_cell(a, v[k]) >> a.shift], 1u);
   }
   for (; i < a.rows; i++) {
      const RT v = rec_at(a, rec, (blockIdx.x * a.rows + i) * 256 + threadIdx.x);
      if (!rec_none(v)) atomicAdd(&hist[rel_cell(a, v) >> a.shift], 1u);
   }
   __syncthreads();
   for (int b = threadIdx.x; b < a.nbins; b += blockDim.x)
      if (hist[b]) atomicAdd(&a.bin_total[b], hist[b]);
   for (int sp = threadIdx.x; sp < a.nsuper; sp += blockDim.x) {
      const int b0 = sp << FL3_SUPER_SHIFT, b1 = min(a.nbins, b0 + (1 << FL3_SUPER_SHIFT));
      unsigned c = 0;
      for (int b = b0; b < b1; b++) c += hist[b];
      a.blk_hist[(size_t)blockIdx.x * a.nsuper + sp] = c;
   }
}

// Accumulation into bits-64 (double) buckets.  Each record adds what the
// direct deposit adds for its point (rect.c:492-497 in double: the palette
// entry, blended for a linear palette, times the xform's visibility, and
// 255 times it to the count), rounded to a multiple of 2^-32 and summed in
// LDS as signed 64-bit integers (integer atomics outpace double ones; a
// work item of at most 2^31 / (largest |entry| + 1) records keeps a cell's
// sum below 2^63 in size); the item's sums are then added to the double
// buckets.  Against the per-sample double atomics a cell's sum differs by
// at most 2^-33 per record (about 1e-12 relative), besides the order of the
// additions.
// 4-byte records (stepped palette, visibility 1; each adds exactly 255 to
// the count, which is then a 32-bit record count scaled at the flush) or
// 8-byte records (make_record).  Persistent workgroups take work items as
// in k_bin_accum; a tile with one item is added without atomics.
#define FL3_F64_FIX 4294967296.0    // 2^32: the fixed-point scale of the LDS sums
template <typename RT>
__global__ __launch_bounds__(FL3_ACC_THREADS) void k_bin_accum_f64(BinArgs a) {
   const fl3_job_hdr *h = (const fl3_job_hdr *)a.job;
   const int tile_cells = 1 << a.shift;
   unsigned long long *acc = (unsigned long long *)fl3_lds;   // [5][tile_cells]
   unsigned long long *dfx = acc + 5 * tile_cells;            // [1024] the palette, fixed point
   double *dm = (double *)(dfx + 1024);                       // [1024]
   double *vis = dm + 1024;                                   // [256]
   __shared__ int item_sh;
   const double *pd = (const double *)(a.job + h->off_dmap);
   const fl3_xform_pack *px = (const fl3_xform_pack *)(a.job + h->off_xf);
   for (int e = threadIdx.x; e < 1024; e += blockDim.x) {
      dm[e] = pd[e];
      dfx[e] = (unsigned long long)__double2ll_rn(pd[e] * FL3_F64_FIX);
   }
   // (the walkers carry a visibility as a float)
   for (int e = threadIdx.x; e < 256; e += blockDim.x) vis[e] = e < h->nxf ? (double)(float)px[e].vis : 1.0;
   const int linear = h->palette_mode == 1;
   const unsigned long long hits = (unsigned long long)__double2ll_rn(255.0 * FL3_F64_FIX);
   const long long total_items = a.item_start[a.nbins];
   const RT *sorted = (const RT *)a.final;
   for (;;) {
      __syncthreads();
      if (threadIdx.x == 0) item_sh = (int)atomicAdd(a.item_counter, 1u);
      for (int e = threadIdx.x; e < 5 * tile_cells; e += blockDim.x) acc[e] = 0ull;
      __syncthreads();
      const long long item = item_sh;
      if (item >= total_items) break;
      int lo = 0, hi = a.nbins - 1;
      while (lo < hi) {  // last tile whose first item is <= item
         const int mid = (lo + hi + 1) >> 1;
         if ((long long)a.item_start[mid] <= item) lo = mid;
         else hi = mid - 1;
      }
      const int b = lo;
      const long long k = item - a.item_start[b];
      const long long r0 = (long long)a.bin_start[b] + k * a.chunk;
      const long long r1 = min((long long)a.bin_start[b + 1], r0 + a.chunk);
      const bool single = (a.item_start[b + 1] - a.item_start[b]) == 1;
      const unsigned tb = (unsigned)b << a.shift;   // the tile's first cell in the band
      const unsigned cell0 = a.cell_base + tb;
      for (long long rb = r0; rb < r1; rb += FL3_ACC_UNROLL * blockDim.x) {
         RT v[FL3_ACC_UNROLL];
         for (int u = 0; u < FL3_ACC_UNROLL; u++) {
            const long long r = rb + u * blockDim.x + threadIdx.x;
            v[u] = r < r1 ? sorted[r] : (RT)FL3_REC_NONE;
         }
         for (int u = 0; u < FL3_ACC_UNROLL; u++) {
            if (rec_none(v[u])) continue;
            const unsigned local = rel_cell(a, v[u]) - tb;
            unsigned long long d[5];
            if (sizeof(RT) == 4) {
               // (every 4-byte record adds 255 hits: the count channel is a
               // 32-bit record count, times 255 at the flush)
               const unsigned long long *e = dfx + ((unsigned)v[u] >> 24) * 4;
               for (int ch = 0; ch < 4; ch++) atomicAdd(&acc[ch * tile_cells + local], e[ch]);
               atomicAdd((unsigned *)(acc + 4 * tile_cells) + local, 1u);
               continue;
            } else {
               const unsigned long long x = (unsigned long long)v[u];
               const int ci = (int)((x >> 32) & 0xff), fn = (int)((x >> 40) & 0xff);
               const double w = vis[fn];
               if (!linear && w == 1.0) {
                  for (int ch = 0; ch < 4; ch++) d[ch] = dfx[ci * 4 + ch];
                  d[4] = hits;
               } else {
                  const double *e = dm + ci * 4;
                  double c[4];
                  if (linear) {
                     const double f = (double)((float)(x >> 48) * (1.0f / 65535.0f));
                     for (int ch = 0; ch < 4; ch++) c[ch] = e[ch] * (1.0 - f) + e[4 + ch] * f;
                  } else
                     for (int ch = 0; ch < 4; ch++) c[ch] = e[ch];
                  for (int ch = 0; ch < 4; ch++)
                     d[ch] = (unsigned long long)__double2ll_rn((w == 1.0 ? c[ch] : w * c[ch]) * FL3_F64_FIX);
                  d[4] = w == 1.0 ? hits : (unsigned long long)__double2ll_rn(w * 255.0 * FL3_F64_FIX);
               }
            }
            for (int ch = 0; ch < 5; ch++) atomicAdd(&acc[ch * tile_cells + local], d[ch]);
         }
      }
      __syncthreads();
      double *bk = (double *)a.buckets + (size_t)cell0 * 5;
      const long long ncells = (long long)a.width * a.height;
      const int nvalid = (int)min((long long)tile_cells, ncells - (long long)cell0);
      // a thread's values of the tile (at most 5 x 2048 / FL3_ACC_THREADS),
      // their loads all issued before any is used
      constexpr int K = 5 * 2048 / FL3_ACC_THREADS;
      double val[K];
      for (int k = 0; k < K; k++) {
         const int e = threadIdx.x + k * FL3_ACC_THREADS;
         val[k] = e >= nvalid * 5 ? 0.0
                  : sizeof(RT) == 4 && e % 5 == 4
                      ? (double)((const unsigned *)(acc + 4 * tile_cells))[e / 5] * 255.0
                      : (double)(long long)acc[(e % 5) * tile_cells + e / 5] * (1.0 / FL3_F64_FIX);
      }
      if (single) {
         double old[K];
         for (int k = 0; k < K; k++) old[k] = val[k] != 0.0 ? bk[threadIdx.x + k * FL3_ACC_THREADS] : 0.0;
         for (int k = 0; k < K; k++) old[k] += val[k];
         for (int k = 0; k < K; k++)
            if (val[k] != 0.0) bk[threadIdx.x + k * FL3_ACC_THREADS] = old[k];
      } else
         for (int k = 0; k < K; k++)
            if (val[k] != 0.0) unsafeAtomicAdd(bk + threadIdx.x + k * FL3_ACC_THREADS, val[k]);
   }
}

extern "C" size_t fl3k_bin_f64_lds_bytes(int shift) { return (size_t)40 * ((size_t)1 << shift) + 8 * (1024 + 1024 + 256) + 64; }

extern "C" size_t fl3k_bin_narrow_lds_bytes(int shift) { return (size_t)16 * ((size_t)1 << shift) + 4 * 1024 + 64; }
extern "C" int fl3k_bin_pass_records(void) { return FL3_ACC_UNROLL * FL3_ACC_THREADS; }

extern "C" size_t fl3k_bin_lds_bytes(int shift) {
   return (size_t)20 * ((size_t)1 << shift) + 8 * 1024 + 4 * 256 + 64;
}

extern "C" int fl3k_launch_bin(const BinArgs *a, int ncu, int format, hipStream_t s) {
   hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, s, *a);
   hipLaunchKernelGGL(k_super_off, dim3(a->nsuper), dim3(256), 0, s, *a);
   const unsigned g2 = (unsigned)(((long long)a->walkers * a->rows + FL3_PART2_RECORDS - 1) / FL3_PART2_RECORDS);
   // the staging area of 4-byte records (8-byte records: none)
   const size_t lds2_4 = 4 * (size_t)FL3_PART2_RECORDS, lds2_8 = 0;
   const size_t lds = fl3k_bin_lds_bytes(a->shift);
   const unsigned gacc = (unsigned)ncu * (lds > 65536 ? 1 : 3);
   const int big = lds > 65536;
   const int cpt = (1 << a->shift) / FL3_ACC_THREADS;
   if (format == FL3_BIN_F64_WIDE || format == FL3_BIN_F64_NARROW) {
      // double buckets: 8- or 4-byte records, accumulated in double
      const size_t lf = fl3k_bin_f64_lds_bytes(a->shift);
      const unsigned gf = (unsigned)ncu * (lf > 81920 ? 1 : (lf > 54000 ? 2 : 3));
      if (format == FL3_BIN_F64_NARROW) {
         if (a->rows >= FL3_PART1_STAGED_ROWS)
            hipLaunchKernelGGL(k_bin_part1_staged<unsigned>, dim3(a->nblk), dim3(256), 0, s, *a);
         else
            hipLaunchKernelGGL(k_bin_part1<unsigned>, dim3(a->nblk), dim3(256), 0, s, *a);
         if (g2) hipLaunchKernelGGL(k_bin_part2<unsigned>, dim3(g2), dim3(256), lds2_4, s, *a);
         if (lf > 65536)
            (void)hipFuncSetAttribute((const void *)k_bin_accum_f64<unsigned>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lf);
         hipLaunchKernelGGL(k_bin_accum_f64<unsigned>, dim3(gf), dim3(FL3_ACC_THREADS), lf, s, *a);
      } else {
         hipLaunchKernelGGL(k_bin_part1<unsigned long long>, dim3(a->nblk), dim3(256), 0, s, *a);
         if (g2) hipLaunchKernelGGL(k_bin_part2<unsigned long long>, dim3(g2), dim3(256), lds2_8, s, *a);
         if (lf > 65536)
            (void)hipFuncSetAttribute((const void *)k_bin_accum_f64<unsigned long long>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lf);
         hipLaunchKernelGGL(k_bin_accum_f64<unsigned long long>, dim3(gf), dim3(FL3_ACC_THREADS), lf, s, *a);
      }
      return hipGetLastError() != hipSuccess;
   }
   if ((format == 1 || format == 2) && a->round > 0 && (cpt == 2 || cpt == 4)) {
      const bool job = format == 2;
      if (job) {
         hipLaunchKernelGGL(k_bin_part1<unsigned long long>, dim3(a->nblk), dim3(256), 0, s, *a);
         if (g2) hipLaunchKernelGGL(k_bin_part2<unsigned long long>, dim3(g2), dim3(256), lds2_8, s, *a);
      } else {
         if (a->rows >= FL3_PART1_STAGED_ROWS)
            hipLaunchKernelGGL(k_bin_part1_staged<unsigned>, dim3(a->nblk), dim3(256), 0, s, *a);
         else
            hipLaunchKernelGGL(k_bin_part1<unsigned>, dim3(a->nblk), dim3(256), 0, s, *a);
         if (g2) hipLaunchKernelGGL(k_bin_part2<unsigned>, dim3(g2), dim3(256), lds2_4, s, *a);
      }
      const size_t ln = fl3k_bin_narrow_lds_bytes(a->shift);
      const unsigned gn = (unsigned)ncu * (ln > 65536 ? 2 : 3);
      const void *kf = job ? (cpt == 4 ? (const void *)k_bin_accum_narrow<4, true> : (const void *)k_bin_accum_narrow<2, true>)
                           : (cpt == 4 ? (const void *)k_bin_accum_narrow<4, false> : (const void *)k_bin_accum_narrow<2, false>);
      if (ln > 65536) (void)hipFuncSetAttribute(kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ln);
      if (job && cpt == 4)
         hipLaunchKernelGGL((k_bin_accum_narrow<4, true>), dim3(gn), dim3(FL3_ACC_THREADS), ln, s, *a);
      else if (job)
         hipLaunchKernelGGL((k_bin_accum_narrow<2, true>), dim3(gn), dim3(FL3_ACC_THREADS), ln, s, *a);
      else if (cpt == 4)
         hipLaunchKernelGGL((k_bin_accum_narrow<4, false>), dim3(gn), dim3(FL3_ACC_THREADS), ln, s, *a);
      else
         hipLaunchKernelGGL((k_bin_accum_narrow<2, false>), dim3(gn), dim3(FL3_ACC_THREADS), ln, s, *a);
   } else if (format == 1) {
      if (a->rows >= FL3_PART1_STAGED_ROWS)
         hipLaunchKernelGGL(k_bin_part1_staged<unsigned>, dim3(a->nblk), dim3(256), 0, s, *a);
      else
         hipLaunchKernelGGL(k_bin_part1<unsigned>, dim3(a->nblk), dim3(256), 0, s, *a);
      if (g2) hipLaunchKernelGGL(k_bin_part2<unsigned>, dim3(g2), dim3(256), lds2_4, s, *a);
      if (big)
         (void)hipFuncSetAttribute((const void *)k_bin_accum<unsigned, false>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      hipLaunchKernelGGL((k_bin_accum<unsigned, false>), dim3(gacc), dim3(FL3_ACC_THREADS), lds, s, *a);
   } else {
      hipLaunchKernelGGL(k_bin_part1<unsigned long long>, dim3(a->nblk), dim3(256), 0, s, *a);
      if (g2) hipLaunchKernelGGL(k_bin_part2<unsigned long long>, dim3(g2), dim3(256), lds2_8, s, *a);
      if (format == 2) {
         if (big)
            (void)hipFuncSetAttribute((const void *)k_bin_accum<unsigned long long, true>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
         hipLaunchKernelGGL((k_bin_accum<unsigned long long, true>), dim3(gacc), dim3(FL3_ACC_THREADS), lds, s, *a);
      } else {
         if (big)
            (void)hipFuncSetAttribute((const void *)k_bin_accum<unsigned long long, false>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
         hipLaunchKernelGGL((k_bin_accum<unsigned long long, false>), dim3(gacc), dim3(FL3_ACC_THREADS), lds, s, *a);
      }
   }
   return hipGetLastError() != hipSuccess;
}

// Banded histograms, first level: the walkers' 8-byte records grouped by
// super-tile (the bands) into a->sorted: scans of the walkers' counts and the
// first sort pass only.
extern "C" int fl3k_launch_bin_split(const BinArgs *a, int narrow_out, hipStream_t s) {
   hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, s, *a);
   hipLaunchKernelGGL(k_super_off, dim3(a->nsuper), dim3(256), 0, s, *a);
   if (narrow_out)
      hipLaunchKernelGGL((k_bin_part1<unsigned long long, unsigned>), dim3(a->nblk), dim3(256), 0, s, *a);
   else
      hipLaunchKernelGGL(k_bin_part1<unsigned long long>, dim3(a->nblk), dim3(256), 0, s, *a);
   return hipGetLastError() != hipSuccess;
}

// Banded histograms, second level: per-tile and per-block counts of one
// band's 8-byte records (a->rec, a->rec_count of them, as a->nblk blocks of
// a->rows x 256), before fl3k_launch_bin sorts and accumulates them.
extern "C" int fl3k_launch_bin_count(const BinArgs *a, int narrow, hipStream_t s) {
   const size_t l = sizeof(unsigned) * (size_t)a->nbins;   // at most FL3_MAX_BINS: 32 KiB
   if (narrow)
      hipLaunchKernelGGL(k_bin_count<unsigned>, dim3(a->nblk), dim3(256), l, s, *a);
   else
      hipLaunchKernelGGL(k_bin_count<unsigned long long>, dim3(a->nblk), dim3(256), l, s, *a);
   return hipGetLastError() != hipSuccess;
}

// A plain copy of n16 16-byte words, one per thread: the reference kernel of
// known traffic for calibrating hardware byte counters.
__global__ __launch_bounds__(256) void k_copy_probe(const uint4 *src, uint4 *dst, long long n16) {
   const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   if (i < n16) dst[i] = src[i];
}

extern "C" int fl3k_launch_copy_probe(const void *src, void *dst, long long n16, hipStream_t s) {
   hipLaunchKernelGGL(k_copy_probe, dim3((unsigned)((n16 + 255) / 256)), dim3(256), 0, s, (const uint4 *)src,
                      (uint4 *)dst, n16);
   return hipGetLastError() != hipSuccess;
}
