// fl3_kernels_fast.hip - FP32 production walkers (compiled with FP
// contraction on).
#include "fl3_core.h"
#include "fl3_launch.h"


// Production walkers.  Walker w plots nsamples/W (+1 for the first
// nsamples%W walkers) points after `fuse` unplotted steps from a random
// start, exactly the per-sub-batch procedure of rect.c:393-400 with one long
// sub-batch per walker.
template <int BITS, bool EXACT, bool RECORDS>
__global__ __launch_bounds__(256) void k_iter_fast(IterArgs a) {
   unsigned char *lds = fl3_lds;
   GenL<float> G;
   load_genome<float>(a.job, lds, G);
   __syncthreads();
   const long long W = (long long)gridDim.x * blockDim.x;
   const long long wid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   const long long n = a.nsamples / W + (wid < a.nsamples % W ? 1 : 0);
   WalkerRng rng;
   float p[4], q[4];
   int lastxf = 0, consec = 0;
   uint32_t *ws = a.wstate;
   if (a.resume) {
      for (int k = 0; k < 4; k++) p[k] = __uint_as_float(ws[k * W + wid]);
      rng.s0 = ws[4 * W + wid];
      rng.s1 = ws[5 * W + wid];
      rng.s2 = ws[6 * W + wid];
      rng.s3 = ws[7 * W + wid];
      lastxf = (int)ws[8 * W + wid];
      consec = (int)ws[9 * W + wid];
   } else {
      rng.seed(a.key0, a.key1, (uint64_t)wid);
      p[0] = u11<float>(rng);
      p[1] = u11<float>(rng);
      p[2] = u01<float>(rng);
      p[3] = u01<float>(rng);
   }
   uint32_t bad = 0;
   const bool use_final = G.final_index >= 0;
   for (long long i = a.resume ? 0 : -a.fuse; i < n;) {
      const uint32_t r = rng.next();
      const int fn = select_xform(G, G.chaos_enable ? lastxf : 0, r & 16383u);
      if (apply_xform<float>(G, fn, p, q, rng)) {
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
         if (G.final_opacity == 1.0f || u01<float>(rng) < G.final_opacity) {
            apply_xform<float>(G, G.final_index, p, q, rng);
            q[3] = p[3];
         }
      }
      if (i >= 0) {
         if (RECORDS)
            a.rec[i * W + wid] = make_record<float>(G, q, fn);
         else
            deposit<float, BITS, EXACT>(G, q, a.buckets, a.sat);
      }
      i++;
   }
   if (RECORDS)
      for (long long i = n; i < a.rows_out; i++) a.rec[i * W + wid] = FL3_REC_NONE;
   if (ws) {
      for (int k = 0; k < 4; k++) ws[k * W + wid] = __float_as_uint(p[k]);
      ws[4 * W + wid] = rng.s0;
      ws[5 * W + wid] = rng.s1;
      ws[6 * W + wid] = rng.s2;
      ws[7 * W + wid] = rng.s3;
      ws[8 * W + wid] = (uint32_t)lastxf;
      ws[9 * W + wid] = (uint32_t)consec;
   }
   add_bad(a.badvals, bad);
}

extern "C" int fl3k_launch_iter_fast(const IterArgs *a, unsigned grid, size_t lds, hipStream_t s, int bits,
                                     int exact) {
   if (a->rec)
      hipLaunchKernelGGL((k_iter_fast<33, true, true>), dim3(grid), dim3(256), lds, s, *a);
   else if (bits == 64)
      hipLaunchKernelGGL((k_iter_fast<64, true, false>), dim3(grid), dim3(256), lds, s, *a);
   else if (exact)
      hipLaunchKernelGGL((k_iter_fast<33, true, false>), dim3(grid), dim3(256), lds, s, *a);
   else
      hipLaunchKernelGGL((k_iter_fast<33, false, false>), dim3(grid), dim3(256), lds, s, *a);
   return hipGetLastError() != hipSuccess;
}

// ------------------------------------------------------------------ binned deposit
//
// The walker kernel above writes one record per plotted sample instead of
// five global atomics.  The records are then grouped by tile of 1 << shift
// cells with a counting sort (count, scan, scatter), and each tile's records
// are accumulated in LDS and added to the histogram once.  Integer sums
// commute, so the histogram is the one the atomic deposit would produce;
// the tile flush applies the reference's saturation exactly.

__global__ __launch_bounds__(256) void k_bin_count(BinArgs a) {
   unsigned *hist = (unsigned *)fl3_lds;
   for (int b = threadIdx.x; b < a.nbins; b += blockDim.x) hist[b] = 0;
   __syncthreads();
   for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x; r < a.n;
        r += (long long)gridDim.x * blockDim.x) {
      const unsigned c = (unsigned)a.rec[r];
      if (c != FL3_REC_NONE) atomicAdd(&hist[c >> a.shift], 1u);
   }
   __syncthreads();
   for (int b = threadIdx.x; b < a.nbins; b += blockDim.x)
      if (hist[b]) atomicAdd(&a.bin_total[b], hist[b]);
}

// One workgroup: exclusive scans of the tile counts and of the work items
// (ceil(count / chunk) per tile).
__global__ __launch_bounds__(1024) void k_bin_scan(BinArgs a) {
   __shared__ unsigned long long part[1024], ipart[1024];
   const int t = threadIdx.x, per = (a.nbins + 1023) / 1024;
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
      a.item_start[b] = (unsigned)irun;
      run += a.bin_total[b];
      irun += (a.bin_total[b] + a.chunk - 1) / a.chunk;
   }
   if (t == 1023) {
      a.bin_start[a.nbins] = (unsigned)part[1023];
      a.item_start[a.nbins] = (unsigned)ipart[1023];
   }
}

__global__ __launch_bounds__(256) void k_bin_scatter(BinArgs a) {
   unsigned *hist = (unsigned *)fl3_lds, *base = hist + a.nbins;
   for (int b = threadIdx.x; b < a.nbins; b += blockDim.x) hist[b] = 0;
   __syncthreads();
   for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x; r < a.n;
        r += (long long)gridDim.x * blockDim.x) {
      const unsigned c = (unsigned)a.rec[r];
      if (c != FL3_REC_NONE) atomicAdd(&hist[c >> a.shift], 1u);
   }
   __syncthreads();
   for (int b = threadIdx.x; b < a.nbins; b += blockDim.x) {
      const unsigned h = hist[b];
      base[b] = h ? a.bin_start[b] + atomicAdd(&a.bin_fill[b], h) : 0u;
      hist[b] = 0;
   }
   __syncthreads();
   for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x; r < a.n;
        r += (long long)gridDim.x * blockDim.x) {
      const unsigned long long v = a.rec[r];
      const unsigned c = (unsigned)v;
      if (c == FL3_REC_NONE) continue;
      const unsigned b = c >> a.shift;
      a.sorted[base[b] + atomicAdd(&hist[b], 1u)] = v;
   }
}

// Persistent workgroups take work items (tile, chunk of its records) from a
// counter.  The tile's 5 channels live in LDS; a tile that has one work item
// is added to the histogram without atomics.
__global__ __launch_bounds__(256) void k_bin_accum(BinArgs a) {
   const fl3_job_hdr *h = (const fl3_job_hdr *)a.job;
   const int tile_cells = 1 << a.shift;
   uint32_t *tile = (uint32_t *)fl3_lds;                       // tile_cells x 5
   uint32_t *dmapu = tile + 5 * tile_cells;               // 1024
   float *dmapf = (float *)(dmapu + 1024);                // 1024
   float *vis = dmapf + 1024;                             // nxf
   __shared__ int item_sh;
   const double *pd = (const double *)(a.job + h->off_dmap);
   const fl3_xform_pack *px = (const fl3_xform_pack *)(a.job + h->off_xf);
   for (int e = threadIdx.x; e < 1024; e += blockDim.x) {
      dmapu[e] = (uint32_t)pd[e];
      dmapf[e] = (float)pd[e];
   }
   for (int e = threadIdx.x; e < h->nxf; e += blockDim.x) vis[e] = (float)px[e].vis;
   const int linear = h->palette_mode == 1;
   const long long total_items = a.item_start[a.nbins];
   for (;;) {
      __syncthreads();
      if (threadIdx.x == 0) item_sh = (int)atomicAdd(a.item_counter, 1u);
      for (int e = threadIdx.x; e < 5 * tile_cells; e += blockDim.x) tile[e] = 0;
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
      const unsigned cell0 = (unsigned)b << a.shift;
      for (long long r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
         const unsigned long long v = a.sorted[r];
         const unsigned local = (unsigned)v - cell0;
         const int ci = (int)((v >> 32) & 0xff), fn = (int)((v >> 40) & 0xff);
         const float vv = vis[fn];
         uint32_t d[5];
         if (linear) {
            const float f = (float)(v >> 48) * (1.0f / 65535.0f);
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
         uint32_t *t = tile + local * 5;
         for (int c = 0; c < 5; c++)
            if (d[c]) atomicAdd(t + c, d[c]);
      }
      __syncthreads();
      uint32_t *bk = (uint32_t *)a.buckets + (size_t)cell0 * 5;
      const long long ncells = (long long)a.width * a.height;
      const int nvalid = (int)min((long long)tile_cells, ncells - (long long)cell0);
      for (int e = threadIdx.x; e < nvalid * 5; e += blockDim.x) {
         const uint32_t v = tile[e];
         if (!v) continue;
         if (single) {
            const unsigned long long s = (unsigned long long)bk[e] + v;
            bk[e] = s > 0xffffffffull ? 0xffffffffu : (uint32_t)s;
         } else
            sat_add(bk + e, v);
      }
   }
}

extern "C" size_t fl3k_bin_lds_bytes(int shift) {
   return sizeof(uint32_t) * 5 * ((size_t)1 << shift) + 8 * 1024 + 4 * 256 + 64;
}

extern "C" int fl3k_launch_bin(const BinArgs *a, int ncu, hipStream_t s) {
   const unsigned g = (unsigned)min(1024LL, max(1LL, (a->n + 256LL * 64 - 1) / (256LL * 64)));
   hipLaunchKernelGGL(k_bin_count, dim3(g), dim3(256), sizeof(unsigned) * a->nbins, s, *a);
   hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, s, *a);
   hipLaunchKernelGGL(k_bin_scatter, dim3(g), dim3(256), 2 * sizeof(unsigned) * a->nbins, s, *a);
   const size_t lds = fl3k_bin_lds_bytes(a->shift);
   if (lds > 65536) (void)hipFuncSetAttribute((const void *)k_bin_accum, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
   hipLaunchKernelGGL(k_bin_accum, dim3(ncu * (lds > 65536 ? 1 : 3)), dim3(256), lds, s, *a);
   return hipGetLastError() != hipSuccess;
}
