// fl3_kernels_fast.hip - FP32 production walkers (compiled with FP
// contraction on).
#include "fl3_core.h"
#include "fl3_launch.h"


// Production walkers.  Walker w plots nsamples/W (+1 for the first
// nsamples%W walkers) points after `fuse` unplotted steps from a random
// start, exactly the per-sub-batch procedure of rect.c:393-400 with one long
// sub-batch per walker.
template <int BITS, bool EXACT>
__global__ __launch_bounds__(256) void k_iter_fast(IterArgs a) {
   unsigned char *lds = fl3_lds;
   GenL<float> G;
   load_genome<float>(a.job, lds, G);
   __syncthreads();
   const long long W = (long long)gridDim.x * blockDim.x;
   const long long wid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
   const long long n = a.nsamples / W + (wid < a.nsamples % W ? 1 : 0);
   WalkerRng rng;
   rng.seed(a.key0, a.key1, (uint64_t)wid);
   float p[4], q[4];
   p[0] = u11<float>(rng);
   p[1] = u11<float>(rng);
   p[2] = u01<float>(rng);
   p[3] = u01<float>(rng);
   int lastxf = 0, consec = 0;
   uint32_t bad = 0;
   const bool use_final = G.final_index >= 0;
   for (long long i = -a.fuse; i < n;) {
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
      if (i >= 0) deposit<float, BITS, EXACT>(G, q, a.buckets, a.sat);
      i++;
   }
   add_bad(a.badvals, bad);
}

extern "C" int fl3k_launch_iter_fast(const IterArgs *a, unsigned grid, size_t lds, hipStream_t s, int bits,
                                     int exact) {
   if (bits == 64)
      hipLaunchKernelGGL((k_iter_fast<64, true>), dim3(grid), dim3(256), lds, s, *a);
   else if (exact)
      hipLaunchKernelGGL((k_iter_fast<33, true>), dim3(grid), dim3(256), lds, s, *a);
   else
      hipLaunchKernelGGL((k_iter_fast<33, false>), dim3(grid), dim3(256), lds, s, *a);
   return hipGetLastError() != hipSuccess;
}
