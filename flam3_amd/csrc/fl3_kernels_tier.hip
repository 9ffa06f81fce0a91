// fl3_kernels_tier.hip - FP32 walker kernels restricted to a variation
// tier (compiled with FP contraction on, like fl3_kernels_fast.hip).  A
// genome whose variations all fall in one tier runs a kernel without the
// code and register demand of the other cases.
#include "fl3_walker.h"
#include "fl3_launch.h"

template <int TIER, bool WAVE>
static void launch_tier_mode(const IterArgs *a, unsigned grid, size_t lds, hipStream_t s, int exact) {
   if (a->rec_mode == FL3_REC_JOB)
      hipLaunchKernelGGL((k_iter_fast<33, true, FL3_REC_JOB, TIER, WAVE>), dim3(grid), dim3(256), lds, s, *a);
   else if (a->rec_mode == FL3_REC_NARROW)
      hipLaunchKernelGGL((k_iter_fast<33, true, FL3_REC_NARROW, TIER, WAVE>), dim3(grid), dim3(256), lds, s, *a);
   else if (a->rec_mode == FL3_REC_WIDE)
      hipLaunchKernelGGL((k_iter_fast<33, true, FL3_REC_WIDE, TIER, WAVE>), dim3(grid), dim3(256), lds, s, *a);
   else if (exact)
      hipLaunchKernelGGL((k_iter_fast<33, true, FL3_REC_DIRECT, TIER, WAVE>), dim3(grid), dim3(256), lds, s, *a);
   else
      hipLaunchKernelGGL((k_iter_fast<33, false, FL3_REC_DIRECT, TIER, WAVE>), dim3(grid), dim3(256), lds, s, *a);
}

template <int TIER>
static void launch_tier(const IterArgs *a, unsigned grid, size_t lds, hipStream_t s, int exact) {
   if (a->wave_xf)
      launch_tier_mode<TIER, true>(a, grid, lds, s, exact);
   else
      launch_tier_mode<TIER, false>(a, grid, lds, s, exact);
}

// Records (binned deposit, either format) or direct bits-33 deposits (fast,
// or saturating with exact).
extern "C" int fl3k_launch_iter_tier(const IterArgs *a, unsigned grid, size_t lds, hipStream_t s, int tier,
                                     int exact) {
   switch (tier) {
   case 0: launch_tier<0>(a, grid, lds, s, exact); break;
   case 1: launch_tier<1>(a, grid, lds, s, exact); break;
   case 2: launch_tier<2>(a, grid, lds, s, exact); break;
   case 3: launch_tier<3>(a, grid, lds, s, exact); break;
   case 4: launch_tier<4>(a, grid, lds, s, exact); break;
   case 5: launch_tier<5>(a, grid, lds, s, exact); break;
   default: return 1;
   }
   return hipGetLastError() != hipSuccess;
}

extern "C" int fl3k_launch_apply_fp32_tier(const unsigned char *job, int xi, const double *pts, long long n,
                                          double *out, unsigned long long key, size_t lds, hipStream_t s,
                                          int tier) {
   const unsigned grid = (unsigned)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
   switch (tier) {
#define FL3_APPLY_CASE(T) \
   case T: hipLaunchKernelGGL(k_apply_fp32<T>, dim3(grid), dim3(256), lds, s, job, xi, pts, n, out, key); break;
   FL3_APPLY_CASE(0) FL3_APPLY_CASE(1) FL3_APPLY_CASE(2) FL3_APPLY_CASE(3) FL3_APPLY_CASE(4) FL3_APPLY_CASE(5)
#undef FL3_APPLY_CASE
   default: return 1;
   }
   return hipGetLastError() != hipSuccess;
}
