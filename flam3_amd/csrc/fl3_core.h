// fl3_core.h - device building blocks shared by the kernel translation
// units: math overloads, random streams, the genome in LDS, the variation
// library, xform application and the histogram deposit.  FP64 overloads call
// the correctly rounded library functions; FP32 ones trade accuracy for
// throughput where the result only feeds the statistics of a render.
#ifndef FL3_CORE_H
#define FL3_CORE_H
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include "fl3_pack.h"
#include "fl3_kargs.h"

// ------------------------------------------------------------------ math

// FP32: hardware transcendentals with explicit range reduction.  sin/cos
// take the angle in revolutions reduced to [-0.5, 0.5]; pow is only ever
// called with a non-negative base by the variations.
__device__ __forceinline__ float rev_reduce(float x) {
   const float r = x * 0.15915494309189535f;
   return r - __builtin_rintf(r);
}
__device__ __forceinline__ float m_sin(float x) { return __builtin_amdgcn_sinf(rev_reduce(x)); }
__device__ __forceinline__ float m_cos(float x) { return __builtin_amdgcn_cosf(rev_reduce(x)); }
__device__ __forceinline__ void m_sincos(float x, float *s, float *c) {
   const float r = rev_reduce(x);
   *s = __builtin_amdgcn_sinf(r);
   *c = __builtin_amdgcn_cosf(r);
}
__device__ __forceinline__ float m_tan(float x) {
   const float r = rev_reduce(x);
   return __builtin_amdgcn_sinf(r) / __builtin_amdgcn_cosf(r);
}
// atan2 from an odd polynomial in t = min/max of |x|, |y| (degree 13, least
// squares on Chebyshev nodes with relative reweighting, fitted for this
// library: |error| < 2.5e-7 rad on [0, 1]) and the octant identities
// atan2 = pi/2 - atan(t), pi - ..., sign of y.
__device__ __forceinline__ float m_atan2(float y, float x) {
   const float ax = fabsf(x), ay = fabsf(y);
   const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
   // (mn <= mx: at the origin both are 0 and t is 0 * 1e30)
   const float t = mn * __builtin_amdgcn_rcpf(fmaxf(mx, 1e-30f));
   const float s = t * t;
   float p = 0.00681174332f;
   p = p * s - 0.0336040774f;
   p = p * s + 0.0796235211f;
   p = p * s - 0.13233335f;
   p = p * s + 0.198078143f;
   p = p * s - 0.33317368f;
   p = p * s + 0.999996112f;
   float r = p * t;
   if (ay > ax) r = 1.5707963267948966f - r;
   if (x < 0.0f) r = 3.141592653589793f - r;
   return copysignf(r, y);
}
// Square root and quotient from the hardware instructions (v_sqrt_f32,
// v_rcp_f32, 1 ulp each): a correctly rounded FP32 square root takes about
// fifteen instructions and an IEEE division about ten, where these take one
// and two; the walkers' FP32 error against the FP64 reference is some 1e-7
// relative either way.  The FP64 forms below are exact.
__device__ __forceinline__ float m_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float m_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float m_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
__device__ __forceinline__ float m_log(float x) { return __builtin_amdgcn_logf(x) * 0.6931471805599453f; }
__device__ __forceinline__ float m_log10(float x) { return __builtin_amdgcn_logf(x) * 0.30102999566398120f; }
__device__ __forceinline__ float m_pow(float x, float y) {
   if (x > 0.0f) return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
   return powf(x, y);
}
__device__ __forceinline__ float m_fmod(float x, float y) { return fmodf(x, y); }
__device__ __forceinline__ float m_floor(float x) { return __builtin_floorf(x); }
__device__ __forceinline__ float m_sinh(float x) {
   const float e = m_exp(x);
   return 0.5f * (e - 1.0f / e);
}
__device__ __forceinline__ float m_cosh(float x) {
   const float e = m_exp(x);
   return 0.5f * (e + 1.0f / e);
}
__device__ __forceinline__ float m_acos(float x) { return acosf(x); }
__device__ __forceinline__ float m_rint(float x) { return __builtin_rintf(x); }
__device__ __forceinline__ float m_fabs(float x) { return __builtin_fabsf(x); }

// FP64: the correctly rounded library functions.
__device__ __forceinline__ double m_sin(double x) { return sin(x); }
__device__ __forceinline__ double m_cos(double x) { return cos(x); }
__device__ __forceinline__ void m_sincos(double x, double *s, double *c) { sincos(x, s, c); }
__device__ __forceinline__ double m_tan(double x) { return tan(x); }
__device__ __forceinline__ double m_atan2(double y, double x) { return atan2(y, x); }
__device__ __forceinline__ double m_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ double m_div(double a, double b) { return a / b; }
__device__ __forceinline__ double m_pow(double x, double y) { return pow(x, y); }
__device__ __forceinline__ double m_exp(double x) { return exp(x); }
__device__ __forceinline__ double m_log(double x) { return log(x); }
__device__ __forceinline__ double m_log10(double x) { return log10(x); }
__device__ __forceinline__ double m_fmod(double x, double y) { return fmod(x, y); }
__device__ __forceinline__ double m_floor(double x) { return floor(x); }
__device__ __forceinline__ double m_sinh(double x) { return sinh(x); }
__device__ __forceinline__ double m_cosh(double x) { return cosh(x); }
__device__ __forceinline__ double m_acos(double x) { return acos(x); }
__device__ __forceinline__ double m_rint(double x) { return rint(x); }
__device__ __forceinline__ double m_fabs(double x) { return fabs(x); }

// NaN or magnitude above 1e10 (variations.c:22): a NaN fails the
// comparison, so one test of the magnitude covers both.
template <typename R>
__device__ __forceinline__ bool fl3_bad(R x) {
   return !(m_fabs(x) <= R(1e10));
}

// ------------------------------------------------------------------ RNGs

// ISAAC with RANDSIZL = 4 on 32-bit words.  The refill only ever uses the
// low 32 bits of the reference's 64-bit words (isaac.c:17-23), so once the
// state has been seeded on the host (where the 64-bit seeding scramble is
// reproduced) this is the reference stream exactly.
struct IsaacRng {
   uint32_t rsl[16], mem[16], a, b, c, cnt;

   __device__ void load(const uint32_t *s) {
      cnt = s[0];
      for (int i = 0; i < 16; i++) rsl[i] = s[1 + i];
      for (int i = 0; i < 16; i++) mem[i] = s[17 + i];
      a = s[33]; b = s[34]; c = s[35];
   }
   __device__ void store(uint32_t *s) const {
      s[0] = cnt;
      for (int i = 0; i < 16; i++) s[1 + i] = rsl[i];
      for (int i = 0; i < 16; i++) s[17 + i] = mem[i];
      s[33] = a; s[34] = b; s[35] = c;
   }
   __device__ void refill() {
      uint32_t aa = a, bb = b + (++c);
      for (int i = 0; i < 16; i++) {
         uint32_t x = mem[i], mixv;
         switch (i & 3) {
            case 0: mixv = aa << 13; break;
            case 1: mixv = aa >> 6; break;
            case 2: mixv = aa << 2; break;
            default: mixv = aa >> 16; break;
         }
         aa = (aa ^ mixv) + mem[(i + 8) & 15];
         uint32_t y = mem[(x >> 2) & 15] + aa + bb;
         mem[i] = y;
         bb = mem[(y >> 6) & 15] + x;  // ind(mm, y >> RANDSIZL): ind shifts by 2 again
         rsl[i] = bb;
      }
      a = aa;
      b = bb;
   }
   __device__ __forceinline__ uint32_t next() {
      if (cnt-- == 0) {
         refill();
         cnt = 15;
      }
      return rsl[cnt];
   }
};

// Per-walker production stream: xoshiro128** (Blackman & Vigna), state
// derived with splitmix64 from the 128-bit key drawn from the frame's ISAAC
// generator and the walker index.
struct WalkerRng {
   uint32_t s0, s1, s2, s3;

   __device__ static uint64_t splitmix(uint64_t &x) {
      uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
      return z ^ (z >> 31);
   }
   __device__ void seed(uint64_t k0, uint64_t k1, uint64_t id) {
      uint64_t x = k0 ^ (id * 0xD1B54A32D192ED03ULL);
      uint64_t y = k1 + id;
      uint64_t a = splitmix(x) ^ splitmix(y);
      uint64_t b = splitmix(x) ^ splitmix(y);
      s0 = (uint32_t)a; s1 = (uint32_t)(a >> 32);
      s2 = (uint32_t)b; s3 = (uint32_t)(b >> 32);
      if ((s0 | s1 | s2 | s3) == 0) s0 = 1;
   }
   __device__ __forceinline__ uint32_t next() {
      uint32_t r = s1 * 5u;
      r = ((r << 7) | (r >> 25)) * 9u;
      uint32_t t = s1 << 9;
      s2 ^= s0; s3 ^= s1; s1 ^= s2; s0 ^= s3;
      s2 ^= t;
      s3 = (s3 << 11) | (s3 >> 21);
      return r;
   }
};

// 28-bit uniforms as in flam3.c:2519-2525.
template <typename R, class RNG>
__device__ __forceinline__ R u01(RNG &r) {
   return (R)(r.next() & 0xfffffffu) / (R)0xfffffff;
}
template <typename R, class RNG>
__device__ __forceinline__ R u11(RNG &r) {
   return (R)((int)(r.next() & 0xfffffffu) - 0x7ffffff) / (R)0x7ffffff;
}

// ------------------------------------------------------------------ genome in LDS

// Dynamic LDS of every kernel that holds a genome.
extern __shared__ __align__(16) unsigned char fl3_lds[];

template <typename R>
struct XfL {
   R c[6], post[6];
   R color, cspeed, vis, opacity, preblur;
   int flags;
   R p[FL3_NPARAM];
};

template <typename R>
struct GenL {
   const fl3_job_hdr *h;   // global, read with uniform loads
   const XfL<R> *xf;       // LDS
   const R *w;             // LDS [nxf * nunion]
   const int *starts;      // LDS [nrows * nstd]
   const int *uid;         // LDS [nunion]
   const R *dmap;          // LDS [256*4]
   const uint32_t *dmapu;  // LDS [256*4] floor(dmap) for unit visibility
   const double *dmapd;    // LDS [256*4] FP64 copy (FP64 histograms)
   const uint8_t *vlist;   // LDS [nxf * nunion] union slots with nonzero weight, ascending
   const uint8_t *vcount;  // LDS [nxf]
   // global, read with scalar loads: per xform the count of nonzero weights,
   // then (variation id, float weight bits) pairs (fl3_job_hdr::off_vl)
   const __attribute__((address_space(4))) int32_t *vlc;
   int nxf, nstd, final_index, chaos_enable, nunion, cmap_size, palette_mode, rotate, width, height;
   R b0, b1, b2, b3, rot00, rot01, rot10, rot11, rc0, rc1, ws0, wb0s0, hs1, hb1s1, final_opacity;
};

// LDS image of a genome: xforms, weights, selection breakpoints, union ids,
// variation lists, then (only for kernels that deposit colours themselves)
// the palette in three forms.
template <typename R>
static size_t lds_bytes_for(const fl3_job_hdr *h, bool palette = true) {
   size_t n = 0;
   n += sizeof(XfL<R>) * h->nxf;
   n += sizeof(R) * h->nxf * (size_t)(h->nunion > 0 ? h->nunion : 1);
   n += sizeof(int) * (size_t)h->nrows * h->nstd;
   n += sizeof(int) * FL3_NVAR;
   n += (size_t)h->nxf * (size_t)(h->nunion > 0 ? h->nunion : 1) + (size_t)h->nxf;
   n = (n + 15) & ~(size_t)15;
   if (palette) n += sizeof(R) * 1024 + sizeof(uint32_t) * 1024 + sizeof(double) * 1024;
   return n + 64;
}

// Copy the job into LDS, converting to R.  All threads participate.
// Without `palette` the palette is left out (the LDS image is then
// lds_bytes_for(h, false) long).
template <typename R>
__device__ void load_genome(const unsigned char *job, unsigned char *lds, GenL<R> &G, bool palette = true) {
   const fl3_job_hdr *h = (const fl3_job_hdr *)job;
   G.h = h;
   G.nxf = h->nxf; G.nstd = h->nstd; G.final_index = h->final_index;
   G.chaos_enable = h->chaos_enable; G.nunion = h->nunion; G.cmap_size = h->cmap_size;
   G.palette_mode = h->palette_mode; G.rotate = h->rotate; G.width = h->width; G.height = h->height;
   G.b0 = (R)h->bounds[0]; G.b1 = (R)h->bounds[1]; G.b2 = (R)h->bounds[2]; G.b3 = (R)h->bounds[3];
   G.rot00 = (R)h->rot[0]; G.rot01 = (R)h->rot[1]; G.rot10 = (R)h->rot[2]; G.rot11 = (R)h->rot[3];
   G.rc0 = (R)h->rot_center[0]; G.rc1 = (R)h->rot_center[1];
   G.ws0 = (R)h->ws0; G.wb0s0 = (R)h->wb0s0; G.hs1 = (R)h->hs1; G.hb1s1 = (R)h->hb1s1;
   G.final_opacity = (R)h->final_opacity;

   size_t off = 0;
   XfL<R> *xf = (XfL<R> *)(lds + off); off += sizeof(XfL<R>) * G.nxf;
   R *w = (R *)(lds + off); off += sizeof(R) * G.nxf * (size_t)(G.nunion > 0 ? G.nunion : 1);
   int *st = (int *)(lds + off); off += sizeof(int) * (size_t)h->nrows * G.nstd;
   int *uid = (int *)(lds + off); off += sizeof(int) * FL3_NVAR;
   uint8_t *vl = (uint8_t *)(lds + off); off += (size_t)G.nxf * (G.nunion > 0 ? G.nunion : 1);
   uint8_t *vc = (uint8_t *)(lds + off); off += (size_t)G.nxf;
   off = (off + 15) & ~(size_t)15;
   R *dmap = (R *)(lds + off); off += sizeof(R) * 1024;
   uint32_t *dmapu = (uint32_t *)(lds + off); off += sizeof(uint32_t) * 1024;
   double *dmapd = (double *)(lds + off);

   const fl3_xform_pack *px = (const fl3_xform_pack *)(job + h->off_xf);
   const double *pw = (const double *)(job + h->off_w);
   const int *ps = (const int *)(job + h->off_starts);
   const double *pd = (const double *)(job + h->off_dmap);
   const int tid = threadIdx.x, nt = blockDim.x;

   // xforms: one element per thread-step, flattening each pack to doubles.
   const int per = (int)(sizeof(fl3_xform_pack) / sizeof(double));
   for (int e = tid; e < G.nxf * per; e += nt) {
      int x = e / per, k = e % per;
      const double *src = (const double *)&px[x];
      XfL<R> &d = xf[x];
      if (k < 6) d.c[k] = (R)src[k];
      else if (k < 12) d.post[k - 6] = (R)src[k];
      else if (k == 12) d.color = (R)src[k];
      else if (k == 13) d.cspeed = (R)src[k];
      else if (k == 14) d.vis = (R)src[k];
      else if (k == 15) d.opacity = (R)src[k];
      else if (k == 16) d.preblur = (R)src[k];
      else if (k == 17) d.flags = px[x].flags;
      else d.p[k - 18] = (R)src[k];
   }
   for (int e = tid; e < G.nxf * G.nunion; e += nt) w[e] = (R)pw[e];
   for (int e = tid; e < h->nrows * G.nstd; e += nt) st[e] = ps[e];
   for (int e = tid; e < FL3_NVAR; e += nt) uid[e] = h->union_ids[e];
   if (palette)
      for (int e = tid; e < 1024; e += nt) {
         double v = pd[e];
         dmap[e] = (R)v;
         dmapu[e] = (uint32_t)v;
         dmapd[e] = v;
      }
   __syncthreads();
   for (int x = tid; x < G.nxf; x += nt) {
      int k = 0;
      for (int u = 0; u < G.nunion; u++)
         if (w[x * G.nunion + u] != R(0)) vl[x * G.nunion + k++] = (uint8_t)u;
      vc[x] = (uint8_t)k;
   }
   G.xf = xf; G.w = w; G.starts = st; G.uid = uid; G.dmap = dmap; G.dmapu = dmapu; G.dmapd = dmapd;
   G.vlist = vl; G.vcount = vc;
   G.vlc = (const __attribute__((address_space(4))) int32_t *)(job + h->off_vl);
}

#include "fl3_variations.inc"

// Xform choice: the entry of the reference's 16384-slot table at `slot`,
// from the row's step breakpoints (flam3.c:252-255).
template <typename R>
__device__ __forceinline__ int select_xform(const GenL<R> &G, int row, uint32_t slot) {
   const int *st = G.starts + row * G.nstd;
   int fn = 0;
   for (int j = 1; j < G.nstd; j++) fn += (slot >= (uint32_t)st[j]) ? 1 : 0;
   return fn;
}

// One application of xform fn to p (variations.c:2129-2399).  Returns true
// when the result is a bad value; q[0..1] then hold a random point.
// Variation tiers of the walker kernels.  Tier 0 holds the classic ids
// 0-13 (linear .. julia); tier t in 1..FL3_NBLOCKS adds block t-1 of the
// remaining ids; FL3_TIER_FULL holds all of them.
#define FL3_CLASSIC 14
#define FL3_BLOCK 17
#define FL3_NBLOCKS 5
#define FL3_TIER_FULL (FL3_NBLOCKS + 1)
__host__ __device__ constexpr int fl3_tier_lo(int t) { return FL3_CLASSIC + (t - 1) * FL3_BLOCK; }
__host__ __device__ constexpr int fl3_tier_hi(int t) {
   return FL3_CLASSIC + t * FL3_BLOCK < FL3_NVAR ? FL3_CLASSIC + t * FL3_BLOCK : FL3_NVAR;
}

template <typename R, class RNG, int TIER>
__device__ __forceinline__ void fl3_var_tier(int v, R w, const XfL<R> &X, Pre<R> &f, RNG &rng) {
   if constexpr (TIER == FL3_TIER_FULL) {
      fl3_var<R, RNG>(v, w, X, f, rng);
   } else if constexpr (TIER == 0) {
      fl3_var<R, RNG, 0, FL3_CLASSIC>(v, w, X, f, rng);
   } else {
      if (v < FL3_CLASSIC)
         fl3_var<R, RNG, 0, FL3_CLASSIC>(v, w, X, f, rng);
      else
         fl3_var<R, RNG, fl3_tier_lo(TIER), fl3_tier_hi(TIER)>(v, w, X, f, rng);
   }
}

// The variation inputs made opaque at each step of a variation loop: they
// do not change inside the loop, and without this the compiler hoists the
// invariant part of every case of the switch (sines, cosines, roots of all
// the tier's variations) out of the loop and evaluates them all at every
// iteration of the walker, used or not.
template <typename R>
__device__ __forceinline__ void opaque_pre(Pre<R> &f) {
   if constexpr (sizeof(R) == 4)
      asm volatile("" : "+v"(f.tx), "+v"(f.ty), "+v"(f.sumsq), "+v"(f.sq), "+v"(f.atan), "+v"(f.sina),
                   "+v"(f.cosa), "+v"(f.atanyx));
}

// UNI: fn is the same in every lane of the wavefront (the final xform), so
// the loop runs over that xform's own variations only.
template <typename R, class RNG, int TIER = FL3_TIER_FULL, bool UNI = false>
__device__ bool apply_xform(const GenL<R> &G, int fn, const R *p, R *q, RNG &rng) {
   const XfL<R> &X = G.xf[fn];
   Pre<R> f;
   R s1 = X.cspeed;
   if constexpr (sizeof(R) == 4)   // (the same blend, p + s (c - p), in one fused multiply-add)
      q[2] = fmaf(s1, X.color - p[2], p[2]);
   else
      q[2] = s1 * X.color + (R(1.0) - s1) * p[2];
   q[3] = X.vis;
   if constexpr (sizeof(R) == 4) {   // (two fused multiply-adds per coordinate)
      f.tx = fmaf(X.c[0], p[0], fmaf(X.c[2], p[1], X.c[4]));
      f.ty = fmaf(X.c[1], p[0], fmaf(X.c[3], p[1], X.c[5]));
   } else {
      f.tx = X.c[0] * p[0] + X.c[2] * p[1] + X.c[4];
      f.ty = X.c[1] * p[0] + X.c[3] * p[1] + X.c[5];
   }
   if (X.preblur != R(0)) {  // pre_blur :1480
      R g0 = u01<R>(rng), g1 = u01<R>(rng), g2 = u01<R>(rng), g3 = u01<R>(rng);
      R rndG = X.preblur * (g0 + g1 + g2 + g3 - R(2.0));
      R rndA = u01<R>(rng) * R(2.0) * R(M_PI), sA, cA;
      m_sincos(rndA, &sA, &cA);
      f.tx += rndG * cA;
      f.ty += rndG * sA;
   }
   f.sumsq = f.tx * f.tx + f.ty * f.ty;
   f.sq = m_sqrt(f.sumsq);
   f.atan = f.sina = f.cosa = f.atanyx = R(0);
   const int flags = X.flags;
   if (flags & FL3_PRE_ATAN_XY) f.atan = m_atan2(f.tx, f.ty);
   if (flags & FL3_PRE_ANGLES) {
      f.sina = m_div(f.tx, f.sq);
      f.cosa = m_div(f.ty, f.sq);
   }
   if (flags & FL3_PRE_ATAN_YX) f.atanyx = m_atan2(f.ty, f.tx);
   f.p0 = R(0);
   f.p1 = R(0);
   const R *wrow = G.w + fn * G.nunion;
   // The xform's LDS offset is made opaque inside the loop so each
   // variation loads its parameters where it uses them instead of the
   // compiler hoisting every case's loads (and registers) out of the loop.
   const uint32_t xoff = (uint32_t)((const unsigned char *)&X - fl3_lds);
   if (UNI) {
      const int fu = __builtin_amdgcn_readfirstlane(fn);
      if constexpr (sizeof(R) == 4) {
         // the xform's variation ids and weights through scalar loads of
         // the job's list (the same floats, in the same order, as the LDS
         // tables hold)
         const __attribute__((address_space(4))) int32_t *vl = G.vlc + fu * (2 + 2 * G.nunion);
         const int nv = vl[0];
         if constexpr (TIER == 0) {
            // Of the classic variations only julia draws (one word per
            // application): its word is drawn before the loop, so the lane's
            // stream does not change inside the variation switch (where its
            // state would be merged after every case).
            struct WordRng {
               uint32_t w;
               __device__ uint32_t next() { return w; }
            } wr;
            wr.w = (flags & FL3_DRAWS_JULIA) ? rng.next() : 0u;
            // (the classic cases read no xform parameters, so the xform
            // needs no opaque offset here; the first four pairs come in one
            // scalar load before the loop -- past the list's end, its
            // padding -- so that the loop waits for no load)
            const __attribute__((address_space(4))) int32_t *pv = vl + 2;
            int pr[8];
            for (int q = 0; q < 8; q++) pr[q] = pv[q];
            for (int k = 0; k < nv; k++) {
               const int v = k < 4 ? pr[2 * k] : pv[2 * k];
               const int wb = k < 4 ? pr[2 * k + 1] : pv[2 * k + 1];
               opaque_pre(f);
               fl3_var_tier<R, WordRng, TIER>(v, __int_as_float(wb), X, f, wr);
            }
         } else
            for (int k = 0; k < nv; k++) {
               const int v = vl[2 + 2 * k];
               const R w = __int_as_float(vl[3 + 2 * k]);
               uint32_t o = xoff;
               asm volatile("" : "+v"(o));
               opaque_pre(f);
               fl3_var_tier<R, RNG, TIER>(v, w, *(const XfL<R> *)(fl3_lds + o), f, rng);
            }
      } else {
         const int nv = __builtin_amdgcn_readfirstlane(G.vcount[fu]);
         for (int k = 0; k < nv; k++) {
            const int u = __builtin_amdgcn_readfirstlane((int)G.vlist[fu * G.nunion + k]);
            const int v = __builtin_amdgcn_readfirstlane(G.uid[u]);
            const R w = wrow[u];
            uint32_t o = xoff;
            asm volatile("" : "+v"(o));
            opaque_pre(f);
            fl3_var_tier<R, RNG, TIER>(v, w, *(const XfL<R> *)(fl3_lds + o), f, rng);
         }
      }
   } else {
      for (int u = 0; u < G.nunion; u++) {
         const int v = __builtin_amdgcn_readfirstlane(G.uid[u]);
         const R w = wrow[u];
         uint32_t o = xoff;
         asm volatile("" : "+v"(o));
         opaque_pre(f);
         const R s0 = f.p0, s1 = f.p1;
         fl3_var_tier<R, RNG, TIER>(v, w, *(const XfL<R> *)(fl3_lds + o), f, rng);
         if (sizeof(R) == 4) {  // FP32 cases ran in every lane: keep the lanes that use the variation
            f.p0 = w != R(0) ? f.p0 : s0;
            f.p1 = w != R(0) ? f.p1 : s1;
         }
      }
   }
   if (flags & FL3_HAS_POST) {
      if constexpr (sizeof(R) == 4) {
         q[0] = fmaf(X.post[0], f.p0, fmaf(X.post[2], f.p1, X.post[4]));
         q[1] = fmaf(X.post[1], f.p0, fmaf(X.post[3], f.p1, X.post[5]));
      } else {
         q[0] = X.post[0] * f.p0 + X.post[2] * f.p1 + X.post[4];
         q[1] = X.post[1] * f.p0 + X.post[3] * f.p1 + X.post[5];
      }
   } else {
      q[0] = f.p0;
      q[1] = f.p1;
   }
   if (fl3_bad(q[0]) || fl3_bad(q[1])) {
      q[0] = u11<R>(rng);
      q[1] = u11<R>(rng);
      return true;
   }
   return false;
}

// ------------------------------------------------------------------ deposit

__device__ __forceinline__ void sat_add(uint32_t *p, uint32_t d) {
   uint32_t old = atomicAdd(p, d);
   if (old > 0xffffffffu - d) atomicExch(p, 0xffffffffu);
}

// Colour of a point from the palette (rect.c:468-498).
template <typename R>
__device__ __forceinline__ void point_color(const GenL<R> &G, R c, int &ci, R &frac) {
   R di = c * (R)G.cmap_size;
   ci = (int)di;
   const int m1 = G.cmap_size - 1;
   if (G.palette_mode == 1) {
      if (ci < 0) { ci = 0; frac = R(0); }
      else if (ci >= m1) { ci = m1 - 1; frac = R(1.0); }
      else frac = di - (R)ci;
   } else {
      if (ci < 0) ci = 0;
      else if (ci >= m1) ci = m1;
      frac = R(0);
   }
}

// Deposit one plotted point (rect.c:419-516).  BITS 32/33: uint32 cells of
// 5 channels with truncated, saturating adds; 64: double cells.  The cell
// index is computed like the reference (truncation of ws0*x - wb0s0) with
// contraction off, and cells outside the histogram are dropped.
template <typename R, int BITS, bool EXACT>
__device__ __forceinline__ void deposit(const GenL<R> &G, const R *q, void *buckets, uint32_t *sat) {
#pragma clang fp contract(off)
   R x = q[0], y = q[1];
   if (G.rotate) {
      R x0 = x - G.rc0, y0 = y - G.rc1;
      x = x0 * G.rot00 + y0 * G.rot01 + G.rc0;
      y = x0 * G.rot10 + y0 * G.rot11 + G.rc1;
   }
   if (!(x >= G.b0 && y >= G.b1 && x <= G.b2 && y <= G.b3)) return;
   const R vis = q[3];
   if (vis == R(0)) return;
   const long ix = (long)(int)(G.ws0 * x - G.wb0s0);
   const long iy = (long)(int)(G.hs1 * y - G.hb1s1);
   const long cell = ix + (long)G.width * iy;
   if (cell < 0 || cell >= (long)G.width * G.height) return;
   int ci;
   R frac;
   point_color(G, q[2], ci, frac);
   if (BITS == 64) {
      double *b = (double *)buckets + cell * 5;
      const double *dm = G.dmapd + ci * 4;
      double v = (double)vis, col[4];
      for (int k = 0; k < 4; k++) {
         col[k] = dm[k];
         if (G.palette_mode == 1) col[k] = dm[k] * (1.0 - (double)frac) + dm[4 + k] * (double)frac;
      }
      if (vis == R(1.0)) {
         for (int k = 0; k < 4; k++) unsafeAtomicAdd(b + k, col[k]);
         unsafeAtomicAdd(b + 4, 255.0);
      } else {
         for (int k = 0; k < 4; k++) unsafeAtomicAdd(b + k, v * col[k]);
         unsafeAtomicAdd(b + 4, v * 255.0);
      }
      return;
   }
   uint32_t *b = (uint32_t *)buckets + cell * 5;
   uint32_t d[5];
   if (G.palette_mode == 1) {
      const R *dm = G.dmap + ci * 4;
      for (int k = 0; k < 4; k++) {
         R col = dm[k] * (R(1.0) - frac) + dm[4 + k] * frac;
         d[k] = (uint32_t)(vis == R(1.0) ? col : vis * col);
      }
   } else if (vis == R(1.0)) {
      const uint32_t *du = G.dmapu + ci * 4;
      for (int k = 0; k < 4; k++) d[k] = du[k];
   } else {
      const R *dm = G.dmap + ci * 4;
      for (int k = 0; k < 4; k++) d[k] = (uint32_t)(vis * dm[k]);
   }
   d[4] = (vis == R(1.0)) ? 255u : (uint32_t)(vis * R(255.0));
   if (EXACT) {
      for (int k = 0; k < 5; k++) sat_add(b + k, d[k]);
   } else {
      // Colour deltas never exceed the count delta (palette <= 1), so only
      // the count channel can reach UINT_MAX first; if it does, the caller
      // repeats the batch in exact mode.
      for (int k = 0; k < 4; k++) atomicAdd(b + k, d[k]);
      uint32_t old = atomicAdd(b + 4, d[4]);
      if (old > 0xffffffffu - d[4]) *sat = 1u;
   }
}

// ------------------------------------------------------------------ binned deposit records
//
// A plotted point reduced to what its deposit needs: bits 0-31 cell index
// (FL3_REC_NONE: nothing to deposit), 32-39 palette index, 40-47 index of
// the xform whose visibility the point carries, 48-63 the linear-palette
// blend fraction in 16-bit fixed point.
#define FL3_REC_NONE 0xffffffffu

// STEPPED: the caller knows the palette is stepped (4-byte records), so the
// palette index is clamped without the linear mode's branches.
template <typename R, bool STEPPED = false>
__device__ __forceinline__ unsigned long long make_record(const GenL<R> &G, const R *q, int fn) {
#pragma clang fp contract(off)
   R x = q[0], y = q[1];
   if (__builtin_amdgcn_readfirstlane(G.rotate)) {   // (the same in every lane: a scalar branch)
      R x0 = x - G.rc0, y0 = y - G.rc1;
      x = x0 * G.rot00 + y0 * G.rot01 + G.rc0;
      y = x0 * G.rot10 + y0 * G.rot11 + G.rc1;
   }
   // (4-byte records carry points of visibility 1 only: no visibility test)
   if (!(x >= G.b0 && y >= G.b1 && x <= G.b2 && y <= G.b3) || (!STEPPED && q[3] == R(0))) return FL3_REC_NONE;
   // the reference's linear index ix + width * iy (rect.c:440-441) in 32-bit
   // unsigned arithmetic: histograms here stay below 2^31 cells, so a
   // negative index wraps above the cell count and is dropped like one past
   // the end.  Inside the bounds iy is in [0, height] (a truncation towards
   // zero of at least -1 ulp), below 2^24, so a 24-bit multiply forms the
   // same product.
   const int ix = (int)(G.ws0 * x - G.wb0s0);
   const int iy = (int)(G.hs1 * y - G.hb1s1);
   const unsigned cell = (unsigned)ix + __umul24((unsigned)G.width, (unsigned)iy);
   if (cell >= (unsigned)G.width * (unsigned)G.height) return FL3_REC_NONE;
   if constexpr (STEPPED) {   // point_color's stepped branch
      const int c = (int)(q[2] * (R)G.cmap_size), m1 = G.cmap_size - 1;
      return (unsigned long long)cell | ((unsigned long long)(c < 0 ? 0 : (c >= m1 ? m1 : c)) << 32);
   }
   int ci;
   R frac;
   point_color(G, q[2], ci, frac);
   R f = frac < R(0) ? R(0) : (frac > R(1) ? R(1) : frac);
   const unsigned long long qf = (unsigned long long)(f * R(65535.0) + R(0.5));
   return (unsigned long long)cell | ((unsigned long long)ci << 32) | ((unsigned long long)fn << 40) | (qf << 48);
}

// Sum of per-walker bad-value counts: wave reduction, one atomic per wave.
__device__ __forceinline__ void add_bad(unsigned long long *dst, uint32_t bad) {
   unsigned long long v = bad;
   for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
   if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, v);
}

#endif
