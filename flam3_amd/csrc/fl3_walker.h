// fl3_walker.h - the FP32 production walker kernel, instantiated per
// variation tier by fl3_kernels_fast.hip (full set) and fl3_kernels_tier.hip
// (the classic set and the classic set plus one block).
#ifndef FL3_WALKER_H
#define FL3_WALKER_H
#include "fl3_core.h"

// Record formats of the binned deposit (RECORDS template argument).
#define FL3_REC_DIRECT 0   // no records: deposit with atomics
#define FL3_REC_WIDE 1     // 8-byte records (make_record)
#define FL3_REC_NARROW 2   // 4-byte records: cell | palette index << 24
#define FL3_REC_JOB 3      // 8-byte records: cell, palette index, xform, job
                           // (or cell and the palette entry's four bytes)

// Records: row i of walker (workgroup b, thread t) at (b * rows_out + i) *
// 256 + t, so a workgroup's records are one contiguous block.
// Production walkers.  Walker (lane) w plots nsamples/W (+1 for the first
// nsamples%W walkers) points after `fuse` unplotted steps from a random
// start, the per-sub-batch procedure of rect.c:393-400 with one long
// sub-batch per walker; with WAVE (genomes without xaos) the xform is drawn
// per wavefront and the workgroup's points change lanes after every step.  In
// record mode each workgroup also counts its records per tile in LDS and
// stores them per super-tile at blk_hist[block][super-tile].
// Steps between two exchanges of the points among a workgroup's lanes in
// the one-xform-per-wavefront mode (a power of two).
#ifndef FL3_PERM_EVERY
#define FL3_PERM_EVERY 1u
#endif

// Waves per SIMD the walker kernel is compiled for: the classic tier (the
// variations of most genomes) at 8, which holds it to 64 VGPRs (a few
// spills; 7 waves at its natural 67-78 VGPRs ran 11% slower per launch),
// the larger tiers at their natural register counts.
template <int TIER>
constexpr int walker_waves() { return TIER == 0 ? 8 : 1; }

template <int BITS, bool EXACT, int RECORDS, int TIER, bool WAVE>
__global__ __launch_bounds__(256, walker_waves<TIER>()) void k_iter_fast(IterArgs a) {
   unsigned char *lds = fl3_lds;
   GenL<float> G;
   // independent renders: direct deposits, or job or 4-byte records in one
   // cell space
   const bool dmode = RECORDS != FL3_REC_WIDE && a.descs;
   const int djob = dmode ? a.blk_job[blockIdx.x] : 0;
   const JobDesc *dsc = dmode ? a.descs + djob : nullptr;
   const int nj = !dsc && a.njobs > 1 ? a.njobs : 1;
   const int jb = nj > 1 ? (int)(blockIdx.x % nj) : 0;
   // record modes leave the colours to the tile accumulation: no palette in LDS
   load_genome<float>(dsc ? a.jobs + dsc->off : (nj > 1 ? a.jobs + a.job_off[jb] : a.job), lds, G,
                      RECORDS == FL3_REC_DIRECT);
   unsigned *bh = (unsigned *)(lds + a.hist_off);
   if (RECORDS)
      for (int b = threadIdx.x; b < a.nbins; b += blockDim.x) bh[b] = 0;
   uint32_t *pal = (uint32_t *)(lds + a.pal_off);
   if (RECORDS == FL3_REC_JOB && a.pal8) {
      const uint32_t *src = a.pal8 + (size_t)(dsc ? djob : jb) * 256;
      for (int e = threadIdx.x; e < 256; e += blockDim.x) pal[e] = src[e];
   }
   __syncthreads();
   // W, wid: the record matrix and walker-state layout over all jobs (with
   // independent renders the launch may cover only the columns from
   // workgroup blk0 on); Wj, wj: this walker's job-local count and index
   const long long W = a.wtot ? a.wtot : (long long)gridDim.x * blockDim.x;
   const long long blk = (long long)a.blk0 + blockIdx.x;
   const long long wid = blk * blockDim.x + threadIdx.x;
   const long long Wj = dsc ? (long long)dsc->nblocks * blockDim.x : (long long)(gridDim.x / nj) * blockDim.x;
   const long long wj = dsc ? (long long)(blockIdx.x - dsc->block0) * blockDim.x + threadIdx.x
                            : (long long)(blockIdx.x / nj) * blockDim.x + threadIdx.x;
   long long nsj = dsc ? dsc->nsamples : a.nsamples;
   if (!dsc && a.job_ns) {   // this chunk's share of the job's own count
      const long long rem = a.job_ns[jb] - a.ns_row0 * Wj;
      nsj = rem < 0 ? 0 : (rem < nsj ? rem : nsj);
   }
   long long n = nsj / Wj + (wj < nsj % Wj ? 1 : 0);
   // independent renders with records run in chunks of rows [row0, row0 + rows_out)
   if (dsc && RECORDS) n = n - a.row0 < 0 ? 0 : (n - a.row0 < a.rows_out ? n - a.row0 : a.rows_out);
   const unsigned long long cell0 = dsc ? (unsigned long long)dsc->cell0 : 0ull;
   // deposit target and counters: the job's render
   void *const buckets = !dsc ? a.buckets
                         : BITS == 64 ? (void *)((double *)a.buckets + dsc->cell0 * 5)
                                      : (void *)((uint32_t *)a.buckets + dsc->cell0 * 5);
   unsigned long long *const badp = dsc ? a.badvals + 2 * dsc->render : a.badvals;
   uint32_t *const satp = dsc ? (uint32_t *)(a.badvals + 2 * dsc->render + 1) : a.sat;
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
      if (dsc)
         rng.seed(dsc->key0, dsc->key1, (uint64_t)wj);
      else if (nj > 1)
         rng.seed(a.keys[2 * jb], a.keys[2 * jb + 1], (uint64_t)wj);
      else
         rng.seed(a.key0, a.key1, (uint64_t)wid);
      p[0] = u11<float>(rng);
      p[1] = u11<float>(rng);
      p[2] = u01<float>(rng);
      p[3] = u01<float>(rng);
   }
   uint32_t bad = 0;
   const bool use_final = G.final_index >= 0;
   // the record (or deposit) of plotted point v at row i
   auto plot = [&](long long i, const float *v, int fn) {
      if (RECORDS == FL3_REC_WIDE) {
         const unsigned long long r = make_record<float>(G, v, fn);
         a.rec[(blk * a.rows_out + i) * 256 + threadIdx.x] = r;
         if ((unsigned)r != FL3_REC_NONE) atomicAdd(&bh[(unsigned)r >> a.bin_shift], 1u);
      } else if (RECORDS == FL3_REC_JOB) {
         // independent renders: the render's first cell moves the cell
         // into the launch's cell space, the job is the descriptor's
         // (palettes as bytes: every job's palette is stepped and every
         // visibility 1, the 4-byte records' case)
         const unsigned long long r = a.pal8 ? make_record<float, true>(G, v, fn) : make_record<float>(G, v, fn);
         const unsigned cell = (unsigned)r == FL3_REC_NONE ? FL3_REC_NONE : (unsigned)(r + cell0);
         const unsigned long long hi = a.pal8 ? (unsigned long long)pal[(r >> 32) & 0xff] << 32
                                              : (r & 0xffff00000000ull) | ((unsigned long long)(dsc ? djob : jb) << 48);
         a.rec[(blk * a.rows_out + i) * 256 + threadIdx.x] =
             cell == FL3_REC_NONE ? (unsigned long long)FL3_REC_NONE : (hi | cell);
         if (cell != FL3_REC_NONE) atomicAdd(&bh[cell >> a.bin_shift], 1u);
      } else if (RECORDS == FL3_REC_NARROW) {
         // (independent renders sharing one palette: the render's first
         // cell moves the cell into the launch's cell space)
         const unsigned long long r = make_record<float, true>(G, v, 0);   // (4-byte records: a stepped palette)
         const unsigned cell = (unsigned)r == FL3_REC_NONE ? FL3_REC_NONE : (unsigned)(r + cell0);
         ((unsigned *)a.rec)[(blk * a.rows_out + i) * 256 + threadIdx.x] =
             cell == FL3_REC_NONE ? FL3_REC_NONE : (cell | ((unsigned)(r >> 32) & 0xffu) << 24);
         if (cell != FL3_REC_NONE) atomicAdd(&bh[cell >> a.bin_shift], 1u);
      } else
         deposit<float, BITS, EXACT>(G, v, buckets, satp);
   };
   if constexpr (WAVE) {
      // One xform per wavefront and step (drawn from a hash of the
      // wavefront's seed, the step and the draw within the step, in scalar
      // registers), so a wavefront evaluates that xform's variations only; after
      // every step the workgroup's 256 points move to other lanes through a
      // random permutation (slot = m (t ^ x) + o mod 256, m odd, from a hash
      // of the job's key, the job-local workgroup and the step), so the
      // xforms a point meets are independent draws with the xforms'
      // probabilities, as in the per-lane chaos game.  Lanes keep their own
      // streams and sample counts; a lane past its count keeps moving points
      // without plotting until the workgroup is done.
      float *xb = (float *)(lds + a.shuf_off);
      const uint64_t kk = dsc ? dsc->key0 : (nj > 1 ? a.keys[2 * jb] : a.key0);
      const uint32_t bjob = dsc ? (uint32_t)(blockIdx.x - dsc->block0)
                                : (nj > 1 ? (uint32_t)(blockIdx.x / nj) : (uint32_t)blk);
      const uint32_t hseed = ((uint32_t)kk ^ (uint32_t)(kk >> 32) * 0x85ebca6bu) ^ bjob * 0xc2b2ae35u;
      // the wavefront's xform draws: draw k of step st is
      // fmix32(wseed ^ (st * 0xcc9e2d51 + k * 0x1b873593)), the 32-bit
      // finaliser of MurmurHash3 (Appleby), over a seed mixed from the
      // workgroup's and the wavefront's index (so it depends on the job-local
      // workgroup, like the permutation, and not on the launch)
      auto fmix32 = [](uint32_t h) {
         h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
         return h;
      };
      const uint32_t wseed = fmix32(hseed + 0x632be5abu * ((uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) + 1u));
      auto wdraw = [&](uint32_t step, uint32_t k) { return fmix32(wseed ^ (step * 0xcc9e2d51u + k * 0x1b873593u)); };
      // the xform at the first slot of each 64-slot bucket of the 14-bit
      // selection slot: a draw starts there and steps over the few
      // breakpoints inside its bucket (select_xform's count, in fewer steps)
      unsigned short *sel = (unsigned short *)(xb + 1536);
      {
         const uint32_t b0 = threadIdx.x * 64u;
         int f = 0;
         for (int j = 1; j < G.nstd; j++) f += b0 >= (uint32_t)G.starts[j] ? 1 : 0;
         sel[threadIdx.x] = (unsigned short)f;
      }
      __syncthreads();
      // (the step count and the workgroup's step bound are the same in
      // every lane: read from the first lane, they live in scalar
      // registers, and so does the permutation's hash below)
      uint32_t st = a.resume ? (uint32_t)__builtin_amdgcn_readfirstlane(lastxf) : 0u;
      // every lane takes one step at a time, so i is the same in all lanes
      // and the workgroup runs until its first walker (the one with the
      // most samples) is done
      long long nwg = nsj / Wj + ((wj - (long long)threadIdx.x) < nsj % Wj ? 1 : 0);
      if (dsc && RECORDS) nwg = nwg - a.row0 < 0 ? 0 : (nwg - a.row0 < a.rows_out ? nwg - a.row0 : a.rows_out);
      nwg = (long long)(((unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane((int)(nwg >> 32)) << 32) |
                        (uint32_t)__builtin_amdgcn_readfirstlane((int)nwg));
      long long i = a.resume ? 0 : -a.fuse;
      for (;;) {
         const uint32_t xslot = wdraw(st, 0u) & 16383u;
         int fs = __builtin_amdgcn_readfirstlane(sel[xslot >> 6]);
         while (fs + 1 < G.nstd && xslot >= (uint32_t)__builtin_amdgcn_readfirstlane(G.starts[fs + 1])) fs++;
         int fn = fs;
         // A bad value (flam3.c:257-270): the point restarts from the
         // random replacement apply_xform left in q and the step is retried
         // with a new xform, up to four times; the fifth bad value in a row
         // is kept.  The retries stay inside the step (the lanes without a
         // bad value wait), each retry's xform the step's next wavefront
         // draw, so every lane still plots exactly one point per step and
         // chunked launches follow the single launch.
         // Bad values are counted on plotted steps only: the reference's
         // count holds its fuse steps too (15 per 10000-sample sub-batch,
         // 0.15% of its steps), but a walker fuses 100 steps, which for
         // small renders would be a large share of the count.
         bool pend = apply_xform<float, WalkerRng, TIER, true>(G, fn, p, q, rng);
         int nbad = 0;
         uint32_t kdraw = 0;
         while (__builtin_amdgcn_ballot_w64(pend)) {
            // (the retry's xform is drawn outside the branch, like the
            // step's own draw)
            const uint32_t rs = wdraw(st, ++kdraw) & 16383u;
            int f2 = __builtin_amdgcn_readfirstlane(sel[rs >> 6]);
            while (f2 + 1 < G.nstd && rs >= (uint32_t)__builtin_amdgcn_readfirstlane(G.starts[f2 + 1])) f2++;
            if (pend) {
               if (i >= 0 && i < n) bad++;
               if (++nbad < 5) {
                  p[0] = q[0]; p[1] = q[1]; p[2] = q[2]; p[3] = q[3];
                  fn = f2;
                  pend = apply_xform<float, WalkerRng, TIER, true>(G, f2, p, q, rng);
               } else
                  pend = false;
            }
         }
         p[0] = q[0]; p[1] = q[1]; p[2] = q[2]; p[3] = q[3];
         // the final xform's point is only plotted, never iterated on: the
         // fuse steps skip it (the reference applies it there and drops the
         // result, flam3.c:280-291)
         if (use_final && i >= 0) {
            if (G.final_opacity == 1.0f || u01<float>(rng) < G.final_opacity) {
               apply_xform<float, WalkerRng, TIER, true>(G, G.final_index, p, q, rng);
               q[3] = p[3];
            }
         }
         if (i >= 0 && i < n) plot(i, q, fn);
         i++;
         // the points change lanes after every FL3_PERM_EVERY-th step (the
         // same steps in every lane: st counts the workgroup's steps)
         const uint32_t stp = st++;
         if ((stp & (FL3_PERM_EVERY - 1u)) == FL3_PERM_EVERY - 1u) {
            const uint32_t pc = stp / FL3_PERM_EVERY;
            uint32_t h = pc * 0x9E3779B9u ^ hseed;
            h ^= h >> 16; h *= 0x7feb352du; h ^= h >> 15; h *= 0x846ca68bu; h ^= h >> 16;
            const unsigned slot = ((h | 1u) * (threadIdx.x ^ ((h >> 8) & 255u)) + (h >> 16)) & 255u;
            float *buf = xb + (pc & 1u) * 768;
            buf[slot] = p[0];
            buf[256 + slot] = p[1];
            buf[512 + slot] = p[2];
            // (two buffers: a wavefront writes buffer pc & 1 only after every
            // wavefront has passed the barrier behind its reads of it)
            __syncthreads();
            p[0] = buf[threadIdx.x];
            p[1] = buf[256 + threadIdx.x];
            p[2] = buf[512 + threadIdx.x];
         }
         if (i >= nwg) break;
      }
      lastxf = (int)st;   // the step count is what resumes here
   } else {
      for (long long i = a.resume ? 0 : -a.fuse; i < n;) {
         const uint32_t r = rng.next();
         const int fn = select_xform(G, G.chaos_enable ? lastxf : 0, r & 16383u);
         if (apply_xform<float, WalkerRng, TIER>(G, fn, p, q, rng)) {
            consec++;
            if (i >= 0) bad++;
            if (consec < 5) {
               p[0] = q[0]; p[1] = q[1]; p[2] = q[2]; p[3] = q[3];
               continue;
            }
            consec = 0;
         } else
            consec = 0;
         lastxf = fn + 1;
         p[0] = q[0]; p[1] = q[1]; p[2] = q[2]; p[3] = q[3];
         if (use_final && i >= 0) {   // (not in the fuse steps, as above)
            if (G.final_opacity == 1.0f || u01<float>(rng) < G.final_opacity) {
               apply_xform<float, WalkerRng, TIER, true>(G, G.final_index, p, q, rng);
               q[3] = p[3];
            }
         }
         if (i >= 0) plot(i, q, fn);
         i++;
      }
   }
   if (RECORDS == FL3_REC_WIDE || RECORDS == FL3_REC_JOB)
      for (long long i = n; i < a.rows_out; i++) a.rec[(blk * a.rows_out + i) * 256 + threadIdx.x] = FL3_REC_NONE;
   if (RECORDS == FL3_REC_NARROW)
      for (long long i = n; i < a.rows_out; i++)
         ((unsigned *)a.rec)[(blk * a.rows_out + i) * 256 + threadIdx.x] = FL3_REC_NONE;
   if (RECORDS) {
      // tile counts summed over workgroups, and this workgroup's counts per
      // super-tile (FL3_SUPER tiles) for the first sort pass
      __syncthreads();
      for (int b = threadIdx.x; b < a.nbins; b += blockDim.x) {
         const unsigned c = bh[b];
         if (c) atomicAdd(&a.bin_total[b], c);
      }
      const int nsuper = (a.nbins + (1 << FL3_SUPER_SHIFT) - 1) >> FL3_SUPER_SHIFT;
      for (int sp = threadIdx.x; sp < nsuper; sp += blockDim.x) {
         const int b0 = sp << FL3_SUPER_SHIFT, b1 = min(a.nbins, b0 + (1 << FL3_SUPER_SHIFT));
         unsigned c = 0;
         for (int b = b0; b < b1; b++) c += bh[b];
         a.blk_hist[(size_t)blk * nsuper + sp] = c;
      }
   }
   if (ws) {
      for (int k = 0; k < 4; k++) ws[k * W + wid] = __float_as_uint(p[k]);
      ws[4 * W + wid] = rng.s0;
      ws[5 * W + wid] = rng.s1;
      ws[6 * W + wid] = rng.s2;
      ws[7 * W + wid] = rng.s3;
      ws[8 * W + wid] = (uint32_t)lastxf;
      ws[9 * W + wid] = (uint32_t)consec;
   }
   add_bad(badp, bad);
}

// One FP32 application of xform xi to each of n points (x, y, colour) in the
// walkers' arithmetic: out = (x, y, colour, 1 if the result was a bad value).
// Random variations draw from a walker stream keyed by (key, point index).
template <int TIER>
__global__ __launch_bounds__(256) void k_apply_fp32(const unsigned char *job, int xi, const double *pts,
                                                    long long n, double *out, unsigned long long key) {
   unsigned char *lds = fl3_lds;
   GenL<float> G;
   load_genome<float>(job, lds, G);
   __syncthreads();
   for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
        i += (long long)gridDim.x * blockDim.x) {
      WalkerRng rng;
      rng.seed(key, ~key, (uint64_t)i);
      float p[4] = {(float)pts[4 * i], (float)pts[4 * i + 1], (float)pts[4 * i + 2], (float)pts[4 * i + 3]}, q[4];
      const bool bad = apply_xform<float, WalkerRng, TIER>(G, xi, p, q, rng);
      out[4 * i] = q[0];
      out[4 * i + 1] = q[1];
      out[4 * i + 2] = q[2];
      out[4 * i + 3] = bad ? 1.0 : 0.0;
   }
}

#endif
