// fl3_launch.h - host-callable launchers of the kernels in
// fl3_kernels_fast.hip and fl3_kernels_exact.hip.  Each returns non-zero if
// the launch failed.
#ifndef FL3_LAUNCH_H
#define FL3_LAUNCH_H
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include "fl3_kargs.h"

extern "C" {
int fl3k_launch_iter_fast(const IterArgs *a, unsigned grid, size_t lds, hipStream_t s, int bits, int exact,
                          int tier);
/* tier in [0, FL3_TIER_FULL): kernels restricted to that variation tier */
int fl3k_launch_iter_tier(const IterArgs *a, unsigned grid, size_t lds, hipStream_t s, int tier, int exact);
int fl3k_launch_apply_fp32(const unsigned char *job, int xi, const double *pts, long long n, double *out,
                           unsigned long long key, size_t lds, hipStream_t s, int tier);
int fl3k_launch_apply_fp32_tier(const unsigned char *job, int xi, const double *pts, long long n, double *out,
                                unsigned long long key, size_t lds, hipStream_t s, int tier);
/* format: 0 8-byte records, 1 4-byte records, 2 8-byte job records;
 * FL3_BIN_F64_WIDE / FL3_BIN_F64_NARROW: 8- / 4-byte records into double buckets */
int fl3k_launch_bin(const BinArgs *a, int ncu, int format, hipStream_t s);
/* banded histograms: the first level (scans and the band pass, writing
 * 4-byte band records when narrow_out), the second level's counting pass */
int fl3k_launch_bin_split(const BinArgs *a, int narrow_out, hipStream_t s);
int fl3k_launch_bin_count(const BinArgs *a, int narrow, hipStream_t s);
size_t fl3k_bin_f64_lds_bytes(int shift);
int fl3k_bin_pass_records(void);   // records one accumulation workgroup loads per pass
int fl3k_launch_copy_probe(const void *src, void *dst, long long n16, hipStream_t s);
size_t fl3k_bin_lds_bytes(int shift);
int fl3k_launch_replay(const IterArgs *a, unsigned grid, size_t lds, hipStream_t s, int bits);
int fl3k_launch_points(const PointsArgs *a, unsigned grid, size_t lds, hipStream_t s);
int fl3k_launch_preview(const unsigned char *job, int xi, double range, int numvals, int depth, double *result,
                        uint32_t *isaac, size_t lds, hipStream_t s);
int fl3k_launch_deposit(const unsigned char *job, const double *pts, long long n, void *buckets, uint32_t *sat,
                        unsigned grid, size_t lds, hipStream_t s, int bits);
int fl3k_launch_de_tiled(const DEArgs *a, hipStream_t s, int bits);
int fl3k_launch_de_many(const DEArgs *as, const unsigned *prep_first, unsigned prep_blocks,
                        const unsigned *gather_first, unsigned gather_blocks, int n, hipStream_t s,
                        int *blkmax, long nblk, size_t gather_lds);
int fl3k_launch_f64_to_f32(const double *src, float *dst, long n, hipStream_t s);
int fl3k_launch_logscale(const void *b, void *acc, long n, double k1, double k2, hipStream_t s, int bits);
int fl3k_launch_earlyclip(void *acc, long n, const TMArgs *a, hipStream_t s, int bits);
int fl3k_launch_tonemap(const TMArgs *a, hipStream_t s, int bits);
int fl3k_launch_tonemap_many(const TMArgs *as, const unsigned *first, unsigned blocks, int n, hipStream_t s,
                             int bits);
}
#endif
