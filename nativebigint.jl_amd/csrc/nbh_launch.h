// nbh_launch.h -- host launchers for the kernels in nbh_kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "nbh_internal.h"

namespace nbh {

// Threads of an LDS-tiled launch whose largest tile holds `points` points
// (one register sub-problem of 16 points per thread), 64..1024.
inline int tile_threads(int64_t points) {
    int64_t t = (points / 16 + 63) / 64 * 64;
    return (int)(t < 64 ? 64 : t > 1024 ? 1024 : t);
}

// Column groups (GroupDesc::col != 0): each thread carries one column of the
// group's S rows through all its stages in registers, no LDS.  Their
// lds_bytes entry is one of these markers (no LDS tile is this small), which
// picks the column kernel sized for S <= 16 or for 16 < S <= 36.
constexpr size_t COL_SMALL = 1;
constexpr size_t COL_BIG = 2;
constexpr int COL_THREADS = 256;  // columns per tile at most

// An LDS group launch takes the six-wave build when LDS admits more than 16
// waves per CU (four per SIMD, the unrestricted build's VGPR limit).
inline bool group_six_waves(size_t lds_bytes, int threads) {
    return (int64_t)(LDS_PER_CU / (lds_bytes > 0 ? lds_bytes : 1)) * (threads / 64) > 16;
}

// Blocks per CU of an LDS group (mid: false) or mid launch: the lesser of
// what LDS admits and what the build's waves per CU admit.
inline int64_t lds_blocks_per_cu(bool group, size_t lds_bytes, int threads) {
    const int cap = group && group_six_waves(lds_bytes, threads) ? 24 : 16;
    const int64_t by_lds = (int64_t)(LDS_PER_CU / (lds_bytes > 0 ? lds_bytes : 1));
    const int64_t by_waves = cap / ((threads + 63) / 64);
    return by_lds < by_waves ? by_lds : by_waves;
}

// Tile-driven launches: `tiles` is a device tile list of ntiles entries, or
// NULL to generate ntiles tiles of job 0 from `gen`.
hipError_t launch_group(const DevJob *jobs, const Tile *tiles, const TileGen &gen, int64_t ntiles, int fwd,
                        size_t lds_bytes, int threads, hipStream_t s);
hipError_t launch_mid(const DevJob *jobs, const Tile *tiles, const TileGen &gen, int64_t ntiles, size_t lds_bytes,
                      int threads, hipStream_t s);
// One single product in one workgroup (k_small): `threads` threads, a's and
// b's arrays of ld doubles per prime, twiddle tables of twl doubles per prime,
// nseg 4-limb result segments and cnt staged Garner words; single: every
// stage its own pass (else radix-4 pairs).
hipError_t launch_small(const DevJob &J, int threads, int ld, int twl, int nseg, int cnt, int single,
                        size_t lds_bytes, hipStream_t s);
// bmin: the smallest chunk width among the launch's jobs (sizes the staging LDS)
// lookback: one job (no tile list) of at most UNPACK_LB_MAX_CHUNKS chunks
// whose flags its mid launch cleared (TILE_CLEAR); no k_carry_fix after it.
hipError_t launch_unpack(const DevJob *jobs, const Tile *tiles, const TileGen &gen, int64_t ntiles, int bmin,
                         hipStream_t s, bool lookback = false);
hipError_t launch_carry_fix(const DevJob *jobs, const Tile *tiles, const TileGen &gen, int64_t ntiles,
                            hipStream_t s);
hipError_t launch_expand_tiles(const TileRange *ranges, int64_t nranges, Tile *tiles, hipStream_t s);
// One workgroup per cofactor-step output (sum, or |difference| and sign);
// lens: some jobs carry a length output (initialised by a launch before).
hipError_t launch_combine(const CombineJob *jobs, int64_t count, hipStream_t s, bool lens = false);
// One workgroup per product below the NTT thresholds (schoolbook).
hipError_t launch_basecase(const BaseJob *jobs, int64_t count, hipStream_t s);
hipError_t launch_basecase_one(const BaseJob &j, hipStream_t s);
// One workgroup per Barrett reduction: subtraction and the two corrections.
hipError_t launch_barrett_fix(const BarrettFixJob *jobs, int64_t count, hipStream_t s);
hipError_t launch_twiddle_stage(int prime, double *out, int m, int64_t Q, int64_t NoverL, const double *tlo,
                                const double *thi, int h, hipStream_t s);
hipError_t launch_pointwise(int prime, double *xa, const double *xb, int64_t n, hipStream_t s);
hipError_t launch_pack(double *x, int64_t N, const uint64_t *limbs, int64_t n, int b, int64_t nch, hipStream_t s);
hipError_t launch_fill_u64(uint64_t *p, int64_t n, uint64_t v, hipStream_t s);
hipError_t launch_top_limbs(const DevJob *jobs, int64_t count, uint64_t *out, hipStream_t s);
// Raise the dynamic-LDS ceiling of the LDS-tiled kernels to `bytes`.
hipError_t set_lds_limit(size_t bytes);

}  // namespace nbh
