// BM3D kernels for CDNA4 (gfx950): block matching, group filter with
// weighted fixed-point aggregation, and finish.
//
// Main path (the shapes whose working sets fit the LDS budget):
//   match_run_kernel<B, SMALL> one workgroup per run of 8 references on one
//   match_run_mid_kernel<B>    grid row; the union of their search windows is
//                              staged in LDS once; each half-wave owns one
//                              reference and each lane 2 x 7 tiles of
//                              candidates (one tile: SMALL; up to three: mid,
//                              keys held in registers); then the G-1 smallest
//                              keys (dist, y, x) by a threshold and a rank
//                              count (rounds of a half-wave min as fallback).
//   filter_tile_kernel<B, W>   one workgroup per tile of references; each wave
//                              filters whole groups (2-D DCT, Haar along the
//                              group, hard threshold or Wiener, inverse) and
//                              adds them into the tile's int64 fixed-point
//                              numerator sums in LDS (ds_add_u64: exact and
//                              order-free); the tile writes those sums as its
//                              partials and one weight per group.
//   finish_tiles_kernel        per pixel, the numerator partials of the tiles
//                              covering it, summed exactly; the weights rebuilt
//                              from the match records; then num / den or the
//                              source.
// General path (any legal shape):
//   match_wave_kernel<B, S>    one wave per reference, window in LDS (or read
//                              from global when it does not fit).
//   filter_atomic_kernel<B, W> one wave per group, device-scope int64 atomics
//                              into frame-sized sums (den at the corner).
//   finish_corner_kernel<B>    box-sums the corner weights, divides.
// Workgroups are dealt round-robin over the 8 XCDs; every kernel remaps its
// block index so that each XCD works on a contiguous part of the frame (its
// L2 then holds the overlapping windows, patches and partial rows).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

#include "bm3d_jobs.hpp"

namespace nsship {

// True when an armed test hook refuses this launch (device_mem.cpp); the
// launch wrappers then return hipErrorLaunchFailure without launching.
bool launch_fault();

// Orthonormal DCT-II matrices (row = frequency) for B in {2,4,8,12,16,32},
// concatenated; filled from the host (double cosines rounded to float, as
// src/cpu/bm/dct8.cpp:71-89 DctTable).
__constant__ float c_dct[4 + 16 + 64 + 144 + 256 + 1024];

template <int B>
__device__ constexpr int dct_offset() {
    return B == 2 ? 0 : B == 4 ? 4 : B == 8 ? 20 : B == 12 ? 84 : B == 16 ? 228 : 484;
}

inline constexpr uint32_t kTaken = 0xffffffffu;
inline constexpr uint32_t kNonFinite = 0x7f800000u;
inline constexpr int kXcds = 8;

// Block index such that each XCD (blockIdx % 8 on dispatch) covers a
// contiguous range of logical blocks; a bijection on [0, gridDim.x).
__device__ inline int xcd_block() {
    const int nb = static_cast<int>(gridDim.x), b = static_cast<int>(blockIdx.x);
    const int q = nb / kXcds, rem = nb % kXcds;
    const int xcd = b % kXcds, slot = b / kXcds;
    return xcd < rem ? xcd * (q + 1) + slot : rem * (q + 1) + (xcd - rem) * q + slot;
}

// Orders the calling wave's LDS accesses (one wave's DS operations execute
// in issue order; this keeps the compiler from moving accesses across).
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// min(v, v of the lane DPP control `ctrl` selects); lanes whose source is
// outside the row (or rows outside row_mask) keep v.
template <int CTRL, int ROW_MASK>
__device__ inline unsigned long long dpp_min_u64(unsigned long long v) {
    const int lo = static_cast<int>(static_cast<uint32_t>(v));
    const int hi = static_cast<int>(static_cast<uint32_t>(v >> 32));
    const int olo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROW_MASK, 0xf, false);
    const int ohi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROW_MASK, 0xf, false);
    const unsigned long long o = (static_cast<unsigned long long>(static_cast<uint32_t>(ohi)) << 32) |
                                 static_cast<uint32_t>(olo);
    return o < v ? o : v;
}

// Minimum over each half-wave (lanes 0-31 and 32-63), returned to every lane
// of that half. Must be reached by all 64 lanes. DPP row_shr 1, 2, 4, 8 leaves
// each 16-lane row's minimum in its last lane; row_bcast15 folds row 0 into
// row 1 and row 2 into row 3; lanes 31 and 63 then hold the two minima.
__device__ inline unsigned long long half_min_u64(unsigned long long v) {
    v = dpp_min_u64<0x111, 0xf>(v);
    v = dpp_min_u64<0x112, 0xf>(v);
    v = dpp_min_u64<0x114, 0xf>(v);
    v = dpp_min_u64<0x118, 0xf>(v);
    v = dpp_min_u64<0x142, 0xa>(v);
    const int lo = static_cast<int>(static_cast<uint32_t>(v));
    const int hi = static_cast<int>(static_cast<uint32_t>(v >> 32));
    const uint32_t lo0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(lo, 31));
    const uint32_t hi0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(hi, 31));
    const uint32_t lo1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(lo, 63));
    const uint32_t hi1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(hi, 63));
    const bool upper = threadIdx.x & 32;
    return (static_cast<unsigned long long>(upper ? hi1 : hi0) << 32) | (upper ? lo1 : lo0);
}

// min(v, v of the lane DPP control `ctrl` selects); lanes whose source is
// outside the row (or rows outside row_mask) see the identity ~0, so the
// move folds into the min as one DPP instruction.
template <int CTRL, int ROW_MASK>
__device__ inline uint32_t dpp_min_u32(uint32_t v) {
    const int o = __builtin_amdgcn_update_dpp(-1, static_cast<int>(v), CTRL, ROW_MASK, 0xf, false);
    return min(v, static_cast<uint32_t>(o));
}

// Half-wave minimum of 32-bit values (the DPP steps of half_min_u64).
__device__ inline uint32_t half_min_u32(uint32_t v) {
    v = dpp_min_u32<0x111, 0xf>(v);
    v = dpp_min_u32<0x112, 0xf>(v);
    v = dpp_min_u32<0x114, 0xf>(v);
    v = dpp_min_u32<0x118, 0xf>(v);
    v = dpp_min_u32<0x142, 0xa>(v);
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 31));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
    return (threadIdx.x & 32) ? hi : lo;
}

// Half-wave minimum of 64-bit keys (dist << 32 | index) in two 32-bit
// passes: the smallest dist, then the smallest index among the lanes holding
// it. Equal to half_min_u64's result; all 64 lanes must call.
__device__ inline unsigned long long half_min_key(unsigned long long v) {
    const uint32_t d = static_cast<uint32_t>(v >> 32);
    const uint32_t md = half_min_u32(d);
    const uint32_t mc = half_min_u32(d == md ? static_cast<uint32_t>(v) : 0xffffffffu);
    return (static_cast<unsigned long long>(md) << 32) | mc;
}

// Minimum over the whole wave, returned to every lane (same DPP steps as
// half_min_u64, then the smaller of lanes 31 and 63). All 64 lanes must call.
__device__ inline unsigned long long wave_min_u64(unsigned long long v) {
    v = dpp_min_u64<0x111, 0xf>(v);
    v = dpp_min_u64<0x112, 0xf>(v);
    v = dpp_min_u64<0x114, 0xf>(v);
    v = dpp_min_u64<0x118, 0xf>(v);
    v = dpp_min_u64<0x142, 0xa>(v);
    const int lo = static_cast<int>(static_cast<uint32_t>(v));
    const int hi = static_cast<int>(static_cast<uint32_t>(v >> 32));
    const unsigned long long a =
        (static_cast<unsigned long long>(static_cast<uint32_t>(__builtin_amdgcn_readlane(hi, 31))) << 32) |
        static_cast<uint32_t>(__builtin_amdgcn_readlane(lo, 31));
    const unsigned long long b =
        (static_cast<unsigned long long>(static_cast<uint32_t>(__builtin_amdgcn_readlane(hi, 63))) << 32) |
        static_cast<uint32_t>(__builtin_amdgcn_readlane(lo, 63));
    return a < b ? a : b;
}

__device__ inline int wave_sum_i32(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ inline float wave_sum_f32(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <typename Table>
__device__ inline int find_job(const Table& tab, int block) {
    int j = 0;
    while (j + 1 < tab.n && block >= tab.job[j + 1].wg_begin) ++j;
    return j;
}

__device__ inline uint32_t dist_key(float d) {
    const uint32_t k = __float_as_uint(d);
    return (k & kNonFinite) == kNonFinite ? kNonFinite : k;
}

// floor(c / d) for 0 <= c < 2^20 and 1 <= d <= 256, from inv = 1.f / d:
// (c + 1/2) / d lies at least 1/(2d) away from an integer, far more than the
// float rounding of the product.
__device__ inline int div_small(int c, float inv) {
    return static_cast<int>((static_cast<float>(c) + 0.5f) * inv);
}

__device__ inline int ref_x(const Grid& g, int i, int width, int block) { return min(i * g.step, width - block); }
__device__ inline int ref_y(const Grid& g, int j, int height, int block) {
    return min((j + g.row0) * g.step, height - block);
}

// ---------------------------------------------------------------------------
// Block matching (src/cpu/bm/matcher.hpp:247-283 collect_spatial_coords:
// window [c - r, c + r] clamped to the plane, raster scan, self excluded,
// reference first; order src/cpu/bm/matcher.hpp:27-57 match_less).
// ---------------------------------------------------------------------------

// The value of lane (lane ^ J) within each group of 32 lanes: DPP quad
// permutes for J = 1, 2, a bit-mode ds_swizzle for J = 4, 8, 16.
template <int J>
__device__ inline uint32_t xor_lane(uint32_t v) {
    if constexpr (J == 1) {
        return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0xb1, 0xf, 0xf, false));
    } else if constexpr (J == 2) {
        return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x4e, 0xf, 0xf, false));
    } else {
        return static_cast<uint32_t>(__builtin_amdgcn_ds_swizzle(static_cast<int>(v), 0x1f | (J << 10)));
    }
}

// One compare-exchange stage (block K, distance J) of a bitonic sort that
// orders each 32-lane half ascending: a lane keeps the smaller value where
// its J bit equals its K bit (K = 32: the whole half ascends).
// One compare-exchange stage (block K, distance J) of a bitonic sort that
// orders each 32-lane half ascending: a lane keeps the smaller value where
// its J bit equals its K bit (K = 32: the whole half ascends).
template <int K, int J>
__device__ inline unsigned long long bitonic_u64(unsigned long long v) {
    const int lane = threadIdx.x & 31;
    const uint32_t plo = xor_lane<J>(static_cast<uint32_t>(v));
    const uint32_t phi = xor_lane<J>(static_cast<uint32_t>(v >> 32));
    const unsigned long long p = (static_cast<unsigned long long>(phi) << 32) | plo;
    const bool keep_min = ((lane & J) == 0) == (K == 32 || (lane & K) == 0);
    return keep_min == (p < v) ? p : v;
}

// Ascending sort of the 32 keys of each half-wave (lane i of the half holds
// the i-th smallest). All 64 lanes must call.
__device__ inline unsigned long long sort32_u64(unsigned long long v) {
    v = bitonic_u64<2, 1>(v);
    v = bitonic_u64<4, 2>(v);
    v = bitonic_u64<4, 1>(v);
    v = bitonic_u64<8, 4>(v);
    v = bitonic_u64<8, 2>(v);
    v = bitonic_u64<8, 1>(v);
    v = bitonic_u64<16, 8>(v);
    v = bitonic_u64<16, 4>(v);
    v = bitonic_u64<16, 2>(v);
    v = bitonic_u64<16, 1>(v);
    v = bitonic_u64<32, 16>(v);
    v = bitonic_u64<32, 8>(v);
    v = bitonic_u64<32, 4>(v);
    v = bitonic_u64<32, 2>(v);
    return bitonic_u64<32, 1>(v);
}

// Lane tiles (2 x 7 candidates) a half-wave needs per lane for a whole
// window of the range.
__host__ __device__ constexpr int run_tiles_per_lane(int range) {
    return (((2 * range + 1 + 6) / 7) * ((2 * range + 1 + 1) / 2) + 31) / 32;
}

// Windows whose lane tiles fit one per lane of a half-wave (a side of up to
// 20: ranges up to 9) keep each lane's keys in registers through a workgroup
// barrier; the window keys and the gathered keys then overlay the staged
// region in LDS. Up to kRunMidTiles tiles per lane (ranges up to 16) do the
// same with the keys of every tile in registers and a smaller gather area
// (kRunMidGather keys per reference), so that a 33 x 33 window's workgroup
// fits four to a CU. Larger windows keep both areas.
__host__ __device__ constexpr bool run_keys_in_regs(int range) { return run_tiles_per_lane(range) <= 1; }
inline constexpr int kRunMidTiles = 3;
inline constexpr int kRunMidGather = 64;
__host__ __device__ constexpr bool run_keys_mid(int range) {
    return run_tiles_per_lane(range) > 1 && run_tiles_per_lane(range) <= kRunMidTiles;
}

// Floats of the staged region the run matcher's LDS layout reserves: the 8
// windows' span + 6 padding columns with the pitch rounded up to 2 mod 4, by
// the windows' height + 1 row (every workgroup's region fits).
__host__ __device__ constexpr int run_region_floats(int block, int step, int range) {
    return (((kRunRefs - 1) * step + 2 * range + block + 9) * (2 * range + block + 1) + 3) & ~3;
}

// SSD of the 2 x 7 candidates of lane tile (dy0, dx0) against the reference
// patch at rp, in block-row order (one fmaf per pixel, the order every
// matcher uses): region rows dy0 + py + i, columns dx0 .. dx0 + B + 5.
// REUSE (large windows, several tiles per lane): row dy0 + py + 1 serves
// block rows py and py + 1 and is read once, and the 14 differences of a
// reference pixel are formed before their products; it costs registers, so
// the small-window form (one tile per lane, 8 waves per SIMD) reads each
// row pair per block row instead.
template <int B, bool REUSE>
__device__ inline void run_tile_ssd(const float* rp, const float* base, int RW, float (&acc)[2][7]) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 7; ++j) acc[i][j] = 0.f;
    if constexpr (!REUSE) {
        for (int py = 0; py < B; ++py) {
            float rrow[B];
#pragma unroll
            for (int px = 0; px < B; ++px) rrow[px] = rp[py * RW + px];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const float* rowp = base + (py + i) * RW;
                float v[B + 6];
#pragma unroll
                for (int c = 0; c < B + 6; ++c) v[c] = rowp[c];
#pragma unroll
                for (int j = 0; j < 7; ++j)
#pragma unroll
                    for (int px = 0; px < B; ++px) {
                        const float d = rrow[px] - v[j + px];
                        acc[i][j] = fmaf(d, d, acc[i][j]);
                    }
            }
        }
    } else {
        float cur[B + 6];
#pragma unroll
        for (int c = 0; c < B + 6; ++c) cur[c] = base[c];
#pragma unroll 2
        for (int py = 0; py < B; ++py) {
            float rrow[B], nxt[B + 6];
#pragma unroll
            for (int px = 0; px < B; ++px) rrow[px] = rp[py * RW + px];
#pragma unroll
            for (int c = 0; c < B + 6; ++c) nxt[c] = base[(py + 1) * RW + c];
#pragma unroll
            for (int px = 0; px < B; ++px) {
                float d0[7], d1[7];
#pragma unroll
                for (int j = 0; j < 7; ++j) {
                    d0[j] = rrow[px] - cur[j + px];
                    d1[j] = rrow[px] - nxt[j + px];
                }
#pragma unroll
                for (int j = 0; j < 7; ++j) {
                    acc[0][j] = fmaf(d0[j], d0[j], acc[0][j]);
                    acc[1][j] = fmaf(d1[j], d1[j], acc[1][j]);
                }
            }
#pragma unroll
            for (int c = 0; c < B + 6; ++c) cur[c] = nxt[c];
        }
    }
}

// Run matcher: references i0 .. i0+7 of grid row j share one LDS region.
// NT = 1: every job's window takes one lane tile per lane
// (run_keys_in_regs); NT = kRunMidTiles: up to that many (run_keys_mid),
// keys in registers; NT = 0: keys in LDS as they are scored.
template <int B, int NT>
__device__ __forceinline__ void match_run_body(const MatchTable& tab) {
    constexpr bool SMALL = NT == 1;
    constexpr int NG = NT == kRunMidTiles ? kRunMidGather : kRunGather;  // gathered keys per reference
    extern __shared__ __align__(16) float lds[];
    const int lb = xcd_block();
    const MatchJob& jb = tab.job[find_job(tab, lb)];
    const int run = lb - jb.wg_begin;
    const int rpr = (jb.grid.nx + kRunRefs - 1) / kRunRefs;
    const int row = run / rpr, i0 = (run - row * rpr) * kRunRefs;
    const int nrun = min(kRunRefs, jb.grid.nx - i0);
    const int r = jb.range;
    const int cy = ref_y(jb.grid, row, jb.height, B);
    const int top = max(cy - r, 0), bottom = min(cy + r, jb.height - B);
    const int rx0 = max(ref_x(jb.grid, i0, jb.width, B) - r, 0);
    const int rx1 = min(ref_x(jb.grid, i0 + nrun - 1, jb.width, B) + r, jb.width - B) + B - 1;
    // Lane tiles of 2 x 7 candidates (2 rows, 7 columns): a 19 x 19 window
    // is 10 x 3 tiles, 448 lane slots for 361 candidates. Tile columns 7 tx
    // spread the lanes over the banks; the region keeps 6 columns after the
    // windows and 1 row below (candidates of an overhanging tile are
    // discarded), with a row pitch = 2 mod 4.
    const int need = rx1 - rx0 + 1 + 6;
    const int RW = need + (6 - need % 4) % 4;
    const int RH = bottom - top + B + 1;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // Rows wave, wave + 4, ... by 64-column strips: the row address is
    // wave-uniform and each lane's column is clamped once per strip; cells
    // outside the plane repeat its edge (only discarded candidates of
    // overhanging tiles read them).
    // Each load goes straight to LDS (global_load_lds: lane l's word lands
    // at the wave-uniform row address + 4 l), so the staging holds no
    // registers and issues no LDS stores; the barrier below waits for them.
    // The row (and so its address) is wave-uniform: scalar arithmetic, the
    // lane's clamped column the only per-lane offset.
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    for (int c0 = 0; c0 < RW; c0 += 64) {
        const int xx = c0 + lane;
        if (xx < RW) {
            const int gx = min(rx0 + xx, jb.width - 1);
            for (int yy = wv; yy < RH; yy += 4) {
                const float* row = jb.plane + static_cast<size_t>(min(top + yy, jb.height - 1)) * jb.stride;
                __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(row + gx),
                                                 (__attribute__((address_space(3))) void*)(lds + yy * RW + c0), 4, 0,
                                                 0);
            }
        }
    }
    const int wmax = 2 * r + 1;
    constexpr bool regs = NT > 0;
    // Window keys (wmax^2 per reference) and up to NG gathered keys
    // per reference. Keys waiting in registers: both over the region. Else
    // the window keys after the region and the gathered keys over it (or
    // after the window keys when the region is smaller).
    // (The offsets follow the reserved region size, the same in every
    // workgroup of the launch.)
    const int region = run_region_floats(B, jb.grid.step, r);
    uint32_t* dist = reinterpret_cast<uint32_t*>(lds + (regs ? 0 : region));
    unsigned long long* surv =
        regs || region * 4 < kRunRefs * NG * 8
            ? reinterpret_cast<unsigned long long*>(dist + kRunRefs * wmax * wmax)
            : reinterpret_cast<unsigned long long*>(lds);
    __syncthreads();

    const int half = threadIdx.x >> 5, hl = threadIdx.x & 31;
    const int cx = ref_x(jb.grid, i0 + half, jb.width, B);
    const int left = max(cx - r, 0), right = min(cx + r, jb.width - B);
    const int ww = right - left + 1, wh = bottom - top + 1;
    const int ncand = ww * wh;
    uint32_t* dk = dist + half * wmax * wmax;
    const int tw = (ww + 6) / 7, th = (wh + 1) >> 1;
    const int self = (cy - top) * ww + (cx - left);
    const float* rp = lds + (cy - top) * RW + (cx - rx0);
    const float* wbase = lds + (left - rx0);
    if (NT > 1) {
        // up to NT tiles per lane (t = hl, hl + 32, ...): their keys wait in
        // registers for the barrier, then go to the key area over the region
        uint32_t kr[NT > 1 ? NT : 1][2][7];
        const int nt = half < nrun ? tw * th : 0;
#pragma unroll
        for (int q = 0; q < NT; ++q) {
            const int t = hl + 32 * q;
            if (t < nt) {
                const int ty = t / tw, tx = t - ty * tw;
                float acc[2][7];
                run_tile_ssd<B, true>(rp, wbase + ty * 2 * RW + tx * 7, RW, acc);
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 7; ++j) kr[q][i][j] = dist_key(acc[i][j]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NT; ++q) {
            const int t = hl + 32 * q;
            if (t < nt) {
                const int ty = t / tw, tx = t - ty * tw;
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 7; ++j) {
                        const int dy = ty * 2 + i, dx = tx * 7 + j, c = dy * ww + dx;
                        if (dy < wh && dx < ww) dk[c] = c == self ? kTaken : kr[q][i][j];
                    }
            }
        }
    } else if (regs) {
        // one tile per lane: its keys wait in registers for the barrier,
        // then go to the key area over the region
        const int ty = hl / tw, tx = hl - ty * tw;
        const int dy0 = ty * 2, dx0 = tx * 7;
        const bool mine = half < nrun && hl < tw * th;
        float acc[2][7];
        if (mine) run_tile_ssd<B, !SMALL>(rp, wbase + dy0 * RW + dx0, RW, acc);
        __syncthreads();
        if (mine) {
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 7; ++j) {
                    const int dy = dy0 + i, dx = dx0 + j, c = dy * ww + dx;
                    if (dy < wh && dx < ww) dk[c] = c == self ? kTaken : dist_key(acc[i][j]);
                }
        }
    } else {
        if (half < nrun) {
            for (int t = hl; t < tw * th; t += 32) {
                const int ty = t / tw, tx = t - ty * tw;
                const int dy0 = ty * 2, dx0 = tx * 7;
                float acc[2][7];
                run_tile_ssd<B, !SMALL>(rp, wbase + dy0 * RW + dx0, RW, acc);
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 7; ++j) {
                        const int dy = dy0 + i, dx = dx0 + j, c = dy * ww + dx;
                        if (dy < wh && dx < ww) dk[c] = c == self ? kTaken : dist_key(acc[i][j]);
                    }
            }
        }
        __syncthreads();  // the region is read; gathered keys may overlay it
    }
    wave_sync();
    // Selection, wave-uniform: both half-waves take part (a half with fewer
    // candidates, or no reference, idles) so that cross-lane steps see all
    // 64 lanes. Keys order as (dist, window index); lane hl owns the keys
    // c = hl mod 32.
    const int K = half < nrun ? min(jb.group - 1, ncand - 1) : 0;
    const int kmax = max(__builtin_amdgcn_readlane(K, 0), __builtin_amdgcn_readlane(K, 32));
    const float inv_ww = 1.f / static_cast<float>(ww);
    const int ref_id = row * jb.grid.nx + i0 + half;
    MatchRec* out = jb.out + static_cast<size_t>(ref_id) * jb.group;
    if (half < nrun && hl == 0) {
        out[0] = MatchRec{cx, cy, jb.t0, 0u};
        jb.counts[ref_id] = 1 + K;
    }
    // Threshold selection (K <= 31). T = the K-th smallest of the half's 32
    // lane minima: distinct candidates, so at least K keys are <= T and the
    // K winners all are. The keys <= T (self excluded) are gathered in LDS
    // and each is ranked against the others; rank k is record 1 + k. When
    // more than NG keys are <= T in either half the wave takes the
    // rounds below instead.
    if (kmax <= 31) {
        const int nc = K > 0 ? ncand : 0;
        const int nc_max = max(__builtin_amdgcn_readlane(nc, 0), __builtin_amdgcn_readlane(nc, 32));
        auto key_at = [&](int c) -> unsigned long long {
            const uint32_t d = c < nc ? dk[c] : kTaken;
            return d == kTaken ? ~0ull : (static_cast<unsigned long long>(d) << 32) | static_cast<uint32_t>(c);
        };
        unsigned long long lmin = ~0ull;
        for (int c0 = 0; c0 < nc_max; c0 += 128) {
            unsigned long long k[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) k[u] = key_at(c0 + 32 * u + hl);
#pragma unroll
            for (int u = 0; u < 4; ++u) lmin = k[u] < lmin ? k[u] : lmin;
        }
        const unsigned long long sorted = sort32_u64(lmin);
        const int k0 = __builtin_amdgcn_readlane(K, 0), k1 = __builtin_amdgcn_readlane(K, 32);
        const int s0 = max(k0 - 1, 0), s1 = 32 + max(k1 - 1, 0);
        const unsigned long long t0 =
            (static_cast<unsigned long long>(static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(sorted >> 32), s0))) << 32) |
            static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(sorted), s0));
        const unsigned long long t1 =
            (static_cast<unsigned long long>(static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(sorted >> 32), s1))) << 32) |
            static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(sorted), s1));
        // ~0 (self, no candidate) is the largest key: clamping T below it
        // leaves it out of the gather
        const unsigned long long T = min((threadIdx.x & 32) ? t1 : t0, ~0ull - 1);
        const uint32_t below = (1u << hl) - 1u;  // lanes of this half below this one
        unsigned long long* sv = surv + half * NG;
        int n = 0;  // keys gathered so far by this half
        for (int c0 = 0; c0 < nc_max; c0 += 128) {
            unsigned long long k[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) k[u] = key_at(c0 + 32 * u + hl);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const bool take = k[u] <= T;
                const unsigned long long m = __ballot(take);
                if (m) {
                    const uint32_t mh = (threadIdx.x & 32) ? static_cast<uint32_t>(m >> 32) : static_cast<uint32_t>(m);
                    const int slot = n + __popc(mh & below);
                    if (take && slot < NG) sv[slot] = k[u];
                    n += __popc(mh);
                }
            }
        }
        const int n0 = __builtin_amdgcn_readlane(n, 0), n1 = __builtin_amdgcn_readlane(n, 32);
        if (n0 <= NG && n1 <= NG) {
            wave_sync();
            // lane hl ranks the gathered keys hl, hl + 32, ... of its half
            const int nmax = max(n0, n1), qn = (nmax + 31) >> 5;
            unsigned long long own[NG / 32];
            int rank[NG / 32];
#pragma unroll
            for (int q = 0; q < NG / 32; ++q) {
                own[q] = q < qn && hl + 32 * q < n ? sv[hl + 32 * q] : ~0ull;
                rank[q] = 0;
            }
            for (int j0 = 0; j0 < nmax; j0 += 8) {
                unsigned long long o[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) o[u] = j0 + u < n ? sv[j0 + u] : ~0ull;
#pragma unroll
                for (int q = 0; q < NG / 32; ++q)
                    if (q < qn) {
#pragma unroll
                        for (int u = 0; u < 8; ++u) rank[q] += o[u] < own[q] ? 1 : 0;
                    }
            }
#pragma unroll
            for (int q = 0; q < NG / 32; ++q)
                if (hl + 32 * q < n && rank[q] < K) {
                    const int c = static_cast<int>(static_cast<uint32_t>(own[q]));
                    const int dy = div_small(c, inv_ww), dx = c - dy * ww;
                    out[1 + rank[q]] = MatchRec{left + dx, top + dy, jb.t0, static_cast<uint32_t>(own[q] >> 32)};
                }
            return;
        }
    }
    // Each lane keeps the four smallest of its keys (window indices c = hl
    // mod 32) sorted in registers; a round's winner is the half-wave minimum
    // of the lanes' heads, and only the winning lane advances (rebuilding its
    // four from LDS, where given-away keys are marked taken, after its fourth).
    unsigned long long best4[4];
    auto insert = [&](uint32_t d, int c) {
        const unsigned long long x = (static_cast<unsigned long long>(d) << 32) | static_cast<uint32_t>(c);
        if (d != kTaken && x < best4[3]) {
            best4[3] = x;
            if (best4[3] < best4[2]) { const unsigned long long t = best4[2]; best4[2] = best4[3]; best4[3] = t; }
            if (best4[2] < best4[1]) { const unsigned long long t = best4[1]; best4[1] = best4[2]; best4[2] = t; }
            if (best4[1] < best4[0]) { const unsigned long long t = best4[0]; best4[0] = best4[1]; best4[1] = t; }
        }
    };
    auto rebuild = [&]() {
        best4[0] = best4[1] = best4[2] = best4[3] = ~0ull;
        // four keys read before any is inserted (their LDS loads overlap)
        for (int c0 = hl; c0 < ncand; c0 += 128) {
            uint32_t d[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) d[u] = c0 + 32 * u < ncand ? dk[c0 + 32 * u] : kTaken;
#pragma unroll
            for (int u = 0; u < 4; ++u) insert(d[u], c0 + 32 * u);
        }
    };
    if (K > 0) {
        rebuild();
    } else {
        best4[0] = best4[1] = best4[2] = best4[3] = ~0ull;
    }
    // Round k's winner is kept by lane k of the half (k < 32) and the
    // records are written together after the rounds.
    int given = 0;
    unsigned long long won = ~0ull;
    for (int k = 0; k < kmax; ++k) {
        const bool live = k < K;
        const unsigned long long head = live ? best4[0] : ~0ull;
        const uint32_t hd = static_cast<uint32_t>(head >> 32);
        const uint32_t md = half_min_u32(hd);
        // The lanes holding the smallest distance: one per half (the usual
        // case) gives the winner directly; a tie between lanes is broken by
        // the smallest window index, as match_tie_less does.
        const unsigned long long at_min = __ballot(live && hd == md);
        const uint32_t lo_mask = static_cast<uint32_t>(at_min), hi_mask = static_cast<uint32_t>(at_min >> 32);
        uint32_t mc;
        if (__popc(lo_mask) <= 1 && __popc(hi_mask) <= 1) {
            const int hc = static_cast<int>(static_cast<uint32_t>(head));
            const uint32_t c0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(hc, lo_mask ? __builtin_ctz(lo_mask) : 0));
            const uint32_t c1 =
                static_cast<uint32_t>(__builtin_amdgcn_readlane(hc, hi_mask ? 32 + __builtin_ctz(hi_mask) : 32));
            mc = (threadIdx.x & 32) ? c1 : c0;
        } else {
            mc = half_min_u32(hd == md ? static_cast<uint32_t>(head) : 0xffffffffu);
        }
        const unsigned long long best = (static_cast<unsigned long long>(md) << 32) | mc;
        const int c = static_cast<int>(mc);
        if (hl == k) won = best;
        if (live && k >= 32 && hl == 0) {
            const int dy = div_small(c, inv_ww), dx = c - dy * ww;
            out[1 + k] = MatchRec{left + dx, top + dy, jb.t0, md};
        }
        if (live && hl == (c & 31)) {
            dk[c] = kTaken;
            best4[0] = best4[1];
            best4[1] = best4[2];
            best4[2] = best4[3];
            best4[3] = ~0ull;
            if (++given == 4) {
                given = 0;
                rebuild();
            }
        }
    }
    if (hl < K && hl < 32) {
        const int c = static_cast<int>(static_cast<uint32_t>(won));
        const int dy = div_small(c, inv_ww), dx = c - dy * ww;
        out[1 + hl] = MatchRec{left + dx, top + dy, jb.t0, static_cast<uint32_t>(won >> 32)};
    }
}

// SMALL: the one-tile-per-lane layout; else keys in LDS as they are scored.
template <int B, bool SMALL>
__global__ __launch_bounds__(256) void match_run_kernel(MatchTable tab) {
    match_run_body<B, SMALL ? 1 : 0>(tab);
}

// Up to kRunMidTiles lane tiles per lane, their keys in registers.
template <int B>
__global__ __launch_bounds__(256) void match_run_mid_kernel(MatchTable tab) {
    match_run_body<B, kRunMidTiles>(tab);
}

// General matcher: one wave per reference.
template <int B, bool STAGE>
__global__ __launch_bounds__(64) void match_wave_kernel(MatchTable tab) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    const int lb = xcd_block();
    const MatchJob& jb = tab.job[find_job(tab, lb)];
    const int ref = lb - jb.wg_begin;
    if (ref >= jb.grid.nx * jb.grid.ny) return;
    const int cx = ref_x(jb.grid, ref % jb.grid.nx, jb.width, B);
    const int cy = ref_y(jb.grid, ref / jb.grid.nx, jb.height, B);
    const int r = jb.range;
    const int top = max(cy - r, 0), bottom = min(cy + r, jb.height - B);
    const int left = max(cx - r, 0), right = min(cx + r, jb.width - B);
    const int ww = right - left + 1, wh = bottom - top + 1;
    const int ncand = ww * wh;
    const int rw = ww + B - 1, rh = wh + B - 1;

    const float* reg;
    int rs;
    uint32_t* dk;
    if constexpr (STAGE) {
        // straight to LDS (global_load_lds: lane l's word at the uniform
        // address lreg + yy rw + x0 + 4 l), every row's loads in flight
        // together; the barrier below waits for them
        float* lreg = reinterpret_cast<float*>(smem);
        for (int yy = 0; yy < rh; ++yy) {
            const float* row = jb.plane + static_cast<size_t>(top + yy) * jb.stride + left;
            for (int x0 = 0; x0 < rw; x0 += 64)
                if (x0 + lane < rw)
                    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(row + x0 + lane),
                                                     (__attribute__((address_space(3))) void*)(lreg + yy * rw + x0), 4,
                                                     0, 0);
        }
        reg = lreg;
        rs = rw;
        dk = reinterpret_cast<uint32_t*>(smem) + ((rw * rh + 3) & ~3);
    } else {
        reg = jb.plane + static_cast<size_t>(top) * jb.stride + left;
        rs = jb.stride;
        dk = reinterpret_cast<uint32_t*>(smem);
    }
    __syncthreads();

    const float* refp = reg + (cy - top) * rs + (cx - left);
    const int self = (cy - top) * ww + (cx - left);
    for (int c = lane; c < ncand; c += 64) {
        const int dy = c / ww, dx = c - dy * ww;
        const float* cp = reg + dy * rs + dx;
        float acc = 0.f;
        for (int py = 0; py < B; ++py) {
#pragma unroll
            for (int px = 0; px < B; ++px) {
                const float d = refp[py * rs + px] - cp[py * rs + px];
                acc = fmaf(d, d, acc);
            }
        }
        dk[c] = c == self ? kTaken : dist_key(acc);
    }
    __syncthreads();

    const int K = min(jb.group - 1, ncand - 1);
    MatchRec* out = jb.out + static_cast<size_t>(ref) * jb.group;
    if (lane == 0) out[0] = MatchRec{cx, cy, jb.t0, 0u};
    for (int i = 0; i < K; ++i) {
        unsigned long long best = ~0ull;
        for (int c = lane; c < ncand; c += 64) {
            const uint32_t d = dk[c];
            const unsigned long long key = (static_cast<unsigned long long>(d) << 32) | static_cast<uint32_t>(c);
            best = (d != kTaken && key < best) ? key : best;
        }
        best = wave_min_u64(best);
        const int c = static_cast<int>(static_cast<uint32_t>(best));
        if (lane == (c & 63)) {
            dk[c] = kTaken;
            const int dy = c / ww, dx = c - dy * ww;
            out[1 + i] = MatchRec{left + dx, top + dy, jb.t0, static_cast<uint32_t>(best >> 32)};
        }
        __syncthreads();
    }
    if (lane == 0) jb.counts[ref] = 1 + K;
}

// ---------------------------------------------------------------------------
// MFMA matcher (B = 8, G <= 32): the role of the reference's tensor-core
// matcher (src/cuda/common/match.cu:1165-1620, spatial_match_tensor), built
// for CDNA4. A workgroup of four waves takes a block of 4 x 4 references.
// The union of their windows is staged in LDS as fp16 values
// h = fp16((v - mu) 2^S), S the block's power-of-two scale (the largest
// |v - mu| then sits in [2^12, 2^13)). Dot products of reference and
// candidate patches run on
// v_mfma_f32_16x16x32_f16 (16 references x 16 candidates of one row per
// tile), giving d' = |h_r|^2 + |h_c|^2 - 2 h_r.h_c for every pair; wave w
// takes the candidate rows w mod 4. Two passes over the tiles:
//   A. per reference, the four smallest upper bounds U of each lane's
//      candidate column (16 lanes hold disjoint columns): the K-th smallest
//      of those 64 distinct candidates bounds the K-th exact key from above;
//   B. every candidate whose lower bound L does not exceed it is listed (at
//      most 128 per reference for G <= 16, 256 for G <= 32).
// The listed candidates are then scored in the FP32 order of the other
// matchers (rows, then columns, one fmaf per pixel) and ranked by
// (dist, y, x): the records are those of match_run_kernel bit for bit. A
// reference whose list overflows, or a block holding a value fp16 cannot
// carry (non-finite, or |v - mu| >= 8192 so that S would be negative),
// scores its whole window exactly.
//
// Error bound, in the scaled units (every distance times 2^2S; u = 2^-24):
//   per element, |(v - mu) 2^S - h| <= 2^-11 |h| (1 + 2^-11) for a normal h,
//     and < 2^-14 where h was flushed to 0 (below fp16's smallest normal);
//     over a patch, E_c^2 <= kappa (|h_c|^2 + 1), kappa = 1.002 2^-22;
//   |d' - d''| <= C1 (|h_r|^2 + |h_c|^2), d'' the exact SSD of the fp16
//     patches (squares exact, 64-term FP32 sums and the MFMA's FP32
//     accumulation, about 200 u), C1 = 512 u;
//   by the triangle inequality sqrt(d*) lies within E_r + E_c of
//     sqrt(d''), d* the real SSD; with (a + b)^2 <= (1 + t) a^2 +
//     (1 + 1/t) b^2, (a - b)^2 >= (1 - t) a^2 - (1/t - 1) b^2 (t = 2^-6)
//     and E^2 <= 2 E_r^2 + 2 E_c^2 both bounds are linear in (c, |h_c|^2),
//     c = |h_r|^2 - 2 h_r.h_c the MFMA tile: one FMA per pair;
//   the FP32 key differs from d* by at most 66 u d* (taken as 1e-4).
// ---------------------------------------------------------------------------

typedef _Float16 mm_h16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 mm_h16x2 __attribute__((ext_vector_type(2)));
typedef float mm_f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t mm_u32x4 __attribute__((ext_vector_type(4)));

inline constexpr float kMmC1 = 3.0517578125e-05f;    // 512 * 2^-24
inline constexpr float kMmTiny = 6.103515625e-05f;   // 2^-14, fp16's smallest normal
inline constexpr float kMmInf = __builtin_inff();
inline constexpr float kMmT = 0.015625f;              // t = 2^-6
inline constexpr float kMmKappa = 2.3889e-07f;        // 1.002 * 2^-22, rounded up
inline constexpr float kMmGammaU = 2.f * kMmKappa * (1.f + 1.f / kMmT);  // E^2 weight, upper bound
inline constexpr float kMmGammaL = 2.f * kMmKappa / kMmT;                // E^2 weight, lower bound
// Per-pair slopes in |h_c|^2 of the bounds U / (1 + t) and L / (1 - t) (c has
// slope 1): (1 + C1) + gamma_U / (1 + t) and (1 - C1) - gamma_L / (1 - t).
inline constexpr float kMmSlopeU = 1.f + kMmC1 + kMmGammaU / (1.f + kMmT) + 1e-6f;
inline constexpr float kMmSlopeL = 1.f - kMmC1 - kMmGammaL / (1.f - kMmT) - 1e-6f;

// fp16 bits of a scaled value (0 below 2^-14).
__device__ inline uint32_t mm_half(float a) {
    const _Float16 h = fabsf(a) < kMmTiny ? static_cast<_Float16>(0.f) : static_cast<_Float16>(a);
    return __builtin_bit_cast(uint16_t, h);
}

// Eight fp16 values at region row `row`, columns col .. col + 7 (two per
// word; odd columns realigned with a funnel shift).
__device__ inline mm_h16x8 mm_frag(const uint32_t* reg, int pitch, int row, int col) {
    const uint32_t* p = reg + row * pitch + (col >> 1);
    const uint32_t w0 = p[0], w1 = p[1], w2 = p[2], w3 = p[3], w4 = p[4];
    const uint32_t sh = static_cast<uint32_t>(col & 1) * 16u;
    const mm_u32x4 o = {__builtin_amdgcn_alignbit(w1, w0, sh), __builtin_amdgcn_alignbit(w2, w1, sh),
                        __builtin_amdgcn_alignbit(w3, w2, sh), __builtin_amdgcn_alignbit(w4, w3, sh)};
    return __builtin_bit_cast(mm_h16x8, o);
}

__device__ inline float mm_sq(mm_h16x8 f) {
    float s = __builtin_amdgcn_fdot2(mm_h16x2{f[0], f[1]}, mm_h16x2{f[0], f[1]}, 0.f, false);
    s = __builtin_amdgcn_fdot2(mm_h16x2{f[2], f[3]}, mm_h16x2{f[2], f[3]}, s, false);
    s = __builtin_amdgcn_fdot2(mm_h16x2{f[4], f[5]}, mm_h16x2{f[4], f[5]}, s, false);
    return __builtin_amdgcn_fdot2(mm_h16x2{f[6], f[7]}, mm_h16x2{f[6], f[7]}, s, false);
}

// Sum over lanes n, n + 16, n + 32, n + 48 (the four 16-lane rows), equal in
// all four: rows swapped pairwise, then halves (additions commute).
__device__ inline float mm_rows_sum(float v) {
    const uint32_t b = __float_as_uint(v);
    const auto p = __builtin_amdgcn_permlane16_swap(b, b, false, false);
    const float s = __uint_as_float(p[0]) + __uint_as_float(p[1]);
    const uint32_t c = __float_as_uint(s);
    const auto q = __builtin_amdgcn_permlane32_swap(c, c, false, false);
    return __uint_as_float(q[0]) + __uint_as_float(q[1]);
}

__device__ inline float wave_max_f32(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

// v of the lane DPP control CTRL selects (controls that stay inside a row).
template <int CTRL>
__device__ inline float dpp_f32(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

// The value of lane l (uniform).
__device__ inline float lane_f32(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Sum and maximum over the wave in a fixed order, the result in every lane:
// each 16-lane row by DPP (quad_perm [1,0,3,2] and [2,3,0,1], row_half_mirror,
// row_mirror), then the four rows' results. All 64 lanes must call.
__device__ inline float wave_sum_dpp(float v) {
    v += dpp_f32<0xb1>(v);
    v += dpp_f32<0x4e>(v);
    v += dpp_f32<0x141>(v);
    v += dpp_f32<0x140>(v);
    return (lane_f32(v, 0) + lane_f32(v, 16)) + (lane_f32(v, 32) + lane_f32(v, 48));
}

__device__ inline float wave_max_dpp(float v) {
    v = fmaxf(v, dpp_f32<0xb1>(v));
    v = fmaxf(v, dpp_f32<0x4e>(v));
    v = fmaxf(v, dpp_f32<0x141>(v));
    v = fmaxf(v, dpp_f32<0x140>(v));
    return fmaxf(fmaxf(lane_f32(v, 0), lane_f32(v, 16)), fmaxf(lane_f32(v, 32), lane_f32(v, 48)));
}

// The FP32 key of the other matchers for B = 8: rows, then columns, one fmaf
// per pixel.
__device__ inline uint32_t mm_exact(const float* plane, size_t stride, int rx, int ry, int x, int y) {
    const float* a = plane + static_cast<size_t>(ry) * stride + rx;
    const float* b = plane + static_cast<size_t>(y) * stride + x;
    float acc = 0.f;
    for (int py = 0; py < 8; ++py) {
#pragma unroll
        for (int px = 0; px < 8; ++px) {
            const float d = a[py * stride + px] - b[py * stride + px];
            acc = fmaf(d, d, acc);
        }
    }
    return dist_key(acc);
}

// Region row pitch in words: the smallest p >= 8 nt + 5 (16 nt + 8 halves
// and the fifth word of an odd-column read) at which the four 16-lane rows
// of a fragment read (each within 9 consecutive words, rows p apart) fall on
// disjoint banks, for 32 banks (two rows per half-wave) and for 64.
__host__ __device__ inline int mm_pitch(int nt) {
    for (int p = 8 * nt + 5;; ++p) {
        const int a = p % 32;
        if (a < 9 || a > 23) continue;
        const int o[4] = {0, p % 64, (2 * p) % 64, (3 * p) % 64};
        bool ok = true;
        for (int i = 0; i < 4; ++i)
            for (int j = i + 1; j < 4; ++j) {
                const int d = ((o[i] - o[j]) % 64 + 64) % 64;
                ok = ok && d >= 9 && d <= 55;
            }
        if (ok) return p;  // p = 16 mod 64 always qualifies
    }
}

struct MmWin {
    int cx, cy, left, right, top, bottom;
};

__device__ inline MmWin mm_win(const MatchJob& jb, int i, int j) {
    MmWin w;
    w.cx = ref_x(jb.grid, i, jb.width, 8);
    w.cy = ref_y(jb.grid, j, jb.height, 8);
    w.left = max(w.cx - jb.range, 0);
    w.right = min(w.cx + jb.range, jb.width - 8);
    w.top = max(w.cy - jb.range, 0);
    w.bottom = min(w.cy + jb.range, jb.height - 8);
    return w;
}

// Minimum over the 16 lanes of each DPP row (rotations by 8, 4, 2, 1),
// returned to every lane of the row.
__device__ inline float row_min_f32(float v) {
    v = fminf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x128, 0xf, 0xf, false)));
    v = fminf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x124, 0xf, 0xf, false)));
    v = fminf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x122, 0xf, 0xf, false)));
    return fminf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x121, 0xf, 0xf, false)));
}

// d into the ascending four s (the largest of the five dropped).
__device__ inline void top4_insert(float (&s)[4], float d) {
    // top down, so that each entry is overwritten after its last use
    s[3] = __builtin_amdgcn_fmed3f(s[2], s[3], d);
    s[2] = __builtin_amdgcn_fmed3f(s[1], s[2], d);
    s[1] = __builtin_amdgcn_fmed3f(s[0], s[1], d);
    s[0] = fminf(s[0], d);
}

// bf16 bits of v rounded towards +inf (an upper bound survives the exchange).
__device__ inline uint16_t bf16_up(float v) {
    const uint32_t b = __float_as_uint(v);
    const uint32_t hi = b >> 16;
    return static_cast<uint16_t>(((b & 0xffffu) != 0u && (b >> 31) == 0u) ? hi + 1u : hi);
}

// LIST: listed candidates per reference (128 for G <= 16, 256 for G <= 32).
template <int LIST>
__global__ __launch_bounds__(256) void match_mfma_kernel(MatchTable tab) {
    extern __shared__ __align__(16) uint32_t mm_lds[];
    const int tid = threadIdx.x, lane = tid & 63, g = lane >> 4, n = lane & 15;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: row loops branch on scalars
    const int lb = xcd_block();
    const MatchJob& jb = tab.job[find_job(tab, lb)];
    const int blk = lb - jb.wg_begin;
    const int nbx = (jb.grid.nx + 3) >> 2;
    const int by = blk / nbx, bx = blk - by * nbx;
    const int i0 = 4 * bx, j0 = 4 * by;
    const int ncol = min(4, jb.grid.nx - i0), nrow = min(4, jb.grid.ny - j0);
    const float* plane = jb.plane;
    const size_t stride = static_cast<size_t>(jb.stride);
    const int G = jb.group;

    const MmWin w00 = mm_win(jb, i0, j0), wlast = mm_win(jb, i0 + ncol - 1, j0 + nrow - 1);
    const int ux0 = w00.left, ux1 = wlast.right, uy0 = w00.top, uy1 = wlast.bottom;
    const int nt = (ux1 - ux0 + 16) >> 4;
    const int UW = 16 * nt;
    const int RH = uy1 - uy0 + 8;
    const int pitch = mm_pitch(nt);
    const int reg_words = max(RH * pitch, 16 * LIST);
    uint32_t* reg = mm_lds;
    // One area serves, in turn, the staging reductions and the pass-A
    // exchange (xw), the candidate lists (pass B, ranking), and the
    // fallback's keys.
    constexpr int kUnion = 8 * LIST > 2048 ? 8 * LIST : 2048;           // words
    uint32_t* xw = mm_lds + reg_words;
    uint16_t* xch = reinterpret_cast<uint16_t*>(xw);                    // [wave][j][lane][4] bf16
    uint16_t* list = reinterpret_cast<uint16_t*>(xw);                   // [16][LIST]
    int* cnt = reinterpret_cast<int*>(mm_lds + reg_words + kUnion);     // 16
    int* pre = cnt + 16;                                                // 17 (+1 keeps 8-byte alignment)
    int* kmv = pre + 18;                                                // 16
    float* wl = reinterpret_cast<float*>(kmv + 16);                     // 16

    // Per-reference counts K = min(G - 1, window - 1); counters cleared.
    if (tid < 16) {
        const int c = tid >> 2, j = tid & 3;
        int k = 0;
        if (c < ncol && j < nrow) {
            const MmWin w = mm_win(jb, i0 + c, j0 + j);
            k = min(G - 1, (w.right - w.left + 1) * (w.bottom - w.top + 1) - 1);
        }
        kmv[tid] = k;
        cnt[tid] = 0;
    }

    // mu: the mean of the first reference's patch (every wave alike). Two
    // passes over the region, eight words per thread in flight each: the
    // largest |v - mu| (non-finite values make it NaN), then the scaled
    // fp16 values.
    const float mu =
        wave_sum_f32(plane[static_cast<size_t>(w00.cy + (lane >> 3)) * stride + w00.cx + (lane & 7)]) *
        (1.f / 64.f);
    const float inv_pitch = 1.f / static_cast<float>(pitch);
    const int words = RH * pitch;
    auto load8 = [&](int base, float (&v0)[8], float (&v1)[8]) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int wd = base + tid + 256 * i;
            v0[i] = mu;
            v1[i] = mu;
            if (wd < words) {
                const int yy = div_small(wd, inv_pitch), xw2 = wd - yy * pitch;
                const float* row = plane + static_cast<size_t>(uy0 + yy) * stride;
                const int x = ux0 + 2 * xw2;
                if (x < jb.width) v0[i] = row[x];
                if (x + 1 < jb.width) v1[i] = row[x + 1];
            }
        }
    };
    float amax = fabsf(mu - mu);  // NaN when mu is
    for (int base = 0; base < words; base += 8 * 256) {
        float v0[8], v1[8];
        load8(base, v0, v1);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const float a = fabsf(v0[i] - mu), b = fabsf(v1[i] - mu);
            amax = (a > amax || a != a) ? a : amax;
            amax = (b > amax || b != b) ? b : amax;
        }
    }
    {
        const float wm = wave_max_f32(amax);
        const bool wn = __builtin_amdgcn_ballot_w64(amax != amax) != 0;
        if (lane == 0) {
            reinterpret_cast<float*>(xw)[wave] = wm;
            xw[4 + wave] = wn ? 1u : 0u;
        }
    }
    __syncthreads();
    const float bmax = fmaxf(fmaxf(reinterpret_cast<float*>(xw)[0], reinterpret_cast<float*>(xw)[1]),
                             fmaxf(reinterpret_cast<float*>(xw)[2], reinterpret_cast<float*>(xw)[3]));
    // bmax < 2^e: S = 13 - e puts it below 2^13 (|-2 h| < 2^15 fits fp16)
    int ex = 0;
    (void)frexpf(bmax, &ex);
    const int S = bmax > 0.f ? min(13 - ex, 40) : 40;
    const bool anybad = (xw[4] | xw[5] | xw[6] | xw[7]) != 0u || !(bmax < 8192.f) || S < 0;
    __syncthreads();
    if (!anybad) {
        const float sc = ldexpf(1.f, S);
        for (int base = 0; base < words; base += 8 * 256) {
            float v0[8], v1[8];
            load8(base, v0, v1);
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int wd = base + tid + 256 * i;
                if (wd < words) reg[wd] = mm_half((v0[i] - mu) * sc) | (mm_half((v1[i] - mu) * sc) << 16);
            }
        }
    }
    __syncthreads();

    if (!anybad) {
        // A operand: reference n (column n >> 2, row n & 3), patch rows g
        // and g + 4.
        const int acol = n >> 2, arow = n & 3;
        mm_h16x8 A0 = {}, A1 = {};
        if (acol < ncol && arow < nrow) {
            const MmWin w = mm_win(jb, i0 + acol, j0 + arow);
            A0 = mm_frag(reg, pitch, w.cy - uy0 + g, w.cx - ux0);
            A1 = mm_frag(reg, pitch, w.cy - uy0 + 4 + g, w.cx - ux0);
        }
        const float nrA = mm_rows_sum(mm_sq(A0) + mm_sq(A1));  // |h_r|^2 of reference n
        // The reference operand enters as -2 h_r (exact in fp16) and the
        // accumulator starts at |h_r|^2: a tile then holds |h_r|^2 - 2 h_r.h_c.
        A0 *= static_cast<_Float16>(-2.f);
        A1 *= static_cast<_Float16>(-2.f);
        // C layout: lane (g, n) holds references 4 g + j (column g, row j)
        // against candidate n of the tile.
        mm_f32x4 nr;
#pragma unroll
        for (int j = 0; j < 4; ++j) nr[j] = __shfl(nrA, 4 * g + j, 64);
        const bool cval = g < ncol;
        const MmWin cw = mm_win(jb, i0 + min(g, ncol - 1), j0);
        const int ccx = cw.cx, cleft = cval ? cw.left : (1 << 30), cright = cval ? cw.right : -1;
        int rtop[4], rbot[4], rcy[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            // uniform: the row tests below branch on scalars
            const MmWin w = mm_win(jb, i0, j0 + min(j, nrow - 1));
            rcy[j] = __builtin_amdgcn_readfirstlane(w.cy);
            rtop[j] = __builtin_amdgcn_readfirstlane(j < nrow ? w.top : (1 << 30));
            rbot[j] = __builtin_amdgcn_readfirstlane(j < nrow ? w.bottom : -1);
        }

        // Each wave takes the candidate rows y = uy0 + wave (mod 4).
        float s[4][4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int k = 0; k < 4; ++k) s[j][k] = kMmInf;
        float wv[4] = {0.f, 0.f, 0.f, 0.f};
        // One row step over NC tile columns at once: independent chains
        // (fragment loads, norms, two MFMAs, epilogue) for the scheduler to
        // interleave.
        auto pass_cols = [&](auto surv_tag, auto nc_tag, int t0) {
            constexpr bool SURV = decltype(surv_tag)::value;
            constexpr int NC = decltype(nc_tag)::value;
            int col[NC];
            bool xv[NC], xs[NC];
#pragma unroll
            for (int q = 0; q < NC; ++q) {
                col[q] = 16 * (t0 + q) + n;
                const int x = ux0 + col[q];
                xv[q] = x >= cleft && x <= cright;
                xs[q] = x == ccx;
            }
            int y = uy0 + wave;
            if (y > uy1) return;
            mm_h16x8 F0[NC];
            float T0[NC];
#pragma unroll
            for (int q = 0; q < NC; ++q) {
                F0[q] = mm_frag(reg, pitch, y - uy0 + g, col[q]);
                T0[q] = mm_rows_sum(mm_sq(F0[q]));
            }
            for (; y <= uy1; y += 4) {
                mm_h16x8 F1[NC];
                float T1[NC];
                mm_f32x4 c[NC];
#pragma unroll
                for (int q = 0; q < NC; ++q) {
                    F1[q] = mm_frag(reg, pitch, y - uy0 + 4 + g, col[q]);
                    T1[q] = mm_rows_sum(mm_sq(F1[q]));
                    c[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A0, F0[q], nr, 0, 0, 0);
                    c[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A1, F1[q], c[q], 0, 0, 0);
                }
#pragma unroll
                for (int q = 0; q < NC; ++q) {
                    const float ncx = xv[q] ? T0[q] + T1[q] : kMmInf;
                    bool sv[4] = {false, false, false, false};
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        // reference row j's window holds row y: uniform
                        if (y >= rtop[j] && y <= rbot[j]) {
                            // pass A: U / (1 + t) up to a per-reference
                            // constant; pass B: L / (1 - t) likewise
                            float d = fmaf(ncx, SURV ? kMmSlopeL : kMmSlopeU, c[q][j]);
                            if (y == rcy[j]) d = xs[q] ? kMmInf : d;  // the reference itself
                            if constexpr (SURV)
                                sv[j] = d <= wv[j];
                            else
                                top4_insert(s[j], d);
                        }
                    }
                    if constexpr (SURV) {
                        if (__builtin_amdgcn_ballot_w64(sv[0] || sv[1] || sv[2] || sv[3]) != 0) {
                            const int u = (y - uy0) * UW + col[q];
#pragma unroll
                            for (int j = 0; j < 4; ++j) {
                                if (sv[j]) {
                                    const int m = 4 * g + j;
                                    const int k = atomicAdd(&cnt[m], 1);
                                    if (k < LIST) list[m * LIST + k] = static_cast<uint16_t>(u);
                                }
                            }
                        }
                    }
                    F0[q] = F1[q];
                    T0[q] = T1[q];
                }
            }
        };
        auto pass = [&](auto surv_tag) {
            int t = 0;
            for (; t + 2 <= nt; t += 2) pass_cols(surv_tag, std::integral_constant<int, 2>{}, t);
            if (t < nt) pass_cols(surv_tag, std::integral_constant<int, 1>{}, t);
        };

        pass(std::false_type{});
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int k = 0; k < 4; ++k) xch[((wave * 4 + j) * 64 + lane) * 4 + k] = bf16_up(s[j][k]);
        __syncthreads();
        // Wave w: references of row j = w. Each lane merges the four waves'
        // four smallest of its candidate column; the K-th smallest over the
        // row's 16 lanes (64 distinct candidates) by rounds of a row minimum.
        {
            float m4[4] = {kMmInf, kMmInf, kMmInf, kMmInf};
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    top4_insert(m4, __uint_as_float(static_cast<uint32_t>(xch[((v * 4 + wave) * 64 + lane) * 4 + k])
                                                    << 16));
            const int K = kmv[4 * g + wave];
            int taken = 0;
            bool found = K == 0;
            float td = kMmInf;
            for (int round = 0; round < 32; ++round) {
                if (__builtin_amdgcn_ballot_w64(!found) == 0) break;
                const float mn = row_min_f32(m4[0]);
                const bool eq = m4[0] == mn;
                const unsigned long long bal = __builtin_amdgcn_ballot_w64(eq);
                taken += __popcll((bal >> (16 * g)) & 0xffffull);
                if (eq) {
                    m4[0] = m4[1];
                    m4[1] = m4[2];
                    m4[2] = m4[3];
                    m4[3] = kMmInf;
                }
                if (!found && taken >= K) {
                    found = true;
                    td = mn;
                }
            }
            if (n == 0) {
                const float nrm = wave == 0 ? nr[0] : wave == 1 ? nr[1] : wave == 2 ? nr[2] : nr[3];
                float wt = -kMmInf;
                if (K > 0) {
                    if (!(td < kMmInf)) {
                        wt = 3.402823466e38f;  // every finite d' (masked pairs are +inf)
                    } else {
                        // U = (1 + t) td + beta_U bounds the K-th key;
                        // a candidate survives while its L = (1 - t) v -
                        // beta_L does not exceed U.
                        const float betaU = (1.f + kMmT) * kMmC1 * nrm + kMmGammaU * (nrm + 2.f);
                        const float betaL = (1.f - kMmT) * kMmC1 * nrm + kMmGammaL * (nrm + 2.f);
                        const float U = ((1.f + kMmT) * td + betaU) * 1.0001f;
                        wt = (U * 1.0001f + betaL) / (1.f - kMmT) * 1.00001f;
                    }
                }
                wl[4 * g + wave] = wt;
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 4; ++j) wv[j] = wl[4 * g + j];
        pass(std::true_type{});
    }
    __syncthreads();

    // Listed candidates: exact keys (into the region's words), then ranks.
    if (tid == 0) {
        int acc = 0;
        for (int m = 0; m < 16; ++m) {
            pre[m] = acc;
            const int c = cnt[m];
            acc += (!anybad && kmv[m] > 0 && c <= LIST) ? c : 0;
        }
        pre[16] = acc;
    }
    __syncthreads();
    const int total = pre[16];
    const float inv_uw = 1.f / static_cast<float>(UW);
    uint32_t* dk = reg;
    auto item = [&](int q, int& m, int& sl) {
        m = 0;
#pragma unroll
        for (int b = 8; b >= 1; b >>= 1)
            if (m + b < 16 && q >= pre[m + b]) m += b;
        sl = q - pre[m];
    };
    for (int q = tid; q < total; q += 256) {
        int m, sl;
        item(q, m, sl);
        const int u = list[m * LIST + sl];
        const int uy = div_small(u, inv_uw), ux = u - uy * UW;
        const MmWin w = mm_win(jb, i0 + (m >> 2), j0 + (m & 3));
        dk[q] = mm_exact(plane, stride, w.cx, w.cy, ux0 + ux, uy0 + uy);
    }
    __syncthreads();
    for (int q = tid; q < total; q += 256) {
        int m, sl;
        item(q, m, sl);
        const int c = pre[m + 1] - pre[m];
        const uint32_t d = dk[q];
        const int u = list[m * LIST + sl];
        int rank = 0;
        for (int s2 = 0; s2 < c; ++s2) {
            const uint32_t d2 = dk[pre[m] + s2];
            const int u2 = list[m * LIST + s2];
            rank += (d2 < d || (d2 == d && u2 < u)) ? 1 : 0;
        }
        if (rank < kmv[m]) {
            const int uy = div_small(u, inv_uw), ux = u - uy * UW;
            const int ref_id = (j0 + (m & 3)) * jb.grid.nx + i0 + (m >> 2);
            jb.out[static_cast<size_t>(ref_id) * G + 1 + rank] = MatchRec{ux0 + ux, uy0 + uy, jb.t0, d};
        }
    }
    if (tid < 16 && (tid >> 2) < ncol && (tid & 3) < nrow) {
        const MmWin w = mm_win(jb, i0 + (tid >> 2), j0 + (tid & 3));
        const int ref_id = (j0 + (tid & 3)) * jb.grid.nx + i0 + (tid >> 2);
        jb.out[static_cast<size_t>(ref_id) * G] = MatchRec{w.cx, w.cy, jb.t0, 0u};
        jb.counts[ref_id] = 1 + kmv[tid];
    }

    __syncthreads();  // the lists' area becomes the fallback's keys
    // Whole windows, exactly, for the references the bound could not serve
    // (wave w takes references w, w + 4, ...): chunks of 64 candidates merged
    // by rank into the running G - 1 best.
    unsigned long long* ck = reinterpret_cast<unsigned long long*>(xw) + wave * 128;  // 64
    unsigned long long* bk = ck + 64;                                                  // 32
    unsigned long long* nbk = bk + 32;                                                 // 32
    for (int m = wave; m < 16; m += 4) {
        const int K = kmv[m];
        if (K == 0 || !(anybad || cnt[m] > LIST)) continue;
        const MmWin w = mm_win(jb, i0 + (m >> 2), j0 + (m & 3));
        const int ww = w.right - w.left + 1, ncand = ww * (w.bottom - w.top + 1);
        const int self = (w.cy - w.top) * ww + (w.cx - w.left);
        const float inv_ww = 1.f / static_cast<float>(ww);
        int nb = 0;
        for (int base = 0; base < ncand; base += 64) {
            const int c = base + lane;
            unsigned long long key = ~0ull;
            if (c < ncand && c != self) {
                const int dy = div_small(c, inv_ww), dx = c - dy * ww;
                key = (static_cast<unsigned long long>(mm_exact(plane, stride, w.cx, w.cy, w.left + dx, w.top + dy))
                       << 32) |
                      static_cast<uint32_t>(c);
            }
            ck[lane] = key;
            wave_sync();
            int rank = 0, brank = 0;
            unsigned long long bkey = ~0ull;
            for (int q = 0; q < 64; ++q) rank += ck[q] < key ? 1 : 0;
            for (int j = 0; j < nb; ++j) rank += bk[j] < key ? 1 : 0;
            if (lane < nb) {
                bkey = bk[lane];
                brank = lane;
                for (int q = 0; q < 64; ++q) brank += ck[q] < bkey ? 1 : 0;
            }
            const int nvalid = __popcll(__builtin_amdgcn_ballot_w64(key != ~0ull));
            wave_sync();
            if (key != ~0ull && rank < K) nbk[rank] = key;
            if (lane < nb && brank < K) nbk[brank] = bkey;
            nb = min(K, nb + nvalid);
            wave_sync();
            if (lane < nb) bk[lane] = nbk[lane];
            wave_sync();
        }
        if (lane < K) {
            const unsigned long long key = bk[lane];
            const int c = static_cast<int>(static_cast<uint32_t>(key));
            const int dy = div_small(c, inv_ww), dx = c - dy * ww;
            const int ref_id = (j0 + (m & 3)) * jb.grid.nx + i0 + (m >> 2);
            jb.out[static_cast<size_t>(ref_id) * G + 1 + lane] =
                MatchRec{w.left + dx, w.top + dy, jb.t0, static_cast<uint32_t>(key >> 32)};
        }
        wave_sync();
    }
}

// ---------------------------------------------------------------------------
// Group filter (src/cpu/bm/dct8.cpp:926-1133 Bm3dFilterGroup).
// ---------------------------------------------------------------------------

// Factor of position j of the in-place orthonormal Haar of n samples
// (src/cpu/bm/dct8.cpp:580-591 HaarScale).
__device__ inline float haar_scale(int n, int j) {
    int e;
    if (j == 0) {
        e = 31 - __builtin_clz(static_cast<unsigned>(n));
    } else {
        e = __builtin_ctz(static_cast<unsigned>(j)) + 1;
    }
    float f = 1.f;
    for (int i = 0; i < e / 2; ++i) f *= 0.5f;
    return (e & 1) ? f * 0.70710678118654752f : f;
}

// In-place orthonormal Haar along one line at base[i * stride], i < n
// (src/cpu/bm/dct8.cpp:600-625 HaarLines).
__device__ inline void haar_forward(float* base, int n, int stride) {
    for (int h = 1; h < n; h <<= 1)
        for (int i = 0; i < n; i += 2 * h) {
            const float a = base[i * stride], b = base[(i + h) * stride];
            base[i * stride] = a + b;
            base[(i + h) * stride] = a - b;
        }
    for (int j = 0; j < n; ++j) base[j * stride] *= haar_scale(n, j);
}

__device__ inline void haar_inverse(float* base, int n, int stride) {
    for (int j = 0; j < n; ++j) base[j * stride] *= haar_scale(n, j);
    for (int h = n >> 1; h >= 1; h >>= 1)
        for (int i = 0; i < n; i += 2 * h) {
            const float a = base[i * stride], b = base[(i + h) * stride];
            base[i * stride] = a + b;
            base[(i + h) * stride] = a - b;
        }
}

// DCT of the lines of every patch of a cube (patch stride ps floats), by the
// 64 lanes of one wave. ROWS: lines along x (sample stride 1); else along y.
template <int B, bool ROWS>
__device__ inline void dct_lines(float* cube, int group, int ps, bool inverse, int lane) {
    if constexpr (B > 1) {
        const float* m = c_dct + dct_offset<B>();
        const int lines = group * B;
        for (int L = lane; L < lines; L += 64) {
            const int p = L / B, q = L - p * B;
            float* base = cube + p * ps + (ROWS ? q * B : q);
            constexpr int ss = ROWS ? 1 : B;
            float v[B], o[B];
#pragma unroll
            for (int i = 0; i < B; ++i) v[i] = base[i * ss];
            if (inverse) {
#pragma unroll
                for (int k = 0; k < B; ++k) {
                    float acc = 0.f;
#pragma unroll
                    for (int i = 0; i < B; ++i) acc = fmaf(m[i * B + k], v[i], acc);
                    o[k] = acc;
                }
            } else {
#pragma unroll
                for (int k = 0; k < B; ++k) {
                    float acc = 0.f;
#pragma unroll
                    for (int i = 0; i < B; ++i) acc = fmaf(m[k * B + i], v[i], acc);
                    o[k] = acc;
                }
            }
#pragma unroll
            for (int i = 0; i < B; ++i) base[i * ss] = o[i];
        }
    }
}

// Gathers the k matched patches of a group into an LDS cube (patch stride
// B*B + 1), zero beyond k up to the configured group size G. src_of(t) and
// wref_of(t) give the plane of temporal slot t (the record's t).
template <int B, bool WIENER, typename SrcOf, typename WrefOf>
__device__ inline void gather_group_wave(const MatchRec* rec, int k, int G, int stride, int wref_stride, SrcOf src_of,
                                         WrefOf wref_of, float* cube, float* rc, int lane) {
    constexpr int A = B * B;
    constexpr int PS = A + 1;
    for (int j = 0; j < G; ++j) {
        int mx = 0, my = 0, mt = 0;
        if (j < k) {
            mx = rec[j].x;
            my = rec[j].y;
            mt = rec[j].t;
        }
        const float* sp = j < k ? src_of(mt) : nullptr;
        const float* wp = (WIENER && j < k) ? wref_of(mt) : nullptr;
        for (int e = lane; e < A; e += 64) {
            const int py = e / B, px = e - (e / B) * B;
            float v = 0.f, w = 0.f;
            if (j < k) {
                v = sp[static_cast<size_t>(my + py) * stride + mx + px];
                if constexpr (WIENER) w = wp[static_cast<size_t>(my + py) * wref_stride + mx + px];
            }
            cube[j * PS + e] = v;
            if constexpr (WIENER) rc[j * PS + e] = w;
        }
    }
    wave_sync();
}

// One group by one wave: gather (spatial: every patch from the job's own
// planes), then the transform. Returns the aggregation weight.
template <int B, bool WIENER>
__device__ inline float transform_group_wave(int G, float sigma, float* cube, float* rc, int lane);

template <int B, bool WIENER>
__device__ inline float filter_group_wave(const FilterJob& jb, const MatchRec* rec, int k, float* cube, float* rc,
                                          int lane) {
    gather_group_wave<B, WIENER>(
        rec, k, jb.group, jb.stride, jb.wref_stride, [&](int) { return jb.src; }, [&](int) { return jb.wref; }, cube,
        rc, lane);
    return transform_group_wave<B, WIENER>(jb.group, jb.sigma, cube, rc, lane);
}

// Forward 3-D transform of the cube (and of the Wiener reference cube),
// shrink, inverse; the filtered patches stay in `cube`. Returns the
// aggregation weight, the same in every lane.
template <int B, bool WIENER>
__device__ inline float transform_group_wave(int G, float sigma, float* cube, float* rc, int lane) {
    constexpr int A = B * B;
    constexpr int PS = A + 1;
    dct_lines<B, true>(cube, G, PS, false, lane);
    if constexpr (WIENER) dct_lines<B, true>(rc, G, PS, false, lane);
    wave_sync();
    dct_lines<B, false>(cube, G, PS, false, lane);
    if constexpr (WIENER) dct_lines<B, false>(rc, G, PS, false, lane);
    wave_sync();

    int kept = 0;
    float w2 = 0.f;
    const float thr = 2.7f * sigma;  // include/nss/params.hpp:46 kBmHardLambda
    const float sig2 = sigma * sigma;
    for (int f = lane; f < A; f += 64) {
        haar_forward(cube + f, G, PS);
        if constexpr (WIENER) haar_forward(rc + f, G, PS);
        for (int j = 0; j < G; ++j) {
            float& c = cube[j * PS + f];
            const bool dc = (j == 0 && f == 0);
            if constexpr (WIENER) {
                float w = 1.f;
                if (!dc) {
                    const float rr = rc[j * PS + f];
                    const float r2 = rr * rr;
                    const float den = r2 + sig2;
                    w = den > 0.f ? r2 / den : 0.f;
                }
                c *= w;
                w2 = fmaf(w, w, w2);
            } else {
                if (!dc && fabsf(c) < thr)
                    c = 0.f;
                else
                    ++kept;
            }
        }
        haar_inverse(cube + f, G, PS);
    }
    wave_sync();
    dct_lines<B, false>(cube, G, PS, true, lane);
    wave_sync();
    dct_lines<B, true>(cube, G, PS, true, lane);
    wave_sync();
    if constexpr (WIENER) {
        return 1.f / fmaxf(wave_sum_f32(w2), 1e-12f);
    } else {
        return 1.f / static_cast<float>(max(wave_sum_i32(kept), 1));
    }
}

__device__ inline unsigned long long to_fixed(float v) {
    return static_cast<unsigned long long>(__float2ll_rn(v * 4294967296.0f));
}

__device__ inline void lds_add(unsigned long long* p, unsigned long long v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Tile epilogue: the tile's exact int64 numerator sums as its partials (2^32
// units), so that the finish adds every pixel's sums exactly and the output
// does not depend on how the grid was cut into tiles. The weights are not
// written per pixel: each group's weight goes out once (FilterJob::gweight)
// and the finish box-sums them from the records. pn points at column ax0 & 3
// of the tile's first row (rows of `pitch` entries). Called by every thread
// of the block after the accumulation's barrier.
__device__ inline void partials_out(int aw, int ah, int pitch, const unsigned long long* num, unsigned long long* pn) {
    const int nt = static_cast<int>(blockDim.x);
    for (int e = threadIdx.x; e < aw * ah; e += nt) {
        const int y = e / aw, x = e - y * aw;
        pn[y * pitch + x] = num[e];
    }
}

// Temporal tile epilogue: den = B x B box sum of the corner weights
// (horizontal pass into scratch, then vertical), and num, as the tile's float
// partials (rows of
// `pitch` floats; pn / pd point at column ax0 & 3 of the tile's first row, so
// that x = 0 mod 4 is 16-byte aligned in every tile). Called by every thread
// of the block after the accumulation's barrier; scratch holds
// (ah - B + 1) * aw floats.
template <int B>
__device__ inline void partials_out_float(int aw, int ah, int pitch, const unsigned long long* num,
                                    const unsigned long long* corner, float* scratch, float* pn, float* pd) {
    const int cw = aw - B + 1, ch = ah - B + 1;
    const int nt = static_cast<int>(blockDim.x);
    if (!pd) {  // numerator only (the weights went to the frame's corner plane)
        for (int e = threadIdx.x; e < aw * ah; e += nt) {
            const int y = e / aw, x = e - (e / aw) * aw;
            pn[y * pitch + x] =
                static_cast<float>(static_cast<double>(static_cast<long long>(num[e])) * (1.0 / 4294967296.0));
        }
        return;
    }
    // Horizontal: a thread slides the B-wide window over 8 outputs of a row
    // (exact int64 sums, so the order of the adds does not matter).
    const int segs = (aw + 7) >> 3;
    for (int e = threadIdx.x; e < ch * segs; e += nt) {
        const int y = e / segs, x0 = (e - y * segs) * 8;
        const unsigned long long* row = corner + y * cw;
        unsigned long long w[B + 7];  // corners x0 - B + 1 .. x0 + 7
#pragma unroll
        for (int i = 0; i < B + 7; ++i) {
            const int cx = x0 - B + 1 + i;
            w[i] = (cx >= 0 && cx < cw) ? row[cx] : 0ull;
        }
        unsigned long long sum = 0;
#pragma unroll
        for (int i = 0; i < B - 1; ++i) sum += w[i];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            sum += w[B - 1 + j];
            if (x0 + j < aw)
                scratch[y * aw + x0 + j] =
                    static_cast<float>(static_cast<double>(static_cast<long long>(sum)) * (1.0 / 4294967296.0));
            sum -= w[j];
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < aw * ah; e += nt) {
        const int y = e / aw, x = e - (e / aw) * aw;
        float d = 0.f;
#pragma unroll
        for (int i = 0; i < B; ++i) {
            const int cy = y - i;
            if (cy >= 0 && cy < ch) d += scratch[cy * aw + x];
        }
        pn[y * pitch + x] = static_cast<float>(static_cast<double>(static_cast<long long>(num[e])) * (1.0 / 4294967296.0));
        pd[y * pitch + x] = d;
    }
}

__device__ inline void tile_partials(const FilterJob& jb, int tile, int ax0, const unsigned long long* num) {
    const size_t base = static_cast<size_t>(tile) * jb.tiles.pitch * jb.tiles.ah + (ax0 & 3);
    partials_out(jb.tiles.aw, jb.tiles.ah, jb.tiles.pitch, num, jb.part_num + base);
}

// Tile path: numerator sums of a tile of groups in LDS, then exact partials;
// each group's weight to gweight. LDS: num [ah][aw] int64 | per-wave cubes.
template <int B, bool WIENER>
__global__ __launch_bounds__(256) void filter_tile_kernel(FilterTable tab) {
    extern __shared__ __align__(16) unsigned long long acc[];
    constexpr int A = B * B;
    constexpr int PS = A + 1;
    const int lb = xcd_block();
    const FilterJob& jb = tab.job[find_job(tab, lb)];
    const TileGeom& tg = jb.tiles;
    const int tile = lb - jb.wg_begin;  // partials index
    const int tx = (tile + jb.tile0) % tg.ntx, ty = (tile + jb.tile0) / tg.ntx;
    const int aw = tg.aw, ah = tg.ah;
    const int ax0 = tg.ax0(tx, jb.width, B, jb.grid.step, jb.range);
    const int ay0 = tg.ay0(ty, jb.height, B, jb.grid.step, jb.range);
    unsigned long long* num = acc;
    float* scratch = reinterpret_cast<float*>(acc + aw * ah);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int e = threadIdx.x; e < aw * ah; e += 256) acc[e] = 0ull;
    __syncthreads();

    const int G = jb.group;
    const int cube_floats = G * PS * (WIENER ? 2 : 1);
    float* cube = scratch + wave * cube_floats;
    float* rc = cube + G * PS;
    const int i_begin = tx * tg.tw, j_begin = ty * tg.th;
    const int ni = min(tg.tw, jb.grid.nx - i_begin), nj = min(tg.th, jb.grid.ny - j_begin);
    for (int g = wave; g < ni * nj; g += 4) {
        const int gj = g / ni, gi = g - gj * ni;
        const int rid = (j_begin + gj - jb.grid.row0) * jb.grid.nx + i_begin + gi;
        const int k = jb.counts[rid];
        const MatchRec* rec = jb.matches + static_cast<size_t>(rid) * G;
        const float w = filter_group_wave<B, WIENER>(jb, rec, k, cube, rc, lane);
        const unsigned long long wfix = to_fixed(w);
        for (int j = 0; j < k; ++j) {
            const int lx = rec[j].x - ax0, ly = rec[j].y - ay0;
            for (int e = lane; e < A; e += 64) {
                const int py = e / B, px = e - (e / B) * B;
                lds_add(num + (ly + py) * aw + lx + px, to_fixed(w * cube[j * PS + e]));
            }
        }
        if (lane == 0) jb.gweight[rid] = wfix;
        wave_sync();
    }
    __syncthreads();
    tile_partials(jb, tile, ax0, num);
}

// ---------------------------------------------------------------------------
// Block-8 group filter in registers (G in {8, 16, 32}). One wave per group;
// three layouts of the G x 8 x 8 cube over the 64 lanes, with LDS transposes
// between them (per-wave buffer T, address g*72 + row*9 + col):
//   rows    lane (gb, y): rows y of patches g = gb + 8m, 8 samples over x
//   columns lane (gb, kx): columns kx of the same patches, 8 samples over y
//   group   lane (ky, kx): coefficient (ky, kx) of all G patches
// Row DCT, column DCT and the group-axis Haar + shrink are then lane-local.
// ---------------------------------------------------------------------------
inline constexpr int kTs = 72;  // patch stride in the transpose buffer
inline constexpr int kTr = 9;   // row stride

// 8-point orthonormal DCT-II (forward) or its transpose, DCT-III (inverse),
// by the even/odd split of the definition X_k = a_k sum_n x_n cos(pi (2n+1) k / 16):
// X_{2m} is the 4-point transform of x_n + x_{7-n} (split once more into
// sums and differences), X_{2m+1} the 4 x 4 cosine matrix times
// x_n - x_{7-n}; the inverse runs the same steps transposed. 36 operations
// instead of 64 multiply-adds; the constants are cos(j pi / 16) / 2 and
// 1 / sqrt(8) rounded from double.
template <bool INVERSE>
__device__ inline void dct8(float v[8]) {
    constexpr float k0 = 0.353553391f;  // 1/sqrt(8) = cos(4 pi/16) / 2
    constexpr float c1 = 0.490392640f, c2 = 0.461939766f, c3 = 0.415734806f;
    constexpr float c5 = 0.277785117f, c6 = 0.191341716f, c7 = 0.0975451610f;
    if (!INVERSE) {
        const float a0 = v[0] + v[7], a1 = v[1] + v[6], a2 = v[2] + v[5], a3 = v[3] + v[4];
        const float b0 = v[0] - v[7], b1 = v[1] - v[6], b2 = v[2] - v[5], b3 = v[3] - v[4];
        const float e0 = a0 + a3, e1 = a1 + a2, d0 = a0 - a3, d1 = a1 - a2;
        v[0] = k0 * (e0 + e1);
        v[4] = k0 * (e0 - e1);
        v[2] = fmaf(c2, d0, c6 * d1);
        v[6] = fmaf(c6, d0, -c2 * d1);
        v[1] = fmaf(c1, b0, fmaf(c3, b1, fmaf(c5, b2, c7 * b3)));
        v[3] = fmaf(c3, b0, fmaf(-c7, b1, fmaf(-c1, b2, -c5 * b3)));
        v[5] = fmaf(c5, b0, fmaf(-c1, b1, fmaf(c7, b2, c3 * b3)));
        v[7] = fmaf(c7, b0, fmaf(-c5, b1, fmaf(c3, b2, -c1 * b3)));
    } else {
        const float s = k0 * (v[0] + v[4]), t = k0 * (v[0] - v[4]);
        const float p = fmaf(c2, v[2], c6 * v[6]), q = fmaf(c6, v[2], -c2 * v[6]);
        const float e0 = s + p, e3 = s - p, e1 = t + q, e2 = t - q;
        const float o0 = fmaf(c1, v[1], fmaf(c3, v[3], fmaf(c5, v[5], c7 * v[7])));
        const float o1 = fmaf(c3, v[1], fmaf(-c7, v[3], fmaf(-c1, v[5], -c5 * v[7])));
        const float o2 = fmaf(c5, v[1], fmaf(-c1, v[3], fmaf(c7, v[5], c3 * v[7])));
        const float o3 = fmaf(c7, v[1], fmaf(-c5, v[3], fmaf(c3, v[5], -c1 * v[7])));
        v[0] = e0 + o0;
        v[7] = e0 - o0;
        v[1] = e1 + o1;
        v[6] = e1 - o1;
        v[2] = e2 + o2;
        v[5] = e2 - o2;
        v[3] = e3 + o3;
        v[4] = e3 - o3;
    }
}

// rows (or columns) layout -> the transposed layout within each patch.
template <int M>
__device__ inline void transpose_rc(float (&v)[M][8], float* T, int lane) {
    const int gb = lane >> 3, q = lane & 7;
#pragma unroll
    for (int m = 0; m < M; ++m)
#pragma unroll
        for (int i = 0; i < 8; ++i) T[(gb + 8 * m) * kTs + q * kTr + i] = v[m][i];
    wave_sync();
#pragma unroll
    for (int m = 0; m < M; ++m)
#pragma unroll
        for (int i = 0; i < 8; ++i) v[m][i] = T[(gb + 8 * m) * kTs + i * kTr + q];
    wave_sync();
}

// columns layout (lane (gb, kx) holds ky = 0..7) -> group layout.
template <int G>
__device__ inline void columns_to_group(float (&v)[G / 8][8], float (&c)[G], float* T, int lane) {
    const int gb = lane >> 3, kx = lane & 7;
#pragma unroll
    for (int m = 0; m < G / 8; ++m)
#pragma unroll
        for (int ky = 0; ky < 8; ++ky) T[(gb + 8 * m) * kTs + ky * kTr + kx] = v[m][ky];
    wave_sync();
    const int ky = lane >> 3;
#pragma unroll
    for (int g = 0; g < G; ++g) c[g] = T[g * kTs + ky * kTr + kx];
    wave_sync();
}

template <int G>
__device__ inline void group_to_columns(float (&c)[G], float (&v)[G / 8][8], float* T, int lane) {
    const int ky = lane >> 3, kx = lane & 7;
#pragma unroll
    for (int g = 0; g < G; ++g) T[g * kTs + ky * kTr + kx] = c[g];
    wave_sync();
    const int gb = lane >> 3;
#pragma unroll
    for (int m = 0; m < G / 8; ++m)
#pragma unroll
        for (int k = 0; k < 8; ++k) v[m][k] = T[(gb + 8 * m) * kTs + k * kTr + kx];
    wave_sync();
}

template <int G>
__device__ inline void haar_forward_reg(float (&c)[G]) {
#pragma unroll
    for (int h = 1; h < G; h <<= 1)
#pragma unroll
        for (int i = 0; i < G; i += 2 * h) {
            const float a = c[i], b = c[i + h];
            c[i] = a + b;
            c[i + h] = a - b;
        }
#pragma unroll
    for (int j = 0; j < G; ++j) c[j] *= haar_scale(G, j);
}

template <int G>
__device__ inline void haar_inverse_reg(float (&c)[G]) {
#pragma unroll
    for (int j = 0; j < G; ++j) c[j] *= haar_scale(G, j);
#pragma unroll
    for (int h = G >> 1; h >= 1; h >>= 1)
#pragma unroll
        for (int i = 0; i < G; i += 2 * h) {
            const float a = c[i], b = c[i + h];
            c[i] = a + b;
            c[i + h] = a - b;
        }
}

// Forward: rows (spatial) -> group-layout 3-D coefficients.
template <int G>
__device__ inline void forward_b8(float (&v)[G / 8][8], float (&c)[G], float* T, int lane) {
#pragma unroll
    for (int m = 0; m < G / 8; ++m) dct8<false>(v[m]);
    transpose_rc<G / 8>(v, T, lane);
#pragma unroll
    for (int m = 0; m < G / 8; ++m) dct8<false>(v[m]);
    columns_to_group<G>(v, c, T, lane);
    haar_forward_reg<G>(c);
}

// Gathers the rows this lane owns (zero beyond k).
template <int G>
__device__ inline void gather_rows_b8(const float* plane, int stride, const MatchRec* rec, int k, int lane,
                                      float (&v)[G / 8][8]) {
    const int gb = lane >> 3, y = lane & 7;
#pragma unroll
    for (int m = 0; m < G / 8; ++m) {
        const int g = gb + 8 * m;
        if (g < k) {
            const float* p = plane + static_cast<size_t>(rec[g].y + y) * stride + rec[g].x;
#pragma unroll
            for (int x = 0; x < 8; ++x) v[m][x] = p[x];
        } else {
#pragma unroll
            for (int x = 0; x < 8; ++x) v[m][x] = 0.f;
        }
    }
}

// Tile kernel for block 8: LDS = num [ah][aw] int64 | match records of the
// tile | per-wave transpose buffers.
template <int G, bool WIENER>
__global__ __launch_bounds__(256) void filter_tile_b8_kernel(FilterTable tab) {
    extern __shared__ __align__(16) unsigned long long acc[];
    constexpr int B = 8;
    constexpr int M = G / 8;
    const int lb = xcd_block();
    const FilterJob& jb = tab.job[find_job(tab, lb)];
    const TileGeom& tg = jb.tiles;
    const int tile = lb - jb.wg_begin;  // partials index
    const int tx = (tile + jb.tile0) % tg.ntx, ty = (tile + jb.tile0) / tg.ntx;
    const int aw = tg.aw, ah = tg.ah;
    const int ax0 = tg.ax0(tx, jb.width, B, jb.grid.step, jb.range);
    const int ay0 = tg.ay0(ty, jb.height, B, jb.grid.step, jb.range);
    const int i_begin = tx * tg.tw, j_begin = ty * tg.th;
    const int ni = min(tg.tw, jb.grid.nx - i_begin), nj = min(tg.th, jb.grid.ny - j_begin);
    const int ngroups = ni * nj;
    unsigned long long* num = acc;
    MatchRec* recs = reinterpret_cast<MatchRec*>(acc + aw * ah);
    int* cnt = reinterpret_cast<int*>(recs + tg.tw * tg.th * G);
    float* scratch = reinterpret_cast<float*>(cnt + ((tg.tw * tg.th + 3) & ~3));
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float* T = scratch + wave * (G * kTs);
    for (int e = threadIdx.x; e < aw * ah; e += 256) acc[e] = 0ull;
    // The tile's match records and counts, row by row of the grid.
    for (int e = threadIdx.x; e < ngroups * G; e += 256) {
        const int g = e / G, j = e - g * G;
        const int gj = g / ni, gi = g - gj * ni;
        const int rid = (j_begin + gj - jb.grid.row0) * jb.grid.nx + i_begin + gi;
        recs[e] = jb.matches[static_cast<size_t>(rid) * G + j];
    }
    for (int g = threadIdx.x; g < ngroups; g += 256) {
        const int gj = g / ni, gi = g - gj * ni;
        cnt[g] = jb.counts[(j_begin + gj - jb.grid.row0) * jb.grid.nx + i_begin + gi];
    }
    __syncthreads();

    const float thr = 2.7f * jb.sigma;  // include/nss/params.hpp:46 kBmHardLambda
    const float sig2 = jb.sigma * jb.sigma;
    float v[M][8], vr[M][8];
    int g = wave;
    if (g < ngroups) {
        gather_rows_b8<G>(jb.src, jb.stride, recs + g * G, cnt[g], lane, v);
        if constexpr (WIENER) gather_rows_b8<G>(jb.wref, jb.wref_stride, recs + g * G, cnt[g], lane, vr);
    }
    for (; g < ngroups; g += 4) {
        const int k = cnt[g];
        const MatchRec* rec = recs + g * G;
        float c[G];
        float r[G];
        forward_b8<G>(v, c, T, lane);
        if constexpr (WIENER) forward_b8<G>(vr, r, T, lane);
        // Prefetch the next group's patches while this one is filtered.
        const int gn = g + 4;
        if (gn < ngroups) {
            gather_rows_b8<G>(jb.src, jb.stride, recs + gn * G, cnt[gn], lane, v);
            if constexpr (WIENER) gather_rows_b8<G>(jb.wref, jb.wref_stride, recs + gn * G, cnt[gn], lane, vr);
        }
        int kept = 0;
        float w2 = 0.f;
#pragma unroll
        for (int j = 0; j < G; ++j) {
            const bool dc = (lane == 0 && j == 0);
            if constexpr (WIENER) {
                float w = 1.f;
                if (!dc) {
                    const float r2 = r[j] * r[j];
                    const float den = r2 + sig2;
                    w = den > 0.f ? r2 / den : 0.f;
                }
                c[j] *= w;
                w2 = fmaf(w, w, w2);
            } else {
                if (!dc && fabsf(c[j]) < thr)
                    c[j] = 0.f;
                else
                    ++kept;
            }
        }
        haar_inverse_reg<G>(c);
        float u[M][8];
        group_to_columns<G>(c, u, T, lane);
#pragma unroll
        for (int m = 0; m < M; ++m) dct8<true>(u[m]);
        transpose_rc<M>(u, T, lane);
#pragma unroll
        for (int m = 0; m < M; ++m) dct8<true>(u[m]);
        float w;
        if constexpr (WIENER) {
            w = 1.f / fmaxf(wave_sum_f32(w2), 1e-12f);
        } else {
            w = 1.f / static_cast<float>(max(wave_sum_i32(kept), 1));
        }
        const int gb = lane >> 3, y = lane & 7;
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const int p = gb + 8 * m;
            if (p < k) {
                const int lx = rec[p].x - ax0, ly = rec[p].y - ay0;
                unsigned long long* row = num + (ly + y) * aw + lx;
#pragma unroll
                for (int x = 0; x < 8; ++x) lds_add(row + x, to_fixed(w * u[m][x]));
            }
        }
        if (lane == 0) {
            const int gj = g / ni, gi = g - gj * ni;
            jb.gweight[(j_begin + gj - jb.grid.row0) * jb.grid.nx + i_begin + gi] = to_fixed(w);
        }
    }
    __syncthreads();
    tile_partials(jb, tile, ax0, num);
}

// ---------------------------------------------------------------------------
// Block-8 group filter on MFMA (groups of 8 or 16), in bf16x3 or fp16 with
// FP32 accumulation. One wave filters a unit of 16 patches (two groups of 8
// or one of 16):
//   2-D DCT   four patches at a time as a 16 x 16 tile Q = [[P0, P1], [P2, P3]]
//             (pixel rows 8 pr + y, columns 8 pc + x): Z = bd(D) (Q bd(D^T))
//             with bd(M) = diag(M, M), two 16x16x16 products per tile; the
//             first product reads Q from memory in the A layout, its C-layout
//             output is the B operand of the second.
//   Haar      coefficient-major through LDS, then per 16 coefficients one
//             16x16x16 product with the exact +-1 butterfly matrix (block
//             diagonal for two groups of 8), scaled in FP32 to the
//             orthonormal transform.
//   shrink    hard threshold or Wiener in FP32 on the product's output.
//   inverse   the same steps reversed: Haar^T product, LDS back to the tile
//             layout, bd(D^T) (Z bd(D)).
// bf16x3 splits each operand v = hi + lo (both bf16, round to nearest even)
// and sums lo*hi + hi*lo + hi*hi; fp16 scales each group's pixels by a power
// of two to max |x| in [0.5, 1) (exact, undone in FP32) before rounding.
// Matching, weights and the fixed-point aggregation are those of the FP32
// kernels. MFMA 16x16x16 operand layouts (lane l, j = 0..3):
//   A[i][k]: i = l % 16, k = 4 (l / 16) + j;  B[k][n] and C[k][n]: n = l % 16,
//   k = 4 (l / 16) + j. So A(M) holds the entries of B(M^T).
// ---------------------------------------------------------------------------
typedef _Float16 h16x4 __attribute__((ext_vector_type(4)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

inline constexpr int kFormBf16x3 = 1;
inline constexpr int kFormF16 = 2;

template <int FORM>
struct Operand;
template <>
struct Operand<kFormBf16x3> {
    s16x4 hi, lo;
};
template <>
struct Operand<kFormF16> {
    h16x4 v;
};

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 b16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 h16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// Two values rounded to nearest even into one dword (v_cvt_pk_bf16_f32 /
// v_cvt_pk_f16_f32).
__device__ inline uint32_t pack_bf16(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, b16x2));
}
__device__ inline uint32_t pack_f16(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, h16x2));
}

// bf16x3 keeps the rounding residual as the lo part; the hi values are read
// back from the packed bits (low half << 16, high half masked).
template <int FORM>
__device__ inline Operand<FORM> operand(f32x4 v) {
    Operand<FORM> o;
    if constexpr (FORM == kFormBf16x3) {
        const uint32_t h0 = pack_bf16(v[0], v[1]), h1 = pack_bf16(v[2], v[3]);
        const uint32_t l0 = pack_bf16(v[0] - __uint_as_float(h0 << 16), v[1] - __uint_as_float(h0 & 0xffff0000u));
        const uint32_t l1 = pack_bf16(v[2] - __uint_as_float(h1 << 16), v[3] - __uint_as_float(h1 & 0xffff0000u));
        o.hi = __builtin_bit_cast(s16x4, u32x2{h0, h1});
        o.lo = __builtin_bit_cast(s16x4, u32x2{l0, l1});
    } else {
        o.v = __builtin_bit_cast(h16x4, u32x2{pack_f16(v[0], v[1]), pack_f16(v[2], v[3])});
    }
    return o;
}

// acc + a b (bf16x3 drops lo * lo only).
template <int FORM>
__device__ inline f32x4 mma(const Operand<FORM>& a, const Operand<FORM>& b, f32x4 acc) {
    if constexpr (FORM == kFormBf16x3) {
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.lo, b.hi, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.hi, b.lo, acc, 0, 0, 0);
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.hi, b.hi, acc, 0, 0, 0);
    } else {
        return __builtin_amdgcn_mfma_f32_16x16x16f16(a.v, b.v, acc, 0, 0, 0);
    }
}

// a b for an A operand that is exact in the operand type (a.lo = 0).
template <int FORM>
__device__ inline f32x4 mma_exact_a(const Operand<FORM>& a, const Operand<FORM>& b) {
    const f32x4 z = {0.f, 0.f, 0.f, 0.f};
    if constexpr (FORM == kFormBf16x3) {
        const f32x4 t = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.hi, b.lo, z, 0, 0, 0);
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.hi, b.hi, t, 0, 0, 0);
    } else {
        return __builtin_amdgcn_mfma_f32_16x16x16f16(a.v, b.v, z, 0, 0, 0);
    }
}

// Lane's A-operand entries of bd(D) (transposed = false) or bd(D^T).
__device__ inline f32x4 dct_bd_entries(bool transposed, int lane) {
    const float* d = c_dct + dct_offset<8>();
    const int i = lane & 15;
    f32x4 v;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int k = 4 * (lane >> 4) + j;
        v[j] = (i >> 3) != (k >> 3) ? 0.f : transposed ? d[(k & 7) * 8 + (i & 7)] : d[(i & 7) * 8 + (k & 7)];
    }
    return v;
}

// Entry (j, g) of the unnormalised in-place Haar butterfly matrix: row 0 sums
// all samples; row j > 0 with lowest set bit h is the sum over [j - h, j)
// minus the sum over [j, j + h) (the level-h difference haar_forward writes
// to position j).
__device__ inline float haar_entry(int j, int g) {
    if (j == 0) return 1.f;
    const int h = j & -j;
    return (g >= j - h && g < j) ? 1.f : (g >= j && g < j + h) ? -1.f : 0.f;
}

// Lane's A-operand entries of the unit's Haar matrix (block diagonal for two
// groups of 8) or of its transpose.
template <int G>
__device__ inline f32x4 haar_unit_entries(bool transposed, int lane) {
    const int i = lane & 15;
    f32x4 v;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int k = 4 * (lane >> 4) + j;
        const int r = transposed ? k : i, c = transposed ? i : k;
        v[j] = (G == 8 && (r >> 3) != (c >> 3)) ? 0.f : haar_entry(r & (G - 1), c & (G - 1));
    }
    return v;
}

// Coefficient-major transpose buffer of one wave: 64 coefficients x 16
// patches (dwords) at a pitch of 20, patch quads permuted per coefficient
// (XOR by 4 ((coef >> 4) & 3)), so that the b32 scatters / gathers and the
// b128 accesses below are at most 2-way bank conflicts.
inline constexpr int kMfmaBuf = 64 * 20;

__device__ inline int mfma_slot(int coef, int patch) { return coef * 20 + (patch ^ (4 * ((coef >> 4) & 3))); }

// Forward 3-D transform of a unit. x[q] = lane's pixels of tile q in the A
// layout, already scaled. Out: raw Haar products out[t][j] = coefficient
// 16 t + (lane & 15) of Haar row 4 (lane >> 4) + j (not yet normalised).
// fp16: dlo holds the DCT matrix's fp16 rounding residual, so the matrix
// enters as hi + lo and only the data operand is rounded.
template <int FORM>
__device__ inline void forward_unit(const f32x4 (&x)[4], const Operand<FORM>& dct, const Operand<FORM>& dlo,
                                    const Operand<FORM>& haar, float* T, int lane, f32x4 (&out)[4]) {
    const f32x4 z0 = {0.f, 0.f, 0.f, 0.f};
    const int h = (lane >> 4) & 1, kx = lane & 7, pc = (lane >> 3) & 1, pr = lane >> 5;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const Operand<FORM> xq = operand<FORM>(x[q]);
        f32x4 y = mma<FORM>(xq, dct, z0);  // Q bd(D^T)
        if constexpr (FORM == kFormF16) y = mma<FORM>(xq, dlo, y);
        const Operand<FORM> yq = operand<FORM>(y);
        f32x4 z = mma<FORM>(dct, yq, z0);  // bd(D) Q bd(D^T)
        if constexpr (FORM == kFormF16) z = mma<FORM>(dlo, yq, z);
#pragma unroll
        for (int j = 0; j < 4; ++j) T[mfma_slot(8 * (4 * h + j) + kx, 4 * q + 2 * pr + pc)] = z[j];
    }
    wave_sync();
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const f32x4 c = *reinterpret_cast<const f32x4*>(T + mfma_slot(16 * t + (lane & 15), 4 * (lane >> 4)));
        out[t] = mma_exact_a<FORM>(haar, operand<FORM>(c));
    }
    wave_sync();
}

// Inverse of a unit: c[t] = Haar-domain coefficients (layout of forward_unit's
// output, in the operand scale and before the butterfly factors are undone).
// Out: xo[q][j] = pixel (y = 4 ((lane >> 4) & 1) + j, x = lane & 7) of patch
// 4 q + 2 (lane >> 5) + ((lane >> 3) & 1).
template <int FORM>
__device__ inline void inverse_unit(const f32x4 (&c)[4], const Operand<FORM>& dctt, const Operand<FORM>& dtlo,
                                    const Operand<FORM>& haart, float* T, int lane, f32x4 (&xo)[4]) {
    const f32x4 z0 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const f32x4 u = mma_exact_a<FORM>(haart, operand<FORM>(c[t]));
        *reinterpret_cast<f32x4*>(T + mfma_slot(16 * t + (lane & 15), 4 * (lane >> 4))) = u;
    }
    wave_sync();
    const int ky = lane & 7, pr = (lane >> 3) & 1, pc = lane >> 5, kx0 = 4 * ((lane >> 4) & 1);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        f32x4 z;
#pragma unroll
        for (int j = 0; j < 4; ++j) z[j] = T[mfma_slot(8 * ky + kx0 + j, 4 * q + 2 * pr + pc)];
        const Operand<FORM> zq = operand<FORM>(z);
        f32x4 v = mma<FORM>(zq, dctt, z0);  // Z bd(D)
        if constexpr (FORM == kFormF16) v = mma<FORM>(zq, dtlo, v);
        const Operand<FORM> vq = operand<FORM>(v);
        xo[q] = mma<FORM>(dctt, vq, z0);  // bd(D^T) Z bd(D)
        if constexpr (FORM == kFormF16) xo[q] = mma<FORM>(dtlo, vq, xo[q]);
    }
    wave_sync();
}

// Lane's pixels of a unit in the A layout (zero for patches beyond a group's
// count or groups beyond the tile): tile q, patch 4 q + 2 ((lane >> 3) & 1) +
// (lane >> 5), row lane & 7, columns 4 ((lane >> 4) & 1) + j. Patch positions
// as x | y << 16.
template <int G>
__device__ inline void load_unit(const float* plane, int stride, const uint32_t* pos, const int* cnt, int ngroups,
                                 int unit, int lane, f32x4 (&x)[4]) {
    const int pr = (lane >> 3) & 1, pc = lane >> 5, y = lane & 7, x0 = 4 * ((lane >> 4) & 1);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int p = 4 * q + 2 * pr + pc;
        const int g = G == 8 ? 2 * unit + (p >> 3) : unit, idx = p & (G - 1);
        if (g < ngroups && idx < cnt[g]) {
            const uint32_t xy = pos[g * G + idx];
            const float* src = plane + static_cast<size_t>((xy >> 16) + y) * stride + (xy & 0xffffu) + x0;
#pragma unroll
            for (int j = 0; j < 4; ++j) x[q][j] = src[j];
        } else {
            x[q] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
    }
}

// Power-of-two scale taking max |x| of a group to [0.5, 1); 1 for a zero or
// non-finite maximum.
__device__ inline float group_scale(float m) {
    if (!(m > 0.f) || !(m <= 3.4e38f)) return 1.f;
    return ldexpf(1.f, -__builtin_amdgcn_frexp_expf(m));
}

// The fp16 rounding residual of four values (the lo half of a split operand).
__device__ inline f32x4 f16_residual(f32x4 v) {
    const h16x4 h = __builtin_convertvector(v, h16x4);
    return v - __builtin_convertvector(h, f32x4);
}

// sum over g < k of the unnormalised Haar butterfly entry (j, g): the Haar
// coefficient of a constant over the k real patches of a group, before the
// orthonormal factor (row 0 sums k; a difference row j with lowest set bit h
// the part of [j - h, j) below k minus the part of [j, j + h) below k).
__device__ inline int haar_row_count(int j, int k) {
    if (j == 0) return k;
    const int h = j & -j;
    const int plus = min(max(min(j, k) - (j - h), 0), h);
    const int minus = min(max(min(j + h, k) - j, 0), h);
    return plus - minus;
}

// MFMA forms: each group's pixels are centred on their mean mu over the
// group's real patches (padding patches stay zero) and scaled by a power of
// two taking max |x - mu| to [0.5, 1) (fp16's range; exact, undone in FP32),
// so the low-precision operands carry the texture and the noise, not the
// group's level; the level's contribution (2-D DC coefficients only) is added
// back in FP32. m = max |x - mu| (unscaled). Groups: q >> 1 for two groups
// of 8, one group of 16.
template <int G>
__device__ inline void unit_centre(f32x4 (&x)[4], const int* cnt, int ngroups, int unit, int lane, float (&mu)[2],
                                   float (&s)[2], float (&m)[2]) {
    const int pr = (lane >> 3) & 1, pc = lane >> 5;
    bool real[4];
    float sum[2] = {0.f, 0.f};
    int k[2] = {0, 0};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int p = 4 * q + 2 * pr + pc;
        const int g = G == 8 ? 2 * unit + (p >> 3) : unit, idx = p & (G - 1);
        real[q] = g < ngroups && idx < cnt[g];
        if (real[q]) sum[G == 8 ? q >> 1 : 0] += (x[q][0] + x[q][1]) + (x[q][2] + x[q][3]);
    }
#pragma unroll
    for (int gs = 0; gs < (G == 8 ? 2 : 1); ++gs) {
        const int g = G == 8 ? 2 * unit + gs : unit;
        k[gs] = g < ngroups ? cnt[g] : 0;
        const float tot = wave_sum_dpp(sum[gs]);
        mu[gs] = k[gs] > 0 ? tot / static_cast<float>(64 * k[gs]) : 0.f;
    }
    if constexpr (G == 16) mu[1] = mu[0];
    float mx[2] = {0.f, 0.f};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int gs = G == 8 ? q >> 1 : 0;
        if (real[q]) x[q] -= mu[gs];
#pragma unroll
        for (int j = 0; j < 4; ++j) mx[gs] = fmaxf(mx[gs], fabsf(x[q][j]));
    }
    m[0] = wave_max_dpp(mx[0]);
    m[1] = G == 8 ? wave_max_dpp(mx[1]) : m[0];
    s[0] = group_scale(m[0]);
    s[1] = group_scale(m[1]);
#pragma unroll
    for (int q = 0; q < 4; ++q) x[q] *= s[G == 8 ? q >> 1 : 0];
}

// FP32 value, in data units, of coefficient ci = 8 u + v (u along y) of Haar
// row jr of unit group gs, from the lane-held centred and scaled pixels xc
// (A layout) and the DCT table d (LDS): sum over the group's patches p of
// haar(jr, p) D[u][y] D[v][x] xc_p[y][x], times the Haar factor and 1 / s,
// plus the centre's share corr. Every argument but xc, d and lane is
// wave-uniform; every lane returns the value.
template <int G>
__device__ inline float exact_coef(const f32x4 (&xc)[4], const float* d, int ci, int jr, int gs, float inv_s,
                                   float corr, int lane) {
    const int u = ci >> 3, v = ci & 7;
    const float dy = d[u * 8 + (lane & 7)];
    const f32x4 dx = *reinterpret_cast<const f32x4*>(d + v * 8 + 4 * ((lane >> 4) & 1));
    float acc = 0.f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (G == 8 && (q >> 1) != gs) continue;
        const int p = 4 * q + 2 * ((lane >> 3) & 1) + (lane >> 5);
        const float r = fmaf(dx[3], xc[q][3], fmaf(dx[2], xc[q][2], fmaf(dx[1], xc[q][1], dx[0] * xc[q][0])));
        acc = fmaf(haar_entry(jr, p & (G - 1)) * dy, r, acc);
    }
    return fmaf(wave_sum_dpp(acc) * haar_scale(G, jr), inv_s, corr);
}

// A hard-threshold decision of an MFMA form is re-made in FP32 when the
// coefficient lies within band * max |x - mu| of the threshold. fp16: the
// measured error of a coefficient stays below 5.2 * 2^-11 max |x - mu|
// over two million coefficients of sigma-20 groups (the data rounding, 2^-12
// relative, carried through the orthonormal transforms; the DCT matrix
// enters split hi + lo), band 6 * 2^-11. bf16x3: each operand is hi + lo
// with a residual near 2^-17 relative, band 8 * 2^-16.
template <int FORM>
__device__ constexpr float redecide_band() {
    return FORM == kFormF16 ? 6.f * (1.f / 2048.f) : 8.f * (1.f / 65536.f);
}

template <int G, bool WIENER, int FORM>
__global__ __launch_bounds__(256) void filter_tile_mfma_kernel(FilterTable tab) {
    extern __shared__ __align__(16) unsigned long long acc[];
    constexpr int B = 8;
    static_assert(G == 8 || G == 16, "units of 16 patches");
    const int lb = xcd_block();
    const FilterJob& jb = tab.job[find_job(tab, lb)];
    const TileGeom& tg = jb.tiles;
    const int tile = lb - jb.wg_begin;  // partials index
    const int tx = (tile + jb.tile0) % tg.ntx, ty = (tile + jb.tile0) / tg.ntx;
    const int aw = tg.aw, ah = tg.ah;
    const int ax0 = tg.ax0(tx, jb.width, B, jb.grid.step, jb.range);
    const int ay0 = tg.ay0(ty, jb.height, B, jb.grid.step, jb.range);
    const int i_begin = tx * tg.tw, j_begin = ty * tg.th;
    const int ni = min(tg.tw, jb.grid.nx - i_begin), nj = min(tg.th, jb.grid.ny - j_begin);
    const int ngroups = ni * nj;
    const int nunits = G == 8 ? (ngroups + 1) / 2 : ngroups;
    unsigned long long* num = acc;
    // Patch positions (x | y << 16) and counts of the tile's groups, then the
    // transpose buffers, 16-byte aligned.
    uint32_t* recs = reinterpret_cast<uint32_t*>(acc + ((aw * ah + 1) & ~1));
    int* cnt = reinterpret_cast<int*>(recs + ((tg.tw * tg.th * G + 3) & ~3));
    float* scratch = reinterpret_cast<float*>(cnt + ((tg.tw * tg.th + 3) & ~3));
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float* T = scratch + wave * kMfmaBuf;
    for (int e = threadIdx.x; e < aw * ah; e += 256) acc[e] = 0ull;
    for (int e = threadIdx.x; e < ngroups * G; e += 256) {
        const int g = e / G, j = e - g * G;
        const int gj = g / ni, gi = g - gj * ni;
        const MatchRec r =
            jb.matches[static_cast<size_t>((j_begin + gj - jb.grid.row0) * jb.grid.nx + i_begin + gi) * G + j];
        recs[e] = static_cast<uint32_t>(r.x) | (static_cast<uint32_t>(r.y) << 16);
    }
    for (int g = threadIdx.x; g < ngroups; g += 256) {
        const int gj = g / ni, gi = g - gj * ni;
        cnt[g] = jb.counts[(j_begin + gj - jb.grid.row0) * jb.grid.nx + i_begin + gi];
    }
    __syncthreads();

    const f32x4 de = dct_bd_entries(false, lane), dte = dct_bd_entries(true, lane);
    const Operand<FORM> dct = operand<FORM>(de);
    const Operand<FORM> dctt = operand<FORM>(dte);
    const Operand<FORM> dlo = operand<FORM>(FORM == kFormF16 ? f16_residual(de) : de);
    const Operand<FORM> dtlo = operand<FORM>(FORM == kFormF16 ? f16_residual(dte) : dte);
    const Operand<FORM> haar = operand<FORM>(haar_unit_entries<G>(false, lane));
    const Operand<FORM> haart = operand<FORM>(haar_unit_entries<G>(true, lane));
    // The DCT table for the FP32 re-decisions of the hard threshold.
    float* dtab = scratch + 4 * kMfmaBuf;
    if (!WIENER && threadIdx.x < 64) dtab[threadIdx.x] = c_dct[dct_offset<8>() + threadIdx.x];
    const float thr = 2.7f * jb.sigma;  // include/nss/params.hpp:46 kBmHardLambda
    const float sig2 = jb.sigma * jb.sigma;
    // Haar rows of this lane in the coefficient layout: 4 (lane >> 4) + j; for
    // two groups of 8 the group is lane >> 5.
    const int hgroup = G == 8 ? lane >> 5 : 0;
    float hs[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) hs[j] = haar_scale(G, (4 * (lane >> 4) + j) & (G - 1));
    if (!WIENER) __syncthreads();

    f32x4 x[4], xr[4];
    int u = wave;
    if (u < nunits) {
        load_unit<G>(jb.src, jb.stride, recs, cnt, ngroups, u, lane, x);
        if constexpr (WIENER) load_unit<G>(jb.wref, jb.wref_stride, recs, cnt, ngroups, u, lane, xr);
    }
    for (; u < nunits; u += 4) {
        float s[2], sr[2], mu[2], mur[2], mx[2], mxr[2];
        unit_centre<G>(x, cnt, ngroups, u, lane, mu, s, mx);
        f32x4 c[4], r[4];
        forward_unit<FORM>(x, dct, dlo, haar, T, lane, c);
        if constexpr (WIENER) {
            unit_centre<G>(xr, cnt, ngroups, u, lane, mur, sr, mxr);
            forward_unit<FORM>(xr, dct, dlo, haar, T, lane, r);
        }
        // Hard threshold: the centred, scaled pixels, kept for the FP32
        // re-decisions after the next unit's loads replace x.
        f32x4 xc[4];
        if constexpr (!WIENER) {
#pragma unroll
            for (int q = 0; q < 4; ++q) xc[q] = x[q];
        }
        const int un = u + 4;
        if (un < nunits) {
            load_unit<G>(jb.src, jb.stride, recs, cnt, ngroups, un, lane, x);
            if constexpr (WIENER) load_unit<G>(jb.wref, jb.wref_stride, recs, cnt, ngroups, un, lane, xr);
        }
        // The centres' share of each 2-D DC coefficient (coefficient 0 of
        // every Haar row; zero on the other lanes).
        float corr[4] = {0.f, 0.f, 0.f, 0.f}, corr_r[4] = {0.f, 0.f, 0.f, 0.f};
        {
            if ((lane & 15) == 0) {
                const int g = G == 8 ? 2 * u + hgroup : u;
                const int k = g < ngroups ? cnt[g] : 0;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float e = hs[j] * static_cast<float>(haar_row_count((4 * (lane >> 4) + j) & (G - 1), k)) * 8.f;
                    corr[j] = e * mu[hgroup];
                    corr_r[j] = e * mur[hgroup];
                }
            }
        }
        // Shrink in FP32 on the orthonormal coefficients of the unscaled data.
        const float inv_s = 1.f / s[hgroup];
        float vv[4][4];
        bool band[4][4];
        float w2 = 0.f;
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool dc = t == 0 && (lane & 15) == 0 && ((4 * (lane >> 4) + j) & (G - 1)) == 0;
                const float cj = t == 0 ? corr[j] : 0.f;
                float v = fmaf(c[t][j] * hs[j], inv_s, cj);
                band[t][j] = false;
                if constexpr (WIENER) {
                    const float rv = fmaf(r[t][j] * hs[j], 1.f / sr[hgroup], t == 0 ? corr_r[j] : 0.f);
                    float w = 1.f;
                    if (!dc) {
                        const float r2 = rv * rv;
                        const float den = r2 + sig2;
                        w = den > 0.f ? r2 / den : 0.f;
                    }
                    v *= w;
                    w2 = fmaf(w, w, w2);
                } else {
                    band[t][j] = !dc && fabsf(fabsf(v) - thr) <= redecide_band<FORM>() * mx[hgroup];
                }
                vv[t][j] = v;
            }
        bool any_band = false;
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) any_band |= band[t][j];
        if (!WIENER && __ballot(any_band)) {
            // Coefficients whose MFMA value cannot decide the threshold: their
            // FP32 value, one coefficient at a time by the whole wave.
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    unsigned long long pend = __ballot(band[t][j]);
                    while (pend) {
                        const int L = __builtin_ctzll(pend);
                        pend &= pend - 1;
                        const int gs = G == 8 ? L >> 5 : 0;
                        const float cl = t == 0 ? __shfl(corr[j], L, 64) : 0.f;
                        const float e = exact_coef<G>(xc, dtab, 16 * t + (L & 15), (4 * (L >> 4) + j) & (G - 1), gs,
                                                      1.f / s[gs], cl, lane);
                        if (lane == L) vv[t][j] = e;
                    }
                }
        }
        int kept = 0;
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float v = vv[t][j];
                if constexpr (!WIENER) {
                    const bool dc = t == 0 && (lane & 15) == 0 && ((4 * (lane >> 4) + j) & (G - 1)) == 0;
                    if (!dc && fabsf(v) < thr)
                        v = 0.f;
                    else
                        ++kept;
                }
                // back to the operand scale without the centre's share,
                // butterfly factor applied again (the inverse orthonormal
                // Haar is H^T diag(factors))
                c[t][j] = (v - (t == 0 ? corr[j] : 0.f)) * hs[j] * s[hgroup];
            }
        float wg[2];
        if constexpr (WIENER) {
            float sum = w2;
#pragma unroll
            for (int off = 16; off >= 1; off >>= 1) sum += __shfl_xor(sum, off, 64);
            if constexpr (G == 16) sum += __shfl_xor(sum, 32, 64);
            wg[0] = 1.f / fmaxf(__shfl(sum, 0, 64), 1e-12f);
            wg[1] = 1.f / fmaxf(__shfl(sum, 32, 64), 1e-12f);
        } else {
            int sum = kept;
#pragma unroll
            for (int off = 16; off >= 1; off >>= 1) sum += __shfl_xor(sum, off, 64);
            if constexpr (G == 16) sum += __shfl_xor(sum, 32, 64);
            wg[0] = 1.f / static_cast<float>(max(__shfl(sum, 0, 64), 1));
            wg[1] = 1.f / static_cast<float>(max(__shfl(sum, 32, 64), 1));
        }
        f32x4 xo[4];
        inverse_unit<FORM>(c, dctt, dtlo, haart, T, lane, xo);
        const int pr = lane >> 5, pc = (lane >> 3) & 1, px = lane & 7, y0 = 4 * ((lane >> 4) & 1);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int p = 4 * q + 2 * pr + pc;
            const int g = G == 8 ? 2 * u + (p >> 3) : u, idx = p & (G - 1);
            const int gsel = G == 8 ? (q >> 1) : 0;
            if (g < ngroups && idx < cnt[g]) {
                const uint32_t xy = recs[g * G + idx];
                const int lx = static_cast<int>(xy & 0xffffu) - ax0, ly = static_cast<int>(xy >> 16) - ay0;
                const float w = wg[gsel];
                const float wq = w / s[gsel];
                const float wm = w * mu[gsel];  // the centre, added back
                unsigned long long* cell = num + (ly + y0) * aw + lx + px;
#pragma unroll
                for (int j = 0; j < 4; ++j) lds_add(cell + j * aw, to_fixed(fmaf(wq, xo[q][j], wm)));
            }
        }
        // the unit's group weights (two groups of 8 or one of 16)
        if (lane < (G == 8 ? 2 : 1)) {
            const int g = G == 8 ? 2 * u + lane : u;
            if (g < ngroups) {
                const int gj = g / ni, gi = g - gj * ni;
                jb.gweight[(j_begin + gj - jb.grid.row0) * jb.grid.nx + i_begin + gi] = to_fixed(lane ? wg[1] : wg[0]);
            }
        }
    }
    __syncthreads();
    tile_partials(jb, tile, ax0, num);
}

// Atomic path: one wave per group, frame-sized fixed-point sums (den at the
// patch corner).
template <int B, bool WIENER>
__device__ inline void atomic_group(const FilterJob& jb, int g, float* cube, int lane) {
    constexpr int A = B * B;
    constexpr int PS = A + 1;
    const int G = jb.group;
    const int k = jb.counts[g];
    const MatchRec* rec = jb.matches + static_cast<size_t>(g) * G;
    float* rc = cube + G * PS;
    const float weight = filter_group_wave<B, WIENER>(jb, rec, k, cube, rc, lane);
    const unsigned long long wfix = to_fixed(weight);
    for (int j = 0; j < k; ++j) {
        const int mx = rec[j].x, my = rec[j].y;
        for (int e = lane; e < A; e += 64) {
            const int py = e / B, px = e - (e / B) * B;
            atomicAdd(jb.num + static_cast<size_t>(my + py) * jb.width + mx + px, to_fixed(weight * cube[j * PS + e]));
        }
        if (lane == 0) atomicAdd(jb.den + static_cast<size_t>(my) * jb.width + mx, wfix);
    }
}

template <int B, bool WIENER>
__global__ __launch_bounds__(64) void filter_atomic_kernel(FilterTable tab) {
    extern __shared__ __align__(16) float lds_cube[];
    const int lane = threadIdx.x;
    if (tab.cube) {
        // Cubes beyond the LDS: workgroup b takes groups b, b + gridDim.x, ...
        // with its own cube in global memory (the same arithmetic; a wave's
        // global stores and loads stay in order on its CU).
        const FilterJob& last = tab.job[tab.n - 1];
        const int total = last.wg_begin + last.nref;
        float* cube = tab.cube + static_cast<size_t>(blockIdx.x) * tab.cube_floats;
        for (int gi = blockIdx.x; gi < total; gi += gridDim.x) {
            const FilterJob& jb = tab.job[find_job(tab, gi)];
            atomic_group<B, WIENER>(jb, gi - jb.wg_begin, cube, lane);
            wave_sync();
        }
        return;
    }
    const int lb = xcd_block();
    const FilterJob& jb = tab.job[find_job(tab, lb)];
    const int g = lb - jb.wg_begin;
    if (g >= jb.nref) return;
    atomic_group<B, WIENER>(jb, g, lds_cube, lane);
}

// ---------------------------------------------------------------------------
// Temporal: predictive search (src/cpu/bm/matcher.hpp:319-373) and the
// group filter with per-slot sources and accumulators.
// ---------------------------------------------------------------------------
__device__ inline unsigned long long pack_pos(int t, int y, int x) {
    return (static_cast<unsigned long long>(t) << 26) | (static_cast<unsigned long long>(y) << 13) |
           static_cast<unsigned long long>(x);
}

// One wave per reference. The spatial matches (slot radius) seed the search:
// per direction, layer dt scores the union of (2 ps_range + 1)^2 windows
// around the previous layer's winners (every union position once), keeps its
// min(ps_num, G) smallest (dist, y, x), and those feed the global G-1
// smallest (dist, t, y, x) with the spatial matches. Layer 1's centres are
// the first ps_num spatial matches. When the layer's windows fit the staging
// buffer, their pixels are copied to LDS once and each lane scores three
// horizontally adjacent candidates from one register row per block row
// (same sums, same order as scoring from the frame).
// LDS: reference patch | layer chunk keys | best, next best | centres | merged
// | staged windows.
template <int B>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) void predictive_kernel(PredTable tab) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    // XCD-contiguous references: neighbours share their frames' rows in L2.
    const int lb = xcd_block();
    const PredJob& jb = tab.job[find_job(tab, lb)];
    const int ref = lb - jb.wg_begin;
    if (ref >= jb.nref) return;
    const int G = jb.group;
    const Slots& sl = jb.slots;
    const int t0 = sl.radius;
    MatchRec* rec = jb.recs + static_cast<size_t>(ref) * G;
    const int n = jb.counts[ref];
    const int cx = rec[0].x, cy = rec[0].y;
    float* refp = reinterpret_cast<float*>(smem);
    unsigned long long* cand = reinterpret_cast<unsigned long long*>(refp + ((B * B + 3) & ~3));
    unsigned long long* best = reinterpret_cast<unsigned long long*>(cand + tab.chunk);
    unsigned long long* nbest = best + kMaxGroup;
    int* centers = reinterpret_cast<int*>(nbest + kMaxGroup);
    unsigned long long* merged = reinterpret_cast<unsigned long long*>(centers + kMaxGroup);
    float* region = reinterpret_cast<float*>(smem + tab.region_offset);
    {
        const float* p0 = jb.match.frame(sl.frame(t0));
        for (int e = lane; e < B * B; e += 64) refp[e] = p0[static_cast<size_t>(cy + e / B) * jb.stride + cx + e % B];
    }
    for (int i = lane; i < n - 1; i += 64)
        merged[i] = (static_cast<unsigned long long>(rec[1 + i].dist) << 32) | pack_pos(t0, rec[1 + i].y, rec[1 + i].x);
    int nm = n - 1;
    const int vb = sl.valid_begin(), ve = sl.valid_end();
    const int kl = min(jb.ps_num, G);
    const int seeds = min(n, jb.ps_num);
    const int ps = jb.ps_range;
    const int maxx = jb.width - B, maxy = jb.height - B;
    __syncthreads();
    for (int sign = -1; sign <= 1; sign += 2) {
        if (lane < seeds) centers[lane] = (rec[lane].y << 13) | rec[lane].x;
        int nc = seeds;
        __syncthreads();
        for (int dt = 1; dt <= sl.radius; ++dt) {
            const int t = t0 + sign * dt;
            if (t < vb || t >= ve) break;
            const float* pt = jb.match.frame(sl.frame(t));
            // The union of the centres' (2 ps + 1)^2 windows (clamped to the
            // plane), enumerated window by window: a cell belongs to the first
            // window that contains it, so every union position is scored once.
            const int W = 2 * ps + 1, WW = W * W;
            const int cells = nc * WW;
            auto ssd = [&](const float* base, int stride) {
                float acc = 0.f;
                for (int py = 0; py < B; ++py) {
                    const float* row = base + py * stride;
#pragma unroll
                    for (int px = 0; px < B; ++px) {
                        const float d = refp[py * B + px] - row[px];
                        acc = fmaf(d, d, acc);
                    }
                }
                return acc;
            };
            auto member_of = [&](int w, int y, int x) {
                bool m = y >= 0 && y <= maxy && x >= 0 && x <= maxx;
                for (int q = 0; q < w && m; ++q) {
                    const int cq = centers[q];
                    m = !(abs(y - (cq >> 13)) <= ps && abs(x - (cq & 8191)) <= ps);
                }
                return m;
            };
            const int R = W + B - 1, P = pred_region_pitch(W, B), SW = (W + 2) / 3;
            const bool staged = cells <= tab.chunk && nc * R * P <= tab.region;
            if (staged) {
                // Window w's pixels (rows y_w - ps .., columns x_w - ps ..) at
                // region[w R P + j P + i]. Columns from R on and pixels off the
                // plane are read only for discarded candidates (positions off
                // the plane are not members; a member's patch lies in rows and
                // columns below R), so the loads leave them as they are.
                // Lanes cover 64 / P rows of P columns per step (P <= 64), or
                // one row in steps of 64 columns.
                const int rows_per = P <= 64 ? 64 / P : 1;
                const int lr = P <= 64 ? lane / P : 0, li = P <= 64 ? lane - lr * P : lane;
                for (int w = 0; w < nc; ++w) {
                    const int cw = centers[w];
                    const int y0 = (cw >> 13) - ps, x0 = (cw & 8191) - ps;
                    float* dst = region + w * (R * P);
                    if (lr >= rows_per) continue;
                    if (P <= 64 && R <= 8 * rows_per) {
                        // one column per lane, up to 8 rows, each load
                        // straight to LDS (global_load_lds: lane l's word at
                        // the wave-uniform address dst + q rows_per P + 4 l,
                        // which is row j = lr + q rows_per, column li);
                        // the barrier below waits for them
#pragma unroll
                        for (int q = 0; q < 8; ++q) {
                            const int j = lr + q * rows_per, y = y0 + j, x = x0 + li;
                            if (j < R && li < R && y >= 0 && y < jb.height && x >= 0 && x < jb.width)
                                __builtin_amdgcn_global_load_lds(
                                    (const __attribute__((address_space(1))) void*)(pt + static_cast<size_t>(y) * jb.stride + x),
                                    (__attribute__((address_space(3))) void*)(dst + q * rows_per * P), 4, 0, 0);
                        }
                        continue;
                    }
                    // the same loads for any pitch: lane l's word at
                    // dst + j P + i = (dst + j P + i - l) + 4 l, the base
                    // alike for every lane of a step
                    for (int j = lr; j < R; j += rows_per)
                        for (int i = li; i < P; i += (P <= 64 ? P : 64)) {
                            const int y = y0 + j, x = x0 + i;
                            if (i < R && y >= 0 && y < jb.height && x >= 0 && x < jb.width)
                                __builtin_amdgcn_global_load_lds(
                                    (const __attribute__((address_space(1))) void*)(pt + static_cast<size_t>(y) * jb.stride + x),
                                    (__attribute__((address_space(3))) void*)(dst + j * P + i - lane), 4, 0, 0);
                        }
                }
                __syncthreads();
                const int strips = nc * W * SW;
                for (int sidx = lane; sidx < strips; sidx += 64) {
                    const int w = sidx / (W * SW), o = sidx - w * (W * SW);
                    const int dy = o / SW, dx0 = 3 * (o - dy * SW);
                    const float* rb = region + w * (R * P) + dy * P + dx0;
                    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
#pragma unroll 1
                    for (int py = 0; py < B; ++py) {
                        float v[B + 2];
#pragma unroll
                        for (int q = 0; q < B + 2; ++q) v[q] = rb[py * P + q];
#pragma unroll
                        for (int px = 0; px < B; ++px) {
                            const float rv = refp[py * B + px];
                            const float d0 = rv - v[px], d1 = rv - v[px + 1], d2 = rv - v[px + 2];
                            a0 = fmaf(d0, d0, a0);
                            a1 = fmaf(d1, d1, a1);
                            a2 = fmaf(d2, d2, a2);
                        }
                    }
                    const int cw = centers[w];
                    const int y = (cw >> 13) - ps + dy;
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        const int dx = dx0 + q;
                        if (dx < W) {
                            const int x = (cw & 8191) - ps + dx;
                            const float a = q == 0 ? a0 : q == 1 ? a1 : a2;
                            cand[w * WW + dy * W + dx] =
                                member_of(w, y, x)
                                    ? (static_cast<unsigned long long>(dist_key(a)) << 32) | pack_pos(0, y, x)
                                    : ~0ull;
                        }
                    }
                }
            }
            int nb = 0;
            for (int base = 0; base < cells; base += tab.chunk) {
                const int cnt = min(tab.chunk, cells - base);
                for (int c = lane; c < cnt && !staged; c += 64) {
                    const int cell = base + c;
                    const int w = cell / WW, o = cell - w * WW;
                    const int dy = o / W;
                    const int cw = centers[w];
                    const int y = (cw >> 13) - ps + dy, x = (cw & 8191) - ps + (o - dy * W);
                    unsigned long long key = ~0ull;
                    if (member_of(w, y, x))
                        key = (static_cast<unsigned long long>(dist_key(ssd(pt + static_cast<size_t>(y) * jb.stride + x,
                                                                            jb.stride)))
                               << 32) |
                              pack_pos(0, y, x);
                    cand[c] = key;
                }
                __syncthreads();
                // The chunk's candidates and the running best: the kl smallest
                // keys (unique: one per position), each taken by its owner lane.
                int nn = 0;
                for (int k = 0; k < kl; ++k) {
                    unsigned long long m = ~0ull, mb = ~0ull;
                    int mi = 0, bi = 0;
                    for (int c = lane; c < cnt; c += 64)
                        if (cand[c] < m) {
                            m = cand[c];
                            mi = c;
                        }
                    for (int i = lane; i < nb; i += 64)
                        if (best[i] < mb) {
                            mb = best[i];
                            bi = i;
                        }
                    const unsigned long long mine = m < mb ? m : mb;
                    const unsigned long long win = wave_min_u64(mine);
                    if (win == ~0ull) break;
                    if (mine == win) {
                        if (m < mb)
                            cand[mi] = ~0ull;
                        else
                            best[bi] = ~0ull;
                    }
                    if (lane == 0) nbest[nn] = win;
                    ++nn;
                    __syncthreads();
                }
                for (int i = lane; i < nn; i += 64) best[i] = nbest[i];
                nb = nn;
                __syncthreads();
            }
            // Layer winners (ascending): into the merged list, and the next centres.
            for (int i = lane; i < nb; i += 64) {
                const unsigned long long b = best[i];
                merged[nm + i] = (b & 0xffffffff00000000ull) | pack_pos(t, static_cast<int>((b >> 13) & 8191),
                                                                       static_cast<int>(b & 8191));
                centers[i] = static_cast<int>(b & 0x3ffffff);
            }
            nm += nb;
            nc = nb;
            __syncthreads();
            if (nc == 0) break;
        }
    }
    // The G-1 smallest merged keys in ascending order: keys are unique (one
    // per (t, y, x)), so each key's rank is the count of smaller keys.
    const int K = min(G - 1, nm);
    for (int i = lane; i < nm; i += 64) {
        const unsigned long long m = merged[i];
        int rank = 0;
        for (int j = 0; j < nm; ++j) rank += merged[j] < m ? 1 : 0;
        if (rank < K)
            rec[1 + rank] = MatchRec{static_cast<int>(m & 8191), static_cast<int>((m >> 13) & 8191),
                                     static_cast<int>((m >> 26) & 63), static_cast<uint32_t>(m >> 32)};
    }
    if (lane == 0) jb.counts[ref] = 1 + K;
}

// One wave per group: patches from their slots' frames, sums into the
// accumulators of those frames.
template <int B, bool WIENER>
__device__ inline void temporal_group(const TFilterJob& jb, int g, float* cube, int lane) {
    constexpr int A = B * B;
    constexpr int PS = A + 1;
    const int G = jb.group;
    const int k = jb.counts[g];
    const MatchRec* rec = jb.matches + static_cast<size_t>(g) * G;
    float* rc = cube + G * PS;
    gather_group_wave<B, WIENER>(
        rec, k, G, jb.stride, jb.wref_stride, [&](int t) { return jb.src.frame(jb.slots.frame(t)); },
        [&](int t) { return jb.wref.frame(jb.slots.frame(t)); }, cube, rc, lane);
    const float weight = transform_group_wave<B, WIENER>(G, jb.sigma, cube, rc, lane);
    const unsigned long long wfix = to_fixed(weight);
    const size_t plane = static_cast<size_t>(jb.width) * jb.height;
    for (int j = 0; j < k; ++j) {
        const int mx = rec[j].x, my = rec[j].y;
        const size_t slot = static_cast<size_t>(jb.slots.frame(rec[j].t) % jb.acc_ring) * plane;
        unsigned long long* num = jb.num + slot;
        for (int e = lane; e < A; e += 64) {
            const int py = e / B, px = e - (e / B) * B;
            atomicAdd(num + static_cast<size_t>(my + py) * jb.width + mx + px, to_fixed(weight * cube[j * PS + e]));
        }
        if (lane == 0) atomicAdd(jb.den + slot + static_cast<size_t>(my) * jb.width + mx, wfix);
    }
}

template <int B, bool WIENER>
__global__ __launch_bounds__(64) void filter_temporal_kernel(TFilterTable tab) {
    extern __shared__ __align__(16) float lds_cube[];
    const int lane = threadIdx.x;
    if (tab.cube) {
        // Cubes beyond the LDS, as in filter_atomic_kernel.
        const TFilterJob& last = tab.job[tab.n - 1];
        const int total = last.wg_begin + last.nref;
        float* cube = tab.cube + static_cast<size_t>(blockIdx.x) * tab.cube_floats;
        for (int gi = blockIdx.x; gi < total; gi += gridDim.x) {
            const TFilterJob& jb = tab.job[find_job(tab, gi)];
            temporal_group<B, WIENER>(jb, gi - jb.wg_begin, cube, lane);
            wave_sync();
        }
        return;
    }
    const int lb = xcd_block();
    const TFilterJob& jb = tab.job[find_job(tab, lb)];
    const int g = lb - jb.wg_begin;
    if (g >= jb.nref) return;
    temporal_group<B, WIENER>(jb, g, lds_cube, lane);
}

// Block 8, G in {8, 16, 32}: four groups per workgroup, a wave per group with
// the cube in registers (the spatial register kernel's transforms); patches
// from their slot's frame, device-scope int64 atomics into the sums of that
// frame (den at the patch corner).
template <int G, bool WIENER>
__global__ __launch_bounds__(256) void filter_temporal_b8_kernel(TFilterTable tab) {
    extern __shared__ __align__(16) float tbuf[];
    constexpr int M = G / 8;
    const int lb = xcd_block();
    const TFilterJob& jb = tab.job[find_job(tab, lb)];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = (lb - jb.wg_begin) * 4 + wave;
    if (g >= jb.nref) return;  // whole wave; no workgroup barrier below
    float* T = tbuf + wave * (G * kTs);
    const int k = jb.counts[g];
    const MatchRec* rec = jb.matches + static_cast<size_t>(g) * G;
    const int gb = lane >> 3, y = lane & 7;
    float v[M][8], vr[M][8];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int p = gb + 8 * m;
        if (p < k) {
            const MatchRec r = rec[p];
            const int f = jb.slots.frame(r.t);
            const size_t o = static_cast<size_t>(r.y + y) * jb.stride + r.x;
            const float* sp = jb.src.frame(f) + o;
#pragma unroll
            for (int x = 0; x < 8; ++x) v[m][x] = sp[x];
            if constexpr (WIENER) {
                const float* wp = jb.wref.frame(f) + static_cast<size_t>(r.y + y) * jb.wref_stride + r.x;
#pragma unroll
                for (int x = 0; x < 8; ++x) vr[m][x] = wp[x];
            }
        } else {
#pragma unroll
            for (int x = 0; x < 8; ++x) v[m][x] = vr[m][x] = 0.f;
        }
    }
    float c[G], r[G];
    forward_b8<G>(v, c, T, lane);
    if constexpr (WIENER) forward_b8<G>(vr, r, T, lane);
    const float thr = 2.7f * jb.sigma;  // include/nss/params.hpp:46 kBmHardLambda
    const float sig2 = jb.sigma * jb.sigma;
    int kept = 0;
    float w2 = 0.f;
#pragma unroll
    for (int j = 0; j < G; ++j) {
        const bool dc = (lane == 0 && j == 0);
        if constexpr (WIENER) {
            float w = 1.f;
            if (!dc) {
                const float r2 = r[j] * r[j];
                const float den = r2 + sig2;
                w = den > 0.f ? r2 / den : 0.f;
            }
            c[j] *= w;
            w2 = fmaf(w, w, w2);
        } else {
            if (!dc && fabsf(c[j]) < thr)
                c[j] = 0.f;
            else
                ++kept;
        }
    }
    haar_inverse_reg<G>(c);
    float u[M][8];
    group_to_columns<G>(c, u, T, lane);
#pragma unroll
    for (int m = 0; m < M; ++m) dct8<true>(u[m]);
    transpose_rc<M>(u, T, lane);
#pragma unroll
    for (int m = 0; m < M; ++m) dct8<true>(u[m]);
    float w;
    if constexpr (WIENER) {
        w = 1.f / fmaxf(wave_sum_f32(w2), 1e-12f);
    } else {
        w = 1.f / static_cast<float>(max(wave_sum_i32(kept), 1));
    }
    const unsigned long long wfix = to_fixed(w);
    const size_t plane = static_cast<size_t>(jb.width) * jb.height;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int p = gb + 8 * m;
        if (p < k) {
            const MatchRec rr = rec[p];
            const size_t slot = static_cast<size_t>(jb.slots.frame(rr.t) % jb.acc_ring) * plane;
            unsigned long long* row = jb.num + slot + static_cast<size_t>(rr.y + y) * jb.width + rr.x;
#pragma unroll
            for (int x = 0; x < 8; ++x) atomicAdd(row + x, to_fixed(w * u[m][x]));
            if (y == 0) atomicAdd(jb.den + slot + static_cast<size_t>(rr.y) * jb.width + rr.x, wfix);
        }
    }
}

// Temporal tile path, block 8, G in {8, 16, 32}: 8 waves per workgroup, a
// wave per group with the register cube; every slot's contributions go into
// that slot's LDS sums (apron = the slot's reach around the tile), then out as
// the centre's float partials, slot by slot. The LDS (well over half a CU's)
// leaves one workgroup per CU, so the kernel is persistent: workgroup L (in
// XCD-contiguous order) filters tiles [L T / n, (L + 1) T / n) of the T tiles,
// and while one tile's partials are written it loads the next tile's records
// and each wave's first patches; within a tile, a wave loads its next
// group's patches while the current one is filtered.
// LDS: match records | counts | per slot num [ah][aw] and corners [ch][cw]
// (int64) | per-wave transpose buffers (also the box sum).
inline constexpr int kTTileWaves = 8;

struct SlotLds {
    int off;       // int64 offset of the slot's num (corners follow)
    int ax0, ay0;  // apron origin
    int aw, ah;
};

struct TileAt {
    int job, tile, tx, ty, i_begin, j_begin, ni, ngroups;
};

__device__ inline TileAt ttile_at(const TTileTable& tab, int lt) {
    TileAt a;
    a.job = find_job(tab, lt);
    const TTileJob& jb = tab.job[a.job];
    const TileGeom& tg = jb.tiles[0];
    a.tile = lt - jb.wg_begin;
    a.tx = a.tile % tg.ntx;
    a.ty = a.tile / tg.ntx;
    a.i_begin = a.tx * tg.tw;
    a.j_begin = a.ty * tg.th;
    a.ni = min(tg.tw, jb.grid.nx - a.i_begin);
    a.ngroups = a.ni * min(tg.th, jb.grid.ny - a.j_begin);
    return a;
}

// This lane's rows (patch gb + 8 m, row y) of a group with k patches, from
// their slots' frames (zero beyond k); the Wiener reference rows too.
template <int G, bool WIENER>
__device__ inline void ttile_gather(const TTileJob& jb, const MatchRec* rec, int k, int gb, int y,
                                    float (&v)[G / 8][8], float (&vr)[G / 8][8]) {
#pragma unroll
    for (int m = 0; m < G / 8; ++m) {
        const int p = gb + 8 * m;
        if (p < k) {
            const MatchRec r = rec[p];
            const int f = jb.slots.frame(r.t);
            const float* sp = jb.src.frame(f) + static_cast<size_t>(r.y + y) * jb.stride + r.x;
#pragma unroll
            for (int x = 0; x < 8; ++x) v[m][x] = sp[x];
            if constexpr (WIENER) {
                const float* wp = jb.wref.frame(f) + static_cast<size_t>(r.y + y) * jb.wref_stride + r.x;
#pragma unroll
                for (int x = 0; x < 8; ++x) vr[m][x] = wp[x];
            }
        } else {
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                v[m][x] = 0.f;
                if constexpr (WIENER) vr[m][x] = 0.f;
            }
        }
    }
}

template <int G, bool WIENER>
__global__ __launch_bounds__(512) void filter_temporal_tile_kernel(TTileTable tab, int total_tiles) {
    extern __shared__ __align__(16) unsigned long long smem64[];
    constexpr int B = 8;
    constexpr int M = G / 8;
    const int L = xcd_block(), nb = static_cast<int>(gridDim.x);
    const int lo = static_cast<int>(static_cast<long long>(L) * total_tiles / nb);
    const int hi = static_cast<int>(static_cast<long long>(L + 1) * total_tiles / nb);
    if (lo >= hi) return;
    // Groups per tile (one tile shape per launch; tw * th * G <= 512, one
    // record per thread).
    const int cap = tab.job[0].tiles[0].tw * tab.job[0].tiles[0].th;
    MatchRec* recs = reinterpret_cast<MatchRec*>(smem64);
    int* cnt = reinterpret_cast<int*>(recs + cap * G);
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(cnt + ((cap + 3) & ~3));
    __shared__ SlotLds sl[kMaxTileSlots];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int gb = lane >> 3, y = lane & 7;

    // Thread e's record (e < ngroups G) and count (e < ngroups) of tile a,
    // as four record words and the count.
    struct Fetched {
        int x, y, t;
        uint32_t dist;
        int count;
    };
    auto fetch = [&](const TileAt& a) {
        const int e = threadIdx.x;
        const TTileJob& jb = tab.job[a.job];
        Fetched f{0, 0, 0, 0u, 0};
        if (e < a.ngroups * G) {
            const int g = e / G, j = e - g * G;
            const int gj = g / a.ni, gi = g - gj * a.ni;
            const MatchRec r = jb.matches[static_cast<size_t>((a.j_begin + gj) * jb.grid.nx + a.i_begin + gi) * G + j];
            f.x = r.x;
            f.y = r.y;
            f.t = r.t;
            f.dist = r.dist;
        }
        if (e < a.ngroups) {
            const int gj = e / a.ni, gi = e - gj * a.ni;
            f.count = jb.counts[(a.j_begin + gj) * jb.grid.nx + a.i_begin + gi];
        }
        return f;
    };
    auto publish = [&](const TileAt& a, const Fetched& f) {
        const int e = threadIdx.x;
        if (e < a.ngroups * G) recs[e] = MatchRec{f.x, f.y, f.t, f.dist};
        if (e < a.ngroups) cnt[e] = f.count;
    };
    auto gather = [&](const TileAt& a, int g, float (&v)[M][8], float (&vr)[M][8]) {
        ttile_gather<G, WIENER>(tab.job[a.job], recs + g * G, cnt[g], gb, y, v, vr);
    };

    TileAt cur = ttile_at(tab, lo);
    publish(cur, fetch(cur));
    __syncthreads();
    float v[M][8], vr[M][8];
    if (wave < cur.ngroups) gather(cur, wave, v, vr);
    for (int lt = lo; lt < hi; ++lt) {
        const TTileJob& jb = tab.job[cur.job];
        const int ns = jb.nslots;
        // per slot: the numerator sums, then (den partials only) the corner sums
        const int cs = jb.corners ? 0 : 1;
        int total = 0;
        for (int t = 0; t < ns; ++t) {
            const TileGeom& g = jb.tiles[t];
            total += g.aw * g.ah + cs * (g.aw - B + 1) * (g.ah - B + 1);
        }
        if (static_cast<int>(threadIdx.x) < ns) {
            const int t = static_cast<int>(threadIdx.x);
            int o = 0;
            for (int q = 0; q < t; ++q)
                o += jb.tiles[q].aw * jb.tiles[q].ah + cs * (jb.tiles[q].aw - B + 1) * (jb.tiles[q].ah - B + 1);
            const TileGeom& g = jb.tiles[t];
            sl[t] = SlotLds{o, g.ax0(cur.tx, jb.width, B, jb.grid.step, jb.range[t]),
                            g.ay0(cur.ty, jb.height, B, jb.grid.step, jb.range[t]), g.aw, g.ah};
        }
        float* scratch = reinterpret_cast<float*>(acc + ((total + 1) & ~1));
        float* T = scratch + wave * (G * kTs);
        for (int e = threadIdx.x; e < total; e += 512) acc[e] = 0ull;
        __syncthreads();

        const float thr = 2.7f * jb.sigma;  // include/nss/params.hpp:46 kBmHardLambda
        const float sig2 = jb.sigma * jb.sigma;
        for (int g = wave; g < cur.ngroups; g += kTTileWaves) {
            const int k = cnt[g];
            const MatchRec* rec = recs + g * G;
            float c[G], r[G];
            forward_b8<G>(v, c, T, lane);
            if constexpr (WIENER) forward_b8<G>(vr, r, T, lane);
            if (g + kTTileWaves < cur.ngroups) gather(cur, g + kTTileWaves, v, vr);
            int kept = 0;
            float w2 = 0.f;
#pragma unroll
            for (int j = 0; j < G; ++j) {
                const bool dc = (lane == 0 && j == 0);
                if constexpr (WIENER) {
                    float w = 1.f;
                    if (!dc) {
                        const float r2 = r[j] * r[j];
                        const float den = r2 + sig2;
                        w = den > 0.f ? r2 / den : 0.f;
                    }
                    c[j] *= w;
                    w2 = fmaf(w, w, w2);
                } else {
                    if (!dc && fabsf(c[j]) < thr)
                        c[j] = 0.f;
                    else
                        ++kept;
                }
            }
            haar_inverse_reg<G>(c);
            float u[M][8];
            group_to_columns<G>(c, u, T, lane);
#pragma unroll
            for (int m = 0; m < M; ++m) dct8<true>(u[m]);
            transpose_rc<M>(u, T, lane);
#pragma unroll
            for (int m = 0; m < M; ++m) dct8<true>(u[m]);
            float w;
            if constexpr (WIENER) {
                w = 1.f / fmaxf(wave_sum_f32(w2), 1e-12f);
            } else {
                w = 1.f / static_cast<float>(max(wave_sum_i32(kept), 1));
            }
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const int p = gb + 8 * m;
                if (p < k) {
                    const MatchRec rr = rec[p];
                    const SlotLds q = sl[rr.t];
                    const int lx = rr.x - q.ax0, ly = rr.y - q.ay0;
                    // The predictive search keeps every slot-t match within the
                    // slot's reach; a patch outside the apron is dropped rather
                    // than written out of bounds.
                    if (lx >= 0 && ly >= 0 && lx <= q.aw - B && ly <= q.ah - B) {
                        unsigned long long* row = acc + q.off + (ly + y) * q.aw + lx;
#pragma unroll
                        for (int x = 0; x < 8; ++x) lds_add(row + x, to_fixed(w * u[m][x]));
                        if (y == 0) {
                            if (jb.corners)
                                atomicAdd(jb.corners +
                                              static_cast<size_t>(jb.slots.frame(rr.t) % jb.acc_ring) * jb.width *
                                                  jb.height +
                                              static_cast<size_t>(rr.y) * jb.width + rr.x,
                                          to_fixed(w));
                            else
                                lds_add(acc + q.off + q.aw * q.ah + ly * (q.aw - B + 1) + lx, to_fixed(w));
                        }
                    }
                }
            }
        }
        __syncthreads();
        // The next tile's records: loaded now, stored once the first slot's
        // partials are out (the records area is idle until then).
        const bool more = lt + 1 < hi;
        const TileAt nxt = more ? ttile_at(tab, lt + 1) : cur;
        const Fetched nf = more ? fetch(nxt) : Fetched{0, 0, 0, 0u, 0};
        for (int t = 0; t < ns; ++t) {
            const SlotLds q = sl[t];
            const TileGeom& g = jb.tiles[t];
            const size_t area = static_cast<size_t>(g.pitch) * g.ah;
            float* pn = jb.part + jb.slot_off[t] + static_cast<size_t>(cur.tile) * area + (q.ax0 & 3);
            float* pd = jb.corners ? nullptr : pn + static_cast<size_t>(g.ntx) * g.nty * area;
            partials_out_float<B>(q.aw, q.ah, g.pitch, acc + q.off, acc + q.off + q.aw * q.ah, scratch, pn, pd);
            if (t == 0 && more) publish(nxt, nf);
            __syncthreads();
            if (t == 0 && more && wave < nxt.ngroups) gather(nxt, wave, v, vr);
        }
        cur = nxt;
    }
}

// Temporal finish: frame n from the partials of every centre that reaches
// it (sources in ascending centre order), 64 x 16 pixels per block, four per
// thread, as finish_tiles_kernel.
__global__ __launch_bounds__(256) void finish_temporal_tiles_kernel(TFinishTable tab) {
    extern __shared__ __align__(16) unsigned long long tsm[];
    const int lb = xcd_block();
    const TFinishJob& jb = tab.job[find_job(tab, lb)];
    const int t = lb - jb.wg_begin;
    const int bx0 = (t % jb.tiles_x) * 64, by0 = (t / jb.tiles_x) * 16;
    const int x0 = bx0 + (threadIdx.x & 15) * 4;
    const int y = by0 + (threadIdx.x >> 4);
    const int B = jb.block, s = jb.step;
    long long dw[4] = {0, 0, 0, 0};
    if (jb.corners) {
        // den: the B x B box sum of the frame's corner plane (corners
        // (x - B + 1 .. x, y - B + 1 .. y)), staged in LDS, exact
        const int CW = 64 + B - 1, CH = 16 + B - 1;
        const int cx0 = bx0 - B + 1, cy0 = by0 - B + 1;
        unsigned long long* ct = tsm;            // [CH][CW]
        unsigned long long* hs = tsm + CH * CW;  // [CH][64]
        // eight loads per thread in flight before their stores
        for (int e0 = threadIdx.x; e0 < CH * CW; e0 += 8 * 256) {
            unsigned long long v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + 256 * u;
                const int r = e / CW, c = e - r * CW;
                const int py = cy0 + r, px = cx0 + c;
                v[u] = e < CH * CW && py >= 0 && px >= 0 && py < jb.height && px < jb.width
                           ? jb.corners[static_cast<size_t>(py) * jb.width + px]
                           : 0ull;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (e0 + 256 * u < CH * CW) ct[e0 + 256 * u] = v[u];
        }
        __syncthreads();
        for (int e = threadIdx.x; e < CH * 64; e += 256) {
            const int row = e >> 6, x = e & 63;
            const unsigned long long* cr = ct + row * CW + x;
            unsigned long long sum = 0;
            for (int b = 0; b < B; ++b) sum += cr[b];
            hs[e] = sum;
        }
        __syncthreads();
        unsigned long long* den = ct;  // [16][64]
        for (int e = threadIdx.x; e < 16 * 64; e += 256) {
            const unsigned long long* hc = hs + e;
            unsigned long long sum = 0;
            for (int a = 0; a < B; ++a) sum += hc[a * 64];
            den[e] = sum;
        }
        __syncthreads();
        const unsigned long long* dr = den + (y - by0) * 64 + (x0 - bx0);
#pragma unroll
        for (int q = 0; q < 4; ++q) dw[q] = static_cast<long long>(dr[q]);
    }
    if (x0 >= jb.width || y >= jb.height) return;
    const int xr = min(x0 + 3, jb.width - 1);
    f32x4 n = {0.f, 0.f, 0.f, 0.f}, d = {0.f, 0.f, 0.f, 0.f};
    for (int q = 0; q < jb.nsrc; ++q) {
        const TSource& src = jb.src[q];
        const TileGeom& tg = src.tiles;
        const int r = src.range;
        int txh = min(tg.ntx - 1, (xr + r) / (tg.tw * s));
        while (txh + 1 < tg.ntx && tg.ax0(txh + 1, jb.width, B, s, r) <= xr) ++txh;
        int tyh = min(tg.nty - 1, (y + r) / (tg.th * s));
        while (tyh + 1 < tg.nty && tg.ay0(tyh + 1, jb.height, B, s, r) <= y) ++tyh;
        const size_t area = static_cast<size_t>(tg.pitch) * tg.ah;
        for (int ty = tyh; ty >= 0; --ty) {
            const int oy = y - tg.ay0(ty, jb.height, B, s, r);
            if (oy >= tg.ah) break;
            for (int tx = txh; tx >= 0; --tx) {
                const int ax0 = tg.ax0(tx, jb.width, B, s, r);
                if (ax0 + tg.aw <= x0) break;
                const size_t o = (static_cast<size_t>(ty) * tg.ntx + tx) * area + static_cast<size_t>(oy) * tg.pitch +
                                 (x0 - ax0 + (ax0 & 3));
                const f32x4 pn = *reinterpret_cast<const f32x4*>(src.num + o);
                const f32x4 pd = jb.corners ? f32x4{0.f, 0.f, 0.f, 0.f} : *reinterpret_cast<const f32x4*>(src.den + o);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const bool in = static_cast<unsigned>(x0 + i - ax0) < static_cast<unsigned>(tg.aw);
                    n[i] += in ? pn[i] : 0.f;
                    d[i] += in ? pd[i] : 0.f;
                }
            }
        }
    }
    if (jb.corners)
#pragma unroll
        for (int i = 0; i < 4; ++i) d[i] = static_cast<float>(static_cast<double>(dw[i]) * (1.0 / 4294967296.0));
    float* dst = jb.dst + static_cast<size_t>(y) * jb.width;
    const float* fb = jb.fallback + static_cast<size_t>(y) * jb.width;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (x0 + i < jb.width) dst[x0 + i] = d[i] > 1e-12f ? n[i] / d[i] : fb[x0 + i];
}

// ---------------------------------------------------------------------------
// Finish (src/cpu/bm/agg.cpp:17-42 AggFinish): dst = den > 1e-12 ? num/den : src.
// ---------------------------------------------------------------------------

// Tile path: 64 x 16 pixels per block, four consecutive pixels of a row per
// thread; per pixel, the partials of every tile whose rectangle covers it
// (tiles in descending (ty, tx) order), read as aligned vectors per tile for
// the four pixels. EXACT: the spatial path's int64 numerator partials, summed
// exactly, and the weights from the records: every patch of a reference
// whose window reaches the block adds its group's weight (gweight) at its
// top-left corner in an LDS corner tile, which is box-summed B x B; then one
// conversion and num / den in double rounded to float (the atomic path's
// formula, finish_corner_kernel). Otherwise the temporal slot partials
// (floats) in slice mode.
template <bool EXACT>
__global__ __launch_bounds__(256) void finish_tiles_kernel(FinishTable tab) {
    extern __shared__ __align__(16) unsigned long long fsm[];
    // XCD-contiguous blocks: neighbouring blocks read the same tile partial
    // lines through one L2.
    const int lb = xcd_block();
    const FinishJob& jb = tab.job[find_job(tab, lb)];
    const int t = lb - jb.wg_begin;
    const int bx0 = (t % jb.tiles_x) * 64, by0 = jb.py0 + (t / jb.tiles_x) * 16;
    const int x0 = bx0 + (threadIdx.x & 15) * 4;
    const int y = by0 + (threadIdx.x >> 4);
    const int B = jb.block, s = jb.step, r = jb.range;
    long long dw[4] = {0, 0, 0, 0};  // EXACT: the weights' sums
    if constexpr (EXACT) {
        const int CW = 64 + B - 1, CH = 16 + B - 1;
        const int cx0 = bx0 - B + 1, cy0 = by0 - B + 1;  // corner tile origin
        unsigned long long* ct = fsm;                     // [CH][CW] corner weights
        unsigned long long* hs = fsm + CH * CW;           // [CH][64] horizontal B-wide sums
        for (int e = threadIdx.x; e < CH * CW; e += 256) ct[e] = 0ull;
        __syncthreads();
        // The references whose windows (reference +- range, clamped) reach
        // the corner tile, of the grid rows the records hold.
        const Grid& g = jb.grid;
        const int lx = cx0 - r, ux = bx0 + 63 + r, ly = cy0 - r, uy = by0 + 15 + r;
        const int i_lo = min(max(lx, 0) / s, g.nx - 1);
        const int i_hi = jb.width - B <= ux ? g.nx - 1 : min(g.nx - 1, ux / s);
        const int a_lo = max(max(ly, 0) / s, g.row0);
        const int a_hi = min(jb.height - B <= uy ? g.row0 + g.ny - 1 : uy / s, g.row0 + g.ny - 1);
        const int ni = i_hi - i_lo + 1, nj = a_hi - a_lo + 1;
        const int G = jb.group;
        const int pairs = ni > 0 && nj > 0 ? ni * nj * G : 0;
        // four (reference, record) pairs per thread and step, every load of
        // them issued before any is used (a reference holds G records; the
        // ones at or beyond its count are read and not used)
        for (int e0 = threadIdx.x; e0 < pairs; e0 += 1024) {
            int cnt[4], kk[4], ox[4], oy[4];
            unsigned long long wgt[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = min(e0 + 256 * u, pairs - 1);
                const int ref = e / G, k = e - ref * G;
                const int rj = ref / ni, ri = ref - rj * ni;
                const int rid = (a_lo - g.row0 + rj) * g.nx + i_lo + ri;
                cnt[u] = jb.counts[rid];
                const MatchRec m = jb.matches[static_cast<size_t>(rid) * G + k];
                wgt[u] = jb.gweight[rid];
                kk[u] = e0 + 256 * u < pairs ? k : G;
                ox[u] = m.x - cx0;
                oy[u] = m.y - cy0;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (kk[u] < cnt[u] && static_cast<unsigned>(ox[u]) < static_cast<unsigned>(CW) &&
                    static_cast<unsigned>(oy[u]) < static_cast<unsigned>(CH))
                    lds_add(ct + oy[u] * CW + ox[u], wgt[u]);
        }
        __syncthreads();
        // hs[row][x] = sum of ct[row][x .. x + B - 1], then the B-high sums
        // into the corner tile's area, den[yr][x]: a wave's lanes take 64
        // consecutive x of one row (bank-conflict-free reads and writes)
        for (int e = threadIdx.x; e < CH * 64; e += 256) {
            const int row = e >> 6, x = e & 63;
            const unsigned long long* cr = ct + row * CW + x;
            unsigned long long sum = 0;
            for (int b = 0; b < B; ++b) sum += cr[b];
            hs[e] = sum;
        }
        __syncthreads();
        unsigned long long* den = ct;  // [16][64]
        for (int e = threadIdx.x; e < 16 * 64; e += 256) {
            const unsigned long long* hc = hs + e;
            unsigned long long sum = 0;
            for (int a = 0; a < B; ++a) sum += hc[a * 64];
            den[e] = sum;
        }
        __syncthreads();
        // the weight at (x, y): corners (x - B + 1 .. x, y - B + 1 .. y)
        const unsigned long long* dr = den + (y - by0) * 64 + (x0 - bx0);
#pragma unroll
        for (int q = 0; q < 4; ++q) dw[q] = static_cast<long long>(dr[q]);
    }
    if (x0 >= jb.width || y >= jb.height) return;
    const TileGeom& tg = jb.tiles;
    const int xr = min(x0 + 3, jb.width - 1);
    int txh = min(tg.ntx - 1, (xr + r) / (tg.tw * s));
    while (txh + 1 < tg.ntx && tg.ax0(txh + 1, jb.width, B, s, r) <= xr) ++txh;
    int tyh = min(tg.nty - 1, (y + r) / (tg.th * s));
    while (tyh + 1 < tg.nty && tg.ay0(tyh + 1, jb.height, B, s, r) <= y) ++tyh;
    // a batch of tile rows: only its tiles, at partials index ty - ty0
    const bool batch = EXACT && jb.acc_num;
    const int ty_lo = batch ? jb.ty0 : 0;
    if (batch) tyh = min(tyh, jb.ty1 - 1);
    using Acc = typename std::conditional<EXACT, long long, float>::type;
    Acc n[4] = {0, 0, 0, 0}, d[4] = {0, 0, 0, 0};
    const size_t area = static_cast<size_t>(tg.pitch) * tg.ah;
    // The covering tiles are the rectangle [tyl, tyh] x [txl, txh]; walked
    // in descending (ty, tx) order, two tiles' partials are loaded before
    // either is added, so their memory latencies overlap (four at a time
    // need 106 VGPRs and ran slower).
    int tyl = tyh + 1;
    while (tyl - 1 >= ty_lo && y - tg.ay0(tyl - 1, jb.height, B, s, r) < tg.ah) --tyl;
    int txl = txh + 1;
    while (txl - 1 >= 0 && tg.ax0(txl - 1, jb.width, B, s, r) + tg.aw > x0) --txl;
    const int nx = txh - txl + 1;
    const int count = tyl <= tyh ? (tyh - tyl + 1) * nx : 0;
    for (int k0 = 0; k0 < count; k0 += 2) {
        Acc pn[2][4], pd[2][4];
        int lo[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int k = k0 + u;
            lo[u] = x0 + 4;  // no pixel of the four inside
#pragma unroll
            for (int i = 0; i < 4; ++i) pn[u][i] = pd[u][i] = Acc(0);
            if (k < count) {
                const int ty = tyh - k / nx, tx = txh - k % nx;
                const int oy = y - tg.ay0(ty, jb.height, B, s, r);
                const int ax0 = tg.ax0(tx, jb.width, B, s, r);
                lo[u] = ax0;
                // column of x0: a multiple of 4 in [0, pitch - 4]
                const size_t o = (static_cast<size_t>(ty - ty_lo) * tg.ntx + tx) * area +
                                 static_cast<size_t>(oy) * tg.pitch + (x0 - ax0 + (ax0 & 3));
                if constexpr (EXACT) {
                    const longlong2* qn =
                        static_cast<const longlong2*>(__builtin_assume_aligned(jb.exact_num + o, 32));
                    const longlong2 n0 = qn[0], n1 = qn[1];
                    pn[u][0] = n0.x, pn[u][1] = n0.y, pn[u][2] = n1.x, pn[u][3] = n1.y;
                } else {
                    const f32x4 vn = *reinterpret_cast<const f32x4*>(jb.part_num + o);
                    const f32x4 vd = *reinterpret_cast<const f32x4*>(jb.part_den + o);
#pragma unroll
                    for (int i = 0; i < 4; ++i) pn[u][i] = vn[i], pd[u][i] = vd[i];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const bool in = k0 + u < count && static_cast<unsigned>(x0 + i - lo[u]) < static_cast<unsigned>(tg.aw);
                n[i] += in ? pn[u][i] : Acc(0);
                d[i] += in ? pd[u][i] : Acc(0);
            }
    }
    if constexpr (EXACT) {
#pragma unroll
        for (int i = 0; i < 4; ++i) d[i] = dw[i];
        if (batch) {
            long long* an = jb.acc_num + static_cast<size_t>(y) * jb.width;
            long long* ad = jb.acc_den + static_cast<size_t>(y) * jb.width;
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (x0 + i < jb.width) {
                    an[x0 + i] += n[i];
                    ad[x0 + i] += d[i];
                }
            return;
        }
    }
    float* dst = jb.dst + static_cast<size_t>(y) * jb.dstride;
    if (jb.dst_den) {
        float* dd = jb.dst_den + static_cast<size_t>(y) * jb.dstride;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (x0 + i < jb.width) {
                if constexpr (EXACT) {
                    dst[x0 + i] = static_cast<float>(static_cast<double>(n[i]) * (1.0 / 4294967296.0));
                    dd[x0 + i] = static_cast<float>(static_cast<double>(d[i]) * (1.0 / 4294967296.0));
                } else {
                    dst[x0 + i] = n[i];
                    dd[x0 + i] = d[i];
                }
            }
        return;
    }
    const float* src = jb.src + static_cast<size_t>(y) * jb.sstride;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (x0 + i < jb.width) {
            if constexpr (EXACT)
                dst[x0 + i] = d[i] > 0 ? static_cast<float>(static_cast<double>(n[i]) / static_cast<double>(d[i]))
                                       : src[x0 + i];
            else
                dst[x0 + i] = d[i] > 1e-12f ? n[i] / d[i] : src[x0 + i];
        }
}

// Atomic path: den kept at each patch's top-left corner, box-summed here.
template <int B>
__global__ __launch_bounds__(256) void finish_corner_kernel(FinishTable tab) {
    constexpr int T = 16;
    constexpr int TS = T + B - 1;
    __shared__ unsigned long long corner[TS * TS];
    __shared__ unsigned long long colsum[T * TS];
    const int lb = xcd_block();
    const FinishJob& jb = tab.job[find_job(tab, lb)];
    const int tile = lb - jb.wg_begin;
    const int x0 = (tile % jb.tiles_x) * T, y0 = (tile / jb.tiles_x) * T;
    const int tid = threadIdx.x;
    for (int e = tid; e < TS * TS; e += 256) {
        const int yy = e / TS, xx = e - (e / TS) * TS;
        const int gy = y0 - (B - 1) + yy, gx = x0 - (B - 1) + xx;
        corner[e] = (gy >= 0 && gx >= 0 && gy < jb.height && gx < jb.width)
                        ? jb.den[static_cast<size_t>(gy) * jb.width + gx]
                        : 0ull;
    }
    __syncthreads();
    for (int e = tid; e < T * TS; e += 256) {
        const int oy = e / TS, xx = e - (e / TS) * TS;
        unsigned long long s = 0;
#pragma unroll
        for (int i = 0; i < B; ++i) s += corner[(oy + i) * TS + xx];
        colsum[e] = s;
    }
    __syncthreads();
    const int oy = tid / T, ox = tid % T;
    const int y = y0 + oy, x = x0 + ox;
    if (y < jb.height && x < jb.width) {
        unsigned long long d = 0;
#pragma unroll
        for (int i = 0; i < B; ++i) d += colsum[oy * TS + ox + i];
        const long long n = static_cast<long long>(jb.num[static_cast<size_t>(y) * jb.width + x]);
        const long long dd = static_cast<long long>(d);
        if (jb.dst_den) {
            jb.dst[static_cast<size_t>(y) * jb.dstride + x] =
                static_cast<float>(static_cast<double>(n) * (1.0 / 4294967296.0));
            jb.dst_den[static_cast<size_t>(y) * jb.dstride + x] =
                static_cast<float>(static_cast<double>(dd) * (1.0 / 4294967296.0));
            return;
        }
        jb.dst[static_cast<size_t>(y) * jb.dstride + x] =
            dd > 0 ? static_cast<float>(static_cast<double>(n) / static_cast<double>(dd))
                   : jb.src[static_cast<size_t>(y) * jb.sstride + x];
    }
}

// The output from frame-sized exact sums with the den per pixel (the
// memory-limited plan's tile buckets, whose batches added their partials
// into acc_num / acc_den): the division of finish_tiles_kernel. 64 x 4
// pixels per block.
__global__ __launch_bounds__(256) void finish_sums_kernel(FinishTable tab) {
    const int lb = xcd_block();
    const FinishJob& jb = tab.job[find_job(tab, lb)];
    const int t = lb - jb.wg_begin;
    const int x = (t % jb.tiles_x) * 64 + (threadIdx.x & 63);
    const int y = (t / jb.tiles_x) * 4 + (threadIdx.x >> 6);
    if (x >= jb.width || y >= jb.height) return;
    const size_t o = static_cast<size_t>(y) * jb.width + x;
    const long long n = jb.acc_num[o], d = jb.acc_den[o];
    jb.dst[static_cast<size_t>(y) * jb.dstride + x] =
        d > 0 ? static_cast<float>(static_cast<double>(n) / static_cast<double>(d))
              : jb.src[static_cast<size_t>(y) * jb.sstride + x];
}

// nss.VAggregate's reduction of one plane (src/cpu/bm/agg.cpp:159-201
// vaggregate_target): sum(num) / sum(den) over the contributions in their
// order, the source where sum(den) <= 1e-12. 64 x 4 pixels per block.
__global__ __launch_bounds__(256) void vaggregate_kernel(VAggJob jb) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= jb.width || y >= jb.height) return;
    float n = 0.f, d = 0.f;
    for (int k = 0; k < jb.count; ++k) {
        const size_t o = static_cast<size_t>(y) * jb.stride[k] + x;
        n += jb.num[k][o];
        d += jb.den[k][o];
    }
    jb.dst[static_cast<size_t>(y) * jb.dstride + x] =
        d > 1e-12f ? n / d : jb.src[static_cast<size_t>(y) * jb.sstride + x];
}

hipError_t launch_vaggregate(const VAggJob& jb, hipStream_t s) {
    if (launch_fault()) return hipErrorLaunchFailure;
    const dim3 grid((jb.width + 63) / 64, (jb.height + 3) / 4);
    vaggregate_kernel<<<grid, 256, 0, s>>>(jb);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void zero_kernel(ZeroTable tab) {
    const ZeroJob& jb = tab.job[find_job(tab, static_cast<int>(blockIdx.x))];
    const long long base = static_cast<long long>(static_cast<int>(blockIdx.x) - jb.wg_begin) * 1024;
    for (int i = threadIdx.x; i < 1024; i += 256) {
        const long long o = base + i;
        if (o < jb.count) jb.p[o] = 0ull;
    }
}

// ---------------------------------------------------------------------------
// WNNM: weighted nuclear-norm shrinkage of each group (src/cpu/wnnm/shrink.cpp:13-49,
// src/cpu/finishers.hpp:12-36, src/cpu/common/sv_shrink.cpp:9-24). The group
// matrix A has m = B^2 rows and k matched patches as columns, optionally
// centred row by row. Its spectrum comes from the k x k Gram matrix
// A^T A = V diag(s^2) V^T by cyclic Jacobi in round-robin order: every round
// rotates N/2 disjoint index pairs, so the whole N x N matrix updates at once,
// one entry per lane and step. The shrunk group is X = A V diag(s'/s) V^T
// (the product src/cpu/batch.cpp:490-520 forms), added back to the centre.
// ---------------------------------------------------------------------------
inline constexpr int kWnnmSweeps = 16;
// A pair rotates while |g_pq| > tol sqrt(g_pp g_qq) (src/cpu/wnnm/numerics.hpp:10)
// and |g_pq| is above FP32 rounding noise of the matrix, floor * trace: a
// rotation leaves g_pq at a few ulps of the largest entries, which the
// relative test alone can keep seeing for a pair of a large and a tiny
// eigenvalue. An off-diagonal at the floor moves the eigenvalues by about
// g_pq^2 / gap.
inline constexpr float kWnnmTol = 1e-6f;
inline constexpr float kWnnmFloor = 0x1p-21f;

// Round r of the round-robin pairing of the first n indices (n even):
// n - 1 is paired with r, and i with (2r - i) mod (n - 1) otherwise; over
// r = 0 .. n - 2 every pair meets once. Indices from n on pair with
// themselves. The schedule depends on the group's own column count only, so
// a group gives the same bits whatever size it is padded to.
__device__ inline int rr_partner(int i, int r, int n) {
    if (i >= n) return i;
    if (i == n - 1) return r;
    if (i == r) return n - 1;
    int j = 2 * r - i;
    if (j < 0) j += n - 1;
    else if (j >= n - 1) j -= n - 1;
    return j;
}

// One wave per reference. N = 8, 16 or 32 >= group; indices k .. N - 1 are
// zero columns that never rotate. LDS: A (N columns of m + 1) | Gram, later
// the shrink operator W (N x N + 1) | V (N x N + 1) | per-index rotation
// coefficients, diagonal steps (later the singular values) and gains.
template <int N, bool NCSR>
__device__ __forceinline__ void gram_group_body(const WnnmTable& tab) {
    extern __shared__ __align__(16) float wn[];
    constexpr int GP = N + 1;
    constexpr int E = N * N / 64;
    constexpr int LOG_N = N == 8 ? 3 : N == 16 ? 4 : 5;
    const int lane = threadIdx.x;
    const int lb = xcd_block();
    const WnnmJob& jb = tab.job[find_job(tab, lb)];
    const int g = lb - jb.wg_begin;
    if (g >= jb.nref) return;
    const int k = min(jb.counts[g], N);
    if (k <= 0) return;
    const int B = jb.block, m = B * B, P = m + 1;
    const MatchRec* rec = jb.matches + static_cast<size_t>(g) * jb.group;
    float* A = wn;
    float* G = A + N * P;
    float* V = G + N * GP;
    float* cself = V + N * GP;
    float* cpart = cself + N;
    float* tstep = cpart + N;
    float* gain = tstep + N;

    // The lane's pixels e = lane + 64 q of a patch (m <= 256).
    int soff[4], noff[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int e = lane + 64 * q, ey = e / B, ex = e - ey * B;
        soff[q] = ey * jb.stride + ex;
        noff[q] = ey * jb.width + ex;
    }
    for (int j = 0; j < N; ++j) {
        if (j < k) {
            const float* plane = jb.temporal ? jb.ring.frame(jb.slots.frame(rec[j].t)) : jb.src;
            const float* base = plane + static_cast<size_t>(rec[j].y) * jb.stride + rec[j].x;
            float v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = lane + 64 * q < m ? base[soff[q]] : 0.f;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (lane + 64 * q < m) A[j * P + lane + 64 * q] = v[q];
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (lane + 64 * q < m) A[j * P + lane + 64 * q] = 0.f;
        }
    }
    // Row centring (src/cpu/common/group_center.cpp:15-58): the mean of each
    // pixel over the k patches, subtracted; only the lane's own rows.
    float mean[4] = {0.f, 0.f, 0.f, 0.f};
    if (jb.residual) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = lane + 64 * q;
            if (e < m) {
                float sum = 0.f;
                for (int j = 0; j < k; ++j) sum += A[j * P + e];
                mean[q] = sum / static_cast<float>(k);
                for (int j = 0; j < k; ++j) A[j * P + e] -= mean[q];
            }
        }
    }
    __syncthreads();

    // Gram matrix and V = I. G[a][b] and G[b][a] take the same fmaf chain,
    // so the matrix starts exactly symmetric.
#pragma unroll
    for (int t = 0; t < E; ++t) {
        const int idx = lane + 64 * t, a = idx >> LOG_N, b = idx & (N - 1);
        float acc = 0.f;
        if (a < k && b < k)
            for (int r = 0; r < m; ++r) acc = fmaf(A[a * P + r], A[b * P + r], acc);
        G[a * GP + b] = acc;
        V[a * GP + b] = a == b ? 1.f : 0.f;
    }
    __syncthreads();
    float trace = 0.f;
    for (int j = 0; j < k; ++j) trace += G[j * GP + j];
    const float floor_abs = kWnnmFloor * trace;

    const int ke = k + (k & 1);  // the schedule's size: k rounded up to even
    if constexpr (N == 8) {
        // One entry per lane, in registers: lane (a, b) = 8a + b holds G[a][b]
        // and V[a][b]; the same rotations as the LDS form below, with the
        // neighbours' entries read across lanes (ds_bpermute) and no barriers.
        // Each lane forms its row index's rotation; column b's is row b's.
        const int a = lane >> 3, b = lane & 7;
        float g = G[a * GP + b], v = a == b ? 1.f : 0.f;
        for (int sweep = 0; sweep < kWnnmSweeps; ++sweep) {
            bool rotated = false;
            for (int r = 0; r < ke - 1; ++r) {
                const int pa = rr_partner(a, r, ke), pb = rr_partner(b, r, ke);
                const int p0 = min(a, pa), q0 = max(a, pa);
                const float app = __shfl(g, p0 * 9, 64), aqq = __shfl(g, q0 * 9, 64);
                const float apq = __shfl(g, p0 * 8 + q0, 64);
                float c = 1.f, sn = 0.f, tn = 0.f;
                bool rot = false;
                if (q0 < k && fabsf(apq) > floor_abs && apq * apq > kWnnmTol * kWnnmTol * fabsf(app * aqq)) {
                    const float delta = (aqq - app) * 0.5f;
                    const float root = __builtin_amdgcn_sqrtf(fmaf(delta, delta, apq * apq));
                    tn = apq * __builtin_amdgcn_rcpf(delta + copysignf(root, delta));
                    c = __builtin_amdgcn_rsqf(fmaf(tn, tn, 1.f));
                    sn = tn * c;
                    rot = true;
                }
                if (!__any(rot)) continue;
                rotated = true;
                const float ca = c, sa = a == p0 ? -sn : sn, ta = a == p0 ? -tn : tn;
                const float cb = __shfl(ca, b * 8, 64), sb = __shfl(sa, b * 8, 64);
                const float g_apb = __shfl(g, a * 8 + pb, 64), v_apb = __shfl(v, a * 8 + pb, 64);
                const float g_pab = __shfl(g, pa * 8 + b, 64), g_papb = __shfl(g, pa * 8 + pb, 64);
                const float vgen = ca * fmaf(cb, g, sb * g_apb) + sa * fmaf(cb, g_pab, sb * g_papb);
                const float vdiag = fmaf(ta, g_apb, g);  // at a == b, pb == pa
                const float vpair = sa != 0.f ? 0.f : g;
                const float gn = a == b ? vdiag : (b == pa ? vpair : vgen);
                v = fmaf(cb, v, sb * v_apb);
                g = gn;
            }
            if (!rotated) break;
        }
        G[a * GP + b] = g;
        V[a * GP + b] = v;
        __syncthreads();
    } else {
    for (int sweep = 0; sweep < kWnnmSweeps; ++sweep) {
        bool rotated = false;
        for (int r = 0; r < ke - 1; ++r) {
            // Rotation of pair (p, q), p < q, zeroing g_pq: J = [c s; -s c],
            // G' = J^T G J, V' = V J, t = tan(theta) the smaller root of
            // t^2 + 2 zeta t - 1 = 0 with zeta = (g_qq - g_pp) / (2 g_pq).
            bool rot = false;
            if (lane < N) {
                const int i = lane, o = rr_partner(i, r, ke);
                const int p0 = min(i, o), q0 = max(i, o);
                float c = 1.f, sn = 0.f, tn = 0.f;
                if (q0 < k) {
                    const float app = G[p0 * GP + p0], aqq = G[q0 * GP + q0], apq = G[p0 * GP + q0];
                    // |g_pq| > tol sqrt(g_pp g_qq), squared. Gram entries of
                    // plane values stay far from overflow in the squares, so
                    // the tangent needs no scaling; v_rcp / v_sqrt (1 ulp)
                    // keep the rotation orthogonal to a few ulps.
                    if (fabsf(apq) > floor_abs && apq * apq > kWnnmTol * kWnnmTol * fabsf(app * aqq)) {
                        const float delta = (aqq - app) * 0.5f;
                        const float root = __builtin_amdgcn_sqrtf(fmaf(delta, delta, apq * apq));
                        tn = apq * __builtin_amdgcn_rcpf(delta + copysignf(root, delta));
                        c = __builtin_amdgcn_rsqf(fmaf(tn, tn, 1.f));
                        sn = tn * c;
                        rot = true;
                    }
                }
                cself[i] = c;
                cpart[i] = i == p0 ? -sn : sn;
                tstep[i] = i == p0 ? -tn : tn;
                rotated |= rot;
            }
            __syncthreads();
            // A round where no pair rotates leaves G and V as they are; the
            // next round's coefficients overwrite these unread.
            if (!__any(rot)) continue;
            float ng[E], nv[E];
            // A lane's entries share their column (64 is a multiple of N).
            const int b = lane & (N - 1), pb = rr_partner(b, r, ke);
            const float cb = cself[b], sb = cpart[b];
#pragma unroll
            for (int t = 0; t < E; ++t) {
                const int a = (lane + 64 * t) >> LOG_N;
                const int pa = rr_partner(a, r, ke);
                const float ca = cself[a], sa = cpart[a];
                const float gab = G[a * GP + b];
                // Selected, not branched: the three cases share the wave.
                const float vgen =
                    ca * fmaf(cb, gab, sb * G[a * GP + pb]) + sa * fmaf(cb, G[pa * GP + b], sb * G[pa * GP + pb]);
                const float vdiag = fmaf(tstep[a], G[a * GP + pa], gab);
                const float vpair = sa != 0.f ? 0.f : gab;
                ng[t] = a == b ? vdiag : (b == pa ? vpair : vgen);
                nv[t] = fmaf(cb, V[a * GP + b], sb * V[a * GP + pb]);
            }
            __syncthreads();
#pragma unroll
            for (int t = 0; t < E; ++t) {
                const int a = (lane + 64 * t) >> LOG_N;
                G[a * GP + b] = ng[t];
                V[a * GP + b] = nv[t];
            }
            __syncthreads();
        }
        if (!__any(rotated)) break;
    }
    }

    // Spectrum: s_j and its place in descending order (ties by index).
    if (lane < N) tstep[lane] = lane < k ? sqrtf(fmaxf(G[lane * GP + lane], 0.f)) : 0.f;
    if constexpr (NCSR) {
        // Column weights exp(-dist / h), h = max(2 m sigma^2, 1e-12), the
        // reference column at distance 0 (src/cpu/ncsr/centralize.cpp:60-88).
        const float sg = jb.sigma;
        const float h = fmaxf(2.f * static_cast<float>(m) * sg * sg, 1e-12f);
        if (lane < N) cpart[lane] = lane < k ? expf(-(lane > 0 ? __uint_as_float(rec[lane].dist) : 0.f) / h) : 0.f;
    }
    __syncthreads();
    int place = 0;
    if (lane < k) {
        const float sj = tstep[lane];
        for (int i = 0; i < k; ++i) {
            const float si = tstep[i];
            place += (si > sj || (si == sj && i < lane)) ? 1 : 0;
        }
    }
    const int rank = min(m, k);
    int kept = 0;
    if constexpr (NCSR) {
        // Code row i of the centred group's PCA, b_ij = s_i V[j][i]: its
        // weighted mean beta, the second moment over the k columns about it,
        // tau = 2 sqrt(2) sigma^2 / (sqrt(moment / k) + 1e-12), the row
        // soft-thresholded about beta (src/cpu/finishers.hpp:53-80,
        // src/cpu/ncsr/centralize.cpp:90-141), kept for the rank largest
        // values with s > 0 and stored over s_i: C[i][j]. The Gram area is
        // free (the spectrum is in tstep).
        const float sg = jb.sigma;
        float wsum = 0.f;
        for (int j = 0; j < k; ++j) wsum += cpart[j];
        const bool unit = !(wsum > 0.f);
        const float inverse = 1.f / (unit ? static_cast<float>(k) : wsum);
        if (lane < k) {
            const int i = lane;
            const float si = tstep[i];
            if (!(sg > 0.f)) {
                for (int j = 0; j < k; ++j) G[i * GP + j] = V[j * GP + i];
            } else if (place < rank && si > 0.f) {
                float beta = 0.f;
                for (int j = 0; j < k; ++j) beta = fmaf(unit ? 1.f : cpart[j], si * V[j * GP + i], beta);
                beta *= inverse;
                float moment = 0.f;
                for (int j = 0; j < k; ++j) {
                    const float e = si * V[j * GP + i] - beta;
                    moment = fmaf(e, e, moment);
                }
                const float tau = 2.8284271247461903f * sg * sg / (sqrtf(moment / static_cast<float>(k)) + 1e-12f);
                const float inv_s = 1.f / si;
                for (int j = 0; j < k; ++j) {
                    const float d = si * V[j * GP + i] - beta, a = fabsf(d);
                    G[i * GP + j] = ((a > tau ? copysignf(a - tau, d) : 0.f) + beta) * inv_s;
                }
            } else {
                for (int j = 0; j < k; ++j) G[i * GP + j] = 0.f;
            }
        }
        __syncthreads();
        // M = V C (M[a][b] = sum_i V[a][i] C[i][b]), then over the Gram area.
        float mv[E];
#pragma unroll
        for (int t = 0; t < E; ++t) {
            const int idx = lane + 64 * t, a = idx >> LOG_N, b = idx & (N - 1);
            float w = 0.f;
            if (a < k && b < k)
                for (int i = 0; i < k; ++i) w = fmaf(V[a * GP + i], G[i * GP + b], w);
            mv[t] = w;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < E; ++t) {
            const int idx = lane + 64 * t;
            G[(idx >> LOG_N) * GP + (idx & (N - 1))] = mv[t];
        }
        __syncthreads();
    } else {
        // Places below `start` stay; later ones shrink to
        // (s + sqrt(s^2 - c)) / 2 while s^2 > c, with c = 8 sqrt(2k) sigma^2.
        bool keep = false;
        float gj = 0.f;
        if (lane < k) {
            const float sj = tstep[lane];
            const int start = min(jb.residual ? 0 : 1, rank);
            const float constant = 8.f * sqrtf(2.f * static_cast<float>(k)) * jb.sigma * jb.sigma;
            if (place < start) {
                gj = 1.f;
                keep = true;
            } else if (place < rank) {
                const float tmp = sj * sj - constant;
                if (tmp > 0.f) {
                    gj = ((sj + sqrtf(tmp)) * 0.5f) / sj;
                    keep = true;
                }
            }
        }
        if (lane < N) gain[lane] = gj;
        kept = static_cast<int>(__popcll(__ballot(keep)));
        __syncthreads();
        // W = V diag(gain) V^T over the Gram area (G is no longer read).
#pragma unroll
        for (int t = 0; t < E; ++t) {
            const int idx = lane + 64 * t, a = idx >> LOG_N, b = idx & (N - 1);
            float w = 0.f;
            if (a < k && b < k)
                for (int i = 0; i < k; ++i) w = fmaf(V[a * GP + i] * gain[i], V[b * GP + i], w);
            G[a * GP + b] = w;
        }
        __syncthreads();
    }

    // X = A W (A M for NCSR) plus the centre, weighted into the sums: WNNM
    // 1 / kept when adaptive and kept > 0, else 1 (src/cuda/wnnm/kernels.cu:84-105);
    // NCSR 1.
    const float wt = !NCSR && jb.adaptive && kept > 0 ? 1.f / static_cast<float>(kept) : 1.f;
    const unsigned long long wfix = to_fixed(wt);
    for (int j = 0; j < k; ++j) {
        const size_t slice = jb.temporal ? static_cast<size_t>(rec[j].t) * jb.sum_slot : 0;
        unsigned long long* cell = jb.num + slice + static_cast<size_t>(rec[j].y) * jb.width + rec[j].x;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = lane + 64 * q;
            if (e < m) {
                float x = 0.f;
                for (int i = 0; i < k; ++i) x = fmaf(A[i * P + e], G[i * GP + j], x);
                atomicAdd(cell + noff[q], to_fixed(wt * (x + mean[q])));
            }
        }
        if (lane == 0) atomicAdd(jb.den + slice + static_cast<size_t>(rec[j].y) * jb.width + rec[j].x, wfix);
    }
}

template <int N>
__global__ __launch_bounds__(64) void wnnm_group_kernel(WnnmTable tab) {
    gram_group_body<N, false>(tab);
}

// NCSR on the same eigenbasis (src/cuda/ncsr/kernels.cu:48-190): the rows
// are always centred and the shrink works on the code rows.
template <int N>
__global__ __launch_bounds__(64) void ncsr_group_kernel(WnnmTable tab) {
    gram_group_body<N, true>(tab);
}

// NCSR rounds after the first: est += delta (noisy - est)
// (src/cpu/common/iter_regularize.cpp:15-39). 64 x 4 pixels per block.
__global__ __launch_bounds__(256) void regularize_kernel(float* est, int est_stride, const float* noisy,
                                                         int noisy_stride, int width, int height, float delta) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= width || y >= height) return;
    float* e = est + static_cast<size_t>(y) * est_stride + x;
    const float n = noisy[static_cast<size_t>(y) * noisy_stride + x];
    *e = fmaf(delta, n - *e, *e);
}

// ---------------------------------------------------------------------------
// Host-side sizing and launchers.
// ---------------------------------------------------------------------------

size_t match_wave_lds_bytes(int block, int range, bool stage) {
    const int w = 2 * range + 1;
    const size_t cand = static_cast<size_t>(w) * w * 4;
    if (!stage) return cand;
    const size_t reg = static_cast<size_t>(w + block - 1) * (w + block - 1);
    return ((reg + 3) & ~static_cast<size_t>(3)) * 4 + cand;
}

static int mm_list(int group) { return group > 16 ? 256 : 128; }

size_t match_mfma_lds_bytes(int step, int range, int group) {
    // the union of a 4 x 4 block's windows spans at most 3 step + 2 range + 1
    // candidate columns and rows
    const int span = 3 * step + 2 * range + 1;
    const int nt = (span + 15) >> 4;
    const size_t list = static_cast<size_t>(mm_list(group));
    const size_t reg = std::max(static_cast<size_t>(span + 7) * mm_pitch(nt), 16 * list);
    return (reg + std::max(8 * list, static_cast<size_t>(2048)) + 66) * 4;
}

bool match_mfma_fits(int block, int group, int step, int range, int width, int height) {
    return block == 8 && group >= 1 && group <= 32 && width >= 8 && height >= 8 &&
           match_mfma_lds_bytes(step, range, group) <= 160 * 1024;
}

int match_workgroups(int kind, const Grid& g) {
    if (kind == 3) return ((g.nx + 3) >> 2) * ((g.ny + 3) >> 2);
    if (kind == 2) return g.ny * ((g.nx + kRunRefs - 1) / kRunRefs);
    return g.nx * g.ny;
}

// LDS of the run matcher in a layout: NT = 1 or kRunMidTiles (keys in
// registers; the window keys and the gathered keys both over the region) or
// 0 (the window keys after the region, the gathered keys over it, or after
// the window keys when the region is smaller).
static size_t run_lds_layout(int block, int step, int range, int nt) {
    const size_t w = 2 * static_cast<size_t>(range) + 1;
    const size_t region = static_cast<size_t>(run_region_floats(block, step, range)) * 4;
    const size_t keys = kRunRefs * w * w * 4;
    const size_t gather = kRunRefs * static_cast<size_t>(nt == kRunMidTiles ? kRunMidGather : kRunGather) * 8;
    return nt > 0 ? std::max(region, keys + gather) : region + keys + (region < gather ? gather : 0);
}

// The layout a job of this range takes alone.
static int run_layout(int range) { return run_keys_in_regs(range) ? 1 : run_keys_mid(range) ? kRunMidTiles : 0; }

size_t match_run_lds_bytes(int block, int step, int range) {
    return run_lds_layout(block, step, range, run_layout(range));
}

size_t filter_atomic_lds_bytes(int block, int group, bool wiener) {
    return static_cast<size_t>(group) * (block * block + 1) * 4 * (wiener ? 2 : 1);
}

size_t filter_tile_b8_lds_bytes(int group, const TileGeom& t) {
    const size_t aw = t.aw, ah = t.ah;
    const size_t acc = aw * ah * 8;
    const size_t recs = static_cast<size_t>(t.tw) * t.th * group * sizeof(MatchRec) +
                        ((static_cast<size_t>(t.tw) * t.th + 3) & ~static_cast<size_t>(3)) * 4;
    const size_t tbuf = 4 * static_cast<size_t>(group) * kTs * 4;
    return acc + recs + tbuf;
}

size_t filter_tile_mfma_lds_bytes(int group, const TileGeom& t) {
    const size_t aw = t.aw, ah = t.ah;
    const size_t acc = ((aw * ah + 1) & ~static_cast<size_t>(1)) * 8;
    const size_t recs = ((static_cast<size_t>(t.tw) * t.th * group + 3) & ~static_cast<size_t>(3)) * 4 +
                        ((static_cast<size_t>(t.tw) * t.th + 3) & ~static_cast<size_t>(3)) * 4;
    const size_t tbuf = 4 * static_cast<size_t>(kMfmaBuf) * 4 + 64 * 4;  // + the fp16 form's DCT table
    return acc + recs + tbuf;
}

size_t filter_tile_lds_bytes(int block, int group, bool wiener, const TileGeom& t) {
    const size_t aw = t.aw, ah = t.ah;
    const size_t acc = aw * ah * 8;
    return acc + 4 * filter_atomic_lds_bytes(block, group, wiener);
}

template <typename K>
static void allow_lds(K kernel, size_t lds) {
    if (lds > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>(lds));
}

template <int B>
static hipError_t launch_match_b(const MatchTable& t, int total_wg, size_t lds, int kind, hipStream_t s) {
    if (kind == 3) {
        if constexpr (B == 8) {
            // One list length for the whole launch (the largest group's), so
            // the LDS is sized here for every job at that length.
            int gmax = 0;
            for (int j = 0; j < t.n; ++j) gmax = std::max(gmax, t.job[j].group);
            for (int j = 0; j < t.n; ++j)
                lds = std::max(lds, match_mfma_lds_bytes(t.job[j].grid.step, t.job[j].range, gmax));
            if (lds > 160 * 1024) return hipErrorInvalidValue;
            if (mm_list(gmax) == 256) {
                allow_lds(&match_mfma_kernel<256>, lds);
                match_mfma_kernel<256><<<total_wg, 256, lds, s>>>(t);
            } else {
                allow_lds(&match_mfma_kernel<128>, lds);
                match_mfma_kernel<128><<<total_wg, 256, lds, s>>>(t);
            }
        } else {
            return hipErrorInvalidValue;
        }
    } else if (kind == 2) {
        if constexpr (B <= 12) {
            // the layout every job's window takes (the one of the largest
            // range); the LDS for the layout launched
            int nt = 1;
            for (int j = 0; j < t.n; ++j) {
                const int l = run_layout(t.job[j].range);
                nt = nt == 0 || l == 0 ? 0 : std::max(nt, l);
            }
            lds = 0;
            for (int j = 0; j < t.n; ++j)
                lds = std::max(lds, run_lds_layout(B, t.job[j].grid.step, t.job[j].range, nt));
            if (nt == 1) {
                allow_lds(&match_run_kernel<B, true>, lds);
                match_run_kernel<B, true><<<total_wg, 256, lds, s>>>(t);
            } else if (nt == kRunMidTiles) {
                allow_lds(&match_run_mid_kernel<B>, lds);
                match_run_mid_kernel<B><<<total_wg, 256, lds, s>>>(t);
            } else {
                allow_lds(&match_run_kernel<B, false>, lds);
                match_run_kernel<B, false><<<total_wg, 256, lds, s>>>(t);
            }
        } else {
            return hipErrorInvalidValue;
        }
    } else if (kind == 1) {
        allow_lds(&match_wave_kernel<B, true>, lds);
        match_wave_kernel<B, true><<<total_wg, 64, lds, s>>>(t);
    } else {
        allow_lds(&match_wave_kernel<B, false>, lds);
        match_wave_kernel<B, false><<<total_wg, 64, lds, s>>>(t);
    }
    return hipGetLastError();
}

// Block sizes outside the BM3D set (the group filters take any block from 1
// to 16): the wave matcher only, window in LDS (kind 1) or from global (0).
template <int B>
static hipError_t launch_match_wave(const MatchTable& t, int total_wg, size_t lds, int kind, hipStream_t s) {
    if (kind == 1) {
        allow_lds(&match_wave_kernel<B, true>, lds);
        match_wave_kernel<B, true><<<total_wg, 64, lds, s>>>(t);
    } else if (kind == 0) {
        allow_lds(&match_wave_kernel<B, false>, lds);
        match_wave_kernel<B, false><<<total_wg, 64, lds, s>>>(t);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// kind: 3 MFMA matcher (B = 8), 2 run matcher, 1 wave matcher with the
// window in LDS, 0 from global.
hipError_t launch_match(int block, const MatchTable& t, int total_wg, size_t lds, int kind, hipStream_t s) {
    if (launch_fault()) return hipErrorLaunchFailure;
    switch (block) {
        case 3: return launch_match_wave<3>(t, total_wg, lds, kind, s);
        case 5: return launch_match_wave<5>(t, total_wg, lds, kind, s);
        case 6: return launch_match_wave<6>(t, total_wg, lds, kind, s);
        case 7: return launch_match_wave<7>(t, total_wg, lds, kind, s);
        case 9: return launch_match_wave<9>(t, total_wg, lds, kind, s);
        case 10: return launch_match_wave<10>(t, total_wg, lds, kind, s);
        case 11: return launch_match_wave<11>(t, total_wg, lds, kind, s);
        case 13: return launch_match_wave<13>(t, total_wg, lds, kind, s);
        case 14: return launch_match_wave<14>(t, total_wg, lds, kind, s);
        case 15: return launch_match_wave<15>(t, total_wg, lds, kind, s);
        case 1: return launch_match_b<1>(t, total_wg, lds, kind, s);
        case 2: return launch_match_b<2>(t, total_wg, lds, kind, s);
        case 4: return launch_match_b<4>(t, total_wg, lds, kind, s);
        case 8: return launch_match_b<8>(t, total_wg, lds, kind, s);
        case 12: return launch_match_b<12>(t, total_wg, lds, kind, s);
        case 16: return launch_match_b<16>(t, total_wg, lds, kind, s);
        case 32: return launch_match_b<32>(t, total_wg, lds, kind, s);
        default: return hipErrorInvalidValue;
    }
}

template <int B>
static hipError_t launch_filter_b(const FilterTable& t, int total_wg, size_t lds, bool wiener, bool tiles,
                                  hipStream_t s) {
    if (tiles) {
        if (wiener) {
            allow_lds(&filter_tile_kernel<B, true>, lds);
            filter_tile_kernel<B, true><<<total_wg, 256, lds, s>>>(t);
        } else {
            allow_lds(&filter_tile_kernel<B, false>, lds);
            filter_tile_kernel<B, false><<<total_wg, 256, lds, s>>>(t);
        }
    } else {
        // a table with global cubes runs kCubeSlots workgroups (no LDS cube)
        const int grid = t.cube ? std::min(total_wg, t.cube_slots) : total_wg;
        if (t.cube) lds = 0;
        if (wiener) {
            allow_lds(&filter_atomic_kernel<B, true>, lds);
            filter_atomic_kernel<B, true><<<grid, 64, lds, s>>>(t);
        } else {
            allow_lds(&filter_atomic_kernel<B, false>, lds);
            filter_atomic_kernel<B, false><<<grid, 64, lds, s>>>(t);
        }
    }
    return hipGetLastError();
}

template <int G>
static hipError_t launch_filter_b8(const FilterTable& t, int total_wg, size_t lds, bool wiener, hipStream_t s) {
    if (wiener) {
        allow_lds(&filter_tile_b8_kernel<G, true>, lds);
        filter_tile_b8_kernel<G, true><<<total_wg, 256, lds, s>>>(t);
    } else {
        allow_lds(&filter_tile_b8_kernel<G, false>, lds);
        filter_tile_b8_kernel<G, false><<<total_wg, 256, lds, s>>>(t);
    }
    return hipGetLastError();
}

template <int G, int FORM>
static hipError_t launch_filter_mfma(const FilterTable& t, int total_wg, size_t lds, bool wiener, hipStream_t s) {
    if (wiener) {
        allow_lds(&filter_tile_mfma_kernel<G, true, FORM>, lds);
        filter_tile_mfma_kernel<G, true, FORM><<<total_wg, 256, lds, s>>>(t);
    } else {
        allow_lds(&filter_tile_mfma_kernel<G, false, FORM>, lds);
        filter_tile_mfma_kernel<G, false, FORM><<<total_wg, 256, lds, s>>>(t);
    }
    return hipGetLastError();
}

// fast_group: 8, 16 or 32 selects the block-8 register kernel (tile path);
// form 1 (bf16x3) or 2 (fp16) with fast_group 8 or 16 the MFMA kernel.
hipError_t launch_filter(int block, const FilterTable& t, int total_wg, size_t lds, bool wiener, bool tiles,
                         int fast_group, int form, hipStream_t s) {
    if (launch_fault()) return hipErrorLaunchFailure;
    if (tiles && block == 8 && form) {
        if (fast_group == 8) return form == 1 ? launch_filter_mfma<8, kFormBf16x3>(t, total_wg, lds, wiener, s)
                                              : launch_filter_mfma<8, kFormF16>(t, total_wg, lds, wiener, s);
        if (fast_group == 16) return form == 1 ? launch_filter_mfma<16, kFormBf16x3>(t, total_wg, lds, wiener, s)
                                               : launch_filter_mfma<16, kFormF16>(t, total_wg, lds, wiener, s);
        return hipErrorInvalidValue;
    }
    if (tiles && block == 8 && fast_group) {
        switch (fast_group) {
            case 8: return launch_filter_b8<8>(t, total_wg, lds, wiener, s);
            case 16: return launch_filter_b8<16>(t, total_wg, lds, wiener, s);
            case 32: return launch_filter_b8<32>(t, total_wg, lds, wiener, s);
            default: return hipErrorInvalidValue;
        }
    }
    switch (block) {
        case 1: return launch_filter_b<1>(t, total_wg, lds, wiener, tiles, s);
        case 2: return launch_filter_b<2>(t, total_wg, lds, wiener, tiles, s);
        case 4: return launch_filter_b<4>(t, total_wg, lds, wiener, tiles, s);
        case 8: return launch_filter_b<8>(t, total_wg, lds, wiener, tiles, s);
        case 12: return launch_filter_b<12>(t, total_wg, lds, wiener, tiles, s);
        case 16: return launch_filter_b<16>(t, total_wg, lds, wiener, tiles, s);
        case 32: return launch_filter_b<32>(t, total_wg, lds, wiener, tiles, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_finish_sums(const FinishTable& t, int total_wg, hipStream_t s) {
    if (launch_fault()) return hipErrorLaunchFailure;
    finish_sums_kernel<<<total_wg, 256, 0, s>>>(t);
    return hipGetLastError();
}

size_t finish_tiles_lds_bytes(int block) {
    const size_t ch = 16 + block - 1, cw = 64 + block - 1;
    return (ch * cw + ch * 64) * 8;
}

hipError_t launch_finish(int block, const FinishTable& t, int total_wg, bool tiles, hipStream_t s) {
    if (launch_fault()) return hipErrorLaunchFailure;
    if (tiles) {
        if (t.n && t.job[0].exact_num) {
            // the corner tile and its horizontal sums of the largest block
            int bmax = 1;
            for (int k = 0; k < t.n; ++k) bmax = std::max(bmax, t.job[k].block);
            const size_t lds = finish_tiles_lds_bytes(bmax);
            allow_lds(&finish_tiles_kernel<true>, lds);
            finish_tiles_kernel<true><<<total_wg, 256, lds, s>>>(t);
        } else {
            finish_tiles_kernel<false><<<total_wg, 256, 0, s>>>(t);
        }
        return hipGetLastError();
    }
    switch (block) {
        case 1: finish_corner_kernel<1><<<total_wg, 256, 0, s>>>(t); break;
        case 2: finish_corner_kernel<2><<<total_wg, 256, 0, s>>>(t); break;
        case 4: finish_corner_kernel<4><<<total_wg, 256, 0, s>>>(t); break;
        case 8: finish_corner_kernel<8><<<total_wg, 256, 0, s>>>(t); break;
        case 12: finish_corner_kernel<12><<<total_wg, 256, 0, s>>>(t); break;
        case 16: finish_corner_kernel<16><<<total_wg, 256, 0, s>>>(t); break;
        case 32: finish_corner_kernel<32><<<total_wg, 256, 0, s>>>(t); break;
        case 3: finish_corner_kernel<3><<<total_wg, 256, 0, s>>>(t); break;
        case 5: finish_corner_kernel<5><<<total_wg, 256, 0, s>>>(t); break;
        case 6: finish_corner_kernel<6><<<total_wg, 256, 0, s>>>(t); break;
        case 7: finish_corner_kernel<7><<<total_wg, 256, 0, s>>>(t); break;
        case 9: finish_corner_kernel<9><<<total_wg, 256, 0, s>>>(t); break;
        case 10: finish_corner_kernel<10><<<total_wg, 256, 0, s>>>(t); break;
        case 11: finish_corner_kernel<11><<<total_wg, 256, 0, s>>>(t); break;
        case 13: finish_corner_kernel<13><<<total_wg, 256, 0, s>>>(t); break;
        case 14: finish_corner_kernel<14><<<total_wg, 256, 0, s>>>(t); break;
        case 15: finish_corner_kernel<15><<<total_wg, 256, 0, s>>>(t); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_zero(const ZeroTable& t, int total_wg, hipStream_t s) {
    if (launch_fault()) return hipErrorLaunchFailure;
    zero_kernel<<<total_wg, 256, 0, s>>>(t);
    return hipGetLastError();
}

size_t predictive_table_lds(int block, PredTable& t, int radius) {
    // Candidate keys: one pass for a layer's union when it fits kPredChunk.
    // Merged candidates: (G - 1) spatial + min(ps_num, G) per layer and direction.
    int chunk = 64, region = 0;
    size_t merged = 0;
    for (int j = 0; j < t.n; ++j) {
        const PredJob& p = t.job[j];
        const int kl = std::min(p.ps_num, p.group);
        const int w = 2 * p.ps_range + 1, cells = kl * w * w;
        chunk = std::max(chunk, std::min(kPredChunk, (cells + 63) / 64 * 64));
        merged = std::max(merged, static_cast<size_t>(p.group - 1 + 2 * radius * kl));
        const int need = kl * (w + block - 1) * pred_region_pitch(w, block);
        if (cells <= kPredChunk && need <= kPredRegionCap) region = std::max(region, need);
    }
    const size_t head = ((static_cast<size_t>(block) * block + 3) & ~static_cast<size_t>(3)) * 4 + chunk * 8 +
                        2 * kMaxGroup * 8 + kMaxGroup * 4 + merged * 8;
    t.chunk = chunk;
    t.region = region;
    t.region_offset = static_cast<int>((head + 15) & ~static_cast<size_t>(15));
    return t.region_offset + static_cast<size_t>(region) * 4;
}

hipError_t launch_predictive(int block, const PredTable& t, int total_wg, size_t lds, hipStream_t s) {
    if (launch_fault()) return hipErrorLaunchFailure;
    switch (block) {
        case 1: allow_lds(&predictive_kernel<1>, lds); predictive_kernel<1><<<total_wg, 64, lds, s>>>(t); break;
        case 2: allow_lds(&predictive_kernel<2>, lds); predictive_kernel<2><<<total_wg, 64, lds, s>>>(t); break;
        case 4: allow_lds(&predictive_kernel<4>, lds); predictive_kernel<4><<<total_wg, 64, lds, s>>>(t); break;
        case 8: allow_lds(&predictive_kernel<8>, lds); predictive_kernel<8><<<total_wg, 64, lds, s>>>(t); break;
        case 12: allow_lds(&predictive_kernel<12>, lds); predictive_kernel<12><<<total_wg, 64, lds, s>>>(t); break;
        case 16: allow_lds(&predictive_kernel<16>, lds); predictive_kernel<16><<<total_wg, 64, lds, s>>>(t); break;
        case 32: allow_lds(&predictive_kernel<32>, lds); predictive_kernel<32><<<total_wg, 64, lds, s>>>(t); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int B>
static hipError_t launch_tfilter_b(const TFilterTable& t, int total_wg, size_t lds, bool wiener, hipStream_t s) {
    const int grid = t.cube ? std::min(total_wg, t.cube_slots) : total_wg;
    if (t.cube) lds = 0;
    if (wiener) {
        allow_lds(&filter_temporal_kernel<B, true>, lds);
        filter_temporal_kernel<B, true><<<grid, 64, lds, s>>>(t);
    } else {
        allow_lds(&filter_temporal_kernel<B, false>, lds);
        filter_temporal_kernel<B, false><<<grid, 64, lds, s>>>(t);
    }
    return hipGetLastError();
}

size_t filter_temporal_b8_lds_bytes(int group) { return 4 * static_cast<size_t>(group) * kTs * 4; }

// Dynamic LDS of filter_temporal_tile_kernel for per-slot geometries g[0..ns).
size_t filter_temporal_tile_lds_bytes(int group, const TileGeom* g, int ns, bool corner_planes) {
    size_t acc = 0, boxsum = 0;
    for (int t = 0; t < ns; ++t) {
        acc += static_cast<size_t>(g[t].aw) * g[t].ah;
        if (corner_planes) continue;  // no corner sums, no den box sum
        acc += static_cast<size_t>(g[t].aw - 7) * (g[t].ah - 7);
        boxsum = std::max(boxsum, static_cast<size_t>(g[t].ah - 7) * g[t].aw * 4);
    }
    acc = ((acc + 1) & ~static_cast<size_t>(1)) * 8;
    const size_t recs = static_cast<size_t>(g[0].tw) * g[0].th * group * sizeof(MatchRec) +
                        ((static_cast<size_t>(g[0].tw) * g[0].th + 3) & ~static_cast<size_t>(3)) * 4;
    const size_t tbuf = static_cast<size_t>(kTTileWaves) * group * kTs * 4;
    return recs + acc + std::max(tbuf, boxsum);
}

// Persistent grid: as many workgroups as the device holds at once (at most
// one per tile).
template <typename Kernel>
static hipError_t launch_persistent_ttile(Kernel k, const TTileTable& t, int tiles, size_t lds, hipStream_t s) {
    allow_lds(k, lds);
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 512, lds);
    if (e != hipSuccess) return e;
    const int grid = std::max(1, std::min(tiles, std::max(cus, 1) * std::max(per_cu, 1)));
    k<<<grid, 512, lds, s>>>(t, tiles);
    return hipGetLastError();
}

template <int G>
static hipError_t launch_ttile(const TTileTable& t, int tiles, size_t lds, bool wiener, hipStream_t s) {
    if (t.job[0].tiles[0].tw * t.job[0].tiles[0].th * G > 512) return hipErrorInvalidValue;
    return wiener ? launch_persistent_ttile(&filter_temporal_tile_kernel<G, true>, t, tiles, lds, s)
                  : launch_persistent_ttile(&filter_temporal_tile_kernel<G, false>, t, tiles, lds, s);
}

hipError_t launch_filter_temporal_tiles(int group, const TTileTable& t, int total_wg, size_t lds, bool wiener,
                                        hipStream_t s) {
    if (launch_fault()) return hipErrorLaunchFailure;
    switch (group) {
        case 8: return launch_ttile<8>(t, total_wg, lds, wiener, s);
        case 16: return launch_ttile<16>(t, total_wg, lds, wiener, s);
        case 32: return launch_ttile<32>(t, total_wg, lds, wiener, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_finish_temporal_tiles(const TFinishTable& t, int total_wg, hipStream_t s) {
    if (launch_fault()) return hipErrorLaunchFailure;
    // the corner tile and its horizontal sums when the weights come from corner planes
    int bmax = 0;
    for (int k = 0; k < t.n; ++k)
        if (t.job[k].corners) bmax = std::max(bmax, t.job[k].block);
    const size_t lds = bmax ? finish_tiles_lds_bytes(bmax) : 0;
    if (lds) {
        allow_lds(&finish_temporal_tiles_kernel, lds);
        finish_temporal_tiles_kernel<<<total_wg, 256, lds, s>>>(t);
        return hipGetLastError();
    }
    finish_temporal_tiles_kernel<<<total_wg, 256, 0, s>>>(t);
    return hipGetLastError();
}

template <int G>
static hipError_t launch_tfilter_b8(const TFilterTable& t, int total_wg, size_t lds, bool wiener, hipStream_t s) {
    if (wiener) {
        allow_lds(&filter_temporal_b8_kernel<G, true>, lds);
        filter_temporal_b8_kernel<G, true><<<total_wg, 256, lds, s>>>(t);
    } else {
        allow_lds(&filter_temporal_b8_kernel<G, false>, lds);
        filter_temporal_b8_kernel<G, false><<<total_wg, 256, lds, s>>>(t);
    }
    return hipGetLastError();
}

// fast_group 8, 16 or 32 (block 8): four groups per workgroup, register cube.
hipError_t launch_filter_temporal(int block, const TFilterTable& t, int total_wg, size_t lds, bool wiener,
                                  int fast_group, hipStream_t s) {
    if (launch_fault()) return hipErrorLaunchFailure;
    if (block == 8 && fast_group) {
        switch (fast_group) {
            case 8: return launch_tfilter_b8<8>(t, total_wg, lds, wiener, s);
            case 16: return launch_tfilter_b8<16>(t, total_wg, lds, wiener, s);
            case 32: return launch_tfilter_b8<32>(t, total_wg, lds, wiener, s);
            default: return hipErrorInvalidValue;
        }
    }
    switch (block) {
        case 1: return launch_tfilter_b<1>(t, total_wg, lds, wiener, s);
        case 2: return launch_tfilter_b<2>(t, total_wg, lds, wiener, s);
        case 4: return launch_tfilter_b<4>(t, total_wg, lds, wiener, s);
        case 8: return launch_tfilter_b<8>(t, total_wg, lds, wiener, s);
        case 12: return launch_tfilter_b<12>(t, total_wg, lds, wiener, s);
        case 16: return launch_tfilter_b<16>(t, total_wg, lds, wiener, s);
        case 32: return launch_tfilter_b<32>(t, total_wg, lds, wiener, s);
        default: return hipErrorInvalidValue;
    }
}

static int wnnm_pad(int group) { return group <= 8 ? 8 : group <= 16 ? 16 : 32; }

size_t wnnm_lds_bytes(int block, int group) {
    const size_t n = wnnm_pad(group), m = static_cast<size_t>(block) * block;
    return 4 * (n * (m + 1) + 2 * n * (n + 1) + 4 * n);
}

hipError_t launch_regularize(float* est, int est_stride, const float* noisy, int noisy_stride, int width, int height,
                             float delta, hipStream_t s) {
    if (launch_fault()) return hipErrorLaunchFailure;
    const dim3 grid((width + 63) / 64, (height + 3) / 4);
    regularize_kernel<<<grid, 256, 0, s>>>(est, est_stride, noisy, noisy_stride, width, height, delta);
    return hipGetLastError();
}

template <int N>
static hipError_t launch_gram_groups(const WnnmTable& t, int total_wg, size_t lds, bool ncsr, hipStream_t s) {
    if (ncsr) {
        allow_lds(&ncsr_group_kernel<N>, lds);
        ncsr_group_kernel<N><<<total_wg, 64, lds, s>>>(t);
    } else {
        allow_lds(&wnnm_group_kernel<N>, lds);
        wnnm_group_kernel<N><<<total_wg, 64, lds, s>>>(t);
    }
    return hipGetLastError();
}

hipError_t launch_wnnm(const WnnmTable& t, int total_wg, bool ncsr, hipStream_t s) {
    if (launch_fault()) return hipErrorLaunchFailure;
    int gmax = 0, bmax = 0;
    for (int j = 0; j < t.n; ++j) {
        if (t.job[j].group < 1 || t.job[j].group > 32 || t.job[j].block < 1 || t.job[j].block > 16)
            return hipErrorInvalidValue;
        gmax = std::max(gmax, t.job[j].group);
        bmax = std::max(bmax, t.job[j].block);
    }
    const size_t lds = wnnm_lds_bytes(bmax, gmax);
    switch (wnnm_pad(gmax)) {
        case 8: return launch_gram_groups<8>(t, total_wg, lds, ncsr, s);
        case 16: return launch_gram_groups<16>(t, total_wg, lds, ncsr, s);
        default: return launch_gram_groups<32>(t, total_wg, lds, ncsr, s);
    }
}

hipError_t upload_dct_tables(const float* host, size_t count) {
    if (count != sizeof(c_dct) / sizeof(float)) return hipErrorInvalidValue;
    return hipMemcpyToSymbol(HIP_SYMBOL(c_dct), host, sizeof(c_dct), 0, hipMemcpyHostToDevice);
}

}  // namespace nsship
