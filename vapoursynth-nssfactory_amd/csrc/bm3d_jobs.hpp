// Launch descriptors shared by the HIP kernels and the host driver.
//
// One launch serves up to kMaxJobs planes (of one frame, or of several frames
// in a batch) that share a block size: the grid is the concatenation of each
// job's workgroups, and a workgroup finds its job by scanning wg_begin. The
// tables travel as kernel arguments (no descriptor copies on the stream).
#pragma once

#include <stdint.h>

namespace nsship {

inline constexpr int kMaxJobs = 16;
inline constexpr int kMaxGroup = 64;   // include/nss/params.hpp:39 kBmMaxGroup
inline constexpr int kMaxRadius = 16;  // include/nss/params.hpp:44 kBmMaxRadius
inline constexpr int kMaxSlots = 2 * kMaxRadius + 1;
inline constexpr int kRunRefs = 8;     // references per workgroup of the run matcher
inline constexpr int kRunGather = 128; // keys at or below a reference's threshold held for ranking

// A match record: x, y, temporal slot, and the distance's bits (the float
// SSD; non-finite distances collapse to +inf bits, so they rank equal and
// fall back to the position order).
struct MatchRec {
    int x, y, t;
    uint32_t dist;
};

// Reference grid of one plane: positions (i*step, j*step) clamped to
// (width - block, height - block) (src/gpuhost/geometry.hpp:17-26).
struct Grid {
    int nx, ny, step;
    int row0 = 0;  // first grid row (a batch of rows of a larger grid: nx x ny references from row row0)
};

struct MatchJob {
    const float* plane;  // the plane searched (the reference's own frame)
    int width, height, stride;
    Grid grid;
    int range;   // bm_range
    int group;   // configured group size; the record stride
    int t0;      // slot recorded for spatial matches
    int wg_begin;
    MatchRec* out;   // [nref][group]
    int* counts;     // [nref]
};

struct MatchTable {
    int n;
    MatchJob job[kMaxJobs];
};

// Tiles of the reference grid for LDS aggregation: tile (tx, ty) holds the
// references i in [tx*tw, tx*tw + tw), j in [ty*th, ty*th + th). Its sums
// cover the rectangle anchored at (min(tx*tw*step, width - block) - range,
// min(ty*th*step, height - block) - range) of aw x ah pixels, which holds
// every patch its references can match.
struct TileGeom {
    int tw, th;        // references per tile
    int ntx, nty;      // tiles
    int aw, ah;        // accumulation rectangle
    int pitch;         // row pitch of a tile's float partials: a multiple of 4,
                       // >= aw + 6; column x - ax0 + (ax0 & 3) holds pixel x
    __host__ __device__ int ax0(int tx, int width, int block, int step, int range) const {
        const int x = tx * tw * step;
        return (x < width - block ? x : width - block) - range;
    }
    __host__ __device__ int ay0(int ty, int height, int block, int step, int range) const {
        const int y = ty * th * step;
        return (y < height - block ? y : height - block) - range;
    }
};

struct FilterJob {
    const float* src;   // filtered plane (noisy)
    const float* wref;  // Wiener reference plane (the basic estimate); null in the hard stage
    int width, height, stride, wref_stride;
    Grid grid;
    int range;
    int nref;
    int group;
    float sigma;        // effective sigma
    int wg_begin;
    const MatchRec* matches;
    const int* counts;
    // Frame-sized fixed-point sums (the atomic path): numerator at every
    // pixel, weight at each patch's top-left corner, 2^32 units.
    unsigned long long* num;
    unsigned long long* den;
    // Per-tile partial numerator sums (the tile path): [tile][ah][pitch]
    // int64, 2^32 units, exact (the finish adds them exactly), and each
    // group's weight (2^32 units) at its record index, from which the finish
    // rebuilds the weight sums.
    TileGeom tiles;
    unsigned long long* part_num;
    unsigned long long* gweight;
    // A batch of tile rows (memory-limited plan): workgroup w filters tile
    // tile0 + w of the frame and writes partials [w]; the records hold the
    // grid rows from grid.row0 on. Both 0 for a whole frame.
    int tile0;
};

// Groups whose cube does not fit the LDS (block 32 with group 64, or group
// 32 with the Wiener reference) keep it in global memory: kCubeSlots
// workgroups each walk the groups with a cube of cube_floats of its own.
inline constexpr int kCubeSlots = 512;

struct FilterTable {
    int n;
    FilterJob job[kMaxJobs];
    float* cube;      // null: one group per workgroup, the cube in LDS
    int cube_floats;  // floats per workgroup's cube
    int cube_slots;   // cubes at `cube` (at most kCubeSlots workgroups)
};

struct FinishJob {
    const unsigned long long* num;  // atomic path
    const unsigned long long* den;
    const long long* exact_num;     // spatial tile path: exact int64 numerator partials
    // ... and the weights: every record j < count of the references of
    // `grid` (rows grid.row0 .. grid.row0 + grid.ny - 1, records at
    // ((row - grid.row0) * grid.nx + column) * group) adds gweight of its
    // reference at the patch's top-left corner, box-summed block x block.
    const MatchRec* matches;
    const int* counts;
    const unsigned long long* gweight;
    Grid grid;
    int group;
    const float* part_num;          // temporal tile path: float slot partials
    const float* part_den;
    TileGeom tiles;
    int block, step, range;
    const float* src;   // fallback where no patch landed
    float* dst;
    // Slice output (legacy temporal): when set, dst receives the summed num
    // and dst_den the summed den (no division, no fallback).
    float* dst_den;
    int width, height, sstride, dstride;
    int tiles_x;        // output blocks per row (64 x 16 tile path, 16 x 16 atomic path)
    int wg_begin;
    // A batch of tile rows [ty0, ty1) (memory-limited plan): the exact
    // partials of its tiles (tile ty at index ty - ty0) are added into the
    // frame sums acc_num / acc_den over pixel rows from py0 on, and no output
    // is written. acc_num null: a whole frame.
    long long* acc_num;
    long long* acc_den;
    int ty0, ty1, py0;
};

struct FinishTable {
    int n;
    FinishJob job[kMaxJobs];
};

// Frames of a clip held on the device: frame f at frames + (f % ring) * frame_stride.
struct FrameRing {
    const float* frames;
    long long frame_stride;
    int ring;
    __host__ __device__ const float* frame(int f) const {
        return frames + static_cast<long long>(f % ring) * frame_stride;
    }
};

// Temporal slots of centre frame c: slot t is frame
// c + clamp(t - radius, -c, nframes - 1 - c) (src/frontend/temporal.hpp:19-21);
// slots outside [max(0, radius - c), radius + min(radius + 1, nframes - c))
// repeat an edge frame and take no part in matching
// (src/host/filter_bm3d.cpp:96-97).
struct Slots {
    int center, nframes, radius;
    __host__ __device__ int frame(int t) const {
        int d = t - radius;
        if (d < -center) d = -center;
        if (d > nframes - 1 - center) d = nframes - 1 - center;
        return center + d;
    }
    __host__ __device__ int valid_begin() const { return radius - center > 0 ? radius - center : 0; }
    __host__ __device__ int valid_end() const {
        return radius + (radius + 1 < nframes - center ? radius + 1 : nframes - center);
    }
};

// Predictive temporal search of one plane, after the spatial search on the
// centre frame (src/cpu/bm/matcher.hpp:319-373 collect_temporal).
struct PredJob {
    FrameRing match;  // frames searched
    int width, height, stride;
    int nref;
    int group, ps_num, ps_range;
    Slots slots;
    int wg_begin;
    MatchRec* recs;   // in: spatial matches (t = radius); out: the temporal group
    int* counts;
};

// Predictive search LDS limits: candidate keys per pass, and the staged
// windows' floats (each window's (2 ps + 1 + B - 1)^2 pixels at a row pitch
// that covers the last strip of three candidates).
inline constexpr int kPredChunk = 1024;
inline constexpr int kPredRegionCap = 8192;
constexpr int pred_region_pitch(int w, int block) { return (3 * ((w + 2) / 3) + block - 1) | 1; }

struct PredTable {
    int n;
    int chunk;   // candidate keys held in LDS per pass (a multiple of 64)
    int region;  // floats of LDS for the staged windows (0: windows read from the frame)
    int region_offset;  // bytes from the start of LDS to the staged windows
    PredJob job[kMaxJobs];
};

// Temporal group filter: patches from their slot's frame, sums into the
// accumulator of the frame they came from (a ring of frame-sized int64 sums).
struct TFilterJob {
    FrameRing src, wref;
    int width, height, stride, wref_stride;
    int nref, group;
    float sigma;
    Slots slots;
    int wg_begin;
    const MatchRec* matches;
    const int* counts;
    unsigned long long* num;  // frame f's sums at num + (f % acc_ring) * width * height
    unsigned long long* den;
    int acc_ring;
};

struct TFilterTable {
    int n;
    TFilterJob job[kMaxJobs];
    float* cube;  // as FilterTable::cube
    int cube_floats;
    int cube_slots;
};

// Temporal tile path (block 8, radius <= 2): every slot's contributions of a
// tile of references accumulate in LDS over that slot's apron (range +
// |t - r| * ps_range around the tile's references, the predictive search's
// reach), then go out as float partials of the centre.
inline constexpr int kMaxTileSlots = 5;

struct TTileJob {
    FrameRing src, wref;
    int width, height, stride, wref_stride;
    Grid grid;
    int nslots;
    TileGeom tiles[kMaxTileSlots];  // per slot; tw, th, ntx, nty shared
    int range[kMaxTileSlots];
    float* part;                    // this centre's partials
    long long slot_off[kMaxTileSlots];  // floats: slot t's num at part + slot_off[t], den after it
    // Rolling output: each group's weight goes, by an int64 device atomic,
    // to the top-left corner of each of its patches in the frame-sized corner
    // plane of the patch's frame (frame f at corners + (f % acc_ring) W H),
    // and the den partials are not written. Null: den partials (legacy).
    unsigned long long* corners;
    int acc_ring;
    int nref, group;
    float sigma;
    Slots slots;
    int wg_begin;
    const MatchRec* matches;
    const int* counts;
};

struct TTileTable {
    int n;
    TTileJob job[3];
};

// Frame n's finish from the partials of the centres that reach it.
struct TSource {
    const float* num;
    const float* den;
    TileGeom tiles;
    int range;
};

struct TFinishJob {
    TSource src[kMaxTileSlots];
    int nsrc;
    int block, step;
    // The frame's corner plane of group weights (rolling output): den is its
    // B x B box sum, exact, instead of the sources' den partials.
    const unsigned long long* corners;
    const float* fallback;  // the source frame
    float* dst;
    int width, height;
    int tiles_x;  // 64 x 16 output blocks per row
    int wg_begin;
};

// Reduction of legacy temporal slices (nss.VAggregate): contribution k's
// num and den rows of one plane, summed in k order.
inline constexpr int kVAggMax = 2 * kMaxRadius + 1;
struct VAggJob {
    const float* num[kVAggMax];
    const float* den[kVAggMax];
    int stride[kVAggMax];
    int count;
    const float* src;
    float* dst;
    int width, height, sstride, dstride;
};

struct TFinishTable {
    int n;
    TFinishJob job[3];
};

// WNNM group shrinkage (src/cuda/wnnm/kernels.hpp:23-36): one wave per
// reference of one plane; patches from src at the matched positions, sums in
// the atomic path's frame-sized fixed-point layout (den at patch corners).
struct WnnmJob {
    const float* src;
    int width, height, stride;
    int nref, group, block;
    float sigma;        // user sigma / 255
    int residual, adaptive;
    int wg_begin;
    const MatchRec* matches;  // nref * group slots
    const int* counts;
    unsigned long long* num;
    unsigned long long* den;
    // Temporal (radius > 0): patch j comes from the frame of its slot t =
    // rec.t, ring.frame(slots.frame(t)), and adds into num / den + t * sum_slot.
    int temporal;
    FrameRing ring;
    Slots slots;
    long long sum_slot;
};

struct WnnmTable {
    int n;
    WnnmJob job[kMaxJobs];
};

struct ZeroJob {
    unsigned long long* p;
    long long count;
    int wg_begin;
};

struct ZeroTable {
    int n;
    ZeroJob job[2 * kMaxJobs];
};

}  // namespace nsship
