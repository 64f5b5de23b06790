// Host-side state shared by the filter instance (driver.cpp) and the serving
// loop (server.cpp): launch plans, per-frame device buffers, the batched
// stage runner.
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/nss_hip.h"
#include "bm3d_jobs.hpp"

namespace nsship {

// Device memory of the library (device_mem.cpp): hipMalloc / hipFree with a
// count of the live allocations; an armed test hook makes the n-th next
// allocation fail with hipErrorOutOfMemory.
hipError_t dev_alloc(void** p, size_t bytes);
template <typename T>
hipError_t dev_alloc(T** p, size_t bytes) {
    void* v = nullptr;
    const hipError_t e = dev_alloc(&v, bytes);
    *p = static_cast<T*>(v);
    return e;
}
void dev_free(void* p);
// True when an armed test hook refuses this kernel launch (the launch
// wrappers then return hipErrorLaunchFailure without launching).
bool launch_fault();

hipError_t launch_match(int block, const MatchTable& t, int total_wg, size_t lds, int kind, hipStream_t s);
hipError_t launch_filter(int block, const FilterTable& t, int total_wg, size_t lds, bool wiener, bool tiles,
                         int fast_group, int form, hipStream_t s);
hipError_t launch_finish(int block, const FinishTable& t, int total_wg, bool tiles, hipStream_t s);
// Output from frame-sized exact sums with per-pixel den (64 x 4 pixels per
// workgroup; tiles_x = ceil(width / 64)).
hipError_t launch_finish_sums(const FinishTable& t, int total_wg, hipStream_t s);
hipError_t launch_zero(const ZeroTable& t, int total_wg, hipStream_t s);
hipError_t upload_dct_tables(const float* host, size_t count);
size_t match_wave_lds_bytes(int block, int range, bool stage);
size_t match_run_lds_bytes(int block, int step, int range);
size_t match_mfma_lds_bytes(int step, int range, int group);
// The MFMA matcher serves B = 8, G <= 32 when its LDS fits.
bool match_mfma_fits(int block, int group, int step, int range, int width, int height);
// Workgroups of a match launch of `kind` over one plane's grid.
int match_workgroups(int kind, const Grid& g);
// Process-wide matcher choice (nss_hip_set_match_path): 0 automatic, 1 the
// FP32 matchers only, 2 the MFMA matcher wherever it fits.
int match_path();
// Automatic choice: the MFMA matcher also where the run matcher fits, when
// the window is wide against the step and the stage has at least
// kMfmaMinRefs references.
bool match_mfma_pays(int step, int range, int group);
constexpr long long kMfmaMinRefs = 16384;
size_t filter_atomic_lds_bytes(int block, int group, bool wiener);
size_t filter_tile_lds_bytes(int block, int group, bool wiener, const TileGeom& t);
size_t filter_tile_b8_lds_bytes(int group, const TileGeom& t);
size_t filter_tile_mfma_lds_bytes(int group, const TileGeom& t);
// Sets the table's LDS layout (chunk, staged-window capacity) from its jobs;
// returns the bytes per workgroup.
size_t predictive_table_lds(int block, PredTable& t, int radius);
hipError_t launch_predictive(int block, const PredTable& t, int total_wg, size_t lds, hipStream_t s);
hipError_t launch_filter_temporal(int block, const TFilterTable& t, int total_wg, size_t lds, bool wiener,
                                  int fast_group, hipStream_t s);
size_t filter_temporal_b8_lds_bytes(int group);
// corner_planes: the weights go to frame corner planes (rolling output), so
// the workgroup holds no corner sums.
size_t filter_temporal_tile_lds_bytes(int group, const TileGeom* g, int ns, bool corner_planes);
hipError_t launch_filter_temporal_tiles(int group, const TTileTable& t, int total_wg, size_t lds, bool wiener,
                                        hipStream_t s);
hipError_t launch_finish_temporal_tiles(const TFinishTable& t, int total_wg, hipStream_t s);
hipError_t launch_vaggregate(const VAggJob& jb, hipStream_t s);
// WNNM group shrinkage (group <= 32, block <= 16).
size_t wnnm_lds_bytes(int block, int group);
// ncsr: the NCSR code shrink on the same eigenbasis instead of WNNM's.
hipError_t launch_wnnm(const WnnmTable& t, int total_wg, bool ncsr, hipStream_t s);
hipError_t launch_regularize(float* est, int est_stride, const float* noisy, int noisy_stride, int width, int height,
                             float delta, hipStream_t s);
// The matcher kind for one plane (3 MFMA, 2 run, 1 wave with the window in
// LDS, 0 wave from global) and its LDS bytes.
int choose_match_kind(int block, int step, int range, int group, int width, int height, long long nref,
                      size_t* lds);

extern thread_local std::string g_last_error;
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

// Restores the caller's current device on scope exit.
struct DeviceGuard {
    int prev = -1;
    hipError_t err = hipSuccess;
    explicit DeviceGuard(int device) {
        err = hipGetDevice(&prev);
        if (err == hipSuccess && prev != device) err = hipSetDevice(device);
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

constexpr size_t kLdsLimit = 160 * 1024;  // LDS per workgroup on gfx950
// Tile aggregation aims at two workgroups per CU; a one-reference tile may
// take the whole LDS.
constexpr size_t kTileLdsBudget = 80 * 1024;
constexpr size_t kRunLdsBudget = 64 * 1024;

int ensure_tables(int device);
Grid make_grid(int width, int height, int block, int step);
TileGeom make_tiles(const Grid& g, int tw, int th, int block, int range);
bool allowed_block(int b);
bool allowed_group(int g);

struct StagePlane {
    float sigma;
    int block, group, step, range, ps_num, ps_range;
};

// Launch plan of one block-size bucket of a stage.
struct Bucket {
    int block = 0;
    int planes[3]{};
    int nplanes = 0;
    int match_kind = 0;  // 2 run, 1 wave + LDS window, 0 wave + global window
    size_t match_lds = 0;
    bool tiles = false;  // tile aggregation, else frame-sized atomics
    TileGeom geom[3]{};  // per plane of the bucket (tile path)
    size_t filter_lds = 0;
    int fast_group = 0;  // block 8 with one group size in {8, 16, 32}: the register kernel
    int form = 0;        // 1 bf16x3, 2 fp16: the MFMA kernel (block 8, groups of 8 or 16)
    int cube_floats = 0;  // atomic path with the group cube in global memory (floats per cube); 0: LDS
};

struct Stage {
    StagePlane plane[3];
    bool wiener;
    Bucket bucket[7];
    int nbuckets;
};

// Event pairs around launches while tracing is on (one vector per kind).
// Calls on one instance may come from several threads: mu guards the pool
// and the spans.
struct Tracer {
    std::mutex mu;
    bool on = false;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> spans[NSS_HIP_KERNEL_KINDS];

    hipEvent_t take() {
        if (used == pool.size()) {
            hipEvent_t e = nullptr;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            pool.push_back(e);
        }
        return pool[used++];
    }
    void reset() {
        used = 0;
        for (auto& v : spans) v.clear();
    }
    ~Tracer() {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
};

// Device buffers one radius-0 frame needs while it is processed: match lists,
// tile partials (or frame-sized sums on the atomic path) and the basic
// estimate of a two-stage filter.
struct FrameBuffers {
    unsigned long long* num[3]{};  // atomic path sums
    unsigned long long* den[3]{};
    unsigned long long* part[3]{};  // tile path partials: num then den (int64, exact)
    MatchRec* matches[3]{};
    int* counts[3]{};
    float* basic[3]{};
    float* cube = nullptr;  // cube_slots global group cubes (the instance's cube_floats each)
};

// One of an instance's num_streams frame slots: a frame's buffers and the
// event recorded after the last work queued on them, which the next lessee's
// stream waits for (so calls on different HIP streams never overlap on one
// slot's buffers).
struct FrameSlot {
    FrameBuffers buf;
    hipEvent_t done = nullptr;
    bool used = false;  // done has been recorded
};

}  // namespace nsship

struct nss_hip_bm3d {
    nss_hip_bm3d_params p{};
    nss_hip_video_info vi{};
    nsship::Tracer tracer;
    int device = 0;
    int nplanes = 0;
    int pw[3]{}, ph[3]{};
    nsship::Stage stages[2]{};
    int nstages = 0;
    size_t part_elems[3]{};  // a plane's tile numerator partials, int64 entries (the group weights follow)
    size_t cube_floats = 0;  // largest global group cube of any bucket (0: every cube fits the LDS)
    int cube_slots = nsship::kCubeSlots;  // global cubes per frame slot (fewer under a memory limit)
    // memory_limit_mb below the normal plan's need (radius 0): every stage
    // runs on frame-sized int64 sums, matching and filtering the grid rows in
    // this many batches (0: the normal plan)
    int batches = 0;
    // num_streams frame slots (radius 0; one for temporal instances). A
    // process call leases the slot free the longest for the time it takes to
    // queue the frame, sleeping while all are leased
    // (src/cuda/runtime/stream_pool.hpp:1-4 SlotPool::acquire).
    std::vector<nsship::FrameSlot> slots;
    std::mutex slot_mu;
    std::condition_variable slot_cv;
    std::deque<int> free_slots;
    // Temporal push / pop: one call at a time, each ordered after the last on
    // the device (roll_done), whatever stream it comes on.
    std::mutex roll_mu;
    hipEvent_t roll_done = nullptr;
    bool roll_used = false;
    // Rolling temporal state (radius > 0): frames enter in clip order and
    // stay on the device in rings (frame f at slot f % ring); each stage keeps
    // frame-sized int64 sums for the 2r+1 frames its current centre touches.
    struct Rolling {
        int N = 0, r = 0, pushed = 0;
        int src_n = 0, ref_n = 0, basic_n = 0, out_n = 0, acc_n = 0;
        float* src[3]{};
        float* ref[3]{};
        float* basic[3]{};
        float* out[3]{};
        unsigned long long* num[2][3]{};
        unsigned long long* den[2][3]{};
        int next_center[2]{};
        int finished[2]{};
        int out_head = 0;
        // A push failed part-way: a centre's contributions may be in the sums
        // already, so no later push is served (frames finished before it
        // still pop).
        bool failed = false;
        // Legacy temporal output: the last stage emits, per centre, its 2r+1
        // num/den slice pairs (out_rows = 2 (2r + 1) plane heights per output
        // frame) instead of finished frames (out_rows = 1).
        bool legacy = false;
        int out_rows = 1;
        // Tile path of a stage (block 8 register kernel, radius <= 2): the
        // tile partials of the last 2r+1 centres per plane, centre c at
        // part + (c % (2r+1)) * part_center.
        bool tiles[2]{};
        int tw[2]{}, th[2]{};
        float* part[2][3]{};
        size_t part_center[2][3]{};
        size_t tile_lds[2]{};
    } roll;
};

namespace nsship {

// One radius-0 frame of a batch: its instance (plan), planes, buffers and
// effective sigma per stage and plane.
struct FrameRef {
    nss_hip_bm3d* h = nullptr;
    const float* src[3]{};
    int src_stride[3]{};
    const float* ref[3]{};
    int ref_stride[3]{};
    float* dst[3]{};
    int dst_stride[3]{};
    FrameBuffers buf;
    float sigma[2][3]{};
};

// Device bytes of one frame's buffers for instance h (radius 0), and of all
// the instance holds (frame buffers plus rolling rings and sums).
size_t frame_buffer_bytes(const nss_hip_bm3d* h);
// Frame slots an instance holds: num_streams at radius 0, one otherwise.
int frame_slot_count(const nss_hip_bm3d* h);
size_t instance_bytes(const nss_hip_bm3d* h);
hipError_t alloc_frame_buffers(const nss_hip_bm3d* h, FrameBuffers& fb);
void free_frame_buffers(FrameBuffers& fb);
// The same buffer set laid out from base (each buffer 256-byte aligned);
// returns the bytes it spans. base null: the size only.
size_t carve_frame_buffers(const nss_hip_bm3d* h, char* base, FrameBuffers& fb);
// The host part of an instance: checks, plane sizes, stage plans, partial
// sizes and the rolling plan; no device memory.
int plan_instance(nss_hip_bm3d* h, const nss_hip_bm3d_params* params, const nss_hip_video_info* info, int device);
// Fits a radius-0 instance's frame buffers into `bytes` per frame slot: the
// normal plan when it fits, else the batched plan with the fewest batches
// that fits (same output bits). False when none does (the MFMA forms have
// no batched plan).
bool fit_frame_buffers(nss_hip_bm3d* h, size_t bytes);
FrameRef frame_ref(nss_hip_bm3d* h, const float* const* src, const int* src_stride, const float* const* ref,
                   const int* ref_stride, float* const* dst, const int* dst_stride);
int run_batch(const FrameRef* frames, int nframes, Tracer* tr, hipStream_t s);

}  // namespace nsship
