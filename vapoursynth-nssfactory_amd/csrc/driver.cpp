// Host driver behind the C ABI: filter instances, per-frame stage
// orchestration and the batched matcher entry point.
//
// One frame (src/host/filter_bm3d.cpp:359-494 bm3dGetFrame): for every stage
// (basic then Wiener when final=1, src/frontend/bm3d_args.cpp:186-221), planes
// with sigma 0 are copied; the others are grouped by block size, and each
// group runs on the caller's stream: match, filter + aggregate, finish (plus
// a clear of the frame-sized sums on the atomic path). The launch plan of
// every (stage, block size) is fixed at creation. The basic estimate stays
// on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "instance.hpp"

namespace nsship {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(NSS_HIP_RESOURCE_FAILURE, std::string("nss_hip: ") + what + ": " + hipGetErrorString(e));
}

// Per device, once: the DCT matrices in constant memory.
std::mutex g_init_mutex;
bool g_dct_ready[64] = {};

int ensure_tables(int device) {
    std::lock_guard<std::mutex> lock(g_init_mutex);
    if (device < 0 || device >= 64) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: device index out of range");
    if (g_dct_ready[device]) return NSS_HIP_OK;
    std::vector<float> t;
    const double pi = 3.14159265358979323846264338327950288;
    for (int n : {2, 4, 8, 12, 16, 32})
        for (int k = 0; k < n; ++k)
            for (int i = 0; i < n; ++i)
                t.push_back(static_cast<float>(std::sqrt((k == 0 ? 1.0 : 2.0) / n) * std::cos(pi * (i + 0.5) * k / n)));
    const hipError_t e = upload_dct_tables(t.data(), t.size());
    if (e != hipSuccess) return hip_fail(e, "DCT table upload");
    g_dct_ready[device] = true;
    return NSS_HIP_OK;
}

int choose_match_kind(int block, int step, int range, int group, int width, int height, long long nref,
                      size_t* lds) {
    // The run matcher is built for the BM3D block sizes up to 12; other sizes
    // (the group filters' 3, 5-7, 9-11, 13-15) take the wave matcher.
    const bool run_block = block == 1 || block == 2 || block == 4 || block == 8 || block == 12;
    const bool run = run_block && match_run_lds_bytes(block, step, range) <= kRunLdsBudget;
    if (match_path() != 1 && match_mfma_fits(block, group, step, range, width, height) &&
        (match_path() == 2 || !run || (match_mfma_pays(step, range, group) && nref >= kMfmaMinRefs))) {
        *lds = match_mfma_lds_bytes(step, range, group);
        return 3;
    }
    if (run) {
        *lds = match_run_lds_bytes(block, step, range);
        return 2;
    }
    if (match_wave_lds_bytes(block, range, true) <= kLdsLimit) {
        *lds = match_wave_lds_bytes(block, range, true);
        return 1;
    }
    *lds = match_wave_lds_bytes(block, range, false);
    return 0;
}

Grid make_grid(int width, int height, int block, int step) {
    Grid g{0, 0, step};
    if (width < block || height < block || step < 1) return g;
    g.nx = (width - block + step - 1) / step + 1;
    g.ny = (height - block + step - 1) / step + 1;
    return g;
}

TileGeom make_tiles(const Grid& g, int tw, int th, int block, int range) {
    TileGeom t{};
    t.tw = tw;
    t.th = th;
    t.ntx = (g.nx + tw - 1) / tw;
    t.nty = (g.ny + th - 1) / th;
    t.aw = (tw - 1) * g.step + 2 * range + block;
    t.pitch = (t.aw + 6 + 3) & ~3;
    t.ah = (th - 1) * g.step + 2 * range + block;
    return t;
}

std::atomic<int> g_match_path{0};

// The MFMA matcher scores every candidate of a 4 x 4 block's union window
// (side 3 step + 2 range + 1) for all 16 references; the run matcher scores
// each reference's own window. Measured on 1080p planes (block 8, group 8
// and 16), the MFMA matcher is faster once the window side exceeds about
// 3.3 steps (range >= 13 at step 8, >= 10 at step 4 or 3) and the frame has
// enough references to fill the device (a 256 x 256 frame at step 3 does
// not). The rule holds at group 32 as well (1080p YUV420, step 3, range 19,
// groups 16 / 32: 7.1 ms per frame with both stages on the MFMA matcher
// against 8.7 with the run matcher on the group-32 stage).
bool match_mfma_pays(int step, int range, int group) {
    return group <= 32 && range >= 10 && 3 * (2 * range + 1) > 10 * step;
}

int match_path() { return g_match_path.load(std::memory_order_relaxed); }

bool allowed_block(int b) { return b == 1 || b == 2 || b == 4 || b == 8 || b == 12 || b == 16 || b == 32; }
bool allowed_group(int g) { return g == 1 || g == 2 || g == 4 || g == 8 || g == 16 || g == 32 || g == 64; }

}  // namespace nsship

using namespace nsship;

namespace {

int check_stage_plane(const StagePlane& s, int w, int h) {
    if (!allowed_block(s.block) || !allowed_group(s.group) || s.block > w || s.block > h || s.step < 1 ||
        s.step > s.block || s.range < 1 || s.range > 64)
        return fail(NSS_HIP_INVALID_INPUT, "nss_hip: invalid resolved BM3D parameters");
    return NSS_HIP_OK;
}

// Tile shapes (references per tile, x by y) in the order the plan tries them;
// the autotuner times every one that fits.
constexpr int kTileShapes[][2] = {{8, 4}, {4, 4}, {4, 2}, {8, 2}, {16, 2}, {2, 2}, {1, 1}};

// Sets bucket bk to tiles of tw x th references if every plane's tile fits the
// LDS budget (the whole LDS for a one-reference tile); false otherwise.
bool apply_tile_shape(const nss_hip_bm3d* h, const Stage& st, Bucket& bk, int tw, int th) {
    const int b = bk.block;
    size_t lds = 0;
    TileGeom g[3]{};
    for (int k = 0; k < bk.nplanes; ++k) {
        const int i = bk.planes[k];
        const StagePlane& sp = st.plane[i];
        g[k] = make_tiles(make_grid(h->pw[i], h->ph[i], b, sp.step), tw, th, b, sp.range);
        lds = std::max(lds, bk.form         ? filter_tile_mfma_lds_bytes(sp.group, g[k])
                            : bk.fast_group ? filter_tile_b8_lds_bytes(sp.group, g[k])
                                            : filter_tile_lds_bytes(b, sp.group, st.wiener, g[k]));
    }
    const size_t budget = (tw == 1 && th == 1) ? kLdsLimit : kTileLdsBudget;
    if (lds > budget) return false;
    // A tile of one reference writes (and the finish reads back) its whole
    // accumulation rectangle: beyond four group cubes of pixels, the atomic
    // path's direct sums move less.
    if (tw == 1 && th == 1)
        for (int k = 0; k < bk.nplanes; ++k)
            if (static_cast<size_t>(g[k].aw) * g[k].ah > 4u * st.plane[bk.planes[k]].group * b * b) return false;
    bk.tiles = true;
    bk.filter_lds = lds;
    for (int k = 0; k < bk.nplanes; ++k) bk.geom[k] = g[k];
    return true;
}

// Chooses the kernels of every (stage, block size) bucket.
int plan_stage(nss_hip_bm3d* h, Stage& st) {
    st.nbuckets = 0;
    for (int b : {1, 2, 4, 8, 12, 16, 32}) {
        Bucket bk;
        bk.block = b;
        for (int i = 0; i < h->nplanes; ++i)
            if (st.plane[i].sigma != 0.f && st.plane[i].block == b) bk.planes[bk.nplanes++] = i;
        if (!bk.nplanes) continue;
        size_t run_lds = 0, wave_lds = 0, mfma_lds = 0;
        bool mfma = match_path() != 1, mfma_pays = true;
        long long nref = 0;
        for (int k = 0; k < bk.nplanes; ++k) {
            const StagePlane& sp = st.plane[bk.planes[k]];
            const int i = bk.planes[k];
            run_lds = std::max(run_lds, match_run_lds_bytes(b, sp.step, sp.range));
            wave_lds = std::max(wave_lds, match_wave_lds_bytes(b, sp.range, true));
            mfma_lds = std::max(mfma_lds, match_mfma_lds_bytes(sp.step, sp.range, sp.group));
            mfma &= match_mfma_fits(b, sp.group, sp.step, sp.range, h->pw[i], h->ph[i]);
            mfma_pays &= match_mfma_pays(sp.step, sp.range, sp.group);
            const Grid gr = make_grid(h->pw[i], h->ph[i], b, sp.step);
            nref += static_cast<long long>(gr.nx) * gr.ny;
        }
        const bool run = b <= 12 && run_lds <= kRunLdsBudget;
        if (mfma && (match_path() == 2 || !run || (mfma_pays && nref >= kMfmaMinRefs))) {
            bk.match_kind = 3;
            bk.match_lds = mfma_lds;
        } else if (run) {
            bk.match_kind = 2;
            bk.match_lds = run_lds;
        } else if (wave_lds <= kLdsLimit) {
            bk.match_kind = 1;
            bk.match_lds = wave_lds;
        } else {
            bk.match_kind = 0;
            bk.match_lds = 0;
            for (int k = 0; k < bk.nplanes; ++k)
                bk.match_lds = std::max(bk.match_lds, match_wave_lds_bytes(b, st.plane[bk.planes[k]].range, false));
        }
        int fast = 0;
        if (b == 8) {
            fast = st.plane[bk.planes[0]].group;
            for (int k = 1; k < bk.nplanes; ++k)
                if (st.plane[bk.planes[k]].group != fast) fast = 0;
            if (fast != 8 && fast != 16 && fast != 32) fast = 0;
        }
        // (the MFMA kernel keeps patch positions as 16-bit x and y)
        bool small = true;
        for (int k = 0; k < bk.nplanes; ++k) small &= h->pw[bk.planes[k]] <= 65535 && h->ph[bk.planes[k]] <= 65535;
        const int form = ((fast == 8 || fast == 16) && small) ? h->p.transform : 0;
        bk.fast_group = fast;
        bk.form = form;
        // The FP32 register filter is latency bound, so it starts from small
        // tiles (more workgroups per CU): on 1080p YUV420 (B8, G8) 4 x 4
        // references filter fastest in both stages (hard threshold 4 x 4
        // 46.5, 4 x 2 47.1, 8 x 2 48.0, 8 x 4 49.3 us; stage averages with
        // the numerator-only partials). The autotuner refines per input.
        bool placed = false;
        if (fast && !form) placed = apply_tile_shape(h, st, bk, 4, 4);
        for (const auto& sh : kTileShapes)
            if (placed || (placed = apply_tile_shape(h, st, bk, sh[0], sh[1]))) break;
        if (!bk.tiles) {
            bk.fast_group = bk.form = 0;
            int gmax = 0;
            for (int k = 0; k < bk.nplanes; ++k) {
                gmax = std::max(gmax, st.plane[bk.planes[k]].group);
                bk.filter_lds =
                    std::max(bk.filter_lds, filter_atomic_lds_bytes(b, st.plane[bk.planes[k]].group, st.wiener));
            }
            // A cube beyond the LDS (block 32 with group 64, or group 32 with
            // the Wiener reference) lives in global memory, as the
            // reference's device path keeps it (src/cuda/bm3d/kernels.cu:248-343).
            if (bk.filter_lds > kLdsLimit) {
                bk.filter_lds = 0;
                bk.cube_floats = ((st.wiener ? 2 : 1) * gmax * (b * b + 1) + 63) & ~63;
                h->cube_floats = std::max(h->cube_floats, static_cast<size_t>(bk.cube_floats));
            }
        }
        st.bucket[st.nbuckets++] = bk;
    }
    return NSS_HIP_OK;
}

void free_buffers(nss_hip_bm3d* h) {
    for (auto& sl : h->slots) {
        free_frame_buffers(sl.buf);
        if (sl.done) (void)hipEventDestroy(sl.done);
        sl.done = nullptr;
    }
    h->slots.clear();
    if (h->roll_done) (void)hipEventDestroy(h->roll_done);
    h->roll_done = nullptr;
    auto& R = h->roll;
    for (int i = 0; i < 3; ++i) {
        for (void* p : {static_cast<void*>(R.src[i]), static_cast<void*>(R.ref[i]), static_cast<void*>(R.basic[i]),
                        static_cast<void*>(R.out[i]), static_cast<void*>(R.num[0][i]), static_cast<void*>(R.den[0][i]),
                        static_cast<void*>(R.num[1][i]), static_cast<void*>(R.den[1][i]), static_cast<void*>(R.part[0][i]),
                        static_cast<void*>(R.part[1][i])})
            if (p) dev_free(p);
        R.src[i] = R.ref[i] = R.basic[i] = R.out[i] = nullptr;
        R.num[0][i] = R.den[0][i] = R.num[1][i] = R.den[1][i] = nullptr;
        R.part[0][i] = R.part[1][i] = nullptr;
    }
}

// Per-plane buffer needs of an instance: match records and counts (the
// largest stage), tile numerator partials (int64 entries) and the group
// weights after them (one per reference of the plane's grid), frame-sized
// sums (atomic path), basic estimate.
struct PlaneNeeds {
    size_t recs = 0, nref = 0, part = 0, gw = 0;
    bool atomic = false, basic = false;
};

// Grid rows one batch of a memory-limited plan matches on plane i in stage
// st: the rows of a batch of tile rows for a tile bucket (plane i's own
// geometry, or plane 0's with chroma, where plane 0 holds the groups),
// ceil(ny / batches) otherwise.
int batch_grid_rows(const nss_hip_bm3d* h, const Stage& st, int i, const Grid& g) {
    for (int bi = 0; bi < st.nbuckets; ++bi) {
        const Bucket& bk = st.bucket[bi];
        for (int k = 0; k < bk.nplanes; ++k)
            if (bk.planes[k] == i && bk.tiles) {
                const TileGeom& tg = bk.geom[k];
                return std::min(g.ny, (tg.nty + h->batches - 1) / h->batches * tg.th);
            }
    }
    return (g.ny + h->batches - 1) / h->batches;
}

PlaneNeeds plane_needs(const nss_hip_bm3d* h, int i) {
    PlaneNeeds n;
    const nss_hip_bm3d_params& p = h->p;
    const bool rolling = p.radius > 0;
    for (int s = 0; s < h->nstages; ++s) {
        const Stage& st = h->stages[s];
        const StagePlane& sp = st.plane[i];
        bool any = false;
        for (int q = 0; q < h->nplanes; ++q) any |= st.plane[q].sigma != 0.f;
        // With chroma, plane 0 holds the shared groups whenever any plane is filtered.
        const bool matched = p.chroma ? (i == 0 && any) : sp.sigma != 0.f;
        if (matched) {
            const Grid g = make_grid(h->pw[i], h->ph[i], sp.block, sp.step);
            // a memory-limited plan holds the records of one batch of grid rows
            const int rows = h->batches > 0 ? batch_grid_rows(h, st, i, g) : g.ny;
            const size_t nref = static_cast<size_t>(g.nx) * rows;
            n.recs = std::max(n.recs, nref * sp.group);
            n.nref = std::max(n.nref, nref);
        }
        if (sp.sigma == 0.f || rolling) continue;
        for (int bi = 0; bi < st.nbuckets; ++bi)
            for (int k = 0; k < st.bucket[bi].nplanes; ++k) {
                if (st.bucket[bi].planes[k] != i) continue;
                if (st.bucket[bi].tiles) {
                    // a memory-limited plan: one batch of tile rows, and the
                    // frame sums its batches add into
                    const TileGeom& tg = st.bucket[bi].geom[k];
                    const Grid g = make_grid(h->pw[i], h->ph[i], sp.block, sp.step);
                    n.gw = std::max(n.gw, static_cast<size_t>(g.nx) * g.ny);
                    const int rows = h->batches > 0 ? (tg.nty + h->batches - 1) / h->batches : tg.nty;
                    n.part = std::max(n.part, static_cast<size_t>(tg.ntx) * rows * tg.pitch * tg.ah);
                    if (h->batches > 0) n.atomic = true;
                } else {
                    n.atomic = true;
                }
            }
    }
    n.basic = p.final_stage && !rolling;
    return n;
}

}  // namespace

namespace nsship {

// The global group cubes of one frame slot (every bucket's launch uses at
// most cube_slots cubes of the instance's largest size).
size_t cube_bytes(const nss_hip_bm3d* h) { return static_cast<size_t>(h->cube_slots) * h->cube_floats * sizeof(float); }

int frame_slot_count(const nss_hip_bm3d* h) { return h->p.radius > 0 ? 1 : std::max(1, h->p.num_streams); }

size_t frame_buffer_bytes(const nss_hip_bm3d* h) {
    size_t total = 0;
    for (int i = 0; i < h->nplanes; ++i) {
        const PlaneNeeds n = plane_needs(h, i);
        const size_t npx = static_cast<size_t>(h->pw[i]) * h->ph[i];
        const size_t part = h->part_elems[i] ? h->part_elems[i] + n.gw : 0;
        total += n.recs * sizeof(MatchRec) + n.nref * sizeof(int) + part * sizeof(unsigned long long) +
                 (n.atomic ? 16 * npx : 0) + (n.basic ? 4 * npx : 0);
    }
    return total + cube_bytes(h);
}

// Rolling state of plane i (rings and sums), from the plan.
size_t rolling_bytes(const nss_hip_bm3d* h, int i, size_t* cached, size_t* output) {
    const auto& R = h->roll;
    if (h->p.radius == 0) return 0;
    const size_t npx = static_cast<size_t>(h->pw[i]) * h->ph[i];
    size_t c = npx * 4 * static_cast<size_t>(R.src_n + R.ref_n + R.basic_n);
    const size_t o = npx * 4 * static_cast<size_t>(R.out_n) * R.out_rows;
    for (int s = 0; s < h->nstages; ++s) {
        if (h->stages[s].plane[i].sigma == 0.f) continue;
        // tile path: the centres' partials, and for rolling output the
        // frames' corner planes of group weights
        c += R.tiles[s] ? R.part_center[s][i] * (2 * static_cast<size_t>(R.r) + 1) * sizeof(float) +
                              (R.legacy ? 0 : npx * 8 * static_cast<size_t>(R.acc_n))
                        : npx * 16 * static_cast<size_t>(R.acc_n);
    }
    if (cached) *cached += c;
    if (output) *output += o;
    return c + o;
}

size_t instance_bytes(const nss_hip_bm3d* h) {
    size_t total = frame_buffer_bytes(h) * frame_slot_count(h);
    for (int i = 0; i < h->nplanes; ++i) total += rolling_bytes(h, i, nullptr, nullptr);
    return total;
}

void free_frame_buffers(FrameBuffers& fb) {
    for (int i = 0; i < 3; ++i) {
        for (void* p : {static_cast<void*>(fb.num[i]), static_cast<void*>(fb.den[i]), static_cast<void*>(fb.part[i]),
                        static_cast<void*>(fb.matches[i]), static_cast<void*>(fb.counts[i]),
                        static_cast<void*>(fb.basic[i])})
            if (p) dev_free(p);
    }
    if (fb.cube) dev_free(fb.cube);
    fb = FrameBuffers{};
}

hipError_t alloc_frame_buffers(const nss_hip_bm3d* h, FrameBuffers& fb) {
    fb = FrameBuffers{};
    hipError_t e = hipSuccess;
    for (int i = 0; i < h->nplanes && e == hipSuccess; ++i) {
        const PlaneNeeds n = plane_needs(h, i);
        const size_t npx = static_cast<size_t>(h->pw[i]) * h->ph[i];
        if (n.recs) {
            e = dev_alloc(&fb.matches[i], n.recs * sizeof(MatchRec));
            if (e == hipSuccess) e = dev_alloc(&fb.counts[i], n.nref * sizeof(int));
        }
        if (e == hipSuccess && n.atomic) {
            e = dev_alloc(&fb.num[i], npx * 8);
            if (e == hipSuccess) e = dev_alloc(&fb.den[i], npx * 8);
        }
        if (e == hipSuccess && h->part_elems[i])
            e = dev_alloc(&fb.part[i], (h->part_elems[i] + n.gw) * sizeof(unsigned long long));
        if (e == hipSuccess && n.basic) e = dev_alloc(&fb.basic[i], npx * 4);
    }
    if (e == hipSuccess && h->cube_floats) e = dev_alloc(&fb.cube, cube_bytes(h));
    if (e != hipSuccess) free_frame_buffers(fb);
    return e;
}

size_t carve_frame_buffers(const nss_hip_bm3d* h, char* base, FrameBuffers& fb) {
    fb = FrameBuffers{};
    size_t off = 0;
    auto take = [&](size_t bytes) -> char* {
        char* p = base ? base + off : nullptr;
        off += (bytes + 255) & ~static_cast<size_t>(255);
        return p;
    };
    for (int i = 0; i < h->nplanes; ++i) {
        const PlaneNeeds n = plane_needs(h, i);
        const size_t npx = static_cast<size_t>(h->pw[i]) * h->ph[i];
        if (n.recs) {
            fb.matches[i] = reinterpret_cast<MatchRec*>(take(n.recs * sizeof(MatchRec)));
            fb.counts[i] = reinterpret_cast<int*>(take(n.nref * sizeof(int)));
        }
        if (n.atomic) {
            fb.num[i] = reinterpret_cast<unsigned long long*>(take(npx * 8));
            fb.den[i] = reinterpret_cast<unsigned long long*>(take(npx * 8));
        }
        if (h->part_elems[i])
            fb.part[i] = reinterpret_cast<unsigned long long*>(take((h->part_elems[i] + n.gw) * sizeof(unsigned long long)));
        if (n.basic) fb.basic[i] = reinterpret_cast<float*>(take(npx * 4));
    }
    if (h->cube_floats) fb.cube = reinterpret_cast<float*>(take(cube_bytes(h)));
    return off;
}

bool fit_frame_buffers(nss_hip_bm3d* h, size_t bytes) {
    if (frame_buffer_bytes(h) <= bytes) return true;
    if (h->p.radius > 0) return false;
    // The batched plan (src/cuda/common/group_driver.cpp:481-520 batches the
    // work to fit the limit): the same kernels, over batches of tile rows (of
    // grid rows on the atomic path), every batch adding its exact sums into
    // frame-sized int64 sums, so the finish sees the sums of the whole-frame
    // plan.
    int max_rows = 1;
    for (int s = 0; s < h->nstages; ++s)
        for (int bi = 0; bi < h->stages[s].nbuckets; ++bi) {
            const Bucket& bk = h->stages[s].bucket[bi];
            for (int k = 0; k < bk.nplanes; ++k) {
                const int i = bk.planes[k];
                max_rows = std::max(max_rows, bk.tiles ? bk.geom[k].nty
                                                      : make_grid(h->pw[i], h->ph[i], bk.block,
                                                                  h->stages[s].plane[i].step).ny);
            }
        }
    // the global group cubes (if any) in halving counts once batching alone
    // does not fit: the slots only set how many groups run at once
    for (int slots = kCubeSlots; slots >= 1; slots /= 2) {
        h->cube_slots = slots;
        for (int nb = 1; nb <= max_rows; ++nb) {
            h->batches = nb;
            for (int i = 0; i < h->nplanes; ++i) h->part_elems[i] = plane_needs(h, i).part;
            if (frame_buffer_bytes(h) <= bytes) return true;
        }
        if (!h->cube_floats) break;
    }
    return false;
}

}  // namespace nsship

namespace {

// Event-traced launch on stream s (kind: 0 clear, 1 match, 2 filter, 3 finish).
template <typename F>
hipError_t traced(Tracer* tr, int kind, hipStream_t s, F&& launch) {
    if (!tr || !tr->on) return launch();
    std::lock_guard<std::mutex> lock(tr->mu);
    hipEvent_t a = tr->take(), z = tr->take();
    if (a) (void)hipEventRecord(a, s);
    const hipError_t err = launch();
    if (a && z) {
        (void)hipEventRecord(z, s);
        tr->spans[kind].emplace_back(a, z);
    }
    return err;
}

template <typename F>
hipError_t traced(nss_hip_bm3d* h, int kind, hipStream_t s, F&& launch) {
    return traced(&h->tracer, kind, s, static_cast<F&&>(launch));
}

// Planes of one frame for one stage: match on `match`, filter `src` (Wiener
// reference `wref` in a Wiener stage), finish into `dst`.
struct StageIO {
    const float* src[3]{};
    int src_stride[3]{};
    const float* match[3]{};
    int match_stride[3]{};
    const float* wref[3]{};
    int wref_stride[3]{};
    float* dst[3]{};
    int dst_stride[3]{};
};

// Stage k's planes of a frame (src/host/filter_bm3d.cpp:1086-1110): final=1
// runs basic into the instance's basic planes, then Wiener matched on and
// referenced to them; a ref clip alone gives one Wiener stage on the ref.
StageIO stage_io(const FrameRef& f, int k) {
    const nss_hip_bm3d* h = f.h;
    StageIO io;
    for (int i = 0; i < h->nplanes; ++i) {
        io.src[i] = f.src[i];
        io.src_stride[i] = f.src_stride[i];
        io.dst[i] = f.dst[i];
        io.dst_stride[i] = f.dst_stride[i];
        io.match[i] = f.src[i];
        io.match_stride[i] = f.src_stride[i];
        if (h->p.final_stage) {
            if (k == 0) {
                io.dst[i] = f.buf.basic[i];
                io.dst_stride[i] = h->pw[i];
            } else {
                io.match[i] = io.wref[i] = f.buf.basic[i];
                io.match_stride[i] = io.wref_stride[i] = h->pw[i];
            }
        } else if (h->stages[0].wiener) {
            io.match[i] = io.wref[i] = f.ref[i];
            io.match_stride[i] = io.wref_stride[i] = f.ref_stride[i];
        }
    }
    return io;
}

// Jobs of one bucket shape gathered over the frames of a batch, launched as
// one clear / match / filter / finish sequence.
struct LaunchGroup {
    int block = 0, match_kind = 0, fast_group = 0, form = 0, cube_floats = 0;
    bool tiles = false, wiener = false;
    size_t match_lds = 0, filter_lds = 0;
    float* cube = nullptr;  // global group cubes (the first frame's), with cube_floats
    int cube_slots = 0;
    ZeroTable zt{};
    MatchTable mt{};
    FilterTable ft{};
    FinishTable fn{};
    int zwg = 0, mwg = 0, fwg = 0, nwg = 0;

    bool fits(const Bucket& bk, bool w) const {
        return block == bk.block && match_kind == bk.match_kind && fast_group == bk.fast_group && form == bk.form &&
               tiles == bk.tiles && wiener == w && cube_floats == bk.cube_floats;
    }
    bool empty() const { return ft.n == 0 && mt.n == 0; }
};

int flush_group(LaunchGroup& g, Tracer* tr, hipStream_t s) {
    if (g.empty()) return NSS_HIP_OK;
    const int b = g.block;
    hipError_t e = hipSuccess;
    if (!g.tiles && g.zt.n) {
        e = traced(tr, 0, s, [&] { return launch_zero(g.zt, g.zwg, s); });
        if (e != hipSuccess) return hip_fail(e, "clear launch");
    }
    if (g.mt.n) {
        e = traced(tr, 1, s, [&] { return launch_match(b, g.mt, g.mwg, g.match_lds, g.match_kind, s); });
        if (e != hipSuccess) return hip_fail(e, "match launch");
    }
    if (g.ft.n) {
        g.ft.cube = g.cube_floats ? g.cube : nullptr;
        g.ft.cube_floats = g.cube_floats;
        g.ft.cube_slots = g.cube_slots;
        e = traced(tr, 2, s, [&] {
            return launch_filter(b, g.ft, g.fwg, g.filter_lds, g.wiener, g.tiles, g.fast_group, g.form, s);
        });
        if (e != hipSuccess) return hip_fail(e, "filter launch");
        e = traced(tr, 3, s, [&] { return launch_finish(b, g.fn, g.nwg, g.tiles, s); });
        if (e != hipSuccess) return hip_fail(e, "finish launch");
    }
    const LaunchGroup fresh{};
    g.zt = fresh.zt;
    g.mt = fresh.mt;
    g.ft = fresh.ft;
    g.fn = fresh.fn;
    g.zwg = g.mwg = g.fwg = g.nwg = 0;
    g.match_lds = g.filter_lds = 0;
    g.cube = nullptr;
    return NSS_HIP_OK;
}

// Appends one frame's jobs of bucket bk (stage st, sigma per plane) to g.
void add_bucket_jobs(LaunchGroup& g, const FrameRef& f, const Stage& st, const Bucket& bk, const StageIO& io,
                     const float* sigma) {
    nss_hip_bm3d* h = f.h;
    const int b = bk.block;
    g.match_lds = std::max(g.match_lds, bk.match_lds);
    g.filter_lds = std::max(g.filter_lds, bk.filter_lds);
    if (bk.cube_floats && !g.cube) {
        g.cube = f.buf.cube;
        g.cube_slots = h->cube_slots;
    }
    // CBM3D (chroma=1): the groups are matched once, on plane 0 of the match
    // planes, and every filtered plane uses them (src/host/filter_bm3d.cpp:451-486).
    int mplanes[3], nmp = 0;
    if (h->p.chroma)
        mplanes[nmp++] = 0;
    else
        for (int k = 0; k < bk.nplanes; ++k) mplanes[nmp++] = bk.planes[k];
    for (int k = 0; k < nmp; ++k) {
        const int i = mplanes[k];
        const StagePlane& sp = st.plane[i];
        const Grid gr = make_grid(h->pw[i], h->ph[i], b, sp.step);
        MatchJob& m = g.mt.job[g.mt.n++];
        m.plane = io.match[i];
        m.width = h->pw[i];
        m.height = h->ph[i];
        m.stride = io.match_stride[i];
        m.grid = gr;
        m.range = sp.range;
        m.group = sp.group;
        m.t0 = 0;
        m.wg_begin = g.mwg;
        m.out = f.buf.matches[i];
        m.counts = f.buf.counts[i];
        g.mwg += match_workgroups(bk.match_kind, gr);
    }
    for (int k = 0; k < bk.nplanes; ++k) {
        const int i = bk.planes[k];
        const int mi = h->p.chroma ? 0 : i;
        const StagePlane& sp = st.plane[i];
        const int w = h->pw[i], hh = h->ph[i];
        const Grid gr = make_grid(w, hh, b, sp.step);
        const int nref = gr.nx * gr.ny;
        if (!bk.tiles) {
            const long long npx = static_cast<long long>(w) * hh;
            for (unsigned long long* buf : {f.buf.num[i], f.buf.den[i]}) {
                ZeroJob& z = g.zt.job[g.zt.n++];
                z.p = buf;
                z.count = npx;
                z.wg_begin = g.zwg;
                g.zwg += static_cast<int>((npx + 1023) / 1024);
            }
        }
        FilterJob& fj = g.ft.job[g.ft.n++];
        fj.src = io.src[i];
        fj.wref = st.wiener ? io.wref[i] : nullptr;
        fj.width = w;
        fj.height = hh;
        fj.stride = io.src_stride[i];
        fj.wref_stride = st.wiener ? io.wref_stride[i] : 0;
        fj.grid = gr;
        fj.range = sp.range;
        fj.nref = nref;
        fj.group = sp.group;
        fj.sigma = sigma[i];
        fj.wg_begin = g.fwg;
        fj.matches = f.buf.matches[mi];
        fj.counts = f.buf.counts[mi];
        fj.num = f.buf.num[i];
        fj.den = f.buf.den[i];
        fj.tiles = bk.geom[k];
        fj.part_num = f.buf.part[i];
        // the group weights after the numerator partials
        fj.gweight = f.buf.part[i] ? f.buf.part[i] + h->part_elems[i] : nullptr;
        g.fwg += bk.tiles ? bk.geom[k].ntx * bk.geom[k].nty : nref;

        FinishJob& n = g.fn.job[g.fn.n++];
        n.num = f.buf.num[i];
        n.den = f.buf.den[i];
        n.exact_num = reinterpret_cast<const long long*>(fj.part_num);
        n.matches = fj.matches;
        n.counts = fj.counts;
        n.gweight = fj.gweight;
        n.grid = gr;
        n.group = sp.group;
        n.tiles = bk.geom[k];
        n.block = b;
        n.step = sp.step;
        n.range = sp.range;
        n.src = io.src[i];
        n.dst = io.dst[i];
        n.width = w;
        n.height = hh;
        n.sstride = io.src_stride[i];
        n.dstride = io.dst_stride[i];
        // finish blocks: 64 x 16 pixels on the tile path, 16 x 16 otherwise
        n.tiles_x = bk.tiles ? (w + 63) / 64 : (w + 15) / 16;
        n.wg_begin = g.nwg;
        g.nwg += n.tiles_x * ((hh + 15) / 16);
    }
}

}  // namespace

namespace nsship {

// A tile bucket of a memory-limited instance: per batch of tile rows, the
// grid rows of those tiles are matched (records of the batch only), the
// tiles filtered (the kernels of the whole-frame plan, partials of the batch
// only), and the batch's partials added into the frame sums; then the output
// from the sums. The sums equal the whole-frame finish's, so do the bits.
int run_limited_tiles(const FrameRef& f, const Stage& st, const Bucket& bk, const StageIO& io, const float* sigma,
                      Tracer* tr, hipStream_t s) {
    nss_hip_bm3d* h = f.h;
    const int nb = h->batches, b = bk.block;
    // the tile rows of batch `batch` of bucket plane q: [tr0, tr1)
    auto rows_of = [&](int q, int batch, int* tr0, int* tr1) {
        const TileGeom& tg = bk.geom[q];
        const int t = (tg.nty + nb - 1) / nb;
        *tr0 = std::min(tg.nty, batch * t);
        *tr1 = std::min(tg.nty, *tr0 + t);
    };
    for (int batch = 0; batch < nb; ++batch) {
        MatchTable mt{};
        int mwg = 0;
        for (int q = 0; q < bk.nplanes; ++q) {
            const int i = bk.planes[q];
            // with chroma the groups are matched once, on plane 0 (the same
            // grid and tiles as every plane of a 4:4:4 clip)
            if (h->p.chroma && q > 0) break;
            const int mi = h->p.chroma ? 0 : i;
            const StagePlane& sp = st.plane[mi];
            int tr0, tr1;
            rows_of(q, batch, &tr0, &tr1);
            if (tr0 >= tr1) continue;
            Grid g = make_grid(h->pw[mi], h->ph[mi], b, sp.step);
            g.row0 = tr0 * bk.geom[q].th;
            g.ny = std::min(g.ny, tr1 * bk.geom[q].th) - g.row0;
            MatchJob& m = mt.job[mt.n++];
            m.plane = io.match[mi];
            m.width = h->pw[mi];
            m.height = h->ph[mi];
            m.stride = io.match_stride[mi];
            m.grid = g;
            m.range = sp.range;
            m.group = sp.group;
            m.t0 = 0;
            m.wg_begin = mwg;
            m.out = f.buf.matches[mi];
            m.counts = f.buf.counts[mi];
            mwg += match_workgroups(bk.match_kind, g);
        }
        if (!mt.n) continue;
        hipError_t e = traced(tr, 1, s, [&] { return launch_match(b, mt, mwg, bk.match_lds, bk.match_kind, s); });
        if (e != hipSuccess) return hip_fail(e, "match launch");
        FilterTable ft{};
        FinishTable fn{};
        int fwg = 0, nwg = 0;
        for (int q = 0; q < bk.nplanes; ++q) {
            const int i = bk.planes[q];
            const int mi = h->p.chroma ? 0 : i;
            const StagePlane& sp = st.plane[i];
            const TileGeom& tg = bk.geom[q];
            int tr0, tr1;
            rows_of(q, batch, &tr0, &tr1);
            if (tr0 >= tr1) continue;
            Grid g = make_grid(h->pw[i], h->ph[i], b, sp.step);
            g.row0 = tr0 * tg.th;
            FilterJob& fj = ft.job[ft.n++];
            fj.src = io.src[i];
            fj.wref = st.wiener ? io.wref[i] : nullptr;
            fj.width = h->pw[i];
            fj.height = h->ph[i];
            fj.stride = io.src_stride[i];
            fj.wref_stride = st.wiener ? io.wref_stride[i] : 0;
            fj.grid = g;
            fj.range = sp.range;
            fj.nref = g.nx * g.ny;
            fj.group = sp.group;
            fj.sigma = sigma[i];
            fj.wg_begin = fwg;
            fj.matches = f.buf.matches[mi];
            fj.counts = f.buf.counts[mi];
            fj.tiles = tg;
            fj.part_num = f.buf.part[i];
            fj.gweight = f.buf.part[i] + h->part_elems[i];
            fj.tile0 = tr0 * tg.ntx;
            fwg += (tr1 - tr0) * tg.ntx;
            // the pixel rows the batch's tiles reach
            const int py0 = std::max(0, tg.ay0(tr0, h->ph[i], b, sp.step, sp.range));
            const int py1 = std::min(h->ph[i], tg.ay0(tr1 - 1, h->ph[i], b, sp.step, sp.range) + tg.ah);
            FinishJob& n = fn.job[fn.n++];
            n.exact_num = reinterpret_cast<const long long*>(fj.part_num);
            // the weights of the batch's references (grid rows of its tiles)
            n.matches = fj.matches;
            n.counts = fj.counts;
            n.gweight = fj.gweight;
            n.grid = g;
            n.grid.ny = std::min(g.ny, tr1 * tg.th) - g.row0;
            n.group = sp.group;
            n.tiles = tg;
            n.block = b;
            n.step = sp.step;
            n.range = sp.range;
            n.src = io.src[i];
            n.dst = io.dst[i];
            n.width = h->pw[i];
            n.height = h->ph[i];
            n.sstride = io.src_stride[i];
            n.dstride = io.dst_stride[i];
            n.acc_num = reinterpret_cast<long long*>(f.buf.num[i]);
            n.acc_den = reinterpret_cast<long long*>(f.buf.den[i]);
            n.ty0 = tr0;
            n.ty1 = tr1;
            n.py0 = py0;
            n.tiles_x = (h->pw[i] + 63) / 64;
            n.wg_begin = nwg;
            nwg += n.tiles_x * ((py1 - py0 + 15) / 16);
        }
        if (!ft.n) continue;
        e = traced(tr, 2, s, [&] {
            return launch_filter(b, ft, fwg, bk.filter_lds, st.wiener, true, bk.fast_group, bk.form, s);
        });
        if (e != hipSuccess) return hip_fail(e, "filter launch");
        e = traced(tr, 3, s, [&] { return launch_finish(b, fn, nwg, true, s); });
        if (e != hipSuccess) return hip_fail(e, "finish launch");
    }
    FinishTable fn{};
    int nwg = 0;
    for (int q = 0; q < bk.nplanes; ++q) {
        const int i = bk.planes[q];
        FinishJob& n = fn.job[fn.n++];
        n.src = io.src[i];
        n.dst = io.dst[i];
        n.width = h->pw[i];
        n.height = h->ph[i];
        n.sstride = io.src_stride[i];
        n.dstride = io.dst_stride[i];
        n.acc_num = reinterpret_cast<long long*>(f.buf.num[i]);
        n.acc_den = reinterpret_cast<long long*>(f.buf.den[i]);
        n.tiles_x = (h->pw[i] + 63) / 64;
        n.wg_begin = nwg;
        nwg += n.tiles_x * ((h->ph[i] + 3) / 4);
    }
    const hipError_t e = traced(tr, 3, s, [&] { return launch_finish_sums(fn, nwg, s); });
    return e == hipSuccess ? NSS_HIP_OK : hip_fail(e, "finish launch");
}

// One frame of a memory-limited instance (h->batches > 0): per stage, the
// filtered planes' frame sums are cleared, then every bucket runs in
// batches (tile rows on the tile path, grid rows on the atomic path), each
// batch's references matched with the batch's records only and their groups
// filtered into the sums; the finish runs once per bucket.
int run_limited(const FrameRef& f, Tracer* tr, hipStream_t s) {
    nss_hip_bm3d* h = f.h;
    const int nb = h->batches;
    for (int k = 0; k < h->nstages; ++k) {
        const Stage& st = h->stages[k];
        const StageIO io = stage_io(f, k);
        ZeroTable zt{};
        int zwg = 0;
        for (int i = 0; i < h->nplanes; ++i) {
            if (st.plane[i].sigma == 0.f) {
                const hipError_t e =
                    hipMemcpy2DAsync(io.dst[i], static_cast<size_t>(io.dst_stride[i]) * 4, io.src[i],
                                     static_cast<size_t>(io.src_stride[i]) * 4, static_cast<size_t>(h->pw[i]) * 4,
                                     h->ph[i], hipMemcpyDeviceToDevice, s);
                if (e != hipSuccess) return hip_fail(e, "plane copy");
                continue;
            }
            const long long npx = static_cast<long long>(h->pw[i]) * h->ph[i];
            for (unsigned long long* buf : {f.buf.num[i], f.buf.den[i]}) {
                ZeroJob& z = zt.job[zt.n++];
                z.p = buf;
                z.count = npx;
                z.wg_begin = zwg;
                zwg += static_cast<int>((npx + 1023) / 1024);
            }
        }
        if (zt.n) {
            const hipError_t e = traced(tr, 0, s, [&] { return launch_zero(zt, zwg, s); });
            if (e != hipSuccess) return hip_fail(e, "clear launch");
        }
        for (int bi = 0; bi < st.nbuckets; ++bi) {
            const Bucket& bk = st.bucket[bi];
            const int b = bk.block;
            if (bk.tiles) {
                const int rc = run_limited_tiles(f, st, bk, io, f.sigma[k], tr, s);
                if (rc != NSS_HIP_OK) return rc;
                continue;
            }
            int mplanes[3], nmp = 0;
            if (h->p.chroma)
                mplanes[nmp++] = 0;
            else
                for (int q = 0; q < bk.nplanes; ++q) mplanes[nmp++] = bk.planes[q];
            for (int batch = 0; batch < nb; ++batch) {
                // The batch's grid rows of each plane: [batch R, batch R + R)
                // clipped, R = ceil(ny / nb); records indexed within the batch.
                auto batch_grid = [&](int i) {
                    Grid g = make_grid(h->pw[i], h->ph[i], b, st.plane[i].step);
                    const int rows = (g.ny + nb - 1) / nb;
                    g.row0 = std::min(g.ny, batch * rows);
                    g.ny = std::min(g.ny, g.row0 + rows) - g.row0;
                    return g;
                };
                MatchTable mt{};
                int mwg = 0;
                for (int q = 0; q < nmp; ++q) {
                    const int i = mplanes[q];
                    const StagePlane& sp = st.plane[i];
                    const Grid g = batch_grid(i);
                    if (g.ny <= 0) continue;
                    MatchJob& m = mt.job[mt.n++];
                    m.plane = io.match[i];
                    m.width = h->pw[i];
                    m.height = h->ph[i];
                    m.stride = io.match_stride[i];
                    m.grid = g;
                    m.range = sp.range;
                    m.group = sp.group;
                    m.t0 = 0;
                    m.wg_begin = mwg;
                    m.out = f.buf.matches[i];
                    m.counts = f.buf.counts[i];
                    mwg += match_workgroups(bk.match_kind, g);
                }
                if (!mt.n) continue;
                hipError_t e = traced(tr, 1, s, [&] { return launch_match(b, mt, mwg, bk.match_lds, bk.match_kind, s); });
                if (e != hipSuccess) return hip_fail(e, "match launch");
                int fwg = 0;
                {
                    FilterTable ft{};
                    for (int q = 0; q < bk.nplanes; ++q) {
                        const int i = bk.planes[q];
                        const int mi = h->p.chroma ? 0 : i;
                        const StagePlane& sp = st.plane[i];
                        const Grid g = batch_grid(i);
                        if (g.ny <= 0) continue;
                        FilterJob& fj = ft.job[ft.n++];
                        fj.src = io.src[i];
                        fj.wref = st.wiener ? io.wref[i] : nullptr;
                        fj.width = h->pw[i];
                        fj.height = h->ph[i];
                        fj.stride = io.src_stride[i];
                        fj.wref_stride = st.wiener ? io.wref_stride[i] : 0;
                        fj.grid = g;
                        fj.range = sp.range;
                        fj.nref = g.nx * g.ny;
                        fj.group = sp.group;
                        fj.sigma = f.sigma[k][i];
                        fj.wg_begin = fwg;
                        fj.matches = f.buf.matches[mi];
                        fj.counts = f.buf.counts[mi];
                        fj.num = f.buf.num[i];
                        fj.den = f.buf.den[i];
                        fwg += fj.nref;
                    }
                    if (!ft.n) continue;
                    if (bk.cube_floats) {
                        ft.cube = f.buf.cube;
                        ft.cube_floats = bk.cube_floats;
                        ft.cube_slots = h->cube_slots;
                    }
                    e = traced(tr, 2, s, [&] {
                        return launch_filter(b, ft, fwg, bk.filter_lds, st.wiener, false, 0, 0, s);
                    });
                }
                if (e != hipSuccess) return hip_fail(e, "filter launch");
            }
            FinishTable fn{};
            int nwg = 0;
            for (int q = 0; q < bk.nplanes; ++q) {
                const int i = bk.planes[q];
                const StagePlane& sp = st.plane[i];
                FinishJob& n = fn.job[fn.n++];
                n.num = f.buf.num[i];
                n.den = f.buf.den[i];
                n.block = b;
                n.step = sp.step;
                n.range = sp.range;
                n.src = io.src[i];
                n.dst = io.dst[i];
                n.width = h->pw[i];
                n.height = h->ph[i];
                n.sstride = io.src_stride[i];
                n.dstride = io.dst_stride[i];
                n.tiles_x = (h->pw[i] + 15) / 16;
                n.wg_begin = nwg;
                nwg += n.tiles_x * ((h->ph[i] + 15) / 16);
            }
            const hipError_t e = traced(tr, 3, s, [&] { return launch_finish(b, fn, nwg, false, s); });
            if (e != hipSuccess) return hip_fail(e, "finish launch");
        }
    }
    return NSS_HIP_OK;
}

// Both stages of a batch of radius-0 frames: per stage round, every frame's
// planes with sigma 0 are copied and the jobs of equal bucket shape share one
// launch sequence (at most kMaxJobs planes per table, else several). Frames
// of memory-limited instances run on their own, batch by batch.
int run_batch(const FrameRef* frames, int nframes, Tracer* tr, hipStream_t s) {
    std::vector<FrameRef> plain;
    for (int q = 0; q < nframes; ++q) {
        if (frames[q].h->batches > 0) {
            const int rc = run_limited(frames[q], tr, s);
            if (rc != NSS_HIP_OK) return rc;
        } else {
            plain.push_back(frames[q]);
        }
    }
    if (plain.size() != static_cast<size_t>(nframes)) {
        frames = plain.data();
        nframes = static_cast<int>(plain.size());
    }
    int rounds = 0;
    for (int q = 0; q < nframes; ++q) rounds = std::max(rounds, frames[q].h->nstages);
    std::vector<LaunchGroup> groups;
    for (int k = 0; k < rounds; ++k) {
        groups.clear();
        for (int q = 0; q < nframes; ++q) {
            const FrameRef& f = frames[q];
            nss_hip_bm3d* h = f.h;
            if (k >= h->nstages) continue;
            const Stage& st = h->stages[k];
            const StageIO io = stage_io(f, k);
            for (int i = 0; i < h->nplanes; ++i) {
                if (st.plane[i].sigma != 0.f) continue;
                const hipError_t e =
                    hipMemcpy2DAsync(io.dst[i], static_cast<size_t>(io.dst_stride[i]) * 4, io.src[i],
                                     static_cast<size_t>(io.src_stride[i]) * 4, static_cast<size_t>(h->pw[i]) * 4,
                                     h->ph[i], hipMemcpyDeviceToDevice, s);
                if (e != hipSuccess) return hip_fail(e, "plane copy");
            }
            for (int bi = 0; bi < st.nbuckets; ++bi) {
                const Bucket& bk = st.bucket[bi];
                LaunchGroup* g = nullptr;
                for (auto& cand : groups)
                    if (cand.fits(bk, st.wiener)) g = &cand;
                if (!g) {
                    groups.emplace_back();
                    g = &groups.back();
                    g->block = bk.block;
                    g->match_kind = bk.match_kind;
                    g->fast_group = bk.fast_group;
                    g->form = bk.form;
                    g->tiles = bk.tiles;
                    g->wiener = st.wiener;
                    g->cube_floats = bk.cube_floats;
                }
                const int nmp = h->p.chroma ? 1 : bk.nplanes;
                if (g->mt.n + nmp > kMaxJobs || g->ft.n + bk.nplanes > kMaxJobs) {
                    const int rc = flush_group(*g, tr, s);
                    if (rc != NSS_HIP_OK) return rc;
                }
                add_bucket_jobs(*g, f, st, bk, io, f.sigma[k]);
            }
        }
        for (auto& g : groups) {
            const int rc = flush_group(g, tr, s);
            if (rc != NSS_HIP_OK) return rc;
        }
    }
    return NSS_HIP_OK;
}

// The instance's sigmas for frame planes src / ref / dst (the caller sets
// the frame's buffers).
FrameRef frame_ref(nss_hip_bm3d* h, const float* const* src, const int* src_stride, const float* const* ref,
                   const int* ref_stride, float* const* dst, const int* dst_stride) {
    FrameRef f{};
    f.h = h;
    for (int i = 0; i < h->nplanes; ++i) {
        f.src[i] = src[i];
        f.src_stride[i] = src_stride[i];
        f.ref[i] = ref ? ref[i] : nullptr;
        f.ref_stride[i] = ref_stride ? ref_stride[i] : 0;
        f.dst[i] = dst[i];
        f.dst_stride[i] = dst_stride[i];
    }
    for (int k = 0; k < h->nstages; ++k)
        for (int i = 0; i < 3; ++i) f.sigma[k][i] = h->stages[k].plane[i].sigma;
    return f;
}

}  // namespace nsship

namespace {

// ---------------------------------------------------------------------------
// Rolling temporal processing (src/host/filter_bm3d.cpp:730-851): the centre
// frames of a stage are processed in ascending order; centre c adds into the
// sums of frames c-r .. c+r (clamped to the clip), and frame n is finished
// once every centre that reaches it (up to min(n + r, N - 1)) is done.
// ---------------------------------------------------------------------------

FrameRing ring_of(float* base, int ring, int w, int hgt) {
    return FrameRing{base, static_cast<long long>(w) * hgt, ring};
}

// Match frames of stage s: the user's ref clip when given, the basic
// estimate in the Wiener stage of final=1, else the source.
FrameRing match_ring(nss_hip_bm3d* h, int s, int i) {
    auto& R = h->roll;
    if (h->p.final_stage && s == 1) return ring_of(R.basic[i], R.basic_n, h->pw[i], h->ph[i]);
    if (h->p.has_ref) return ring_of(R.ref[i], R.ref_n, h->pw[i], h->ph[i]);
    return ring_of(R.src[i], R.src_n, h->pw[i], h->ph[i]);
}

FrameRing wref_ring(nss_hip_bm3d* h, int s, int i) {
    auto& R = h->roll;
    if (h->p.final_stage && s == 1) return ring_of(R.basic[i], R.basic_n, h->pw[i], h->ph[i]);
    if (h->p.has_ref) return ring_of(R.ref[i], R.ref_n, h->pw[i], h->ph[i]);
    return FrameRing{nullptr, 0, 1};
}

// Slot geometries of the temporal tile path for plane i of stage s: slot t
// reaches range + |t - r| * ps_range around its references (the predictive
// search's seeds are spatial matches, each layer widens by ps_range).
void slot_geoms(const nss_hip_bm3d* h, int s, int i, int tw, int th, TileGeom* g, int* range) {
    const StagePlane& sp = h->stages[s].plane[i];
    const int r = h->p.radius;
    const Grid gr = make_grid(h->pw[i], h->ph[i], 8, sp.step);
    for (int t = 0; t < 2 * r + 1; ++t) {
        range[t] = sp.range + (t > r ? t - r : r - t) * sp.ps_range;
        g[t] = make_tiles(gr, tw, th, 8, range[t]);
    }
}

// Floats of one centre's partials of plane i (every slot's num, and its den
// after it for legacy output; rolling output keeps the weights in corner
// planes), and slot offsets.
size_t centre_partials(const TileGeom* g, int ns, long long* slot_off, bool den) {
    size_t total = 0;
    for (int t = 0; t < ns; ++t) {
        if (slot_off) slot_off[t] = static_cast<long long>(total);
        total += (den ? 2 : 1) * static_cast<size_t>(g[t].ntx) * g[t].nty * g[t].pitch * g[t].ah;
    }
    return total;
}

// The temporal tile path for stage s when it has one bucket of block 8 with
// one group size in {8, 16, 32}, radius <= 2 and a tile shape whose slot
// sums fit the LDS (one workgroup per CU); else the atomic kernels.
void plan_temporal_tiles(nss_hip_bm3d* h, int s) {
    auto& R = h->roll;
    const Stage& st = h->stages[s];
    R.tiles[s] = false;
    if (R.r < 1 || 2 * R.r + 1 > kMaxTileSlots || st.nbuckets != 1 || st.bucket[0].block != 8) return;
    const Bucket& bk = st.bucket[0];
    const int G = st.plane[bk.planes[0]].group;
    for (int k = 1; k < bk.nplanes; ++k)
        if (st.plane[bk.planes[k]].group != G) return;
    if (G != 8 && G != 16 && G != 32) return;
    static const int shapes[][2] = {{4, 4}, {4, 2}, {2, 2}, {2, 1}, {1, 1}};
    for (const auto& sh : shapes) {
        if (sh[0] * sh[1] * G > 512) continue;  // one match record per thread of the tile kernel
        size_t lds = 0;
        for (int k = 0; k < bk.nplanes; ++k) {
            TileGeom g[kMaxTileSlots];
            int range[kMaxTileSlots];
            slot_geoms(h, s, bk.planes[k], sh[0], sh[1], g, range);
            lds = std::max(lds, filter_temporal_tile_lds_bytes(G, g, 2 * R.r + 1, !R.legacy));
        }
        if (lds + 1024 <= kLdsLimit) {
            R.tiles[s] = true;
            R.tw[s] = sh[0];
            R.th[s] = sh[1];
            R.tile_lds[s] = lds;
            return;
        }
    }
}

int temporal_center(nss_hip_bm3d* h, int s, int c, hipStream_t st) {
    auto& R = h->roll;
    const Stage& stg = h->stages[s];
    const Slots slots{c, R.N, R.r};
    for (int bi = 0; bi < stg.nbuckets; ++bi) {
        const Bucket& bk = stg.bucket[bi];
        const int b = bk.block;
        MatchTable mt{};
        PredTable pt{};
        TFilterTable ft{};
        int mwg = 0, pwg = 0, fwg = 0;
        size_t plds = 0, flds = 0;
        int mplanes[3], nmp = 0;
        if (h->p.chroma)
            mplanes[nmp++] = 0;
        else
            for (int k = 0; k < bk.nplanes; ++k) mplanes[nmp++] = bk.planes[k];
        bool predictive = false;
        for (int k = 0; k < nmp; ++k) {
            const int i = mplanes[k];
            const StagePlane& sp = stg.plane[i];
            const Grid g = make_grid(h->pw[i], h->ph[i], b, sp.step);
            const FrameRing mr = match_ring(h, s, i);
            MatchJob& m = mt.job[mt.n++];
            m.plane = mr.frame(c);
            m.width = h->pw[i];
            m.height = h->ph[i];
            m.stride = h->pw[i];
            m.grid = g;
            m.range = sp.range;
            m.group = sp.group;
            m.t0 = R.r;
            m.wg_begin = mwg;
            m.out = h->slots[0].buf.matches[i];
            m.counts = h->slots[0].buf.counts[i];
            mwg += match_workgroups(bk.match_kind, g);
            // group 1 keeps the reference alone (src/cpu/bm/matcher.hpp:323).
            if (sp.group > 1) {
                PredJob& p = pt.job[pt.n++];
                p.match = mr;
                p.width = h->pw[i];
                p.height = h->ph[i];
                p.stride = h->pw[i];
                p.nref = g.nx * g.ny;
                p.group = sp.group;
                p.ps_num = sp.ps_num;
                p.ps_range = sp.ps_range;
                p.slots = slots;
                p.wg_begin = pwg;
                p.recs = h->slots[0].buf.matches[i];
                p.counts = h->slots[0].buf.counts[i];
                pwg += p.nref;
                predictive = true;
            }
        }
        // Block 8 with one group size in {8, 16, 32}: the register kernel,
        // four groups per workgroup.
        int fast = 0;
        if (b == 8) {
            fast = stg.plane[bk.planes[0]].group;
            for (int k = 1; k < bk.nplanes; ++k)
                if (stg.plane[bk.planes[k]].group != fast) fast = 0;
            if (fast != 8 && fast != 16 && fast != 32) fast = 0;
        }
        for (int k = 0; k < bk.nplanes; ++k) {
            const int i = bk.planes[k];
            const int mi = h->p.chroma ? 0 : i;
            const StagePlane& sp = stg.plane[i];
            const Grid g = make_grid(h->pw[i], h->ph[i], b, sp.step);
            TFilterJob& f = ft.job[ft.n++];
            f.src = ring_of(R.src[i], R.src_n, h->pw[i], h->ph[i]);
            f.wref = wref_ring(h, s, i);
            f.width = h->pw[i];
            f.height = h->ph[i];
            f.stride = h->pw[i];
            f.wref_stride = h->pw[i];
            f.nref = g.nx * g.ny;
            f.group = sp.group;
            f.sigma = sp.sigma;
            f.slots = slots;
            f.wg_begin = fwg;
            f.matches = h->slots[0].buf.matches[mi];
            f.counts = h->slots[0].buf.counts[mi];
            f.num = R.num[s][i];
            f.den = R.den[s][i];
            f.acc_ring = R.acc_n;
            fwg += fast ? (f.nref + 3) / 4 : f.nref;
            flds = std::max(flds, fast ? filter_temporal_b8_lds_bytes(sp.group)
                                       : filter_atomic_lds_bytes(b, sp.group, stg.wiener));
        }
        if (bk.cube_floats && !fast) {
            ft.cube = h->slots[0].buf.cube;
            ft.cube_floats = bk.cube_floats;
            ft.cube_slots = h->cube_slots;
        }
        hipError_t e = traced(h, 1, st, [&] { return launch_match(b, mt, mwg, bk.match_lds, bk.match_kind, st); });
        if (e != hipSuccess) return hip_fail(e, "match launch");
        if (predictive) {
            plds = predictive_table_lds(b, pt, R.r);
            e = traced(h, 1, st, [&] { return launch_predictive(b, pt, pwg, plds, st); });
            if (e != hipSuccess) return hip_fail(e, "predictive match launch");
        }
        if (R.tiles[s]) {
            TTileTable tt{};
            int twg = 0;
            const int ns = 2 * R.r + 1;
            for (int k = 0; k < bk.nplanes; ++k) {
                const int i = bk.planes[k];
                const TFilterJob& f = ft.job[k];
                TTileJob& j = tt.job[tt.n++];
                j.src = f.src;
                j.wref = f.wref;
                j.width = f.width;
                j.height = f.height;
                j.stride = f.stride;
                j.wref_stride = f.wref_stride;
                j.grid = make_grid(h->pw[i], h->ph[i], 8, stg.plane[i].step);
                j.nslots = ns;
                slot_geoms(h, s, i, R.tw[s], R.th[s], j.tiles, j.range);
                centre_partials(j.tiles, ns, j.slot_off, R.legacy);
                j.part = R.part[s][i] + static_cast<size_t>(c % ns) * R.part_center[s][i];
                j.corners = R.legacy ? nullptr : R.den[s][i];
                j.acc_ring = R.acc_n;
                j.nref = f.nref;
                j.group = f.group;
                j.sigma = f.sigma;
                j.slots = f.slots;
                j.wg_begin = twg;
                j.matches = f.matches;
                j.counts = f.counts;
                twg += j.tiles[0].ntx * j.tiles[0].nty;
            }
            e = traced(h, 2, st, [&] {
                return launch_filter_temporal_tiles(stg.plane[bk.planes[0]].group, tt, twg, R.tile_lds[s], stg.wiener,
                                                    st);
            });
        } else {
            e = traced(h, 2, st, [&] { return launch_filter_temporal(b, ft, fwg, flds, stg.wiener, fast, st); });
        }
        if (e != hipSuccess) return hip_fail(e, "filter launch");
    }
    return NSS_HIP_OK;
}

// Frame n of stage s: num/den (or the source where nothing landed) into the
// basic ring (stage 0 of final=1) or the output ring; its sums are cleared
// for the frame that reuses the accumulator slot.
int temporal_finish(nss_hip_bm3d* h, int s, int n, hipStream_t st) {
    auto& R = h->roll;
    const Stage& stg = h->stages[s];
    const bool to_out = s == h->nstages - 1;
    for (int i = 0; i < h->nplanes; ++i) {
        if (stg.plane[i].sigma != 0.f) continue;
        const size_t npx = static_cast<size_t>(h->pw[i]) * h->ph[i];
        float* dst = to_out ? R.out[i] + static_cast<size_t>(n % R.out_n) * npx
                            : R.basic[i] + static_cast<size_t>(n % R.basic_n) * npx;
        const hipError_t e = hipMemcpyAsync(dst, R.src[i] + static_cast<size_t>(n % R.src_n) * npx, npx * 4,
                                            hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return hip_fail(e, "plane copy");
    }
    if (R.tiles[s]) {
        const Bucket& bk = stg.bucket[0];
        const int ns = 2 * R.r + 1;
        TFinishTable fn{};
        int nwg = 0;
        for (int k = 0; k < bk.nplanes; ++k) {
            const int i = bk.planes[k];
            const size_t npx = static_cast<size_t>(h->pw[i]) * h->ph[i];
            TFinishJob& f = fn.job[fn.n++];
            TileGeom g[kMaxTileSlots];
            int range[kMaxTileSlots];
            long long off[kMaxTileSlots];
            slot_geoms(h, s, i, R.tw[s], R.th[s], g, range);
            centre_partials(g, ns, off, R.legacy);
            f.nsrc = 0;
            for (int c = std::max(0, n - R.r); c <= std::min(R.N - 1, n + R.r); ++c) {
                const int t = n - c + R.r;
                const float* base = R.part[s][i] + static_cast<size_t>(c % ns) * R.part_center[s][i] + off[t];
                TSource& src = f.src[f.nsrc++];
                src.num = base;
                // rolling output: the weights come from the corner plane
                src.den = R.legacy ? base + static_cast<size_t>(g[t].ntx) * g[t].nty * g[t].pitch * g[t].ah : nullptr;
                src.tiles = g[t];
                src.range = range[t];
            }
            f.block = 8;
            f.step = stg.plane[i].step;
            f.corners = R.legacy ? nullptr : R.den[s][i] + static_cast<size_t>(n % R.acc_n) * npx;
            f.fallback = R.src[i] + static_cast<size_t>(n % R.src_n) * npx;
            f.dst = to_out ? R.out[i] + static_cast<size_t>(n % R.out_n) * npx
                           : R.basic[i] + static_cast<size_t>(n % R.basic_n) * npx;
            f.width = h->pw[i];
            f.height = h->ph[i];
            f.tiles_x = (h->pw[i] + 63) / 64;
            f.wg_begin = nwg;
            nwg += f.tiles_x * ((h->ph[i] + 15) / 16);
        }
        hipError_t e = traced(h, 3, st, [&] { return launch_finish_temporal_tiles(fn, nwg, st); });
        if (e != hipSuccess) return hip_fail(e, "finish launch");
        if (R.legacy) return NSS_HIP_OK;
        // the frame's corner planes start empty for the frame that reuses them
        ZeroTable zt{};
        int zwg = 0;
        for (int k = 0; k < fn.n; ++k) {
            const int i = bk.planes[k];
            const long long npx = static_cast<long long>(h->pw[i]) * h->ph[i];
            ZeroJob& z = zt.job[zt.n++];
            z.p = R.den[s][i] + static_cast<size_t>(n % R.acc_n) * npx;
            z.count = npx;
            z.wg_begin = zwg;
            zwg += static_cast<int>((npx + 1023) / 1024);
        }
        e = traced(h, 0, st, [&] { return launch_zero(zt, zwg, st); });
        return e == hipSuccess ? NSS_HIP_OK : hip_fail(e, "clear launch");
    }
    for (int bi = 0; bi < stg.nbuckets; ++bi) {
        const Bucket& bk = stg.bucket[bi];
        FinishTable fn{};
        ZeroTable zt{};
        int nwg = 0, zwg = 0;
        for (int k = 0; k < bk.nplanes; ++k) {
            const int i = bk.planes[k];
            const size_t npx = static_cast<size_t>(h->pw[i]) * h->ph[i];
            const size_t slot = static_cast<size_t>(n % R.acc_n) * npx;
            float* dst = to_out ? R.out[i] + static_cast<size_t>(n % R.out_n) * npx
                                : R.basic[i] + static_cast<size_t>(n % R.basic_n) * npx;
            FinishJob& f = fn.job[fn.n++];
            f.num = R.num[s][i] + slot;
            f.den = R.den[s][i] + slot;
            f.src = R.src[i] + static_cast<size_t>(n % R.src_n) * npx;
            f.dst = dst;
            f.width = h->pw[i];
            f.height = h->ph[i];
            f.sstride = h->pw[i];
            f.dstride = h->pw[i];
            f.tiles_x = (h->pw[i] + 15) / 16;
            f.wg_begin = nwg;
            nwg += f.tiles_x * ((h->ph[i] + 15) / 16);
            for (unsigned long long* p : {R.num[s][i] + slot, R.den[s][i] + slot}) {
                ZeroJob& z = zt.job[zt.n++];
                z.p = p;
                z.count = static_cast<long long>(npx);
                z.wg_begin = zwg;
                zwg += static_cast<int>((npx + 1023) / 1024);
            }
        }
        hipError_t e = traced(h, 3, st, [&] { return launch_finish(bk.block, fn, nwg, false, st); });
        if (e != hipSuccess) return hip_fail(e, "finish launch");
        e = traced(h, 0, st, [&] { return launch_zero(zt, zwg, st); });
        if (e != hipSuccess) return hip_fail(e, "clear launch");
    }
    return NSS_HIP_OK;
}

// Runs every centre that the frames pushed so far allow, in order, and
// finishes the frames they complete.
// Legacy temporal output of centre c (src/host/filter_bm3d.cpp:322-334):
// for each slot t, the num and den sums of c's groups' patches taken from
// slot t's frame, as float planes at rows 2t and 2t + 1 (in plane heights) of
// output frame c; slots outside the clip, and clamped repeats of an edge
// frame, hold zeros. Planes not filtered get the identity contribution of
// temporal_identity (src/frontend/temporal.hpp:23-36): the source in the
// centre slice with weight 1.
int legacy_emit(nss_hip_bm3d* h, int c, hipStream_t st) {
    auto& R = h->roll;
    const int s = h->nstages - 1;
    const Stage& stg = h->stages[s];
    const int ns = 2 * R.r + 1;
    const Slots slots{c, R.N, R.r};
    const int vb = slots.valid_begin(), ve = slots.valid_end();
    auto fat = [&](int i) {
        const size_t npx = static_cast<size_t>(h->pw[i]) * h->ph[i];
        return R.out[i] + static_cast<size_t>(c % R.out_n) * npx * R.out_rows;
    };
    for (int i = 0; i < h->nplanes; ++i) {
        const size_t npx = static_cast<size_t>(h->pw[i]) * h->ph[i];
        float* f = fat(i);
        const bool filtered = stg.plane[i].sigma != 0.f;
        // zeros in every slice the kernels below do not write
        for (int t = 0; t < ns; ++t) {
            if (filtered && (R.tiles[s] || (t >= vb && t < ve))) continue;
            hipError_t e = hipMemsetAsync(f + static_cast<size_t>(2 * t) * npx, 0, 2 * npx * 4, st);
            if (e != hipSuccess) return hip_fail(e, "slice clear");
        }
        if (!filtered) {
            hipError_t e = hipMemcpyAsync(f + static_cast<size_t>(2 * R.r) * npx,
                                          R.src[i] + static_cast<size_t>(c % R.src_n) * npx, npx * 4,
                                          hipMemcpyDeviceToDevice, st);
            if (e == hipSuccess) e = hipMemsetD32Async(f + static_cast<size_t>(2 * R.r + 1) * npx, 0x3f800000u, npx, st);
            if (e != hipSuccess) return hip_fail(e, "identity slice");
        }
    }
    for (int bi = 0; bi < stg.nbuckets; ++bi) {
        const Bucket& bk = stg.bucket[bi];
        FinishTable fn{};
        int nwg = 0;
        auto flush = [&]() -> int {
            if (!fn.n) return NSS_HIP_OK;
            const hipError_t e = traced(h, 3, st, [&] { return launch_finish(bk.block, fn, nwg, R.tiles[s], st); });
            fn = FinishTable{};
            nwg = 0;
            return e == hipSuccess ? NSS_HIP_OK : hip_fail(e, "slice launch");
        };
        for (int k = 0; k < bk.nplanes; ++k) {
            const int i = bk.planes[k];
            const size_t npx = static_cast<size_t>(h->pw[i]) * h->ph[i];
            TileGeom g[kMaxTileSlots];
            int range[kMaxTileSlots];
            long long off[kMaxTileSlots];
            if (R.tiles[s]) {
                slot_geoms(h, s, i, R.tw[s], R.th[s], g, range);
                centre_partials(g, ns, off, R.legacy);
            }
            for (int t = 0; t < ns; ++t) {
                if (!R.tiles[s] && (t < vb || t >= ve)) continue;
                if (fn.n == kMaxJobs) {
                    const int rc = flush();
                    if (rc != NSS_HIP_OK) return rc;
                }
                FinishJob& f = fn.job[fn.n++];
                if (R.tiles[s]) {
                    f.part_num = R.part[s][i] + static_cast<size_t>(c % ns) * R.part_center[s][i] + off[t];
                    f.part_den = f.part_num + static_cast<size_t>(g[t].ntx) * g[t].nty * g[t].pitch * g[t].ah;
                    f.tiles = g[t];
                    f.range = range[t];
                    f.tiles_x = (h->pw[i] + 63) / 64;
                } else {
                    const size_t slot = static_cast<size_t>(slots.frame(t) % R.acc_n) * npx;
                    f.num = R.num[s][i] + slot;
                    f.den = R.den[s][i] + slot;
                    f.tiles_x = (h->pw[i] + 15) / 16;
                }
                f.block = bk.block;
                f.step = stg.plane[i].step;
                f.dst = fat(i) + static_cast<size_t>(2 * t) * npx;
                f.dst_den = f.dst + npx;
                f.width = h->pw[i];
                f.height = h->ph[i];
                f.sstride = f.dstride = h->pw[i];
                f.wg_begin = nwg;
                nwg += f.tiles_x * ((h->ph[i] + 15) / 16);
            }
        }
        const int rc = flush();
        if (rc != NSS_HIP_OK) return rc;
        if (!R.tiles[s])
            for (int k = 0; k < bk.nplanes; ++k) {
                // the next centre starts from empty sums
                const int i = bk.planes[k];
                const size_t bytes = static_cast<size_t>(h->pw[i]) * h->ph[i] * 8 * R.acc_n;
                hipError_t e = hipMemsetAsync(R.num[s][i], 0, bytes, st);
                if (e == hipSuccess) e = hipMemsetAsync(R.den[s][i], 0, bytes, st);
                if (e != hipSuccess) return hip_fail(e, "sum clear");
            }
    }
    return NSS_HIP_OK;
}

int temporal_advance(nss_hip_bm3d* h, hipStream_t st) {
    auto& R = h->roll;
    const int last = h->nstages - 1;
    for (bool progress = true; progress;) {
        progress = false;
        for (int s = 0; s <= last; ++s) {
            const int c = R.next_center[s];
            if (c >= R.N) continue;
            const int need = std::min(c + R.r, R.N - 1);  // newest frame centre c reads
            const bool ready = s == 0 ? R.pushed > need : R.finished[0] > need;
            if (!ready) continue;
            int rc = temporal_center(h, s, c, st);
            if (rc != NSS_HIP_OK) return rc;
            R.next_center[s] = c + 1;
            if (R.legacy && s == last) {
                rc = legacy_emit(h, c, st);
                if (rc != NSS_HIP_OK) return rc;
                R.finished[s] = c + 1;
                progress = true;
                continue;
            }
            while (R.finished[s] < R.N && R.next_center[s] > std::min(R.finished[s] + R.r, R.N - 1)) {
                rc = temporal_finish(h, s, R.finished[s], st);
                if (rc != NSS_HIP_OK) return rc;
                ++R.finished[s];
            }
            progress = true;
        }
    }
    return NSS_HIP_OK;
}

// A lease of one frame slot for the time a call queues its frame: the slot
// free the longest (src/cuda/runtime/stream_pool.hpp SlotPool::acquire, FIFO
// so that frames rotate over the slots), waiting asleep while all are taken.
struct SlotLease {
    nss_hip_bm3d* h;
    int index;
    explicit SlotLease(nss_hip_bm3d* hh) : h(hh) {
        std::unique_lock<std::mutex> lock(h->slot_mu);
        h->slot_cv.wait(lock, [&] { return !h->free_slots.empty(); });
        index = h->free_slots.front();
        h->free_slots.pop_front();
    }
    ~SlotLease() {
        {
            std::lock_guard<std::mutex> lock(h->slot_mu);
            h->free_slots.push_back(index);
        }
        h->slot_cv.notify_one();
    }
    SlotLease(const SlotLease&) = delete;
    SlotLease& operator=(const SlotLease&) = delete;
};

// Every slot at once (the autotuner changes the plan and the buffers): waits
// until no call holds a slot, then holds them all.
struct AllSlotsLease {
    nss_hip_bm3d* h;
    explicit AllSlotsLease(nss_hip_bm3d* hh) : h(hh) {
        std::unique_lock<std::mutex> lock(h->slot_mu);
        h->slot_cv.wait(lock, [&] { return h->free_slots.size() == h->slots.size(); });
        h->free_slots.clear();
    }
    ~AllSlotsLease() {
        {
            std::lock_guard<std::mutex> lock(h->slot_mu);
            for (size_t k = 0; k < h->slots.size(); ++k) h->free_slots.push_back(static_cast<int>(k));
        }
        h->slot_cv.notify_all();
    }
    AllSlotsLease(const AllSlotsLease&) = delete;
    AllSlotsLease& operator=(const AllSlotsLease&) = delete;
};

// One radius-0 frame on slot k, queued on stream s after the slot's previous
// work (which may have been queued on another stream).
int process_on_slot(nss_hip_bm3d* h, int k, const float* const* src, const int* src_stride, const float* const* ref,
                    const int* ref_stride, float* const* dst, const int* dst_stride, hipStream_t s) {
    FrameSlot& sl = h->slots[static_cast<size_t>(k)];
    if (sl.used) {
        const hipError_t e = hipStreamWaitEvent(s, sl.done, 0);
        if (e != hipSuccess) return hip_fail(e, "stream wait");
    }
    FrameRef f = frame_ref(h, src, src_stride, ref, ref_stride, dst, dst_stride);
    f.buf = sl.buf;
    const int rc = run_batch(&f, 1, &h->tracer, s);
    // Recorded whatever run_batch returned: the next lessee orders after
    // everything this call queued.
    const hipError_t e = hipEventRecord(sl.done, s);
    if (e != hipSuccess) return hip_fail(e, "event record");
    sl.used = true;
    return rc;
}

// Temporal calls: one at a time, each ordered on the device after the last.
struct RollTurn {
    nss_hip_bm3d* h;
    hipStream_t s;
    std::lock_guard<std::mutex> lock;
    hipError_t err = hipSuccess;
    RollTurn(nss_hip_bm3d* hh, hipStream_t st) : h(hh), s(st), lock(hh->roll_mu) {
        if (h->roll_used) err = hipStreamWaitEvent(s, h->roll_done, 0);
    }
    ~RollTurn() {
        if (hipEventRecord(h->roll_done, s) == hipSuccess) h->roll_used = true;
    }
};

// Partials of every slot grown to need[i] int64 entries per half, after the
// slots' earlier work (the caller holds all slots).
int grow_partials(nss_hip_bm3d* h, const size_t* need) {
    hipError_t e = hipSuccess;
    for (auto& sl : h->slots)
        if (e == hipSuccess && sl.used) e = hipEventSynchronize(sl.done);
    if (e != hipSuccess) return hip_fail(e, "synchronize");
    // All new buffers first; the old ones (and the plan that uses them) stay
    // when any allocation fails.
    std::vector<unsigned long long*> fresh;
    for (int i = 0; i < h->nplanes && e == hipSuccess; ++i) {
        if (need[i] <= h->part_elems[i]) continue;
        const size_t gw = plane_needs(h, i).gw;
        for (size_t k = 0; k < h->slots.size() && e == hipSuccess; ++k) {
            unsigned long long* p = nullptr;
            e = dev_alloc(&p, (need[i] + gw) * sizeof(unsigned long long));
            if (e == hipSuccess) fresh.push_back(p);
        }
    }
    if (e != hipSuccess) {
        for (auto* p : fresh) dev_free(p);
        return hip_fail(e, "device allocation");
    }
    size_t next = 0;
    for (int i = 0; i < h->nplanes; ++i) {
        if (need[i] <= h->part_elems[i]) continue;
        for (auto& sl : h->slots) {
            dev_free(sl.buf.part[i]);
            sl.buf.part[i] = fresh[next++];
        }
        h->part_elems[i] = need[i];
    }
    return NSS_HIP_OK;
}

}  // namespace

namespace nsship {

// The host part of an instance (the node creation of
// src/host/filter_bm3d.cpp:1143-1163 without its buffers): format checks,
// plane sizes, the stages and their launch plans, the tile partial sizes and,
// for radius > 0, the rolling plan. Allocates no device memory.
int plan_instance(nss_hip_bm3d* h, const nss_hip_bm3d_params* params, const nss_hip_video_info* info, int device) {
    if (info->sample_type != 1 || info->bits_per_sample != 32 || info->width <= 0 || info->height <= 0 ||
        info->num_planes < 1 || info->num_planes > 3)
        return fail(NSS_HIP_INVALID_INPUT, "nss_hip.BM3D: constant Gray/YUV/RGB 32-bit float required");
    if (params->radius != 0 && (info->width > 8191 || info->height > 8191))
        return fail(NSS_HIP_UNSUPPORTED, "nss_hip: radius > 0 supports planes up to 8191 x 8191");
    if (params->radius != 0 && info->num_frames < 1)
        return fail(NSS_HIP_INVALID_INPUT, "nss_hip: radius > 0 needs the clip length");
    h->p = *params;
    h->vi = *info;
    h->device = device;
    h->nplanes = info->num_planes;
    for (int i = 0; i < h->nplanes; ++i) {
        h->pw[i] = i ? info->width >> info->sub_sampling_w : info->width;
        h->ph[i] = i ? info->height >> info->sub_sampling_h : info->height;
    }
    const nss_hip_bm3d_params& p = *params;
    if (p.final_stage) {
        Stage& b = h->stages[h->nstages++];
        b.wiener = false;
        for (int i = 0; i < 3; ++i)
            b.plane[i] = StagePlane{p.sigma_basic[i], p.block_size_basic[i], p.group_size_basic[i],
                                    p.block_step_basic[i], p.bm_range[i], p.ps_num_basic[i], p.ps_range[i]};
    }
    {
        Stage& f = h->stages[h->nstages++];
        f.wiener = p.final_stage || p.has_ref;
        for (int i = 0; i < 3; ++i)
            f.plane[i] = StagePlane{p.sigma[i], p.block_size[i], p.group_size[i], p.block_step[i], p.bm_range[i],
                                    p.ps_num[i], p.ps_range[i]};
    }
    for (int s = 0; s < h->nstages; ++s) {
        for (int i = 0; i < h->nplanes; ++i) {
            if (h->stages[s].plane[i].sigma == 0.f) continue;
            const int rc = check_stage_plane(h->stages[s].plane[i], h->pw[i], h->ph[i]);
            if (rc != NSS_HIP_OK) return rc;
        }
        const int rc = plan_stage(h, h->stages[s]);
        if (rc != NSS_HIP_OK) return rc;
    }
    const bool rolling = p.radius > 0;
    for (int i = 0; i < h->nplanes; ++i) h->part_elems[i] = plane_needs(h, i).part;
    if (rolling) {
        // Ring depths: the oldest frame a pending centre or finish reads is
        // 4r (final=1) or 2r (one stage) behind the newest pushed; the Wiener
        // stage reads basic frames up to 3r behind the newest finished one;
        // 4r + 1 finished frames can wait for pop at the clip's end.
        auto& R = h->roll;
        const int r = p.radius;
        R.N = info->num_frames;
        R.r = r;
        R.src_n = p.final_stage ? 4 * r + 2 : 2 * r + 2;
        R.ref_n = (p.has_ref && !p.final_stage) ? 2 * r + 2 : 0;
        R.basic_n = p.final_stage ? 3 * r + 2 : 0;
        R.out_n = 4 * r + 2;
        R.acc_n = 2 * r + 1;
        R.legacy = !p.temporal_mode;
        R.out_rows = R.legacy ? 2 * (2 * r + 1) : 1;
        for (int s = 0; s < h->nstages; ++s) plan_temporal_tiles(h, s);
        for (int i = 0; i < h->nplanes; ++i)
            for (int s = 0; s < h->nstages; ++s) {
                if (h->stages[s].plane[i].sigma == 0.f || !R.tiles[s]) continue;
                TileGeom g[kMaxTileSlots];
                int range[kMaxTileSlots];
                slot_geoms(h, s, i, R.tw[s], R.th[s], g, range);
                R.part_center[s][i] = centre_partials(g, 2 * r + 1, nullptr, R.legacy);
            }
    }
    return NSS_HIP_OK;
}

}  // namespace nsship

extern "C" {

const char* nss_hip_last_error(void) { return g_last_error.c_str(); }

int nss_hip_backend_info(int device, char* text, int len) {
    if (!text || len < 1) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: no output buffer");
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess) return hip_fail(e, "device count");
    if (device < 0 || device >= count) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: device index out of range");
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return hip_fail(e, "device properties");
    std::snprintf(text, static_cast<size_t>(len), "nss_hip 1.0 device=%d name=%s arch=%s cus=%d hbm_bytes=%zu",
                  device, prop.name, prop.gcnArchName, prop.multiProcessorCount, prop.totalGlobalMem);
    return NSS_HIP_OK;
}

int nss_hip_bm3d_create(const nss_hip_bm3d_params* params, const nss_hip_video_info* info, int device,
                        nss_hip_bm3d** out) {
    if (!params || !info || !out) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: null argument");
    *out = nullptr;
    DeviceGuard guard(device);
    if (guard.err != hipSuccess) return hip_fail(guard.err, "set device");
    auto* h = new (std::nothrow) nss_hip_bm3d();
    if (!h) return fail(NSS_HIP_RESOURCE_FAILURE, "nss_hip: out of host memory");
    int rc = plan_instance(h, params, info, device);
    if (rc == NSS_HIP_OK) rc = ensure_tables(device);
    if (rc != NSS_HIP_OK) {
        delete h;
        return rc;
    }
    const nss_hip_bm3d_params& p = *params;
    const bool rolling = p.radius > 0;
    // memory_limit_mb caps the instance's device memory; a limit too small
    // for the clip is an error at creation, never while frames are filtered
    // (docs/api/cuda/README.md:141, src/cuda/common/group_driver.cpp:491-498).
    // Checked before anything is allocated.
    if (p.memory_limit_mb > 0 && !rolling) {
        // Radius 0: batch the work to fit the limit (docs/api/cuda/README.md:141);
        // an error only when one batch of grid rows does not fit.
        const int n = frame_slot_count(h);
        if (!fit_frame_buffers(h, (static_cast<size_t>(p.memory_limit_mb) << 20) / static_cast<size_t>(n))) {
            const size_t each = frame_buffer_bytes(h);
            delete h;
            return fail(NSS_HIP_INVALID_INPUT, "nss_hip.BM3D: memory_limit_mb is too small for this clip: each of the " +
                                                   std::to_string(n) + " stream(s) needs at least " +
                                                   std::to_string((each >> 20) + 1) + " MiB");
        }
    }
    if (p.memory_limit_mb > 0 && rolling) {
        const size_t need = instance_bytes(h);
        if (need > (static_cast<size_t>(p.memory_limit_mb) << 20)) {
            const int n = frame_slot_count(h);
            const size_t each = need / static_cast<size_t>(n);
            delete h;
            return fail(NSS_HIP_INVALID_INPUT, "nss_hip.BM3D: memory_limit_mb is too small for this clip: each of the " +
                                                   std::to_string(n) + " stream(s) needs at least " +
                                                   std::to_string((each >> 20) + 1) + " MiB");
        }
    }
    h->slots.resize(static_cast<size_t>(frame_slot_count(h)));
    for (size_t k = 0; k < h->slots.size(); ++k) {
        hipError_t ea = alloc_frame_buffers(h, h->slots[k].buf);
        if (ea == hipSuccess) ea = hipEventCreateWithFlags(&h->slots[k].done, hipEventDisableTiming);
        if (ea != hipSuccess) {
            free_buffers(h);
            delete h;
            return hip_fail(ea, "device allocation");
        }
        h->free_slots.push_back(static_cast<int>(k));
    }
    if (rolling) {
        auto& R = h->roll;
        hipError_t e = hipEventCreateWithFlags(&h->roll_done, hipEventDisableTiming);
        for (int i = 0; i < h->nplanes && e == hipSuccess; ++i) {
            const size_t npx = static_cast<size_t>(h->pw[i]) * h->ph[i];
            e = dev_alloc(&R.src[i], npx * 4 * R.src_n);
            if (e == hipSuccess && R.ref_n) e = dev_alloc(&R.ref[i], npx * 4 * R.ref_n);
            if (e == hipSuccess && R.basic_n) e = dev_alloc(&R.basic[i], npx * 4 * R.basic_n);
            if (e == hipSuccess) e = dev_alloc(&R.out[i], npx * 4 * R.out_n * R.out_rows);
            for (int s = 0; s < h->nstages && e == hipSuccess; ++s) {
                if (h->stages[s].plane[i].sigma == 0.f) continue;
                if (R.tiles[s]) {
                    e = dev_alloc(&R.part[s][i], R.part_center[s][i] * (2 * R.r + 1) * sizeof(float));
                    if (e == hipSuccess && !R.legacy) {
                        e = dev_alloc(&R.den[s][i], npx * 8 * R.acc_n);
                        if (e == hipSuccess) e = hipMemset(R.den[s][i], 0, npx * 8 * R.acc_n);
                    }
                    continue;
                }
                e = dev_alloc(&R.num[s][i], npx * 8 * R.acc_n);
                if (e == hipSuccess) e = dev_alloc(&R.den[s][i], npx * 8 * R.acc_n);
                if (e == hipSuccess) e = hipMemset(R.num[s][i], 0, npx * 8 * R.acc_n);
                if (e == hipSuccess) e = hipMemset(R.den[s][i], 0, npx * 8 * R.acc_n);
            }
        }
        if (e != hipSuccess) {
            free_buffers(h);
            delete h;
            return hip_fail(e, "device allocation");
        }
    }
    *out = h;
    return NSS_HIP_OK;
}

int nss_hip_bm3d_resources(nss_hip_bm3d* h, long long* bytes, long long* peak, long long* limit) {
    if (!h || !bytes || !peak || !limit) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: null argument");
    // kinds of include/nss/resources.hpp:11: workspace, cached, pinned,
    // inflight, output, framework
    size_t kind[6] = {frame_buffer_bytes(h) * h->slots.size(), 0, 0, 0, 0, 0};
    for (int i = 0; i < h->nplanes; ++i) rolling_bytes(h, i, &kind[1], &kind[4]);
    size_t total = 0;
    for (int k = 0; k < 6; ++k) {
        bytes[k] = static_cast<long long>(kind[k]);
        total += kind[k];
    }
    *peak = static_cast<long long>(total);  // every buffer lives from creation to destruction
    *limit = h->p.memory_limit_mb > 0 ? static_cast<long long>(h->p.memory_limit_mb) << 20 : 0;
    return NSS_HIP_OK;
}

int nss_hip_bm3d_push(nss_hip_bm3d* h, const float* const* src, const int* src_stride, const float* const* ref,
                      const int* ref_stride, void* stream) {
    if (!h || !src || !src_stride) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: null argument");
    auto& R = h->roll;
    if (h->p.radius == 0) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: push/pop serve radius > 0; use process");
    const bool need_ref = h->p.has_ref && !h->p.final_stage;
    for (int i = 0; i < h->nplanes; ++i) {
        if (!src[i] || src_stride[i] < h->pw[i])
            return fail(NSS_HIP_INVALID_INPUT, "nss_hip: missing plane or stride below width");
        if (need_ref && (!ref || !ref_stride || !ref[i] || ref_stride[i] < h->pw[i]))
            return fail(NSS_HIP_INVALID_INPUT, "nss_hip: ref planes required");
    }
    DeviceGuard guard(h->device);
    if (guard.err != hipSuccess) return hip_fail(guard.err, "set device");
    hipStream_t s = static_cast<hipStream_t>(stream);
    RollTurn turn(h, s);
    if (turn.err != hipSuccess) return hip_fail(turn.err, "stream wait");
    if (R.failed)
        return fail(NSS_HIP_RESOURCE_FAILURE,
                    "nss_hip: an earlier push failed part-way; the instance's rolling state is not consistent "
                    "(create a new instance)");
    if (R.pushed >= R.N) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: every frame of the clip was already pushed");
    const int last = h->nstages - 1;
    const int pending = R.finished[last] - R.out_head;
    const int may_finish = R.pushed + 1 == R.N ? R.N - R.finished[last] : 2;
    if (pending + may_finish > R.out_n)
        return fail(NSS_HIP_RESOURCE_FAILURE, "nss_hip: pop the finished frames before pushing more");
    const int f = R.pushed;
    for (int i = 0; i < h->nplanes; ++i) {
        const size_t npx = static_cast<size_t>(h->pw[i]) * h->ph[i];
        hipError_t e = hipMemcpy2DAsync(R.src[i] + static_cast<size_t>(f % R.src_n) * npx,
                                        static_cast<size_t>(h->pw[i]) * 4, src[i], static_cast<size_t>(src_stride[i]) * 4,
                                        static_cast<size_t>(h->pw[i]) * 4, h->ph[i], hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess && need_ref)
            e = hipMemcpy2DAsync(R.ref[i] + static_cast<size_t>(f % R.ref_n) * npx, static_cast<size_t>(h->pw[i]) * 4,
                                 ref[i], static_cast<size_t>(ref_stride[i]) * 4, static_cast<size_t>(h->pw[i]) * 4,
                                 h->ph[i], hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return hip_fail(e, "frame upload");
    }
    R.pushed = f + 1;
    const int rc = temporal_advance(h, s);
    if (rc != NSS_HIP_OK) R.failed = true;
    return rc;
}

// Random access (rollingGetFrame, src/host/filter_bm3d.cpp:985-1072): frame
// n's chunk of rolling_chunk frames (rolling_chunk_start, :548-550) and the
// source frames its centres read. Rolling output: last-stage centres
// [start - r, end + r], which read frames r further; a first stage adds the
// basic frames those read, 2r further again. Legacy output: the chunk's own
// centres.
static void chunk_window(const nss_hip_bm3d* h, int n, int* start, int* count, int* first, int* last,
                         int* centre0) {
    const auto& R = h->roll;
    const int chunk = std::max(1, h->p.rolling_chunk);
    const int st = (n / chunk) * chunk, cnt = std::min(chunk, R.N - st);
    const int two = h->nstages == 2 ? 1 : 0;
    const int margin = R.r * ((R.legacy ? 1 : 2) + 2 * two);
    *start = st;
    *count = cnt;
    *first = std::max(0, st - margin);
    *last = std::min(R.N - 1, st + cnt - 1 + margin);
    // first centre of each stage (last stage, then the basic one)
    const int c_last = R.legacy ? st : st - R.r;
    centre0[h->nstages - 1] = std::max(0, c_last);
    if (two) centre0[0] = std::max(0, c_last - 2 * R.r);
}

int nss_hip_bm3d_chunk(nss_hip_bm3d* h, int n, int* start, int* count, int* first, int* last) {
    if (!h || !start || !count || !first || !last) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: null argument");
    if (h->p.radius == 0) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: chunks serve radius > 0");
    if (n < 0 || n >= h->roll.N) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: frame index outside the clip");
    int c0[2];
    chunk_window(h, n, start, count, first, last, c0);
    return NSS_HIP_OK;
}

int nss_hip_bm3d_seek(nss_hip_bm3d* h, int n, int* next_push, void* stream) {
    if (!h || !next_push) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: null argument");
    auto& R = h->roll;
    if (h->p.radius == 0) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: seek serves radius > 0");
    if (n < 0 || n >= R.N) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: frame index outside the clip");
    int start, count, first, last, c0[2];
    chunk_window(h, n, &start, &count, &first, &last, c0);
    DeviceGuard guard(h->device);
    if (guard.err != hipSuccess) return hip_fail(guard.err, "set device");
    hipStream_t s = static_cast<hipStream_t>(stream);
    RollTurn turn(h, s);
    if (turn.err != hipSuccess) return hip_fail(turn.err, "stream wait");
    // The next chunk in order carries on from the state as it is.
    if (!R.failed && start == R.out_head && R.pushed >= first) {
        *next_push = R.pushed;
        return NSS_HIP_OK;
    }
    // Else restart at the chunk's first source frame: every centre from the
    // stage's first one runs again. Rolling output also finishes the r
    // frames before a stage's first centre, and every frame before its first
    // complete one (first centre + r): the centres add into their sums, so
    // finishing them clears those sums before the ring reuses them; nothing
    // reads their output.
    for (int st = 0; st < h->nstages; ++st)
        for (int i = 0; i < h->nplanes; ++i) {
            const size_t npx = static_cast<size_t>(h->pw[i]) * h->ph[i];
            for (unsigned long long* p : {R.num[st][i], R.den[st][i]})
                if (p) {
                    const hipError_t e = hipMemsetAsync(p, 0, npx * 8 * R.acc_n, s);
                    if (e != hipSuccess) return hip_fail(e, "clear");
                }
        }
    R.pushed = first;
    for (int st = 0; st < h->nstages; ++st) {
        R.next_center[st] = c0[st];
        R.finished[st] = R.legacy ? c0[st] : std::max(0, c0[st] - R.r);
    }
    R.out_head = start;
    R.failed = false;
    *next_push = first;
    return NSS_HIP_OK;
}

int nss_hip_bm3d_pop(nss_hip_bm3d* h, float* const* dst, const int* dst_stride, int* frame, void* stream) {
    if (!h || !dst || !dst_stride) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: null argument");
    auto& R = h->roll;
    if (h->p.radius == 0) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: push/pop serve radius > 0; use process");
    DeviceGuard guard(h->device);
    if (guard.err != hipSuccess) return hip_fail(guard.err, "set device");
    hipStream_t s = static_cast<hipStream_t>(stream);
    RollTurn turn(h, s);
    if (turn.err != hipSuccess) return hip_fail(turn.err, "stream wait");
    if (R.out_head >= R.finished[h->nstages - 1]) return NSS_HIP_BYPASS;
    const int n = R.out_head;
    for (int i = 0; i < h->nplanes; ++i) {
        if (!dst[i] || dst_stride[i] < h->pw[i])
            return fail(NSS_HIP_INVALID_INPUT, "nss_hip: missing plane or stride below width");
        const size_t fpx = static_cast<size_t>(h->pw[i]) * h->ph[i] * R.out_rows;
        const hipError_t e =
            hipMemcpy2DAsync(dst[i], static_cast<size_t>(dst_stride[i]) * 4, R.out[i] + static_cast<size_t>(n % R.out_n) * fpx,
                             static_cast<size_t>(h->pw[i]) * 4, static_cast<size_t>(h->pw[i]) * 4,
                             static_cast<size_t>(h->ph[i]) * R.out_rows, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return hip_fail(e, "frame download");
    }
    if (frame) *frame = n;
    R.out_head = n + 1;
    return NSS_HIP_OK;
}

int nss_hip_bm3d_slots(const nss_hip_bm3d* handle) { return handle ? 2 * handle->p.radius + 1 : 0; }

int nss_hip_vaggregate(float* dst, int dst_stride, const float* const* nums, const float* const* dens,
                       const int* strides, int count, const float* src, int src_stride, int width, int height,
                       void* stream) {
    if (!dst || !nums || !dens || !strides || !src) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: null argument");
    if (count < 1 || count > kVAggMax)
        return fail(NSS_HIP_INVALID_INPUT, "nss_hip.VAggregate: 1 to 33 contributions per frame");
    if (width < 1 || height < 1 || dst_stride < width || src_stride < width)
        return fail(NSS_HIP_INVALID_INPUT, "nss_hip.VAggregate: invalid plane dimensions or strides");
    VAggJob jb{};
    for (int k = 0; k < count; ++k) {
        if (!nums[k] || !dens[k] || strides[k] < width)
            return fail(NSS_HIP_INVALID_INPUT, "nss_hip.VAggregate: missing contribution or stride below width");
        jb.num[k] = nums[k];
        jb.den[k] = dens[k];
        jb.stride[k] = strides[k];
    }
    jb.count = count;
    jb.src = src;
    jb.dst = dst;
    jb.width = width;
    jb.height = height;
    jb.sstride = src_stride;
    jb.dstride = dst_stride;
    const hipError_t e = launch_vaggregate(jb, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? NSS_HIP_OK : hip_fail(e, "vaggregate launch");
}

int nss_hip_bm3d_process(nss_hip_bm3d* h, const float* const* src, const int* src_stride, const float* const* ref,
                         const int* ref_stride, float* const* dst, const int* dst_stride, int frame, int frame_count,
                         void* stream) {
    (void)frame;
    (void)frame_count;
    if (!h || !src || !src_stride || !dst || !dst_stride) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: null argument");
    if (h->p.radius > 0) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: radius > 0 filters through push/pop");
    for (int i = 0; i < h->nplanes; ++i)
        if (!src[i] || !dst[i] || src_stride[i] < h->pw[i] || dst_stride[i] < h->pw[i])
            return fail(NSS_HIP_INVALID_INPUT, "nss_hip: missing plane or stride below width");
    if (h->p.has_ref && !h->p.final_stage) {
        if (!ref || !ref_stride) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: ref planes required");
        for (int i = 0; i < h->nplanes; ++i)
            if (!ref[i] || ref_stride[i] < h->pw[i])
                return fail(NSS_HIP_INVALID_INPUT, "nss_hip: missing ref plane or stride below width");
    }
    DeviceGuard guard(h->device);
    if (guard.err != hipSuccess) return hip_fail(guard.err, "set device");
    SlotLease lease(h);
    return process_on_slot(h, lease.index, src, src_stride, ref, ref_stride, dst, dst_stride,
                           static_cast<hipStream_t>(stream));
}

int nss_hip_bm3d_autotune(nss_hip_bm3d* h, int repeats, void* stream, double* best_ms) {
    if (!h) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: null handle");
    if (repeats < 1 || repeats > 100) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: repeats must be in [1, 100]");
    if (h->p.radius > 0 || (h->p.has_ref && !h->p.final_stage) || h->batches > 0) return NSS_HIP_OK;  // nothing to tune
    DeviceGuard guard(h->device);
    if (guard.err != hipSuccess) return hip_fail(guard.err, "set device");
    hipStream_t s = static_cast<hipStream_t>(stream);
    AllSlotsLease all(h);
    // Partials large enough for every candidate shape.
    size_t need[3]{};
    for (int k = 0; k < h->nstages; ++k)
        for (int bi = 0; bi < h->stages[k].nbuckets; ++bi) {
            Bucket probe = h->stages[k].bucket[bi];
            if (!probe.tiles) continue;
            for (const auto& sh : kTileShapes)
                if (apply_tile_shape(h, h->stages[k], probe, sh[0], sh[1]))
                    for (int q = 0; q < probe.nplanes; ++q) {
                        const TileGeom& tg = probe.geom[q];
                        need[probe.planes[q]] = std::max(need[probe.planes[q]],
                                                         static_cast<size_t>(tg.ntx) * tg.nty * tg.pitch * tg.ah);
                    }
        }
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "synchronize");
    int rc = grow_partials(h, need);
    if (rc != NSS_HIP_OK) return rc;
    // A synthetic frame: smooth pattern plus pseudo-random noise (sigma 10/255).
    float* src[3]{};
    float* dst[3]{};
    int stride[3]{};
    std::vector<float> host;
    unsigned lcg = 12345u;
    for (int i = 0; i < h->nplanes && rc == NSS_HIP_OK; ++i) {
        const size_t npx = static_cast<size_t>(h->pw[i]) * h->ph[i];
        host.resize(npx);
        for (size_t o = 0; o < npx; ++o) {
            const int x = static_cast<int>(o % h->pw[i]), y = static_cast<int>(o / h->pw[i]);
            lcg = lcg * 1664525u + 1013904223u;
            const float noise = (static_cast<float>(lcg >> 8) / 16777216.f - 0.5f) * (10.f / 255.f) * 3.46f;
            host[o] = 0.5f + 0.25f * std::sin(0.05f * x) * std::cos(0.07f * y) + noise;
        }
        stride[i] = h->pw[i];
        e = dev_alloc(&src[i], npx * 4);
        if (e == hipSuccess) e = dev_alloc(&dst[i], npx * 4);
        if (e == hipSuccess) e = hipMemcpy(src[i], host.data(), npx * 4, hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = hip_fail(e, "autotune buffers");
    }
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (rc == NSS_HIP_OK && (hipEventCreate(&ev[0]) != hipSuccess || hipEventCreate(&ev[1]) != hipSuccess))
        rc = fail(NSS_HIP_RESOURCE_FAILURE, "nss_hip: event create");
    const float* csrc[3] = {src[0], src[1], src[2]};
    auto time_frame = [&](double* ms) -> int {
        int r = process_on_slot(h, 0, csrc, stride, nullptr, nullptr, dst, stride, s);
        if (r != NSS_HIP_OK) return r;
        (void)hipEventRecord(ev[0], s);
        for (int k = 0; k < repeats && r == NSS_HIP_OK; ++k)
            r = process_on_slot(h, 0, csrc, stride, nullptr, nullptr, dst, stride, s);
        (void)hipEventRecord(ev[1], s);
        if (r != NSS_HIP_OK) return r;
        const hipError_t x = hipEventSynchronize(ev[1]);
        if (x != hipSuccess) return hip_fail(x, "autotune timing");
        float t = 0.f;
        (void)hipEventElapsedTime(&t, ev[0], ev[1]);
        *ms = t / repeats;
        return NSS_HIP_OK;
    };
    double best = 0.0;
    if (rc == NSS_HIP_OK) rc = time_frame(&best);
    // One bucket at a time, every shape that fits; the others stay as set.
    for (int k = 0; k < h->nstages && rc == NSS_HIP_OK; ++k)
        for (int bi = 0; bi < h->stages[k].nbuckets && rc == NSS_HIP_OK; ++bi) {
            Bucket& bk = h->stages[k].bucket[bi];
            if (!bk.tiles) continue;
            Bucket chosen = bk;
            for (const auto& sh : kTileShapes) {
                Bucket trial = chosen;
                if (!apply_tile_shape(h, h->stages[k], trial, sh[0], sh[1])) continue;
                bk = trial;
                double ms = 0.0;
                rc = time_frame(&ms);
                if (rc != NSS_HIP_OK) break;
                if (ms < best) {
                    best = ms;
                    chosen = trial;
                }
            }
            bk = chosen;
        }
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    (void)hipStreamSynchronize(s);
    for (int i = 0; i < 3; ++i) {
        if (src[i]) dev_free(src[i]);
        if (dst[i]) dev_free(dst[i]);
    }
    if (rc == NSS_HIP_OK && best_ms) *best_ms = best;
    return rc;
}

int nss_hip_bm3d_set_tiles(nss_hip_bm3d* h, int stage, int tw, int th) {
    if (!h) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: null handle");
    if (stage < 0 || stage >= h->nstages) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: no such stage");
    if (tw < 1 || th < 1 || tw * th > 64) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: tile of 1 to 64 references");
    if (h->p.radius > 0) return fail(NSS_HIP_UNSUPPORTED, "nss_hip: temporal instances keep their tile plan");
    if (h->batches > 0)
        return fail(NSS_HIP_UNSUPPORTED, "nss_hip: an instance batched to fit memory_limit_mb has no tile plan");
    DeviceGuard guard(h->device);
    if (guard.err != hipSuccess) return hip_fail(guard.err, "set device");
    AllSlotsLease all(h);
    Stage& st = h->stages[stage];
    Bucket trial[7];
    size_t need[3]{};
    for (int bi = 0; bi < st.nbuckets; ++bi) {
        trial[bi] = st.bucket[bi];
        if (!trial[bi].tiles) continue;
        if (!apply_tile_shape(h, st, trial[bi], tw, th))
            return fail(NSS_HIP_UNSUPPORTED, "nss_hip: the tile shape does not fit this stage");
        for (int q = 0; q < trial[bi].nplanes; ++q) {
            const TileGeom& tg = trial[bi].geom[q];
            need[trial[bi].planes[q]] = std::max(need[trial[bi].planes[q]],
                                                 static_cast<size_t>(tg.ntx) * tg.nty * tg.pitch * tg.ah);
        }
    }
    const int rc = grow_partials(h, need);
    if (rc != NSS_HIP_OK) return rc;
    for (int bi = 0; bi < st.nbuckets; ++bi) st.bucket[bi] = trial[bi];
    return NSS_HIP_OK;
}

int nss_hip_bm3d_set_tracing(nss_hip_bm3d* h, int on) {
    if (!h) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: null handle");
    h->tracer.on = on != 0;
    return NSS_HIP_OK;
}

int nss_hip_bm3d_kernel_times(nss_hip_bm3d* h, int* launches, double* total_ms, int reset) {
    if (!h || !launches || !total_ms) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: null argument");
    DeviceGuard guard(h->device);
    std::lock_guard<std::mutex> lock(h->tracer.mu);
    for (int k = 0; k < NSS_HIP_KERNEL_KINDS; ++k) {
        launches[k] = static_cast<int>(h->tracer.spans[k].size());
        double sum = 0;
        for (auto& span : h->tracer.spans[k]) {
            hipError_t e = hipEventSynchronize(span.second);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, span.first, span.second);
            if (e != hipSuccess) return hip_fail(e, "event timing");
            sum += ms;
        }
        total_ms[k] = sum;
    }
    if (reset) h->tracer.reset();
    return NSS_HIP_OK;
}

int nss_hip_bm3d_destroy(nss_hip_bm3d* h) {
    if (!h) return NSS_HIP_OK;
    {
        DeviceGuard guard(h->device);
        (void)hipDeviceSynchronize();
        free_buffers(h);
    }
    delete h;
    return NSS_HIP_OK;
}

int nss_hip_set_match_path(int path) {
    if (path < 0 || path > 2) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: match path must be 0, 1 or 2");
    g_match_path.store(path, std::memory_order_relaxed);
    return NSS_HIP_OK;
}

int nss_hip_spatial_match(const float* plane, int width, int height, int stride, int block, int step, int bm_range,
                          int group, int* out_rec, int* out_count, void* stream) {
    if (!plane || !out_rec || !out_count) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: null argument");
    if (!allowed_block(block) || group < 1 || group > kMaxGroup || width < block || height < block || step < 1 ||
        stride < width || bm_range < 0 || bm_range > 64)
        return fail(NSS_HIP_INVALID_INPUT, "nss_hip: invalid match arguments");
    MatchTable mt{};
    MatchJob& m = mt.job[mt.n++];
    m.plane = plane;
    m.width = width;
    m.height = height;
    m.stride = stride;
    m.grid = make_grid(width, height, block, step);
    m.range = bm_range;
    m.group = group;
    m.t0 = 0;
    m.wg_begin = 0;
    m.out = reinterpret_cast<MatchRec*>(out_rec);
    m.counts = out_count;
    const int nref = m.grid.nx * m.grid.ny;
    if (nref == 0) return NSS_HIP_OK;
    size_t lds;
    const int kind = choose_match_kind(block, step, bm_range, group, width, height, nref, &lds);
    const int wg = match_workgroups(kind, m.grid);
    const hipError_t e = launch_match(block, mt, wg, lds, kind, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(e, "match launch");
    return NSS_HIP_OK;
}

int nss_hip_predictive_match(const float* frames, int nframes, int center, int width, int height, int stride,
                             int block, int step, int bm_range, int group, int radius, int ps_num, int ps_range,
                             int* out_rec, int* out_count, void* stream) {
    if (!frames || !out_rec || !out_count) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: null argument");
    if (!allowed_block(block) || group < 1 || group > kMaxGroup || width < block || height < block || step < 1 ||
        step > block || stride < width || bm_range < 0 || bm_range > 64 || radius < 1 || radius > kMaxRadius ||
        ps_num < 1 || ps_num > group || ps_range < 0 || ps_range > 64 || nframes < 1 || center < 0 ||
        center >= nframes)
        return fail(NSS_HIP_INVALID_INPUT, "nss_hip: invalid predictive match arguments");
    if (width > 8191 || height > 8191)
        return fail(NSS_HIP_UNSUPPORTED, "nss_hip: radius > 0 supports planes up to 8191 x 8191");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const FrameRing ring{frames, static_cast<long long>(stride) * height, nframes};
    const Slots slots{center, nframes, radius};
    MatchTable mt{};
    MatchJob& m = mt.job[mt.n++];
    m.plane = ring.frame(slots.frame(radius));
    m.width = width;
    m.height = height;
    m.stride = stride;
    m.grid = make_grid(width, height, block, step);
    m.range = bm_range;
    m.group = group;
    m.t0 = radius;
    m.wg_begin = 0;
    m.out = reinterpret_cast<MatchRec*>(out_rec);
    m.counts = out_count;
    const int nref = m.grid.nx * m.grid.ny;
    if (nref == 0) return NSS_HIP_OK;
    size_t lds;
    const int kind = choose_match_kind(block, step, bm_range, group, width, height, nref, &lds);
    hipError_t e = launch_match(block, mt, match_workgroups(kind, m.grid), lds, kind, s);
    if (e != hipSuccess) return hip_fail(e, "match launch");
    // group 1 keeps the reference alone (src/cpu/bm/matcher.hpp:323).
    if (group == 1) return NSS_HIP_OK;
    PredTable pt{};
    PredJob& pj = pt.job[pt.n++];
    pj.match = ring;
    pj.width = width;
    pj.height = height;
    pj.stride = stride;
    pj.nref = nref;
    pj.group = group;
    pj.ps_num = ps_num;
    pj.ps_range = ps_range;
    pj.slots = slots;
    pj.wg_begin = 0;
    pj.recs = m.out;
    pj.counts = out_count;
    const size_t plds = predictive_table_lds(block, pt, radius);
    e = launch_predictive(block, pt, nref, plds, s);
    return e == hipSuccess ? NSS_HIP_OK : hip_fail(e, "predictive match launch");
}

}  // extern "C"
