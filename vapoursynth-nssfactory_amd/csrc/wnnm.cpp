// nss_hip.WNNM and nss_hip.NCSR behind the C ABI: one spatial frame per
// call, every step on the caller's stream.
//
// Per active plane (src/host/filter_wnnm.cpp:44-160 process_plane_batched at
// radius 0): clear the plane's fixed-point sums, match every raster reference
// on rclip (or src) with the BM3D matchers, shrink each group and add it into
// the sums (wnnm_group_kernel), then finish dst = num / den, or src where no
// patch landed (the atomic path's corner finish). A plane with sigma 0 is
// copied (src/host/filter_wnnm.cpp:236-246).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <new>
#include <string>

#include "instance.hpp"

using namespace nsship;

struct nss_hip_wnnm {
    nss_hip_wnnm_params p{};
    nss_hip_video_info vi{};
    int device = 0;
    bool ncsr = false;  // NCSR: iterations on the estimate, the code shrink
    int iters = 1;
    float delta = 0.f;
    struct Plane {
        int width = 0, height = 0;
        Grid grid{};
        int nref = 0;
        bool active = false;
        MatchRec* recs = nullptr;
        int* counts = nullptr;
        unsigned long long* num = nullptr;  // 2 radius + 1 slices when temporal
        unsigned long long* den = nullptr;
        float* est = nullptr;  // NCSR estimate, width floats per row; a ring like win when temporal
        // Temporal: the window's frames (frame f at (f % (2 radius + 1)) * pixels),
        // and rclip's when one is given.
        float* win = nullptr;
        float* mwin = nullptr;
    } pl[3];
    long long bytes = 0;
    std::mutex mu;  // temporal calls: one at a time (the window and sums are shared)
};

// The NCSR instance is the same state with the NCSR fields set.
struct nss_hip_ncsr : nss_hip_wnnm {};

namespace {

void release(nss_hip_wnnm* h) {
    for (auto& q : h->pl) {
        if (q.recs) dev_free(q.recs);
        if (q.counts) dev_free(q.counts);
        if (q.num) dev_free(q.num);
        if (q.den) dev_free(q.den);
        if (q.est) dev_free(q.est);
        dev_free(q.win);
        dev_free(q.mwin);
        q.win = q.mwin = nullptr;
        q.est = nullptr;
        q.recs = nullptr;
        q.counts = nullptr;
        q.num = nullptr;
        q.den = nullptr;
    }
}

// Buffers of every active plane; h->ncsr and the parameters are set.
int init_instance(nss_hip_wnnm* h, const nss_hip_wnnm_params& p, const nss_hip_video_info* info, int device,
                  const char* name) {
    const std::string ns = std::string("nss_hip.") + name;
    if (info->num_planes < 1 || info->num_planes > 3 || info->width < 1 || info->height < 1)
        return fail(NSS_HIP_INVALID_INPUT, ns + ": invalid video info");
    if (p.radius < 0 || p.radius > kMaxRadius || p.block_size < 1 ||
        p.block_size > 16 || p.group_size < 1 || p.group_size > 32 || p.block_step < 1 ||
        p.block_step > p.block_size || p.bm_range < 1 || p.bm_range > 64 ||
        (p.radius > 0 && (p.ps_num < 1 || p.ps_num > p.group_size || p.ps_range < 1 || p.ps_range > 64)))
        return fail(NSS_HIP_INVALID_INPUT, ns + ": parameters not resolved by the parse call");
    if (p.radius > 0 && (info->num_frames < 1 || info->width > 8191 || info->height > 8191))
        return fail(NSS_HIP_INVALID_INPUT, ns + ": radius > 0 needs num_frames and planes up to 8191 x 8191");
    if (p.radius > 0 && p.block_size != 1 && p.block_size != 2 && p.block_size != 4 && p.block_size != 8 &&
        p.block_size != 12 && p.block_size != 16)
        return fail(NSS_HIP_UNSUPPORTED, ns + ": radius > 0 takes block_size 1, 2, 4, 8, 12 or 16 in nss_hip "
                                              "(the predictive search's sizes)");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess) return hip_fail(e, "device count");
    if (device < 0 || device >= ndev) return fail(NSS_HIP_INVALID_INPUT, ns + ": device_id out of range");
    DeviceGuard dg(device);
    if (dg.err != hipSuccess) return hip_fail(dg.err, "set device");
    h->p = p;
    h->vi = *info;
    h->device = device;
    for (int i = 0; i < info->num_planes; ++i) {
        auto& q = h->pl[i];
        q.width = info->width >> (i ? info->sub_sampling_w : 0);
        q.height = info->height >> (i ? info->sub_sampling_h : 0);
        q.active = p.sigma[i] != 0.f;
        if (!q.active) continue;
        if (q.width < p.block_size || q.height < p.block_size) {
            release(h);
            return fail(NSS_HIP_INVALID_INPUT, "nss: selected plane is smaller than block_size");
        }
        q.grid = make_grid(q.width, q.height, p.block_size, p.block_step);
        q.nref = q.grid.nx * q.grid.ny;
        const size_t pixels = static_cast<size_t>(q.width) * q.height;
        const size_t rec_bytes = static_cast<size_t>(q.nref) * p.group_size * sizeof(MatchRec);
        const size_t cnt_bytes = static_cast<size_t>(q.nref) * sizeof(int);
        const size_t slices = 2 * static_cast<size_t>(p.radius) + 1;
        const size_t sum_bytes = slices * pixels * sizeof(unsigned long long);
        const size_t est_bytes = h->ncsr ? slices * pixels * sizeof(float) : 0;
        const size_t win_bytes = p.radius > 0 ? slices * pixels * sizeof(float) : 0;
        const size_t mwin_bytes = p.radius > 0 && p.has_rclip ? win_bytes : 0;
        h->bytes += static_cast<long long>(rec_bytes + cnt_bytes + 2 * sum_bytes + est_bytes + win_bytes + mwin_bytes);
        if ((e = dev_alloc(&q.recs, rec_bytes)) != hipSuccess || (e = dev_alloc(&q.counts, cnt_bytes)) != hipSuccess ||
            (e = dev_alloc(&q.num, sum_bytes)) != hipSuccess || (e = dev_alloc(&q.den, sum_bytes)) != hipSuccess ||
            (h->ncsr && (e = dev_alloc(&q.est, est_bytes)) != hipSuccess) ||
            (win_bytes && (e = dev_alloc(&q.win, win_bytes)) != hipSuccess) ||
            (mwin_bytes && (e = dev_alloc(&q.mwin, mwin_bytes)) != hipSuccess)) {
            release(h);
            return hip_fail(e, "group filter buffers");
        }
    }
    if (p.memory_limit_mb > 0 && h->bytes > static_cast<long long>(p.memory_limit_mb) << 20) {
        const long long need = (h->bytes + (1ll << 20) - 1) >> 20;
        release(h);
        return fail(NSS_HIP_INVALID_INPUT, ns + ": memory_limit_mb is too small for this clip: the instance needs at "
                                                "least " + std::to_string(need) + " MiB");
    }
    return NSS_HIP_OK;
}

// One round over every active plane: clear the sums, match on mplane[p],
// shrink the groups of gsrc[p], finish into fdst[p] (fallback fsrc[p]).
int group_round(nss_hip_wnnm* h, const float* const* mplane, const int* mstride, const float* const* gsrc,
                const int* gstride, const float* const* fsrc, const int* fsstride, float* const* fdst,
                const int* fdstride, hipStream_t s) {
    const nss_hip_wnnm_params& p = h->p;
    ZeroTable zt{};
    WnnmTable wt{};
    FinishTable ft{};
    int zero_wg = 0, wnnm_wg = 0, finish_wg = 0;
    hipError_t e;
    for (int i = 0; i < h->vi.num_planes; ++i) {
        const auto& q = h->pl[i];
        if (!q.active) continue;
        const long long pixels = static_cast<long long>(q.width) * q.height;
        for (unsigned long long* buf : {q.num, q.den}) {
            ZeroJob& z = zt.job[zt.n++];
            z.p = buf;
            z.count = pixels;
            z.wg_begin = zero_wg;
            zero_wg += static_cast<int>((pixels + 1023) / 1024);
        }
        WnnmJob& w = wt.job[wt.n++];
        w.src = gsrc[i];
        w.width = q.width;
        w.height = q.height;
        w.stride = gstride[i];
        w.nref = q.nref;
        w.group = p.group_size;
        w.block = p.block_size;
        w.sigma = p.sigma[i] / 255.f;
        w.residual = h->ncsr ? 1 : p.residual;
        w.adaptive = p.adaptive;
        w.wg_begin = wnnm_wg;
        w.matches = q.recs;
        w.counts = q.counts;
        w.num = q.num;
        w.den = q.den;
        wnnm_wg += q.nref;
        FinishJob& f = ft.job[ft.n++];
        f.num = q.num;
        f.den = q.den;
        f.block = p.block_size;
        f.src = fsrc[i];
        f.dst = fdst[i];
        f.dst_den = nullptr;
        f.width = q.width;
        f.height = q.height;
        f.sstride = fsstride[i];
        f.dstride = fdstride[i];
        f.tiles_x = (q.width + 15) / 16;
        f.wg_begin = finish_wg;
        finish_wg += f.tiles_x * ((q.height + 15) / 16);
    }
    if (wt.n == 0) return NSS_HIP_OK;
    if ((e = launch_zero(zt, zero_wg, s)) != hipSuccess) return hip_fail(e, "group filter clear");
    for (int i = 0; i < h->vi.num_planes; ++i) {
        const auto& q = h->pl[i];
        if (!q.active) continue;
        MatchTable mt{};
        MatchJob& m = mt.job[mt.n++];
        m.plane = mplane[i];
        m.stride = mstride[i];
        m.width = q.width;
        m.height = q.height;
        m.grid = q.grid;
        m.range = p.bm_range;
        m.group = p.group_size;
        m.t0 = 0;
        m.wg_begin = 0;
        m.out = q.recs;
        m.counts = q.counts;
        size_t lds = 0;
        const int kind =
            choose_match_kind(p.block_size, p.block_step, p.bm_range, p.group_size, q.width, q.height, q.nref, &lds);
        if ((e = launch_match(p.block_size, mt, match_workgroups(kind, q.grid), lds, kind, s)) != hipSuccess)
            return hip_fail(e, "group filter match launch");
    }
    if ((e = launch_wnnm(wt, wnnm_wg, h->ncsr, s)) != hipSuccess) return hip_fail(e, "group filter launch");
    if ((e = launch_finish(p.block_size, ft, finish_wg, false, s)) != hipSuccess)
        return hip_fail(e, "group filter finish launch");
    return NSS_HIP_OK;
}

int check_planes(const nss_hip_wnnm* h, const float* const* src, const int* src_stride, const float* const* rclip,
                 const int* rclip_stride, float* const* dst, const int* dst_stride, const char* name) {
    const std::string ns = std::string("nss_hip.") + name;
    if (!src || !src_stride || !dst || !dst_stride || (rclip && !rclip_stride))
        return fail(NSS_HIP_INVALID_INPUT, ns + ": null argument");
    for (int i = 0; i < h->vi.num_planes; ++i) {
        const auto& q = h->pl[i];
        if (!src[i] || !dst[i] || (rclip && !rclip[i])) return fail(NSS_HIP_INVALID_INPUT, ns + ": null plane pointer");
        if (src_stride[i] < q.width || dst_stride[i] < q.width || (rclip && rclip_stride[i] < q.width))
            return fail(NSS_HIP_INVALID_INPUT, ns + ": plane stride smaller than its width");
    }
    return NSS_HIP_OK;
}

int copy_inactive(const nss_hip_wnnm* h, const float* const* src, const int* src_stride, float* const* dst,
                  const int* dst_stride, hipStream_t s) {
    for (int i = 0; i < h->vi.num_planes; ++i) {
        const auto& q = h->pl[i];
        if (q.active) continue;
        const hipError_t e = hipMemcpy2DAsync(dst[i], static_cast<size_t>(dst_stride[i]) * 4, src[i],
                                              static_cast<size_t>(src_stride[i]) * 4,
                                              static_cast<size_t>(q.width) * 4, q.height, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return hip_fail(e, "plane copy");
    }
    return NSS_HIP_OK;
}

// Every slot's num / den of one contribution as float planes, slot t at
// rows 2t h (num) and (2t + 1) h (den, the corner box sum, no division).
int finish_slices(nss_hip_wnnm* h, float* const* dst, const int* dst_stride, hipStream_t s) {
    const nss_hip_wnnm_params& p = h->p;
    const int ns = 2 * p.radius + 1;
    FinishTable ft{};
    int finish_wg = 0;
    hipError_t e;
    for (int i = 0; i < h->vi.num_planes; ++i) {
        const auto& q = h->pl[i];
        if (!q.active) continue;
        const size_t npx = static_cast<size_t>(q.width) * q.height, rows = static_cast<size_t>(q.height) * dst_stride[i];
        for (int t = 0; t < ns; ++t) {
            if (ft.n == kMaxJobs) {
                if ((e = launch_finish(p.block_size, ft, finish_wg, false, s)) != hipSuccess)
                    return hip_fail(e, "slice launch");
                ft = FinishTable{};
                finish_wg = 0;
            }
            FinishJob& f = ft.job[ft.n++];
            f.num = q.num + t * npx;
            f.den = q.den + t * npx;
            f.block = p.block_size;
            f.src = nullptr;
            f.dst = dst[i] + 2 * t * rows;
            f.dst_den = dst[i] + (2 * t + 1) * rows;
            f.width = q.width;
            f.height = q.height;
            f.sstride = q.width;
            f.dstride = dst_stride[i];
            f.tiles_x = (q.width + 15) / 16;
            f.wg_begin = finish_wg;
            finish_wg += f.tiles_x * ((q.height + 15) / 16);
        }
    }
    if (ft.n && (e = launch_finish(p.block_size, ft, finish_wg, false, s)) != hipSuccess)
        return hip_fail(e, "slice launch");
    return NSS_HIP_OK;
}

// NCSR rounds before the last: the estimate of every slot inside the clip
// becomes num / den, or stays where no patch landed (in place).
int finish_estimates(nss_hip_wnnm* h, const Slots& slots, hipStream_t s) {
    const nss_hip_wnnm_params& p = h->p;
    const int ns = 2 * p.radius + 1;
    FinishTable ft{};
    int finish_wg = 0;
    hipError_t e;
    for (int i = 0; i < h->vi.num_planes; ++i) {
        const auto& q = h->pl[i];
        if (!q.active) continue;
        const size_t npx = static_cast<size_t>(q.width) * q.height;
        for (int t = slots.valid_begin(); t < slots.valid_end(); ++t) {
            if (ft.n == kMaxJobs) {
                if ((e = launch_finish(p.block_size, ft, finish_wg, false, s)) != hipSuccess)
                    return hip_fail(e, "NCSR estimate launch");
                ft = FinishTable{};
                finish_wg = 0;
            }
            float* est = q.est + static_cast<size_t>(slots.frame(t) % ns) * npx;
            FinishJob& f = ft.job[ft.n++];
            f.num = q.num + t * npx;
            f.den = q.den + t * npx;
            f.block = p.block_size;
            f.src = est;
            f.dst = est;
            f.dst_den = nullptr;
            f.width = q.width;
            f.height = q.height;
            f.sstride = q.width;
            f.dstride = q.width;
            f.tiles_x = (q.width + 15) / 16;
            f.wg_begin = finish_wg;
            finish_wg += f.tiles_x * ((q.height + 15) / 16);
        }
    }
    if (ft.n && (e = launch_finish(p.block_size, ft, finish_wg, false, s)) != hipSuccess)
        return hip_fail(e, "NCSR estimate launch");
    return NSS_HIP_OK;
}

// The contribution of centre frame `center`: the fat output of wnnmGetFrame
// (src/host/filter_wnnm.cpp:180-260, process_plane_batched :44-177 with
// emit_fat) and of ncsrGetFrame (src/host/filter_ncsr.cpp:34-190 at radius
// > 0). Per active plane the window's frames go into the instance's ring
// (frame f at f % (2r + 1); distinct frames of one window never share an
// entry); each reference is matched on the centre of the matching ring and
// predictively over the slots inside the clip (collect_temporal, as
// nss.BM3D), its group shrunk from the slot frames and its patches added into
// their slot's sums. NCSR runs its rounds on an estimate ring that starts as
// the window: round i > 0 relaxes every slot inside the clip toward its
// noisy frame, each round matches on the estimate (on rclip in round 0) and,
// but the last, finishes every such slot's estimate in place. The last
// round's sums are written as every slot's num / den; sigma-0 planes carry
// the identity contribution (src/frontend/temporal.hpp:23-36).
int temporal_contribution(nss_hip_wnnm* h, const char* name, int center, const float* const* src,
                          const int* src_stride, const float* const* rclip, const int* rclip_stride,
                          float* const* dst, const int* dst_stride, void* stream) {
    const std::string ns_name = std::string("nss_hip.") + name;
    if (!src || !src_stride || !dst || !dst_stride || (rclip && !rclip_stride))
        return fail(NSS_HIP_INVALID_INPUT, ns_name + ": null argument");
    const nss_hip_wnnm_params& p = h->p;
    if (p.radius == 0) return fail(NSS_HIP_INVALID_INPUT, ns_name + ": the temporal call serves radius > 0");
    const int r = p.radius, ns = 2 * r + 1, np = h->vi.num_planes, nframes = h->vi.num_frames;
    if (center < 0 || center >= nframes) return fail(NSS_HIP_INVALID_INPUT, ns_name + ": centre outside the clip");
    if (p.has_rclip && !rclip) return fail(NSS_HIP_INVALID_INPUT, ns_name + ": rclip planes required");
    if (rclip && !p.has_rclip)
        return fail(NSS_HIP_INVALID_INPUT, ns_name + ": this instance was made without rclip");
    for (int t = 0; t < ns; ++t)
        for (int i = 0; i < np; ++i) {
            const int k = t * np + i;
            if (!src[k] || src_stride[k] < h->pl[i].width || (rclip && (!rclip[k] || rclip_stride[k] < h->pl[i].width)))
                return fail(NSS_HIP_INVALID_INPUT, ns_name + ": missing plane or stride below width");
        }
    for (int i = 0; i < np; ++i)
        if (!dst[i] || dst_stride[i] < h->pl[i].width)
            return fail(NSS_HIP_INVALID_INPUT, ns_name + ": missing output plane or stride below width");
    std::lock_guard<std::mutex> lock(h->mu);
    DeviceGuard dg(h->device);
    if (dg.err != hipSuccess) return hip_fail(dg.err, "set device");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const Slots slots{center, nframes, r};
    const int vb = slots.valid_begin(), ve = slots.valid_end();
    hipError_t e = hipSuccess;
    bool any = false;
    for (int i = 0; i < np && e == hipSuccess; ++i) {
        const auto& q = h->pl[i];
        const size_t w = static_cast<size_t>(q.width), ph = static_cast<size_t>(q.height), npx = w * ph;
        float* fat = dst[i];
        const size_t dpitch = static_cast<size_t>(dst_stride[i]) * 4, rows = ph * dst_stride[i];
        if (!q.active) {
            // identity contribution: the source in the centre slice, weight 1
            e = hipMemset2DAsync(fat, dpitch, 0, w * 4, 2 * ns * ph, s);
            if (e == hipSuccess)
                e = hipMemcpy2DAsync(fat + 2 * r * rows, dpitch, src[r * np + i], static_cast<size_t>(src_stride[r * np + i]) * 4,
                                     w * 4, ph, hipMemcpyDeviceToDevice, s);
            for (size_t y = 0; y < ph && e == hipSuccess; ++y)
                e = hipMemsetD32Async(fat + (2 * r + 1) * rows + y * dst_stride[i], 0x3f800000u, w, s);
            continue;
        }
        any = true;
        for (int t = 0; t < ns && e == hipSuccess; ++t) {
            const size_t at = static_cast<size_t>(slots.frame(t) % ns) * npx;
            const int k = t * np + i;
            e = hipMemcpy2DAsync(q.win + at, w * 4, src[k], static_cast<size_t>(src_stride[k]) * 4, w * 4, ph,
                                 hipMemcpyDeviceToDevice, s);
            if (e == hipSuccess && rclip)
                e = hipMemcpy2DAsync(q.mwin + at, w * 4, rclip[k], static_cast<size_t>(rclip_stride[k]) * 4, w * 4,
                                     ph, hipMemcpyDeviceToDevice, s);
        }
        for (int t = vb; h->ncsr && t < ve && e == hipSuccess; ++t) {
            const size_t at = static_cast<size_t>(slots.frame(t) % ns) * npx;
            e = hipMemcpyAsync(q.est + at, q.win + at, npx * 4, hipMemcpyDeviceToDevice, s);
        }
    }
    if (e != hipSuccess) return hip_fail(e, "window copy");
    if (!any) return NSS_HIP_OK;
    const int rounds = h->ncsr ? h->iters : 1;
    for (int it = 0; it < rounds; ++it) {
        for (int i = 0; it > 0 && i < np; ++i) {
            const auto& q = h->pl[i];
            if (!q.active) continue;
            const size_t npx = static_cast<size_t>(q.width) * q.height;
            for (int t = vb; t < ve; ++t) {
                const size_t at = static_cast<size_t>(slots.frame(t) % ns) * npx;
                if ((e = launch_regularize(q.est + at, q.width, q.win + at, q.width, q.width, q.height, h->delta, s)) !=
                    hipSuccess)
                    return hip_fail(e, "NCSR regularize launch");
            }
        }
        ZeroTable zt{};
        WnnmTable wt{};
        int zero_wg = 0, wnnm_wg = 0;
        for (int i = 0; i < np; ++i) {
            const auto& q = h->pl[i];
            if (!q.active) continue;
            const size_t npx = static_cast<size_t>(q.width) * q.height;
            for (unsigned long long* buf : {q.num, q.den}) {
                ZeroJob& z = zt.job[zt.n++];
                z.p = buf;
                z.count = static_cast<long long>(ns * npx);
                z.wg_begin = zero_wg;
                zero_wg += static_cast<int>((ns * npx + 1023) / 1024);
            }
            WnnmJob& j = wt.job[wt.n++];
            j.src = nullptr;
            j.width = q.width;
            j.height = q.height;
            j.stride = q.width;
            j.nref = q.nref;
            j.group = p.group_size;
            j.block = p.block_size;
            j.sigma = p.sigma[i] / 255.f;
            j.residual = h->ncsr ? 1 : p.residual;
            j.adaptive = p.adaptive;
            j.wg_begin = wnnm_wg;
            j.matches = q.recs;
            j.counts = q.counts;
            j.num = q.num;
            j.den = q.den;
            j.temporal = 1;
            j.ring = FrameRing{h->ncsr ? q.est : q.win, static_cast<long long>(npx), ns};
            j.slots = slots;
            j.sum_slot = static_cast<long long>(npx);
            wnnm_wg += q.nref;
        }
        if ((e = launch_zero(zt, zero_wg, s)) != hipSuccess) return hip_fail(e, "group filter clear");
        // spatial top on the centre of the matching ring, then the predictive
        // layers (nss_hip_predictive_match's two launches)
        for (int i = 0; i < np; ++i) {
            const auto& q = h->pl[i];
            if (!q.active) continue;
            const float* match = it == 0 && rclip ? q.mwin : h->ncsr ? q.est : q.win;
            const FrameRing mring{match, static_cast<long long>(q.width) * q.height, ns};
            MatchTable mt{};
            MatchJob& m = mt.job[mt.n++];
            m.plane = mring.frame(slots.frame(r));
            m.stride = q.width;
            m.width = q.width;
            m.height = q.height;
            m.grid = q.grid;
            m.range = p.bm_range;
            m.group = p.group_size;
            m.t0 = r;
            m.wg_begin = 0;
            m.out = q.recs;
            m.counts = q.counts;
            size_t lds = 0;
            const int kind =
                choose_match_kind(p.block_size, p.block_step, p.bm_range, p.group_size, q.width, q.height, q.nref, &lds);
            if ((e = launch_match(p.block_size, mt, match_workgroups(kind, q.grid), lds, kind, s)) != hipSuccess)
                return hip_fail(e, "group filter match launch");
            if (p.group_size == 1) continue;  // the reference alone (src/cpu/bm/matcher.hpp:323)
            PredTable pt{};
            PredJob& pj = pt.job[pt.n++];
            pj.match = mring;
            pj.width = q.width;
            pj.height = q.height;
            pj.stride = q.width;
            pj.nref = q.nref;
            pj.group = p.group_size;
            pj.ps_num = p.ps_num;
            pj.ps_range = p.ps_range;
            pj.slots = slots;
            pj.wg_begin = 0;
            pj.recs = q.recs;
            pj.counts = q.counts;
            const size_t plds = predictive_table_lds(p.block_size, pt, r);
            if ((e = launch_predictive(p.block_size, pt, q.nref, plds, s)) != hipSuccess)
                return hip_fail(e, "group filter predictive launch");
        }
        if ((e = launch_wnnm(wt, wnnm_wg, h->ncsr, s)) != hipSuccess) return hip_fail(e, "group filter launch");
        const int rc = it + 1 == rounds ? finish_slices(h, dst, dst_stride, s) : finish_estimates(h, slots, s);
        if (rc != NSS_HIP_OK) return rc;
    }
    return NSS_HIP_OK;
}

}  // namespace

extern "C" {

int nss_hip_wnnm_create(const nss_hip_wnnm_params* params, const nss_hip_video_info* info, int device,
                        nss_hip_wnnm** out) {
    if (!params || !info || !out) return fail(NSS_HIP_INVALID_INPUT, "nss_hip.WNNM: null argument");
    *out = nullptr;
    auto* h = new (std::nothrow) nss_hip_wnnm;
    if (!h) return fail(NSS_HIP_RESOURCE_FAILURE, "nss_hip.WNNM: out of host memory");
    const int rc = init_instance(h, *params, info, device, "WNNM");
    if (rc != NSS_HIP_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return NSS_HIP_OK;
}

int nss_hip_wnnm_process(nss_hip_wnnm* h, const float* const* src, const int* src_stride,
                         const float* const* rclip, const int* rclip_stride, float* const* dst,
                         const int* dst_stride, void* stream) {
    if (!h) return fail(NSS_HIP_INVALID_INPUT, "nss_hip.WNNM: null argument");
    int rc = check_planes(h, src, src_stride, rclip, rclip_stride, dst, dst_stride, "WNNM");
    if (rc != NSS_HIP_OK) return rc;
    DeviceGuard dg(h->device);
    if (dg.err != hipSuccess) return hip_fail(dg.err, "set device");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (h->p.radius > 0)
        return fail(NSS_HIP_INVALID_INPUT, "nss_hip.WNNM: radius > 0 filters through nss_hip_wnnm_process_temporal");
    if ((rc = copy_inactive(h, src, src_stride, dst, dst_stride, s)) != NSS_HIP_OK) return rc;
    return group_round(h, rclip ? rclip : src, rclip ? rclip_stride : src_stride, src, src_stride, src, src_stride,
                       dst, dst_stride, s);
}

// src/host/filter_wnnm.cpp:180-260 at radius > 0: the contribution of one
// centre frame (temporal_contribution).
int nss_hip_wnnm_process_temporal(nss_hip_wnnm* h, int center, const float* const* src, const int* src_stride,
                                  const float* const* rclip, const int* rclip_stride, float* const* dst,
                                  const int* dst_stride, void* stream) {
    if (!h) return fail(NSS_HIP_INVALID_INPUT, "nss_hip.WNNM: null argument");
    return temporal_contribution(h, "WNNM", center, src, src_stride, rclip, rclip_stride, dst, dst_stride, stream);
}

int nss_hip_wnnm_destroy(nss_hip_wnnm* h) {
    if (!h) return NSS_HIP_OK;
    {
        DeviceGuard dg(h->device);
        (void)hipDeviceSynchronize();
        release(h);
    }
    delete h;
    return NSS_HIP_OK;
}

int nss_hip_ncsr_create(const nss_hip_ncsr_params* params, const nss_hip_video_info* info, int device,
                        nss_hip_ncsr** out) {
    if (!params || !info || !out) return fail(NSS_HIP_INVALID_INPUT, "nss_hip.NCSR: null argument");
    *out = nullptr;
    if (params->iters < 1 || params->iters > 64 || !(params->delta >= 0.f && params->delta <= 1.f))
        return fail(NSS_HIP_INVALID_INPUT, "nss_hip.NCSR: parameters not resolved by the parse call");
    auto* h = new (std::nothrow) nss_hip_ncsr;
    if (!h) return fail(NSS_HIP_RESOURCE_FAILURE, "nss_hip.NCSR: out of host memory");
    nss_hip_wnnm_params p{};
    for (int i = 0; i < 3; ++i) p.sigma[i] = params->sigma[i];
    p.block_size = params->block_size;
    p.block_step = params->block_step;
    p.group_size = params->group_size;
    p.bm_range = params->bm_range;
    p.radius = params->radius;
    p.temporal_mode = params->temporal_mode;
    p.ps_num = params->ps_num;
    p.ps_range = params->ps_range;
    p.residual = 1;
    p.adaptive = 0;
    p.has_rclip = params->has_rclip;
    p.memory_limit_mb = params->memory_limit_mb;
    p.device_id = params->device_id;
    h->ncsr = true;
    h->iters = params->iters;
    h->delta = params->delta;
    const int rc = init_instance(h, p, info, device, "NCSR");
    if (rc != NSS_HIP_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return NSS_HIP_OK;
}

// src/host/filter_ncsr.cpp:34-190 at radius 0: the estimate starts as the
// noisy plane; round i > 0 relaxes it toward the noisy plane by delta; each
// round matches on the estimate (on rclip in round 0), filters the
// estimate's groups and finishes into the estimate (the last round into dst),
// with the estimate where no patch landed.
int nss_hip_ncsr_process(nss_hip_ncsr* h, const float* const* src, const int* src_stride,
                         const float* const* rclip, const int* rclip_stride, float* const* dst,
                         const int* dst_stride, void* stream) {
    if (!h) return fail(NSS_HIP_INVALID_INPUT, "nss_hip.NCSR: null argument");
    int rc = check_planes(h, src, src_stride, rclip, rclip_stride, dst, dst_stride, "NCSR");
    if (rc != NSS_HIP_OK) return rc;
    DeviceGuard dg(h->device);
    if (dg.err != hipSuccess) return hip_fail(dg.err, "set device");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (h->p.radius > 0)
        return fail(NSS_HIP_INVALID_INPUT, "nss_hip.NCSR: radius > 0 filters through nss_hip_ncsr_process_temporal");
    if ((rc = copy_inactive(h, src, src_stride, dst, dst_stride, s)) != NSS_HIP_OK) return rc;
    const float* est_c[3] = {};
    float* est_m[3] = {};
    int est_stride[3] = {};
    hipError_t e;
    for (int i = 0; i < h->vi.num_planes; ++i) {
        const auto& q = h->pl[i];
        if (!q.active) continue;
        est_c[i] = est_m[i] = q.est;
        est_stride[i] = q.width;
        e = hipMemcpy2DAsync(q.est, static_cast<size_t>(q.width) * 4, src[i], static_cast<size_t>(src_stride[i]) * 4,
                             static_cast<size_t>(q.width) * 4, q.height, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return hip_fail(e, "NCSR estimate copy");
    }
    for (int it = 0; it < h->iters; ++it) {
        if (it > 0) {
            for (int i = 0; i < h->vi.num_planes; ++i) {
                const auto& q = h->pl[i];
                if (!q.active) continue;
                if ((e = launch_regularize(q.est, q.width, src[i], src_stride[i], q.width, q.height, h->delta, s)) !=
                    hipSuccess)
                    return hip_fail(e, "NCSR regularize launch");
            }
        }
        const bool guide = it == 0 && rclip;
        const bool last = it + 1 == h->iters;
        rc = group_round(h, guide ? rclip : est_c, guide ? rclip_stride : est_stride, est_c, est_stride, est_c,
                         est_stride, last ? dst : est_m, last ? dst_stride : est_stride, s);
        if (rc != NSS_HIP_OK) return rc;
    }
    return NSS_HIP_OK;
}

// src/host/filter_ncsr.cpp:34-190 at radius > 0: the contribution of one
// centre frame (temporal_contribution).
int nss_hip_ncsr_process_temporal(nss_hip_ncsr* h, int center, const float* const* src, const int* src_stride,
                                  const float* const* rclip, const int* rclip_stride, float* const* dst,
                                  const int* dst_stride, void* stream) {
    if (!h) return fail(NSS_HIP_INVALID_INPUT, "nss_hip.NCSR: null argument");
    return temporal_contribution(h, "NCSR", center, src, src_stride, rclip, rclip_stride, dst, dst_stride, stream);
}

int nss_hip_ncsr_destroy(nss_hip_ncsr* h) {
    if (!h) return NSS_HIP_OK;
    {
        DeviceGuard dg(h->device);
        (void)hipDeviceSynchronize();
        release(h);
    }
    delete h;
    return NSS_HIP_OK;
}

}  // extern "C"
