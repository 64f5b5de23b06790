// nss_hip.WNNM behind the C ABI: one spatial frame per call, every step on
// the caller's stream.
//
// Per active plane (src/host/filter_wnnm.cpp:44-160 process_plane_batched at
// radius 0): clear the plane's fixed-point sums, match every raster reference
// on rclip (or src) with the BM3D matchers, shrink each group and add it into
// the sums (wnnm_group_kernel), then finish dst = num / den, or src where no
// patch landed (the atomic path's corner finish). A plane with sigma 0 is
// copied (src/host/filter_wnnm.cpp:236-246).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <string>

#include "instance.hpp"

using namespace nsship;

struct nss_hip_wnnm {
    nss_hip_wnnm_params p{};
    nss_hip_video_info vi{};
    int device = 0;
    struct Plane {
        int width = 0, height = 0;
        Grid grid{};
        int nref = 0;
        bool active = false;
        MatchRec* recs = nullptr;
        int* counts = nullptr;
        unsigned long long* num = nullptr;
        unsigned long long* den = nullptr;
    } pl[3];
    long long bytes = 0;
};

namespace {

void release(nss_hip_wnnm* h) {
    for (auto& q : h->pl) {
        if (q.recs) (void)hipFree(q.recs);
        if (q.counts) (void)hipFree(q.counts);
        if (q.num) (void)hipFree(q.num);
        if (q.den) (void)hipFree(q.den);
        q.recs = nullptr;
        q.counts = nullptr;
        q.num = nullptr;
        q.den = nullptr;
    }
}

}  // namespace

extern "C" {

int nss_hip_wnnm_create(const nss_hip_wnnm_params* params, const nss_hip_video_info* info, int device,
                        nss_hip_wnnm** out) {
    if (!params || !info || !out) return fail(NSS_HIP_INVALID_INPUT, "nss_hip.WNNM: null argument");
    *out = nullptr;
    const nss_hip_wnnm_params& p = *params;
    if (info->num_planes < 1 || info->num_planes > 3 || info->width < 1 || info->height < 1)
        return fail(NSS_HIP_INVALID_INPUT, "nss_hip.WNNM: invalid video info");
    if (p.radius != 0 || p.block_size < 1 || p.block_size > 16 || p.group_size < 1 || p.group_size > 32 ||
        p.block_step < 1 || p.block_step > p.block_size || p.bm_range < 1 || p.bm_range > 64)
        return fail(NSS_HIP_INVALID_INPUT, "nss_hip.WNNM: parameters not resolved by nss_hip_wnnm_parse");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess) return hip_fail(e, "device count");
    if (device < 0 || device >= ndev) return fail(NSS_HIP_INVALID_INPUT, "nss_hip.WNNM: device_id out of range");
    DeviceGuard dg(device);
    if (dg.err != hipSuccess) return hip_fail(dg.err, "set device");
    auto* h = new (std::nothrow) nss_hip_wnnm;
    if (!h) return fail(NSS_HIP_RESOURCE_FAILURE, "nss_hip.WNNM: out of host memory");
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
            delete h;
            return fail(NSS_HIP_INVALID_INPUT, "nss: selected plane is smaller than block_size");
        }
        q.grid = make_grid(q.width, q.height, p.block_size, p.block_step);
        q.nref = q.grid.nx * q.grid.ny;
        const size_t pixels = static_cast<size_t>(q.width) * q.height;
        const size_t rec_bytes = static_cast<size_t>(q.nref) * p.group_size * sizeof(MatchRec);
        const size_t cnt_bytes = static_cast<size_t>(q.nref) * sizeof(int);
        const size_t sum_bytes = pixels * sizeof(unsigned long long);
        h->bytes += static_cast<long long>(rec_bytes + cnt_bytes + 2 * sum_bytes);
        if ((e = hipMalloc(&q.recs, rec_bytes)) != hipSuccess || (e = hipMalloc(&q.counts, cnt_bytes)) != hipSuccess ||
            (e = hipMalloc(&q.num, sum_bytes)) != hipSuccess || (e = hipMalloc(&q.den, sum_bytes)) != hipSuccess) {
            release(h);
            delete h;
            return hip_fail(e, "WNNM buffers");
        }
    }
    if (p.memory_limit_mb > 0 && h->bytes > static_cast<long long>(p.memory_limit_mb) << 20) {
        const long long need = (h->bytes + (1ll << 20) - 1) >> 20;
        release(h);
        delete h;
        return fail(NSS_HIP_INVALID_INPUT, "nss_hip.WNNM: memory_limit_mb is too small for this clip: the instance "
                                           "needs at least " + std::to_string(need) + " MiB");
    }
    *out = h;
    return NSS_HIP_OK;
}

int nss_hip_wnnm_process(nss_hip_wnnm* h, const float* const* src, const int* src_stride,
                         const float* const* rclip, const int* rclip_stride, float* const* dst,
                         const int* dst_stride, void* stream) {
    if (!h || !src || !src_stride || !dst || !dst_stride || (rclip && !rclip_stride))
        return fail(NSS_HIP_INVALID_INPUT, "nss_hip.WNNM: null argument");
    const int np = h->vi.num_planes;
    for (int i = 0; i < np; ++i) {
        const auto& q = h->pl[i];
        if (!src[i] || !dst[i] || (rclip && !rclip[i]))
            return fail(NSS_HIP_INVALID_INPUT, "nss_hip.WNNM: null plane pointer");
        if (src_stride[i] < q.width || dst_stride[i] < q.width || (rclip && rclip_stride[i] < q.width))
            return fail(NSS_HIP_INVALID_INPUT, "nss_hip.WNNM: plane stride smaller than its width");
    }
    DeviceGuard dg(h->device);
    if (dg.err != hipSuccess) return hip_fail(dg.err, "set device");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const nss_hip_wnnm_params& p = h->p;
    ZeroTable zt{};
    WnnmTable wt{};
    FinishTable ft{};
    int zero_wg = 0, wnnm_wg = 0, finish_wg = 0;
    hipError_t e;
    for (int i = 0; i < np; ++i) {
        const auto& q = h->pl[i];
        if (!q.active) {
            e = hipMemcpy2DAsync(dst[i], static_cast<size_t>(dst_stride[i]) * 4, src[i],
                                 static_cast<size_t>(src_stride[i]) * 4, static_cast<size_t>(q.width) * 4, q.height,
                                 hipMemcpyDeviceToDevice, s);
            if (e != hipSuccess) return hip_fail(e, "WNNM plane copy");
            continue;
        }
        const long long pixels = static_cast<long long>(q.width) * q.height;
        for (unsigned long long* buf : {q.num, q.den}) {
            ZeroJob& z = zt.job[zt.n++];
            z.p = buf;
            z.count = pixels;
            z.wg_begin = zero_wg;
            zero_wg += static_cast<int>((pixels + 1023) / 1024);
        }
        WnnmJob& w = wt.job[wt.n++];
        w.src = src[i];
        w.width = q.width;
        w.height = q.height;
        w.stride = src_stride[i];
        w.nref = q.nref;
        w.group = p.group_size;
        w.block = p.block_size;
        w.sigma = p.sigma[i] / 255.f;
        w.residual = p.residual;
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
        f.src = src[i];
        f.dst = dst[i];
        f.dst_den = nullptr;
        f.width = q.width;
        f.height = q.height;
        f.sstride = src_stride[i];
        f.dstride = dst_stride[i];
        f.tiles_x = (q.width + 15) / 16;
        f.wg_begin = finish_wg;
        finish_wg += f.tiles_x * ((q.height + 15) / 16);
    }
    if (wt.n == 0) return NSS_HIP_OK;
    if ((e = launch_zero(zt, zero_wg, s)) != hipSuccess) return hip_fail(e, "WNNM clear");
    for (int i = 0; i < np; ++i) {
        const auto& q = h->pl[i];
        if (!q.active) continue;
        MatchTable mt{};
        MatchJob& m = mt.job[mt.n++];
        m.plane = rclip ? rclip[i] : src[i];
        m.stride = rclip ? rclip_stride[i] : src_stride[i];
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
            return hip_fail(e, "WNNM match launch");
    }
    if ((e = launch_wnnm(wt, wnnm_wg, s)) != hipSuccess) return hip_fail(e, "WNNM group launch");
    if ((e = launch_finish(p.block_size, ft, finish_wg, false, s)) != hipSuccess)
        return hip_fail(e, "WNNM finish launch");
    return NSS_HIP_OK;
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

}  // extern "C"
