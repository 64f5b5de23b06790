// No-allocation form of the radius-0 frame filter, for callers that cannot
// allocate device memory: code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc. The frame's working buffers (match
// records and counts, tile partials or frame-sized sums, the basic estimate
// of a two-stage filter) live in one file-scope device array of fixed size
// instead of buffers allocated at creation. Every call on a GPU shares that
// GPU's array: calls that overlap on one GPU (on different streams, or from
// several threads without ordering) race on it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "instance.hpp"

// 192 MiB per device: a 1920 x 1080 YUV420 two-stage frame at the default
// arguments needs about 103 MiB on the tile plan (the int64 numerator
// partials dominate); larger frames take the plan batched over grid rows, whose
// frame-sized sums need 20 bytes per pixel of a two-stage frame.
constexpr size_t kScratchBytes = static_cast<size_t>(192) << 20;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;

std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

// Set-up, once per device with that device current: the device address of
// this device's copy of the array (the array's name is a host-side stand-in)
// and the DCT tables. Later calls reuse the address.
int scratch_on(int device, char** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return nsship::hip_fail(e, "scratch symbol address");
        const int rc = nsship::ensure_tables(device);
        if (rc != NSS_HIP_OK) return rc;
        g_scratch_ptr[device] = static_cast<char*>(p);
    }
    *out = g_scratch_ptr[device];
    return NSS_HIP_OK;
}

}  // namespace

extern "C" int nss_hip_bm3d_process_noalloc(const nss_hip_bm3d_params* params, const nss_hip_video_info* info,
                                            int device, const float* const* src, const int* src_stride,
                                            const float* const* ref, const int* ref_stride, float* const* dst,
                                            const int* dst_stride, void* stream) {
    using namespace nsship;
    if (!params || !info || !src || !src_stride || !dst || !dst_stride)
        return fail(NSS_HIP_INVALID_INPUT, "nss_hip: null argument");
    if (device < 0 || device >= kMaxDevices) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: device index out of range");
    if (params->radius > 0) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: radius > 0 filters through push/pop");
    nss_hip_bm3d h;  // the host plan only; no device buffers
    int rc = plan_instance(&h, params, info, device);
    if (rc != NSS_HIP_OK) return rc;
    for (int i = 0; i < h.nplanes; ++i)
        if (!src[i] || !dst[i] || src_stride[i] < h.pw[i] || dst_stride[i] < h.pw[i])
            return fail(NSS_HIP_INVALID_INPUT, "nss_hip: missing plane or stride below width");
    if (h.p.has_ref && !h.p.final_stage) {
        if (!ref || !ref_stride) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: ref planes required");
        for (int i = 0; i < h.nplanes; ++i)
            if (!ref[i] || ref_stride[i] < h.pw[i])
                return fail(NSS_HIP_INVALID_INPUT, "nss_hip: missing ref plane or stride below width");
    }
    // The frame's buffers in the array: the normal plan, or the plan batched
    // over grid rows (the memory_limit_mb plan, same output bits) when the
    // array, or a smaller memory_limit_mb, is too small for it.
    // (64 KiB held back for the buffers' 256-byte alignment)
    FrameBuffers fb;
    size_t cap = kScratchBytes;
    if (h.p.memory_limit_mb > 0) cap = std::min(cap, static_cast<size_t>(h.p.memory_limit_mb) << 20);
    if (!fit_frame_buffers(&h, cap - (static_cast<size_t>(64) << 10)) ||
        carve_frame_buffers(&h, nullptr, fb) > kScratchBytes)
        return fail(NSS_HIP_UNSUPPORTED, "nss_hip: this frame needs " +
                                             std::to_string((frame_buffer_bytes(&h) >> 20) + 1) +
                                             " MiB of working memory even in batches; the no-allocation entry point "
                                             "holds " + std::to_string(cap >> 20) + " MiB");
    DeviceGuard guard(device);
    if (guard.err != hipSuccess) return hip_fail(guard.err, "set device");
    char* base = nullptr;
    rc = scratch_on(device, &base);
    if (rc != NSS_HIP_OK) return rc;
    carve_frame_buffers(&h, base, fb);
    FrameRef f = frame_ref(&h, src, src_stride, ref, ref_stride, dst, dst_stride);
    f.buf = fb;
    return run_batch(&f, 1, nullptr, static_cast<hipStream_t>(stream));
}
