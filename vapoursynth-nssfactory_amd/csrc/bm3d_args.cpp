// Creation-time argument resolution of <ns>.BM3D behind the C ABI.
//
// Order of checks (the first error is the one reported, and its text is part
// of the interface):
//   1. every int argument fits int32, every float is finite and fits float
//      (src/frontend/validate.cpp:13-46 validate_numeric_args; map keys are
//      visited in sorted order, as VapourSynth keeps them);
//   2. memory_limit_mb in [1, 1048576] when given (src/frontend/validate.hpp:37-45);
//   3. temporal_mode (src/frontend/temporal_args.cpp:11-28), unset = rolling
//      as the device backends default (src/cuda/runtime/context.hpp:26);
//   4. parse_bm3d (src/frontend/bm3d_args.cpp:140-170 with its helpers :42-134);
//   5. rolling arguments with the adaptive cache (src/frontend/temporal_args.cpp:30-58);
//   6. backend tail device_id / num_streams (src/cuda/runtime/context.cpp:23-41).
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nss_hip.h"

namespace nsship {
extern thread_local std::string g_last_error;
}

namespace {

constexpr int kBmBlock = 8, kBmGroup = 8, kBmDefaultStep = 8, kBm3dDefaultRange = 9, kBmMaxRange = 64;
constexpr int kBmDefaultPsNum = 2, kBmDefaultPsRange = 4, kBmMaxRadius = 16;
constexpr float kBmDefaultSigma = 3.0f;
constexpr int kMaxStreams = 16;

struct Invalid : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};

[[noreturn]] void fail(const std::string& ns, const char* message) {
    throw Invalid(ns + ".BM3D: " + message);
}

bool finite_bits(float v) {
    std::uint32_t u;
    std::memcpy(&u, &v, 4);
    return (u & 0x7f800000u) != 0x7f800000u;
}

// The argument map (a VSMap's view): lookups by key; a missing key has -1
// elements, as mapNumElements reports.
class ArgMap {
public:
    ArgMap(const nss_hip_arg* args, int n) {
        for (int i = 0; i < n; ++i)
            if (args[i].key && args[i].count > 0) by_key_[args[i].key] = &args[i];
    }
    const std::map<std::string, const nss_hip_arg*>& entries() const { return by_key_; }
    int num(const char* key) const {
        auto it = by_key_.find(key);
        return it == by_key_.end() ? -1 : it->second->count;
    }
    std::int64_t get_int(const char* key, int idx, int* err) const {
        const nss_hip_arg* a = find(key, idx, err);
        if (!a) return 0;
        if (a->type != 'i' || !a->ints) {
            *err = 1;
            return 0;
        }
        return a->ints[idx];
    }
    double get_float(const char* key, int idx, int* err) const {
        const nss_hip_arg* a = find(key, idx, err);
        if (!a) return 0;
        if (a->type != 'f' || !a->floats) {
            *err = 1;
            return 0;
        }
        return a->floats[idx];
    }
    const char* get_data(const char* key, int* err) const {
        const nss_hip_arg* a = find(key, 0, err);
        if (!a) return nullptr;
        if (a->type != 's' || !a->data) {
            *err = 1;
            return nullptr;
        }
        return a->data;
    }
    const nss_hip_video_info* get_clip(const char* key) const {
        auto it = by_key_.find(key);
        if (it == by_key_.end() || it->second->type != 'c') return nullptr;
        return it->second->clip_info;
    }

private:
    const nss_hip_arg* find(const char* key, int idx, int* err) const {
        *err = 0;
        auto it = by_key_.find(key);
        if (it == by_key_.end() || idx < 0 || idx >= it->second->count) {
            *err = 1;
            return nullptr;
        }
        return it->second;
    }
    std::map<std::string, const nss_hip_arg*> by_key_;
};

// src/frontend/validate.cpp:80-146 map_* helpers.
int map_int(const ArgMap& in, const char* key, int def, int* err_out = nullptr) {
    int err = 0;
    const std::int64_t v = in.get_int(key, 0, &err);
    if (err_out) *err_out = err;
    return err ? def : static_cast<int>(v);
}

void map_float_array(const ArgMap& in, const char* key, float* out, int n, float def) {
    const int m = in.num(key);
    for (int i = 0; i < n; ++i) {
        if (m <= 0) {
            out[i] = def;
            continue;
        }
        int err = 0;
        out[i] = static_cast<float>(in.get_float(key, i < m ? i : m - 1, &err));
        if (err) out[i] = def;
    }
}

void map_int_array(const ArgMap& in, const char* key, int* out, int n, int def) {
    const int m = in.num(key);
    for (int i = 0; i < n; ++i) {
        if (m <= 0) {
            out[i] = def;
            continue;
        }
        int err = 0;
        out[i] = static_cast<int>(in.get_int(key, i < m ? i : m - 1, &err));
        if (err) out[i] = def;
    }
}

void map_inherit_int(const ArgMap& in, const char* key, int* out, int n, int def) {
    const int m = in.num(key);
    for (int i = 0; i < n; ++i) {
        if (m <= 0) {
            out[i] = def;
            continue;
        }
        if (i < m) {
            int err = 0;
            out[i] = static_cast<int>(in.get_int(key, i, &err));
            if (err) out[i] = i == 0 ? def : out[i - 1];
        } else {
            out[i] = out[i - 1];
        }
    }
}

bool allowed_block(int v) { return v == 1 || v == 2 || v == 4 || v == 8 || v == 12 || v == 16 || v == 32; }
bool allowed_group(int v) { return v == 1 || v == 2 || v == 4 || v == 8 || v == 16 || v == 32 || v == 64; }

int plane_width(const nss_hip_video_info& vi, int plane) { return vi.width >> (plane ? vi.sub_sampling_w : 0); }
int plane_height(const nss_hip_video_info& vi, int plane) { return vi.height >> (plane ? vi.sub_sampling_h : 0); }

// src/frontend/validate.cpp:48-57 is_const_32f.
bool is_const_32f(const nss_hip_video_info& vi) {
    if (vi.width < 1 || vi.height < 1 || vi.num_frames < 1 ||
        static_cast<std::uint64_t>(vi.width) * vi.height > static_cast<std::uint64_t>(std::numeric_limits<int>::max()))
        return false;
    if (vi.color_family < 1 || vi.color_family > 3 || vi.num_planes < 1) return false;
    return vi.sample_type == 1 && vi.bits_per_sample == 32;
}

bool same_video(const nss_hip_video_info& a, const nss_hip_video_info& b) {
    return a.width == b.width && a.height == b.height && a.color_family == b.color_family &&
           a.sample_type == b.sample_type && a.bits_per_sample == b.bits_per_sample &&
           a.sub_sampling_w == b.sub_sampling_w && a.sub_sampling_h == b.sub_sampling_h &&
           a.num_planes == b.num_planes && a.num_frames == b.num_frames;
}

// include/nss/contracts.hpp:27-29 NoiseProfile::bm3d_effective.
float bm3d_effective(float user) { return (user * (1.f / 255.f)) * 0.75f; }

void validate_numeric(const ArgMap& in, const std::string& ns) {
    for (const auto& kv : in.entries()) {
        const nss_hip_arg* a = kv.second;
        if (a->type != 'i' && a->type != 'f') continue;
        for (int i = 0; i < a->count; ++i) {
            bool valid;
            if (a->type == 'i') {
                const std::int64_t v = a->ints ? a->ints[i] : 0;
                valid = a->ints && v >= std::numeric_limits<int>::min() && v <= std::numeric_limits<int>::max();
            } else {
                const double v = a->floats ? a->floats[i] : 0.0;
                std::uint64_t bits;
                std::memcpy(&bits, &v, 8);
                const bool finite = (bits & UINT64_C(0x7ff0000000000000)) != UINT64_C(0x7ff0000000000000);
                const double limit = std::numeric_limits<float>::max();
                valid = a->floats && finite && v >= -limit && v <= limit;
            }
            if (!valid) {
                char message[256];
                std::snprintf(message, sizeof(message), "%s: parameter %s[%d] is not representable as a finite %s",
                              ns.c_str(), kv.first.c_str(), i, a->type == 'i' ? "int" : "float");
                throw Invalid(message);
            }
        }
    }
}

// src/frontend/temporal_args.cpp:11-28; returns true for rolling.
bool parse_temporal_mode(const ArgMap& in, int radius, const std::string& ns) {
    int err = 0;
    const char* mode = in.get_data("temporal_mode", &err);
    std::string s;
    if (!err && mode) {
        s = mode;
        for (char& c : s)
            if (c >= 'A' && c <= 'Z') c = static_cast<char>(c - 'A' + 'a');
    }
    if (s == "fused") fail(ns, "temporal_mode=fused is not supported; use rolling or legacy");
    if (!s.empty() && s != "rolling" && s != "legacy") fail(ns, "temporal_mode must be rolling or legacy");
    if (radius <= 0) return false;
    if (s.empty()) return true;  // the device default
    return s == "rolling";
}

// src/frontend/temporal_args.cpp:30-58 with the adaptive cache.
void parse_rolling(const ArgMap& in, int radius, const std::string& ns, nss_hip_bm3d_params& p) {
    p.rolling_chunk = map_int(in, "rolling_chunk", 1);
    int limit_err = 0, chunks_err = 0;
    p.rolling_cache_limit = map_int(in, "rolling_cache_limit", 16, &limit_err);
    p.rolling_cache_chunks = map_int(in, "rolling_cache_chunks", 1, &chunks_err);
    p.rolling_cache_adaptive = 1;
    if (radius < 1) fail(ns, "rolling mode requires radius > 0");
    if (p.rolling_chunk < 1 || p.rolling_chunk > 64) fail(ns, "rolling_chunk must be in [1, 64]");
    if (limit_err && !chunks_err) p.rolling_cache_limit = p.rolling_cache_chunks;
    if (p.rolling_cache_limit < 1 || p.rolling_cache_limit > 64 || p.rolling_cache_chunks < 1 ||
        p.rolling_cache_chunks > 64)
        fail(ns, "rolling_cache_limit must be in [1, 64]");
    if (p.rolling_cache_limit < p.rolling_cache_chunks)
        fail(ns, "rolling_cache_limit must be at least rolling_cache_chunks");
}

// src/frontend/bm3d_args.cpp:42-49 parse_sigma.
void parse_sigma(const ArgMap& in, int np, const std::string& ns, nss_hip_bm3d_params& p, float given[3]) {
    map_float_array(in, "sigma", p.sigma, np, kBmDefaultSigma);
    std::copy_n(p.sigma, 3, given);
    for (int i = 0; i < np; ++i) {
        if (!finite_bits(p.sigma[i]) || p.sigma[i] < 0.f) fail(ns, "sigma must be finite and non-negative");
        if (p.sigma[i] != 0.f) p.sigma[i] = bm3d_effective(p.sigma[i]);
    }
}

// src/frontend/bm3d_args.cpp:53-60 map_adaptive.
void map_adaptive(const ArgMap& in, const char* key, int* out, int np, int fallback, const int* limit) {
    if (in.num(key) <= 0) {
        for (int i = 0; i < np; ++i) out[i] = std::min(fallback, limit[i]);
    } else {
        map_inherit_int(in, key, out, np, fallback);
    }
}

// src/frontend/bm3d_args.cpp:63-79 parse_chroma.
void parse_chroma(const ArgMap& in, const nss_hip_video_info& vi, const std::string& ns, nss_hip_bm3d_params& p) {
    const int chroma = map_int(in, "chroma", 0);
    if (chroma != 0 && chroma != 1) fail(ns, "chroma must be 0 or 1");
    p.chroma = chroma != 0;
    if (!p.chroma) return;
    if (vi.color_family != NSS_HIP_CF_YUV || vi.sub_sampling_w != 0 || vi.sub_sampling_h != 0)
        fail(ns, "chroma requires a YUV 4:4:4 clip");
    for (int i = 1; i < vi.num_planes; ++i) {
        p.block_size[i] = p.block_size[0];
        p.group_size[i] = p.group_size[0];
        p.block_step[i] = p.block_step[0];
        p.bm_range[i] = p.bm_range[0];
        p.ps_num[i] = p.ps_num[0];
        p.ps_range[i] = p.ps_range[0];
    }
}

// src/frontend/bm3d_args.cpp:83-119 parse_basic_stage.
void parse_basic_stage(const ArgMap& in, const nss_hip_video_info& vi, const std::string& ns, bool sigma_basic,
                       bool block_basic, bool group_basic, float sigma_given[3], nss_hip_bm3d_params& p) {
    const int np = vi.num_planes;
    if (sigma_basic) map_float_array(in, "sigma_basic", sigma_given, np, kBmDefaultSigma);
    const bool step_given = in.num("block_step") > 0;
    const bool ps_given = in.num("ps_num") > 0;
    for (int i = 0; i < np; ++i) {
        p.block_size_basic[i] = p.block_size[i];
        p.group_size_basic[i] = p.group_size[i];
    }
    if (block_basic) map_inherit_int(in, "block_size_basic", p.block_size_basic, np, kBmBlock);
    if (group_basic) map_inherit_int(in, "group_size_basic", p.group_size_basic, np, kBmGroup);
    for (int i = 0; i < np; ++i) {
        const int plane = p.chroma ? 0 : i;
        p.block_size_basic[i] = p.block_size_basic[plane];
        p.group_size_basic[i] = p.group_size_basic[plane];
        if (!finite_bits(sigma_given[i]) || sigma_given[i] < 0.f) fail(ns, "sigma_basic must be finite and non-negative");
        p.sigma_basic[i] = sigma_given[i] != 0.f ? bm3d_effective(sigma_given[i]) : 0.f;
        if (!allowed_block(p.block_size_basic[i])) fail(ns, "block_size_basic must be one of 1, 2, 4, 8, 12, 16, 32");
        if (!allowed_group(p.group_size_basic[i])) fail(ns, "group_size_basic must be one of 1, 2, 4, 8, 16, 32, 64");
        if (p.block_size_basic[i] > plane_width(vi, i) || p.block_size_basic[i] > plane_height(vi, i))
            fail(ns, "block_size_basic must not exceed plane dimensions");
        p.block_step_basic[i] = step_given ? p.block_step[i] : std::min(kBmDefaultStep, p.block_size_basic[i]);
        p.ps_num_basic[i] = ps_given ? p.ps_num[i] : std::min(kBmDefaultPsNum, p.group_size_basic[i]);
        if (p.block_step_basic[i] > p.block_size_basic[i]) fail(ns, "block_step must be in [1, block_size_basic]");
        if (p.ps_num_basic[i] > p.group_size_basic[i]) fail(ns, "ps_num must be in [1, group_size_basic]");
    }
}

// src/frontend/bm3d_args.cpp:122-134 check_planes.
void check_planes(const nss_hip_video_info& vi, const std::string& ns, const nss_hip_bm3d_params& p) {
    for (int i = 0; i < vi.num_planes; ++i) {
        if (!allowed_block(p.block_size[i])) fail(ns, "block_size must be one of 1, 2, 4, 8, 12, 16, 32");
        if (!allowed_group(p.group_size[i])) fail(ns, "group_size must be one of 1, 2, 4, 8, 16, 32, 64");
        if (p.block_size[i] > plane_width(vi, i) || p.block_size[i] > plane_height(vi, i))
            fail(ns, "block_size must not exceed plane dimensions");
        if (p.block_step[i] < 1 || p.block_step[i] > p.block_size[i]) fail(ns, "block_step must be in [1, block_size]");
        if (p.ps_num[i] < 1 || p.ps_num[i] > p.group_size[i]) fail(ns, "ps_num must be in [1, group_size]");
        if (p.bm_range[i] < 1 || p.bm_range[i] > kBmMaxRange) fail(ns, "bm_range must be in [1, 64]");
        if (p.ps_range[i] < 0 || p.ps_range[i] > kBmMaxRange) fail(ns, "ps_range must be in [0, 64]");
    }
}

// src/frontend/bm3d_args.cpp:140-170 parse_bm3d.
void parse_bm3d(const ArgMap& in, const nss_hip_video_info& vi, const nss_hip_video_info* ref_vi,
                const std::string& ns, nss_hip_bm3d_params& p) {
    if (!is_const_32f(vi)) fail(ns, "constant Gray/YUV/RGB 32-bit float required");
    if (ref_vi && !same_video(vi, *ref_vi)) fail(ns, "ref must match clip");
    const int np = vi.num_planes;
    for (int i = 0; i < 3; ++i) {
        p.sigma[i] = kBmDefaultSigma;
        p.block_size[i] = kBmBlock;
        p.group_size[i] = kBmGroup;
        p.block_step[i] = kBmDefaultStep;
        p.bm_range[i] = kBm3dDefaultRange;
        p.ps_num[i] = kBmDefaultPsNum;
        p.ps_range[i] = kBmDefaultPsRange;
        p.sigma_basic[i] = 0.f;
        p.block_size_basic[i] = p.group_size_basic[i] = p.block_step_basic[i] = p.ps_num_basic[i] = 0;
    }
    float sigma_given[3];
    parse_sigma(in, np, ns, p, sigma_given);
    map_inherit_int(in, "block_size", p.block_size, np, kBmBlock);
    map_inherit_int(in, "group_size", p.group_size, np, kBmGroup);
    map_adaptive(in, "block_step", p.block_step, np, kBmDefaultStep, p.block_size);
    map_int_array(in, "bm_range", p.bm_range, np, kBm3dDefaultRange);
    map_adaptive(in, "ps_num", p.ps_num, np, kBmDefaultPsNum, p.group_size);
    map_int_array(in, "ps_range", p.ps_range, np, kBmDefaultPsRange);
    p.radius = map_int(in, "radius", 0);
    if (p.radius < 0 || p.radius > kBmMaxRadius) fail(ns, "radius must be in [0, 16]");
    parse_chroma(in, vi, ns, p);
    const int final = map_int(in, "final", 0);
    if (final != 0 && final != 1) fail(ns, "final must be 0 or 1");
    p.final_stage = final != 0;
    const bool sigma_basic = in.num("sigma_basic") > 0;
    const bool block_basic = in.num("block_size_basic") > 0;
    const bool group_basic = in.num("group_size_basic") > 0;
    if (!p.final_stage && (sigma_basic || block_basic || group_basic))
        fail(ns, "sigma_basic, block_size_basic and group_size_basic require final=1");
    if (p.final_stage && ref_vi) fail(ns, "use only one of final and ref");
    if (p.final_stage) parse_basic_stage(in, vi, ns, sigma_basic, block_basic, group_basic, sigma_given, p);
    check_planes(vi, ns, p);
}

// ---- WNNM (src/frontend/wnnm_args.cpp:8-29, include/nss/params.hpp:61-75) ----
constexpr int kWnnmMaxBlock = 16, kWnnmMaxGroup = 32, kWnnmDefaultBlock = 8, kWnnmDefaultStep = 8;
constexpr int kWnnmDefaultGroup = 8, kWnnmDefaultRange = 7, kWnnmDefaultResidual = 0, kWnnmDefaultAdaptive = 1;
constexpr float kWnnmDefaultSigma = 3.0f;

[[noreturn]] void fail_wnnm(const std::string& ns, const std::string& message) {
    throw Invalid(ns + ".WNNM: " + message);
}

void parse_wnnm(const ArgMap& in, const nss_hip_video_info& vi, const std::string& ns, nss_hip_wnnm_params& p) {
    if (!is_const_32f(vi)) fail_wnnm(ns, "constant Gray/YUV/RGB 32-bit float required");
    map_float_array(in, "sigma", p.sigma, vi.num_planes, kWnnmDefaultSigma);
    for (int i = vi.num_planes; i < 3; ++i) p.sigma[i] = 0.f;
    p.block_size = map_int(in, "block_size", kWnnmDefaultBlock);
    p.block_step = map_int(in, "block_step", kWnnmDefaultStep);
    p.group_size = map_int(in, "group_size", kWnnmDefaultGroup);
    p.bm_range = map_int(in, "bm_range", kWnnmDefaultRange);
    p.radius = map_int(in, "radius", 0);
    p.ps_num = map_int(in, "ps_num", kBmDefaultPsNum);
    p.ps_range = map_int(in, "ps_range", kBmDefaultPsRange);
    p.residual = map_int(in, "residual", kWnnmDefaultResidual);
    p.adaptive = map_int(in, "adaptive_aggregation", kWnnmDefaultAdaptive);
    if (p.block_size < 1 || p.block_size > kWnnmMaxBlock || p.group_size < 1 || p.group_size > kWnnmMaxGroup ||
        p.block_step < 1 || p.block_step > p.block_size || p.radius < 0 || p.radius > kBmMaxRadius)
        fail_wnnm(ns, "invalid block_size/group_size/block_step/radius");
    if (p.bm_range < 1 || p.bm_range > kBmMaxRange) fail_wnnm(ns, "bm_range must be in [1, 64]");
    if (p.ps_num < 1 || p.ps_num > p.group_size || p.ps_range < 1 || p.ps_range > kBmMaxRange)
        fail_wnnm(ns, "invalid ps_num/ps_range");
}

// ---- NCSR (src/frontend/ncsr_args.cpp:10-32, include/nss/params.hpp:101-107) ----
constexpr int kNcsrDefaultIters = 2;
constexpr float kNcsrDefaultDelta = 0.1f;

[[noreturn]] void fail_ncsr(const std::string& ns, const std::string& message) {
    throw Invalid(ns + ".NCSR: " + message);
}

void parse_ncsr(const ArgMap& in, const nss_hip_video_info& vi, const std::string& ns, nss_hip_ncsr_params& p) {
    if (!is_const_32f(vi)) fail_ncsr(ns, "constant Gray/YUV/RGB 32-bit float required");
    map_float_array(in, "sigma", p.sigma, vi.num_planes, kWnnmDefaultSigma);
    for (int i = vi.num_planes; i < 3; ++i) p.sigma[i] = 0.f;
    p.block_size = map_int(in, "block_size", kWnnmDefaultBlock);
    p.block_step = map_int(in, "block_step", kWnnmDefaultStep);
    p.group_size = map_int(in, "group_size", kWnnmDefaultGroup);
    p.bm_range = map_int(in, "bm_range", kWnnmDefaultRange);
    p.radius = map_int(in, "radius", 0);
    p.ps_num = map_int(in, "ps_num", kBmDefaultPsNum);
    p.ps_range = map_int(in, "ps_range", kBmDefaultPsRange);
    p.iters = map_int(in, "iters", kNcsrDefaultIters);
    int err = 0;
    const double delta = in.get_float("delta", 0, &err);
    p.delta = err ? kNcsrDefaultDelta : static_cast<float>(delta);
    if (p.block_size < 1 || p.block_size > kWnnmMaxBlock || p.group_size < 1 || p.group_size > kWnnmMaxGroup ||
        p.block_step < 1 || p.block_step > p.block_size || p.radius < 0 || p.radius > kBmMaxRadius)
        fail_ncsr(ns, "invalid block_size/group_size/block_step/radius");
    if (p.iters < 1 || p.iters > 64 || !finite_bits(p.delta) || p.delta < 0.f || p.delta > 1.f)
        fail_ncsr(ns, "invalid iters/delta (iters in [1, 64], delta in [0, 1])");
    if (p.ps_num < 1 || p.ps_num > p.group_size || p.ps_range < 1 || p.ps_range > kBmMaxRange)
        fail_ncsr(ns, "invalid ps_num/ps_range");
    if (p.bm_range < 1 || p.bm_range > kBmMaxRange) fail_ncsr(ns, "bm_range must be in [1, 64]");
}

}  // namespace

extern "C" {

int nss_hip_bm3d_parse(const nss_hip_video_info* clip_info, const nss_hip_video_info* ref_info,
                       const nss_hip_arg* args, int nargs, const char* ns_c, nss_hip_bm3d_params* out) {
    const std::string ns = ns_c ? ns_c : "nss_hip";
    if (!clip_info || !out || (nargs > 0 && !args)) {
        nsship::g_last_error = ns + ".BM3D: null argument";
        return NSS_HIP_INVALID_INPUT;
    }
    try {
        const ArgMap in(args, nargs);
        validate_numeric(in, ns);
        nss_hip_bm3d_params p{};
        int err = 0;
        const std::int64_t mb = in.get_int("memory_limit_mb", 0, &err);
        if (!err && (mb < 1 || mb > 1048576)) throw Invalid(ns + ": memory_limit_mb must be in [1, 1048576]");
        p.memory_limit_mb = err ? 0 : static_cast<int>(mb);
        const int radius = map_int(in, "radius", 0);
        const bool rolling = parse_temporal_mode(in, radius, ns);
        parse_bm3d(in, *clip_info, ref_info, ns, p);
        p.has_ref = ref_info != nullptr;
        p.temporal_mode = rolling ? 1 : 0;
        p.rolling_chunk = 1;
        p.rolling_cache_chunks = 1;
        p.rolling_cache_limit = 16;
        p.rolling_cache_adaptive = 1;
        if (rolling) parse_rolling(in, p.radius, ns, p);
        const std::int64_t device = in.get_int("device_id", 0, &err);
        if (!err && (device < 0 || device > INT32_MAX)) throw Invalid(ns + ".BM3D: device_id must be >= 0");
        p.device_id = err ? 0 : static_cast<int>(device);
        const std::int64_t streams = in.get_int("num_streams", 0, &err);
        if (!err && (streams < 1 || streams > kMaxStreams))
            throw Invalid(ns + ".BM3D: num_streams must be in [1, " + std::to_string(kMaxStreams) + "]");
        p.num_streams = err ? 1 : static_cast<int>(streams);
        const std::int64_t form = in.get_int("transform", 0, &err);
        if (!err && (form < NSS_HIP_TRANSFORM_FP32 || form > NSS_HIP_TRANSFORM_FP16))
            throw Invalid(ns + ".BM3D: transform must be 0 (fp32), 1 (bf16x3) or 2 (fp16)");
        p.transform = err ? NSS_HIP_TRANSFORM_FP32 : static_cast<int>(form);
        *out = p;
        nsship::g_last_error.clear();
        return NSS_HIP_OK;
    } catch (const std::exception& e) {
        nsship::g_last_error = e.what();
        return NSS_HIP_INVALID_INPUT;
    }
}

int nss_hip_bm3d_output_info(const nss_hip_bm3d_params* p, const nss_hip_video_info* clip_info,
                             nss_hip_video_info* out) {
    if (!p || !clip_info || !out) {
        nsship::g_last_error = "nss_hip: null argument";
        return NSS_HIP_INVALID_INPUT;
    }
    *out = *clip_info;
    // Legacy temporal output: 2r+1 num/den slice pairs stacked vertically
    // (src/host/filter_bm3d.cpp:1155-1157, include/nss/checked.hpp:46-49).
    if (p->radius > 0 && !p->temporal_mode) {
        const std::uint64_t h = static_cast<std::uint64_t>(clip_info->height) * (2ull * p->radius + 1) * 2;
        if (h > static_cast<std::uint64_t>(std::numeric_limits<int>::max())) {
            nsship::g_last_error = "nss_hip: invalid temporal dimensions";
            return NSS_HIP_INVALID_INPUT;
        }
        out->height = static_cast<int>(h);
    }
    return NSS_HIP_OK;
}

// <ns>.WNNM's create-time checks in the reference's order: numeric
// representability and memory_limit_mb (src/frontend/validate.hpp:33-45),
// temporal_mode (src/frontend/temporal_args.cpp:11-28), parse_wnnm, rclip
// against clip, validate_group_planes (src/frontend/validate.hpp:19-25).
int nss_hip_wnnm_parse(const nss_hip_video_info* clip_info, const nss_hip_video_info* rclip_info,
                       const nss_hip_arg* args, int nargs, const char* ns_c, nss_hip_wnnm_params* out) {
    const std::string ns = ns_c ? ns_c : "nss_hip";
    if (!clip_info || !out || (nargs > 0 && !args)) {
        nsship::g_last_error = ns + ".WNNM: null argument";
        return NSS_HIP_INVALID_INPUT;
    }
    try {
        const ArgMap in(args, nargs);
        validate_numeric(in, ns);
        nss_hip_wnnm_params p{};
        int err = 0;
        const std::int64_t mb = in.get_int("memory_limit_mb", 0, &err);
        if (!err && (mb < 1 || mb > 1048576)) throw Invalid(ns + ": memory_limit_mb must be in [1, 1048576]");
        p.memory_limit_mb = err ? 0 : static_cast<int>(mb);
        const char* mode = in.get_data("temporal_mode", &err);
        std::string m;
        if (!err && mode) {
            m = mode;
            for (char& c : m)
                if (c >= 'A' && c <= 'Z') c = static_cast<char>(c - 'A' + 'a');
        }
        if (m == "fused") fail_wnnm(ns, "temporal_mode=fused is not supported; use rolling or legacy");
        if (!m.empty() && m != "rolling" && m != "legacy") fail_wnnm(ns, "temporal_mode must be rolling or legacy");
        parse_wnnm(in, *clip_info, ns, p);
        if (rclip_info && !same_video(*clip_info, *rclip_info)) fail_wnnm(ns, "rclip must match clip");
        p.has_rclip = rclip_info != nullptr;
        for (int i = 0; i < clip_info->num_planes; ++i) {
            if (p.sigma[i] < 0.f) throw Invalid("nss: sigma must be non-negative");
            if (p.sigma[i] > 0.f &&
                (plane_width(*clip_info, i) < p.block_size || plane_height(*clip_info, i) < p.block_size))
                throw Invalid("nss: selected plane is smaller than block_size");
        }
        p.temporal_mode = m == "legacy" ? 0 : 1;
        const std::int64_t device = in.get_int("device_id", 0, &err);
        if (!err && (device < 0 || device > INT32_MAX)) fail_wnnm(ns, "device_id must be >= 0");
        p.device_id = err ? 0 : static_cast<int>(device);
        *out = p;
        nsship::g_last_error.clear();
        return NSS_HIP_OK;
    } catch (const std::exception& e) {
        nsship::g_last_error = e.what();
        return NSS_HIP_INVALID_INPUT;
    }
}

// <ns>.NCSR's create-time checks, in the order of nss_hip_wnnm_parse.
int nss_hip_ncsr_parse(const nss_hip_video_info* clip_info, const nss_hip_video_info* rclip_info,
                       const nss_hip_arg* args, int nargs, const char* ns_c, nss_hip_ncsr_params* out) {
    const std::string ns = ns_c ? ns_c : "nss_hip";
    if (!clip_info || !out || (nargs > 0 && !args)) {
        nsship::g_last_error = ns + ".NCSR: null argument";
        return NSS_HIP_INVALID_INPUT;
    }
    try {
        const ArgMap in(args, nargs);
        validate_numeric(in, ns);
        nss_hip_ncsr_params p{};
        int err = 0;
        const std::int64_t mb = in.get_int("memory_limit_mb", 0, &err);
        if (!err && (mb < 1 || mb > 1048576)) throw Invalid(ns + ": memory_limit_mb must be in [1, 1048576]");
        p.memory_limit_mb = err ? 0 : static_cast<int>(mb);
        const char* mode = in.get_data("temporal_mode", &err);
        std::string m;
        if (!err && mode) {
            m = mode;
            for (char& c : m)
                if (c >= 'A' && c <= 'Z') c = static_cast<char>(c - 'A' + 'a');
        }
        if (m == "fused") fail_ncsr(ns, "temporal_mode=fused is not supported; use rolling or legacy");
        if (!m.empty() && m != "rolling" && m != "legacy") fail_ncsr(ns, "temporal_mode must be rolling or legacy");
        parse_ncsr(in, *clip_info, ns, p);
        if (rclip_info && !same_video(*clip_info, *rclip_info)) fail_ncsr(ns, "rclip must match clip");
        p.has_rclip = rclip_info != nullptr;
        for (int i = 0; i < clip_info->num_planes; ++i) {
            if (p.sigma[i] < 0.f) throw Invalid("nss: sigma must be non-negative");
            if (p.sigma[i] > 0.f &&
                (plane_width(*clip_info, i) < p.block_size || plane_height(*clip_info, i) < p.block_size))
                throw Invalid("nss: selected plane is smaller than block_size");
        }
        p.temporal_mode = m == "legacy" ? 0 : 1;
        const std::int64_t device = in.get_int("device_id", 0, &err);
        if (!err && (device < 0 || device > INT32_MAX)) fail_ncsr(ns, "device_id must be >= 0");
        p.device_id = err ? 0 : static_cast<int>(device);
        *out = p;
        nsship::g_last_error.clear();
        return NSS_HIP_OK;
    } catch (const std::exception& e) {
        nsship::g_last_error = e.what();
        return NSS_HIP_INVALID_INPUT;
    }
}

}  // extern "C"
