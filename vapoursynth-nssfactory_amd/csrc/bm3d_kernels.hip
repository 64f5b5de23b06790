// BM3D kernels for CDNA4 (gfx950): block matching, group filter with
// weighted fixed-point aggregation, and finish.
//
//   match_kernel<B, STAGE>   one wave per reference: the clamped search
//                            window (plus the block overhang) staged in LDS,
//                            SSD of every candidate in FP32, then the G-1
//                            smallest keys (dist, y, x) by rounds of a
//                            wave-wide 64-bit min.
//   filter_kernel<B, WIENER> one wave per group: gather the k matched patches
//                            (zero-padded to the configured group size) into
//                            an LDS cube, 2-D DCT per patch, orthonormal Haar
//                            along the group, hard threshold or Wiener
//                            shrinkage, inverse, weight, and scatter-add into
//                            int64 fixed-point num/den planes (order-free,
//                            so results are run-to-run identical).
//   finish_kernel<B>         den is kept at each patch's top-left corner; the
//                            finish box-sums it over B x B and writes
//                            num/den, or the source where nothing landed.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bm3d_jobs.hpp"

namespace nsship {

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

__device__ inline unsigned long long wave_min_u64(unsigned long long v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    return v;
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
__device__ inline int find_job(const Table& tab) {
    int j = 0;
    while (j + 1 < tab.n && static_cast<int>(blockIdx.x) >= tab.job[j + 1].wg_begin) ++j;
    return j;
}

// ---------------------------------------------------------------------------
// Block matching (src/cpu/bm/matcher.hpp:247-283 collect_spatial_coords;
// order src/cpu/bm/matcher.hpp:27-57 match_less).
// ---------------------------------------------------------------------------
template <int B, bool STAGE>
__global__ __launch_bounds__(64) void match_kernel(MatchTable tab) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    const MatchJob& jb = tab.job[find_job(tab)];
    const int ref = static_cast<int>(blockIdx.x) - jb.wg_begin;
    if (ref >= jb.grid.nx * jb.grid.ny) return;
    const int cx = min((ref % jb.grid.nx) * jb.grid.step, jb.width - B);
    const int cy = min((ref / jb.grid.nx) * jb.grid.step, jb.height - B);
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
        float* lreg = reinterpret_cast<float*>(smem);
        for (int yy = 0; yy < rh; ++yy) {
            const float* row = jb.plane + static_cast<size_t>(top + yy) * jb.stride + left;
            for (int xx = lane; xx < rw; xx += 64) lreg[yy * rw + xx] = row[xx];
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
        uint32_t key = __float_as_uint(acc);
        if ((key & kNonFinite) == kNonFinite) key = kNonFinite;
        dk[c] = c == self ? kTaken : key;
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
// Group filter (src/cpu/bm/dct8.cpp:926-1133 Bm3dFilterGroup) and
// aggregation (src/cpu/bm/pack.cpp:42-86 UnpackPatch; fixed point as in the
// design of src/cuda/common/fixed_accumulate.cuh:18-35, den at the corner).
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

// DCT of the lines of every patch of a cube (patch stride ps floats).
// ROWS: lines along x (sample stride 1); else along y (sample stride B).
template <int B, bool ROWS>
__device__ inline void dct_lines(float* cube, int group, int ps, bool inverse) {
    if constexpr (B > 1) {
        const float* m = c_dct + dct_offset<B>();
        const int lines = group * B;
        for (int L = threadIdx.x; L < lines; L += 64) {
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

__device__ inline unsigned long long to_fixed(float v) {
    return static_cast<unsigned long long>(__float2ll_rn(v * 4294967296.0f));
}

template <int B, bool WIENER>
__global__ __launch_bounds__(64) void filter_kernel(FilterTable tab) {
    extern __shared__ __align__(16) float cube[];
    constexpr int A = B * B;
    constexpr int PS = A + 1;
    const int lane = threadIdx.x;
    const FilterJob& jb = tab.job[find_job(tab)];
    const int g = static_cast<int>(blockIdx.x) - jb.wg_begin;
    if (g >= jb.nref) return;
    const int G = jb.group;
    const int k = jb.counts[g];
    const MatchRec* rec = jb.matches + static_cast<size_t>(g) * G;
    float* rc = cube + G * PS;

    for (int j = 0; j < G; ++j) {
        int mx = 0, my = 0;
        if (j < k) {
            mx = rec[j].x;
            my = rec[j].y;
        }
        for (int e = lane; e < A; e += 64) {
            const int py = e / B, px = e - (e / B) * B;
            float v = 0.f, w = 0.f;
            if (j < k) {
                v = jb.src[static_cast<size_t>(my + py) * jb.stride + mx + px];
                if constexpr (WIENER) w = jb.wref[static_cast<size_t>(my + py) * jb.wref_stride + mx + px];
            }
            cube[j * PS + e] = v;
            if constexpr (WIENER) rc[j * PS + e] = w;
        }
    }
    __syncthreads();
    dct_lines<B, true>(cube, G, PS, false);
    if constexpr (WIENER) dct_lines<B, true>(rc, G, PS, false);
    __syncthreads();
    dct_lines<B, false>(cube, G, PS, false);
    if constexpr (WIENER) dct_lines<B, false>(rc, G, PS, false);
    __syncthreads();

    int kept = 0;
    float w2 = 0.f;
    const float thr = 2.7f * jb.sigma;  // include/nss/params.hpp:46 kBmHardLambda
    const float sig2 = jb.sigma * jb.sigma;
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
    __syncthreads();
    dct_lines<B, false>(cube, G, PS, true);
    __syncthreads();
    dct_lines<B, true>(cube, G, PS, true);
    __syncthreads();

    float weight;
    if constexpr (WIENER) {
        weight = 1.f / fmaxf(wave_sum_f32(w2), 1e-12f);
    } else {
        weight = 1.f / static_cast<float>(max(wave_sum_i32(kept), 1));
    }
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

// ---------------------------------------------------------------------------
// Finish (src/cpu/bm/agg.cpp:17-42 AggFinish): dst = den > 0 ? num/den : src,
// den being the B x B box sum of the corner weights.
// ---------------------------------------------------------------------------
template <int B>
__global__ __launch_bounds__(256) void finish_kernel(FinishTable tab) {
    constexpr int T = 16;
    constexpr int TS = T + B - 1;
    __shared__ unsigned long long corner[TS * TS];
    __shared__ unsigned long long colsum[T * TS];
    const FinishJob& jb = tab.job[find_job(tab)];
    const int tile = static_cast<int>(blockIdx.x) - jb.wg_begin;
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
        jb.dst[static_cast<size_t>(y) * jb.dstride + x] =
            dd > 0 ? static_cast<float>(static_cast<double>(n) / static_cast<double>(dd))
                   : jb.src[static_cast<size_t>(y) * jb.sstride + x];
    }
}

__global__ __launch_bounds__(256) void zero_kernel(ZeroTable tab) {
    const ZeroJob& jb = tab.job[find_job(tab)];
    const long long base = static_cast<long long>(static_cast<int>(blockIdx.x) - jb.wg_begin) * 1024;
    for (int i = threadIdx.x; i < 1024; i += 256) {
        const long long o = base + i;
        if (o < jb.count) jb.p[o] = 0ull;
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers.
// ---------------------------------------------------------------------------

size_t match_lds_bytes(int block, int range, bool stage) {
    const int w = 2 * range + 1;
    const size_t cand = static_cast<size_t>(w) * w * 4;
    if (!stage) return cand;
    const size_t reg = static_cast<size_t>(w + block - 1) * (w + block - 1);
    return ((reg + 3) & ~static_cast<size_t>(3)) * 4 + cand;
}

size_t filter_lds_bytes(int block, int group, bool wiener) {
    return static_cast<size_t>(group) * (block * block + 1) * 4 * (wiener ? 2 : 1);
}

template <int B>
static hipError_t launch_match_b(const MatchTable& t, int total_wg, size_t lds, bool stage, hipStream_t s) {
    if (stage) {
        if (lds > 65536)
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&match_kernel<B, true>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        match_kernel<B, true><<<total_wg, 64, lds, s>>>(t);
    } else {
        if (lds > 65536)
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&match_kernel<B, false>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        match_kernel<B, false><<<total_wg, 64, lds, s>>>(t);
    }
    return hipGetLastError();
}

hipError_t launch_match(int block, const MatchTable& t, int total_wg, size_t lds, bool stage, hipStream_t s) {
    switch (block) {
        case 1: return launch_match_b<1>(t, total_wg, lds, stage, s);
        case 2: return launch_match_b<2>(t, total_wg, lds, stage, s);
        case 4: return launch_match_b<4>(t, total_wg, lds, stage, s);
        case 8: return launch_match_b<8>(t, total_wg, lds, stage, s);
        case 12: return launch_match_b<12>(t, total_wg, lds, stage, s);
        case 16: return launch_match_b<16>(t, total_wg, lds, stage, s);
        case 32: return launch_match_b<32>(t, total_wg, lds, stage, s);
        default: return hipErrorInvalidValue;
    }
}

template <int B>
static hipError_t launch_filter_b(const FilterTable& t, int total_wg, size_t lds, bool wiener, hipStream_t s) {
    if (wiener) {
        if (lds > 65536)
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&filter_kernel<B, true>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        filter_kernel<B, true><<<total_wg, 64, lds, s>>>(t);
    } else {
        if (lds > 65536)
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&filter_kernel<B, false>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        filter_kernel<B, false><<<total_wg, 64, lds, s>>>(t);
    }
    return hipGetLastError();
}

hipError_t launch_filter(int block, const FilterTable& t, int total_wg, size_t lds, bool wiener, hipStream_t s) {
    switch (block) {
        case 1: return launch_filter_b<1>(t, total_wg, lds, wiener, s);
        case 2: return launch_filter_b<2>(t, total_wg, lds, wiener, s);
        case 4: return launch_filter_b<4>(t, total_wg, lds, wiener, s);
        case 8: return launch_filter_b<8>(t, total_wg, lds, wiener, s);
        case 12: return launch_filter_b<12>(t, total_wg, lds, wiener, s);
        case 16: return launch_filter_b<16>(t, total_wg, lds, wiener, s);
        case 32: return launch_filter_b<32>(t, total_wg, lds, wiener, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_finish(int block, const FinishTable& t, int total_wg, hipStream_t s) {
    switch (block) {
        case 1: finish_kernel<1><<<total_wg, 256, 0, s>>>(t); break;
        case 2: finish_kernel<2><<<total_wg, 256, 0, s>>>(t); break;
        case 4: finish_kernel<4><<<total_wg, 256, 0, s>>>(t); break;
        case 8: finish_kernel<8><<<total_wg, 256, 0, s>>>(t); break;
        case 12: finish_kernel<12><<<total_wg, 256, 0, s>>>(t); break;
        case 16: finish_kernel<16><<<total_wg, 256, 0, s>>>(t); break;
        case 32: finish_kernel<32><<<total_wg, 256, 0, s>>>(t); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_zero(const ZeroTable& t, int total_wg, hipStream_t s) {
    zero_kernel<<<total_wg, 256, 0, s>>>(t);
    return hipGetLastError();
}

hipError_t upload_dct_tables(const float* host, size_t count) {
    if (count != sizeof(c_dct) / sizeof(float)) return hipErrorInvalidValue;
    return hipMemcpyToSymbol(HIP_SYMBOL(c_dct), host, sizeof(c_dct), 0, hipMemcpyHostToDevice);
}

}  // namespace nsship
