// CPU oracle for the BM3D hot path. TEST INFRASTRUCTURE ONLY: the parity
// tests, the smoke check and the CPU-baseline leg of the benchmark call it as
// the reference results are compared against; the product library never
// links, loads or calls it.
//
// A from-scratch restatement of the reference CPU plugin's nss.BM3D
// algorithm (numeric build NSS_BM_EXPERIMENT=0), in two arithmetic modes:
//   fp32 ("fp32-ref"): float storage and float arithmetic in the reference's
//         order where the order decides a result (threshold compare, Wiener
//         weight, weight sums, raster-ordered aggregation, finish);
//   fp64: the same algorithm in double throughout (error oracle for the
//         reduced-precision GPU forms).
// Each function cites the reference file:line it follows. Exact bits of the
// reference's SIMD summation trees are not reproduced (the reference's own
// contract allows 8e-6 relative on distances and 3e-5 on outputs).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int kMaxGroup = 64;        // include/nss/params.hpp:39 kBmMaxGroup
constexpr float kHardLambda = 2.7f;  // include/nss/params.hpp:46 kBmHardLambda

struct Match {
    int x, y, t;
    float dist;
    std::uint32_t ordinal;
};

bool finite_bits(float v) {
    std::uint32_t u;
    std::memcpy(&u, &v, 4);
    return (u & 0x7f800000u) != 0x7f800000u;
}

// src/cpu/bm/matcher.hpp:27-45 match_tie_less
bool tie_less(const Match& a, const Match& b) {
    const bool as = a.ordinal == 0, bs = b.ordinal == 0;
    if (as != bs) return as;
    if (a.t != b.t) return a.t < b.t;
    if (a.y != b.y) return a.y < b.y;
    if (a.x != b.x) return a.x < b.x;
    return a.ordinal < b.ordinal;
}

// src/cpu/bm/matcher.hpp:47-57 match_less
bool match_less(const Match& a, const Match& b) {
    const bool af = finite_bits(a.dist), bf = finite_bits(b.dist);
    if (af != bf) return af;
    if (af && a.dist != b.dist) return a.dist < b.dist;
    return tie_less(a, b);
}

// Scalar SSD in row-major order (tests/test_matcher.cpp:40-49 scalar_ssd).
template <typename Real>
float ssd(const float* a, int sa, const float* b, int sb, int block) {
    Real sum = 0;
    for (int y = 0; y < block; ++y)
        for (int x = 0; x < block; ++x) {
            const Real d = Real(a[y * sa + x]) - Real(b[y * sb + x]);
            sum += d * d;
        }
    return static_cast<float>(sum);
}

// Keeps the `cap` smallest candidates by match_less, sorted on finish. Same
// result as the reference's StableTopK (src/cpu/bm/matcher.hpp:59-109): the
// order is total over distinct (t, y, x), so the retained set is unique.
struct TopK {
    std::vector<Match> v;
    int cap;
    explicit TopK(int c) : cap(c) {}
    void add(const Match& m) {
        if (cap <= 0) return;
        if (static_cast<int>(v.size()) < cap) {
            v.push_back(m);
            return;
        }
        int worst = 0;
        for (int i = 1; i < static_cast<int>(v.size()); ++i)
            if (match_less(v[worst], v[i])) worst = i;
        if (match_less(m, v[worst])) v[worst] = m;
    }
    int finish() {
        std::sort(v.begin(), v.end(), match_less);
        return static_cast<int>(v.size());
    }
};

// src/cpu/bm/matcher.hpp:247-283 collect_spatial_coords (window clamp, raster
// scan, self skipped, reference first with distance 0).
template <typename Real>
int spatial_match(const float* plane, int stride, int width, int height, int bx, int by, int block, int range,
                  int group, Match* out) {
    if (width < block || height < block || block < 1 || group < 1) return 0;
    const int wanted = std::min(group, kMaxGroup);
    const int max_x = width - block, max_y = height - block;
    const int cx = std::clamp(bx, 0, max_x), cy = std::clamp(by, 0, max_y);
    const int r = std::max(range, 0);
    const int top = std::max(cy - r, 0), bottom = std::min(cy + r, max_y);
    const int left = std::max(cx - r, 0), right = std::min(cx + r, max_x);
    out[0] = Match{cx, cy, 0, 0.f, 0};
    if (wanted == 1) return 1;
    TopK topk(wanted - 1);
    std::uint32_t ordinal = 1;
    const float* refp = plane + static_cast<std::ptrdiff_t>(cy) * stride + cx;
    for (int y = top; y <= bottom; ++y)
        for (int x = left; x <= right; ++x, ++ordinal) {
            if (x == cx && y == cy) continue;
            const float d = ssd<Real>(refp, stride, plane + static_cast<std::ptrdiff_t>(y) * stride + x, stride, block);
            topk.add(Match{x, y, 0, d, ordinal});
        }
    const int n = topk.finish();
    for (int i = 0; i < n; ++i) out[1 + i] = topk.v[i];
    return 1 + n;
}

// src/cpu/bm/matcher.hpp:297-314 assign_temporal_order
void assign_temporal_order(Match* m, int count, int t, int t0, std::uint32_t& next) {
    constexpr std::uint32_t kFrameStride = 1u << 20;
    for (int i = 0; i < count; ++i) {
        const std::uint32_t local = m[i].ordinal;
        m[i].t = t;
        if (t == t0 && local == 0)
            m[i].ordinal = 0;
        else
            m[i].ordinal = static_cast<std::uint32_t>(std::max(t, 0) + 1) * kFrameStride + local;
        next = std::max(next, m[i].ordinal + 1u);
    }
}

// src/cpu/bm/match.cpp:10-33 predictive_match and src/cpu/bm/matcher.hpp:319-373
// collect_temporal: spatial search on t0, then per direction a chain of
// predictive windows (union of (2 ps_range + 1)^2 windows around the previous
// layer's winners, raster order, overlaps visited once).
template <typename Real>
int predictive_match(const float* const* planes, int stride, int ntemp, int width, int height, int bx, int by,
                     int t0, int block, int group, int range, int ps_num, int ps_range, int radius, int valid_begin,
                     int valid_end, Match* out) {
    int n = spatial_match<Real>(planes[t0], stride, width, height, bx, by, block, range, group, out);
    const int cx = std::clamp(bx, 0, width - block), cy = std::clamp(by, 0, height - block);
    const float* reference = planes[t0] + static_cast<std::ptrdiff_t>(cy) * stride + cx;
    std::uint32_t ordinal = 0;
    assign_temporal_order(out, n, t0, t0, ordinal);
    if (n <= 0 || radius == 0 || ntemp == 1 || group == 1) return n;
    Match seeds[kMaxGroup];
    const int seed_count = std::min({n, ps_num, kMaxGroup});
    std::copy_n(out, seed_count, seeds);
    // The global top-k adopts the spatial matches out[1, n) (StableTopK::adopt).
    TopK global(group - 1);
    for (int i = 1; i < n; ++i) global.v.push_back(out[i]);
    const int first = std::max(0, valid_begin);
    const int last = valid_end < 0 ? ntemp : std::min(ntemp, valid_end);
    for (int sign : {-1, 1}) {
        Match centers[kMaxGroup];
        std::copy_n(seeds, seed_count, centers);
        int nc = seed_count;
        for (int dt = 1; dt <= radius; ++dt) {
            const int t = t0 + sign * dt;
            if (t < first || t >= last) break;
            TopK top(std::min(ps_num, group));
            int ymin = height, ymax = -1;
            for (int i = 0; i < nc; ++i) {
                ymin = std::min(ymin, std::max(0, centers[i].y - ps_range));
                ymax = std::max(ymax, std::min(height - block, centers[i].y + ps_range));
            }
            std::uint32_t visit = 1;
            for (int y = ymin; y <= ymax; ++y) {
                std::pair<int, int> iv[kMaxGroup];
                int ni = 0;
                for (int i = 0; i < nc; ++i)
                    if (std::abs(y - centers[i].y) <= ps_range)
                        iv[ni++] = {std::max(0, centers[i].x - ps_range), std::min(width - block, centers[i].x + ps_range)};
                std::sort(iv, iv + ni);
                int done = -1;
                for (int i = 0; i < ni; ++i) {
                    for (int x = std::max(done + 1, iv[i].first); x <= iv[i].second; ++x) {
                        const float d = ssd<Real>(reference, stride,
                                                  planes[t] + static_cast<std::ptrdiff_t>(y) * stride + x, stride, block);
                        Match m{x, y, t, d, visit++};
                        assign_temporal_order(&m, 1, t, t0, ordinal);
                        top.add(m);
                    }
                    done = std::max(done, iv[i].second);
                }
            }
            nc = top.finish();
            for (int i = 0; i < nc; ++i) global.add(top.v[i]);
            std::copy_n(top.v.data(), nc, centers);
            if (nc == 0) break;
        }
    }
    const int k = global.finish();
    for (int i = 0; i < k; ++i) out[1 + i] = global.v[i];
    return 1 + k;
}

// Orthonormal DCT-II matrix, row k = frequency, column i = sample, entries
// computed in double (src/cpu/bm/dct8.cpp:71-89 DctTable).
template <typename Real>
std::vector<Real> dct_matrix(int n) {
    std::vector<Real> m(static_cast<std::size_t>(n) * n);
    const double pi = 3.14159265358979323846264338327950288;
    for (int k = 0; k < n; ++k)
        for (int i = 0; i < n; ++i) {
            const double c = std::sqrt((k == 0 ? 1.0 : 2.0) / n) * std::cos(pi * (i + 0.5) * k / n);
            // fp32 mode keeps the reference's float table; fp64 keeps the double.
            m[static_cast<std::size_t>(k) * n + i] =
                sizeof(Real) == sizeof(float) ? static_cast<Real>(static_cast<float>(c)) : static_cast<Real>(c);
        }
    return m;
}

// One length-n line at base[i * stride]: forward DCT-II or inverse DCT-III
// (src/cpu/bm/dct8.cpp:648-790 DctLines).
template <typename Real>
void dct_line(Real* base, int n, int stride, const Real* m, bool inverse) {
    Real in[64], outv[64];
    for (int i = 0; i < n; ++i) in[i] = base[i * stride];
    for (int k = 0; k < n; ++k) {
        Real acc = 0;
        for (int i = 0; i < n; ++i) acc += (inverse ? m[i * n + k] : m[k * n + i]) * in[i];
        outv[k] = acc;
    }
    for (int i = 0; i < n; ++i) base[i * stride] = outv[i];
}

// src/cpu/bm/dct8.cpp:580-591 HaarScale
double haar_scale(int n, int j) {
    int e = 0;
    if (j == 0) {
        while ((1 << e) < n) ++e;
    } else {
        while (((j >> e) & 1) == 0) ++e;
        ++e;
    }
    double f = 1.0;
    for (int i = 0; i < e / 2; ++i) f *= 0.5;
    return e % 2 ? f * 0.70710678118654752440 : f;
}

// In-place orthonormal Haar of one length-n line (src/cpu/bm/dct8.cpp:600-625
// HaarLines): level h pairs samples h apart; forward scales last, inverse first.
template <typename Real>
void haar_line(Real* base, int n, int stride, bool inverse) {
    if (n <= 1) return;
    auto at = [&](int i) -> Real& { return base[static_cast<std::ptrdiff_t>(i) * stride]; };
    auto scale = [&](int j) {
        return sizeof(Real) == sizeof(float) ? static_cast<Real>(static_cast<float>(haar_scale(n, j)))
                                             : static_cast<Real>(haar_scale(n, j));
    };
    if (inverse) {
        for (int j = 0; j < n; ++j) at(j) = at(j) * scale(j);
        for (int h = n / 2; h >= 1; h >>= 1)
            for (int i = 0; i < n; i += 2 * h) {
                const Real a = at(i), b = at(i + h);
                at(i) = a + b;
                at(i + h) = a - b;
            }
    } else {
        for (int h = 1; h < n; h <<= 1)
            for (int i = 0; i < n; i += 2 * h) {
                const Real a = at(i), b = at(i + h);
                at(i) = a + b;
                at(i + h) = a - b;
            }
        for (int j = 0; j < n; ++j) at(j) = at(j) * scale(j);
    }
}

template <typename Real>
void transform3d(Real* cube, int group, int block, const Real* m, bool inverse) {
    const int area = block * block;
    auto dct2 = [&](bool inv) {
        if (block <= 1) return;
        for (int g = 0; g < group; ++g) {
            Real* p = cube + static_cast<std::size_t>(g) * area;
            for (int y = 0; y < block; ++y) dct_line(p + y * block, block, 1, m, inv);  // rows
            for (int x = 0; x < block; ++x) dct_line(p + x, block, block, m, inv);      // columns
        }
    };
    if (!inverse) {
        dct2(false);
        for (int f = 0; f < area; ++f) haar_line(cube + f, group, area, false);
    } else {
        for (int f = 0; f < area; ++f) haar_line(cube + f, group, area, true);
        dct2(true);
    }
}

// src/cpu/bm/dct8.cpp:926-1133 Bm3dFilterGroup: cube of the configured group
// size, zero beyond k; 2-D DCT per patch, Haar along the group; hard
// threshold |c| < 2.7 sigma killed (DC kept), weight 1/max(kept, 1); or
// Wiener w = R^2/(R^2 + sigma^2) (0 when the denominator is not positive, DC
// 1), weight 1/max(sum w^2, 1e-12); inverse transforms.
template <typename Real>
float filter_group(Real* cube, int group, int k, int block, float sigma, bool wiener, Real* ref_cube) {
    const int area = block * block;
    const std::size_t n = static_cast<std::size_t>(group) * area;
    for (std::size_t i = static_cast<std::size_t>(k) * area; i < n; ++i) cube[i] = 0;
    if (wiener)
        for (std::size_t i = static_cast<std::size_t>(k) * area; i < n; ++i) ref_cube[i] = 0;
    const auto m = dct_matrix<Real>(block);
    transform3d(cube, group, block, m.data(), false);
    if (wiener) transform3d(ref_cube, group, block, m.data(), false);
    int kept = 0;
    Real w2sum = 0;
    if (wiener) {
        const Real sig2 = static_cast<Real>(sigma) * static_cast<Real>(sigma);
        for (std::size_t i = 0; i < n; ++i) {
            Real w = 1;
            if (i != 0) {
                const Real r = ref_cube[i];
                const Real r2 = r * r;
                const Real den = r2 + sig2;
                w = den > 0 ? r2 / den : Real(0);
            }
            cube[i] *= w;
            w2sum += w * w;
        }
    } else {
        const Real thr = static_cast<Real>(kHardLambda * sigma);
        for (std::size_t i = 0; i < n; ++i) {
            if (i != 0 && std::fabs(cube[i]) < thr)
                cube[i] = 0;
            else
                ++kept;
        }
    }
    transform3d(cube, group, block, m.data(), true);
    if (wiener) return static_cast<float>(Real(1) / std::max(w2sum, static_cast<Real>(1e-12f)));
    return static_cast<float>(Real(1) / static_cast<Real>(std::max(kept, 1)));
}

// Raster reference grid with the last row/column clamped
// (src/host/batch_runner.hpp:17-31 append_raster_jobs).
void raster_grid(int width, int height, int block, int step, std::vector<int>& xs, std::vector<int>& ys) {
    xs.clear();
    ys.clear();
    if (width < block || height < block || step < 1) return;
    for (long long by0 = 0; by0 < static_cast<long long>(height) - block + step; by0 += step)
        for (long long bx0 = 0; bx0 < static_cast<long long>(width) - block + step; bx0 += step) {
            xs.push_back(static_cast<int>(std::min<long long>(bx0, width - block)));
            ys.push_back(static_cast<int>(std::min<long long>(by0, height - block)));
        }
}

// src/host/filter_bm3d.cpp:73-339 process_planes_batched for one plane
// (single channel): matches on `match` planes, filters `src` (with `wref` as
// the Wiener reference), aggregates in raster-ordinal order into num/den
// slices (2r+1 of them), then finishes the centre slice
// (src/cpu/bm/agg.cpp:17-42): dst = den > 1e-12 ? num/den : src.
template <typename Real>
int bm3d_plane(const float* const* match, const float* const* src, const float* const* wref, int ntemp, int t0,
               int width, int height, int stride, float* dst, int dstride, float* num_out, float* den_out, float sigma,
               int block, int group, int step, int range, int ps_num, int ps_range, int radius, int wiener, int center,
               int frame_count, int* match_dump) {
    const int slices = 2 * radius + 1;
    const std::size_t plane = static_cast<std::size_t>(width) * height;
    std::vector<Real> num(plane * slices, 0), den(plane * slices, 0);
    std::vector<int> xs, ys;
    raster_grid(width, height, block, step, xs, ys);
    const int valid_begin = std::max(0, radius - center);
    const int valid_end = radius + std::min(radius + 1, frame_count - center);
    const int area = block * block;
    std::vector<Real> cube(static_cast<std::size_t>(group) * area), rcube(cube.size());
    Match ms[kMaxGroup];
    for (std::size_t j = 0; j < xs.size(); ++j) {
        int k;
        if (radius > 0)
            k = predictive_match<Real>(match, stride, ntemp, width, height, xs[j], ys[j], t0, block, group, range,
                                       ps_num, ps_range, radius, valid_begin, valid_end, ms);
        else
            k = spatial_match<Real>(match[t0], stride, width, height, xs[j], ys[j], block, range, group, ms);
        if (match_dump) {
            int* d = match_dump + j * (1 + 3 * kMaxGroup);
            d[0] = k;
            for (int i = 0; i < k; ++i) {
                d[1 + 3 * i] = ms[i].x;
                d[2 + 3 * i] = ms[i].y;
                d[3 + 3 * i] = radius > 0 ? ms[i].t : t0;
            }
        }
        if (k <= 0) return -1;
        for (int i = 0; i < k; ++i) {
            const int t = radius > 0 ? ms[i].t : t0;
            for (int py = 0; py < block; ++py)
                for (int px = 0; px < block; ++px) {
                    const std::size_t off = static_cast<std::size_t>(ms[i].y + py) * stride + ms[i].x + px;
                    cube[static_cast<std::size_t>(i) * area + py * block + px] = src[t][off];
                    if (wiener) rcube[static_cast<std::size_t>(i) * area + py * block + px] = wref[t][off];
                }
        }
        const float w = filter_group<Real>(cube.data(), group, k, block, sigma, wiener != 0, rcube.data());
        for (int i = 0; i < k; ++i) {
            const int t = radius > 0 ? ms[i].t : t0;
            const int sl = std::clamp(t - t0 + radius, 0, slices - 1);
            Real* np = num.data() + static_cast<std::size_t>(sl) * plane;
            Real* dp = den.data() + static_cast<std::size_t>(sl) * plane;
            for (int py = 0; py < block; ++py)
                for (int px = 0; px < block; ++px) {
                    const std::size_t o = static_cast<std::size_t>(ms[i].y + py) * width + ms[i].x + px;
                    const Real v = cube[static_cast<std::size_t>(i) * area + py * block + px];
                    if (sizeof(Real) == sizeof(float))
                        np[o] = static_cast<Real>(std::fma(static_cast<float>(w), static_cast<float>(v),
                                                           static_cast<float>(np[o])));
                    else
                        np[o] += static_cast<Real>(w) * v;
                    dp[o] += static_cast<Real>(w);
                }
        }
    }
    if (num_out && den_out) {
        for (std::size_t i = 0; i < plane * slices; ++i) {
            num_out[i] = static_cast<float>(num[i]);
            den_out[i] = static_cast<float>(den[i]);
        }
    }
    if (dst) {
        const Real* np = num.data() + static_cast<std::size_t>(t0 < slices ? radius : 0) * plane;
        const Real* dp = den.data() + static_cast<std::size_t>(t0 < slices ? radius : 0) * plane;
        for (int y = 0; y < height; ++y)
            for (int x = 0; x < width; ++x) {
                const std::size_t o = static_cast<std::size_t>(y) * width + x;
                const float s = src[t0][static_cast<std::size_t>(y) * stride + x];
                dst[static_cast<std::size_t>(y) * dstride + x] =
                    dp[o] > static_cast<Real>(1e-12f) ? static_cast<float>(np[o] / dp[o]) : s;
            }
    }
    return static_cast<int>(xs.size());
}

}  // namespace

extern "C" {

// Spatial match of one reference; writes count entries of (x, y, t) and dist.
int orc_spatial_match(const float* plane, int stride, int width, int height, int bx, int by, int block, int range,
                      int group, int fp64, int* xyt, float* dist) {
    Match ms[kMaxGroup];
    const int n = fp64 ? spatial_match<double>(plane, stride, width, height, bx, by, block, range, group, ms)
                       : spatial_match<float>(plane, stride, width, height, bx, by, block, range, group, ms);
    for (int i = 0; i < n; ++i) {
        xyt[3 * i] = ms[i].x;
        xyt[3 * i + 1] = ms[i].y;
        xyt[3 * i + 2] = ms[i].t;
        dist[i] = ms[i].dist;
    }
    return n;
}

int orc_predictive_match(const float* const* planes, int stride, int ntemp, int width, int height, int bx, int by,
                         int t0, int block, int group, int range, int ps_num, int ps_range, int radius,
                         int valid_begin, int valid_end, int fp64, int* xyt, float* dist) {
    Match ms[kMaxGroup];
    const int n = fp64 ? predictive_match<double>(planes, stride, ntemp, width, height, bx, by, t0, block, group,
                                                  range, ps_num, ps_range, radius, valid_begin, valid_end, ms)
                       : predictive_match<float>(planes, stride, ntemp, width, height, bx, by, t0, block, group, range,
                                                 ps_num, ps_range, radius, valid_begin, valid_end, ms);
    for (int i = 0; i < n; ++i) {
        xyt[3 * i] = ms[i].x;
        xyt[3 * i + 1] = ms[i].y;
        xyt[3 * i + 2] = ms[i].t;
        dist[i] = ms[i].dist;
    }
    return n;
}

// Group filter on a dense cube of group * block^2 floats (first k patches real).
float orc_filter_group(float* cube, int group, int k, int block, float sigma, int wiener, float* ref_cube, int fp64) {
    const std::size_t n = static_cast<std::size_t>(group) * block * block;
    if (fp64) {
        std::vector<double> c(cube, cube + n), r(n, 0.0);
        if (wiener) r.assign(ref_cube, ref_cube + n);
        const float w = filter_group<double>(c.data(), group, k, block, sigma, wiener != 0, r.data());
        for (std::size_t i = 0; i < n; ++i) cube[i] = static_cast<float>(c[i]);
        return w;
    }
    std::vector<float> r(n, 0.f);
    if (wiener) r.assign(ref_cube, ref_cube + n);
    return filter_group<float>(cube, group, k, block, sigma, wiener != 0, r.data());
}

// One plane of one stage. planes are arrays of ntemp pointers (slot t0 is the
// centre frame); wref may be null when wiener == 0. num/den (2r+1 slices of
// width*height) and match_dump (per reference: count then count * (x, y, t))
// are optional outputs. Returns the number of references, or -1.
int orc_bm3d_plane(const float* const* match, const float* const* src, const float* const* wref, int ntemp, int t0,
                   int width, int height, int stride, float* dst, int dstride, float* num, float* den, float sigma,
                   int block, int group, int step, int range, int ps_num, int ps_range, int radius, int wiener,
                   int center, int frame_count, int fp64, int* match_dump) {
    if (fp64)
        return bm3d_plane<double>(match, src, wref, ntemp, t0, width, height, stride, dst, dstride, num, den, sigma,
                                  block, group, step, range, ps_num, ps_range, radius, wiener, center, frame_count,
                                  match_dump);
    return bm3d_plane<float>(match, src, wref, ntemp, t0, width, height, stride, dst, dstride, num, den, sigma, block,
                             group, step, range, ps_num, ps_range, radius, wiener, center, frame_count, match_dump);
}

// Every reference of one plane's raster grid (raster_grid above): per
// reference `group` slots of (x, y) and distance in match order and its count, as
// spatial_match_batch writes them for WNNM (src/host/filter_wnnm.cpp:96-104).
int orc_spatial_match_plane(const float* plane, int stride, int width, int height, int block, int step, int range,
                            int group, int fp64, int* xy, int* counts, float* dist) {
    std::vector<int> xs, ys;
    raster_grid(width, height, block, step, xs, ys);
    Match ms[kMaxGroup];
    for (std::size_t r = 0; r < xs.size(); ++r) {
        const int n = fp64 ? spatial_match<double>(plane, stride, width, height, xs[r], ys[r], block, range, group, ms)
                           : spatial_match<float>(plane, stride, width, height, xs[r], ys[r], block, range, group, ms);
        counts[r] = n;
        for (int i = 0; i < n; ++i) {
            xy[(r * group + i) * 2] = ms[i].x;
            xy[(r * group + i) * 2 + 1] = ms[i].y;
            dist[r * group + i] = ms[i].dist;
        }
    }
    return static_cast<int>(xs.size());
}

int orc_max_group(void) { return kMaxGroup; }

}  // extern "C"
