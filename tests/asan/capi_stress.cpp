// Exercises the C ABI of nss_hip from several host threads, for a build whose
// host code carries AddressSanitizer (device code is built as usual). Without
// a device it covers argument resolution and the creation error paths; with
// one it also runs concurrent process calls on one handle, a rolling clip, the
// serving loop with client threads, the no-allocation entry point, WNNM, and
// the allocation / launch fault hooks. Any sanitizer report aborts the run;
// every C ABI failure prints its message and makes the exit status non-zero.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/nss_hip.h"

namespace {

std::mutex g_print_mu;
std::atomic<int> g_failures{0};

bool check(int rc, const char* what, int want = NSS_HIP_OK) {
    if (rc == want) return true;
    std::lock_guard<std::mutex> lock(g_print_mu);
    std::fprintf(stderr, "FAIL %s: rc %d (want %d): %s\n", what, rc, want, nss_hip_last_error());
    ++g_failures;
    return false;
}

void expect(bool ok, const char* what) {
    if (ok) return;
    std::lock_guard<std::mutex> lock(g_print_mu);
    std::fprintf(stderr, "FAIL %s\n", what);
    ++g_failures;
}

nss_hip_video_info video(int cf, int sw, int sh, int planes, int w, int h, int frames) {
    return nss_hip_video_info{cf, 1, 32, sw, sh, planes, w, h, frames};
}

const nss_hip_video_info kGray = video(NSS_HIP_CF_GRAY, 0, 0, 1, 64, 48, 6);
const nss_hip_video_info kYuv = video(NSS_HIP_CF_YUV, 1, 1, 3, 96, 64, 1);

// Argument list builder: keeps the value arrays alive next to the entries.
struct Args {
    std::vector<nss_hip_arg> a;
    std::vector<std::vector<int64_t>> ints;
    std::vector<std::vector<double>> floats;
    std::vector<std::string> texts;
    Args& i(const char* k, std::vector<int64_t> v) {
        ints.push_back(std::move(v));
        return push(k, 'i');
    }
    Args& f(const char* k, std::vector<double> v) {
        floats.push_back(std::move(v));
        return push(k, 'f');
    }
    Args& s(const char* k, const char* v) {
        texts.emplace_back(v);
        return push(k, 's');
    }
    Args& push(const char* k, char t) {
        nss_hip_arg x{};
        x.key = k;
        x.type = t;
        a.push_back(x);
        order.push_back(t);
        return *this;
    }
    // pointers resolved once the vectors stop growing
    const nss_hip_arg* data() {
        size_t ii = 0, fi = 0, si = 0;
        for (size_t k = 0; k < a.size(); ++k) {
            if (order[k] == 'i') {
                a[k].ints = ints[ii].data();
                a[k].count = static_cast<int>(ints[ii++].size());
            } else if (order[k] == 'f') {
                a[k].floats = floats[fi].data();
                a[k].count = static_cast<int>(floats[fi++].size());
            } else {
                a[k].data = texts[si].c_str();
                a[k].count = 1;
                ++si;
            }
        }
        return a.data();
    }
    int n() const { return static_cast<int>(a.size()); }
    std::vector<char> order;
};

int parse(const nss_hip_video_info& vi, Args& args, nss_hip_bm3d_params* p) {
    return nss_hip_bm3d_parse(&vi, nullptr, args.data(), args.n(), "nss_hip", p);
}

void argument_paths() {
    nss_hip_bm3d_params p{};
    {
        Args a;
        a.f("sigma", {10.0}).i("final", {1});
        check(parse(kYuv, a, &p), "parse defaults");
    }
    {
        Args a;
        a.f("sigma", {10.0, 5.0, 5.0}).i("block_size", {5});
        check(parse(kYuv, a, &p), "parse block_size 5", NSS_HIP_INVALID_INPUT);
        expect(std::strlen(nss_hip_last_error()) > 0, "error text for block_size 5");
    }
    {
        Args a;
        a.f("sigma", {-1.0});
        check(parse(kGray, a, &p), "parse negative sigma", NSS_HIP_INVALID_INPUT);
    }
    {
        Args a;
        a.f("sigma", {10.0}).i("radius", {1}).s("temporal_mode", "legacy");
        if (check(parse(kGray, a, &p), "parse legacy")) {
            nss_hip_video_info out{};
            check(nss_hip_bm3d_output_info(&p, &kGray, &out), "output info legacy");
            expect(out.height == kGray.height * 2 * 3, "legacy output height");
        }
    }
    {
        Args a;
        a.f("sigma", {10.0}).s("temporal_mode", "sideways");
        check(parse(kGray, a, &p), "parse bad temporal_mode", NSS_HIP_INVALID_INPUT);
    }
    // texts from many threads at once (thread-local error messages)
    std::vector<std::thread> ts;
    for (int t = 0; t < 8; ++t)
        ts.emplace_back([t] {
            for (int k = 0; k < 200; ++k) {
                nss_hip_bm3d_params q{};
                Args a;
                a.f("sigma", {static_cast<double>(k % 7 + 1)}).i("group_size", {(k + t) % 3 == 0 ? 0 : 8});
                const int rc = parse(kGray, a, &q);
                expect(rc == NSS_HIP_OK || std::strlen(nss_hip_last_error()) > 0, "threaded parse");
            }
        });
    for (auto& t : ts) t.join();
    nss_hip_server* srv = nullptr;
    nss_hip_server_config cfg{};
    cfg.max_batch_frames = 0;
    cfg.max_inflight = 2;
    check(nss_hip_server_create(&cfg, &srv), "server with no batch frames", NSS_HIP_INVALID_INPUT);
}

struct Planes {
    std::vector<float*> p;
    std::vector<int> stride;
    explicit Planes(const nss_hip_video_info& vi, float fill) {
        for (int i = 0; i < vi.num_planes; ++i) {
            const int w = i ? vi.width >> vi.sub_sampling_w : vi.width;
            const int h = i ? vi.height >> vi.sub_sampling_h : vi.height;
            std::vector<float> host(static_cast<size_t>(w) * h);
            unsigned x = 1234u + static_cast<unsigned>(i) * 77u + static_cast<unsigned>(fill * 1000.f);
            for (auto& v : host) {
                x = x * 1664525u + 1013904223u;
                v = fill + static_cast<float>(x >> 8) / 16777216.f * 0.2f;
            }
            float* d = nullptr;
            if (hipMalloc(&d, host.size() * 4) != hipSuccess || hipMemcpy(d, host.data(), host.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
                expect(false, "plane upload");
            p.push_back(d);
            stride.push_back(w);
        }
    }
    ~Planes() {
        for (float* d : p) (void)hipFree(d);
    }
    std::vector<float> host(int i, const nss_hip_video_info& vi) const {
        const int w = i ? vi.width >> vi.sub_sampling_w : vi.width;
        const int h = i ? vi.height >> vi.sub_sampling_h : vi.height;
        std::vector<float> out(static_cast<size_t>(w) * h);
        (void)hipMemcpy(out.data(), p[i], out.size() * 4, hipMemcpyDeviceToHost);
        return out;
    }
};

void concurrent_process() {
    nss_hip_bm3d_params p{};
    Args a;
    a.f("sigma", {10.0}).i("final", {1}).i("num_streams", {2});
    if (!check(parse(kYuv, a, &p), "parse two-stage")) return;
    nss_hip_bm3d* h = nullptr;
    if (!check(nss_hip_bm3d_create(&p, &kYuv, 0, &h), "create two-stage")) return;
    Planes src(kYuv, 0.4f), want(kYuv, 0.f);
    check(nss_hip_bm3d_process(h, src.p.data(), src.stride.data(), nullptr, nullptr, want.p.data(),
                               want.stride.data(), 0, 1, nullptr),
          "process serial");
    (void)hipDeviceSynchronize();
    std::vector<std::vector<float>> ref;
    for (int i = 0; i < 3; ++i) ref.push_back(want.host(i, kYuv));
    std::vector<std::thread> ts;
    for (int t = 0; t < 4; ++t)
        ts.emplace_back([&] {
            hipStream_t s = nullptr;
            if (hipStreamCreate(&s) != hipSuccess) return expect(false, "stream create");
            Planes out(kYuv, 0.f);
            for (int k = 0; k < 6; ++k) {
                check(nss_hip_bm3d_process(h, src.p.data(), src.stride.data(), nullptr, nullptr, out.p.data(),
                                           out.stride.data(), 0, 1, s),
                      "process threaded");
                (void)hipStreamSynchronize(s);
                for (int i = 0; i < 3; ++i) expect(out.host(i, kYuv) == ref[i], "threaded output equals serial");
            }
            (void)hipStreamDestroy(s);
        });
    for (auto& t : ts) t.join();
    // a refused launch, then the same frame again
    nss_hip_debug_fail_launch(1);
    Planes out(kYuv, 0.f);
    check(nss_hip_bm3d_process(h, src.p.data(), src.stride.data(), nullptr, nullptr, out.p.data(),
                               out.stride.data(), 0, 1, nullptr),
          "process with a refused launch", NSS_HIP_RESOURCE_FAILURE);
    nss_hip_debug_fail_launch(-1);
    check(nss_hip_bm3d_process(h, src.p.data(), src.stride.data(), nullptr, nullptr, out.p.data(),
                               out.stride.data(), 0, 1, nullptr),
          "process after the refusal");
    (void)hipDeviceSynchronize();
    for (int i = 0; i < 3; ++i) expect(out.host(i, kYuv) == ref[i], "output after a refused launch");
    // the no-allocation entry point gives the same bits
    check(nss_hip_bm3d_process_noalloc(&p, &kYuv, 0, src.p.data(), src.stride.data(), nullptr, nullptr,
                                       out.p.data(), out.stride.data(), nullptr),
          "process_noalloc");
    (void)hipDeviceSynchronize();
    for (int i = 0; i < 3; ++i) expect(out.host(i, kYuv) == ref[i], "no-allocation output");
    check(nss_hip_bm3d_destroy(h), "destroy two-stage");
    // every allocation of a creation failing in turn
    long long base = 0, now = 0;
    nss_hip_debug_live_allocations(&base, nullptr);
    for (int k = 0; k < 64; ++k) {
        nss_hip_debug_fail_alloc(k);
        nss_hip_bm3d* x = nullptr;
        const int rc = nss_hip_bm3d_create(&p, &kYuv, 0, &x);
        nss_hip_debug_fail_alloc(-1);
        nss_hip_debug_live_allocations(&now, nullptr);
        if (rc == NSS_HIP_OK) {
            check(nss_hip_bm3d_destroy(x), "destroy after the last allocation");
            break;
        }
        expect(now == base, "allocations released after a failed creation");
    }
}

void rolling_clip() {
    nss_hip_bm3d_params p{};
    Args a;
    a.f("sigma", {10.0}).i("final", {1}).i("radius", {1});
    if (!check(parse(kGray, a, &p), "parse rolling")) return;
    nss_hip_bm3d* h = nullptr;
    if (!check(nss_hip_bm3d_create(&p, &kGray, 0, &h), "create rolling")) return;
    Planes out(kGray, 0.f);
    int got = 0;
    for (int f = 0; f < kGray.num_frames; ++f) {
        Planes src(kGray, 0.3f + 0.01f * static_cast<float>(f));
        check(nss_hip_bm3d_push(h, src.p.data(), src.stride.data(), nullptr, nullptr, nullptr), "push");
        int frame = -1;
        while (nss_hip_bm3d_pop(h, out.p.data(), out.stride.data(), &frame, nullptr) == NSS_HIP_OK) {
            expect(frame == got, "frames pop in order");
            ++got;
        }
        (void)hipDeviceSynchronize();
    }
    expect(got == kGray.num_frames, "every frame popped");
    check(nss_hip_bm3d_destroy(h), "destroy rolling");
}

void serving_loop() {
    nss_hip_server_config cfg{};
    cfg.device_id = 0;
    cfg.max_batch_frames = 4;
    cfg.max_inflight = 2;
    cfg.use_graphs = 1;
    nss_hip_server* srv = nullptr;
    if (!check(nss_hip_server_create(&cfg, &srv), "server create")) return;
    nss_hip_bm3d_params p0{}, p1{};
    Args a0, a1;
    a0.f("sigma", {10.0}).i("final", {1});
    a1.f("sigma", {10.0}).i("radius", {1});
    int s0 = -1, s1 = -1;
    if (check(parse(kYuv, a0, &p0), "parse stream 0")) check(nss_hip_server_open_stream(srv, &p0, &kYuv, &s0), "open 0");
    if (check(parse(kGray, a1, &p1), "parse stream 1")) check(nss_hip_server_open_stream(srv, &p1, &kGray, &s1), "open 1");
    Planes yuv(kYuv, 0.5f);
    std::vector<std::thread> ts;
    for (int t = 0; t < 3; ++t)
        ts.emplace_back([&, t] {
            Planes out(kYuv, 0.f);
            for (int k = 0; k < 8; ++k) {
                long long id = -1;
                if (!check(nss_hip_server_submit(srv, s0, yuv.p.data(), yuv.stride.data(), out.p.data(),
                                                 out.stride.data(), t == 1 ? 7.0 : 0.0, nullptr, &id),
                           "submit"))
                    return;
                double ms = 0;
                int rc;
                while ((rc = nss_hip_server_query(srv, id, &ms)) == NSS_HIP_BYPASS) std::this_thread::yield();
                check(rc, "query");
            }
        });
    // the rolling stream's frames in order from this thread
    Planes gray_out(kGray, 0.f);
    std::vector<Planes*> keep;
    for (int f = 0; f < kGray.num_frames; ++f) {
        keep.push_back(new Planes(kGray, 0.3f));
        long long id = -1;
        check(nss_hip_server_submit(srv, s1, keep.back()->p.data(), keep.back()->stride.data(), gray_out.p.data(),
                                    gray_out.stride.data(), 0.0, nullptr, &id),
              "submit rolling");
    }
    for (auto& t : ts) t.join();
    check(nss_hip_server_drain(srv), "drain");
    nss_hip_server_stats st{};
    check(nss_hip_server_get_stats(srv, &st), "stats");
    expect(st.completed == st.submitted, "every request completed");
    check(nss_hip_server_destroy(srv), "server destroy");
    for (Planes* x : keep) delete x;
}

void wnnm() {
    nss_hip_wnnm_params p{};
    Args a;
    a.f("sigma", {10.0});
    if (!check(nss_hip_wnnm_parse(&kGray, nullptr, a.data(), a.n(), "nss_hip", &p), "wnnm parse")) return;
    nss_hip_wnnm* h = nullptr;
    if (!check(nss_hip_wnnm_create(&p, &kGray, 0, &h), "wnnm create")) return;
    Planes src(kGray, 0.4f), out(kGray, 0.f);
    check(nss_hip_wnnm_process(h, src.p.data(), src.stride.data(), nullptr, nullptr, out.p.data(), out.stride.data(),
                               nullptr),
          "wnnm process");
    (void)hipDeviceSynchronize();
    check(nss_hip_wnnm_destroy(h), "wnnm destroy");
}

}  // namespace

int main() {
    argument_paths();
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) {
        nss_hip_bm3d_params p{};
        Args a;
        a.f("sigma", {10.0});
        nss_hip_bm3d* h = nullptr;
        if (check(parse(kGray, a, &p), "parse without a device"))
            expect(nss_hip_bm3d_create(&p, &kGray, 0, &h) != NSS_HIP_OK, "create fails without a device");
        std::printf("no device: argument paths only; %d failure(s)\n", g_failures.load());
        return g_failures.load() ? 1 : 0;
    }
    concurrent_process();
    rolling_clip();
    serving_loop();
    wnnm();
    long long live = 0;
    nss_hip_debug_live_allocations(&live, nullptr);
    expect(live == 0, "no device allocation outlives its instance");
    std::printf("device paths done; %d failure(s)\n", g_failures.load());
    return g_failures.load() ? 1 : 0;
}
