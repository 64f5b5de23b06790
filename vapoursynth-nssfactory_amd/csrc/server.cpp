// Serving loop behind nss_hip_server_* (no counterpart in the reference,
// whose filters answer one VapourSynth frame request at a time).
//
// Clients queue frames of several clip streams from any thread. A scheduler
// thread, per iteration:
//   retire  batches whose completion event has passed: their requests are
//           done and their frame slots free again;
//   admit   queued radius-0 frames in arrival order into a free frame slot of
//           their stream, allocating a new slot while the HBM budget allows
//           (idle extra slots of other streams are released first); a frame
//           that does not fit waits, later frames of other streams may pass;
//   launch  the admitted frames as one batch (run_batch: per stage round one
//           launch sequence over the planes of every frame, sizes and sigmas
//           mixed) on one of max_inflight HIP streams, so a batch starts while
//           the previous ones still run; from the second batch of one shape
//           (the same instances and filtered planes, in order) on, the batch
//           is captured as a hipGraph whose executable form is instantiated
//           once per shape and updated with each batch's buffers and sigmas.
// Radius > 0 streams keep their rolling state in their own instance and run
// push / pop on a HIP stream of their own, in frame order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "instance.hpp"

using namespace nsship;
using Clock = std::chrono::steady_clock;

namespace {

constexpr int kMaxBatch = 16;
constexpr int kMaxInflight = 4;
constexpr size_t kGraphCacheLimit = 64;

enum ReqState { kQueued, kLaunched, kDone, kFailed };

// A frame slot: one radius-0 frame's device buffers. Slots live behind
// unique_ptr, so a request keeps a stable pointer to its slot while idle
// slots of the same or another stream are released around it.
struct Slot {
    FrameBuffers buf;
    bool own = false;  // the instance's own buffers (never released)
    bool busy = false;
};

struct Request {
    long long id = 0;
    int stream = 0;
    const float* src[3]{};
    int src_stride[3]{};
    float* dst[3]{};
    int dst_stride[3]{};
    float sigma_user = 0.f;
    int frame = 0;  // radius > 0: index in the clip
    Slot* slot = nullptr;
    hipEvent_t ready = nullptr;  // recorded on the client's stream at submit
    ReqState state = kQueued;
    Clock::time_point t_submit, t_done;
};

struct ClipStream {
    nss_hip_bm3d* h = nullptr;
    int radius = 0;
    size_t slot_bytes = 0;
    std::vector<std::unique_ptr<Slot>> slots;
    hipStream_t hs = nullptr;                  // radius > 0
    int submitted = 0;                          // radius > 0: frames queued so far
    std::map<int, long long> frame_request;    // radius > 0: frame -> request
};

struct Batch {
    hipEvent_t done = nullptr;
    std::vector<long long> reqs;
};

struct GraphEntry {
    hipGraphExec_t exec = nullptr;
    unsigned long long last_use = 0;
};

float effective_sigma(float user) { return (user * (1.f / 255.f)) * 0.75f; }

}  // namespace

struct nss_hip_server {
    nss_hip_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work;  // new requests or stop
    std::condition_variable cv_done;  // completions (drain)
    std::thread worker;
    bool stop = false;
    std::string worker_error;

    std::vector<ClipStream> streams;
    std::unordered_map<long long, Request> requests;
    std::deque<long long> queue;  // arrival order
    std::vector<Batch> inflight;
    std::vector<hipStream_t> hs;  // radius-0 batches, round robin
    int next_hs = 0;
    std::vector<hipEvent_t> event_pool;
    std::unordered_map<std::string, GraphEntry> graphs;
    std::unordered_set<std::string> seen_once;
    unsigned long long use_clock = 0;
    long long next_id = 1;
    nss_hip_server_stats st{};

    hipEvent_t take_event() {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
        return e;
    }

    size_t bytes_in_use() const {
        size_t b = 0;
        for (const auto& c : streams) {
            if (c.radius > 0) {
                b += c.slot_bytes;
                continue;
            }
            b += c.slots.size() * c.slot_bytes;
        }
        return b;
    }

    // --- scheduler (called with mu held) ---------------------------------

    void retire() {
        for (size_t k = 0; k < inflight.size();) {
            const hipError_t q = hipEventQuery(inflight[k].done);
            if (q == hipErrorNotReady) {
                ++k;
                continue;
            }
            const auto now = Clock::now();
            for (long long id : inflight[k].reqs) {
                Request& r = requests[id];
                r.state = q == hipSuccess ? kDone : kFailed;
                r.t_done = now;
                if (r.slot) r.slot->busy = false;
                ++st.completed;
            }
            event_pool.push_back(inflight[k].done);
            inflight.erase(inflight.begin() + static_cast<long>(k));
            cv_done.notify_all();
        }
    }

    // A free slot of stream c, allocating one within the budget (releasing
    // idle extra slots of other streams when needed); null if none.
    Slot* free_slot(int c) {
        ClipStream& cs = streams[c];
        for (auto& sl : cs.slots)
            if (!sl->busy) return sl.get();
        const size_t budget = static_cast<size_t>(cfg.hbm_budget_bytes);
        if (bytes_in_use() + cs.slot_bytes > budget) {
            for (size_t o = 0; o < streams.size() && bytes_in_use() + cs.slot_bytes > budget; ++o) {
                if (o == static_cast<size_t>(c) || streams[o].radius > 0) continue;
                auto& sl = streams[o].slots;
                for (size_t i = sl.size(); i-- > 0;) {
                    if (sl[i]->own || sl[i]->busy) continue;
                    free_frame_buffers(sl[i]->buf);
                    sl.erase(sl.begin() + static_cast<long>(i));
                    ++st.slots_released;
                    if (bytes_in_use() + cs.slot_bytes <= budget) break;
                }
            }
            if (bytes_in_use() + cs.slot_bytes > budget) return nullptr;
        }
        auto s = std::make_unique<Slot>();
        if (alloc_frame_buffers(cs.h, s->buf) != hipSuccess) return nullptr;
        cs.slots.push_back(std::move(s));
        st.bytes_in_use = static_cast<long long>(bytes_in_use());
        st.peak_bytes = std::max(st.peak_bytes, st.bytes_in_use);
        return cs.slots.back().get();
    }

    // Orders stream s after the client work that produced request r's
    // planes, then returns r's readiness event to the pool (a wait takes
    // the event's state at the time of the call).
    bool wait_ready(Request& r, hipStream_t s) {
        if (!r.ready) return true;
        const hipError_t e = hipStreamWaitEvent(s, r.ready, 0);
        event_pool.push_back(r.ready);
        r.ready = nullptr;
        return e == hipSuccess;
    }

    bool launch_temporal(long long id) {
        Request& r = requests[id];
        ClipStream& cs = streams[r.stream];
        if (!wait_ready(r, cs.hs) || nss_hip_bm3d_push(cs.h, r.src, r.src_stride, nullptr, nullptr, cs.hs) != NSS_HIP_OK) {
            r.state = kFailed;
            r.t_done = Clock::now();
            ++st.completed;
            cv_done.notify_all();
            return false;
        }
        r.state = kLaunched;
        Batch b;
        for (;;) {
            const int n = cs.h->roll.out_head;
            if (n >= cs.h->roll.finished[cs.h->nstages - 1]) break;
            const auto it = cs.frame_request.find(n);
            if (it == cs.frame_request.end()) break;
            Request& o = requests[it->second];
            int frame = -1;
            if (nss_hip_bm3d_pop(cs.h, o.dst, o.dst_stride, &frame, cs.hs) != NSS_HIP_OK) break;
            b.reqs.push_back(it->second);
            cs.frame_request.erase(it);
        }
        if (!b.reqs.empty()) {
            b.done = take_event();
            if (!b.done || hipEventRecord(b.done, cs.hs) != hipSuccess) {
                worker_error = "nss_hip: server event record failed";
                return false;
            }
            inflight.push_back(std::move(b));
        }
        return true;
    }

    // A batch's shape: what its launch sequence depends on (the instances of
    // its frames, in order, and which planes of each stage filter). Buffers,
    // strides and sigma values are kernel arguments only.
    static std::string graph_key(const std::vector<FrameRef>& frames) {
        std::string key;
        for (const FrameRef& f : frames) {
            key.append(reinterpret_cast<const char*>(&f.h), sizeof(f.h));
            char flags = 0;
            for (int i = 0; i < 3; ++i) {
                flags = static_cast<char>(flags | (f.sigma[0][i] != 0.f ? 1 << i : 0) |
                                          (f.sigma[1][i] != 0.f ? 8 << i : 0));
            }
            key.push_back(flags);
            key.push_back(f.ref[0] || f.ref[1] || f.ref[2] ? 1 : 0);
        }
        return key;
    }

    // A shape runs direct the first time it is seen. From the second time on
    // the batch is captured into a graph: the first capture is instantiated,
    // later ones hand their arguments (buffers, strides, sigmas) to that
    // executable graph through hipGraphExecUpdate and replay it; an update
    // the runtime refuses (a different launch sequence) re-instantiates.
    int run_graph(const std::vector<FrameRef>& frames, hipStream_t s) {
        std::string key = graph_key(frames);
        auto it = graphs.find(key);
        if (it == graphs.end()) {
            auto seen = seen_once.find(key);
            if (seen == seen_once.end()) {
                if (seen_once.size() >= 4 * kGraphCacheLimit) seen_once.clear();
                seen_once.insert(key);
                return run_batch(frames.data(), static_cast<int>(frames.size()), nullptr, s);
            }
        }
        hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed);
        if (e != hipSuccess) return hip_fail(e, "graph capture");
        const int rc = run_batch(frames.data(), static_cast<int>(frames.size()), nullptr, s);
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(s, &g);
        if (rc != NSS_HIP_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(e, "graph capture");
        bool fresh = it == graphs.end();
        if (!fresh) {
            hipGraphNode_t bad = nullptr;
            hipGraphExecUpdateResult res = hipGraphExecUpdateError;
            e = hipGraphExecUpdate(it->second.exec, g, &bad, &res);
            if (e != hipSuccess || res != hipGraphExecUpdateSuccess) {
                (void)hipGetLastError();  // the refusal is not a failure of the batch
                (void)hipGraphExecDestroy(it->second.exec);
                graphs.erase(it);
                it = graphs.end();
                fresh = true;
            }
        }
        if (fresh) {
            seen_once.erase(key);
            if (graphs.size() >= kGraphCacheLimit) {
                auto lru = graphs.begin();
                for (auto q = graphs.begin(); q != graphs.end(); ++q)
                    if (q->second.last_use < lru->second.last_use) lru = q;
                (void)hipGraphExecDestroy(lru->second.exec);
                graphs.erase(lru);
            }
            GraphEntry ge;
            e = hipGraphInstantiate(&ge.exec, g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(g);
                return hip_fail(e, "graph instantiate");
            }
            it = graphs.emplace(std::move(key), ge).first;
            ++st.graph_captures;
        } else {
            ++st.graph_replays;
        }
        (void)hipGraphDestroy(g);
        it->second.last_use = ++use_clock;
        e = hipGraphLaunch(it->second.exec, s);
        return e == hipSuccess ? NSS_HIP_OK : hip_fail(e, "graph launch");
    }

    // Radius > 0: every queued frame, in order, on its stream's HIP stream.
    bool schedule_temporal() {
        bool any = false;
        for (size_t q = 0; q < queue.size();) {
            Request& r = requests[queue[q]];
            if (streams[r.stream].radius == 0) {
                ++q;
                continue;
            }
            const long long id = queue[q];
            queue.erase(queue.begin() + static_cast<long>(q));
            if (!launch_temporal(id)) return any;
            any = true;
        }
        return any;
    }

    // Radius 0: admits queued frames in arrival order (up to max_batch_frames)
    // and launches them as one batch; true if a batch was launched.
    bool schedule_batch() {
        std::vector<FrameRef> frames;
        std::vector<long long> ids;
        for (size_t q = 0; q < queue.size() && static_cast<int>(frames.size()) < cfg.max_batch_frames;) {
            const long long id = queue[q];
            Request& r = requests[id];
            Slot* slot = free_slot(r.stream);
            if (!slot) {
                ++st.admissions_deferred;
                ++q;
                continue;
            }
            ClipStream& cs = streams[r.stream];
            slot->busy = true;
            r.slot = slot;
            FrameRef f = frame_ref(cs.h, r.src, r.src_stride, nullptr, nullptr, r.dst, r.dst_stride);
            f.buf = slot->buf;
            if (r.sigma_user > 0.f) {
                const float eff = effective_sigma(r.sigma_user);
                for (int k = 0; k < cs.h->nstages; ++k)
                    for (int i = 0; i < cs.h->nplanes; ++i)
                        if (f.sigma[k][i] != 0.f) f.sigma[k][i] = eff;
            }
            frames.push_back(f);
            ids.push_back(id);
            queue.erase(queue.begin() + static_cast<long>(q));
        }
        if (frames.empty()) return false;
        hipStream_t s = hs[static_cast<size_t>(next_hs)];
        next_hs = (next_hs + 1) % static_cast<int>(hs.size());
        bool ready = true;
        for (long long id : ids) ready &= wait_ready(requests[id], s);
        const int rc = !ready ? fail(NSS_HIP_RESOURCE_FAILURE, "nss_hip: server stream wait failed")
                     : cfg.use_graphs ? run_graph(frames, s)
                                      : run_batch(frames.data(), static_cast<int>(frames.size()), nullptr, s);
        Batch b;
        b.reqs = ids;
        b.done = take_event();
        if (rc != NSS_HIP_OK || !b.done || hipEventRecord(b.done, s) != hipSuccess) {
            worker_error = rc != NSS_HIP_OK ? g_last_error : "nss_hip: server event record failed";
            const auto now = Clock::now();
            for (long long id : ids) {
                Request& r = requests[id];
                r.state = kFailed;
                r.t_done = now;
                r.slot->busy = false;
                ++st.completed;
            }
            if (b.done) event_pool.push_back(b.done);
            cv_done.notify_all();
            return false;
        }
        for (long long id : ids) requests[id].state = kLaunched;
        ++st.batches;
        st.batched_frames += static_cast<long long>(ids.size());
        inflight.push_back(std::move(b));
        return true;
    }

    int radius0_inflight() const {
        int n = 0;
        for (const auto& b : inflight)
            if (!b.reqs.empty() && streams[requests.at(b.reqs[0]).stream].radius == 0) ++n;
        return n;
    }

    void loop() {
        (void)hipSetDevice(cfg.device_id);
        std::unique_lock<std::mutex> lock(mu);
        while (!stop) {
            retire();
            bool launched = schedule_temporal();
            if (radius0_inflight() < cfg.max_inflight) launched |= schedule_batch();
            if (launched) continue;
            if (inflight.empty()) {
                cv_work.wait(lock, [&] { return stop || !queue.empty(); });
            } else {
                // Completions are polled: a short sleep without the lock.
                lock.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(20));
                lock.lock();
            }
        }
    }
};

extern "C" {

int nss_hip_server_create(const nss_hip_server_config* config, nss_hip_server** out) {
    if (!config || !out) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: null argument");
    *out = nullptr;
    nss_hip_server_config c = *config;
    if (c.max_batch_frames < 1 || c.max_batch_frames > kMaxBatch)
        return fail(NSS_HIP_INVALID_INPUT, "nss_hip: max_batch_frames must be in [1, 16]");
    if (c.max_inflight < 1 || c.max_inflight > kMaxInflight)
        return fail(NSS_HIP_INVALID_INPUT, "nss_hip: max_inflight must be in [1, 4]");
    if (c.hbm_budget_bytes < 0) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: hbm_budget_bytes must be >= 0");
    DeviceGuard guard(c.device_id);
    if (guard.err != hipSuccess) return hip_fail(guard.err, "set device");
    int rc = ensure_tables(c.device_id);
    if (rc != NSS_HIP_OK) return rc;
    if (c.hbm_budget_bytes == 0) {
        size_t free_b = 0, total_b = 0;
        const hipError_t e = hipMemGetInfo(&free_b, &total_b);
        if (e != hipSuccess) return hip_fail(e, "device memory info");
        c.hbm_budget_bytes = static_cast<long long>(total_b / 4);
    }
    auto* s = new (std::nothrow) nss_hip_server();
    if (!s) return fail(NSS_HIP_RESOURCE_FAILURE, "nss_hip: out of host memory");
    s->cfg = c;
    for (int i = 0; i < c.max_inflight; ++i) {
        hipStream_t h = nullptr;
        const hipError_t e = hipStreamCreateWithFlags(&h, hipStreamNonBlocking);
        if (e != hipSuccess) {
            for (hipStream_t x : s->hs) (void)hipStreamDestroy(x);
            delete s;
            return hip_fail(e, "stream create");
        }
        s->hs.push_back(h);
    }
    s->worker = std::thread([s] { s->loop(); });
    *out = s;
    return NSS_HIP_OK;
}

int nss_hip_server_open_stream(nss_hip_server* s, const nss_hip_bm3d_params* params, const nss_hip_video_info* info,
                               int* stream_id) {
    if (!s || !params || !info || !stream_id) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: null argument");
    if (params->device_id != s->cfg.device_id)
        return fail(NSS_HIP_INVALID_INPUT, "nss_hip: stream device_id differs from the server's");
    if (params->has_ref) return fail(NSS_HIP_UNSUPPORTED, "nss_hip: server streams take no ref clip");
    if (params->radius > 0 && !params->temporal_mode)
        return fail(NSS_HIP_UNSUPPORTED, "nss_hip: server streams return finished frames; use temporal_mode=rolling");
    DeviceGuard guard(s->cfg.device_id);
    if (guard.err != hipSuccess) return hip_fail(guard.err, "set device");
    std::lock_guard<std::mutex> lock(s->mu);
    ClipStream cs;
    cs.radius = params->radius;
    nss_hip_bm3d* h = nullptr;
    // The scheduler gives each frame a slot of its own; the instance keeps
    // one (the stream's first slot).
    nss_hip_bm3d_params one = *params;
    one.num_streams = 1;
    int rc = nss_hip_bm3d_create(&one, info, s->cfg.device_id, &h);
    if (rc != NSS_HIP_OK) return rc;
    cs.h = h;
    cs.slot_bytes = cs.radius > 0 ? instance_bytes(h) : frame_buffer_bytes(h);
    if (s->bytes_in_use() + cs.slot_bytes > static_cast<size_t>(s->cfg.hbm_budget_bytes)) {
        nss_hip_bm3d_destroy(h);
        return fail(NSS_HIP_RESOURCE_FAILURE, "nss_hip: the stream's device state exceeds the server's HBM budget");
    }
    if (cs.radius > 0) {
        const hipError_t e = hipStreamCreateWithFlags(&cs.hs, hipStreamNonBlocking);
        if (e != hipSuccess) {
            nss_hip_bm3d_destroy(h);
            return hip_fail(e, "stream create");
        }
    } else {
        if (s->cfg.autotune) {
            hipStream_t ts = nullptr;
            hipError_t e = hipStreamCreateWithFlags(&ts, hipStreamNonBlocking);
            rc = e == hipSuccess ? nss_hip_bm3d_autotune(h, 3, ts, nullptr) : hip_fail(e, "stream create");
            if (ts) (void)hipStreamDestroy(ts);
            if (rc != NSS_HIP_OK) {
                nss_hip_bm3d_destroy(h);
                return rc;
            }
            cs.slot_bytes = frame_buffer_bytes(h);
        }
        auto own = std::make_unique<Slot>();
        own->buf = h->slots[0].buf;
        own->own = true;
        cs.slots.push_back(std::move(own));
    }
    s->streams.push_back(std::move(cs));
    s->st.bytes_in_use = static_cast<long long>(s->bytes_in_use());
    s->st.peak_bytes = std::max(s->st.peak_bytes, s->st.bytes_in_use);
    *stream_id = static_cast<int>(s->streams.size()) - 1;
    return NSS_HIP_OK;
}

int nss_hip_server_submit(nss_hip_server* s, int stream_id, const float* const* src, const int* src_stride,
                          float* const* dst, const int* dst_stride, double sigma_user, void* ready_stream,
                          long long* request_id) {
    if (!s || !src || !src_stride || !dst || !dst_stride || !request_id)
        return fail(NSS_HIP_INVALID_INPUT, "nss_hip: null argument");
    if (!(sigma_user >= 0.0) || sigma_user > 1e30) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: invalid sigma_user");
    std::lock_guard<std::mutex> lock(s->mu);
    if (!s->worker_error.empty()) return fail(NSS_HIP_RESOURCE_FAILURE, s->worker_error);
    if (stream_id < 0 || stream_id >= static_cast<int>(s->streams.size()))
        return fail(NSS_HIP_INVALID_INPUT, "nss_hip: unknown stream");
    ClipStream& cs = s->streams[stream_id];
    const nss_hip_bm3d* h = cs.h;
    for (int i = 0; i < h->nplanes; ++i)
        if (!src[i] || !dst[i] || src_stride[i] < h->pw[i] || dst_stride[i] < h->pw[i])
            return fail(NSS_HIP_INVALID_INPUT, "nss_hip: missing plane or stride below width");
    if (cs.radius > 0 && cs.submitted >= h->roll.N)
        return fail(NSS_HIP_INVALID_INPUT, "nss_hip: every frame of the clip was already submitted");
    Request r;
    r.id = s->next_id++;
    r.stream = stream_id;
    for (int i = 0; i < h->nplanes; ++i) {
        r.src[i] = src[i];
        r.src_stride[i] = src_stride[i];
        r.dst[i] = dst[i];
        r.dst_stride[i] = dst_stride[i];
    }
    r.sigma_user = static_cast<float>(sigma_user);
    if (ready_stream) {
        DeviceGuard guard(s->cfg.device_id);
        r.ready = s->take_event();
        if (!r.ready || guard.err != hipSuccess ||
            hipEventRecord(r.ready, static_cast<hipStream_t>(ready_stream)) != hipSuccess) {
            if (r.ready) s->event_pool.push_back(r.ready);
            return fail(NSS_HIP_RESOURCE_FAILURE, "nss_hip: readiness event record failed");
        }
    }
    r.t_submit = Clock::now();
    if (cs.radius > 0) {
        r.frame = cs.submitted++;
        cs.frame_request[r.frame] = r.id;
    }
    s->requests.emplace(r.id, r);
    s->queue.push_back(r.id);
    ++s->st.submitted;
    *request_id = r.id;
    s->cv_work.notify_one();
    return NSS_HIP_OK;
}

int nss_hip_server_query(nss_hip_server* s, long long request_id, double* latency_ms) {
    if (!s) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: null argument");
    std::lock_guard<std::mutex> lock(s->mu);
    const auto it = s->requests.find(request_id);
    if (it == s->requests.end()) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: unknown request");
    const Request& r = it->second;
    if (r.state == kFailed)
        return fail(NSS_HIP_RESOURCE_FAILURE,
                    s->worker_error.empty() ? std::string("nss_hip: request failed") : s->worker_error);
    if (r.state != kDone) return NSS_HIP_BYPASS;
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r.t_done - r.t_submit).count();
    return NSS_HIP_OK;
}

int nss_hip_server_drain(nss_hip_server* s) {
    if (!s) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: null argument");
    std::unique_lock<std::mutex> lock(s->mu);
    s->cv_done.wait(lock, [&] { return s->st.completed >= s->st.submitted || !s->worker_error.empty(); });
    if (!s->worker_error.empty()) return fail(NSS_HIP_RESOURCE_FAILURE, s->worker_error);
    return NSS_HIP_OK;
}

int nss_hip_server_get_stats(nss_hip_server* s, nss_hip_server_stats* out) {
    if (!s || !out) return fail(NSS_HIP_INVALID_INPUT, "nss_hip: null argument");
    std::lock_guard<std::mutex> lock(s->mu);
    s->st.bytes_in_use = static_cast<long long>(s->bytes_in_use());
    *out = s->st;
    return NSS_HIP_OK;
}

int nss_hip_server_destroy(nss_hip_server* s) {
    if (!s) return NSS_HIP_OK;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    DeviceGuard guard(s->cfg.device_id);
    for (hipStream_t h : s->hs) {
        (void)hipStreamSynchronize(h);
        (void)hipStreamDestroy(h);
    }
    for (auto& c : s->streams) {
        if (c.hs) {
            (void)hipStreamSynchronize(c.hs);
            (void)hipStreamDestroy(c.hs);
        }
        for (auto& sl : c.slots)
            if (!sl->own) free_frame_buffers(sl->buf);
        nss_hip_bm3d_destroy(c.h);
    }
    for (auto& g : s->graphs) (void)hipGraphExecDestroy(g.second.exec);
    for (auto& b : s->inflight) s->event_pool.push_back(b.done);
    for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
    delete s;
    return NSS_HIP_OK;
}

}  // extern "C"
