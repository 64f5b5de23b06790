"""Serving loop (nss_hip_server_*): frames of several clip streams queued at
once are packed into shared launches; each result must be bitwise the one
the stream's own filter instance gives for that frame alone (same kernels,
same per-frame arithmetic), including per-request sigma, radius-1 streams,
graph replays and admission under a tight HBM budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def _alone(gpu, fmt, w, h, planes, sigma, **kw):
    import torch
    inst = gpu.BM3D(fmt, w, h, sigma=sigma, **kw)
    out = inst([_dev(p) for p in planes])
    torch.cuda.synchronize()
    res = [o.cpu().numpy() for o in out]
    inst.close()
    return res


STREAMS = [("YUV420PS", 96, 64, dict(final=1)), ("YUV420PS", 128, 96, dict(final=1)),
           ("RGBS", 64, 48, dict(final=1)), ("GRAYS", 80, 56, dict(final=0, group_size=16, block_step=4))]


def _frames(fmt, w, h, n, seed):
    import synth
    return [synth.noisy_frame(fmt, w, h, 10, seed=seed + k)[1] for k in range(n)]


@pytest.mark.parametrize("use_graphs", [False, True])
def test_mix_equals_alone(gpu, use_graphs):
    import torch
    # One batch in flight at a time, requests submitted back to back: the
    # ones arriving while a batch runs are packed into the next.
    srv = gpu.Server(max_batch_frames=6, max_inflight=1, use_graphs=use_graphs)
    sids = [srv.open_stream(fmt, w, h, sigma=10, **kw) for fmt, w, h, kw in STREAMS]
    rng = np.random.default_rng(5)
    work = []
    for rep in range(3):
        for s, (fmt, w, h, kw) in enumerate(STREAMS):
            for planes in _frames(fmt, w, h, 2, seed=10 * s + rep):
                sig = float(rng.choice([3, 5, 10, 20]))
                src = [_dev(p) for p in planes]
                out = [torch.empty_like(p) for p in src]
                work.append((s, planes, sig, src, out))
    torch.cuda.synchronize()
    reqs = [(srv.submit(sids[s], src, out, sigma=sig), s, planes, sig, src, out) for s, planes, sig, src, out in work]
    srv.drain()
    st = srv.stats()
    print(f"server stats {st}", flush=True)
    assert st["completed"] == len(reqs) and st["batches"] < len(reqs)
    for rid, s, planes, sig, src, out in reqs:
        assert srv.query(rid) is not None
        fmt, w, h, kw = STREAMS[s]
        want = _alone(gpu, fmt, w, h, planes, sig, **kw)
        for g, w_ in zip(out, want):
            assert np.array_equal(g.cpu().numpy(), w_), (s, sig)
    srv.close()


def test_full_size_streams_mix_equals_alone(gpu):
    """The serving workload's three streams at their full sizes (720x480
    YUV420PS at radius 1, 1920x1080 YUV420PS, 3840x2160 RGBS; two stages,
    per-request sigma from {3, 5, 10, 20}), frames queued at once so the
    radius-0 frames share batches; every result is bitwise the one its
    stream's own instance gives alone."""
    import torch
    import synth
    rng = np.random.default_rng(8)
    srv = gpu.Server(max_batch_frames=4, max_inflight=2)
    n_t = 4
    _, clip = synth.shifted_sequence("YUV420PS", 720, 480, n_t, 5, seed=1, max_step=4)
    tid = srv.open_stream("YUV420PS", 720, 480, n_t, sigma=5, radius=1, final=1)
    hd = srv.open_stream("YUV420PS", 1920, 1080, sigma=10, final=1)
    uhd = srv.open_stream("RGBS", 3840, 2160, sigma=10, final=1)
    work = []
    for n in range(n_t):
        src = [_dev(p) for p in clip[n]]
        work.append(("t", n, tid, src, [torch.empty_like(p) for p in src], 0.0))
    # one generated frame per stream (the 4K RGB generator is slow on the
    # host), submitted three times with different sigmas
    frames = {fmt: [_dev(p) for p in synth.noisy_frame(fmt, w, h, 10, seed=50)[1]]
              for fmt, w, h in (("YUV420PS", 1920, 1080), ("RGBS", 3840, 2160))}
    for k in range(3):
        for sid, fmt in ((hd, "YUV420PS"), (uhd, "RGBS")):
            src = frames[fmt]
            sig = float(rng.choice([3, 5, 10, 20]))
            work.append((fmt, k, sid, src, [torch.empty_like(p) for p in src], sig))
    torch.cuda.synchronize()
    rids = [srv.submit(sid, src, out, sigma=sig) for _, _, sid, src, out, sig in work]
    srv.drain()
    st = srv.stats()
    print(f"full-size mix stats {st}", flush=True)
    assert all(srv.query(r) is not None for r in rids)
    inst = gpu.BM3D("YUV420PS", 720, 480, n_t, sigma=5, radius=1, final=1)
    want_t = [[p.cpu().numpy() for p in f] for f in inst.filter_clip([[_dev(p) for p in f] for f in clip])]
    inst.close()
    for kind, n, sid, src, out, sig in work:
        if kind == "t":
            want = want_t[n]
        else:
            h, w = src[0].shape
            inst = gpu.BM3D(kind, w, h, sigma=sig, final=1)
            res = inst(src)
            torch.cuda.synchronize()
            want = [o.cpu().numpy() for o in res]
            inst.close()
        for i, (g, w_) in enumerate(zip(out, want)):
            assert np.array_equal(g.cpu().numpy(), w_), (kind, n, i, sig)
    srv.close()


def test_graph_replays_match_eager(gpu):
    """One batch shape replayed several times with alternating buffers and
    sigmas (each replay updates the executable graph's arguments), with eager work and
    synchronisations between the replays; every replay equals the eager
    (no-graph) result."""
    import torch
    fmt, w, h, kw = STREAMS[0]
    planes = _frames(fmt, w, h, 2, seed=3)
    src = [[_dev(p) for p in f] for f in planes]
    outs = {}
    for use_graphs in (False, True):
        # one frame per batch: the same two compositions every repetition
        srv = gpu.Server(max_batch_frames=1, max_inflight=1, use_graphs=use_graphs)
        sid = srv.open_stream(fmt, w, h, sigma=10, **kw)
        got = []
        out = [[torch.empty_like(p) for p in f] for f in src]
        for rep in range(4):
            for f in out:
                for p in f:
                    p.fill_(-1.0)
            torch.cuda.synchronize()
            rids = [srv.submit(sid, src[k], out[k], sigma=(5.0, 10.0)[(rep + k) % 2]) for k in range(2)]
            srv.drain()
            assert all(srv.query(r) is not None for r in rids)
            got.append([[p.cpu().numpy() for p in f] for f in out])
            x = torch.randn(1 << 20, device="cuda")
            (x * 2).sum().item()
            torch.cuda.synchronize()
        outs[use_graphs] = got
        st = srv.stats()
        print(f"use_graphs={use_graphs} stats {st}", flush=True)
        if use_graphs:
            # the one shape: direct, captured once, then six updated replays
            assert st["graph_captures"] == 1 and st["graph_replays"] == 6
        srv.close()
    for rep in range(4):
        for fe, fg in zip(outs[False][rep], outs[True][rep]):
            for a, b in zip(fe, fg):
                assert np.array_equal(a, b), rep


def test_temporal_stream_in_mix(gpu):
    import torch
    import synth
    _, clip = synth.shifted_sequence("YUV420PS", 64, 48, 5, 10, seed=2, max_step=2)
    inst = gpu.BM3D("YUV420PS", 64, 48, 5, sigma=10, radius=1, final=1)
    want = [[p.cpu().numpy() for p in f] for f in inst.filter_clip([[_dev(p) for p in f] for f in clip])]
    inst.close()
    srv = gpu.Server(max_batch_frames=4)
    tid = srv.open_stream("YUV420PS", 64, 48, 5, sigma=10, radius=1, final=1)
    sid = srv.open_stream("GRAYS", 72, 56, sigma=10)
    reqs, side = [], []
    for n, f in enumerate(clip):
        src = [_dev(p) for p in f]
        out = [torch.empty_like(p) for p in src]
        g = _dev(synth.noisy_frame("GRAYS", 72, 56, 10, seed=n)[1][0])
        go = torch.empty_like(g)
        torch.cuda.synchronize()
        reqs.append((srv.submit(tid, src, out), src, out))
        side.append((srv.submit(sid, [g], [go]), g, go))
    srv.drain()
    for n, (rid, src, out) in enumerate(reqs):
        assert srv.query(rid) is not None
        for g, w_ in zip(out, want[n]):
            assert np.array_equal(g.cpu().numpy(), w_), n
    srv.close()


def test_budget_admission(gpu):
    """A budget that holds one frame slot per stream: frames wait for slots
    (deferred admissions), all complete, the budget is never exceeded."""
    import torch
    probe = gpu.Server()
    a = probe.open_stream("YUV420PS", 128, 96, sigma=10, final=1)
    b = probe.open_stream("GRAYS", 96, 64, sigma=10)
    base = probe.stats()["bytes_in_use"]
    probe.close()
    srv = gpu.Server(hbm_budget_bytes=base, max_batch_frames=8, max_inflight=2)
    a = srv.open_stream("YUV420PS", 128, 96, sigma=10, final=1)
    b = srv.open_stream("GRAYS", 96, 64, sigma=10)
    reqs = []
    for k in range(6):
        for sid, fmt, w, h in ((a, "YUV420PS", 128, 96), (b, "GRAYS", 96, 64)):
            planes = _frames(fmt, w, h, 1, seed=k)[0]
            src = [_dev(p) for p in planes]
            out = [torch.empty_like(p) for p in src]
            torch.cuda.synchronize()
            reqs.append((srv.submit(sid, src, out), fmt, w, h, planes, out))
    srv.drain()
    st = srv.stats()
    print(f"budget stats {st}", flush=True)
    assert st["completed"] == len(reqs) and st["peak_bytes"] <= base
    for rid, fmt, w, h, planes, out in reqs:
        want = _alone(gpu, fmt, w, h, planes, 10, **({"final": 1} if fmt == "YUV420PS" else {}))
        for g, w_ in zip(out, want):
            assert np.array_equal(g.cpu().numpy(), w_)
    srv.close()


def test_autotune_keeps_parity(gpu, oracle):
    """The autotuned plan (fastest tile shape per bucket) still meets the
    oracle gate; radius-0 two-stage YUV420 and a one-stage group-16 gray."""
    import torch
    import synth
    for fmt, w, h, kw in (("YUV420PS", 160, 120, dict(sigma=10, final=1)),
                          ("GRAYS", 128, 96, dict(sigma=5, group_size=16, block_step=3, bm_range=16))):
        _, noisy = synth.noisy_frame(fmt, w, h, 10, seed=2)
        inst = gpu.BM3D(fmt, w, h, **kw)
        ms = inst.autotune(repeats=2)
        out = inst([_dev(p) for p in noisy])
        torch.cuda.synchronize()
        got = [o.cpu().numpy() for o in out]
        inst.close()
        want = oracle.bm3d_frame(noisy, gpu.parse(fmt, w, h, **kw))
        ps = [oracle.psnr(g, w_) for g, w_ in zip(got, want)]
        print(f"autotune {fmt}: {ms:.3f} ms/frame, PSNR {ps}", flush=True)
        assert ms > 0 and min(ps) >= 60.0


def test_budget_evicts_idle_slots_under_busy_ones(gpu):
    """One stream grows to several frame slots; a second stream then needs a
    slot under a full budget and releases the first stream's idle extras
    while others of them are still busy (one frame per batch, four batches in
    flight, so the slots finish in any order). Every output equals the alone
    result and the eviction path ran."""
    import torch
    probe = gpu.Server()
    a = probe.open_stream("YUV420PS", 160, 120, sigma=10, final=1)
    base_a = probe.stats()["bytes_in_use"]
    b = probe.open_stream("GRAYS", 128, 96, sigma=10)
    base = probe.stats()["bytes_in_use"]
    probe.close()
    budget = base + 3 * base_a  # own slots plus three extra slots of A's size
    srv = gpu.Server(hbm_budget_bytes=budget, max_batch_frames=1, max_inflight=4, use_graphs=False)
    a = srv.open_stream("YUV420PS", 160, 120, sigma=10, final=1)
    b = srv.open_stream("GRAYS", 128, 96, sigma=10)
    specs = {a: ("YUV420PS", 160, 120, {"final": 1}), b: ("GRAYS", 128, 96, {})}
    order = [a] * 6 + [b] * 6 + [a, b] * 4 + [a] * 4 + [b] * 4
    work = []
    for k, sid in enumerate(order):
        fmt, w, h, kw = specs[sid]
        planes = _frames(fmt, w, h, 1, seed=100 + k)[0]
        src = [_dev(p) for p in planes]
        out = [torch.empty_like(p) for p in src]
        work.append((sid, planes, src, out))
    torch.cuda.synchronize()
    # Every request waits for a long kernel on a side stream, so the first
    # four batches hold A's four slots while the rest of the queue arrives;
    # then B's second frame needs a slot that only A's idle extras can free.
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        x = torch.randn(1 << 24, device="cuda")
        for _ in range(40):
            x = x * 1.0001
    reqs = [(srv.submit(sid, src, out, ready_stream=side), sid, planes, out) for sid, planes, src, out in work]
    srv.drain()
    st = srv.stats()
    print(f"eviction stats {st}", flush=True)
    assert st["completed"] == len(reqs) and st["peak_bytes"] <= budget
    assert st["slots_released"] > 0
    for rid, sid, planes, out in reqs:
        fmt, w, h, kw = specs[sid]
        want = _alone(gpu, fmt, w, h, planes, 10, **kw)
        for g, w_ in zip(out, want):
            assert np.array_equal(g.cpu().numpy(), w_), sid
    srv.close()


def test_submit_waits_for_the_producing_stream(gpu):
    """Planes written on a side stream and submitted at once, without a host
    synchronisation: the request waits for that stream's work."""
    import torch
    fmt, w, h, kw = STREAMS[1]
    planes = _frames(fmt, w, h, 1, seed=77)[0]
    host = [torch.from_numpy(np.ascontiguousarray(p, np.float32)).pin_memory() for p in planes]
    want = _alone(gpu, fmt, w, h, planes, 10, **kw)
    srv = gpu.Server(max_batch_frames=2, max_inflight=1)
    sid = srv.open_stream(fmt, w, h, sigma=10, **kw)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        src = [torch.empty(p.shape, dtype=torch.float32, device="cuda") for p in host]
        out = [torch.empty_like(p) for p in src]
        # a long kernel first, so the copies land well after the submit
        x = torch.randn(1 << 24, device="cuda")
        for _ in range(20):
            x = x * 1.0001
        for d, p in zip(src, host):
            d.copy_(p, non_blocking=True)
        rid = srv.submit(sid, src, out, ready_stream=side)
    srv.drain()
    assert srv.query(rid) is not None
    for g, w_ in zip(out, want):
        assert np.array_equal(g.cpu().numpy(), w_)
    srv.close()
