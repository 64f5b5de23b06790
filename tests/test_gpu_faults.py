"""Error paths under injected faults (the library's allocation and launch
hooks): every construction whose n-th device allocation fails reports the
failure and releases everything it allocated; a frame whose n-th kernel
launch is refused reports it, and the instance then filters the next frame
exactly; a refused launch fails the server's request and the server stays
failed (fail-stop) until it is closed. The reference's tests put a fake
allocator and a launch-error hook behind their device layer for the same
purpose."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(p):
    import torch
    return torch.from_numpy(np.ascontiguousarray(p)).cuda()


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _frame(fmt, w, h, sigma, seed):
    import synth
    _, noisy = synth.noisy_frame(fmt, w, h, sigma, seed=seed)
    return [_dev(p) for p in noisy]


# name -> (constructor, frame format, width, height)
def _cases(gpu):
    return {
        "two-stage": lambda: gpu.BM3D("YUV420PS", 96, 64, sigma=10, final=1),
        "two-slots": lambda: gpu.BM3D("YUV420PS", 96, 64, sigma=10, final=1, num_streams=2),
        "memory-limit": lambda: gpu.BM3D("GRAYS", 320, 224, sigma=10, final=1, memory_limit_mb=3),
        "global-cubes": lambda: gpu.BM3D("GRAYS", 96, 96, sigma=20, final=1, block_size=32, group_size=32,
                                         block_step=16, bm_range=8),
        "rolling": lambda: gpu.BM3D("GRAYS", 64, 48, frames=4, sigma=10, final=1, radius=1),
        "legacy": lambda: gpu.BM3D("GRAYS", 64, 48, frames=4, sigma=10, radius=1, temporal_mode="legacy"),
        "wnnm": lambda: gpu.WNNM("GRAYS", 64, 48, sigma=10),
        "wnnm-temporal": lambda: gpu.WNNM("YUV420PS", 64, 48, frames=4, sigma=[10, 0, 10], radius=1),
        "ncsr-temporal": lambda: gpu.NCSR("GRAYS", 64, 48, frames=4, sigma=10, radius=1, rclip=True),
    }


CASE_NAMES = ["two-stage", "two-slots", "memory-limit", "global-cubes", "rolling", "legacy", "wnnm", "wnnm-temporal",
              "ncsr-temporal"]


def _live(gpu):
    return gpu.debug_live_allocations()


@pytest.mark.parametrize("case", CASE_NAMES)
def test_allocation_failures_release_everything(gpu, case):
    make = _cases(gpu)[case]
    base = _live(gpu)
    inst = make()
    n = _live(gpu)[0] - base[0]
    inst.close()
    assert _live(gpu) == base
    assert n > 0
    print(f"{case}: {n} device allocations at creation", flush=True)
    try:
        for k in range(n):
            gpu.debug_fail_alloc(k)
            with pytest.raises(gpu.Error) as err:
                make()
            gpu.debug_fail_alloc(-1)
            assert "out of memory" in str(err.value).lower(), (k, str(err.value))
            assert _live(gpu) == base, (case, k, _live(gpu), base)
    finally:
        gpu.debug_fail_alloc(-1)
    inst = make()  # and the hook is off again
    inst.close()
    assert _live(gpu) == base


def _refused_launches(call):
    """Arms the launch hook at k = 0, 1, ... until a call goes through; every
    refused call must raise. Returns the launches of one call."""
    import nss_hip as gpu
    try:
        for k in range(256):
            gpu.debug_fail_launch(k)
            try:
                call()
            except gpu.Error as e:
                assert "launch" in str(e).lower() or "failure" in str(e).lower(), str(e)
                gpu.debug_fail_launch(-1)
                yield k
                continue
            return
        raise AssertionError("more than 256 launches in one call")
    finally:
        gpu.debug_fail_launch(-1)


@pytest.mark.parametrize("kw", [dict(sigma=10, final=1), dict(sigma=10, final=1, memory_limit_mb=3),
                                dict(sigma=10, transform=1)])
def test_refused_launch_reports_and_the_next_frame_is_exact(gpu, kw):
    fmt, w, h = ("GRAYS", 320, 224) if "memory_limit_mb" in kw else ("YUV420PS", 96, 64)
    src = _frame(fmt, w, h, 10, 41)
    inst = gpu.BM3D(fmt, w, h, **kw)
    want = [_host(o) for o in inst(src)]
    refused = 0
    for k in _refused_launches(lambda: inst(src)):
        refused += 1
        got = [_host(o) for o in inst(src)]
        for g, w_ in zip(got, want):
            assert np.array_equal(g, w_), k
    inst.close()
    print(f"{kw}: {refused} launches refused in turn", flush=True)
    assert refused >= 3


def test_refused_launch_fails_the_server_request(gpu):
    import torch
    base = _live(gpu)
    srv = gpu.Server(max_batch_frames=2, use_graphs=False)
    sid = srv.open_stream("YUV420PS", 96, 64, sigma=10, final=1)
    src = _frame("YUV420PS", 96, 64, 10, 42)
    out = [torch.empty_like(p) for p in src]
    rid = srv.submit(sid, src, out)
    srv.drain()
    assert srv.query(rid) is not None
    try:
        gpu.debug_fail_launch(0)
        rid = srv.submit(sid, src, out)
        with pytest.raises(gpu.Error):
            srv.drain()
        gpu.debug_fail_launch(-1)
        with pytest.raises(gpu.Error):
            srv.query(rid)
        # fail-stop: later submits are refused with the first failure's message
        with pytest.raises(gpu.Error) as err:
            srv.submit(sid, src, out)
        assert "launch" in str(err.value).lower() or "failure" in str(err.value).lower(), str(err.value)
    finally:
        gpu.debug_fail_launch(-1)
        srv.close()
    torch.cuda.synchronize()
    assert _live(gpu) == base


def test_server_allocation_failures_release_everything(gpu):
    base = _live(gpu)

    def make():
        srv = gpu.Server(max_batch_frames=2)
        try:
            srv.open_stream("YUV420PS", 96, 64, sigma=10, final=1)
            srv.open_stream("GRAYS", 64, 48, frames=4, sigma=10, radius=1)
        except Exception:
            srv.close()
            raise
        return srv

    srv = make()
    n = _live(gpu)[0] - base[0]
    srv.close()
    assert _live(gpu) == base and n > 0
    try:
        for k in range(n):
            gpu.debug_fail_alloc(k)
            with pytest.raises(gpu.Error):
                make()
            gpu.debug_fail_alloc(-1)
            assert _live(gpu) == base, (k, _live(gpu), base)
    finally:
        gpu.debug_fail_alloc(-1)


def test_refused_launch_stops_a_rolling_instance(gpu):
    """A push refused part-way may have added some of a centre's patches to
    the rolling sums already: later pushes are refused (fail-stop), frames
    finished before still pop, and a new instance filters the clip exactly."""
    import synth
    _, seq = synth.shifted_sequence("GRAYS", 64, 48, 5, 10, seed=43)
    frames = [[_dev(p) for p in f] for f in seq]

    def make():
        return gpu.BM3D("GRAYS", 64, 48, frames=5, sigma=10, final=1, radius=1)

    inst = make()
    want = [[_host(p) for p in f] for f in inst.filter_clip(frames)]
    inst.close()
    stopped = 0
    for k in (0, 2, 5):
        inst = make()
        popped = []
        try:
            gpu.debug_fail_launch(k)
            failed_at = None
            for i, f in enumerate(frames):
                try:
                    inst.push(f)
                except gpu.Error:
                    failed_at = i
                    break
                while (r := inst.pop()) is not None:
                    popped.append([_host(p) for p in r[1]])
        finally:
            gpu.debug_fail_launch(-1)
        if failed_at is None:
            inst.close()
            continue
        stopped += 1
        if failed_at + 1 < len(frames):
            with pytest.raises(gpu.Error) as err:
                inst.push(frames[failed_at + 1])
            assert "earlier push failed" in str(err.value)
        while (r := inst.pop()) is not None:
            popped.append([_host(p) for p in r[1]])
        for got, exp in zip(popped, want):
            for g, w_ in zip(got, exp):
                assert np.array_equal(g, w_)
        inst.close()
    assert stopped >= 2
    inst = make()
    again = [[_host(p) for p in f] for f in inst.filter_clip(frames)]
    inst.close()
    for got, exp in zip(again, want):
        for g, w_ in zip(got, exp):
            assert np.array_equal(g, w_)


@pytest.mark.parametrize("kind", ["WNNM", "NCSR"])
def test_refused_launch_in_a_temporal_contribution_reports_and_the_next_is_exact(gpu, kind):
    """nss_hip_wnnm_process_temporal and nss_hip_ncsr_process_temporal keep
    no state between calls: after a refused launch the same centre's
    contribution is exact again."""
    import synth
    _, seq = synth.shifted_sequence("GRAYS", 64, 48, 4, 10, seed=44)
    inst = getattr(gpu, kind)("GRAYS", 64, 48, frames=4, sigma=10, radius=1)
    frames = [[_dev(seq[f][0])] for f in inst.slot_frames(1)]
    want = _host(inst.contribution(1, frames)[0])
    refused = 0
    for k in _refused_launches(lambda: inst.contribution(1, frames)):
        refused += 1
        assert np.array_equal(_host(inst.contribution(1, frames)[0]), want), k
    inst.close()
    assert refused >= 4
