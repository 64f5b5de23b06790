"""Concurrent calls on one filter instance (the node's fmParallel contract,
src/cuda/runtime/stream_pool.hpp:1-4, docs/api/cuda/README.md:140): host
threads call nss_hip_bm3d_process on one handle, each on its own HIP stream;
the instance leases one of its num_streams frame slots per call and orders
each slot's users on the device. Every frame must be bitwise the frame the
same instance gives when the calls come one at a time."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


@pytest.mark.parametrize("num_streams", [1, 2])
def test_threads_share_one_handle(gpu, num_streams):
    import torch
    import synth
    fmt, w, h = "YUV420PS", 160, 120
    kw = dict(sigma=10, final=1, num_streams=num_streams)
    frames = [synth.noisy_frame(fmt, w, h, 10, seed=40 + k)[1] for k in range(12)]
    src = [[_dev(p) for p in f] for f in frames]
    inst = gpu.BM3D(fmt, w, h, **kw)
    assert inst.resources()["_NSSResourceBytes"][0] > 0
    # serial reference: one call at a time on the default stream
    serial = []
    for f in src:
        out = inst(f)
        torch.cuda.synchronize()
        serial.append([o.cpu().numpy() for o in out])
    outs = [[torch.full_like(p, -1.0) for p in f] for f in src]
    torch.cuda.synchronize()
    errors = []

    def worker(t):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for k in range(t, len(src), 4):
                    inst(src[k], out=outs[k])
            s.synchronize()
        except Exception as e:  # reported after join
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    torch.cuda.synchronize()
    assert not errors, errors
    for k in range(len(src)):
        for g, want in zip(outs[k], serial[k]):
            assert np.array_equal(g.cpu().numpy(), want), (num_streams, k)
    inst.close()


def test_streams_slot_accounting(gpu):
    """num_streams frame slots: the workspace grows with the slot count, and
    memory_limit_mb is checked against all of them at creation."""
    one = gpu.BM3D("YUV420PS", 160, 120, sigma=10, final=1)
    three = gpu.BM3D("YUV420PS", 160, 120, sigma=10, final=1, num_streams=3)
    w1 = one.resources()["_NSSResourceBytes"][0]
    w3 = three.resources()["_NSSResourceBytes"][0]
    one.close()
    three.close()
    assert w3 == 3 * w1
    mb = (w1 >> 20) + 1
    with pytest.raises(gpu.Error, match="each of the 3 stream"):
        gpu.BM3D("YUV420PS", 160, 120, sigma=10, final=1, num_streams=3, memory_limit_mb=mb)


def test_temporal_calls_across_streams(gpu):
    """push on one stream and pop on another, with no synchronisation
    between them: the instance orders its own calls on the device."""
    import torch
    import synth
    _, clip = synth.shifted_sequence("GRAYS", 96, 64, 6, 10, seed=3, max_step=2)
    dev = [[_dev(p) for p in f] for f in clip]
    inst = gpu.BM3D("GRAYS", 96, 64, 6, sigma=10, radius=1)
    want = [[p.cpu().numpy() for p in f] for f in inst.filter_clip(dev)]
    inst.close()
    inst = gpu.BM3D("GRAYS", 96, 64, 6, sigma=10, radius=1)
    a, b = torch.cuda.Stream(), torch.cuda.Stream()
    got = []
    for f in dev:
        inst.push(f, stream=a)
        while True:
            r = inst.pop(stream=b)
            if r is None:
                break
            got.append(r[1])
    torch.cuda.synchronize()
    inst.close()
    assert len(got) == len(want)
    for g, w_ in zip(got, want):
        for x, y in zip(g, w_):
            assert np.array_equal(x.cpu().numpy(), y)
