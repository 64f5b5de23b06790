"""Temporal (radius > 0, rolling) and CBM3D (chroma=1) on the HIP path
against the oracle, through the C ABI (push/pop for temporal).

Gates as in tests/test_gpu_bm3d.py: >= 60 dB per plane against the oracle
(tests/data/cuda_tolerances_v1.json), bitwise run-to-run identity, and the
reference's own invariants (one-frame clip at maximum radius = spatial;
CBM3D plane 0 = the separate filter's plane 0).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIN_PSNR = 60.0


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def _host(planes):
    import torch
    torch.cuda.synchronize()
    return [p.cpu().numpy() for p in planes]


def _gpu_clip(gpu, fmt, clip, ref_clip=None, **kw):
    h, w = clip[0][0].shape
    inst = gpu.BM3D(fmt, w, h, len(clip), ref=ref_clip is not None, **kw)
    out = inst.filter_clip([[_dev(p) for p in f] for f in clip],
                           None if ref_clip is None else [[_dev(p) for p in f] for f in ref_clip])
    res = [_host(f) for f in out]
    inst.close()
    assert len(res) == len(clip)
    return res


def _oracle_clip(oracle, gpu, fmt, clip, ref_clip=None, **kw):
    h, w = clip[0][0].shape
    p = gpu.parse(fmt, w, h, len(clip), ref=ref_clip is not None, **kw)
    return oracle.bm3d_clip(clip, p, ref_clip=ref_clip)


def _check(oracle, got, want, label):
    for n, (gf, wf) in enumerate(zip(got, want)):
        for i, (g, w_) in enumerate(zip(gf, wf)):
            p = oracle.psnr(g, w_)
            assert p >= MIN_PSNR, (label, n, i, p)


def _moving_clip(fmt, w, h, frames, sigma, seed=0):
    import synth
    _, noisy = synth.shifted_sequence(fmt, w, h, frames, sigma, seed=seed, max_step=2)
    return noisy


@pytest.mark.parametrize("radius", [1, 2])
@pytest.mark.parametrize("final", [0, 1])
def test_temporal_rolling(gpu, oracle, radius, final):
    clip = _moving_clip("YUV420PS", 64, 48, 7, 10, seed=radius)
    kw = dict(sigma=10, radius=radius, final=final)
    got = _gpu_clip(gpu, "YUV420PS", clip, **kw)
    want = _oracle_clip(oracle, gpu, "YUV420PS", clip, **kw)
    print(f"temporal r{radius} final={final}: min PSNR vs oracle "
          f"{min(oracle.psnr(g, w_) for gf, wf in zip(got, want) for g, w_ in zip(gf, wf)):.2f} dB", flush=True)
    _check(oracle, got, want, f"r{radius} final{final}")


def test_temporal_shapes_and_ref(gpu, oracle):
    """Other block / group / predictive settings (staged and unstaged
    predictive windows), and a ref clip (Wiener)."""
    import synth
    clip = _moving_clip("GRAYS", 56, 44, 5, 20, seed=3)
    clean = [[synth.clean_image(56, 44, seed=3)] for _ in range(5)]
    for kw in (dict(block_size=4, block_step=2, group_size=16, ps_num=3, ps_range=2),
               dict(block_size=8, block_step=4, group_size=32, bm_range=6, ps_range=5),
               dict(block_size=16, block_step=8, group_size=4, bm_range=4),
               # windows too large to stage in LDS: scored from the frame in chunks
               dict(block_size=8, block_step=4, group_size=8, bm_range=4, ps_range=20)):
        kw = dict(kw, sigma=20, radius=1)
        _check(oracle, _gpu_clip(gpu, "GRAYS", clip, **kw), _oracle_clip(oracle, gpu, "GRAYS", clip, **kw), str(kw))
        got = _gpu_clip(gpu, "GRAYS", clip, ref_clip=clean, **kw)
        want = _oracle_clip(oracle, gpu, "GRAYS", clip, ref_clip=clean, **kw)
        _check(oracle, got, want, "ref " + str(kw))
    print("temporal shapes ok", flush=True)


def test_temporal_global_cubes(gpu, oracle):
    """Block 32 with group 32 in both stages (the Wiener stage's two cubes
    exceed the LDS and live in global memory) and block 32 with group 64."""
    clip = _moving_clip("GRAYS", 80, 72, 4, 20, seed=5)
    for kw in (dict(block_size=32, block_step=16, group_size=32, bm_range=4, final=1),
               dict(block_size=32, block_step=16, group_size=64, bm_range=6)):
        kw = dict(kw, sigma=20, radius=1)
        got = _gpu_clip(gpu, "GRAYS", clip, **kw)
        want = _oracle_clip(oracle, gpu, "GRAYS", clip, **kw)
        print(f"temporal {kw}: min PSNR vs oracle "
              f"{min(oracle.psnr(g, w_) for gf, wf in zip(got, want) for g, w_ in zip(gf, wf)):.2f} dB", flush=True)
        _check(oracle, got, want, str(kw))


def test_one_frame_at_max_radius_equals_spatial(gpu):
    """tests/test_bm3d_semantics.py:125-135 on the GPU."""
    rng = np.random.default_rng(73)
    noisy = list(rng.uniform(.1, .8, (3, 12, 12)).astype(np.float32))
    kw = dict(sigma=3, block_size=4, group_size=8, block_step=2, bm_range=2)
    inst = gpu.BM3D("RGBS", 12, 12, 1, **kw)
    spatial = _host(inst([_dev(p) for p in noisy]))
    inst.close()
    temporal = _gpu_clip(gpu, "RGBS", [noisy], radius=16, **kw)[0]
    for a, b in zip(spatial, temporal):
        np.testing.assert_allclose(a, b, atol=1e-5)


def test_temporal_run_to_run_identity(gpu):
    clip = _moving_clip("YUV420PS", 64, 48, 5, 10, seed=8)
    a = _gpu_clip(gpu, "YUV420PS", clip, sigma=10, radius=1, final=1)
    b = _gpu_clip(gpu, "YUV420PS", clip, sigma=10, radius=1, final=1)
    for fa, fb in zip(a, b):
        for x, y in zip(fa, fb):
            assert np.array_equal(x, y)


def test_push_pop_protocol(gpu):
    """Frames come out in order; pushing past the clip, or leaving too many
    finished frames unpopped, is an error (never a silent change)."""
    import torch
    clip = _moving_clip("GRAYS", 32, 32, 4, 10, seed=2)
    inst = gpu.BM3D("GRAYS", 32, 32, 4, sigma=10, radius=1)
    order = []
    for f in clip:
        inst.push([_dev(p) for p in f])
        while (r := inst.pop()) is not None:
            order.append(r[0])
    assert order == [0, 1, 2, 3]
    with pytest.raises(gpu.Error, match="already pushed"):
        inst.push([_dev(p) for p in clip[0]])
    inst.close()
    inst = gpu.BM3D("GRAYS", 32, 32, 40, sigma=10, radius=1)
    with pytest.raises(gpu.Error, match="pop the finished frames"):
        for _ in range(40):
            inst.push([_dev(p) for p in clip[0]])
    inst.close()
    torch.cuda.synchronize()


def test_chroma_spatial(gpu, oracle):
    import synth
    _, noisy = synth.noisy_frame("YUV444PS", 72, 56, 10, seed=6)
    for kw in (dict(sigma=10, chroma=1), dict(sigma=[10, 5, 0], chroma=1, final=1),
               dict(sigma=[0, 10, 10], chroma=1, block_step=4, group_size=16)):
        inst = gpu.BM3D("YUV444PS", 72, 56, **kw)
        got = _host(inst([_dev(p) for p in noisy]))
        inst.close()
        want = oracle.bm3d_frame(noisy, gpu.parse("YUV444PS", 72, 56, **kw))
        for i, (g, w_) in enumerate(zip(got, want)):
            assert oracle.psnr(g, w_) >= MIN_PSNR, (kw, i)
    # Plane 0 of CBM3D is the plain filter's plane 0 (tests/test_bm3d_chroma.py).
    a = gpu.BM3D("YUV444PS", 72, 56, sigma=10, chroma=1)
    b = gpu.BM3D("YUV444PS", 72, 56, sigma=10)
    pa = _host(a([_dev(p) for p in noisy]))
    pb = _host(b([_dev(p) for p in noisy]))
    assert np.array_equal(pa[0], pb[0])
    a.close()
    b.close()


def test_chroma_temporal(gpu, oracle):
    clip = _moving_clip("YUV444PS", 48, 40, 5, 10, seed=4)
    kw = dict(sigma=10, chroma=1, radius=1, final=1)
    _check(oracle, _gpu_clip(gpu, "YUV444PS", clip, **kw), _oracle_clip(oracle, gpu, "YUV444PS", clip, **kw),
           "chroma temporal")


def _gpu_legacy(gpu, fmt, clip, **kw):
    h, w = clip[0][0].shape
    inst = gpu.BM3D(fmt, w, h, len(clip), temporal_mode="legacy", **kw)
    assert inst.legacy
    out = inst.filter_clip([[_dev(p) for p in f] for f in clip])
    inst.close()
    assert len(out) == len(clip)
    return out


@pytest.mark.parametrize("final", [0, 1])
@pytest.mark.parametrize("shape", ["tiles", "atomic"])
def test_legacy_slices(gpu, oracle, final, shape):
    """temporal_mode="legacy": per centre the 2r+1 num/den slice pairs
    (src/host/filter_bm3d.cpp:322-334) against the oracle's, >= 60 dB per
    slice relative to its peak; reduced with the device VAggregate they give
    the rolling output (>= 60 dB against the oracle, within float rounding of
    the device's own rolling frames). "tiles" is the block-8 tile path,
    "atomic" a block-4 shape on frame-sized sums."""
    import torch
    fmt, w, h, frames = "YUV420PS", 64, 48, 5
    clip = _moving_clip(fmt, w, h, frames, 10, seed=21 + final)
    kw = dict(sigma=[10, 10, 0], radius=1, final=final)
    if shape == "atomic":
        kw.update(block_size=4, block_step=2, group_size=8, bm_range=4)
    fat = _gpu_legacy(gpu, fmt, clip, **kw)
    p = gpu.parse(fmt, w, h, frames, temporal_mode="legacy", **kw)
    want = oracle.bm3d_clip_legacy(clip, p)
    for c in range(frames):
        for i in range(3):
            g = fat[c][i].cpu().numpy()
            assert g.shape == want[c][i].shape
            for s in range(6):
                ph = g.shape[0] // 6
                a, b = g[s * ph:(s + 1) * ph], want[c][i][s * ph:(s + 1) * ph]
                peak = max(float(np.abs(b).max()), 1e-6)
                if not b.any():
                    assert not a.any(), (c, i, s)
                    continue
                assert oracle.psnr(a / peak, b / peak) >= MIN_PSNR, (shape, final, c, i, s)
    src = [[_dev(q) for q in f] for f in clip]
    red = gpu.VAggregate(fat, src, 1)
    rolling = _gpu_clip(gpu, fmt, clip, **kw)
    want_rolling = _oracle_clip(oracle, gpu, fmt, clip, **kw)
    got = [_host(f) for f in red]
    _check(oracle, got, want_rolling, f"legacy+VAggregate {shape} final={final}")
    for gf, rf in zip(got, rolling):
        for g, r_ in zip(gf, rf):
            np.testing.assert_allclose(g, r_, rtol=0, atol=2e-5)
    # the plane with sigma 0 comes back unchanged
    for n in range(frames):
        assert np.array_equal(got[n][2], clip[n][2])
    torch.cuda.synchronize()


def test_vaggregate_kernel_is_exact(gpu, oracle):
    """The device reduction against the oracle's vaggregate_target on random
    slices with dens at, below and around the 1e-12 cut, three frames at
    radius 1 and a selected plane subset: bit for bit (float32 sums in
    ascending centre order, IEEE division)."""
    rng = np.random.default_rng(5)
    h, w, r, frames = 12, 37, 1, 4
    src, fat = [], []
    for _ in range(frames):
        src.append([rng.uniform(-1, 1, (h, w)).astype(np.float32) for _ in range(2)])
        f = []
        for _ in range(2):
            a = rng.uniform(-2, 2, (6 * h, w)).astype(np.float32)
            for s in range(3):
                d = a[(2 * s + 1) * h:(2 * s + 2) * h]
                d[:] = np.abs(d)
                d[rng.uniform(size=d.shape) < 0.2] = 0
                d[rng.uniform(size=d.shape) < 0.1] = 1e-13
                d[rng.uniform(size=d.shape) < 0.1] = 1
            f.append(a)
        fat.append(f)
    want = oracle.vaggregate(fat, src, r, planes=[1])
    got = gpu.VAggregate([[_dev(q) for q in f] for f in fat], [[_dev(q) for q in f] for f in src], r, planes=[1])
    for n in range(frames):
        for i in range(2):
            assert np.array_equal(_host(got[n])[i], want[n][i]), (n, i)


def test_vaggregate_argument_errors(gpu):
    """src/frontend/vaggregate_args.cpp:13-45 messages."""
    a = [[_dev(np.zeros((6 * 4, 5)))]]
    s = [[_dev(np.zeros((4, 5)))]]
    with pytest.raises(gpu.Error, match="radius must be in"):
        gpu.VAggregate(a, s, 17)
    with pytest.raises(gpu.Error, match="height=src.height"):
        gpu.VAggregate(a, s, 2)
    with pytest.raises(gpu.Error, match="unique valid plane"):
        gpu.VAggregate(a, s, 1, planes=[0, 0])


@pytest.mark.parametrize("kw", [dict(radius=2), dict(radius=3), dict(radius=1, chroma=1),
                                dict(radius=1, block_size=16, block_step=8, group_size=4, bm_range=4)])
def test_legacy_more_shapes(gpu, oracle, kw):
    """Legacy slices reduced on the device against the oracle's rolling
    frames (>= 60 dB per plane): radius 2 (tile path), radius 3 (beyond the
    tile path's slot count: frame-sized sums), CBM3D, and block 16."""
    fmt, w, h, frames = ("YUV444PS" if kw.get("chroma") else "YUV420PS"), 48, 40, 6
    clip = _moving_clip(fmt, w, h, frames, 10, seed=31)
    kw = dict(kw, sigma=10)
    fat = _gpu_legacy(gpu, fmt, clip, **kw)
    red = [_host(f) for f in gpu.VAggregate(fat, [[_dev(q) for q in f] for f in clip], kw["radius"])]
    _check(oracle, red, _oracle_clip(oracle, gpu, fmt, clip, **kw), f"legacy {kw}")


def test_legacy_with_ref_and_identity(gpu, oracle):
    """Legacy output of a Wiener stage with a ref clip against the oracle's
    slices, and bitwise identical on a second run."""
    import synth
    fmt, w, h, frames = "GRAYS", 40, 32, 4
    clip = _moving_clip(fmt, w, h, frames, 20, seed=41)
    ref = [[synth.clean_image(w, h, seed=41)] for _ in range(frames)]
    kw = dict(sigma=20, radius=1, temporal_mode="legacy")
    outs = []
    for _ in range(2):
        inst = gpu.BM3D(fmt, w, h, frames, ref=True, **kw)
        fat = inst.filter_clip([[_dev(p) for p in f] for f in clip], [[_dev(p) for p in f] for f in ref])
        inst.close()
        outs.append([_host(f) for f in fat])
    for a, b in zip(*outs):
        assert np.array_equal(a[0], b[0])
    want = oracle.bm3d_clip_legacy(clip, gpu.parse(fmt, w, h, frames, ref=True, **kw), ref_clip=ref)
    for c in range(frames):
        g, o = outs[0][c][0], want[c][0]
        for s in range(6):
            a, b = g[s * h:(s + 1) * h], o[s * h:(s + 1) * h]
            if not b.any():
                assert not a.any()
                continue
            peak = float(np.abs(b).max())
            assert oracle.psnr(a / peak, b / peak) >= MIN_PSNR, (c, s)


@pytest.mark.parametrize("r,final,chunk,legacy", [(1, 0, 2, False), (1, 1, 3, False), (2, 1, 2, False),
                                                  (2, 0, 1, False), (1, 1, 2, True), (2, 0, 3, True)])
def test_random_access_chunks_equal_the_sequential_clip(gpu, r, final, chunk, legacy):
    """src/host/filter_bm3d.cpp:985-1072 rollingGetFrame: a frame asked for
    out of order is computed with its chunk of rolling_chunk frames from the
    chunk's dependency window; the output does not depend on the order. Chunks
    asked for shuffled, again, and in order (the carried state) are bitwise
    the sequential clip's frames."""
    import random
    import synth
    n_frames = 11
    _, seq = synth.shifted_sequence("GRAYS", 64, 48, n_frames, 10, seed=50 + r + 3 * final + chunk)
    frames = [[_dev(p) for p in f] for f in seq]
    kw = dict(sigma=10, final=final, radius=r, rolling_chunk=chunk)
    if legacy:
        kw["temporal_mode"] = "legacy"
    inst = gpu.BM3D("GRAYS", 64, 48, frames=n_frames, **kw)
    want = [_host(f) for f in inst.filter_clip(frames)]
    inst.close()
    assert len(want) == n_frames
    inst = gpu.BM3D("GRAYS", 64, 48, frames=n_frames, **kw)
    margin = r * ((1 if legacy else 2) + (2 if final else 0))
    if legacy:
        chunk = 1  # legacy output takes no rolling arguments: one centre per request
    starts = list(range(0, n_frames, chunk))
    order = starts[:]
    random.Random(7).shuffle(order)
    order += [order[0], starts[0], starts[min(1, len(starts) - 1)]]  # a repeat, then two chunks in order
    for n in order:
        start, count, first, last = inst.chunk(n)
        assert (start, count) == (n, min(chunk, n_frames - n))
        assert (first, last) == (max(0, n - margin), min(n_frames - 1, n + count - 1 + margin))
        got = inst.get_chunk(n, lambda i: frames[i])
        for i in range(start, start + count):
            for g, w_ in zip(_host(got[i]), want[i]):
                assert np.array_equal(g, w_), (n, i)
    inst.close()
