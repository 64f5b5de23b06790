"""The MFMA matcher (block 8, group <= 32) through the C ABI.

Its records must equal the FP32 matchers' bit for bit (same count, same
x, y, t and distance bits per entry): the MFMA pass only chooses which
candidates to score, and every score is the exact FP32 key. Shapes cover
the headline geometry (step 8, range 9), the one-stage 256x256 geometry
(step 3, range 16, group 16), large windows (range 20, 30, 64), clamped
windows on small planes, partial 4 x 4 blocks at the grid's edges, and the
exact fallbacks: a constant plane (every distance ties, the candidate
lists overflow), exact-tie blocks, and planes holding values fp16 cannot
carry (non-finite, or beyond its range). The reference's tensor matcher is
held to the same bar (src/cuda/common/match.cu:1165-1180: the lane
kernels' matches, bit for bit).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _plane(kind, width, height, seed):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:height, 0:width]
    base = 0.5 + 0.3 * np.sin(x * 0.21 + seed) * np.cos(y * 0.17)
    if kind == "noise":
        v = base + rng.normal(0, 0.05, (height, width))
    elif kind == "lownoise":
        v = base + rng.normal(0, 0.002, (height, width))
    elif kind == "flat":
        v = (((x // 8) + (y // 8) + seed) % 3) * 0.25
    elif kind == "const":
        v = np.full((height, width), 0.5)
    elif kind == "bright":
        v = 3000.0 * base + rng.normal(0, 20.0, (height, width))
    elif kind == "huge":
        v = base + rng.normal(0, 0.05, (height, width))
        v[height // 2, width // 3] = 1e6
    elif kind == "nan":
        v = base + rng.normal(0, 0.05, (height, width))
        v[height // 3, width // 2] = np.nan
    else:
        raise ValueError(kind)
    return v.astype(np.float32)


@pytest.fixture
def paths(gpu):
    yield gpu
    gpu.set_match_path(gpu.MATCH_AUTO)


def _both(gpu, plane, **kw):
    t = _dev(plane)
    out = []
    for path in (gpu.MATCH_FP32, gpu.MATCH_MFMA):
        gpu.set_match_path(path)
        rec, cnt = gpu.spatial_match(t, **kw)
        out.append((_host(rec), _host(cnt)))
    return out


def _assert_same(a, b, label):
    (ra, ca), (rb, cb) = a, b
    assert np.array_equal(ca, cb), (label, np.nonzero(ca != cb)[0][:8])
    for i in range(ca.shape[0]):
        k = int(ca[i])
        if not np.array_equal(ra[i, :k], rb[i, :k]):
            j = int(np.nonzero((ra[i, :k] != rb[i, :k]).any(axis=1))[0][0])
            raise AssertionError((label, i, j, ra[i, j].tolist(), rb[i, j].tolist()))


@pytest.mark.parametrize("kind", ["noise", "lownoise", "flat"])
@pytest.mark.parametrize("width,height,step,rng_,group", [
    (160, 120, 8, 9, 8), (131, 97, 3, 16, 16), (150, 110, 8, 20, 16), (170, 130, 6, 30, 8),
    (200, 180, 8, 64, 16), (96, 80, 5, 12, 4), (64, 48, 2, 9, 2), (45, 37, 7, 9, 16),
    (150, 110, 3, 19, 32), (120, 90, 8, 20, 32), (40, 30, 1, 6, 32)])
def test_mfma_equals_fp32(paths, kind, width, height, step, rng_, group):
    gpu = paths
    label = f"{kind} {width}x{height} s{step} r{rng_} g{group}"
    print("mfma match", label, flush=True)
    plane = _plane(kind, width, height, seed=width + rng_)
    a, b = _both(gpu, plane, block=8, step=step, bm_range=rng_, group=group)
    _assert_same(a, b, label)


@pytest.mark.parametrize("group", [16, 32])
@pytest.mark.parametrize("kind", ["const", "bright", "huge", "nan"])
def test_mfma_fallbacks(paths, kind, group):
    gpu = paths
    label = f"{kind} 120x90 s4 r18 g{group}"
    print("mfma match", label, flush=True)
    plane = _plane(kind, 120, 90, seed=3)
    a, b = _both(gpu, plane, block=8, step=4, bm_range=18, group=group)
    _assert_same(a, b, label)


@pytest.mark.parametrize("width,height", [(8, 8), (9, 13), (16, 8), (23, 11)])
def test_mfma_tiny_planes(paths, width, height):
    gpu = paths
    label = f"tiny {width}x{height}"
    print("mfma match", label, flush=True)
    plane = _plane("noise", width, height, seed=7)
    a, b = _both(gpu, plane, block=8, step=1, bm_range=9, group=16)
    _assert_same(a, b, label)


def test_mfma_against_oracle(paths, oracle):
    """The MFMA path against the CPU oracle's lists (same gate as the FP32
    matchers: only near ties may differ)."""
    from test_gpu_bm3d import _compare_matches, _test_plane
    gpu = paths
    gpu.set_match_path(gpu.MATCH_MFMA)
    for flat in (False, True):
        for block, group, step, rng_ in [(8, 16, 3, 16), (8, 8, 8, 9), (8, 16, 8, 20)]:
            plane = _test_plane(96, 72, 5, flat)
            flips = _compare_matches(gpu, oracle, plane, block, step, rng_, group, flat,
                                     f"mfma b{block} g{group} s{step} r{rng_} flat={flat}")
            print(f"mfma vs oracle s{step} r{rng_} g{group} flat={flat}: near-tie flips {flips}", flush=True)


@pytest.mark.parametrize("fmt,kw", [("YUV420PS", dict(sigma=10, final=1)),
                                    ("GRAYS", dict(sigma=5, block_step=3, group_size=16, bm_range=16)),
                                    ("GRAYS", dict(sigma=20, bm_range=24, group_size=16, final=1)),
                                    ("YUV420PS", dict(sigma=3, final=1, block_step=3, group_size=32,
                                                      group_size_basic=16, bm_range=19)),
                                    # planes with different groups and ranges in one launch: the
                                    # launch's list length (group 32) sized for every plane
                                    ("YUV420PS", dict(sigma=10, group_size=[32, 16, 16], bm_range=[8, 30, 30]))])
def test_mfma_filter_output_identical(paths, fmt, kw):
    """A whole filter call gives the same output bits on either matcher."""
    import synth
    gpu = paths
    _, noisy = synth.noisy_frame(fmt, 160, 120, kw["sigma"], seed=4)
    outs = []
    for path in (gpu.MATCH_FP32, gpu.MATCH_MFMA):
        gpu.set_match_path(path)
        inst = gpu.BM3D(fmt, 160, 120, **kw)
        res = inst([_dev(p) for p in noisy])
        outs.append([_host(o) for o in res])
        inst.close()
    for i, (a, b) in enumerate(zip(*outs)):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (fmt, kw, i)


@pytest.mark.parametrize("kw", [dict(sigma=10, radius=1, bm_range=20, group_size=16, final=1),
                                dict(sigma=10, radius=1, bm_range=20, group_size=8, temporal_mode="legacy")])
def test_mfma_temporal_identical(paths, kw):
    """The temporal path's spatial matches come from the same matcher: the
    clip's frames (rolling) or contribution slices (legacy) are the same bits
    on either matcher."""
    import synth
    gpu = paths
    w, h, n = 160, 120, 5
    _, clip = synth.shifted_sequence("YUV420PS", w, h, n, kw["sigma"], seed=2, max_step=2)
    outs = []
    for path in (gpu.MATCH_FP32, gpu.MATCH_MFMA):
        gpu.set_match_path(path)
        inst = gpu.BM3D("YUV420PS", w, h, n, **kw)
        res = inst.filter_clip([[_dev(p) for p in f] for f in clip])
        outs.append([[_host(p) for p in f] for f in res])
        inst.close()
    assert len(outs[0]) == len(outs[1]) == n
    for fa, fb in zip(*outs):
        for a, b in zip(fa, fb):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), kw


def test_mfma_strided_plane(paths):
    """A plane whose row stride exceeds its width (a view into a wider
    buffer) gives the same records on both matchers."""
    import torch
    gpu = paths
    wide = _plane("noise", 200, 110, seed=11)
    full = torch.from_numpy(wide).cuda()
    view = full[:, 7:7 + 150]  # width 150, stride 200
    assert view.stride(0) == 200
    out = []
    for path in (gpu.MATCH_FP32, gpu.MATCH_MFMA):
        gpu.set_match_path(path)
        rec, cnt = gpu.spatial_match(view, block=8, step=4, bm_range=20, group=16)
        out.append((_host(rec), _host(cnt)))
    _assert_same(out[0], out[1], "strided 150x110 (stride 200) s4 r20 g16")
