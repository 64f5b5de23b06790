"""Parity of the HIP path with the CPU oracle, through the C ABI.

Gates (the reference's own, tests/data/cuda_tolerances_v1.json and
tests/cuda/test_cuda_match.cpp:49-76):
- match lists: same count; an entry may differ only as a near tie
  |a - b| <= 2e-5 max(1, |a|, |b|), and on exact-tie (piecewise-constant)
  images not at all; distances agree within the same bound;
- output: PSNR >= 60 dB per plane against the oracle, and run-to-run
  bitwise identity.
"""
import sys

import numpy as np
import pytest

import closed_form as cf

pytestmark = pytest.mark.gpu

MIN_PSNR = 60.0


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def _host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _near_tie(a, b):
    return abs(a - b) <= 2e-5 * max(1.0, abs(a), abs(b))


def _test_plane(width, height, seed, flat):
    """tests/cuda/test_cuda_match.cpp:33-46: sinusoid + noise, or exact-tie
    blocks of 0 / 0.25 / 0.5."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:height, 0:width]
    if flat:
        return ((((x // 8) + (y // 8) + seed) % 3) * 0.25).astype(np.float32)
    v = 0.5 + 0.3 * np.sin(x * 0.21 + seed) * np.cos(y * 0.17)
    return (v + rng.normal(0, 0.05, (height, width))).astype(np.float32)


def _compare_matches(gpu, oracle, plane, block, step, rng_, group, flat, label):
    import torch
    rec, cnt = gpu.spatial_match(_dev(plane), block=block, step=step, bm_range=rng_, group=group)
    rec = _host(rec)
    cnt = _host(cnt)
    h, w = plane.shape
    grid = oracle.raster_grid(w, h, block, step)
    assert len(grid) == rec.shape[0]
    flips = 0
    for i, (bx, by) in enumerate(grid):
        want = oracle.spatial_match(plane, bx, by, block, rng_, group)
        assert cnt[i] == len(want), (label, i, cnt[i], len(want))
        for j, (x, y, t, d) in enumerate(want):
            gx, gy, gt, gd = rec[i, j]
            gdist = float(np.array(gd, np.int32).view(np.float32))
            if (gx, gy) != (x, y):
                flips += 1
                assert not flat, (label, i, j, (x, y), (gx, gy))
                assert _near_tie(d, gdist), (label, i, j, d, gdist)
            else:
                assert _near_tie(d, gdist), (label, i, j, d, gdist)
    return flips


@pytest.mark.parametrize("flat", [False, True])
@pytest.mark.parametrize("block,group,step,rng_", [(8, 16, 3, 16), (8, 8, 8, 9), (4, 16, 2, 7), (12, 8, 6, 5),
                                                   (16, 4, 8, 6), (2, 32, 1, 3), (1, 8, 1, 2), (32, 2, 16, 4),
                                                   (8, 64, 3, 30), (16, 8, 8, 20), (32, 4, 16, 64),
                                                   (8, 64, 8, 9), (12, 64, 6, 5), (8, 32, 4, 9),
                                                   (4, 8, 4, 12), (12, 16, 6, 11), (8, 32, 5, 13),
                                                   (2, 16, 2, 10), (1, 8, 3, 14)])
def test_match_lists(gpu, oracle, block, group, step, rng_, flat):
    """Covers the run matcher (B <= 12: one lane tile per lane up to range 9,
    two or three with the keys in registers up to range 16, keys in LDS
    beyond; groups up to 32 by the threshold selection, 64 by the rounds),
    the wave matcher with the window in LDS, and with the window read from
    global (B 32, range 64)."""
    print(f"match b{block} g{group} s{step} r{rng_} flat={flat}", flush=True)
    plane = _test_plane(77 if rng_ < 30 else 150, 61 if rng_ < 30 else 140, 7 + block + group, flat)
    flips = _compare_matches(gpu, oracle, plane, block, step, rng_, group, flat, f"b{block}g{group}")
    print(f"near-tie flips {flips}", flush=True)


def test_match_lists_config0(gpu, oracle):
    """256x256 gray, sigma 5, block 8, step 3, group 16, range 16."""
    import synth
    _, noisy = synth.noisy_frame("GRAYS", 256, 256, 5, seed=0)
    flips = _compare_matches(gpu, oracle, noisy[0], 8, 3, 16, 16, False, "config0")
    print(f"config0 near-tie flips {flips}", flush=True)


def _run(gpu, fmt, planes, ref=None, **kw):
    h, w = planes[0].shape
    inst = gpu.BM3D(fmt, w, h, ref=ref is not None, **kw)
    out = inst([_dev(p) for p in planes], ref=None if ref is None else [_dev(p) for p in ref])
    res = [_host(o) for o in out]
    inst.close()
    return res


def _oracle_frame(oracle, gpu, fmt, planes, ref=None, **kw):
    h, w = planes[0].shape
    p = gpu.parse(fmt, w, h, ref=ref is not None, **kw)
    return oracle.bm3d_frame(planes, p, ref=ref)


def test_config0_one_stage(gpu, oracle):
    import synth
    _, noisy = synth.noisy_frame("GRAYS", 256, 256, 5, seed=0)
    kw = dict(sigma=5, block_size=8, block_step=3, group_size=16, bm_range=16)
    got = _run(gpu, "GRAYS", noisy, **kw)
    want = _oracle_frame(oracle, gpu, "GRAYS", noisy, **kw)
    p = oracle.psnr(got[0], want[0])
    print(f"config0 PSNR vs oracle {p:.2f} dB", flush=True)
    assert p >= MIN_PSNR


@pytest.mark.parametrize("ranges", [[9, 14], [14, 20], [16, 12]])
def test_planes_of_different_matcher_layouts_in_one_launch(gpu, oracle, ranges):
    """One match launch covers the three planes of a block size, and its
    layout is the one the largest range needs (one lane tile per lane up to
    range 9, up to three with the keys in registers up to 16, keys in LDS
    beyond). Each plane's output must equal that plane run alone, bit for
    bit, and the oracle's within the PSNR gate."""
    import synth
    _, noisy = synth.noisy_frame("YUV420PS", 160, 120, 10, seed=3)
    kw = dict(sigma=10, block_step=4, group_size=16)
    got = _run(gpu, "YUV420PS", noisy, bm_range=ranges, **kw)
    want = _oracle_frame(oracle, gpu, "YUV420PS", noisy, bm_range=ranges, **kw)
    for i in range(3):
        r = ranges[min(i, len(ranges) - 1)]
        alone = _run(gpu, "GRAYS", [noisy[i]], bm_range=r, **kw)[0]
        print(f"ranges {ranges} plane {i} range {r}", flush=True)
        assert np.array_equal(got[i].view(np.uint32), alone.view(np.uint32)), (ranges, i)
        assert oracle.psnr(got[i], want[i]) >= MIN_PSNR


@pytest.mark.parametrize("fmt", ["YUV420PS", "RGBS", "GRAYS"])
def test_two_stage(gpu, oracle, fmt):
    import synth
    _, noisy = synth.noisy_frame(fmt, 160, 120, 10, seed=1)
    kw = dict(sigma=10, final=1)
    got = _run(gpu, fmt, noisy, **kw)
    want = _oracle_frame(oracle, gpu, fmt, noisy, **kw)
    for i, (g, w_) in enumerate(zip(got, want)):
        p = oracle.psnr(g, w_)
        print(f"{fmt} plane {i} PSNR vs oracle {p:.2f} dB", flush=True)
        assert p >= MIN_PSNR


def test_headline_1080p_yuv420_two_stage(gpu, oracle):
    """The headline workload at its full size: one 1920x1080 YUV420PS frame
    (the sinusoid-and-rectangles luma, smooth chroma), sigma 10, both stages
    at the default arguments, >= 60 dB per plane against the oracle."""
    import synth
    _, noisy = synth.noisy_frame("YUV420PS", 1920, 1080, 10, seed=0)
    kw = dict(sigma=10, final=1)
    got = _run(gpu, "YUV420PS", noisy, **kw)
    want = _oracle_frame(oracle, gpu, "YUV420PS", noisy, **kw)
    for i, (g, w_) in enumerate(zip(got, want)):
        p = oracle.psnr(g, w_)
        d = np.abs(g.astype(np.float64) - w_)
        print(f"1080p YUV420 plane {i}: PSNR vs oracle {p:.2f} dB, max {d.max():.3e}", flush=True)
        assert np.isfinite(g).all() and p >= MIN_PSNR


@pytest.mark.parametrize("block", [1, 2, 4, 8, 12, 16, 32])
def test_shapes_hard_and_wiener(gpu, oracle, block):
    """tests/cuda/test_cuda_bm3d_live.py:33-50: every block x group, basic and
    Wiener (ref given), on a small seeded sinusoid + noise plane."""
    import synth
    h, w = 48, 56
    clean = synth.clean_image(w, h, seed=block)
    noisy = synth.add_noise(clean, 20, seed=block)
    for group in (1, 2, 4, 8, 16, 32, 64):
        # block 32 with group 64, or group 32 and a Wiener reference, keeps
        # the group cube in global memory
        step = max(1, block // 2)
        for wiener in (False, True):
            kw = dict(sigma=20, block_size=block, group_size=group, block_step=step, bm_range=4)
            ref = [clean] if wiener else None
            got = _run(gpu, "GRAYS", [noisy], ref=ref, **kw)
            want = _oracle_frame(oracle, gpu, "GRAYS", [noisy], ref=ref, **kw)
            p = oracle.psnr(got[0], want[0])
            assert p >= MIN_PSNR, (block, group, wiener, p)
    print(f"block {block} ok", flush=True)


@pytest.mark.parametrize("block,group,rng_", [(32, 8, 40), (16, 32, 48), (8, 8, 60)])
def test_large_windows(gpu, oracle, block, group, rng_):
    """Windows too large for tile aggregation in LDS take the frame-sized
    atomic sums (and the wave matcher); same gates."""
    import synth
    h, w = 120, 136
    clean = synth.clean_image(w, h, seed=block)
    noisy = synth.add_noise(clean, 20, seed=block + 1)
    for wiener in (False, True):
        kw = dict(sigma=20, block_size=block, group_size=group, block_step=block // 2, bm_range=rng_)
        ref = [clean] if wiener else None
        got = _run(gpu, "GRAYS", [noisy], ref=ref, **kw)
        want = _oracle_frame(oracle, gpu, "GRAYS", [noisy], ref=ref, **kw)
        p = oracle.psnr(got[0], want[0])
        print(f"large window b{block} g{group} r{rng_} wiener={wiener}: {p:.2f} dB", flush=True)
        assert p >= MIN_PSNR


@pytest.mark.parametrize("block", [1, 2, 4, 8, 12, 16, 32])
def test_single_block_closed_form(gpu, oracle, block):
    """tests/test_bm3d_semantics.py:161-173 on the GPU: one block, one match."""
    rng = np.random.default_rng(59)
    p = rng.uniform(-.02, .03, (block, block)).astype(np.float32)
    ref = (p * np.float32(.8)).astype(np.float32)
    for group in (1, 2, 4, 8, 16, 32, 64):
        for user in (.5, 3, 20, 40):
            for wiener in (False, True):
                got = _run(gpu, "GRAYS", [p], ref=[ref] if wiener else None, sigma=user, block_size=block,
                           block_step=block, group_size=group, bm_range=1)
                want = cf.single_block(p.astype(np.float64), ref.astype(np.float64), group,
                                       oracle.effective_sigma(user), wiener)
                assert np.max(np.abs(got[0] - want)) <= 3e-5, (block, group, user, wiener)


def test_constant_image_identity(gpu):
    img = np.full((64, 80), .3, np.float32)
    got = _run(gpu, "GRAYS", [img], sigma=10, block_step=4)
    np.testing.assert_allclose(got[0], img, atol=1e-6)


def test_sigma_zero_copies(gpu):
    import synth
    _, noisy = synth.noisy_frame("YUV420PS", 64, 48, 10, seed=2)
    got = _run(gpu, "YUV420PS", noisy, sigma=[0, 5, 0])
    assert np.array_equal(got[0], noisy[0]) and np.array_equal(got[2], noisy[2])
    assert not np.array_equal(got[1], noisy[1])


def test_run_to_run_identity(gpu):
    """Fixed-point aggregation: two runs are bitwise identical."""
    import synth
    _, noisy = synth.noisy_frame("YUV420PS", 320, 240, 10, seed=4)
    a = _run(gpu, "YUV420PS", noisy, sigma=10, final=1)
    b = _run(gpu, "YUV420PS", noisy, sigma=10, final=1)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_final_equals_two_calls(gpu):
    """final=1 is the basic filter (the *_basic values) followed by the
    Wiener filter with the basic estimate as ref, bitwise
    (src/frontend/bm3d_args.cpp:161-196)."""
    import synth
    _, noisy = synth.noisy_frame("YUV420PS", 128, 96, 10, seed=12)
    # basic stage: sigma 12, group 8; Wiener: sigma 10, group 16; both step 4
    fused = _run(gpu, "YUV420PS", noisy, final=1, sigma=10, block_step=4, group_size=16, sigma_basic=12,
                 group_size_basic=8)
    basic = _run(gpu, "YUV420PS", noisy, sigma=12, group_size=8, block_step=4)
    wiener = _run(gpu, "YUV420PS", noisy, ref=basic, sigma=10, block_step=4, group_size=16)
    for a, b in zip(fused, wiener):
        assert np.array_equal(a, b)


def test_memory_limit_enforced_and_reported(gpu):
    """tests/cuda/test_cuda_bm3d_live.py:345-369 and
    tests/test_plan01_plugin.py:63-73: _NSSResourceLimit is 0 without
    memory_limit_mb and the limit in bytes with it; the owned bytes stay
    within the limit; a limit too small for the clip fails at creation; the
    smallest limit that creates gives the unlimited output."""
    import synth
    w, h = 640, 360
    _, noisy = synth.noisy_frame("YUV420PS", w, h, 10, seed=14)
    inst = gpu.BM3D("YUV420PS", w, h, sigma=10, final=1)
    r = inst.resources()
    free_out = [_host(o) for o in inst([_dev(p) for p in noisy])]
    inst.close()
    print("resources", r, flush=True)
    assert r["_NSSResourceLimit"] == 0
    assert r["_NSSResourcePeak"] == sum(r["_NSSResourceBytes"]) > (1 << 20)
    assert r["_NSSResourceBytes"][1:] == [0, 0, 0, 0, 0]
    need = (r["_NSSResourcePeak"] >> 20) + 1
    # below one batch of grid rows (the frame-sized sums alone are 5.5 MB)
    with pytest.raises(gpu.Error) as e:
        gpu.BM3D("YUV420PS", w, h, sigma=10, final=1, memory_limit_mb=1)
    assert "memory_limit_mb is too small for this clip" in str(e.value)
    inst = gpu.BM3D("YUV420PS", w, h, sigma=10, final=1, memory_limit_mb=need)
    r2 = inst.resources()
    out = [_host(o) for o in inst([_dev(p) for p in noisy])]
    inst.close()
    assert r2["_NSSResourceLimit"] == need << 20
    assert sum(r2["_NSSResourceBytes"][:5]) <= r2["_NSSResourceLimit"]
    for x, y in zip(free_out, out):
        assert np.array_equal(x, y)
    roll = gpu.BM3D("YUV420PS", w, h, frames=6, sigma=10, radius=1)
    rr = roll.resources()
    roll.close()
    assert rr["_NSSResourceBytes"][1] > 0 and rr["_NSSResourceBytes"][4] > 0
    assert rr["_NSSResourcePeak"] == sum(rr["_NSSResourceBytes"])


def test_memory_limit_does_not_change_output(gpu):
    """tests/cuda/test_cuda_bm3d_live.py:352-369: memory_limit_mb never
    changes the result."""
    import synth
    _, noisy = synth.noisy_frame("YUV420PS", 96, 64, 10, seed=13)
    a = _run(gpu, "YUV420PS", noisy, sigma=10, final=1)
    b = _run(gpu, "YUV420PS", noisy, sigma=10, final=1, memory_limit_mb=1)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


LIMIT_CASES = [
    ("YUV420PS", 480, 272, dict(sigma=10, final=1)),                      # block-8 register buckets
    ("YUV444PS", 320, 224, dict(sigma=10, final=1, chroma=1)),            # CBM3D: groups of plane 0
    ("GRAYS", 400, 240, dict(sigma=20, block_size=4, group_size=16, block_step=2, bm_range=6)),  # atomic kernel
    ("GRAYS", 320, 224, dict(sigma=20, final=1, group_size=32, group_size_basic=16, block_step=4)),
    ("RGBS", 256, 192, dict(sigma=10, final=1, block_size=12, group_size=8, block_step=6)),
    ("GRAYS", 288, 224, dict(sigma=20, final=1, block_size=32, group_size=32, block_step=16, bm_range=8)),  # global cubes
]


@pytest.mark.parametrize("case", range(len(LIMIT_CASES)))
def test_memory_limit_batches_give_the_same_bits(gpu, case):
    """docs/api/cuda/README.md:141: under a memory_limit_mb below what the
    frame buffers need, the work is batched to fit and the output does not
    change. The smallest limit that creates (found by bisection) batches the
    grid rows the most (or, with group cubes in global memory, runs the
    fewest cubes at once); its bits equal the unlimited instance's, its
    buffers stay within the limit, and the matcher runs once per batch."""
    import torch
    import synth
    fmt, w, h, kw = LIMIT_CASES[case]
    _, noisy = synth.noisy_frame(fmt, w, h, kw["sigma"], seed=60 + case)
    src = [_dev(p) for p in noisy]
    free = gpu.BM3D(fmt, w, h, **kw)
    want = [_host(o) for o in free(src)]
    need = (free.resources()["_NSSResourcePeak"] >> 20) + 1
    free.close()

    def creates(mb):
        try:
            gpu.BM3D(fmt, w, h, memory_limit_mb=mb, **kw).close()
            return True
        except gpu.Error as e:
            assert "memory_limit_mb is too small for this clip" in str(e)
            return False
    lo, hi = 1, need  # creates(hi); find the smallest limit that creates
    while lo < hi:
        mid = (lo + hi) // 2
        if creates(mid):
            hi = mid
        else:
            lo = mid + 1
    inst = gpu.BM3D(fmt, w, h, memory_limit_mb=lo, **kw)
    r = inst.resources()
    inst.set_tracing(True)
    inst.kernel_times(reset=True)
    got = [_host(o) for o in inst(src)]
    torch.cuda.synchronize()
    kt = inst.kernel_times(reset=True)
    inst.close()
    print(f"case {case}: unlimited {need} MiB, smallest limit {lo} MiB, owned {sum(r['_NSSResourceBytes']) >> 20} MiB, "
          f"launches {kt}", flush=True)
    assert lo < need and sum(r["_NSSResourceBytes"]) <= lo << 20
    nstages = 2 if kw.get("final") else 1
    if kw.get("block_size", 8) < 32:
        assert kt["match"][0] >= 2 * nstages  # at least two batches per stage
    # (global group cubes: the cube count shrinks first, since the cubes, not
    # the rows' records, dominate the buffers)
    for g, w_ in zip(got, want):
        assert np.array_equal(g, w_)
