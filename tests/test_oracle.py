"""Pins the CPU oracle against the reference's own known answers, closed forms
and brute force (no GPU)."""
import numpy as np
import pytest

import closed_form as cf

BLOCKS = (1, 2, 4, 8, 12, 16, 32)
GROUPS = (1, 2, 4, 8, 16, 32, 64)


def _tol(v):
    # tests/test_bm3d_contract.cpp:70: min(3e-5, 2e-5 + 2e-4 |v|)
    return np.minimum(3e-5, 2e-5 + 2e-4 * np.abs(v))


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("wiener", [False, True])
@pytest.mark.parametrize("block", BLOCKS)
def test_group_filter_closed_form(oracle, block, wiener, fp64):
    """tests/test_bm3d_contract.cpp:60-71: every B x G x {hard, Wiener}, k =
    min(G, 3), random values in [-0.2, 0.8), sigma 0.015."""
    rng = np.random.default_rng(42 + block)
    for group in GROUPS:
        if block == 32 and group == 64:
            continue  # 256 KiB cubes: covered at G <= 32 for B = 32
        k = min(group, 3)
        cube = np.zeros((group, block, block), np.float32)
        ref = np.zeros_like(cube)
        cube[:k] = rng.uniform(-.2, .8, (k, block, block))
        ref[:k] = rng.uniform(-.2, .8, (k, block, block))
        want, wexp = cf.group_filter(cube.astype(np.float64), k, .015, wiener, ref.astype(np.float64))
        got, w = oracle.filter_group(cube, k, .015, wiener, ref, fp64=fp64)
        assert abs(w - wexp) <= 1e-5 * max(1.0, abs(wexp)), (block, group, w, wexp)
        err = np.abs(got[:k] - want[:k])
        assert np.all(err <= _tol(want[:k])), (block, group, float(err.max()))


def test_threshold_neighbours(oracle):
    """tests/test_bm3d_contract.cpp:140-161: a coefficient placed just below,
    at and above 2.7 sigma on the coarsest group-axis Haar detail x patch DC
    is either kept exactly or removed exactly; clear cases decide."""
    sigma = oracle.effective_sigma(3.0)
    thr = np.float32(2.7) * np.float32(sigma)
    basis = np.zeros((8, 8, 8))
    basis[:4] = 1.0 / np.sqrt(8.) / 8
    basis[4:] = -1.0 / np.sqrt(8.) / 8
    for coef in (thr * np.float32(.99), np.nextafter(thr, np.float32(0)), thr,
                 np.nextafter(thr, np.float32(np.inf)), thr * np.float32(1.01)):
        cube = (.25 + float(coef) * basis).astype(np.float32)
        out, w = oracle.filter_group(cube, 8, sigma)
        rec = float(np.sum((out.astype(np.float64) - .25) * basis))
        kept = abs(rec - coef) < 2e-5
        removed = abs(rec) < 2e-5
        assert kept or removed
        if coef < thr - 2e-5:
            assert removed
        if coef > thr + 2e-5:
            assert kept
        assert abs(w - (.5 if kept else 1.)) <= 1e-6
        np.testing.assert_allclose(out, .25 + (coef * basis if kept else 0), atol=2e-5)


def test_hard_threshold_zero_sigma_is_identity(oracle):
    """With sigma 0 nothing is thresholded: the 3-D transform round trip
    returns the cube (DCT-II/III and Haar orthonormality)."""
    rng = np.random.default_rng(1)
    for b in BLOCKS:
        for g in (1, 8, 16):
            cube = rng.uniform(-1, 1, (g, b, b)).astype(np.float32)
            out, w = oracle.filter_group(cube, g, 0.0)
            np.testing.assert_allclose(out, cube, atol=1e-5)
            assert w == pytest.approx(1.0 / (g * b * b))


def test_dc_of_ones(oracle):
    """A constant cube has only its DC: with any threshold above 0 the output
    is the constant (DC of ones = sqrt(N) per axis, tests/test_dct8.cpp)."""
    for b in BLOCKS:
        for g in (1, 4, 16):
            cube = np.full((g, b, b), .37, np.float32)
            out, w = oracle.filter_group(cube, g, .05)
            np.testing.assert_allclose(out, cube, atol=2e-6)
            assert w == 1.0


@pytest.mark.parametrize("block", BLOCKS)
def test_single_block_pipeline(oracle, block):
    """tests/test_bm3d_semantics.py:161-173: an image of exactly one block
    with block_step = block has one reference and one match, so the output is
    the closed form for every B x G x sigma x {hard, Wiener} within 3e-5."""
    rng = np.random.default_rng(59)
    p = rng.uniform(-.02, .03, (block, block)).astype(np.float32)
    ref = (p * np.float32(.8)).astype(np.float32)
    for group in GROUPS:
        if block == 32 and group == 64:
            continue
        for user in (.5, 1, 3, 5, 10, 20, 40):
            s = oracle.effective_sigma(user)
            for wiener in (False, True):
                out = oracle.bm3d_plane(p, s, block=block, group=group, step=block, bm_range=1,
                                        wref=ref if wiener else None)
                want = cf.single_block(p.astype(np.float64), ref.astype(np.float64), group, s, wiener)
                assert np.max(np.abs(out - want)) <= 3e-5, (block, group, user, wiener)


def test_matcher_brute_force(oracle):
    """Positions equal a brute-force FP64 scan + stable sort
    (tests/test_matcher.cpp:51-89); distances within 8e-6 relative."""
    rng = np.random.default_rng(7)
    img = rng.uniform(0, 1, (37, 41)).astype(np.float32)
    for block, rg, group in ((8, 5, 8), (4, 7, 16), (12, 3, 4), (1, 2, 8), (8, 9, 64), (16, 4, 32)):
        for (bx, by) in ((0, 0), (11, 7), (40, 36), (20, 3)):
            got = oracle.spatial_match(img, bx, by, block, rg, group)
            want = cf.brute_force_matches(img, bx, by, block, rg, group)
            assert [(x, y) for x, y, _, _ in got] == [(x, y) for x, y, _ in want]
            for (_, _, _, d), (_, _, dw) in zip(got, want):
                assert abs(d - dw) <= 8e-6 * max(1.0, abs(d), abs(dw))


def test_constant_image_tie_order(oracle):
    """tests/test_matcher.cpp:133-157: on a constant image every distance is
    0 and the order is the reference, then raster (y, x)."""
    img = np.full((20, 24), .5, np.float32)
    got = oracle.spatial_match(img, 10, 9, 8, 7, 8)
    assert len(got) == 8
    want = cf.brute_force_matches(img, 10, 9, 8, 7, 8)
    assert [(x, y) for x, y, _, _ in got] == [(x, y) for x, y, _ in want]
    assert [(x, y) for x, y, _, _ in got][:3] == [(10, 9), (3, 2), (4, 2)]


def test_short_window(oracle):
    """tests/test_matcher.cpp:159-173: a frame of exactly one block has no
    candidate besides the reference."""
    img = np.full((8, 8), .25, np.float32)
    got = oracle.spatial_match(img, 0, 0, 8, 64, 8)
    assert got == [(0, 0, 0, 0.0)]


def test_cross_frame_ssd_known_answers(oracle):
    """tests/test_bm3d_contract.cpp:101-105: frames of 1, 0 and 2 (8x8),
    block 8, radius 1, bm_range 0, ps_range 0: SSD 64 / 0 / 256, and the
    clamped slot is skipped."""
    a = np.full((8, 8), 1, np.float32)
    b = np.full((8, 8), 0, np.float32)
    c = np.full((8, 8), 2, np.float32)
    got = oracle.predictive_match([a, b, c], 0, 0, 1, 8, 8, 0, 2, 0, 1)
    assert len(got) == 3
    for _, _, t, d in got:
        assert d == {0: 64.0, 1: 0.0, 2: 256.0}[t]
    got = oracle.predictive_match([a, b, c], 0, 0, 1, 8, 8, 0, 2, 0, 1, valid_begin=1)
    assert len(got) == 2


def test_predictive_motion(oracle):
    """tests/test_bm3d_contract.cpp:120-125: translated frames; no duplicate
    candidate, SSDs equal a double-precision sum, and the true motion is
    found in every frame."""
    w, h = 40, 24
    frames = []
    for t in range(5):
        shift = (t - 2) * 2
        y, x = np.mgrid[0:h, 0:w]
        frames.append(np.sin((x - shift) * .91 + y * .73).astype(np.float32))
    got = oracle.predictive_match(frames, 16, 8, 2, 4, 16, 2, 2, 2, 2)
    seen = set()
    found = [False] * 5
    for x, y, t, d in got:
        assert (t, y, x) not in seen
        seen.add((t, y, x))
        e = frames[2][8:12, 16:20].astype(np.float64) - frames[t][y:y + 4, x:x + 4]
        assert abs(d - float(np.sum(e * e))) <= 1e-4
        if y == 8 and x == 16 + (t - 2) * 2:
            found[t] = True
    assert all(found)


def test_constant_plane_identity(oracle):
    """A constant image is its own output (every AC coefficient is zero, the
    DC is kept, weights cancel)."""
    img = np.full((40, 48), .3, np.float32)
    for wiener in (False, True):
        out = oracle.bm3d_plane(img, oracle.effective_sigma(10), block=8, group=8, step=3, bm_range=4,
                                wref=img if wiener else None)
        np.testing.assert_allclose(out, img, atol=1e-6)


def test_fp32_fp64_agree(oracle):
    """The two arithmetic modes agree far above the 60 dB gate on noisy input."""
    import synth
    clean, noisy = synth.noisy_frame("GRAYS", 64, 48, 10, seed=3)
    s = oracle.effective_sigma(10)
    a = oracle.bm3d_plane(noisy[0], s, block=8, group=16, step=3, bm_range=8)
    b = oracle.bm3d_plane(noisy[0], s, block=8, group=16, step=3, bm_range=8, fp64=True)
    assert oracle.psnr(a, b) > 80
    assert oracle.psnr(a, clean[0]) > oracle.psnr(noisy[0], clean[0]) + 1.5


def test_one_frame_temporal_equals_spatial(oracle):
    """tests/test_bm3d_semantics.py:125-135: a one-frame clip at maximum
    radius has no other frame to search, so the rolling result is the
    spatial one (hard and Wiener)."""
    import nss_hip
    rng = np.random.default_rng(73)
    noisy = rng.uniform(.1, .8, (3, 12, 12)).astype(np.float32)
    for wiener in (False, True):
        kw = dict(sigma=3, block_size=4, group_size=8, block_step=2, bm_range=2)
        ref = [[p for p in noisy]] if wiener else None
        ps = nss_hip.parse("RGBS", 12, 12, 1, ref=wiener, **kw)
        spatial = oracle.bm3d_clip([list(noisy)], ps, ref_clip=ref)
        pt = nss_hip.parse("RGBS", 12, 12, 1, ref=wiener, radius=16, **kw)
        temporal = oracle.bm3d_clip([list(noisy)], pt, ref_clip=ref)
        for a, b in zip(spatial[0], temporal[0]):
            np.testing.assert_allclose(a, b, atol=1e-5)


def test_chroma_plane0_is_the_separate_filter(oracle):
    """tests/test_bm3d_chroma.py: with chroma=1 the groups are matched on
    plane 0, so plane 0 equals the plain filter's plane 0; planes 1, 2 are
    filtered with plane 0's groups."""
    import nss_hip
    import synth
    _, noisy = synth.noisy_frame("YUV444PS", 48, 40, 10, seed=5)
    pc = nss_hip.parse("YUV444PS", 48, 40, sigma=10, chroma=1)
    pp = nss_hip.parse("YUV444PS", 48, 40, sigma=10)
    c = oracle.bm3d_frame(noisy, pc)
    p = oracle.bm3d_frame(noisy, pp)
    assert np.array_equal(c[0], p[0])
    assert not np.array_equal(c[1], p[1])
    want1 = oracle.bm3d_plane(noisy[1], pc.sigma[1], block=8, group=8, step=8, bm_range=9, match=noisy[0])
    assert np.array_equal(c[1], want1)


def test_static_clip_temporal_denoises_more(oracle):
    """A static clip gives the temporal search exact copies of every patch
    in the neighbouring frames' noise: radius 1 beats the spatial filter."""
    import nss_hip
    import synth
    clean = synth.clean_image(48, 40, seed=9)
    clip = [[synth.add_noise(clean, 10, seed=100 + f)] for f in range(4)]
    p0 = nss_hip.parse("GRAYS", 48, 40, 4, sigma=10)
    p1 = nss_hip.parse("GRAYS", 48, 40, 4, sigma=10, radius=1)
    s = oracle.bm3d_clip(clip, p0)
    t = oracle.bm3d_clip(clip, p1)
    ps = np.mean([oracle.psnr(f[0], clean) for f in s])
    pt = np.mean([oracle.psnr(f[0], clean) for f in t])
    assert pt > ps + 0.5, (ps, pt)


@pytest.mark.parametrize("final", [0, 1])
def test_legacy_slices_reduce_to_rolling(oracle, final):
    """The reference adds rolling contributions in ascending centre order "as
    legacy + VAggregate" (src/host/filter_bm3d.cpp:588, :726): reducing the
    legacy slices with VAggregate's float32 sums gives the rolling frames bit
    for bit, including a plane with sigma 0 (identity slice) and the clip's
    first and last frames (slices outside the clip are zero and unread)."""
    import nss_hip
    import synth
    _, clip = synth.shifted_sequence("YUV420PS", 40, 32, 4, 10, seed=11, max_step=2)
    kw = dict(sigma=[10, 10, 0], radius=1, final=final, block_size=4, block_step=2, group_size=8, bm_range=3)
    rolling = oracle.bm3d_clip(clip, nss_hip.parse("YUV420PS", 40, 32, 4, **kw))
    pl = nss_hip.parse("YUV420PS", 40, 32, 4, temporal_mode="legacy", **kw)
    fat = oracle.bm3d_clip_legacy(clip, pl)
    assert fat[0][0].shape == (32 * 6, 40) and fat[0][1].shape == (16 * 6, 20)
    # slot 0 of the first centre and slot 2 of the last lie outside the clip
    assert not fat[0][0][:2 * 32].any() and not fat[3][0][4 * 32:].any()
    got = oracle.vaggregate(fat, clip, 1)
    for n in range(4):
        for p in range(3):
            assert np.array_equal(got[n][p], rolling[n][p]), (final, n, p)


def test_vaggregate_oracle_known_answers(oracle):
    """vaggregate_target (src/cpu/bm/agg.cpp:159-201) by hand on a 1 x 3 plane,
    radius 1, three frames: sums over the contributing centres, the source
    where the den sum is <= 1e-12, and unselected planes copied."""
    h, w, r = 1, 3, 1
    src = [[np.array([[5., 6., 7.]], np.float32)] for _ in range(3)]
    fat = []
    for c in range(3):
        f = np.zeros((2 * 3 * h, w), np.float32)
        for t in range(3):
            f[2 * t] = [c + 1, 2 * (c + 1), 0]          # num of slot t
            f[2 * t + 1] = [1, 2, 1e-13]                 # den of slot t
        fat.append([f])
    out = oracle.vaggregate(fat, src, r)
    # frame 0: centres 0, 1 -> num (1 + 2, 2 + 4), den (2, 4); third pixel falls back
    np.testing.assert_array_equal(out[0][0], [[1.5, 1.5, 7.]])
    # frame 1: centres 0, 1, 2 -> num (6, 12), den (3, 6)
    np.testing.assert_array_equal(out[1][0], [[2., 2., 7.]])
    assert np.array_equal(oracle.vaggregate(fat, src, r, planes=[])[2][0], src[2][0])


@pytest.mark.parametrize("tie", [False, True])
@pytest.mark.parametrize("group", [4, 8])
def test_temporal_slot_routing_closed_form(oracle, group, tie):
    """Rolling radius 1 on three single-block frames: the oracle's groups and
    the frame each filtered patch is summed into match the closed form
    (tests/closed_form.py temporal_single_block) within 3e-5."""
    import closed_form as cf
    b = 8
    rng = np.random.default_rng(11 + group)
    # values on a 1/64 grid: the patches, their differences and the squared
    # distances are exact in float32, so a tie is an exact tie
    p1 = rng.integers(16, 48, (b, b)) / 64.0
    d = rng.integers(-5, 6, (b, b)) / 64.0
    p0 = p1 + d
    p2 = p1 - d if tie else p1 + rng.integers(-8, 9, (b, b)) / 64.0
    frames = [np.asarray(p, np.float32) for p in (p0, p1, p2)]
    sigma = oracle.effective_sigma(20.0)
    want, groups = cf.temporal_single_block([f.astype(np.float64) for f in frames], group, sigma)
    if tie:
        assert groups[1] == [1, 0, 2]
    import nss_hip
    params = nss_hip.parse("GRAYS", b, b, 3, sigma=20.0, radius=1, block_size=b, block_step=b, group_size=group,
                           bm_range=1, ps_num=1, ps_range=0)
    got = [f[0] for f in oracle.bm3d_clip([[f] for f in frames], params)]
    for n in range(3):
        assert np.abs(got[n] - want[n]).max() <= 3e-5, (n, np.abs(got[n] - want[n]).max())
    # a routing mistake (each patch summed into the centre's own frame) is
    # far outside that tolerance
    wrong, _ = cf.temporal_single_block([f.astype(np.float64) for f in frames], group, sigma, route_to_center=True)
    assert max(np.abs(got[n] - wrong[n]).max() for n in range(3)) > 1e-3
