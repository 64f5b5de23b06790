"""Temporal (predictive) match lists of the HIP path against the oracle,
through nss_hip_predictive_match, as the reference compares its device lists
(tests/cuda/test_cuda_match.cpp:327-345): same counts; an entry may differ
only as a near tie |a - b| <= 2e-5 max(1, |a|, |b|) (none on exact-tie
clips); and the slot routing of the rolling sums pinned by a closed form."""
import numpy as np
import pytest

import closed_form as cf

pytestmark = pytest.mark.gpu


def _near_tie(a, b):
    return abs(a - b) <= 2e-5 * max(1.0, abs(a), abs(b))


def _clip(n, w, h, seed, flat):
    """n frames of a base plane shifted by cumulative integer steps; flat:
    piecewise-constant 0 / 0.25 / 0.5 blocks (exact ties everywhere)."""
    rng = np.random.default_rng(seed)
    H, W = h + 40, w + 40
    y, x = np.mgrid[0:H, 0:W]
    if flat:
        base = ((((x // 8) + (y // 8) + seed) % 3) * 0.25).astype(np.float32)
    else:
        base = (0.5 + 0.3 * np.sin(x * 0.21 + seed) * np.cos(y * 0.17)).astype(np.float32)
    frames, ox, oy = [], 20, 20
    for _ in range(n):
        f = base[oy:oy + h, ox:ox + w].copy()
        if not flat:
            f = (f + rng.normal(0, 0.05, f.shape)).astype(np.float32)
        frames.append(f)
        ox = int(np.clip(ox + rng.integers(-3, 4), 0, 40))
        oy = int(np.clip(oy + rng.integers(-3, 4), 0, 40))
    return frames


CASES = [  # radius, ps_num, ps_range, group, block, step, bm_range
    (1, 2, 4, 8, 8, 8, 9),
    (1, 1, 2, 16, 8, 4, 6),
    (2, 3, 5, 16, 8, 6, 7),
    (2, 2, 3, 32, 8, 8, 5),
    (1, 2, 16, 8, 8, 8, 4),   # windows too wide to stage: scored from the frame
    (2, 2, 2, 8, 4, 4, 5),
    (1, 3, 4, 8, 12, 6, 6),
]


@pytest.mark.parametrize("flat", [False, True])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_predictive_match_lists(gpu, oracle, case, flat):
    import torch
    r, psn, psr, group, block, step, rng_ = CASES[case]
    n, w, h = 5, 72, 56
    frames = _clip(n, w, h, seed=3 + case, flat=flat)
    dev = torch.from_numpy(np.stack(frames)).cuda()
    grid = oracle.raster_grid(w, h, block, step)
    mismatched = checked = 0
    for center in (0, 2, n - 1):
        rec, cnt = gpu.predictive_match(dev, center, r, block=block, step=step, bm_range=rng_, group=group,
                                        ps_num=psn, ps_range=psr)
        torch.cuda.synchronize()
        rec, cnt = rec.cpu().numpy(), cnt.cpu().numpy()
        slots = [center + min(max(t - r, -center), n - 1 - center) for t in range(2 * r + 1)]
        vb, ve = max(0, r - center), r + min(r + 1, n - center)
        planes = [frames[f] for f in slots]
        for i, (bx, by) in enumerate(grid):
            want = oracle.predictive_match(planes, bx, by, r, block, group, rng_, psn, psr, r, vb, ve)
            assert cnt[i] == len(want), (center, i, cnt[i], len(want))
            got = [(int(a), int(b_), int(t), float(np.array(d, np.int32).view(np.float32)))
                   for a, b_, t, d in rec[i, :cnt[i]]]
            checked += 1
            ok = True
            for (x, y, t, d), (gx, gy, gt, gd) in zip(want, got):
                if not _near_tie(d, gd):
                    ok = False
                elif (x, y, t) != (gx, gy, gt) and flat:
                    ok = False
            if not ok:
                # Only on noisy clips: a seed or layer winner chosen
                # differently between near ties leads the next layer's
                # windows elsewhere. Exact-tie clips must agree entry by entry.
                assert not flat, (center, i, want, got)
                mismatched += 1
    print(f"case {case} flat={flat}: {checked} groups, {mismatched} diverged after a near tie", flush=True)
    assert mismatched <= max(1, checked // 100)


@pytest.mark.parametrize("tie", [False, True])
@pytest.mark.parametrize("group", [4, 8])
def test_temporal_slot_routing_closed_form(gpu, oracle, group, tie):
    """Three single-block frames at radius 1: each centre's group and the
    frame every filtered patch is summed into are fixed by the closed form
    (tie: frames 0 and 2 at equal distance from frame 1, ordered by slot)."""
    import torch
    b, sigma_user = 8, 20.0
    rng = np.random.default_rng(11 + group)
    # values on a 1/64 grid: the patches, their differences and the squared
    # distances are exact in float32, so a tie is an exact tie
    p1 = rng.integers(16, 48, (b, b)) / 64.0
    d = rng.integers(-5, 6, (b, b)) / 64.0
    p0 = p1 + d
    p2 = p1 - d if tie else p1 + rng.integers(-8, 9, (b, b)) / 64.0
    frames = [np.asarray(p, np.float32) for p in (p0, p1, p2)]
    sigma = oracle.effective_sigma(sigma_user)
    want, groups = cf.temporal_single_block([f.astype(np.float64) for f in frames], group, sigma)
    assert groups[1] == ([1, 0, 2] if tie else groups[1]) and len(groups[0]) == 2 and len(groups[2]) == 2
    kw = dict(sigma=sigma_user, radius=1, block_size=b, block_step=b, group_size=group, bm_range=1, ps_num=1,
              ps_range=0)
    inst = gpu.BM3D("GRAYS", b, b, 3, **kw)
    got = [f[0].cpu().numpy() for f in inst.filter_clip([[torch.from_numpy(f).cuda()] for f in frames])]
    torch.cuda.synchronize()
    inst.close()
    params = gpu.parse("GRAYS", b, b, 3, **kw)
    orc = [f[0] for f in oracle.bm3d_clip([[f] for f in frames], params)]
    for n in range(3):
        assert np.abs(got[n] - want[n]).max() <= 3e-5, (n, np.abs(got[n] - want[n]).max())
        assert np.abs(orc[n] - want[n]).max() <= 3e-5, (n, np.abs(orc[n] - want[n]).max())
