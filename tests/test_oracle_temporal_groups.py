"""Pins of the temporal group-filter oracles (oracle/oracle.py
wnnm_contribution, ncsr_contribution), on the CPU.

- Slot routing: frames that share one random texture and differ by a
  per-frame offset (0.02 per frame) make the co-located patches of the other
  frames the nearest candidates (distance 64 x 0.0004 against about 10 for
  any other patch), so groups span the slots; at a tiny sigma every group
  comes back as it went in, so each slot's estimate (num / den) must equal
  its own frame wherever a patch landed. A patch added to another slot's
  sums would show that frame's offset. The slots outside the clip stay
  empty, and the slots inside it are reached.
- A one-frame clip: only the centre slot is inside the clip, and its
  num / den is the spatial filter's output (wnnm_plane / ncsr_plane).
"""
import numpy as np
import pytest

import oracle as orc


def offset_clip(n_frames, h=32, w=32, seed=5):
    rng = np.random.default_rng(seed)
    tex = rng.random((h, w)).astype(np.float32)
    return [(tex + np.float32(0.02 * f)).astype(np.float32) for f in range(n_frames)]


def slot_frames(center, radius, n_frames):
    return [center + min(max(t - radius, -center), n_frames - 1 - center) for t in range(2 * radius + 1)]


def contribution(kind, slots, center, n_frames, sigma, **kw):
    fn = orc.wnnm_contribution if kind == "wnnm" else orc.ncsr_contribution
    return fn(slots, center, n_frames, sigma, **kw)


@pytest.mark.parametrize("kind", ["wnnm", "ncsr"])
@pytest.mark.parametrize("radius,center", [(1, 0), (1, 2), (2, 1), (2, 4)])
def test_slot_routing_offset_clip(kind, radius, center):
    n_frames = 5
    clip = offset_clip(n_frames)
    frames = slot_frames(center, radius, n_frames)
    con = contribution(kind, [clip[f] for f in frames], center, n_frames, 0.01, step=4)
    vb, ve = max(0, radius - center), radius + min(radius + 1, n_frames - center)
    for t in range(2 * radius + 1):
        num, den = con[t, 0], con[t, 1]
        if not vb <= t < ve:
            assert not num.any() and not den.any(), t
            continue
        ok = den > 1e-12
        assert ok.mean() > 0.5, (t, ok.mean())
        est = num[ok] / den[ok]
        assert np.abs(est - clip[frames[t]][ok]).max() < 1e-5, t


@pytest.mark.parametrize("kind", ["wnnm", "ncsr"])
def test_one_frame_clip_is_the_spatial_filter(kind):
    rng = np.random.default_rng(9)
    yy, xx = np.mgrid[0:32, 0:40]
    plane = (0.5 + 0.3 * np.sin(xx / 5.0) * np.cos(yy / 4.0) + rng.normal(0, 10 / 255, (32, 40))).astype(np.float32)
    con = contribution(kind, [plane] * 3, 0, 1, 10.0)
    assert not con[[0, 2]].any()
    num, den = con[1, 0], con[1, 1]
    est = np.where(den > 1e-12, num / np.where(den > 1e-12, den, 1.0), plane)
    want = orc.wnnm_plane(plane, 10.0) if kind == "wnnm" else orc.ncsr_plane(plane, 10.0)
    assert np.abs(est - want).max() < 1e-6
