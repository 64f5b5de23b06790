"""Slot routing of the temporal group filters on the device: frames sharing
one random texture with a per-frame offset of 0.02 make the co-located
patches of the other frames the nearest candidates, and at a tiny sigma each
group comes back as it went in, so every slot's estimate (num / den) must be
its own frame wherever a patch landed (a patch summed into another slot
would carry that frame's offset); slots outside the clip stay empty."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["WNNM", "NCSR"])
@pytest.mark.parametrize("radius,center", [(1, 0), (1, 2), (2, 1), (2, 4)])
def test_slot_routing_offset_clip(gpu, kind, radius, center):
    import torch
    from test_oracle_temporal_groups import offset_clip
    n_frames, h, w = 5, 32, 32
    clip = offset_clip(n_frames, h, w)
    inst = getattr(gpu, kind)("GRAYS", w, h, frames=n_frames, sigma=[0.01], radius=radius, block_step=4)
    frames = inst.slot_frames(center)
    fat = inst.contribution(center, [[torch.from_numpy(clip[f]).cuda()] for f in frames])[0]
    torch.cuda.synchronize()
    fat = fat.cpu().numpy().astype(np.float64)
    inst.close()
    vb, ve = max(0, radius - center), radius + min(radius + 1, n_frames - center)
    for t in range(2 * radius + 1):
        num, den = fat[2 * t * h:(2 * t + 1) * h], fat[(2 * t + 1) * h:(2 * t + 2) * h]
        if not vb <= t < ve:
            assert not num.any() and not den.any(), t
            continue
        ok = den > 1e-12
        assert ok.mean() > 0.5, (t, ok.mean())
        err = np.abs(num[ok] / den[ok] - clip[frames[t]][ok]).max()
        print(f"{kind} r{radius} centre {center} slot {t}: max |est - frame| {err:.2e}", flush=True)
        assert err < 1e-4, (t, err)
