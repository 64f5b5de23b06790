"""The output does not depend on the launch plan: every filter tile shape the
stage fits, and the autotuned plan, give the same bits, in all three
transform forms. Tile partials are exact int64 sums (2^32 units) that the
finish adds exactly and converts once, as the reference's fixed-point
accumulation does (src/cuda/common/fixed_accumulate.cuh:18-35), so the
memory limit, the autotuner and batching may change the plan and never the
output (docs/api/cuda/README.md:140-141)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(8, 4), (4, 4), (4, 2), (8, 2), (16, 2), (2, 2), (1, 1), (2, 1), (1, 4)]


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def _run(inst, src):
    import torch
    out = inst(src)
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in out]


CASES = [
    ("YUV420PS", 200, 136, dict(sigma=10, final=1)),
    ("GRAYS", 160, 112, dict(sigma=5, group_size=16, block_step=3, bm_range=16)),
    ("GRAYS", 144, 96, dict(sigma=20, final=1, group_size=32, group_size_basic=16, block_step=4)),
    ("RGBS", 96, 80, dict(sigma=10, block_size=4, group_size=8, block_step=2, bm_range=6)),
    ("GRAYS", 120, 88, dict(sigma=10, block_size=12, group_size=16, block_step=6, final=1)),
]


@pytest.mark.parametrize("transform", [0, 1, 2])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_output_independent_of_tile_plan(gpu, case, transform):
    import synth
    fmt, w, h, kw = CASES[case]
    _, noisy = synth.noisy_frame(fmt, w, h, 10, seed=31 + case)
    src = [_dev(p) for p in noisy]
    inst = gpu.BM3D(fmt, w, h, transform=transform, **kw)
    base = _run(inst, src)
    nstages = 2 if kw.get("final") else 1
    tried = 0
    for stage in range(nstages):
        for tw, th in SHAPES:
            if not inst.set_tiles(stage, tw, th):
                continue
            tried += 1
            got = _run(inst, src)
            for g, b in zip(got, base):
                assert np.array_equal(g, b), (stage, tw, th)
    inst.autotune(repeats=1)
    for g, b in zip(_run(inst, src), base):
        assert np.array_equal(g, b), "autotuned plan"
    inst.close()
    print(f"case {case} transform {transform}: {tried} tile shapes, all identical", flush=True)
    assert tried >= 2 * nstages
