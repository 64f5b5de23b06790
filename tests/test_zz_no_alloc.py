"""The no-allocation frame filter (nss_hip_bm3d_process_noalloc) against the
instance entry point (nss_hip_bm3d_create + nss_hip_bm3d_process), through
the C ABI: the same kernels and plan, so the output bits are equal. Calls in
a row with shrinking inputs (the shared working array keeps the previous
call's contents), and a frame beyond the array's size is refused before any
work is queued."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def _noalloc(gpu, fmt, w, h, planes, **kw):
    import torch
    params = gpu.parse(fmt, w, h, **kw)
    vi = gpu.video_info(fmt, w, h)
    src = [_dev(p) for p in planes]
    out = [torch.full_like(p, float("nan")) for p in src]
    n = len(src)
    VP, IP = ctypes.c_void_p * n, ctypes.c_int * n
    rc = gpu.lib().nss_hip_bm3d_process_noalloc(
        ctypes.byref(params), ctypes.byref(vi), 0, VP(*[p.data_ptr() for p in src]), IP(*[p.stride(0) for p in src]),
        None, None, VP(*[p.data_ptr() for p in out]), IP(*[p.stride(0) for p in out]),
        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == gpu.OK, gpu.last_error()
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in out]


def _instance(gpu, fmt, w, h, planes, **kw):
    import torch
    inst = gpu.BM3D(fmt, w, h, **kw)
    out = inst([_dev(p) for p in planes])
    torch.cuda.synchronize()
    res = [o.cpu().numpy() for o in out]
    inst.close()
    return res


# Largest first: every later call runs on a working array that still holds
# the earlier, larger frame's records, partials and basic estimate.
CASES = [
    # beyond the array on the tile plan: batched over grid rows in it
    ("GRAYS", 3840, 2160, dict(sigma=20, final=1)),
    ("YUV420PS", 1920, 1080, dict(sigma=10, final=1)),
    ("GRAYS", 256, 256, dict(sigma=5, block_step=3, group_size=16, bm_range=16)),
    ("RGBS", 160, 120, dict(sigma=10, final=1, chroma=0, block_size=16, group_size=16, block_step=8, bm_range=12)),
    ("GRAYS", 96, 64, dict(sigma=20, block_size=4, group_size=32, block_step=2, bm_range=5, final=1)),
]


def test_noalloc_equals_instance_shrinking_inputs(gpu):
    import synth
    for k, (fmt, w, h, kw) in enumerate(CASES):
        print(f"noalloc case {k}: {fmt} {w}x{h} {kw}", flush=True)
        _, noisy = synth.noisy_frame(fmt, w, h, kw["sigma"], seed=70 + k)
        want = _instance(gpu, fmt, w, h, noisy, **kw)
        got = _noalloc(gpu, fmt, w, h, noisy, **kw)
        for i, (g, w_) in enumerate(zip(got, want)):
            assert np.isfinite(g).all(), (k, i)
            assert np.array_equal(g, w_), (k, i, float(np.abs(g - w_).max()))


def test_noalloc_twice_in_a_row_second_smaller(gpu):
    """Two calls back to back without a host wait between them: 1080p YUV420
    two-stage, then a smaller gray frame."""
    import synth
    a = ("YUV420PS", 1920, 1080, dict(sigma=10, final=1))
    b = ("GRAYS", 200, 136, dict(sigma=5, group_size=16, block_step=4))
    _, na = synth.noisy_frame(a[0], a[1], a[2], 10, seed=91)
    _, nb = synth.noisy_frame(b[0], b[1], b[2], 5, seed=92)
    want_a = _instance(gpu, a[0], a[1], a[2], na, **a[3])
    want_b = _instance(gpu, b[0], b[1], b[2], nb, **b[3])
    for rep in range(2):
        got_a = _noalloc(gpu, a[0], a[1], a[2], na, **a[3])
        got_b = _noalloc(gpu, b[0], b[1], b[2], nb, **b[3])
        for g, w_ in zip(got_a + got_b, want_a + want_b):
            assert np.array_equal(g, w_), rep


def test_noalloc_refuses_frames_beyond_its_array(gpu):
    import torch
    params = gpu.parse("RGBS", 3840, 2160, sigma=10, final=1)
    vi = gpu.video_info("RGBS", 3840, 2160)
    planes = [torch.zeros(2160, 3840, device="cuda") for _ in range(3)]
    ptrs = (ctypes.c_void_p * 3)(*[p.data_ptr() for p in planes])
    strides = (ctypes.c_int * 3)(3840, 3840, 3840)
    rc = gpu.lib().nss_hip_bm3d_process_noalloc(ctypes.byref(params), ctypes.byref(vi), 0, ptrs, strides, None,
                                                None, ptrs, strides, None)
    torch.cuda.synchronize()
    assert rc == gpu.UNSUPPORTED
    assert "MiB of working memory even in batches" in gpu.last_error()
