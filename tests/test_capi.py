"""The C-ABI library loads and exports every entry point include/nss_hip.h
declares; host-only calls behave without a device."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "nss_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nss_hip_\w+)\s*\(", text)))


def test_exports_every_declared_symbol():
    import nss_hip
    lib = nss_hip.lib()
    names = _declared()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(nss_hip.EXPORTS) == set(names)


def test_parse_defaults_and_error_text():
    import nss_hip
    p = nss_hip.parse("YUV420PS", 1920, 1080, sigma=10, final=1)
    d = p.as_dict()
    assert d["block_size"] == [8, 8, 8] and d["group_size"] == [8, 8, 8] and d["block_step"] == [8, 8, 8]
    assert d["bm_range"] == [9, 9, 9] and d["final_stage"] == 1
    assert d["sigma"][0] == pytest.approx(10 / 255 * 0.75, rel=1e-6)
    with pytest.raises(nss_hip.Error, match=r"^nss_hip\.BM3D: block_size must be one of 1, 2, 4, 8, 12, 16, 32$"):
        nss_hip.parse("GRAYS", 48, 40, block_size=3)
    assert nss_hip.last_error().startswith("nss_hip.BM3D: block_size")


def test_null_arguments_are_errors():
    import nss_hip
    lib = nss_hip.lib()
    assert lib.nss_hip_bm3d_create(None, None, 0, None) == nss_hip.INVALID_INPUT
    assert lib.nss_hip_bm3d_destroy(None) == nss_hip.OK
    assert lib.nss_hip_bm3d_set_tracing(None, 1) == nss_hip.INVALID_INPUT


def test_server_config_validation_without_device():
    """Configuration errors are reported before any device call."""
    import nss_hip
    lib = nss_hip.lib()
    h = ctypes.c_void_p()
    for cfg, msg in (((0, 0, 0, 1, 1, 0), "max_batch_frames must be in [1, 16]"),
                     ((0, 0, 17, 1, 1, 0), "max_batch_frames must be in [1, 16]"),
                     ((0, 0, 4, 0, 1, 0), "max_inflight must be in [1, 4]"),
                     ((0, -1, 4, 2, 1, 0), "hbm_budget_bytes must be >= 0")):
        c = nss_hip.ServerConfig(*cfg)
        assert lib.nss_hip_server_create(ctypes.byref(c), ctypes.byref(h)) == nss_hip.INVALID_INPUT
        assert msg in nss_hip.last_error()
    assert lib.nss_hip_server_create(None, ctypes.byref(h)) == nss_hip.INVALID_INPUT
    assert lib.nss_hip_server_destroy(None) == nss_hip.OK
    assert lib.nss_hip_server_drain(None) == nss_hip.INVALID_INPUT


def test_match_path_and_resources_argument_checks():
    """nss_hip_set_match_path takes 0, 1 or 2 and nothing else; the resource
    query rejects null arguments. Neither touches the device."""
    import nss_hip
    lib = nss_hip.lib()
    for bad in (-1, 3):
        assert lib.nss_hip_set_match_path(bad) == nss_hip.INVALID_INPUT
        assert "match path must be 0, 1 or 2" in nss_hip.last_error()
    for good in (nss_hip.MATCH_FP32, nss_hip.MATCH_MFMA, nss_hip.MATCH_AUTO):
        assert lib.nss_hip_set_match_path(good) == nss_hip.OK
    b = (ctypes.c_longlong * 6)()
    v = ctypes.c_longlong()
    assert lib.nss_hip_bm3d_resources(None, b, ctypes.byref(v), ctypes.byref(v)) == nss_hip.INVALID_INPUT


def test_wnnm_parse_messages_and_null_checks():
    """nss_hip.WNNM's argument errors carry the reference's text
    (src/frontend/wnnm_args.cpp:8-29, src/frontend/validate.hpp:19-25);
    the entry points reject null handles without touching a device."""
    import nss_hip
    lib = nss_hip.lib()
    p = nss_hip.parse_wnnm("YUV420PS", 64, 48)
    assert (p.block_size, p.block_step, p.group_size, p.bm_range, p.residual, p.adaptive) == (8, 8, 8, 7, 0, 1)
    assert list(p.sigma) == [3.0, 3.0, 3.0]
    cases = [
        (dict(block_size=17), "nss_hip.WNNM: invalid block_size/group_size/block_step/radius"),
        (dict(group_size=33), "nss_hip.WNNM: invalid block_size/group_size/block_step/radius"),
        (dict(block_step=9), "nss_hip.WNNM: invalid block_size/group_size/block_step/radius"),
        (dict(bm_range=65), "nss_hip.WNNM: bm_range must be in [1, 64]"),
        (dict(ps_num=9), "nss_hip.WNNM: invalid ps_num/ps_range"),
        (dict(sigma=[-1.0]), "nss: sigma must be non-negative"),
        (dict(temporal_mode="fused"), "nss_hip.WNNM: temporal_mode=fused is not supported; use rolling or legacy"),
    ]
    for kw, text in cases:
        with pytest.raises(nss_hip.Error) as e:
            nss_hip.parse_wnnm("GRAYS", 64, 48, **kw)
        assert str(e.value) == text
    with pytest.raises(nss_hip.Error, match="selected plane is smaller than block_size"):
        nss_hip.parse_wnnm("GRAYS", 6, 6)
    p = nss_hip.parse_wnnm("GRAYS", 64, 48, radius=1)
    assert (p.radius, p.temporal_mode) == (1, 1)
    assert nss_hip.parse_wnnm("GRAYS", 64, 48, radius=2, temporal_mode="legacy").temporal_mode == 0
    assert lib.nss_hip_wnnm_create(None, None, 0, None) == nss_hip.INVALID_INPUT
    assert lib.nss_hip_wnnm_process(None, None, None, None, None, None, None, None) == nss_hip.INVALID_INPUT
    assert lib.nss_hip_wnnm_destroy(None) == nss_hip.OK


def test_ncsr_parse_messages():
    """nss_hip.NCSR's argument errors carry the reference's text
    (src/frontend/ncsr_args.cpp:10-32)."""
    import nss_hip
    lib = nss_hip.lib()
    p = nss_hip.parse_ncsr("GRAYS", 64, 48)
    assert (p.block_size, p.block_step, p.group_size, p.bm_range, p.iters) == (8, 8, 8, 7, 2)
    assert abs(p.delta - 0.1) < 1e-7
    cases = [
        (dict(block_size=17), "nss_hip.NCSR: invalid block_size/group_size/block_step/radius"),
        (dict(iters=65), "nss_hip.NCSR: invalid iters/delta (iters in [1, 64], delta in [0, 1])"),
        (dict(delta=-0.5), "nss_hip.NCSR: invalid iters/delta (iters in [1, 64], delta in [0, 1])"),
        (dict(ps_num=9), "nss_hip.NCSR: invalid ps_num/ps_range"),
        (dict(bm_range=65), "nss_hip.NCSR: bm_range must be in [1, 64]"),
    ]
    for kw, text in cases:
        with pytest.raises(nss_hip.Error) as e:
            nss_hip.parse_ncsr("GRAYS", 64, 48, **kw)
        assert str(e.value) == text
    p = nss_hip.parse_ncsr("GRAYS", 64, 48, radius=2)
    assert (p.radius, p.temporal_mode) == (2, 1)
    assert nss_hip.parse_ncsr("GRAYS", 64, 48, radius=1, temporal_mode="legacy").temporal_mode == 0
    with pytest.raises(nss_hip.Error) as e:
        nss_hip.parse_ncsr("GRAYS", 64, 48, radius=1, temporal_mode="fused")
    assert str(e.value) == "nss_hip.NCSR: temporal_mode=fused is not supported; use rolling or legacy"
    assert lib.nss_hip_ncsr_create(None, None, 0, None) == nss_hip.INVALID_INPUT
    assert lib.nss_hip_ncsr_process(None, None, None, None, None, None, None, None) == nss_hip.INVALID_INPUT
    assert lib.nss_hip_ncsr_process_temporal(None, 0, None, None, None, None, None, None, None) == nss_hip.INVALID_INPUT
    assert lib.nss_hip_ncsr_destroy(None) == nss_hip.OK
