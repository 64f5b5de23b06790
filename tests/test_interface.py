"""Replays the reference's 900 BM3D creation probes (tests/golden/
plugin_interface_bm3d_v1.json, from tests/data/plugin_interface_v1.json)
against nss_hip's argument resolution through the C ABI, with the rules the
reference harness applies to a device backend (tests/test_plugin_interface.py):
- "nss." in an expected error becomes "nss_hip."; an error beginning "nss: "
  may read "nss_hip: ";
- the device default temporal mode is rolling, so a probe that leaves
  temporal_mode unset is made in legacy mode (:112-115), and the default is
  checked separately (:251-263);
- the cache is adaptive: a probe the CPU rejects with "use only one of
  rolling_cache_limit and rolling_cache_chunks" must be accepted (:264-279).
"""
import ast
import json
import os
import re

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "plugin_interface_bm3d_v1.json")
INT_PROBES = (-1, 0, 1, 2, 3, 4, 7, 8, 9, 16, 17, 33, 65, 1000)
FLOAT_PROBES = (-1.0, 0.0, 0.5, 1.0, 3.0, 1000.0)
CLIPS = {"GRAYS": (48, 40), "RGBS": (48, 40), "YUV420PS": (48, 40)}
BOTH = "use only one of rolling_cache_limit and rolling_cache_chunks"


def _golden():
    with open(GOLDEN) as f:
        return json.load(f)


def _split_combo(spec):
    """'a=1,b=[8, 3],c='x'' -> dict (keys are identifiers followed by '=')."""
    starts = [m.start(1) for m in re.finditer(r"(?:^|,)([a-z_]+)=", spec)]
    out = {}
    for i, s in enumerate(starts):
        end = starts[i + 1] - 1 if i + 1 < len(starts) else len(spec)
        key, value = spec[s:end].split("=", 1)
        out[key] = ast.literal_eval(value)
    return out


def _kwargs(spec, types):
    if spec == "defaults":
        return {}
    if spec.startswith("combo:"):
        return _split_combo(spec[len("combo:"):])
    key, value = spec.split("=", 1)
    if value == "pair":
        probes = INT_PROBES if types[key] == "int" else FLOAT_PROBES
        return {key: [probes[4], probes[3]]}
    return {key: ast.literal_eval(value)}


def _outcome(nss_hip, fmt, kwargs):
    w, h = CLIPS[fmt]
    if not kwargs.get("temporal_mode"):
        kwargs = dict(kwargs, temporal_mode="legacy")
    try:
        info = nss_hip.video_info(fmt, w, h, 4)
        p = nss_hip.parse(fmt, w, h, 4, info=info, **kwargs)
        o = nss_hip.output_info(p, info)
        return {"ok": {"format": nss_hip.FORMAT_NAMES[fmt], "frames": o.num_frames, "height": o.height,
                       "width": o.width}}
    except nss_hip.Error as e:
        return {"error": str(e)}


def _same(expected, actual):
    if expected == actual:
        return True
    err = expected.get("error", "")
    if err.startswith("nss.BM3D: ") and actual == {"error": "nss_hip.BM3D: " + err[len("nss.BM3D: "):]}:
        return True
    return err.startswith("nss: ") and actual == {"error": "nss_hip: " + err[len("nss: "):]}


def test_signature_is_reference_plus_backend_tail():
    import nss_hip
    g = _golden()
    assert nss_hip.BM3D_SIGNATURE == (g["signature"] + "device_id:int:opt;num_streams:int:opt;"
                                      "transform:int:opt;")


def test_replay_900_cases():
    import nss_hip
    g = _golden()
    types = nss_hip._TYPES
    assert len(g["cases"]) == 900
    problems = []
    for key, expected in sorted(g["cases"].items()):
        _, fmt, spec = key.split("|", 2)
        actual = _outcome(nss_hip, fmt, _kwargs(spec, types))
        if BOTH in expected.get("error", ""):
            if "ok" not in actual:
                problems.append(f"{key}: adaptive cache arguments rejected: {actual}")
            continue
        if not _same(expected, actual):
            problems.append(f"{key}: {expected} -> {actual}")
    assert not problems, "\n".join(problems[:40]) + f"\n({len(problems)} differences)"


def test_default_temporal_mode_is_rolling():
    """radius > 0 without temporal_mode returns finished frames (the output
    height is the clip's)."""
    import nss_hip
    info = nss_hip.video_info("GRAYS", 48, 40, 4)
    p = nss_hip.parse("GRAYS", 48, 40, 4, info=info, radius=1)
    assert p.temporal_mode == 1
    assert nss_hip.output_info(p, info).height == 40


def test_adaptive_cache_limit_below_chunks_rejected():
    import nss_hip
    with pytest.raises(nss_hip.Error, match="rolling_cache_limit must be at least rolling_cache_chunks"):
        nss_hip.parse("GRAYS", 48, 40, 4, radius=1, temporal_mode="rolling", rolling_cache_chunks=4,
                      rolling_cache_limit=2)


@pytest.mark.parametrize("key,value,msg", [
    ("device_id", -1, "nss_hip.BM3D: device_id must be >= 0"),
    ("num_streams", 0, "nss_hip.BM3D: num_streams must be in [1, 16]"),
    ("num_streams", 17, "nss_hip.BM3D: num_streams must be in [1, 16]"),
    ("transform", 3, "nss_hip.BM3D: transform must be 0 (fp32), 1 (bf16x3) or 2 (fp16)"),
    ("transform", -1, "nss_hip.BM3D: transform must be 0 (fp32), 1 (bf16x3) or 2 (fp16)"),
])
def test_backend_tail(key, value, msg):
    import nss_hip
    with pytest.raises(nss_hip.Error) as e:
        nss_hip.parse("GRAYS", 48, 40, 4, **{key: value})
    assert str(e.value) == msg


def test_numeric_validation_order():
    """src/frontend/validate.cpp:13-46: an int beyond int32 or a non-finite
    float is reported first, by sorted key."""
    import nss_hip
    with pytest.raises(nss_hip.Error) as e:
        nss_hip.parse("GRAYS", 48, 40, 4, sigma=float("nan"), block_size=2 ** 40)
    assert str(e.value) == "nss_hip: parameter block_size[0] is not representable as a finite int"
    with pytest.raises(nss_hip.Error) as e:
        nss_hip.parse("GRAYS", 48, 40, 4, sigma=[3, float("inf")])
    assert str(e.value) == "nss_hip: parameter sigma[1] is not representable as a finite float"
    with pytest.raises(nss_hip.Error) as e:
        nss_hip.parse("GRAYS", 48, 40, 4, memory_limit_mb=0)
    assert str(e.value) == "nss_hip: memory_limit_mb must be in [1, 1048576]"
