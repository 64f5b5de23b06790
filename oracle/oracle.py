"""CPU oracle for the BM3D hot path: ctypes wrapper over ``bm3d_oracle.cpp``.

TEST INFRASTRUCTURE ONLY. Only the tests, the smoke check and the benchmark's
CPU-baseline leg use this module, as the reference that results are compared
against. The product library never imports it.

Frame-level semantics follow the reference CPU plugin:
- per plane, sigma == 0 copies the plane (src/host/filter_bm3d.cpp:402-412);
- ``final=1`` is the basic (hard-threshold) stage with the ``*_basic``
  values, then the Wiener stage with that estimate as ``ref``
  (src/frontend/bm3d_args.cpp:186-221 bm3d_two_nodes);
- matching runs on ``ref`` when one is given (src/host/filter_bm3d.cpp:351,
  process_plane_batched passes refs as the match planes);
- radius > 0 ("rolling"): each centre frame's 2r+1 num/den slices are added
  into their target frames in ascending centre order, then finished
  (src/host/filter_bm3d.cpp:730-851, src/frontend/temporal.hpp:11-36).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_bm3d.so")
_lib = None

MAX_GROUP = 64


def build(force: bool = False) -> str:
    """Compile the oracle with g++ (-O3, no fast-math)."""
    src = os.path.join(_HERE, "bm3d_oracle.cpp")
    if force or not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle_bm3d.so"])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        f32p, i32p = P(ctypes.c_float), P(ctypes.c_int)
        fpp = P(P(ctypes.c_float))
        ci, cf = ctypes.c_int, ctypes.c_float
        L.orc_spatial_match.argtypes = [f32p, ci, ci, ci, ci, ci, ci, ci, ci, ci, i32p, f32p]
        L.orc_spatial_match.restype = ci
        L.orc_predictive_match.argtypes = [fpp, ci, ci, ci, ci, ci, ci, ci, ci, ci, ci, ci, ci, ci, ci, ci, ci,
                                           i32p, f32p]
        L.orc_predictive_match.restype = ci
        L.orc_filter_group.argtypes = [f32p, ci, ci, ci, cf, ci, f32p, ci]
        L.orc_filter_group.restype = cf
        L.orc_bm3d_plane.argtypes = [fpp, fpp, fpp, ci, ci, ci, ci, ci, f32p, ci, f32p, f32p, cf,
                                     ci, ci, ci, ci, ci, ci, ci, ci, ci, ci, ci, i32p]
        L.orc_bm3d_plane.restype = ci
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _ptrs(arrs):
    t = ctypes.POINTER(ctypes.c_float) * len(arrs)
    return t(*[_fp(a) for a in arrs])


def effective_sigma(user: float) -> float:
    """include/nss/contracts.hpp:27-29: (sigma * (1/255)) * 0.75 in float."""
    return float(np.float32(np.float32(user) * np.float32(1.0 / 255.0)) * np.float32(0.75))


def spatial_match(plane, bx, by, block, bm_range, group, fp64=False):
    """Match list of one reference: list of (x, y, t, dist), reference first."""
    plane = np.ascontiguousarray(plane, np.float32)
    h, w = plane.shape
    xyt = np.zeros((MAX_GROUP, 3), np.int32)
    dist = np.zeros(MAX_GROUP, np.float32)
    n = lib().orc_spatial_match(_fp(plane), w, w, h, bx, by, block, bm_range, group, int(fp64), _ip(xyt), _fp(dist))
    return [(int(xyt[i, 0]), int(xyt[i, 1]), int(xyt[i, 2]), float(dist[i])) for i in range(n)]


def predictive_match(planes, bx, by, t0, block, group, bm_range, ps_num, ps_range, radius,
                     valid_begin=0, valid_end=-1, fp64=False):
    planes = [np.ascontiguousarray(p, np.float32) for p in planes]
    h, w = planes[0].shape
    xyt = np.zeros((MAX_GROUP, 3), np.int32)
    dist = np.zeros(MAX_GROUP, np.float32)
    n = lib().orc_predictive_match(_ptrs(planes), w, len(planes), w, h, bx, by, t0, block, group, bm_range, ps_num,
                                   ps_range, radius, valid_begin, valid_end, int(fp64), _ip(xyt), _fp(dist))
    return [(int(xyt[i, 0]), int(xyt[i, 1]), int(xyt[i, 2]), float(dist[i])) for i in range(n)]


def filter_group(cube, k, sigma, wiener=False, ref_cube=None, fp64=False):
    """Filter a (group, block, block) cube in place semantics; returns (cube, weight)."""
    c = np.ascontiguousarray(cube, np.float32).copy()
    g, b, _ = c.shape
    r = np.ascontiguousarray(ref_cube if ref_cube is not None else np.zeros_like(c), np.float32).copy()
    w = lib().orc_filter_group(_fp(c), g, k, b, sigma, int(wiener), _fp(r), int(fp64))
    return c, float(w)


def raster_grid(width, height, block, step):
    """Reference positions (src/host/batch_runner.hpp:17-31)."""
    xs = [min(x, width - block) for x in range(0, width - block + step, step)]
    ys = [min(y, height - block) for y in range(0, height - block + step, step)]
    return [(x, y) for y in ys for x in xs]


@dataclass
class PlaneArgs:
    """One plane's resolved arguments for one stage (effective sigma)."""
    sigma: float
    block: int = 8
    group: int = 8
    step: int = 8
    bm_range: int = 9
    ps_num: int = 2
    ps_range: int = 4


def bm3d_plane(src, sigma, block=8, group=8, step=8, bm_range=9, ps_num=2, ps_range=4, wref=None,
               match=None, fp64=False, want_matches=False):
    """One spatial plane of one stage. src/wref/match: 2-D float32 arrays.

    Matching runs on ``match`` (defaults to ``wref`` when given, else ``src``).
    Returns the finished plane (and the match dump when asked)."""
    return bm3d_plane_temporal([src], 0, sigma, block, group, step, bm_range, ps_num, ps_range, 0,
                               wrefs=None if wref is None else [wref],
                               matches=None if match is None else [match], center=0, frame_count=1,
                               fp64=fp64, want_matches=want_matches)


def bm3d_plane_temporal(srcs, t0, sigma, block, group, step, bm_range, ps_num, ps_range, radius, wrefs=None,
                        matches=None, center=0, frame_count=1, fp64=False, want_matches=False, want_slices=False):
    srcs = [np.ascontiguousarray(s, np.float32) for s in srcs]
    h, w = srcs[0].shape
    wiener = wrefs is not None
    wr = [np.ascontiguousarray(s, np.float32) for s in wrefs] if wiener else srcs
    mt = [np.ascontiguousarray(s, np.float32) for s in matches] if matches is not None else (wr if wiener else srcs)
    slices = 2 * radius + 1
    dst = np.zeros((h, w), np.float32)
    num = np.zeros((slices, h, w), np.float32) if want_slices else None
    den = np.zeros((slices, h, w), np.float32) if want_slices else None
    nref = len(raster_grid(w, h, block, step))
    dump = np.zeros((nref, 1 + 3 * MAX_GROUP), np.int32) if want_matches else None
    rc = lib().orc_bm3d_plane(_ptrs(mt), _ptrs(srcs), _ptrs(wr), len(srcs), t0, w, h, w, _fp(dst), w,
                              _fp(num) if want_slices else None, _fp(den) if want_slices else None,
                              sigma, block, group, step, bm_range, ps_num, ps_range, radius, int(wiener),
                              center, frame_count, int(fp64), _ip(dump) if want_matches else None)
    if rc < 0:
        raise RuntimeError("oracle: matching failed for an active group")
    out = [dst]
    if want_matches:
        out.append(dump)
    if want_slices:
        out += [num, den]
    return out[0] if len(out) == 1 else tuple(out)


def _stage_args(params, plane, basic):
    p = params
    if basic:
        return dict(sigma=p.sigma_basic[plane], block=p.block_size_basic[plane], group=p.group_size_basic[plane],
                    step=p.block_step_basic[plane], bm_range=p.bm_range[plane], ps_num=p.ps_num_basic[plane],
                    ps_range=p.ps_range[plane])
    return dict(sigma=p.sigma[plane], block=p.block_size[plane], group=p.group_size[plane],
                step=p.block_step[plane], bm_range=p.bm_range[plane], ps_num=p.ps_num[plane],
                ps_range=p.ps_range[plane])


def _is_final(params):
    """final=1 (named final_stage in the C ABI's parameter struct)."""
    return bool(getattr(params, "final_stage", getattr(params, "final", False)))


def _stage_frame(planes, params, basic, wref, match, fp64):
    """One spatial stage over a frame's planes. With chroma (CBM3D) every
    plane is matched on plane 0 of ``match`` (src/host/filter_bm3d.cpp:451-486)."""
    chroma = bool(getattr(params, "chroma", False))
    out = []
    for i, p in enumerate(planes):
        a = _stage_args(params, i, basic)
        if a["sigma"] == 0:
            out.append(np.array(p, np.float32, copy=True))
            continue
        out.append(bm3d_plane(p, wref=None if wref is None else wref[i], match=match[0 if chroma else i],
                              fp64=fp64, **a))
    return out


def bm3d_frame(planes, params, ref=None, fp64=False):
    """Spatial (radius 0) nss.BM3D on one frame: list of 2-D planes in, list out.

    ``params`` carries resolved per-plane values (effective sigma) with the
    attribute names of Bm3dParams (include/nss/params/bm3d.hpp:15-31); ``ref``
    is the planes of the ref clip's frame when one is given."""
    if params.radius != 0:
        raise ValueError("bm3d_frame is the spatial path; use bm3d_clip for radius > 0")
    if _is_final(params):
        ref = _stage_frame(planes, params, True, None, planes, fp64)
    return _stage_frame(planes, params, False, ref, ref if ref is not None else planes, fp64)


def _temporal_stage(clip, params, basic, wref_clip, match_clip, fp64):
    """Rolling temporal stage over a whole clip ([frame][plane] 2-D arrays)."""
    n_frames = len(clip)
    r = params.radius
    chroma = bool(getattr(params, "chroma", False))
    out = [[None] * len(clip[0]) for _ in clip]
    for p in range(len(clip[0])):
        a = _stage_args(params, p, basic)
        frames = [f[p] for f in clip]
        if a["sigma"] == 0:
            for n in range(n_frames):
                out[n][p] = np.array(frames[n], np.float32, copy=True)
            continue
        mplane = 0 if chroma else p
        h, w = frames[0].shape
        dt = np.float64 if fp64 else np.float32
        acc_num = [np.zeros((h, w), dt) for _ in range(n_frames)]
        acc_den = [np.zeros((h, w), dt) for _ in range(n_frames)]
        for c in range(n_frames):
            slot_frames = [c + min(max(t - r, -c), n_frames - 1 - c) for t in range(2 * r + 1)]
            srcs = [frames[f] for f in slot_frames]
            wr = [wref_clip[f][p] for f in slot_frames] if wref_clip is not None else None
            mt = [match_clip[f][mplane] for f in slot_frames]
            _, num, den = bm3d_plane_temporal(srcs, r, a["sigma"], a["block"], a["group"], a["step"], a["bm_range"],
                                              a["ps_num"], a["ps_range"], r, wrefs=wr, matches=mt, center=c,
                                              frame_count=n_frames, fp64=fp64, want_slices=True)
            for t in range(2 * r + 1):
                target = c + t - r
                if 0 <= target < n_frames:
                    acc_num[target] += num[t]
                    acc_den[target] += den[t]
        for n in range(n_frames):
            d = acc_den[n]
            src = frames[n].astype(dt)
            res = np.where(d > 1e-12, acc_num[n] / np.where(d > 1e-12, d, 1), src)
            out[n][p] = res.astype(np.float32)
    return out


def bm3d_clip(clip, params, ref_clip=None, fp64=False):
    """nss.BM3D over a clip ([frame][plane] 2-D arrays); radius > 0 gives the
    rolling (finished) frames."""
    if params.radius == 0:
        return [bm3d_frame(f, params, ref=None if ref_clip is None else ref_clip[n], fp64=fp64)
                for n, f in enumerate(clip)]
    if _is_final(params):
        basic = _temporal_stage(clip, params, True, None, clip, fp64)
        return _temporal_stage(clip, params, False, basic, basic, fp64)
    if ref_clip is not None:
        return _temporal_stage(clip, params, False, ref_clip, ref_clip, fp64)
    return _temporal_stage(clip, params, False, None, clip, fp64)


def _legacy_stage(clip, params, wref_clip, match_clip, fp64):
    """Legacy temporal output of the last stage: per centre frame c and plane,
    the 2r+1 num/den slice pairs stacked vertically (slot t's num at rows
    2 t h, den at (2 t + 1) h; src/host/filter_bm3d.cpp:322-334), zeros for
    slots whose frame lies outside the clip; planes with sigma 0 get the
    identity contribution (src/frontend/temporal.hpp:23-36)."""
    n_frames = len(clip)
    r = params.radius
    ns = 2 * r + 1
    chroma = bool(getattr(params, "chroma", False))
    out = [[None] * len(clip[0]) for _ in clip]
    for p in range(len(clip[0])):
        a = _stage_args(params, p, False)
        frames = [f[p] for f in clip]
        h, w = frames[0].shape
        mplane = 0 if chroma else p
        for c in range(n_frames):
            fat = np.zeros((2 * ns * h, w), np.float32)
            if a["sigma"] == 0:
                fat[2 * r * h:(2 * r + 1) * h] = frames[c]
                fat[(2 * r + 1) * h:(2 * r + 2) * h] = 1.0
                out[c][p] = fat
                continue
            slot_frames = [c + min(max(t - r, -c), n_frames - 1 - c) for t in range(ns)]
            srcs = [frames[f] for f in slot_frames]
            wr = [wref_clip[f][p] for f in slot_frames] if wref_clip is not None else None
            mt = [match_clip[f][mplane] for f in slot_frames]
            _, num, den = bm3d_plane_temporal(srcs, r, a["sigma"], a["block"], a["group"], a["step"], a["bm_range"],
                                              a["ps_num"], a["ps_range"], r, wrefs=wr, matches=mt, center=c,
                                              frame_count=n_frames, fp64=fp64, want_slices=True)
            for t in range(ns):
                if 0 <= c + t - r < n_frames:
                    fat[2 * t * h:(2 * t + 1) * h] = num[t]
                    fat[(2 * t + 1) * h:(2 * t + 2) * h] = den[t]
            out[c][p] = fat
    return out


def bm3d_clip_legacy(clip, params, ref_clip=None, fp64=False):
    """nss.BM3D with radius > 0 and temporal_mode "legacy" over a clip: the
    contribution frames of the centres in order. With final=1 the basic
    stage runs rolling and feeds the Wiener stage, which emits the slices
    (src/frontend/bm3d_args.cpp:186-221)."""
    if params.radius < 1:
        raise ValueError("legacy temporal output needs radius > 0")
    if _is_final(params):
        basic = _temporal_stage(clip, params, True, None, clip, fp64)
        return _legacy_stage(clip, params, basic, basic, fp64)
    if ref_clip is not None:
        return _legacy_stage(clip, params, ref_clip, ref_clip, fp64)
    return _legacy_stage(clip, params, None, clip, fp64)


def vaggregate(fat_clip, src_clip, radius, planes=None):
    """nss.VAggregate (src/host/filter_vaggregate.cpp:29-100 with
    vaggregate_target, src/cpu/bm/agg.cpp:159-201): frame n of plane p is the
    float32 sum over c = max(0, n - r) .. min(N - 1, n + r), ascending, of
    slice n - c + r's num, over the same sum of den, or the source where that
    den sum is <= 1e-12; planes not selected are the source."""
    n_frames = len(src_clip)
    rows = 2 * (2 * radius + 1)
    out = []
    for n in range(n_frames):
        frame = []
        for p, s in enumerate(src_clip[n]):
            s = np.asarray(s, np.float32)
            if planes is not None and p not in planes:
                frame.append(s.copy())
                continue
            h = s.shape[0]
            num = np.zeros_like(s)
            den = np.zeros_like(s)
            for c in range(max(0, n - radius), min(n_frames - 1, n + radius) + 1):
                f = np.asarray(fat_clip[c][p], np.float32)
                assert f.shape == (h * rows, s.shape[1])
                sl = n - c + radius
                num = num + f[2 * sl * h:(2 * sl + 1) * h]
                den = den + f[(2 * sl + 1) * h:(2 * sl + 2) * h]
            ok = den > np.float32(1e-12)
            frame.append(np.where(ok, num / np.where(ok, den, np.float32(1)), s).astype(np.float32))
        out.append(frame)
    return out


@dataclass
class Params:
    """Mirror of Bm3dParams with the reference defaults (effective sigma)."""
    sigma: list = field(default_factory=lambda: [effective_sigma(3)] * 3)
    block_size: list = field(default_factory=lambda: [8] * 3)
    group_size: list = field(default_factory=lambda: [8] * 3)
    block_step: list = field(default_factory=lambda: [8] * 3)
    bm_range: list = field(default_factory=lambda: [9] * 3)
    ps_num: list = field(default_factory=lambda: [2] * 3)
    ps_range: list = field(default_factory=lambda: [4] * 3)
    radius: int = 0
    chroma: bool = False
    final: bool = False
    sigma_basic: list = field(default_factory=lambda: [0.0] * 3)
    block_size_basic: list = field(default_factory=lambda: [0] * 3)
    group_size_basic: list = field(default_factory=lambda: [0] * 3)
    block_step_basic: list = field(default_factory=lambda: [0] * 3)
    ps_num_basic: list = field(default_factory=lambda: [0] * 3)


def psnr(a, b, peak=1.0):
    """10 log10(peak^2 / MSE) in float64 (tests/cuda/test_cuda_bm3d_live.py:204-206)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    mse = float(np.mean((a - b) ** 2))
    return float("inf") if mse == 0 else 10.0 * np.log10(peak * peak / mse)


# ---------------------------------------------------------------------------
# WNNM (the sibling group filter on the same matcher), restated in float64.
# Parity pins: the reference's zero-rank plugin test (its known answer, sigma
# 255 with residual 1), sigma 0 (every singular value kept whole, so the
# output is the input), hand-computed shrink values, and a per-group scalar
# restatement through the Gram eigenproblem instead of the SVD.
# ---------------------------------------------------------------------------
WNNM_DEFAULTS = dict(block=8, step=8, group=8, bm_range=7, residual=0, adaptive=1)  # include/nss/params.hpp:61-75


def match_plane(plane, block, step, bm_range, group, fp64=False, want_dist=False):
    """Every raster reference's matches: (xy[nref, group, 2], counts[nref]),
    and their distances [nref, group] with want_dist
    (src/host/filter_wnnm.cpp:96-104 spatial_match_batch on the match plane)."""
    plane = np.ascontiguousarray(plane, np.float32)
    h, w = plane.shape
    L = lib()
    if not getattr(L, "_plane_types", False):
        P = ctypes.POINTER
        ci = ctypes.c_int
        L.orc_spatial_match_plane.argtypes = [P(ctypes.c_float), ci, ci, ci, ci, ci, ci, ci, ci, P(ci), P(ci),
                                              P(ctypes.c_float)]
        L.orc_spatial_match_plane.restype = ci
        L._plane_types = True
    nref = len(raster_grid(w, h, block, step))
    xy = np.zeros((max(nref, 1), group, 2), np.int32)
    counts = np.zeros(max(nref, 1), np.int32)
    dist = np.zeros((max(nref, 1), group), np.float32)
    n = L.orc_spatial_match_plane(_fp(plane), w, w, h, block, step, bm_range, group, int(fp64), _ip(xy), _ip(counts),
                                  _fp(dist))
    assert n == nref
    return (xy[:nref], counts[:nref], dist[:nref]) if want_dist else (xy[:nref], counts[:nref])


def sv_shrink(s, constant, start):
    """src/cpu/common/sv_shrink.cpp:9-24 on a descending spectrum: values
    before `start` stay; each later value s with s^2 - constant > 0 becomes
    (s + sqrt(s^2 - constant)) / 2, and the walk stops at the first that is
    not. Returns (shrunk values, kept count)."""
    s = np.array(s, np.float64)
    k = min(max(start, 0), len(s))
    while k < len(s):
        tmp = s[k] * s[k] - constant
        if not tmp > 0.0:
            break
        s[k] = (s[k] + np.sqrt(tmp)) * 0.5
        k += 1
    s[k:] = 0.0
    return s, k


def wnnm_group(A, sigma, residual, adaptive):
    """One group matrix A (m = block^2 rows, n = matched patches as columns):
    src/cpu/wnnm/shrink.cpp:13-49 wnnm_shrink with src/cpu/finishers.hpp:12-36
    finish_wnnm. Optional row centering (src/cpu/common/group_center.cpp:15-58),
    economy SVD, sv_shrink with constant 8 sqrt(2n) sigma^2 from place
    residual ? 0 : 1, X = U S' V^T, centre added back. Returns (X, weight):
    1 / kept when adaptive and kept > 0, else 1."""
    A = np.array(A, np.float64)
    m, n = A.shape
    mean = A.mean(axis=1, keepdims=True) if residual else np.zeros((m, 1))
    U, s, Vt = np.linalg.svd(A - mean, full_matrices=False)
    constant = 8.0 * np.sqrt(2.0 * n) * sigma * sigma
    shrunk, kept = sv_shrink(s, constant, 0 if residual else 1)
    X = (U * shrunk) @ Vt + mean
    weight = 1.0 / kept if adaptive and kept > 0 else 1.0
    return X, weight


def wnnm_plane(src, sigma_user, block=8, step=8, group=8, bm_range=7, residual=0, adaptive=1, match=None,
               want_matches=False):
    """One spatial plane of nss.WNNM (src/host/filter_wnnm.cpp:44-160
    process_plane_batched, radius 0): match on `match` (rclip) or `src`,
    pack each reference's k matches from `src` (src/cpu/bm/pack.cpp:23-40),
    shrink, add weight * X into num and weight into den over each patch
    (aggregate_add), then dst = den > 1e-12 ? num / den : src
    (src/cpu/bm/agg.cpp:17-42). sigma is the user value (divided by 255,
    src/cuda/wnnm/filter.cpp:51)."""
    src = np.ascontiguousarray(src, np.float32)
    h, w = src.shape
    sigma = float(np.float32(sigma_user) / np.float32(255.0))
    mplane = src if match is None else np.ascontiguousarray(match, np.float32)
    xy, counts = match_plane(mplane, block, step, bm_range, group)
    num = np.zeros((h, w), np.float64)
    den = np.zeros((h, w), np.float64)
    dy, dx = np.divmod(np.arange(block * block), block)
    s64 = src.astype(np.float64)
    for r in range(len(counts)):
        k = int(counts[r])
        if k <= 0:
            continue
        cols = [s64[xy[r, j, 1] + dy, xy[r, j, 0] + dx] for j in range(k)]
        X, wt = wnnm_group(np.stack(cols, axis=1), sigma, residual, adaptive)
        for j in range(k):
            num[xy[r, j, 1] + dy, xy[r, j, 0] + dx] += wt * X[:, j]
            den[xy[r, j, 1]:xy[r, j, 1] + block, xy[r, j, 0]:xy[r, j, 0] + block] += wt
    out = np.where(den > 1e-12, num / np.where(den > 1e-12, den, 1.0), s64).astype(np.float32)
    return (out, xy, counts) if want_matches else out


def wnnm_contribution(src_slots, center, n_frames, sigma_user, block=8, step=8, group=8, bm_range=7, ps_num=2,
                      ps_range=4, residual=0, adaptive=1, match_slots=None):
    """One plane of nss.WNNM's fat output at radius r > 0 for centre frame
    `center` of an n_frames clip (src/host/filter_wnnm.cpp:44-177
    process_plane_batched with emit_fat): src_slots / match_slots are the
    planes of slots 0 .. 2r (temporal_slot_frame). Each raster reference is
    matched on the centre of match_slots (or src_slots) and predictively over
    the slots inside the clip (predictive_match_batch, :99-102), its k patches
    packed from their slot's frame (:116-121), shrunk, and added with the
    group weight into the num / den of their slot (:140-160). Returns an
    array (2r + 1, 2, h, w): slot t's num and den."""
    ns = len(src_slots)
    r = (ns - 1) // 2
    srcs = [np.ascontiguousarray(p, np.float32) for p in src_slots]
    mpl = srcs if match_slots is None else [np.ascontiguousarray(p, np.float32) for p in match_slots]
    h, w = srcs[0].shape
    sigma = float(np.float32(sigma_user) / np.float32(255.0))
    vb = max(0, r - center)
    ve = r + min(r + 1, n_frames - center)
    out = np.zeros((ns, 2, h, w), np.float64)
    dy, dx = np.divmod(np.arange(block * block), block)
    s64 = [p.astype(np.float64) for p in srcs]
    for bx, by in raster_grid(w, h, block, step):
        m = predictive_match(mpl, bx, by, r, block, group, bm_range, ps_num, ps_range, r, vb, ve)
        cols = [s64[t][y + dy, x + dx] for x, y, t, _ in m]
        X, wt = wnnm_group(np.stack(cols, axis=1), sigma, residual, adaptive)
        for j, (x, y, t, _) in enumerate(m):
            out[t, 0, y + dy, x + dx] += wt * X[:, j]
            out[t, 1, y:y + block, x:x + block] += wt
    return out


def wnnm_frame(planes, sigma, block=8, step=8, group=8, bm_range=7, residual=0, adaptive=1, rclip=None):
    """All planes of one frame; a plane with sigma 0 is copied
    (src/host/filter_wnnm.cpp:200-260)."""
    out = []
    for i, p in enumerate(planes):
        if sigma[i] == 0:
            out.append(np.array(p, np.float32))
            continue
        out.append(wnnm_plane(p, sigma[i], block, step, group, bm_range, residual, adaptive,
                              None if rclip is None else rclip[i]))
    return out


# ---------------------------------------------------------------------------
# NCSR (the second sibling group filter), restated in float64.
# ---------------------------------------------------------------------------
NCSR_DEFAULTS = dict(block=8, step=8, group=8, bm_range=7, iters=2, delta=0.1)  # include/nss/params.hpp:101-107


def ncsr_group(A, dist, sigma):
    """One group (m rows, n = k columns) and its match distances (self
    first): src/cpu/ncsr/centralize.cpp:143-176 NcsrFilterGroup with
    src/cpu/finishers.hpp:53-80 finish_ncsr_codes. PCA of the row-centred
    group (U, codes B = U^T (A - mean), the min(m, n) leading directions),
    column weights exp(-dist / h) with h = max(2 m sigma^2, 1e-12) and the
    self column at distance 0 (centralize.cpp:60-88), per code row the
    weighted mean beta, the second moment over the actual columns about it,
    tau = 2 sqrt(2) sigma^2 / (sqrt(moment / n) + 1e-12), the row
    soft-thresholded about beta (centralize.cpp:90-141), then
    X = U B' + mean. sigma 0 returns the group unchanged. Weight 1."""
    A = np.array(A, np.float64)
    m, n = A.shape
    if not sigma > 0:
        return A.copy()
    mean = A.mean(axis=1, keepdims=True)
    C = A - mean
    U, s, Vt = np.linalg.svd(C, full_matrices=False)
    r = min(m, n)
    B = U[:, :r].T @ C
    h = max(2.0 * m * sigma * sigma, 1e-12)
    d = np.array(dist, np.float64)[:n].copy()
    d[0] = 0.0
    w = np.exp(-d / h)
    wsum = w.sum()
    if not wsum > 0:
        w, wsum = np.ones(n), float(n)
    for i in range(r):
        if not s[i] > 0:
            B[i] = 0.0
            continue
        row = B[i]
        beta = float(np.dot(w, row)) / wsum
        moment = float(np.sum((row - beta) ** 2))
        tau = 2.8284271247461903 * sigma * sigma / (np.sqrt(moment / n) + 1e-12)
        e = row - beta
        B[i] = np.sign(e) * np.maximum(np.abs(e) - tau, 0.0) + beta
    return U[:, :r] @ B + mean


def ncsr_plane(src, sigma_user, block=8, step=8, group=8, bm_range=7, iters=2, delta=0.1, match=None):
    """One spatial plane of nss.NCSR (src/host/filter_ncsr.cpp:34-190 at
    radius 0): the estimate starts as the noisy plane; round i > 0 first
    relaxes it toward the noisy plane, est += delta (noisy - est)
    (src/cpu/common/iter_regularize.cpp:15-39); each round matches on the
    estimate (on rclip in round 0), filters the estimate's groups
    (src/cpu/ncsr/centralize.cpp:423-540 ncsr_run_groups), aggregates with
    weight 1 and finishes est = den > 0 ? num / den : est."""
    noisy = np.ascontiguousarray(src, np.float32).astype(np.float64)
    h, w = noisy.shape
    sigma = float(np.float32(sigma_user) / np.float32(255.0))
    est = noisy.copy()
    dy, dx = np.divmod(np.arange(block * block), block)
    for it in range(iters):
        if it > 0:
            est = est + delta * (noisy - est)
        mplane = match if (it == 0 and match is not None) else est
        xy, counts, dist = match_plane(np.asarray(mplane, np.float32), block, step, bm_range, group, want_dist=True)
        num = np.zeros((h, w))
        den = np.zeros((h, w))
        for r in range(len(counts)):
            k = int(counts[r])
            cols = [est[xy[r, j, 1] + dy, xy[r, j, 0] + dx] for j in range(k)]
            X = ncsr_group(np.stack(cols, axis=1), dist[r, :k], sigma)
            for j in range(k):
                num[xy[r, j, 1] + dy, xy[r, j, 0] + dx] += X[:, j]
                den[xy[r, j, 1]:xy[r, j, 1] + block, xy[r, j, 0]:xy[r, j, 0] + block] += 1.0
        est = np.where(den > 1e-12, num / np.where(den > 1e-12, den, 1.0), est)
    return est.astype(np.float32)


def ncsr_contribution(src_slots, center, n_frames, sigma_user, block=8, step=8, group=8, bm_range=7, ps_num=2,
                      ps_range=4, iters=2, delta=0.1, match_slots=None):
    """One plane of nss.NCSR's fat output at radius r > 0 for centre frame
    `center` of an n_frames clip (src/host/filter_ncsr.cpp:34-190): every
    slot t keeps an estimate and its noisy plane (:124-133); round i > 0
    relaxes each estimate toward its noisy plane (:154-159); each round
    matches every raster reference on match_slots in round 0 (else the
    estimates) by the predictive search over the slots inside the clip,
    packs the patches from their slot's estimate (ncsr_run_groups,
    src/cpu/ncsr/centralize.cpp:423-500), filters with the match distances
    and adds each patch with weight 1 into its slot's num / den (:162-167);
    every round but the last finishes each estimate (:171-175). Returns the
    last round's sums as an array (2r + 1, 2, h, w)."""
    ns = len(src_slots)
    r = (ns - 1) // 2
    noisy = [np.ascontiguousarray(p, np.float32).astype(np.float64) for p in src_slots]
    h, w = noisy[0].shape
    sigma = float(np.float32(sigma_user) / np.float32(255.0))
    vb = max(0, r - center)
    ve = r + min(r + 1, n_frames - center)
    est = [p.copy() for p in noisy]
    dy, dx = np.divmod(np.arange(block * block), block)
    for it in range(iters):
        if it > 0:
            est = [e + delta * (y - e) for e, y in zip(est, noisy)]
        mpl = match_slots if (it == 0 and match_slots is not None) else est
        mpl = [np.asarray(p, np.float32) for p in mpl]
        out = np.zeros((ns, 2, h, w), np.float64)
        for bx, by in raster_grid(w, h, block, step):
            m = predictive_match(mpl, bx, by, r, block, group, bm_range, ps_num, ps_range, r, vb, ve)
            cols = [est[t][y + dy, x + dx] for x, y, t, _ in m]
            X = ncsr_group(np.stack(cols, axis=1), np.array([d for *_, d in m], np.float64), sigma)
            for j, (x, y, t, _) in enumerate(m):
                out[t, 0, y + dy, x + dx] += X[:, j]
                out[t, 1, y:y + block, x:x + block] += 1.0
        if it + 1 == iters:
            return out
        est = [np.where(out[t, 1] > 1e-12, out[t, 0] / np.where(out[t, 1] > 1e-12, out[t, 1], 1.0), est[t])
               for t in range(ns)]
