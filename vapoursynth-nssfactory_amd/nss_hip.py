"""nss_hip: Python host layer over the C ABI of libnss_hip.so.

Mirrors the reference plugin's ``nss.BM3D`` interface for callers without
VapourSynth: the same argument names, defaults, per-plane short-array
inheritance and error texts (resolved by the C++ parser behind
``nss_hip_bm3d_parse``), applied to frames held as device tensors.

    inst = BM3D("YUV420PS", 1920, 1080, sigma=10, final=1)
    out_planes = inst(planes)          # list of torch.cuda float32 2-D tensors

PyTorch only provides device memory and the stream; every computation runs in
the HIP kernels of libnss_hip.so. A missing library is an error: there is no
fallback path.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnss_hip.so")
NAMESPACE = "nss_hip"

# The reference signature (src/frontend/bm3d_args.hpp:8-12) plus the device
# backend tail (src/cuda/runtime/context.hpp:17) and this library's transform
# argument (no counterpart in the reference: the MFMA forms).
BM3D_SIGNATURE = (
    "clip:vnode;ref:vnode:opt;sigma:float[]:opt;block_size:int[]:opt;group_size:int[]:opt;"
    "block_step:int[]:opt;bm_range:int[]:opt;radius:int:opt;ps_num:int[]:opt;ps_range:int[]:opt;chroma:int:opt;"
    "final:int:opt;sigma_basic:float[]:opt;block_size_basic:int[]:opt;group_size_basic:int[]:opt;"
    "temporal_mode:data:opt;rolling_chunk:int:opt;rolling_cache_chunks:int:opt;rolling_cache_limit:int:opt;"
    "memory_limit_mb:int:opt;"
    "device_id:int:opt;num_streams:int:opt;"
    "transform:int:opt;"
)

# Group-transform arithmetic (the transform argument).
TRANSFORM_FP32, TRANSFORM_BF16X3, TRANSFORM_FP16 = 0, 1, 2

OK, BYPASS, INVALID_INPUT, NUMERICAL_FAILURE, UNSUPPORTED, RESOURCE_FAILURE = range(6)

# name: (color family, sub-sampling w, h, planes); VapourSynth API4 values.
FORMATS = {
    "GRAYS": (1, 0, 0, 1),
    "RGBS": (2, 0, 0, 3),
    "YUV420PS": (3, 1, 1, 3),
    "YUV422PS": (3, 1, 0, 3),
    "YUV444PS": (3, 0, 0, 3),
}
FORMAT_NAMES = {"GRAYS": "GrayS", "RGBS": "RGBS", "YUV420PS": "YUV420PS", "YUV422PS": "YUV422PS",
                "YUV444PS": "YUV444PS"}


class Error(Exception):
    """Raised with the library's message (the text nss.BM3D would report)."""


class VideoInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("color_family", "sample_type", "bits_per_sample", "sub_sampling_w",
                                            "sub_sampling_h", "num_planes", "width", "height", "num_frames")]


class Arg(ctypes.Structure):
    _fields_ = [("key", ctypes.c_char_p), ("type", ctypes.c_char), ("count", ctypes.c_int),
                ("ints", ctypes.POINTER(ctypes.c_int64)), ("floats", ctypes.POINTER(ctypes.c_double)),
                ("data", ctypes.c_char_p), ("clip_info", ctypes.POINTER(VideoInfo))]


def _arr(t, n):
    return t * n


class Params(ctypes.Structure):
    _fields_ = [("sigma", _arr(ctypes.c_float, 3)), ("block_size", _arr(ctypes.c_int, 3)),
                ("group_size", _arr(ctypes.c_int, 3)), ("block_step", _arr(ctypes.c_int, 3)),
                ("bm_range", _arr(ctypes.c_int, 3)), ("ps_num", _arr(ctypes.c_int, 3)),
                ("ps_range", _arr(ctypes.c_int, 3)), ("radius", ctypes.c_int), ("chroma", ctypes.c_int),
                ("final_stage", ctypes.c_int), ("sigma_basic", _arr(ctypes.c_float, 3)),
                ("block_size_basic", _arr(ctypes.c_int, 3)), ("group_size_basic", _arr(ctypes.c_int, 3)),
                ("block_step_basic", _arr(ctypes.c_int, 3)), ("ps_num_basic", _arr(ctypes.c_int, 3)),
                ("has_ref", ctypes.c_int), ("temporal_mode", ctypes.c_int), ("rolling_chunk", ctypes.c_int),
                ("rolling_cache_chunks", ctypes.c_int), ("rolling_cache_limit", ctypes.c_int),
                ("rolling_cache_adaptive", ctypes.c_int), ("memory_limit_mb", ctypes.c_int),
                ("device_id", ctypes.c_int), ("num_streams", ctypes.c_int), ("transform", ctypes.c_int)]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


class WnnmParams(ctypes.Structure):
    _fields_ = [("sigma", _arr(ctypes.c_float, 3)), ("block_size", ctypes.c_int), ("block_step", ctypes.c_int),
                ("group_size", ctypes.c_int), ("bm_range", ctypes.c_int), ("radius", ctypes.c_int),
                ("ps_num", ctypes.c_int), ("ps_range", ctypes.c_int), ("residual", ctypes.c_int),
                ("adaptive", ctypes.c_int), ("has_rclip", ctypes.c_int), ("memory_limit_mb", ctypes.c_int),
                ("device_id", ctypes.c_int), ("temporal_mode", ctypes.c_int)]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


class NcsrParams(ctypes.Structure):
    _fields_ = [("sigma", _arr(ctypes.c_float, 3)), ("block_size", ctypes.c_int), ("block_step", ctypes.c_int),
                ("group_size", ctypes.c_int), ("bm_range", ctypes.c_int), ("radius", ctypes.c_int),
                ("ps_num", ctypes.c_int), ("ps_range", ctypes.c_int), ("iters", ctypes.c_int),
                ("delta", ctypes.c_float), ("has_rclip", ctypes.c_int), ("memory_limit_mb", ctypes.c_int),
                ("device_id", ctypes.c_int),
                ("temporal_mode", ctypes.c_int)]

    as_dict = WnnmParams.as_dict


_lib = None

EXPORTS = ("nss_hip_last_error", "nss_hip_backend_info", "nss_hip_bm3d_parse", "nss_hip_bm3d_output_info",
           "nss_hip_bm3d_create", "nss_hip_bm3d_slots", "nss_hip_bm3d_process", "nss_hip_bm3d_destroy",
           "nss_hip_spatial_match", "nss_hip_set_match_path", "nss_hip_bm3d_set_tracing",
           "nss_hip_bm3d_kernel_times", "nss_hip_bm3d_push", "nss_hip_bm3d_pop", "nss_hip_server_create", "nss_hip_server_open_stream", "nss_hip_server_submit",
           "nss_hip_server_query", "nss_hip_server_drain", "nss_hip_server_get_stats", "nss_hip_server_destroy",
           "nss_hip_bm3d_autotune", "nss_hip_vaggregate", "nss_hip_bm3d_resources", "nss_hip_wnnm_parse",
           "nss_hip_wnnm_create", "nss_hip_wnnm_process", "nss_hip_wnnm_destroy", "nss_hip_ncsr_parse",
           "nss_hip_ncsr_create", "nss_hip_ncsr_process", "nss_hip_ncsr_destroy", "nss_hip_bm3d_set_tiles",
           "nss_hip_predictive_match", "nss_hip_bm3d_process_noalloc", "nss_hip_debug_fail_alloc",
           "nss_hip_debug_fail_launch", "nss_hip_debug_live_allocations", "nss_hip_bm3d_chunk", "nss_hip_bm3d_seek",
           "nss_hip_wnnm_process_temporal", "nss_hip_ncsr_process_temporal")
# Contribution identity of legacy temporal frames (frame properties the
# VapourSynth glue stamps; include/nss/contracts.hpp:7-20,
# src/frontend/contribution.hpp:12-27).
FAT_VERSION, FAT_LAYOUT, MODEL_BM3D, MODEL_VERSION_BM3D, NOISE_PROFILE_BM3D = 2, 1, 1, 3, 1
KERNEL_KINDS = ("clear", "match", "filter", "finish")


class ServerConfig(ctypes.Structure):
    _fields_ = [("device_id", ctypes.c_int), ("hbm_budget_bytes", ctypes.c_longlong),
                ("max_batch_frames", ctypes.c_int), ("max_inflight", ctypes.c_int), ("use_graphs", ctypes.c_int),
                ("autotune", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_longlong) for n in ("submitted", "completed", "batches", "batched_frames",
                                                 "graph_replays", "graph_captures", "bytes_in_use", "peak_bytes",
                                                 "admissions_deferred", "slots_released")]


def lib():
    """The loaded libnss_hip.so; raises if it was not built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        vp, ci = ctypes.c_void_p, ctypes.c_int
        L.nss_hip_last_error.restype = ctypes.c_char_p
        L.nss_hip_last_error.argtypes = []
        L.nss_hip_backend_info.argtypes = [ci, ctypes.c_char_p, ci]
        L.nss_hip_bm3d_parse.argtypes = [P(VideoInfo), P(VideoInfo), P(Arg), ci, ctypes.c_char_p, P(Params)]
        L.nss_hip_bm3d_output_info.argtypes = [P(Params), P(VideoInfo), P(VideoInfo)]
        L.nss_hip_bm3d_create.argtypes = [P(Params), P(VideoInfo), ci, P(vp)]
        L.nss_hip_bm3d_slots.argtypes = [vp]
        L.nss_hip_bm3d_process.argtypes = [vp, P(vp), P(ci), P(vp), P(ci), P(vp), P(ci), ci, ci, vp]
        L.nss_hip_bm3d_destroy.argtypes = [vp]
        L.nss_hip_bm3d_process_noalloc.argtypes = [P(Params), P(VideoInfo), ci, P(vp), P(ci), P(vp), P(ci), P(vp),
                                                   P(ci), vp]
        L.nss_hip_spatial_match.argtypes = [vp, ci, ci, ci, ci, ci, ci, ci, vp, vp, vp]
        L.nss_hip_set_match_path.argtypes = [ci]
        L.nss_hip_predictive_match.argtypes = [vp, ci, ci, ci, ci, ci, ci, ci, ci, ci, ci, ci, ci, vp, vp, vp]
        L.nss_hip_bm3d_set_tracing.argtypes = [vp, ci]
        L.nss_hip_bm3d_kernel_times.argtypes = [vp, P(ci), P(ctypes.c_double), ci]
        L.nss_hip_bm3d_push.argtypes = [vp, P(vp), P(ci), P(vp), P(ci), vp]
        L.nss_hip_bm3d_pop.argtypes = [vp, P(vp), P(ci), P(ci), vp]
        L.nss_hip_bm3d_autotune.argtypes = [vp, ci, vp, P(ctypes.c_double)]
        L.nss_hip_bm3d_set_tiles.argtypes = [vp, ci, ci, ci]
        L.nss_hip_vaggregate.argtypes = [vp, ci, P(vp), P(vp), P(ci), ci, vp, ci, ci, ci, vp]
        L.nss_hip_bm3d_resources.argtypes = [vp, P(ctypes.c_longlong), P(ctypes.c_longlong),
                                             P(ctypes.c_longlong)]
        L.nss_hip_server_create.argtypes = [P(ServerConfig), P(vp)]
        L.nss_hip_server_open_stream.argtypes = [vp, P(Params), P(VideoInfo), P(ci)]
        L.nss_hip_server_submit.argtypes = [vp, ci, P(vp), P(ci), P(vp), P(ci), ctypes.c_double, vp,
                                            P(ctypes.c_longlong)]
        L.nss_hip_server_query.argtypes = [vp, ctypes.c_longlong, P(ctypes.c_double)]
        L.nss_hip_server_drain.argtypes = [vp]
        L.nss_hip_server_get_stats.argtypes = [vp, P(ServerStats)]
        L.nss_hip_server_destroy.argtypes = [vp]
        L.nss_hip_wnnm_parse.argtypes = [P(VideoInfo), P(VideoInfo), P(Arg), ci, ctypes.c_char_p, P(WnnmParams)]
        L.nss_hip_wnnm_create.argtypes = [P(WnnmParams), P(VideoInfo), ci, P(vp)]
        L.nss_hip_wnnm_process.argtypes = [vp, P(vp), P(ci), P(vp), P(ci), P(vp), P(ci), vp]
        L.nss_hip_wnnm_destroy.argtypes = [vp]
        for fn in (L.nss_hip_wnnm_process_temporal, L.nss_hip_ncsr_process_temporal):
            fn.argtypes = [vp, ci, P(vp), P(ci), P(vp), P(ci), P(vp), P(ci), vp]
        L.nss_hip_ncsr_parse.argtypes = [P(VideoInfo), P(VideoInfo), P(Arg), ci, ctypes.c_char_p, P(NcsrParams)]
        L.nss_hip_ncsr_create.argtypes = [P(NcsrParams), P(VideoInfo), ci, P(vp)]
        L.nss_hip_ncsr_process.argtypes = [vp, P(vp), P(ci), P(vp), P(ci), P(vp), P(ci), vp]
        L.nss_hip_ncsr_destroy.argtypes = [vp]
        L.nss_hip_bm3d_chunk.argtypes = [vp, ci, P(ci), P(ci), P(ci), P(ci)]
        L.nss_hip_bm3d_seek.argtypes = [vp, ci, P(ci), vp]
        L.nss_hip_debug_fail_alloc.argtypes = [ci]
        L.nss_hip_debug_fail_launch.argtypes = [ci]
        L.nss_hip_debug_live_allocations.argtypes = [P(ctypes.c_longlong), P(ctypes.c_longlong)]
        for name in EXPORTS[1:]:
            getattr(L, name).restype = ci
        _lib = L
    return _lib


def last_error() -> str:
    return lib().nss_hip_last_error().decode()


def debug_fail_alloc(n: int) -> None:
    """Test hook: the n-th next device allocation of the library fails
    (0: the next one); n < 0 disarms."""
    _check(lib().nss_hip_debug_fail_alloc(int(n)))


def debug_fail_launch(n: int) -> None:
    """Test hook: the n-th next kernel launch is refused; n < 0 disarms."""
    _check(lib().nss_hip_debug_fail_launch(int(n)))


def debug_live_allocations():
    """(count, bytes) of the device allocations the library holds now."""
    c, b = ctypes.c_longlong(), ctypes.c_longlong()
    _check(lib().nss_hip_debug_live_allocations(ctypes.byref(c), ctypes.byref(b)))
    return c.value, b.value


def _check(rc):
    if rc != OK:
        raise Error(last_error())


def video_info(fmt: str, width: int, height: int, frames: int = 1) -> VideoInfo:
    cf, sw, sh, np_ = FORMATS[fmt]
    return VideoInfo(cf, 1, 32, sw, sh, np_, width, height, frames)


def _signature_types(signature=None):
    types = {}
    for item in filter(None, (signature or BM3D_SIGNATURE).split(";")):
        name, kind, *_ = item.split(":")
        types[name] = kind.rstrip("[]")
    return types


_TYPES = _signature_types()
# nss.WNNM's argument list (src/frontend/wnnm_args.hpp:10-13 kWnnmSignature
# with NSS_TEMPORAL_SIGNATURE, src/frontend/temporal_args.hpp:20-21) and the
# device backend's device_id.
WNNM_SIGNATURE = ("clip:vnode;sigma:float[]:opt;block_size:int:opt;block_step:int:opt;group_size:int:opt;"
                  "bm_range:int:opt;radius:int:opt;ps_num:int:opt;ps_range:int:opt;residual:int:opt;"
                  "adaptive_aggregation:int:opt;rclip:vnode:opt;memory_limit_mb:int:opt;"
                  "temporal_mode:data:opt;rolling_chunk:int:opt;rolling_cache_chunks:int:opt;"
                  "rolling_cache_limit:int:opt;device_id:int:opt;")
_WNNM_TYPES = _signature_types(WNNM_SIGNATURE)
# nss.NCSR's (src/frontend/ncsr_args.hpp signature: WNNM's group arguments
# with iters and delta in place of residual and adaptive_aggregation).
NCSR_SIGNATURE = ("clip:vnode;sigma:float[]:opt;block_size:int:opt;block_step:int:opt;group_size:int:opt;"
                  "bm_range:int:opt;radius:int:opt;ps_num:int:opt;ps_range:int:opt;iters:int:opt;delta:float:opt;"
                  "rclip:vnode:opt;memory_limit_mb:int:opt;"
                  "temporal_mode:data:opt;rolling_chunk:int:opt;rolling_cache_chunks:int:opt;"
                  "rolling_cache_limit:int:opt;device_id:int:opt;")
_NCSR_TYPES = _signature_types(NCSR_SIGNATURE)


def _build_args(kwargs, types=None, name="BM3D"):
    """kwargs -> Arg array, values converted to the signature's type (as
    VapourSynth's invoke does); the buffers are kept alive by the return."""
    types = types or _TYPES
    keep = []
    args = []
    for key, value in kwargs.items():
        if value is None:
            continue
        kind = types.get(key)
        if kind is None:
            raise Error(f"{NAMESPACE}.{name}: no argument named {key}")
        vals = list(value) if isinstance(value, (list, tuple)) else [value]
        if kind == "int":
            buf = (ctypes.c_int64 * len(vals))(*[int(v) for v in vals])
            keep.append(buf)
            args.append(Arg(key.encode(), b"i", len(vals), buf, None, None, None))
        elif kind == "float":
            buf = (ctypes.c_double * len(vals))(*[float(v) for v in vals])
            keep.append(buf)
            args.append(Arg(key.encode(), b"f", len(vals), None, buf, None, None))
        elif kind == "data":
            data = str(vals[0]).encode()
            keep.append(data)
            args.append(Arg(key.encode(), b"s", 1, None, None, data, None))
        else:
            raise Error(f"{NAMESPACE}.{name}: argument {key} is a clip")
    arr = (Arg * max(1, len(args)))(*args)
    return arr, len(args), keep


def parse(fmt, width, height, frames=1, ref=False, info=None, ref_info=None, **kwargs) -> Params:
    """Resolve nss.BM3D arguments; raises Error with the reference's text."""
    vi = info if info is not None else video_info(fmt, width, height, frames)
    rvi = ref_info if ref_info is not None else (vi if ref else None)
    arr, n, keep = _build_args(kwargs)
    out = Params()
    rc = lib().nss_hip_bm3d_parse(ctypes.byref(vi), ctypes.byref(rvi) if rvi is not None else None, arr, n,
                                  NAMESPACE.encode(), ctypes.byref(out))
    del keep
    _check(rc)
    return out


def output_info(params: Params, info: VideoInfo) -> VideoInfo:
    out = VideoInfo()
    _check(lib().nss_hip_bm3d_output_info(ctypes.byref(params), ctypes.byref(info), ctypes.byref(out)))
    return out


def backend_info(device: int = 0) -> str:
    buf = ctypes.create_string_buffer(512)
    _check(lib().nss_hip_backend_info(device, buf, 512))
    return buf.value.decode()


def _check_out(name, out, dims):
    """Output tensors: one per plane, row-contiguous 2-D float32 on the
    device, of the plane's shape (a smaller buffer would be written past
    its end)."""
    import torch
    if len(out) != len(dims):
        raise Error(f"{NAMESPACE}.{name}: expected {len(dims)} output planes, got {len(out)}")
    for i, p in enumerate(out):
        if p.dtype != torch.float32 or not p.is_cuda or p.dim() != 2 or p.stride(1) != 1:
            raise Error(f"{NAMESPACE}.{name}: output plane {i} must be a row-contiguous 2-D float32 CUDA tensor")
        if tuple(p.shape) != tuple(dims[i]):
            raise Error(f"{NAMESPACE}.{name}: output plane {i} has shape {tuple(p.shape)}, expected {tuple(dims[i])}")


def _stream_handle(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class BM3D:
    """One filter instance (fixed format and size) on one device.

    Calling it filters a frame: ``planes`` is a list of 2-D float32 CUDA
    tensors (one per plane); ``ref`` likewise when the instance was made with
    ``ref=True``. Returns new output tensors (or writes ``out`` when given).
    Work is queued on the current torch stream.
    """

    def __init__(self, fmt, width, height, frames=1, ref=False, **kwargs):
        self.fmt = fmt
        self.info = video_info(fmt, width, height, frames)
        self.params = parse(fmt, width, height, frames, ref=ref, info=self.info, **kwargs)
        self.out_info = output_info(self.params, self.info)
        self.device = self.params.device_id
        h = ctypes.c_void_p()
        _check(lib().nss_hip_bm3d_create(ctypes.byref(self.params), ctypes.byref(self.info), self.device,
                                         ctypes.byref(h)))
        self._h = h
        cf, sw, sh, n = FORMATS[fmt]
        self.plane_dims = [(height >> (sh if i else 0), width >> (sw if i else 0)) for i in range(n)]
        # radius > 0 with temporal_mode "legacy": pop returns contribution
        # frames of 2 (2r + 1) plane heights (num/den slice pairs).
        self.legacy = self.params.radius > 0 and self.params.temporal_mode == 0
        rows = 2 * (2 * self.params.radius + 1) if self.legacy else 1
        self.out_dims = [(hh * rows, ww) for hh, ww in self.plane_dims]

    def autotune(self, repeats=3, stream=None):
        """Times every filter tile shape that fits (per bucket) on a synthetic
        frame and keeps the fastest plan; returns its ms per frame."""
        ms = ctypes.c_double(0)
        _check(lib().nss_hip_bm3d_autotune(self._h, repeats, _stream_handle(stream), ctypes.byref(ms)))
        return ms.value

    def set_tiles(self, stage, tw, th):
        """Filter tiles of tw x th references for stage (0 first); False when
        the shape does not fit. The output does not change with the shape."""
        rc = lib().nss_hip_bm3d_set_tiles(self._h, stage, tw, th)
        if rc == UNSUPPORTED:
            return False
        _check(rc)
        return True

    def set_tracing(self, on=True):
        _check(lib().nss_hip_bm3d_set_tracing(self._h, int(on)))

    def kernel_times(self, reset=True):
        """{kind: (launches, total_ms)} since the last reset (waits for the events)."""
        n = (ctypes.c_int * 4)()
        ms = (ctypes.c_double * 4)()
        _check(lib().nss_hip_bm3d_kernel_times(self._h, n, ms, int(reset)))
        return {k: (n[i], ms[i]) for i, k in enumerate(KERNEL_KINDS)}

    def resources(self):
        """The frame properties a VapourSynth glue stamps for this instance
        (src/frontend/ownership.hpp:103-115): _NSSResourceBytes (workspace,
        cached, pinned, inflight, output, framework), _NSSResourcePeak and
        _NSSResourceLimit (0 without memory_limit_mb)."""
        b = (ctypes.c_longlong * 6)()
        peak, limit = ctypes.c_longlong(), ctypes.c_longlong()
        _check(lib().nss_hip_bm3d_resources(self._h, b, ctypes.byref(peak), ctypes.byref(limit)))
        return {"_NSSResourceBytes": list(b), "_NSSResourcePeak": peak.value, "_NSSResourceLimit": limit.value}

    def _check_planes(self, planes, what="plane"):
        import torch
        n = len(self.plane_dims)
        if len(planes) != n:
            raise Error(f"{NAMESPACE}.BM3D: expected {n} {what}s, got {len(planes)}")
        for i, p in enumerate(planes):
            if p.dtype != torch.float32 or not p.is_cuda or p.dim() != 2 or p.stride(1) != 1:
                raise Error(f"{NAMESPACE}.BM3D: {what} {i} must be a row-contiguous 2-D float32 CUDA tensor")
            if tuple(p.shape) != self.plane_dims[i]:
                raise Error(f"{NAMESPACE}.BM3D: {what} {i} has shape {tuple(p.shape)}, expected {self.plane_dims[i]}")

    def push(self, planes, ref=None, stream=None):
        """radius > 0: the clip's next frame (its planes as device tensors)."""
        self._check_planes(planes)
        if ref is not None:
            self._check_planes(ref, "ref plane")
        n = len(planes)
        src = (ctypes.c_void_p * n)(*[p.data_ptr() for p in planes])
        sst = (ctypes.c_int * n)(*[p.stride(0) for p in planes])
        rp = (ctypes.c_void_p * n)(*[p.data_ptr() for p in ref]) if ref is not None else None
        rst = (ctypes.c_int * n)(*[p.stride(0) for p in ref]) if ref is not None else None
        _check(lib().nss_hip_bm3d_push(self._h, src, sst, rp, rst, _stream_handle(stream)))

    def pop(self, out=None, stream=None):
        """radius > 0: (frame index, planes) of the next finished frame, or None."""
        import torch
        n = len(self.plane_dims)
        if out is None:
            out = [torch.empty(d, dtype=torch.float32, device=torch.device("cuda", self.device))
                   for d in self.out_dims]
        _check_out("BM3D", out, self.out_dims)
        dp = (ctypes.c_void_p * n)(*[p.data_ptr() for p in out])
        dst = (ctypes.c_int * n)(*[p.stride(0) for p in out])
        frame = ctypes.c_int(-1)
        rc = lib().nss_hip_bm3d_pop(self._h, dp, dst, ctypes.byref(frame), _stream_handle(stream))
        if rc == BYPASS:
            return None
        _check(rc)
        return frame.value, out

    def chunk(self, n):
        """radius > 0: (start, count, first, last) of frame n's chunk and the
        source frames its computation reads."""
        v = [ctypes.c_int() for _ in range(4)]
        _check(lib().nss_hip_bm3d_chunk(self._h, int(n), *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def seek(self, n, stream=None):
        """radius > 0: makes frame n's chunk the next one popped; returns the
        frame to push next."""
        nxt = ctypes.c_int(-1)
        _check(lib().nss_hip_bm3d_seek(self._h, int(n), ctypes.byref(nxt), _stream_handle(stream)))
        return nxt.value

    def get_chunk(self, n, source, ref_source=None, stream=None):
        """radius > 0, random access: the finished frames of frame n's chunk
        as {index: planes}, pushing source(i) (and ref_source(i)) for the
        frames the chunk reads (what a glue's getFrame does on a chunk-cache
        miss). Frames behind the chunk that come out are included."""
        start, count, first, last = self.chunk(n)
        out = {}

        def drain():
            while True:
                r = self.pop(stream=stream)
                if r is None:
                    return
                out[r[0]] = r[1]

        for i in range(self.seek(n, stream=stream), last + 1):
            self.push(source(i), None if ref_source is None else ref_source(i), stream=stream)
            drain()
        drain()
        missing = [i for i in range(start, start + count) if i not in out]
        if missing:
            raise Error(f"{NAMESPACE}.BM3D: chunk frames {missing} did not come out")
        return out

    def filter_clip(self, frames, refs=None, stream=None):
        """radius > 0: every frame of a clip (list of per-frame plane lists)
        through push/pop; returns the finished frames in order (legacy
        temporal mode: the contribution frames of the centres in order)."""
        out = []
        for i, f in enumerate(frames):
            self.push(f, None if refs is None else refs[i], stream=stream)
            while True:
                r = self.pop(stream=stream)
                if r is None:
                    break
                out.append(r[1])
        return out

    def close(self):
        if getattr(self, "_h", None):
            lib().nss_hip_bm3d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __call__(self, planes, ref=None, out=None, frame=0, frame_count=1, stream=None):
        import torch
        n = len(self.plane_dims)
        self._check_planes(planes)
        if ref is not None:
            self._check_planes(ref, "ref plane")
        if out is None:
            out = [torch.empty(self.plane_dims[i], dtype=torch.float32, device=planes[i].device) for i in range(n)]
        _check_out("BM3D", out, self.plane_dims)
        VP = ctypes.c_void_p * len(planes)
        IP = ctypes.c_int * len(planes)
        src = VP(*[p.data_ptr() for p in planes])
        sst = IP(*[p.stride(0) for p in planes])
        if ref is not None:
            rp = VP(*[p.data_ptr() for p in ref])
            rst = IP(*[p.stride(0) for p in ref])
        else:
            rp, rst = None, None
        dp = (ctypes.c_void_p * n)(*[p.data_ptr() for p in out])
        dst = (ctypes.c_int * n)(*[p.stride(0) for p in out])
        _check(lib().nss_hip_bm3d_process(self._h, src, sst, rp, rst, dp, dst, frame, frame_count,
                                          _stream_handle(stream)))
        return out


def parse_wnnm(fmt, width, height, frames=1, rclip=False, info=None, **kwargs) -> WnnmParams:
    """Resolve nss.WNNM arguments; raises Error with the reference's text."""
    vi = info if info is not None else video_info(fmt, width, height, frames)
    arr, n, keep = _build_args(kwargs, _WNNM_TYPES, "WNNM")
    out = WnnmParams()
    rc = lib().nss_hip_wnnm_parse(ctypes.byref(vi), ctypes.byref(vi) if rclip else None, arr, n,
                                  NAMESPACE.encode(), ctypes.byref(out))
    del keep
    _check(rc)
    return out


def parse_ncsr(fmt, width, height, frames=1, rclip=False, info=None, **kwargs) -> NcsrParams:
    """Resolve nss.NCSR arguments; raises Error with the reference's text."""
    vi = info if info is not None else video_info(fmt, width, height, frames)
    arr, n, keep = _build_args(kwargs, _NCSR_TYPES, "NCSR")
    out = NcsrParams()
    rc = lib().nss_hip_ncsr_parse(ctypes.byref(vi), ctypes.byref(vi) if rclip else None, arr, n,
                                  NAMESPACE.encode(), ctypes.byref(out))
    del keep
    _check(rc)
    return out


class WNNM:
    """nss.WNNM on the device (src/cuda/wnnm/filter.cpp:16-70): one instance
    per format and size. Calling it filters one frame: ``planes`` is a list
    of 2-D float32 CUDA tensors; ``rclip`` (instances made with
    ``rclip=True``) guides the matching only. At radius > 0, ``contribution``
    gives a centre frame's num/den slices (the fat output)."""

    _NAME = "WNNM"

    def _parse(self, fmt, width, height, frames, rclip, **kwargs):
        return parse_wnnm(fmt, width, height, frames, rclip=rclip, info=self.info, **kwargs)

    def _call(self, what, *args):
        return getattr(lib(), f"nss_hip_{self._NAME.lower()}_{what}")(*args)

    def __init__(self, fmt, width, height, frames=1, rclip=False, **kwargs):
        self.fmt = fmt
        self.info = video_info(fmt, width, height, frames)
        self.params = self._parse(fmt, width, height, frames, rclip, **kwargs)
        self.device = self.params.device_id
        h = ctypes.c_void_p()
        _check(self._call("create", ctypes.byref(self.params), ctypes.byref(self.info), self.device,
                          ctypes.byref(h)))
        self._h = h
        cf, sw, sh, n = FORMATS[fmt]
        self.plane_dims = [(height >> (sh if i else 0), width >> (sw if i else 0)) for i in range(n)]

    def _check_planes(self, planes, what="plane"):
        import torch
        n = len(self.plane_dims)
        if len(planes) != n:
            raise Error(f"{NAMESPACE}.{self._NAME}: expected {n} {what}s, got {len(planes)}")
        for i, p in enumerate(planes):
            if p.dtype != torch.float32 or not p.is_cuda or p.dim() != 2 or p.stride(1) != 1:
                raise Error(f"{NAMESPACE}.{self._NAME}: {what} {i} must be a row-contiguous 2-D float32 CUDA tensor")
            if tuple(p.shape) != self.plane_dims[i]:
                raise Error(f"{NAMESPACE}.{self._NAME}: {what} {i} has shape {tuple(p.shape)}, "
                            f"expected {self.plane_dims[i]}")

    def __call__(self, planes, rclip=None, out=None, stream=None):
        import torch
        n = len(self.plane_dims)
        self._check_planes(planes)
        if rclip is not None:
            if not self.params.has_rclip:
                raise Error(f"{NAMESPACE}.{self._NAME}: this instance was made without rclip")
            self._check_planes(rclip, "rclip plane")
        if out is None:
            out = [torch.empty(self.plane_dims[i], dtype=torch.float32, device=planes[i].device) for i in range(n)]
        _check_out(self._NAME, out, self.plane_dims)
        VP, IP = ctypes.c_void_p * n, ctypes.c_int * n
        src = VP(*[p.data_ptr() for p in planes])
        sst = IP(*[p.stride(0) for p in planes])
        rp = VP(*[p.data_ptr() for p in rclip]) if rclip is not None else None
        rst = IP(*[p.stride(0) for p in rclip]) if rclip is not None else None
        dp = VP(*[p.data_ptr() for p in out])
        dst = IP(*[p.stride(0) for p in out])
        _check(self._call("process", self._h, src, sst, rp, rst, dp, dst, _stream_handle(stream)))
        return out

    def slot_frames(self, center):
        """Clip frames of slots 0 .. 2 radius around `center`
        (src/frontend/temporal.hpp:19-21)."""
        r, n = self.params.radius, self.info.num_frames
        return [center + min(max(t - r, -center), n - 1 - center) for t in range(2 * r + 1)]

    def contribution(self, center, frames, rclip=None, out=None, stream=None):
        """radius > 0: centre frame `center`'s contribution from `frames`, the
        planes of its 2 radius + 1 slot frames (slot_frames order), matched
        on `rclip`'s likewise: per plane, 2 (2 radius + 1) plane heights, slot
        t's num at rows 2t h and den at (2t + 1) h."""
        import torch
        r = self.params.radius
        ns, n = 2 * r + 1, len(self.plane_dims)
        if len(frames) != ns or (rclip is not None and len(rclip) != ns):
            raise Error(f"{NAMESPACE}.{self._NAME}: expected {ns} slot frames")
        if rclip is not None and not self.params.has_rclip:
            raise Error(f"{NAMESPACE}.{self._NAME}: this instance was made without rclip")
        for f in frames:
            self._check_planes(f)
        if rclip is not None:
            for f in rclip:
                self._check_planes(f, "rclip plane")
        dims = [(2 * ns * hh, ww) for hh, ww in self.plane_dims]
        if out is None:
            out = [torch.empty(d, dtype=torch.float32, device=frames[0][0].device) for d in dims]
        _check_out(self._NAME, out, dims)
        VP, IP = ctypes.c_void_p, ctypes.c_int
        src = (VP * (ns * n))(*[p.data_ptr() for f in frames for p in f])
        sst = (IP * (ns * n))(*[p.stride(0) for f in frames for p in f])
        rp = (VP * (ns * n))(*[p.data_ptr() for f in rclip for p in f]) if rclip is not None else None
        rst = (IP * (ns * n))(*[p.stride(0) for f in rclip for p in f]) if rclip is not None else None
        dp = (VP * n)(*[p.data_ptr() for p in out])
        dst = (IP * n)(*[p.stride(0) for p in out])
        _check(self._call("process_temporal", self._h, int(center), src, sst, rp, rst, dp, dst,
                          _stream_handle(stream)))
        return out

    def close(self):
        if getattr(self, "_h", None):
            self._call("destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class NCSR(WNNM):
    """nss.NCSR on the device (src/cuda/ncsr/filter.cpp:16-80): the WNNM
    instance's interface with NCSR's arguments (iters, delta)."""

    _NAME = "NCSR"

    def _parse(self, fmt, width, height, frames, rclip, **kwargs):
        return parse_ncsr(fmt, width, height, frames, rclip=rclip, info=self.info, **kwargs)


def grid_size(width, height, block, step):
    """(nx, ny) of the reference grid (src/host/batch_runner.hpp:17-31)."""
    if width < block or height < block or step < 1:
        return 0, 0
    return (width - block + step - 1) // step + 1, (height - block + step - 1) // step + 1


MATCH_AUTO, MATCH_FP32, MATCH_MFMA = 0, 1, 2


def set_match_path(path):
    """Matcher choice for the process: MATCH_AUTO, MATCH_FP32 (the FP32
    matchers only) or MATCH_MFMA (the MFMA matcher wherever it applies: block
    8, group <= 32). Every choice returns the same records."""
    _check(lib().nss_hip_set_match_path(int(path)))


def spatial_match(plane, block=8, step=8, bm_range=9, group=8, stream=None):
    """Batched matcher over one plane's grid. plane: 2-D float32 CUDA tensor.
    Returns (rec int32 [nref, group, 4] = x, y, t, dist bits; count int32 [nref])."""
    import torch
    h, w = plane.shape
    nx, ny = grid_size(w, h, block, step)
    rec = torch.zeros((nx * ny, group, 4), dtype=torch.int32, device=plane.device)
    cnt = torch.zeros((nx * ny,), dtype=torch.int32, device=plane.device)
    _check(lib().nss_hip_spatial_match(ctypes.c_void_p(plane.data_ptr()), w, h, plane.stride(0), block, step,
                                       bm_range, group, ctypes.c_void_p(rec.data_ptr()),
                                       ctypes.c_void_p(cnt.data_ptr()), _stream_handle(stream)))
    return rec, cnt


def predictive_match(frames, center, radius, block=8, step=8, bm_range=9, group=8, ps_num=2, ps_range=4,
                     stream=None):
    """Batched temporal matcher over one plane's grid. frames: 3-D float32
    CUDA tensor [nframes, height, width] (the clip's planes), centre frame
    `center`. Returns (rec int32 [nref, group, 4] = x, y, slot, dist bits;
    count int32 [nref])."""
    import torch
    n, h, w = frames.shape
    if frames.dtype != torch.float32 or not frames.is_cuda or frames.stride(2) != 1 or \
            frames.stride(0) != frames.stride(1) * h:
        raise Error(f"{NAMESPACE}: frames must be a contiguous [frames, height, width] float32 CUDA tensor")
    nx, ny = grid_size(w, h, block, step)
    rec = torch.zeros((nx * ny, group, 4), dtype=torch.int32, device=frames.device)
    cnt = torch.zeros((nx * ny,), dtype=torch.int32, device=frames.device)
    _check(lib().nss_hip_predictive_match(ctypes.c_void_p(frames.data_ptr()), n, center, w, h, frames.stride(1),
                                          block, step, bm_range, group, radius, ps_num, ps_range,
                                          ctypes.c_void_p(rec.data_ptr()), ctypes.c_void_p(cnt.data_ptr()),
                                          _stream_handle(stream)))
    return rec, cnt


def VAggregate(clip, src, radius, planes=None, stream=None):
    """nss.VAggregate on the device (frame logic src/host/filter_vaggregate.cpp:29-100,
    argument checks src/frontend/vaggregate_args.cpp:13-45): ``clip`` holds the
    legacy contribution frames (per frame, per plane a row-contiguous 2-D
    float32 CUDA tensor of 2 (2 radius + 1) plane heights), ``src`` the
    source frames. Output frame n is sum(num) / sum(den) over the
    contribution frames c = max(0, n - radius) .. min(N - 1, n + radius),
    slice n - c + radius, added in ascending c, or the source where nothing
    contributed; planes not in ``planes`` are copied from src."""
    import torch
    ns = f"{NAMESPACE}.VAggregate"
    if not 0 <= radius <= 16:
        raise Error(f"{ns}: radius must be in [0, 16]")
    nframes = len(src)
    nplanes = len(src[0]) if nframes else 0
    if planes is None:
        sel = [True] * nplanes
    else:
        if len(set(planes)) != len(planes) or any(not 0 <= q < nplanes for q in planes):
            raise Error(f"{ns}: planes must contain unique valid plane indices")
        sel = [q in planes for q in range(nplanes)]
    rows = 2 * (2 * radius + 1)
    shape_msg = f"{ns}: clip must match src format, width, and frame count, with height=src.height*(2*radius+1)*2"
    if len(clip) != nframes:
        raise Error(shape_msg)
    for n in range(nframes):
        if len(clip[n]) != nplanes:
            raise Error(shape_msg)
        for q in range(nplanes):
            a, b = clip[n][q], src[n][q]
            if a.shape != (b.shape[0] * rows, b.shape[1]) or a.dtype != torch.float32 or b.dtype != torch.float32:
                raise Error(shape_msg)
            if a.stride(1) != 1 or b.stride(1) != 1:
                raise Error(f"{ns}: planes must be row-contiguous")
    vp, ci = ctypes.c_void_p, ctypes.c_int
    out = []
    for n in range(nframes):
        first, last = max(0, n - radius), min(nframes - 1, n + radius)
        frame = []
        for q in range(nplanes):
            s = src[n][q]
            if not sel[q]:
                frame.append(s.clone())
                continue
            h, w = s.shape
            dst = torch.empty_like(s)
            nums, dens, strides = [], [], []
            for c in range(first, last + 1):
                sl, f = n - c + radius, clip[c][q]
                nums.append(f.data_ptr() + 2 * sl * h * f.stride(0) * 4)
                dens.append(f.data_ptr() + (2 * sl + 1) * h * f.stride(0) * 4)
                strides.append(f.stride(0))
            k = len(nums)
            _check(lib().nss_hip_vaggregate(dst.data_ptr(), dst.stride(0), (vp * k)(*nums), (vp * k)(*dens),
                                            (ci * k)(*strides), k, s.data_ptr(), s.stride(0), w, h,
                                            _stream_handle(stream)))
            frame.append(dst)
        out.append(frame)
    return out


class Server:
    """Continuous-batching serving loop (nss_hip_server_*): clip streams of
    any format, frames queued as they arrive, packed into shared launches by
    a native scheduler thread under an HBM budget."""

    def __init__(self, device=0, hbm_budget_bytes=0, max_batch_frames=8, max_inflight=2, use_graphs=True,
                 autotune=False):
        cfg = ServerConfig(device, int(hbm_budget_bytes), max_batch_frames, max_inflight, int(bool(use_graphs)),
                           int(bool(autotune)))
        h = ctypes.c_void_p()
        _check(lib().nss_hip_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self.device = device
        self._dims = []
        # request id -> (planes, out): the tensors stay referenced until the
        # request is reported done, so the caching allocator cannot hand
        # their memory to another tensor while the scheduler still uses it.
        self._held = {}

    def open_stream(self, fmt, width, height, frames=1, **kwargs):
        """A clip stream with nss.BM3D arguments; returns its id."""
        kwargs.setdefault("device_id", self.device)
        params = parse(fmt, width, height, frames, **kwargs)
        info = video_info(fmt, width, height, frames)
        sid = ctypes.c_int(-1)
        _check(lib().nss_hip_server_open_stream(self._h, ctypes.byref(params), ctypes.byref(info), ctypes.byref(sid)))
        _, sw, sh, n = FORMATS[fmt]
        self._dims.append([(height >> (sh if i else 0), width >> (sw if i else 0)) for i in range(n)])
        return sid.value

    def submit(self, stream, planes, out, sigma=0.0, ready_stream=None):
        """Queues one frame (device planes); `out` planes are written when
        the request completes. The request waits for the work queued so far
        on `ready_stream` (default: torch's current stream), where the planes
        were produced. Returns its id."""
        import torch
        dims = self._dims[stream]
        for what, ps in (("plane", planes), ("output plane", out)):
            if len(ps) != len(dims):
                raise Error(f"{NAMESPACE}.Server: expected {len(dims)} {what}s")
            for i, p in enumerate(ps):
                if p.dtype != torch.float32 or not p.is_cuda or p.stride(1) != 1 or tuple(p.shape) != dims[i]:
                    raise Error(f"{NAMESPACE}.Server: {what} {i} must be a row-contiguous float32 CUDA tensor "
                                f"of shape {dims[i]}")
        n = len(planes)
        src = (ctypes.c_void_p * n)(*[p.data_ptr() for p in planes])
        sst = (ctypes.c_int * n)(*[p.stride(0) for p in planes])
        dst = (ctypes.c_void_p * n)(*[p.data_ptr() for p in out])
        dstr = (ctypes.c_int * n)(*[p.stride(0) for p in out])
        rid = ctypes.c_longlong(-1)
        _check(lib().nss_hip_server_submit(self._h, stream, src, sst, dst, dstr, float(sigma),
                                           _stream_handle(ready_stream), ctypes.byref(rid)))
        self._held[rid.value] = (list(planes), list(out))
        return rid.value

    def query(self, rid):
        """Latency in ms once the request completed, else None."""
        lat = ctypes.c_double(0)
        rc = lib().nss_hip_server_query(self._h, rid, ctypes.byref(lat))
        if rc == BYPASS:
            return None
        self._held.pop(rid, None)
        _check(rc)
        return lat.value

    def drain(self):
        _check(lib().nss_hip_server_drain(self._h))
        self._held.clear()

    def stats(self):
        st = ServerStats()
        _check(lib().nss_hip_server_get_stats(self._h, ctypes.byref(st)))
        return {n: getattr(st, n) for n, _ in st._fields_}

    def close(self):
        if self._h:
            lib().nss_hip_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
