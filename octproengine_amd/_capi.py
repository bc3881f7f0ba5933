"""ctypes binding of the C ABI in include/ope_hip.h (libope_hip.so).

The library is loaded from this package directory; there is no fallback:
if it is missing or fails to load, importing this module raises.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, c_char_p, c_float, c_int, c_int32, c_int64, c_size_t,
                    c_void_p)
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libope_hip.so"

OK, EINVAL, EUNSUP, ERUNTIME, ENOMEM, ESTATE = 0, -1, -2, -3, -4, -5

# ope::DataType values (include/types.h)
UINT8, UINT16, UINT32, INT8, INT16, INT32, FLOAT32 = 1, 2, 3, 5, 6, 7, 9
DTYPE_CODES = {np.dtype(np.uint8): UINT8, np.dtype(np.uint16): UINT16,
               np.dtype(np.uint32): UINT32, np.dtype(np.int8): INT8,
               np.dtype(np.int16): INT16, np.dtype(np.int32): INT32,
               np.dtype(np.float32): FLOAT32}
LINEAR, CUBIC, LANCZOS = 0, 1, 2
LANCZOS4 = 3  # C ABI only: Lanczos with a = 4 (8 taps)
# sample offset modes (ope_hip_set_sample_offset)
OFFSET_OFF, OFFSET_AUTO, OFFSET_FIXED = 0, 1, 2

#: Every function include/ope_hip.h declares, with its ctypes signature.
SIGNATURES = {
    "ope_hip_version": (c_char_p, []),
    "ope_hip_device_count": (c_int, [POINTER(c_int)]),
    "ope_hip_create": (c_int, [c_int, c_void_p, c_int, POINTER(c_void_p)]),
    "ope_hip_destroy": (c_int, [c_void_p]),
    "ope_hip_last_error": (c_char_p, [c_void_p]),
    "ope_hip_set_params": (c_int, [c_void_p, c_void_p]),
    "ope_hip_set_resampling_curve": (c_int, [c_void_p, c_void_p, c_size_t]),
    "ope_hip_set_window_curve": (c_int, [c_void_p, c_void_p, c_size_t]),
    "ope_hip_set_dispersion_curve": (c_int, [c_void_p, c_void_p, c_size_t]),
    "ope_hip_set_fpn_profile": (c_int, [c_void_p, c_void_p, c_size_t]),
    "ope_hip_get_fpn_profile": (c_int, [c_void_p, c_void_p, c_size_t, POINTER(c_size_t)]),
    "ope_hip_request_fpn": (c_int, [c_void_p]),
    "ope_hip_set_bg_profile": (c_int, [c_void_p, c_void_p, c_size_t]),
    "ope_hip_get_bg_profile": (c_int, [c_void_p, c_void_p, c_size_t, POINTER(c_size_t)]),
    "ope_hip_request_bg_recording": (c_int, [c_void_p]),
    "ope_hip_set_sample_offset": (c_int, [c_void_p, c_int, c_float]),
    "ope_hip_get_sample_offset": (c_int, [c_void_p, POINTER(c_int), POINTER(c_float)]),
    "ope_hip_process_device": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p]),
    "ope_hip_submit_host": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p]),
    "ope_hip_synchronize": (c_int, [c_void_p]),
    "ope_hip_op_min_variance_mean": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "ope_hip_op_min_variance_mean_device": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p,
                                                    c_void_p]),
    "ope_hip_op_convert_input": (c_int, [c_void_p, c_void_p, c_int, c_int64, c_int, c_void_p]),
    "ope_hip_op_klinearization": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "ope_hip_op_windowing": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "ope_hip_op_dispersion": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "ope_hip_op_ifft_magnitude": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "ope_hip_op_fpn_removal": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p]),
    "ope_hip_host_alloc": (c_int, [c_size_t, POINTER(c_void_p)]),
    "ope_hip_host_free": (c_int, [c_void_p]),
    "ope_hip_autofocus_loss": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    "ope_hip_autofocus_grad": (c_int, [c_void_p, c_void_p, c_int64, c_void_p, ctypes.c_double,
                                       ctypes.c_double, c_void_p, c_void_p]),
    "ope_hip_autotune": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    "ope_hip_set_launch_variant": (c_int, [c_void_p, c_int, c_int, c_int]),
    "ope_hip_graph_create": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int64, POINTER(c_void_p)]),
    "ope_hip_graph_launch": (c_int, [c_void_p, c_void_p]),
    "ope_hip_graph_last_error": (c_char_p, [c_void_p]),
    "ope_hip_graph_destroy": (c_int, [c_void_p]),
    "ope_hip_ragged_create": (c_int, [c_void_p, c_int, c_void_p, c_int, c_int, POINTER(c_void_p)]),
    "ope_hip_ragged_run": (c_int, [c_void_p, c_void_p]),
    "ope_hip_ragged_latencies": (c_int, [c_void_p, c_void_p]),
    "ope_hip_ragged_last_error": (c_char_p, [c_void_p]),
    "ope_hip_ragged_destroy": (c_int, [c_void_p]),
}


FRAME_CB = ctypes.CFUNCTYPE(None, c_void_p, c_void_p)
INPUT_CB = ctypes.CFUNCTYPE(None, c_void_p)


class Params(Structure):
    """ope_hip_params (include/ope_hip.h)."""
    _fields_ = [
        ("signal_length", c_int32), ("ascans_per_bscan", c_int32),
        ("bscans_per_buffer", c_int32), ("input_type", c_int32),
        ("resampling_enabled", c_int32), ("interpolation", c_int32),
        ("windowing_enabled", c_int32), ("dispersion_enabled", c_int32),
        ("dc_removal_enabled", c_int32), ("dc_window_size", c_int32),
        ("log_scaling", c_int32), ("grayscale_min", c_float), ("grayscale_max", c_float),
        ("multiplicator", c_float), ("addend", c_float), ("fpn_removal", c_int32),
        ("fpn_bscan_count", c_int32), ("fpn_continuous", c_int32), ("bg_removal", c_int32),
        ("bg_weight", c_float), ("bg_offset", c_float), ("bitshift", c_int32), ("bscan_flip", c_int32),
        ("sinusoidal_correction", c_int32),
    ]


class Request(Structure):
    """ope_hip_request (include/ope_hip.h)."""
    _fields_ = [("d_in", c_void_p), ("d_out", c_void_p), ("n_ascans", c_int64),
                ("signal_length", c_int32), ("reserved", c_int32)]


def make_params(**kw) -> Params:
    """Params with the reference's defaults (processorconfiguration.h:33-172)."""
    d = dict(signal_length=1024, ascans_per_bscan=512, bscans_per_buffer=1, input_type=UINT16,
             resampling_enabled=0, interpolation=LINEAR, windowing_enabled=0,
             dispersion_enabled=0, dc_removal_enabled=0, dc_window_size=64, log_scaling=1,
             grayscale_min=30.0, grayscale_max=100.0, multiplicator=1.0, addend=0.0,
             fpn_removal=0, fpn_bscan_count=1, fpn_continuous=0, bg_removal=0,
             bg_weight=1.0, bg_offset=0.0, bitshift=0, bscan_flip=0, sinusoidal_correction=0)
    unknown = set(kw) - set(d)
    if unknown:
        raise TypeError(f"unknown parameters: {sorted(unknown)}")
    d.update(kw)
    p = Params()
    for k, v in d.items():
        setattr(p, k, int(v) if isinstance(v, (bool, np.bool_)) else v)
    return p


def _load() -> ctypes.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run `python -m octproengine_amd.build`")
    lib = ctypes.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | ctypes.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class HipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def device_count() -> int:
    n = c_int(0)
    lib.ope_hip_device_count(ctypes.byref(n))
    return n.value


class Context:
    """One ope_hip_ctx: the device-side state of a HIP processing backend."""

    def __init__(self, params: Params, device: int = 0, n_slots: int = 2):
        self.params = params
        h = c_void_p()
        rc = lib.ope_hip_create(device, ctypes.byref(params), n_slots, ctypes.byref(h))
        if rc != OK:
            raise HipError(rc, lib.ope_hip_last_error(None).decode())
        self._h = h

    # -- plumbing
    def _check(self, rc: int) -> None:
        if rc != OK:
            raise HipError(rc, lib.ope_hip_last_error(self._h).decode())

    def close(self) -> None:
        if getattr(self, "_h", None) and lib is not None:
            lib.ope_hip_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self) -> int:
        return self._h.value

    # -- configuration
    def set_params(self, params: Params) -> None:
        self._check(lib.ope_hip_set_params(self._h, ctypes.byref(params)))
        self.params = params

    def set_resampling_curve(self, curve) -> None:
        c = _f32(curve)
        self._check(lib.ope_hip_set_resampling_curve(self._h, c.ctypes.data, c.size))

    def set_window_curve(self, curve) -> None:
        c = _f32(curve)
        self._check(lib.ope_hip_set_window_curve(self._h, c.ctypes.data, c.size))

    def set_dispersion_curve(self, cos_sin) -> None:
        c = _f32(cos_sin)
        self._check(lib.ope_hip_set_dispersion_curve(self._h, c.ctypes.data, c.size))

    def set_fpn_profile(self, interleaved) -> None:
        c = _f32(interleaved)
        self._check(lib.ope_hip_set_fpn_profile(self._h, c.ctypes.data, c.size // 2))

    def get_fpn_profile(self) -> np.ndarray:
        n = c_size_t(0)
        half = self.params.signal_length // 2
        buf = np.zeros(2 * half, dtype=np.float32)
        self._check(lib.ope_hip_get_fpn_profile(self._h, buf.ctypes.data, half, ctypes.byref(n)))
        return buf[: 2 * n.value]

    def request_fpn(self) -> None:
        self._check(lib.ope_hip_request_fpn(self._h))

    def set_bg_profile(self, profile) -> None:
        c = _f32(profile)
        self._check(lib.ope_hip_set_bg_profile(self._h, c.ctypes.data, c.size))

    def get_bg_profile(self) -> np.ndarray:
        n = c_size_t(0)
        half = self.params.signal_length // 2
        buf = np.zeros(half, dtype=np.float32)
        self._check(lib.ope_hip_get_bg_profile(self._h, buf.ctypes.data, half, ctypes.byref(n)))
        return buf[: n.value]

    def request_bg_recording(self) -> None:
        self._check(lib.ope_hip_request_bg_recording(self._h))

    # -- processing
    def process_device(self, d_in: int, d_out: int, n_ascans: int, stream: int = 0) -> None:
        """Device pointers (ints), asynchronous on `stream` (hipStream_t as int)."""
        self._check(lib.ope_hip_process_device(self._h, c_void_p(d_in), c_void_p(d_out),
                                               int(n_ascans), c_void_p(stream or None)))

    def min_variance_mean(self, rows: np.ndarray, segments: int = 8) -> np.ndarray:
        """rows: complex [height, width] -> interleaved float32 [2*width]."""
        rows = np.asarray(rows)
        h, w = rows.shape
        inter = np.empty((h, w, 2), dtype=np.float32)
        inter[..., 0] = rows.real
        inter[..., 1] = rows.imag
        out = np.zeros(2 * w, dtype=np.float32)
        self._check(lib.ope_hip_op_min_variance_mean(self._h, inter.ctypes.data, w, h, segments,
                                                     out.ctypes.data))
        return out

    def min_variance_mean_device(self, d_input: int, width: int, height: int, segments: int, d_out: int,
                                 stream: int = 0) -> None:
        """The allocation-free form on device memory: d_input interleaved
        complex [height][width][2], d_out 2*width floats; asynchronous."""
        self._check(lib.ope_hip_op_min_variance_mean_device(self._h, c_void_p(d_input), width, height, segments,
                                                            c_void_p(d_out), c_void_p(stream or None)))

    def submit_host(self, host_in: int, host_out: int, on_done=None, on_input_free=None) -> None:
        """Host frame: H2D, chain, D2H on the context's streams; returns once
        enqueued.  Callbacks run on a HIP runtime thread."""
        done = FRAME_CB(lambda user, out: on_done()) if on_done else FRAME_CB()
        free = INPUT_CB(lambda user: on_input_free()) if on_input_free else INPUT_CB()
        self._keep = getattr(self, "_keep", [])[-16:] + [done, free]  # alive until fired
        self._check(lib.ope_hip_submit_host(self._h, c_void_p(host_in), c_void_p(host_out),
                                            ctypes.cast(free, c_void_p), None,
                                            ctypes.cast(done, c_void_p), None))

    def synchronize(self) -> None:
        self._check(lib.ope_hip_synchronize(self._h))

    def set_sample_offset(self, mode: int, value: float = 0.0) -> None:
        self._check(lib.ope_hip_set_sample_offset(self._h, int(mode), float(value)))

    def get_sample_offset(self):
        """(mode, offset in use; 0 while undetermined)."""
        mode, value = c_int(0), c_float(0.0)
        self._check(lib.ope_hip_get_sample_offset(self._h, ctypes.byref(mode), ctypes.byref(value)))
        return int(mode.value), float(value.value)

    def autotune(self, d_in: int = 0, d_out: int = 0, n_ascans: int = 0, stream: int = 0):
        """Times the chain's launch variants on d_in -> d_out (the context's
        own frame slot when 0) and keeps the fastest for this configuration.
        Returns ({as, sched, per_cu}, microseconds per launch)."""
        choice = np.zeros(3, dtype=np.int32)
        best = ctypes.c_float(0.0)
        self._check(lib.ope_hip_autotune(self._h, c_void_p(d_in or None), c_void_p(d_out or None), int(n_ascans),
                                         c_void_p(stream or None), choice.ctypes.data, ctypes.byref(best)))
        return {"ascans_per_lane": int(choice[0]), "small_radix_pass": int(choice[1]),
                "workgroups_per_cu": int(choice[2])}, float(best.value)

    def set_launch_variant(self, ascans_per_lane: int, small_radix_pass: int, workgroups_per_cu: int = 0) -> None:
        self._check(lib.ope_hip_set_launch_variant(self._h, ascans_per_lane, small_radix_pass, workgroups_per_cu))

    # -- autofocus (theta = d2, d3, c0, c1, c2, c3)
    def autofocus_loss(self, d_in: int, n_ascans: int, theta, stream: int = 0):
        """(loss, S, sum I log I) of the entropy of the frame at d_in."""
        th = np.ascontiguousarray(theta, dtype=np.float64)
        out = np.zeros(3, dtype=np.float64)
        self._check(lib.ope_hip_autofocus_loss(self._h, c_void_p(d_in), int(n_ascans), th.ctypes.data,
                                               out.ctypes.data, c_void_p(stream or None)))
        return float(out[0]), float(out[1]), float(out[2])

    def autofocus_grad(self, d_in: int, n_ascans: int, theta, sum_i: float, loss: float,
                       stream: int = 0) -> np.ndarray:
        th = np.ascontiguousarray(theta, dtype=np.float64)
        out = np.zeros(6, dtype=np.float64)
        self._check(lib.ope_hip_autofocus_grad(self._h, c_void_p(d_in), int(n_ascans), th.ctypes.data,
                                               float(sum_i), float(loss), out.ctypes.data,
                                               c_void_p(stream or None)))
        return out


class RaggedBatch:
    """ope_hip_ragged: requests of mixed signal length in one persistent
    launch.  `contexts` supply the chain of each length; `requests` is a
    sequence of (d_in, d_out, n_ascans, signal_length) with device pointers
    as ints."""

    def __init__(self, contexts, requests, chunk_ascans: int = 0):
        self.contexts = list(contexts)  # kept alive: the batch points at their tables
        handles = (c_void_p * len(self.contexts))(*[c._h for c in self.contexts])
        reqs = (Request * max(len(requests), 1))()
        for i, (d_in, d_out, n, length) in enumerate(requests):
            reqs[i] = Request(c_void_p(d_in), c_void_p(d_out), int(n), int(length), 0)
        self.n_reqs = len(requests)
        h = c_void_p()
        rc = lib.ope_hip_ragged_create(handles, len(self.contexts), reqs, self.n_reqs, int(chunk_ascans),
                                       ctypes.byref(h))
        if rc != OK:
            raise HipError(rc, lib.ope_hip_ragged_last_error(None).decode())
        self._h = h

    def _check(self, rc: int) -> None:
        if rc != OK:
            raise HipError(rc, lib.ope_hip_ragged_last_error(self._h).decode())

    def run(self, stream: int = 0) -> None:
        self._check(lib.ope_hip_ragged_run(self._h, c_void_p(stream or None)))

    def latencies_us(self) -> np.ndarray:
        out = np.zeros(max(self.n_reqs, 1), dtype=np.float64)
        self._check(lib.ope_hip_ragged_latencies(self._h, out.ctypes.data))
        return out[:self.n_reqs]

    def close(self) -> None:
        if getattr(self, "_h", None) and lib is not None:
            lib.ope_hip_ragged_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()


class FrameGraph:
    """ope_hip_graph: the chain over a fixed ring of device frames captured
    into one HIP graph (steady state only)."""

    def __init__(self, context: Context, d_in, d_out, ascans_per_frame: int):
        self.context = context
        n = len(d_in)
        ins = (c_void_p * n)(*[c_void_p(p) for p in d_in])
        outs = (c_void_p * n)(*[c_void_p(p) for p in d_out])
        h = c_void_p()
        rc = lib.ope_hip_graph_create(context._h, ins, outs, n, int(ascans_per_frame), ctypes.byref(h))
        if rc != OK:
            raise HipError(rc, lib.ope_hip_graph_last_error(None).decode())
        self._h = h

    def launch(self, stream: int = 0) -> None:
        rc = lib.ope_hip_graph_launch(self._h, c_void_p(stream or None))
        if rc != OK:
            raise HipError(rc, lib.ope_hip_graph_last_error(self._h).decode())

    def close(self) -> None:
        if getattr(self, "_h", None) and lib is not None:
            lib.ope_hip_graph_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()
