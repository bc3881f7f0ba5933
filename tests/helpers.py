"""Shared test helpers: run the HIP chain through the C ABI, run the oracle,
and compare them with the accuracy metrics of the parity bar.

Parity bar (float32 chain vs the float64 oracle on the same float32 tables):
  * linear output: max |GPU - FP64| / per-A-scan peak  <= 1e-4
  * log output in dB: max |GPU - FP64| over bins within 60 dB of the A-scan
    peak <= 1e-3 dB
  * every bin of log output: the linear magnitude recovered from dB
    (10^(dB/20)) has max |GPU - FP64| / per-A-scan peak <= 1e-4, the same
    bound as the linear output.  (Pointwise dB error at noise-floor bins far
    below the peak is large for any float32 chain, the reference's included.)
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from oracle import chain64 as O

LIN_TOL = 1e-4
DB_TOL = 1e-3


def np_to_code(dtype) -> int:
    from octproengine_amd import _capi
    return _capi.DTYPE_CODES[np.dtype(dtype)]


def chain_params(p: O.ChainParams, dtype=np.uint16):
    """ope_hip_params from an oracle ChainParams."""
    from octproengine_amd import _capi
    return _capi.make_params(
        signal_length=p.signal_length, ascans_per_bscan=p.ascans_per_bscan,
        bscans_per_buffer=p.bscans_per_buffer, input_type=np_to_code(dtype),
        resampling_enabled=p.resampling, interpolation=p.interpolation,
        windowing_enabled=p.windowing, dispersion_enabled=p.dispersion,
        dc_removal_enabled=p.dc_removal, dc_window_size=p.dc_window,
        log_scaling=p.log_scaling, grayscale_min=p.grayscale_min,
        grayscale_max=p.grayscale_max, multiplicator=p.multiplicator, addend=p.addend,
        fpn_removal=p.fpn_removal, fpn_bscan_count=p.fpn_bscan_count,
        fpn_continuous=p.fpn_continuous, bg_removal=p.bg_removal, bg_weight=p.bg_weight,
        bg_offset=p.bg_offset, bitshift=p.bitshift, bscan_flip=p.bscan_flip,
        sinusoidal_correction=p.sinusoidal)


def to_device(raw: np.ndarray):
    import torch
    a = np.ascontiguousarray(raw)
    if a.dtype == np.uint16:
        a = a.view(np.int16)
    elif a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(a).cuda()


class GpuChain:
    """A C-ABI context plus the curves, for frame-by-frame processing."""

    def __init__(self, p: O.ChainParams, curve, window, phasor, dtype=np.uint16, n_slots=2):
        from octproengine_amd import _capi
        self.p = p
        self.ctx = _capi.Context(chain_params(p, dtype), 0, n_slots)
        self.ctx.set_resampling_curve(curve)
        self.ctx.set_window_curve(window)
        self.ctx.set_dispersion_curve(phasor)

    def frame(self, raw: np.ndarray) -> np.ndarray:
        import torch
        n = self.p.signal_length
        a = raw.reshape(-1, n).shape[0]
        d_in = to_device(raw.reshape(-1, n))
        d_out = torch.empty((a, n // 2), dtype=torch.float32, device="cuda")
        stream = torch.cuda.current_stream().cuda_stream
        self.ctx.process_device(d_in.data_ptr(), d_out.data_ptr(), a, stream)
        torch.cuda.synchronize()
        return d_out.cpu().numpy()

    def close(self):
        self.ctx.close()


@dataclass
class Tables:
    curve: np.ndarray
    window: np.ndarray
    phasor: np.ndarray


def standard_tables(n: int, window_type=O.HANN, center=0.5, fill=0.95, d2=5e-6, d3=-2e-9,
                    custom_curve=True) -> Tables:
    from octproengine_amd import synthetic as S
    curve = S.resampling_curve(n) if custom_curve else O.generate_resampling_curve(n, (0.5, n, -100, 50))
    window = O.generate_window(n, window_type, center, fill)
    phase = S.dispersion_phase(n, d2 * (2048.0 / n) ** 2, d3 * (2048.0 / n) ** 3)
    return Tables(curve, window, O.phase_to_phasor(phase))


def db_params(p: O.ChainParams) -> O.ChainParams:
    """Same chain with an identity grayscale map, so log output is dB."""
    p.log_scaling = True
    p.grayscale_min, p.grayscale_max, p.multiplicator, p.addend = 0.0, 1.0, 1.0, 0.0
    return p


def lin_params(p: O.ChainParams) -> O.ChainParams:
    p.log_scaling = False
    p.grayscale_min, p.grayscale_max, p.multiplicator, p.addend = 0.0, 1.0, 1.0, 0.0
    return p


def cpu32(p: O.ChainParams, raw, t: Tables, fpn=None, dtype_code=None) -> np.ndarray:
    """The float32 C restatement (oracle/cpu_chain.c) through ctypes."""
    lib = cpu_lib()
    n = p.signal_length
    raw = np.ascontiguousarray(raw)
    a = raw.size // n
    code = {np.dtype(np.uint8): 1, np.dtype(np.uint16): 2, np.dtype(np.uint32): 3,
            np.dtype(np.float32): 9}[raw.dtype]
    P = CpuParams(n, a, code, int(p.resampling), int(p.interpolation), int(p.windowing),
                  int(p.dispersion), int(p.dc_removal), int(p.dc_window), int(p.log_scaling),
                  p.grayscale_min, p.grayscale_max, p.multiplicator, p.addend,
                  0 if fpn is None else len(fpn) // 2, len(t.phasor))
    out = np.zeros((a, n // 2), dtype=np.float32)
    f = None if fpn is None else np.ascontiguousarray(fpn, dtype=np.float32)
    lib.ope_cpu_process(ctypes.byref(P), raw.ctypes.data, _fp(t.curve), _fp(t.window), _fp(t.phasor),
                        None if f is None else f.ctypes.data, out.ctypes.data, 1)
    return out


def _fp(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.ctypes.data_as(ctypes.c_void_p)


class CpuParams(ctypes.Structure):
    _fields_ = [("signal_length", ctypes.c_int32), ("n_ascans", ctypes.c_int32),
                ("dtype", ctypes.c_int32), ("resampling", ctypes.c_int32),
                ("interpolation", ctypes.c_int32), ("windowing", ctypes.c_int32),
                ("dispersion", ctypes.c_int32), ("dc_removal", ctypes.c_int32),
                ("dc_window", ctypes.c_int32), ("log_scaling", ctypes.c_int32),
                ("grayscale_min", ctypes.c_float), ("grayscale_max", ctypes.c_float),
                ("multiplicator", ctypes.c_float), ("addend", ctypes.c_float),
                ("fpn_pairs", ctypes.c_int32), ("phasor_len", ctypes.c_int32)]


_CPU = None


def cpu_lib():
    global _CPU
    if _CPU is None:
        from pathlib import Path
        path = Path(__file__).resolve().parents[1] / "oracle" / "libope_cpu.so"
        if not path.exists():
            import subprocess
            subprocess.run(["make", "-C", str(path.parent), "libope_cpu.so"], check=True)
        _CPU = ctypes.CDLL(str(path))
        _CPU.ope_cpu_process.argtypes = [ctypes.c_void_p] * 7 + [ctypes.c_int]
        _CPU.ope_cpu_process.restype = ctypes.c_int
        _CPU.ope_cpu_min_variance_mean.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                   ctypes.c_int, ctypes.c_void_p]
        _CPU.ope_cpu_ifft.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        _CPU.ope_cpu_spectra.argtypes = [ctypes.c_void_p] * 6 + [ctypes.c_int, ctypes.c_void_p]
    return _CPU


def db_error(test_db, ref_db, within=60.0) -> float:
    return O.db_error_near_peak(np.asarray(test_db, np.float64), np.asarray(ref_db, np.float64), within)


def db_lin_error(test_db, ref_db) -> float:
    """Peak-normalised linear-magnitude error recovered from dB outputs
    (|X| is proportional to 10^(dB/20)); covers every bin."""
    t = 10.0 ** (np.asarray(test_db, np.float64) / 20.0)
    r = 10.0 ** (np.asarray(ref_db, np.float64) / 20.0)
    return O.peak_normalized_lin_error(t, r)


def lin_error(test_lin, ref_lin) -> float:
    return O.peak_normalized_lin_error(np.asarray(test_lin, np.float64), np.asarray(ref_lin, np.float64))
