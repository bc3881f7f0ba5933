"""FP64 NumPy restatement of OCTproEngine's CPU processing chain.

TEST INFRASTRUCTURE ONLY.  This module is the oracle that the HIP backend is
checked against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker.
No product code path imports it.

Every function restates one step of the reference CPU backend in float64
arithmetic, consuming the exact float32 tables the backends receive
(resampling positions, window, interleaved cos/sin phasor), so that the
difference between a float32 implementation and this oracle measures
arithmetic error only.  Citations are ``path:line`` in the reference tree.

Pinning (what checks each function, see tests/test_oracle.py):
  convert_input            exact integer -> float (closed form)
  rolling_average_dc       brute-force window mean per sample
  klin_linear              exact on affine data; equals numpy.interp inside
  klin_cubic               reproduces quadratics exactly; y(t=0) = y1
  klin_lanczos             interpolating (integer positions return samples);
                           brute-force closed form of sinc*sinc/8
  ifft_unnormalized        brute-force O(N^2) DFT, Parseval, single-tone peak
  min_variance_mean        brute-force segment statistics on tiny inputs
  log/lin scale            closed forms; auto-min/max per A-scan
  generate_* curves        bit-exact against curves written by the
                           reference's own processorconfiguration.cpp
                           (oracle/_ref, fixtures in tests/golden/)
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# Interpolation / window enums, same values as the reference
# (include/processorconfiguration.h:17-31).
LINEAR, CUBIC, LANCZOS = 0, 1, 2
# Lanczos with a = 4: the reference's kernel (cpu_kernels.tpp:148-162) at the
# other window size it takes; its kLinearizationLanczos fixes a = 8 (:174)
LANCZOS4 = 3
HANN, GAUSS, SINE, LANCZOS_WIN, RECTANGULAR, FLAT_TOP = 0, 1, 2, 3, 4, 5

FPN_SEGMENTS = 8  # src/backends/cpu/cpu_backend.cpp:306


# ---------------------------------------------------------------------------
# Per-A-scan stages (src/backends/cpu/cpu_kernels.tpp)
# ---------------------------------------------------------------------------

def convert_input(raw: np.ndarray, ieee_float32: bool = True, bitshift: bool = False) -> np.ndarray:
    """Raw samples -> float64 (cpu_kernels.tpp:24-54).

    The reference reads by bit depth (<=8 -> uint8, <=16 -> uint16, else
    uint32).  For FLOAT32 input the reference reinterprets the bits as uint32
    (cpu_kernels.tpp:47-52 with types.cpp:15); this build reads IEEE floats
    (``ieee_float32=True``), a documented divergence.
    ``bitshift`` is the CUDA backend's input option
    (cuda_kernels.cu:45-70): 8/16-bit samples >> 4, 32-bit integers / 2^32;
    it leaves IEEE float input unchanged.
    """
    if raw.dtype == np.float32 and not ieee_float32:
        raw = raw.view(np.uint32)
    if bitshift and raw.dtype != np.float32:
        if raw.dtype.itemsize <= 2:
            return (raw.astype(np.uint32) >> 4).astype(np.float64)
        return raw.astype(np.float64) / 4294967296.0
    return raw.astype(np.float64)


def rolling_average_dc(x: np.ndarray, window: int) -> np.ndarray:
    """Rolling-average DC removal along the last axis (cpu_kernels.tpp:56-82).

    For sample i the window is [max(i-(W-1), 0), min(i+W, N-1)] inclusive
    (note the asymmetric upper bound of the reference).  W == 0 follows the
    reference's unsigned arithmetic: the lower bound is 0 and the upper is i.
    """
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[-1]
    i = np.arange(n)
    if window >= 1:
        start = np.where(i >= window - 1, i - (window - 1), 0)
        end = np.minimum(i + window, n - 1)
    elif window == 0:
        start = np.zeros(n, dtype=np.int64)
        end = i.copy()
    else:
        raise ValueError("window must be >= 0")
    cs = np.concatenate([np.zeros(x.shape[:-1] + (1,)), np.cumsum(x, axis=-1)], axis=-1)
    s = cs[..., end + 1] - cs[..., start]
    return x - s / (end - start + 1)


def _trunc_int(v: np.ndarray) -> np.ndarray:
    """C ``static_cast<int>(float)``: truncation toward zero."""
    return np.trunc(np.asarray(v, dtype=np.float64)).astype(np.int64)


def klin_linear(x: np.ndarray, curve: np.ndarray) -> np.ndarray:
    """Linear k-linearization (cpu_kernels.tpp:84-111).

    n1 = int(r), n2 = n1+1, both clamped to N-1 from above; the fractional
    part is taken against the clamped n1.
    """
    n = x.shape[-1]
    r = np.asarray(curve, dtype=np.float32).astype(np.float64)
    n1 = _trunc_int(r)
    n2 = n1 + 1
    n1 = np.minimum(n1, n - 1)
    n2 = np.minimum(n2, n - 1)
    y1 = x[..., n1]
    y2 = x[..., n2]
    return y1 + (y2 - y1) * (r - n1)


def cubic_hermite(y0, y1, y2, y3, t):
    """Catmull-Rom spline (cpu_kernels.tpp:486-502)."""
    a = -y0 + 3.0 * (y1 - y2) + y3
    b = 2.0 * y0 - 5.0 * y1 + 4.0 * y2 - y3
    c = -y0 + y2
    return 0.5 * t * (a * t * t + b * t + c) + y1


def klin_cubic(x: np.ndarray, curve: np.ndarray) -> np.ndarray:
    """Cubic k-linearization (cpu_kernels.tpp:113-146).

    Taps n0 = |n1-1|, n1, n1+1, n1+2, each clamped to N-1 from above.
    """
    n = x.shape[-1]
    r = np.asarray(curve, dtype=np.float32).astype(np.float64)
    n1 = _trunc_int(r)
    n0 = np.abs(n1 - 1)
    n2 = n1 + 1
    n3 = n2 + 1
    n0, n1, n2, n3 = (np.minimum(v, n - 1) for v in (n0, n1, n2, n3))
    return cubic_hermite(x[..., n0], x[..., n1], x[..., n2], x[..., n3], r - n1)


def lanczos_kernel(d: np.ndarray, a: int = 8) -> np.ndarray:
    """sinc(d)*sinc(d/a) with the reference's guards (cpu_kernels.tpp:148-162)."""
    d = np.asarray(d, dtype=np.float64)
    out = np.zeros_like(d)
    small = np.abs(d) < 1e-5
    inside = (~small) & (np.abs(d) < a)
    pd = math.pi * d[inside]
    out[inside] = (np.sin(pd) / pd) * (np.sin(pd / a) / (pd / a))
    out[small] = 1.0
    return out


def klin_lanczos(x: np.ndarray, curve: np.ndarray, a: int = 8) -> np.ndarray:
    """Lanczos-a k-linearization, 2a taps i = -a+1..a (cpu_kernels.tpp:164-196;
    the reference uses a = 8)."""
    n = x.shape[-1]
    r = np.asarray(curve, dtype=np.float32).astype(np.float64)
    n0 = _trunc_int(r)
    y = np.zeros(x.shape[:-1] + (r.shape[0],))
    for i in range(-a + 1, a + 1):
        idx = np.clip(n0 + i, 0, n - 1)
        y = y + x[..., idx] * lanczos_kernel(r - idx, a)
    return y


def apply_window(y: np.ndarray, window: np.ndarray) -> np.ndarray:
    """Real window multiply (cpu_kernels.tpp:227-235)."""
    return y * np.asarray(window, dtype=np.float32).astype(np.float64)


def dispersion(y: np.ndarray, phasor_interleaved: np.ndarray) -> np.ndarray:
    """Real x complex phasor (cpu_kernels.tpp:198-225, cpu_backend.cpp:492-501).

    The reference silently skips the step if the phasor length differs from
    the A-scan length (cpu_kernels.tpp:205-208).
    """
    p = np.asarray(phasor_interleaved, dtype=np.float32).astype(np.float64)
    c = p[0::2] + 1j * p[1::2]
    if c.shape[0] != y.shape[-1]:
        return y.astype(np.complex128)
    return y * c


def ifft_unnormalized(s: np.ndarray) -> np.ndarray:
    """X[k] = sum_j s[j] exp(+2 pi i j k / N), no 1/N (cpu_backend.cpp:385-391,
    cpu_kernels.tpp:238-267 with the normalisation commented out at :265)."""
    return np.fft.ifft(s, axis=-1, norm="forward")


def min_variance_mean(rows: np.ndarray, segments: int = FPN_SEGMENTS) -> np.ndarray:
    """Per-column complex mean of the minimum-variance segment
    (cpu_kernels.tpp:402-458).  ``rows`` is complex [height, width].

    Returns interleaved float32 [width*2].  The statistics follow the
    reference's mixed precision: sums in float64 of float32 samples, factor
    = float32(1/segWidth), means rounded to float32, variance rounded to
    float32, strict ``<`` so the first minimum wins.
    """
    rows = np.asarray(rows)
    height, width = rows.shape
    if height == 0 or width <= 0 or segments <= 0:
        return np.zeros(0, dtype=np.float32)
    seg_w = height // segments
    with np.errstate(divide="ignore", invalid="ignore"):
        factor = np.float64(np.float32(1.0) / np.float32(seg_w)) if seg_w else np.float64(np.inf)
        re = rows.real.astype(np.float32).astype(np.float64)
        im = rows.imag.astype(np.float32).astype(np.float64)
        best_var = np.full(width, np.inf, dtype=np.float32)
        best_x = np.zeros(width, dtype=np.float32)
        best_y = np.zeros(width, dtype=np.float32)
        for s in range(segments):
            sl = slice(s * seg_w, (s + 1) * seg_w)
            sx = re[sl].sum(axis=0)
            sy = im[sl].sum(axis=0)
            sxx = (re[sl] ** 2 + im[sl] ** 2).sum(axis=0)
            mx = (sx * factor).astype(np.float32)
            my = (sy * factor).astype(np.float32)
            var = (sxx * factor - (mx.astype(np.float64) ** 2 + my.astype(np.float64) ** 2)).astype(np.float32)
            take = var < best_var
            best_var = np.where(take, var, best_var)
            best_x = np.where(take, mx, best_x)
            best_y = np.where(take, my, best_y)
    out = np.empty(2 * width, dtype=np.float32)
    out[0::2] = best_x
    out[1::2] = best_y
    return out


def mean_aline_subtraction(X: np.ndarray, mean_interleaved: np.ndarray | None) -> np.ndarray:
    """X[k] -= m[k] for k < min(N/2, pairs) (cpu_kernels.tpp:460-479)."""
    if mean_interleaved is None or len(mean_interleaved) == 0:
        return X
    m = np.asarray(mean_interleaved, dtype=np.float32).astype(np.float64)
    half = X.shape[-1] // 2
    use = min(half, m.shape[0] // 2)
    X = X.copy()
    X[..., :use] -= m[0:2 * use:2] + 1j * m[1:2 * use:2]
    return X


def log_scale_and_truncate(X, coeff, vmin, vmax, addend, auto_minmax):
    """10*log10(|X|^2/(N/2)) mapped to grayscale, k < N/2
    (cpu_kernels.tpp:301-341).  ``auto_minmax`` takes min/max per A-scan."""
    half = X.shape[-1] // 2
    m2 = np.abs(X[..., :half]) ** 2
    with np.errstate(divide="ignore"):
        v = 10.0 * np.log10(m2 / half)
    if auto_minmax:
        vmin = v.min(axis=-1, keepdims=True)
        vmax = v.max(axis=-1, keepdims=True)
        rng = vmax - vmin
        rng = np.where(rng < 1e-6, 1.0, rng)
    else:
        rng = vmax - vmin
        if rng < 1e-6:
            rng = 1.0
    return coeff * ((v - vmin) / rng + addend)


def linear_scale_and_truncate(X, coeff, vmin, vmax, addend):
    """|X|/(N/2) mapped to grayscale, k < N/2 (cpu_kernels.tpp:343-378)."""
    half = X.shape[-1] // 2
    mag = np.abs(X[..., :half]) / half
    rng = vmax - vmin
    if rng < 1e-6:
        rng = 1.0
    return coeff * ((mag - vmin) / rng + addend)


def record_background(out: np.ndarray) -> np.ndarray:
    """Mean over all A-scans of the processed frame (cpu_backend.cpp:105-118)."""
    return out.reshape(-1, out.shape[-1]).mean(axis=0)


def subtract_background(out, bg, weight, offset):
    """out -= weight*bg + offset (cpu_kernels.tpp:380-400)."""
    return out - (weight * np.asarray(bg, dtype=np.float64) + offset)


# ---------------------------------------------------------------------------
# Frame orchestration (src/backends/cpu/cpu_backend.cpp:169-357)
# ---------------------------------------------------------------------------

@dataclass
class ChainParams:
    """Scalar/flag snapshot of ProcessorConfiguration used by processData
    (include/processorconfiguration.h:33-172)."""
    signal_length: int
    ascans_per_bscan: int
    bscans_per_buffer: int = 1
    resampling: bool = False
    interpolation: int = LINEAR
    windowing: bool = False
    dispersion: bool = False
    dc_removal: bool = False
    dc_window: int = 64
    log_scaling: bool = True
    grayscale_min: float = 30.0
    grayscale_max: float = 100.0
    multiplicator: float = 1.0
    addend: float = 0.0
    fpn_removal: bool = False
    fpn_bscan_count: int = 1
    fpn_continuous: bool = False
    bg_removal: bool = False
    bg_weight: float = 1.0
    bg_offset: float = 0.0
    bitshift: bool = False
    bscan_flip: bool = False
    sinusoidal: bool = False


@dataclass
class ChainState:
    """Cross-frame state of the CPU backend: the FPN accumulation
    (cpu_backend.cpp:54-59,305-357) and the background profile
    (cpu_backend.cpp:65,105-118)."""
    fpn_rows: list = field(default_factory=list)
    fpn_count: int = 0
    fpn_profile: np.ndarray | None = None
    fpn_requested: bool = False
    bg_profile: np.ndarray | None = None
    bg_requested: bool = False


def spectra(raw: np.ndarray, p: ChainParams, curve, window, phasor, ieee_float32=True) -> np.ndarray:
    """Complex IFFT output per A-scan, before FPN (cpu_backend.cpp:184-243)."""
    n = p.signal_length
    x = convert_input(raw, ieee_float32, p.bitshift).reshape(-1, n)
    if p.dc_removal:
        x = rolling_average_dc(x, p.dc_window)
    if p.resampling:
        if p.interpolation == LINEAR:
            y = klin_linear(x, curve)
        elif p.interpolation == CUBIC:
            y = klin_cubic(x, curve)
        elif p.interpolation == LANCZOS4:
            y = klin_lanczos(x, curve, a=4)
        else:
            y = klin_lanczos(x, curve)
    else:
        y = x
    if p.windowing:
        y = apply_window(y, window)
    s = dispersion(y, phasor) if p.dispersion else y.astype(np.complex128)
    return ifft_unnormalized(s)


def fpn_update(state: ChainState, X: np.ndarray, p: ChainParams) -> None:
    """computeFixedPatternNoiseIfRequested (cpu_backend.cpp:305-357).

    Only the first ``required`` accumulated A-scans ever enter a
    determination, so only those are kept.
    """
    required = p.fpn_bscan_count * p.ascans_per_bscan
    keep = max(0, required - len(state.fpn_rows))
    state.fpn_rows.extend(list(X[:keep]))
    state.fpn_count += X.shape[0]
    if not (state.fpn_requested or p.fpn_continuous):
        return
    if state.fpn_count < required:
        return
    rows = np.asarray(state.fpn_rows[:required])
    mean = min_variance_mean(rows, FPN_SEGMENTS)
    state.fpn_profile = mean[: 2 * (p.signal_length // 2)].copy()
    if p.fpn_continuous:
        state.fpn_count = min(state.fpn_count, required)
    else:
        state.fpn_rows = []
        state.fpn_count = 0
    state.fpn_requested = False


def process_frame(raw, p: ChainParams, curve, window, phasor, state: ChainState | None = None,
                  ieee_float32=True) -> np.ndarray:
    """One frame through processData (cpu_backend.cpp:169-302) in float64.

    Returns [A, N/2] float64.
    """
    if state is None:
        state = ChainState()
    X = spectra(raw, p, curve, window, phasor, ieee_float32)
    fpn_update(state, X, p)
    if p.fpn_removal and state.fpn_profile is not None and len(state.fpn_profile):
        X = mean_aline_subtraction(X, state.fpn_profile)
    if p.log_scaling:
        out = log_scale_and_truncate(X, p.multiplicator, p.grayscale_min, p.grayscale_max,
                                     p.addend, p.grayscale_min == p.grayscale_max)
    else:
        out = linear_scale_and_truncate(X, p.multiplicator, p.grayscale_min, p.grayscale_max, p.addend)
    if p.bscan_flip:
        out = bscan_flip(out, p.ascans_per_bscan)
    if p.sinusoidal:
        out = sinusoidal_correction(out, p.ascans_per_bscan, sinusoidal_curve(p.ascans_per_bscan))
    if p.bg_removal:
        if state.bg_requested:
            state.bg_profile = record_background(out)
            state.bg_requested = False
        if state.bg_profile is not None and len(state.bg_profile):
            out = subtract_background(out, state.bg_profile, p.bg_weight, p.bg_offset)
    return out


# ---------------------------------------------------------------------------
# Steps of the CUDA backend only (the CPU backend ignores these flags)
# ---------------------------------------------------------------------------

def bscan_flip(out: np.ndarray, ascans_per_bscan: int) -> np.ndarray:
    """cuda_bscanFlip (cuda_kernels.cu:553-571) as launched by
    cuda_backend.cu:689-694: the A-scan order of B-scans 0, 2, 4, ... is
    reversed.  The launch covers the first half of the frame's samples, so
    with an odd B-scan count the last (even-indexed) B-scan is left as is."""
    out = np.array(out, copy=True)
    a = ascans_per_bscan
    b_count = out.shape[0] // a
    for b in range(0, b_count, 2):
        if b + 1 < b_count:
            out[b * a:(b + 1) * a] = out[b * a:(b + 1) * a][::-1]
    return out


def sinusoidal_curve(ascans_per_bscan: int) -> np.ndarray:
    """fillSinusoidalScanCorrectionCurve (cuda_kernels.cu:433-437):
    (A / pi) * acos(1 - 2k/A), acos in float, the product in double."""
    a = ascans_per_bscan
    k = np.arange(a, dtype=np.float64)
    arg = (1.0 - (2.0 * k.astype(np.float32).astype(np.float64)) / float(np.float32(a))).astype(np.float32)
    return ((float(np.float32(a)) / np.pi) * np.arccos(arg).astype(np.float64)).astype(np.float32)


def sinusoidal_correction(out: np.ndarray, ascans_per_bscan: int, curve: np.ndarray) -> np.ndarray:
    """sinusoidalScanCorrection (cuda_kernels.cu:408-431): A-scan k of B-scan
    l becomes the linear interpolation at scan position curve[k] between
    A-scans floor(curve[k]) and floor(curve[k]) + 1 (flat buffer indexing);
    the buffer's last A-scan is not rewritten."""
    flat = np.asarray(out, dtype=np.float64)
    rows, width = flat.shape
    a = ascans_per_bscan
    res = flat.copy()
    for i in range(rows - 1):
        k, l = i % a, i // a
        x = float(curve[k])
        xi = int(x)
        r0 = min(xi + l * a, rows - 1)
        r1 = min(r0 + 1, rows - 1)
        res[i] = flat[r0] + (flat[r1] - flat[r0]) * (x - xi)
    return res


# ---------------------------------------------------------------------------
# Host curve generation (src/core/processorconfiguration.cpp:438-589,
# src/core/processor.cpp:77-99).  These emulate the reference's float32
# arithmetic step by step, so they are compared bit-exactly with curves the
# reference's own code writes.
# ---------------------------------------------------------------------------

f32 = np.float32


def generate_resampling_curve(n: int, c) -> np.ndarray:
    """processorconfiguration.cpp:438-459 (float Horner, clamp [7, N-9])."""
    size_m1 = f32(n) - f32(1.0)  # (size - 1.0f): size_t converted to float
    c0 = f32(c[0])
    c1 = f32(f32(c[1]) / size_m1)
    c2 = f32(f32(c[2]) / f32(size_m1 * size_m1))
    c3 = f32(f32(c[3]) / f32(f32(size_m1 * size_m1) * size_m1))
    x = np.arange(n, dtype=np.float32)
    val = c0 + x * (c1 + x * (c2 + x * c3))
    val = val.astype(np.float32)
    lo, hi = f32(7.0), f32(n - 9)
    return np.where(val < lo, lo, np.where(val > hi, hi, val)).astype(np.float32)


def generate_dispersion_phase(n: int, d) -> np.ndarray:
    """processorconfiguration.cpp:461-482.  ``i`` is size_t, promoted to float
    in each product; the ``factor`` field is not applied (as the reference)."""
    denom = f32(n - 1)
    k0 = f32(d[0])
    k1 = f32(f32(d[1]) / denom)
    k2 = f32(f32(d[2]) / f32(denom * denom))
    k3 = f32(f32(d[3]) / f32(f32(denom * denom) * denom))
    i = np.arange(n, dtype=np.float32)
    return (k0 + i * (k1 + i * (k2 + i * k3))).astype(np.float32)


def phase_to_phasor(phase: np.ndarray) -> np.ndarray:
    """Interleaved (cos, sin) in float (processor.cpp:91-97)."""
    ph = np.asarray(phase, dtype=np.float32)
    out = np.empty(2 * ph.shape[0], dtype=np.float32)
    out[0::2] = np.cos(ph)
    out[1::2] = np.sin(ph)
    return out


def generate_window(n: int, wtype: int, center: float, fill: float) -> np.ndarray:
    """processorconfiguration.cpp:484-589."""
    center = f32(center)
    fill = f32(fill)
    if center > 1.0:
        center = f32(1.0)
    if center < 0.0:
        center = f32(0.0)
    width = int(np.uint32(f32(fill * f32(n))))
    cpos = int(np.uint32(f32(center * f32(n))))
    min_pos = int(np.int32(np.uint32((cpos - width // 2) & 0xFFFFFFFF)))
    out = np.zeros(n, dtype=np.float32)
    i = np.arange(n)
    if wtype == GAUSS:
        xi = (i - cpos).astype(np.float32)
        xn = (xi / (f32(n) - f32(1.0))) / fill
        # expf is correctly rounded in glibc: evaluate in double, round once.
        arg = (f32(-10.0) * (xn * xn)).astype(np.float32)
        return np.exp(arg.astype(np.float64)).astype(np.float32)
    xi = (i - min_pos).astype(np.float32)
    xn = (xi / (f32(width) - f32(1.0))).astype(np.float32)
    inside = ~((xn > f32(0.999)) | (xn < f32(0.0001)))
    xd = xn.astype(np.float64)
    if wtype == HANN:
        val = (0.5 * (1.0 - np.cos(2.0 * math.pi * xd))).astype(np.float32)
    elif wtype == SINE:
        val = np.sin(math.pi * xd).astype(np.float32)
    elif wtype == LANCZOS_WIN:
        arg = (f32(2.0) * xn - f32(1.0)).astype(np.float32)
        ad = arg.astype(np.float64)
        with np.errstate(divide="ignore", invalid="ignore"):
            val = np.where(arg == 0, 1.0, np.sin(math.pi * ad) / (math.pi * ad)).astype(np.float32)
    elif wtype == RECTANGULAR:
        val = np.ones(n, dtype=np.float32)
    elif wtype == FLAT_TOP:
        a0, a1, a2, a3, a4 = (f32(0.215578948), f32(0.416631580), f32(0.277263158),
                              f32(0.083578947), f32(0.006947368))
        val = (a0 - a1 * np.cos(2.0 * math.pi * xd).astype(np.float32)
               + a2 * np.cos(4.0 * math.pi * xd).astype(np.float32)
               - a3 * np.cos(6.0 * math.pi * xd).astype(np.float32)
               + a4 * np.cos(8.0 * math.pi * xd).astype(np.float32)).astype(np.float32)
    else:
        raise ValueError("unknown window type")
    out[inside] = val[inside]
    return out


def adjust_custom_curve(curve, target: int) -> np.ndarray:
    """Zero-pad or truncate a custom curve (processorconfiguration.cpp:268-326)."""
    c = np.asarray(curve, dtype=np.float32)
    if c.shape[0] >= target:
        return c[:target].copy()
    return np.concatenate([c, np.zeros(target - c.shape[0], dtype=np.float32)])


# ---------------------------------------------------------------------------
# Accuracy metrics
# ---------------------------------------------------------------------------

def peak_normalized_lin_error(test_lin: np.ndarray, ref_lin: np.ndarray) -> float:
    """max |test - ref| / per-A-scan peak of ref, over all bins."""
    ref = ref_lin.reshape(-1, ref_lin.shape[-1])
    tst = test_lin.reshape(ref.shape)
    peak = np.max(np.abs(ref), axis=-1, keepdims=True)
    peak = np.where(peak > 0, peak, 1.0)
    return float(np.max(np.abs(tst - ref) / peak))


def db_error_near_peak(test_db: np.ndarray, ref_db: np.ndarray, within_db: float = 60.0) -> float:
    """max |test - ref| (dB) over bins within ``within_db`` of the A-scan peak."""
    ref = ref_db.reshape(-1, ref_db.shape[-1])
    tst = test_db.reshape(ref.shape)
    peak = np.max(ref, axis=-1, keepdims=True)
    mask = ref >= peak - within_db
    return float(np.max(np.abs(tst - ref)[mask]))
