"""Pins the FP64 oracle (oracle/chain64.py) and the float32 C restatement
(oracle/cpu_chain.c) before anything is checked against them.

Every check is independent of the oracle's own formulas: closed forms,
brute force on tiny inputs, invariants, and golden curves written by the
reference's own code.  The `planted` tests verify that a plausible mistake
(sign, index, dropped term) in an oracle function makes its check fail.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest

from oracle import chain64 as O
from tests.helpers import CpuParams, Tables, _fp, cpu32, cpu_lib, db_error, db_params

RNG = np.random.default_rng(1234)


# ------------------------------------------------------------------ DFT

def brute_idft(s):
    n = len(s)
    j = np.arange(n)
    return np.array([np.sum(s * np.exp(2j * np.pi * j * k / n)) for k in range(n)])


def check_ifft(fn):
    for n in (8, 16, 64, 100):
        s = RNG.normal(size=n) + 1j * RNG.normal(size=n)
        X = fn(s)
        assert np.allclose(X, brute_idft(s), atol=1e-9 * n)
        # Parseval for the unnormalised transform: sum |X|^2 = N sum |s|^2
        assert math.isclose(np.sum(np.abs(X) ** 2), n * np.sum(np.abs(s) ** 2), rel_tol=1e-12)
    # a cosine at bin z peaks at k = z and N - z
    n, z = 256, 37
    X = fn(np.cos(2 * np.pi * z * np.arange(n) / n))
    assert set(np.argsort(np.abs(X))[-2:]) == {z, n - z}


def test_ifft_definition():
    check_ifft(O.ifft_unnormalized)


def test_planted_ifft_sign_error_is_caught():
    with pytest.raises(AssertionError):
        check_ifft(lambda s: np.fft.fft(s))  # forward sign instead of backward


# ------------------------------------------------------------------ interpolation

def check_linear(fn):
    n = 64
    x = 3.0 * np.arange(n) - 7.0  # affine data
    r = np.linspace(0.0, n - 2.0, 200).astype(np.float32)
    assert np.allclose(fn(x, r), 3.0 * r.astype(np.float64) - 7.0, atol=1e-12)
    # equals numpy.interp on arbitrary data inside the range
    y = RNG.normal(size=n)
    assert np.allclose(fn(y, r), np.interp(r.astype(np.float64), np.arange(n), y), atol=1e-12)


def test_linear_interpolation():
    check_linear(O.klin_linear)
    # clamping at the upper end reproduces the last sample
    x = RNG.normal(size=32)
    assert O.klin_linear(x, np.array([31.0, 40.0], np.float32))[0] == x[31]


def test_planted_linear_index_error_is_caught():
    def bad(x, curve):
        r = np.asarray(curve, np.float64)
        n1 = np.minimum(np.trunc(r).astype(int), len(x) - 1)
        n2 = np.minimum(n1 + 1, len(x) - 1)
        return x[n2] + (x[n2] - x[n1]) * (r - n1)  # wrong base sample
    with pytest.raises(AssertionError):
        check_linear(bad)


def check_cubic(fn):
    n = 64
    j = np.arange(n, dtype=np.float64)
    q = 0.5 * j * j - 3.0 * j + 2.0  # Catmull-Rom reproduces quadratics exactly
    r = np.linspace(1.0, n - 3.0, 301).astype(np.float32)
    rr = r.astype(np.float64)
    assert np.allclose(fn(q, r), 0.5 * rr * rr - 3.0 * rr + 2.0, atol=1e-9)
    y = RNG.normal(size=n)
    ints = np.arange(1, n - 3, dtype=np.float32)
    assert np.allclose(fn(y, ints), y[1:n - 3], atol=1e-15)  # interpolating


def test_cubic_interpolation():
    check_cubic(O.klin_cubic)
    # the reference's n0 = |n1 - 1| at n1 = 0 (cpu_kernels.tpp:127)
    y = RNG.normal(size=16)
    t = 0.25
    got = O.klin_cubic(y, np.array([t], np.float32))[0]
    assert math.isclose(got, O.cubic_hermite(y[1], y[0], y[1], y[2], np.float64(np.float32(t))), rel_tol=1e-14)


def test_planted_cubic_dropped_term_is_caught():
    def bad(x, curve):
        r = np.asarray(curve, np.float32).astype(np.float64)
        n1 = np.trunc(r).astype(int)
        y0, y1, y2, y3 = x[np.abs(n1 - 1)], x[n1], x[n1 + 1], x[np.minimum(n1 + 2, len(x) - 1)]
        t = r - n1
        a = -y0 + 3.0 * (y1 - y2) + y3
        c = -y0 + y2  # the b t^2 term dropped
        return 0.5 * t * (a * t * t + c) + y1
    with pytest.raises(AssertionError):
        check_cubic(bad)


def check_lanczos(fn, a=8):
    # sinc(d) sinc(d/a) in closed form, brute force
    d = np.array([0.0, 1e-6, 0.3, 1.0, 2.5, 3.9, 4.0, 4.5, 7.9, 8.0, 9.0, -3.2])
    want = []
    for v in d:
        if abs(v) < 1e-5:
            want.append(1.0)
        elif abs(v) >= a:
            want.append(0.0)
        else:
            want.append(math.sin(math.pi * v) / (math.pi * v) * math.sin(math.pi * v / a) / (math.pi * v / a))
    assert np.allclose(fn(d), want, atol=1e-15)


def test_lanczos():
    check_lanczos(O.lanczos_kernel)
    # interpolating: integer positions return the samples
    y = RNG.normal(size=64)
    ints = np.arange(8, 56, dtype=np.float32)
    assert np.allclose(O.klin_lanczos(y, ints), y[8:56], atol=1e-12)
    # a constant signal stays (nearly) constant: the kernel sums to ~1
    got = O.klin_lanczos(np.ones(64), np.linspace(10, 50, 77).astype(np.float32))
    assert np.all(np.abs(got - 1.0) < 0.05)


def test_planted_lanczos_wrong_scale_is_caught():
    with pytest.raises(AssertionError):
        check_lanczos(lambda d: O.lanczos_kernel(np.asarray(d) * 1.01))


def test_lanczos4():
    """a = 4 (the extra 8-tap mode): closed form, interpolating at integers,
    a constant stays nearly constant, and the taps are n0-3 .. n0+4: a
    sample 4 or more positions away never enters."""
    check_lanczos(lambda d: O.lanczos_kernel(d, 4), a=4)
    y = RNG.normal(size=64)
    ints = np.arange(8, 56, dtype=np.float32)
    assert np.allclose(O.klin_lanczos(y, ints, a=4), y[8:56], atol=1e-12)
    got = O.klin_lanczos(np.ones(64), np.linspace(10, 50, 77).astype(np.float32), a=4)
    assert np.all(np.abs(got - 1.0) < 0.05)
    for r in (20.0, 20.25, 20.75):
        for far in (16, 25):
            x = np.zeros(64)
            x[far] = 1.0
            assert O.klin_lanczos(x, np.array([r], np.float32), a=4)[0] == 0.0
        # brute force over the 8 taps n0-3 .. n0+4 of the closed form
        x = RNG.normal(size=64)
        n0 = int(r)
        want = sum(x[n0 + i] * (1.0 if abs(r - n0 - i) < 1e-5 else
                                math.sin(math.pi * (r - n0 - i)) / (math.pi * (r - n0 - i))
                                * math.sin(math.pi * (r - n0 - i) / 4) / (math.pi * (r - n0 - i) / 4))
                   for i in range(-3, 5))
        assert math.isclose(O.klin_lanczos(x, np.array([r], np.float32), a=4)[0], want, rel_tol=1e-12,
                            abs_tol=1e-12)


def test_planted_lanczos4_with_the_a8_window_is_caught():
    with pytest.raises(AssertionError):
        check_lanczos(lambda d: O.lanczos_kernel(d, 8), a=4)


# ------------------------------------------------------------------ DC removal, FPN, scaling

def brute_dc(x, w):
    n = len(x)
    out = np.empty(n)
    for i in range(n):
        lo = max(i - (w - 1), 0) if w >= 1 else 0
        hi = min(i + w, n - 1) if w >= 1 else i
        out[i] = x[i] - np.mean(x[lo:hi + 1])
    return out


def test_rolling_average_dc():
    x = RNG.normal(size=50) + 10
    for w in (0, 1, 3, 7, 49, 80):
        assert np.allclose(O.rolling_average_dc(x, w), brute_dc(x, w), atol=1e-12)


def test_planted_dc_window_off_by_one_is_caught():
    x = RNG.normal(size=50) + 10
    bad = brute_dc(x, 8)  # the window of a different size
    assert not np.allclose(O.rolling_average_dc(x, 7), bad, atol=1e-9)


def brute_mvm(rows, segments):
    h, w = rows.shape
    sw = h // segments
    out = np.zeros(2 * w, np.float32)
    for c in range(w):
        best = None
        for s in range(segments):
            seg = rows[s * sw:(s + 1) * sw, c]
            var = np.mean(np.abs(seg) ** 2) - np.abs(np.mean(seg)) ** 2
            if best is None or var < best[0]:
                best = (var, np.mean(seg))
        out[2 * c], out[2 * c + 1] = best[1].real, best[1].imag
    return out


def test_min_variance_mean_brute_force():
    # well-separated segment variances: the float32/float64 mixture of the
    # reference picks the same segment as exact arithmetic
    h, w, segs = 40, 6, 8
    scale = np.repeat(np.array([5.0, 1.0, 3.0, 0.2, 4.0, 2.0, 6.0, 7.0]), h // segs)[:, None]
    rows = (RNG.normal(size=(h, w)) + 1j * RNG.normal(size=(h, w))) * scale + (2 + 1j)
    rows = rows.real.astype(np.float32) + 1j * rows.imag.astype(np.float32)
    got = O.min_variance_mean(rows, segs)
    assert np.allclose(got, brute_mvm(rows, segs), rtol=1e-6, atol=1e-6)
    # and the C restatement agrees bit for bit with the oracle
    lib = cpu_lib()
    inter = np.empty((h, w, 2), np.float32)
    inter[..., 0], inter[..., 1] = rows.real, rows.imag
    out = np.zeros(2 * w, np.float32)
    lib.ope_cpu_min_variance_mean(inter.ctypes.data, w, h, segs, out.ctypes.data)
    assert np.array_equal(out, got)


def test_log_and_linear_scaling_closed_forms():
    X = np.array([[3 + 4j, 0.0, 1.0, 2.0, 0, 0, 0, 0]])
    half = 4
    log = O.log_scale_and_truncate(X, 2.0, 10.0, 60.0, 0.5, False)
    with np.errstate(divide="ignore"):
        v = 10 * np.log10(np.array([25.0, 0.0, 1.0, 4.0]) / half)
    assert np.allclose(log[0], 2.0 * ((v - 10.0) / 50.0 + 0.5))
    assert log[0, 1] == -np.inf  # zero spectrum gives -inf, as the reference
    lin = O.linear_scale_and_truncate(X, 1.5, 0.1, 0.6, 0.0)
    assert np.allclose(lin[0], 1.5 * ((np.array([5.0, 0, 1, 2]) / half - 0.1) / 0.5))
    # guard: range < 1e-6 -> 1; auto min/max per A-scan when min == max
    Y = np.array([[1.0, 2.0, 4.0, 8.0, 0, 0, 0, 0]], complex)
    auto = O.log_scale_and_truncate(Y, 1.0, 5.0, 5.0, 0.0, True)
    assert np.isclose(auto.min(), 0.0) and np.isclose(auto.max(), 1.0)


# ------------------------------------------------------------------ curves vs the reference's own code

def _fixtures():
    from pathlib import Path
    return np.load(Path(__file__).parent / "golden" / "reference_curves.npz")


def test_oracle_curves_match_reference_generator():
    F = _fixtures()
    for n in F["lengths"]:
        for i, c in enumerate(F["params_resample"]):
            assert np.array_equal(O.generate_resampling_curve(int(n), c), F[f"resample_n{n}_{i}"])
        for i, d in enumerate(F["params_dispersion"]):
            assert np.array_equal(O.generate_dispersion_phase(int(n), d), F[f"dispersion_n{n}_{i}"])
        for i, (t, c, f) in enumerate(F["params_window"]):
            got = O.generate_window(int(n), int(t), c, f)
            want = F[f"window_n{n}_{i}"]
            if int(t) == O.GAUSS:  # expf rounding: within 1 ulp
                ulp = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32))
                assert ulp.max() <= 1
            else:
                assert np.array_equal(got, want)


# ------------------------------------------------------------------ C restatement vs FP64

@pytest.mark.parametrize("method", [O.LINEAR, O.CUBIC, O.LANCZOS, O.LANCZOS4])
def test_cpu32_restatement_matches_fp64_oracle(method):
    from octproengine_amd import synthetic as S
    n, a = 512, 12
    p = db_params(O.ChainParams(signal_length=n, ascans_per_bscan=a, resampling=True, interpolation=method,
                                windowing=True, dispersion=True))
    raw = S.fringes(n, a, seed=2)
    t = Tables(S.resampling_curve(n), O.generate_window(n, O.HANN, 0.5, 0.95),
               O.phase_to_phasor(S.dispersion_phase(n, 2e-5, -1e-8)))
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    got = cpu32(p, raw, t)
    assert db_error(got, ref) <= 1e-3


# ---------------------------------------------------------------- CUDA-only steps


def test_bscan_flip_restatement():
    a, width = 5, 3
    for b in (1, 2, 3, 4):
        x = np.arange(a * b * width, dtype=np.float64).reshape(a * b, width)
        y = O.bscan_flip(x, a)
        for k in range(b):
            blk, src = y[k * a:(k + 1) * a], x[k * a:(k + 1) * a]
            flipped = k % 2 == 0 and k + 1 < b  # the launch covers half of the samples
            assert np.array_equal(blk, src[::-1] if flipped else src), (b, k)
        assert np.array_equal(O.bscan_flip(y, a), x)  # an involution
    # planted mistake: flipping the odd B-scans instead must fail the check
    x = np.arange(8 * 2, dtype=np.float64).reshape(8, 2)
    assert not np.array_equal(O.bscan_flip(x, 4)[4:], x[4:][::-1])


def test_sinusoidal_curve_closed_form():
    for a in (4, 64, 1024):
        c = O.sinusoidal_curve(a).astype(np.float64)
        k = np.arange(a)
        assert c[0] == 0.0
        assert np.all(np.diff(c) > 0)
        assert np.allclose(c, a / np.pi * np.arccos(1 - 2 * k / a), rtol=1e-6, atol=1e-4)
        if a % 2 == 0:
            assert abs(c[a // 2] - a / 2) < 1e-4 * a


def test_sinusoidal_correction_reproduces_linear_fields():
    """Linear interpolation along the scan is exact on fields linear in the
    flat A-scan index: row i of B-scan l becomes curve[k] + l*A."""
    a, b, width = 32, 3, 4
    curve = O.sinusoidal_curve(a)
    rows = np.arange(a * b, dtype=np.float64)
    field = np.repeat((2.5 * rows + 1.0)[:, None], width, axis=1)
    out = O.sinusoidal_correction(field, a, curve)
    k, l = rows[:-1] % a, rows[:-1] // a
    want = 2.5 * (curve[k.astype(int)].astype(np.float64) + l * a) + 1.0
    assert np.allclose(out[:-1, 0], want, rtol=0, atol=1e-9)
    assert np.array_equal(out[-1], field[-1])  # the last A-scan is not rewritten


# ------------------------------------------------------------------ FPN accumulation across frames

class ReferenceFpnReplay:
    """Step-by-step replay of CpuBackend::Impl::computeFixedPatternNoiseIfRequested
    (cpu_backend.cpp:305-357): every processed A-scan is pushed onto the
    accumulation; a determination needs a request (or continuous mode) and at
    least bscanCount * ascansPerBscan accumulated A-scans, and uses the first
    that many; continuous mode then erases the newer entries, one-shot mode
    clears everything.  Written from the reference, not from the oracle."""

    def __init__(self):
        self.accumulated = []
        self.count = 0
        self.requested = False
        self.profile = None

    def frame(self, X, bscan_count, ascans_per_bscan, continuous, n):
        for row in X:
            self.accumulated.append(row)
            self.count += 1
        required = bscan_count * ascans_per_bscan
        if not (self.requested or continuous):
            return
        if self.count < required:
            return
        inputs = self.accumulated if self.count == required else self.accumulated[:required]
        mean = O.min_variance_mean(np.asarray(inputs), 8)
        self.profile = np.array([mean[i] for i in range(2 * (n // 2))], dtype=np.float32)
        if continuous:
            if self.count > required:
                del self.accumulated[required:]
                self.count = required
        else:
            self.accumulated = []
            self.count = 0
        self.requested = False


def check_fpn_update(update):
    n, a = 16, 16
    rng = np.random.default_rng(77)
    frames = [rng.normal(0, 1, (2 * a, n)) + 1j * rng.normal(0, 1, (2 * a, n)) + k for k in range(4)]
    # (bscan count, continuous, frames before which a request is made)
    for count, continuous, requests in ((2, False, {0, 3}), (3, True, set()), (2, False, {2}), (1, False, {1, 2})):
        p = O.ChainParams(signal_length=n, ascans_per_bscan=a, bscans_per_buffer=2, fpn_bscan_count=count,
                          fpn_continuous=continuous)
        state, ref = O.ChainState(), ReferenceFpnReplay()
        first = None
        for f, X in enumerate(frames):
            if f in requests:
                state.fpn_requested = True
                ref.requested = True
            update(state, X, p)
            ref.frame(X, count, a, continuous, n)
            if ref.profile is None:
                assert state.fpn_profile is None, (count, continuous, f)
            else:
                assert state.fpn_profile is not None and np.array_equal(state.fpn_profile, ref.profile), \
                    (count, continuous, f)
            if continuous and ref.profile is not None:
                # a continuous profile never leaves the earliest window
                first = ref.profile if first is None else first
                assert np.array_equal(state.fpn_profile, first)


def test_fpn_update_replays_the_reference_accumulation():
    check_fpn_update(O.fpn_update)


def _fpn_update_latest_window(state, X, p):
    """Planted: determination over the most recent `required` A-scans."""
    required = p.fpn_bscan_count * p.ascans_per_bscan
    state.fpn_rows.extend(list(X))
    state.fpn_rows = state.fpn_rows[-required:]
    state.fpn_count += X.shape[0]
    if not (state.fpn_requested or p.fpn_continuous) or state.fpn_count < required:
        return
    state.fpn_profile = O.min_variance_mean(np.asarray(state.fpn_rows), 8)[: 2 * (p.signal_length // 2)].copy()
    if not p.fpn_continuous:
        state.fpn_rows, state.fpn_count = [], 0
    state.fpn_requested = False


def _fpn_update_no_clear(state, X, p):
    """Planted: the one-shot determination keeps its accumulation."""
    keep_requested = state.fpn_requested
    p2 = O.ChainParams(**{**p.__dict__, "fpn_continuous": True}) if keep_requested else p
    O.fpn_update(state, X, p2)


def _fpn_update_bscans_only(state, X, p):
    """Planted: required = bscan count (the ascans-per-B-scan factor dropped)."""
    p2 = O.ChainParams(**{**p.__dict__, "ascans_per_bscan": 1})
    O.fpn_update(state, X, p2)


@pytest.mark.parametrize("planted", [_fpn_update_latest_window, _fpn_update_no_clear, _fpn_update_bscans_only])
def test_planted_fpn_update_mistakes_are_caught(planted):
    with pytest.raises(AssertionError):
        check_fpn_update(planted)


# ------------------------------------------------------------------ post-process background

def check_record_background(fn):
    rng = np.random.default_rng(78)
    b, a, w = 3, 8, 12
    base = rng.integers(-50, 50, w).astype(np.float64)
    # each column's deviations cancel exactly: +d and -d in paired rows
    d = rng.integers(-20, 20, (b * a // 2, w)).astype(np.float64)
    out = base[None, :] + np.concatenate([d, -d], 0)
    out = out[rng.permutation(b * a)]
    assert np.array_equal(fn(out), base)  # mean over every A-scan of every B-scan
    # the reference's float accumulation (cpu_backend.cpp:105-118) on data
    # that is not exactly representable agrees to float rounding
    f = rng.normal(3.0, 1.0, (b * a, w)).astype(np.float32)
    replay = np.zeros(w, dtype=np.float32)
    for i in range(w):
        s = np.float32(0.0)
        for r in range(b * a):
            s = np.float32(s + f[r, i])
        replay[i] = np.float32(s / np.float32(b * a))
    assert np.allclose(fn(f.astype(np.float64)), replay, rtol=1e-6, atol=0)


def test_record_background_is_the_column_mean():
    check_record_background(O.record_background)


@pytest.mark.parametrize("planted", [
    lambda out: out[:8].mean(axis=0),        # first B-scan only (what the CUDA backend records)
    lambda out: np.median(out, axis=0),      # median instead of mean
    lambda out: out.sum(axis=0) / (out.shape[0] - 1),
])
def test_planted_record_background_mistakes_are_caught(planted):
    with pytest.raises(AssertionError):
        check_record_background(planted)


# ------------------------------------------------------------------ dispersion with a mismatched phasor

def check_dispersion(fn):
    rng = np.random.default_rng(79)
    n = 32
    y = rng.normal(0, 1, (3, n))
    ang = rng.uniform(-np.pi, np.pi, n)
    ph = np.empty(2 * n, np.float32)
    ph[0::2], ph[1::2] = np.cos(ang), np.sin(ang)
    c, s = ph[0::2].astype(np.float64), ph[1::2].astype(np.float64)
    got = fn(y, ph)
    for r in range(3):
        for j in range(n):
            assert got[r, j] == complex(y[r, j] * c[j], y[r, j] * s[j])
    # any other length leaves the data real and unchanged (cpu_kernels.tpp:205-208)
    for bad in (np.concatenate([ph, ph[:2]]), ph[:-2], np.zeros(0, np.float32)):
        got = fn(y, bad)
        assert np.array_equal(got.real, y) and not np.any(got.imag)


def test_dispersion_length_rule():
    check_dispersion(O.dispersion)


def _dispersion_truncating(y, ph):
    """Planted: a longer phasor is truncated to N instead of ignored."""
    n = y.shape[-1]
    p = np.asarray(ph, np.float32).astype(np.float64)
    if p.size >= 2 * n:
        p = p[: 2 * n]
    return O.dispersion(y, p.astype(np.float32))


def test_planted_dispersion_mistake_is_caught():
    with pytest.raises(AssertionError):
        check_dispersion(_dispersion_truncating)


# ------------------------------------------------------------------ mean A-line subtraction

def check_mean_subtraction(fn):
    rng = np.random.default_rng(80)
    n = 16
    X = rng.normal(0, 1, (4, n)) + 1j * rng.normal(0, 1, (4, n))
    for pairs in (0, 3, n // 2, n // 2 + 5, n):
        m = rng.normal(0, 1, 2 * pairs).astype(np.float32)
        got = fn(X, m if pairs else np.zeros(0, np.float32))
        use = min(n // 2, pairs)
        for k in range(n):
            want = X[:, k] - (complex(m[2 * k], m[2 * k + 1]) if k < use else 0)
            assert np.array_equal(got[:, k], want), (pairs, k)


def test_mean_subtraction_touches_min_half_pairs():
    check_mean_subtraction(O.mean_aline_subtraction)


@pytest.mark.parametrize("planted", [
    # every pair of the profile, up to N
    lambda X, m: X - np.pad(m[0::2] + 1j * m[1::2], (0, max(0, X.shape[-1] - len(m) // 2)))[: X.shape[-1]]
    if len(m) else X,
    # the conjugate of the profile
    lambda X, m: O.mean_aline_subtraction(X, np.stack([m[0::2], -m[1::2]], -1).reshape(-1)) if len(m) else X,
])
def test_planted_mean_subtraction_mistakes_are_caught(planted):
    with pytest.raises(AssertionError):
        check_mean_subtraction(planted)


# ------------------------------------------------------------------ input conversion with bitshift

def check_bitshift(fn):
    v8 = np.arange(256, dtype=np.uint8)
    assert np.array_equal(fn(v8, True, True), np.array([v // 16 for v in range(256)], np.float64))
    v16 = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    assert np.array_equal(fn(v16, True, True), (np.arange(65536) // 16).astype(np.float64))
    # 32-bit integers: x / 2^32 (cuda_kernels.cu:58-59), an exact scaling
    v32 = np.array([0, 1, 15, 16, 2 ** 24 + 1, 2 ** 31, 2 ** 32 - 1, 123456789], dtype=np.uint32)
    assert [float(x) for x in fn(v32, True, True)] == [math.ldexp(int(x), -32) for x in v32]
    # IEEE floats are not shifted; without the option every width is exact
    f = np.array([0.25, 1.5, 3e4], np.float32)
    assert np.array_equal(fn(f, True, True), f.astype(np.float64))
    assert np.array_equal(fn(v16, True, False), np.arange(65536, dtype=np.float64))
    assert [float(x) for x in fn(v32, True, False)] == [float(int(x)) for x in v32]


def test_bitshift_conversion():
    check_bitshift(O.convert_input)


@pytest.mark.parametrize("planted", [
    lambda raw, ieee, shift: O.convert_input(raw, ieee, False) / 16.0 if shift else O.convert_input(raw, ieee, False),
    lambda raw, ieee, shift: (raw.astype(np.uint64) >> (4 if shift else 0)).astype(np.float64)
    if raw.dtype != np.float32 else raw.astype(np.float64),
])
def test_planted_bitshift_mistakes_are_caught(planted):
    with pytest.raises(AssertionError):
        check_bitshift(planted)
