"""Parity of the fused HIP chain (through the C ABI) with the FP64 oracle and
the float32 restatement of the reference CPU backend.

Each case prints its error figures, so the run log documents the accuracy
the parity bar is checked at.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import chain64 as O
from octproengine_amd import synthetic as S
from tests.helpers import (DB_TOL, LIN_TOL, GpuChain, cpu32, db_error, db_lin_error, db_params,
                           lin_error, lin_params, standard_tables)

pytestmark = pytest.mark.gpu

METHODS = {"none": None, "linear": O.LINEAR, "cubic": O.CUBIC, "lanczos": O.LANCZOS}


def _params(n, a, method, disp, win=True, **kw):
    p = O.ChainParams(signal_length=n, ascans_per_bscan=a)
    p.resampling = METHODS[method] is not None
    p.interpolation = METHODS[method] or 0
    p.windowing = win
    p.dispersion = disp
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("n", [256, 512, 1024, 2048, 4096])
@pytest.mark.parametrize("method", ["none", "linear", "cubic", "lanczos"])
@pytest.mark.parametrize("disp", [False, True])
def test_log_chain_parity(n, method, disp):
    a = 67  # not a multiple of the A-scans per workgroup
    p = db_params(_params(n, a, method, disp))
    raw = S.fringes(n, a, seed=n + 7 * len(method) + int(disp))
    t = standard_tables(n)
    g = GpuChain(p, t.curve, t.window, t.phasor)
    gpu = g.frame(raw)
    g.close()
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    c32 = cpu32(p, raw, t)
    e_pk = db_error(gpu, ref)
    e_pk_cpu = db_error(c32, ref)
    e_lin, e_lin_cpu = db_lin_error(gpu, ref), db_lin_error(c32, ref)
    print(f"N={n} {method} disp={disp}: dB err within 60 dB of peak gpu={e_pk:.2e} cpu32={e_pk_cpu:.2e}; "
          f"all bins, peak-normalised linear gpu={e_lin:.2e} cpu32={e_lin_cpu:.2e}")
    assert e_pk <= DB_TOL
    assert e_lin <= LIN_TOL


@pytest.mark.parametrize("n", [512, 2048])
@pytest.mark.parametrize("method", ["linear", "cubic", "lanczos"])
def test_linear_magnitude_parity(n, method):
    a = 40
    p = lin_params(_params(n, a, method, True))
    raw = S.fringes(n, a, seed=3)
    t = standard_tables(n)
    g = GpuChain(p, t.curve, t.window, t.phasor)
    gpu = g.frame(raw)
    g.close()
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    c32 = cpu32(p, raw, t)
    e_gpu, e_cpu = lin_error(gpu, ref), lin_error(c32, ref)
    print(f"N={n} {method}: peak-normalised linear error gpu={e_gpu:.2e} cpu32={e_cpu:.2e}")
    assert e_gpu <= LIN_TOL


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.uint32, np.float32])
def test_input_types(dtype):
    n, a = 1024, 33
    p = db_params(_params(n, a, "cubic", True))
    raw16 = S.fringes(n, a, seed=11)
    if dtype == np.uint8:
        raw = (raw16 >> 8).astype(np.uint8)
    elif dtype == np.uint32:
        raw = raw16.astype(np.uint32) * 65537
    elif dtype == np.float32:
        raw = (raw16 / 65535.0).astype(np.float32)
    else:
        raw = raw16
    t = standard_tables(n)
    g = GpuChain(p, t.curve, t.window, t.phasor, dtype=dtype)
    gpu = g.frame(raw)
    g.close()
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    e = db_error(gpu, ref)
    print(f"{np.dtype(dtype).name}: dB err near peak {e:.2e}")
    assert e <= DB_TOL


def _typed(raw16, dtype):
    if dtype == np.uint8:
        return (raw16 >> 8).astype(np.uint8)
    if dtype == np.uint32:
        return raw16.astype(np.uint32) * 65537
    if dtype == np.float32:
        return (raw16 / 65535.0).astype(np.float32)
    return raw16


# windows: 0 (cumulative mean), 1, odd, typical, wide, wider than half the
# A-scan (every window clipped) and wider than the A-scan
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.uint32, np.float32])
@pytest.mark.parametrize("window", [0, 1, 7, 64, 200, 1500, 5000])
def test_dc_removal(dtype, window):
    n, a = 2048, 19
    p = db_params(_params(n, a, "cubic", True, dc_removal=True, dc_window=window))
    raw = _typed(S.fringes(n, a, seed=5), dtype)
    t = standard_tables(n)
    g = GpuChain(p, t.curve, t.window, t.phasor, dtype=dtype)
    gpu = g.frame(raw)
    g.close()
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    c32 = cpu32(p, raw, t)
    e, ec = db_error(gpu, ref), db_error(c32, ref)
    print(f"DC W={window} {np.dtype(dtype).name}: dB err near peak gpu={e:.2e} cpu32={ec:.2e}")
    assert e <= DB_TOL


# The banded DC path: kernels with up to 2 resampling taps (none, linear)
# and cubic on the swap schedule at N = 1024 and 2048 (at 4096 cubic takes
# the general path), 8/16-bit samples, N >= 1024 and windows of 1 .. 64
# samples.  Every window residue mod 8
# (each a different placement of the padded prefix blocks), the widest band
# windows, and 0 and 65 on the general path; the clipped windows at both
# ends of every A-scan are part of each case.  Peak-normalised linear error
# and dB error near the peak against the FP64 oracle.
@pytest.mark.parametrize("method", ["none", "linear", "cubic"])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("window", [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 31, 56, 63, 64, 65])
@pytest.mark.parametrize("n", [1024, 2048, 4096])
def test_dc_removal_banded(n, window, dtype, method):
    a = 23
    raw = _typed(S.fringes(n, a, seed=7 * n + window), dtype)
    t = standard_tables(n)
    for make, err, tol in ((lin_params, lin_error, LIN_TOL), (db_params, db_error, DB_TOL)):
        p = make(_params(n, a, method, True, dc_removal=True, dc_window=window))
        g = GpuChain(p, t.curve, t.window, t.phasor, dtype=dtype)
        gpu = g.frame(raw)
        g.close()
        ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
        e = err(gpu, ref)
        print(f"banded DC N={n} {method} W={window} {np.dtype(dtype).name} {make.__name__}: {e:.2e}")
        assert e <= tol


@pytest.mark.parametrize("window", [9, 64])
def test_dc_removal_banded_bitshift(window):
    """The banded path on the shifted 16-bit samples (x >> 4)."""
    n, a = 2048, 17
    raw = S.fringes(n, a, seed=3 + window)
    t = standard_tables(n)
    p = lin_params(_params(n, a, "linear", True, dc_removal=True, dc_window=window, bitshift=True))
    g = GpuChain(p, t.curve, t.window, t.phasor)
    gpu = g.frame(raw)
    g.close()
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    e = lin_error(gpu, ref)
    print(f"banded DC, bitshift, W={window}: peak-normalised linear error {e:.2e}")
    assert e <= LIN_TOL


@pytest.mark.parametrize("n", [256, 512, 1024, 4096])
@pytest.mark.parametrize("window", [0, 3, 64, 300])
def test_dc_removal_lengths_linear(n, window):
    """Every length (one and several A-scans per workgroup, one and several
    wavefronts per A-scan) with linear output: peak-normalised error."""
    a = 37
    p = lin_params(_params(n, a, "cubic", True, dc_removal=True, dc_window=window))
    raw = S.fringes(n, a, seed=n + window)
    t = standard_tables(n)
    g = GpuChain(p, t.curve, t.window, t.phasor)
    gpu = g.frame(raw)
    g.close()
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    e = lin_error(gpu, ref)
    print(f"DC N={n} W={window}: peak-normalised linear error {e:.2e}")
    assert e <= LIN_TOL


def test_grayscale_map():
    n, a = 1024, 20
    raw = S.fringes(n, a, seed=9)
    t = standard_tables(n)
    for gmin, gmax, mult, add in [(30.0, 100.0, 1.0, 0.0), (0.0, 80.0, 2.5, 0.25), (10.0, 10.000001, 1.0, 0.0)]:
        p = _params(n, a, "cubic", True, grayscale_min=gmin, grayscale_max=gmax,
                    multiplicator=mult, addend=add)
        g = GpuChain(p, t.curve, t.window, t.phasor)
        gpu = g.frame(raw)
        g.close()
        ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
        rng = np.float32(gmax) - np.float32(gmin)
        rng = 1.0 if rng < 1e-6 else float(rng)
        to_db = lambda y: (np.asarray(y, np.float64) / mult - add) * rng + gmin
        e = db_error(to_db(gpu), to_db(ref))
        print(f"map ({gmin},{gmax},{mult},{add}): dB err {e:.2e}")
        assert e <= DB_TOL


def test_auto_min_max():
    """grayscale min == max: per-A-scan min/max of the dB values
    (cpu_kernels.tpp:326-335).  Checked exactly against the GPU's own dB
    output (the extremes sit at noise-floor bins, whose float32 error is
    not the subject here); the dB output itself is checked against FP64."""
    n, a = 2048, 24
    raw = S.fringes(n, a, seed=9)
    t = standard_tables(n)
    for gm, mult, add in [(50.0, 1.0, 0.0), (0.0, 3.0, 0.5)]:
        p = _params(n, a, "cubic", True, grayscale_min=gm, grayscale_max=gm, multiplicator=mult,
                    addend=add)
        g = GpuChain(p, t.curve, t.window, t.phasor)
        auto = g.frame(raw).astype(np.float64)
        g.close()
        pdb = db_params(_params(n, a, "cubic", True))
        g = GpuChain(pdb, t.curve, t.window, t.phasor)
        db = g.frame(raw).astype(np.float64)
        g.close()
        mn = db.min(axis=1, keepdims=True)
        mx = db.max(axis=1, keepdims=True)
        want = mult * ((db - mn) / (mx - mn) + add)
        e = float(np.max(np.abs(auto - want)))
        print(f"auto min/max (mult {mult}, add {add}): max abs deviation from the rule {e:.2e}")
        assert e <= 1e-5
        assert db_error(db, O.process_frame(raw, pdb, t.curve, t.window, t.phasor)) <= DB_TOL


def test_fpn_one_shot_and_continuous():
    """FPN accumulation across frames, one-shot and continuous
    (cpu_backend.cpp:305-357).  Rolling-average DC removal is on: without it
    the DC bins make the reference's min-variance choice depend on float32
    rounding of the segment means (see test_min_variance_mean_exact)."""
    n, per_frame = 1024, 48
    t = standard_tables(n)
    for continuous in (False, True):
        p = db_params(_params(n, per_frame, "cubic", True, fpn_removal=True, fpn_bscan_count=2,
                              fpn_continuous=continuous, dc_removal=True, dc_window=32))
        g = GpuChain(p, t.curve, t.window, t.phasor)
        state = O.ChainState()
        if not continuous:
            g.ctx.request_fpn()
            state.fpn_requested = True
        for f in range(4):
            raw = S.fringes(n, per_frame, seed=100 + f)
            gpu = g.frame(raw)
            ref = O.process_frame(raw, p, t.curve, t.window, t.phasor, state)
            e = db_error(gpu, ref)
            prof = g.ctx.get_fpn_profile()
            print(f"FPN continuous={continuous} frame {f}: dB err {e:.2e}, profile pairs {prof.size // 2}")
            if state.fpn_profile is not None:
                ref_prof = state.fpn_profile.astype(np.float64)
                scale = np.max(np.abs(ref_prof)) or 1.0
                assert prof.size == ref_prof.size
                ep = np.max(np.abs(prof - ref_prof)) / scale
                print(f"   profile: max |diff| / max |profile| = {ep:.2e}")
                assert ep <= 1e-4
            else:
                assert prof.size == 0
            assert e <= DB_TOL
        g.close()


def test_min_variance_mean_exact():
    """getMinimumVarianceMean on identical float32 input is bit-exact with the
    oracle (same summation order in double, same float roundings), including
    DC-dominated columns where the selection hinges on those roundings."""
    from octproengine_amd import _capi
    rng = np.random.default_rng(0)
    h, w = 96, 300
    rows = (rng.normal(size=(h, w)) + 1j * rng.normal(size=(h, w))) * 1e3
    rows[:, :5] += 7e6  # DC-like columns
    rows = (rows.real.astype(np.float32) + 1j * rows.imag.astype(np.float32))
    with _capi.Context(_capi.make_params(signal_length=256, ascans_per_bscan=8)) as ctx:
        got = ctx.min_variance_mean(rows, 8)
    want = O.min_variance_mean(rows, 8)
    assert np.array_equal(got, want)


def test_set_fpn_profile_subtracts():
    n, a = 512, 16
    p = db_params(_params(n, a, "linear", False, fpn_removal=True))
    t = standard_tables(n)
    rng = np.random.default_rng(1)
    prof = (rng.normal(size=n) * 1e4).astype(np.float32)  # n/2 pairs
    g = GpuChain(p, t.curve, t.window, t.phasor)
    g.ctx.set_fpn_profile(prof)
    raw = S.fringes(n, a, seed=4)
    gpu = g.frame(raw)
    got = g.ctx.get_fpn_profile()
    g.close()
    state = O.ChainState(fpn_profile=prof)
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor, state)
    assert np.array_equal(got, prof)
    assert db_error(gpu, ref) <= DB_TOL


def test_background_record_and_subtract():
    """Post-process background (cpu_backend.cpp:283-301): the recorded
    profile is the column mean of the frame's unsubtracted output, and the
    output is out - (weight * bg + offset).  Checked against the GPU's own
    unsubtracted output, which is itself checked against FP64."""
    n, a = 1024, 32
    w_, off = 0.75, 0.1
    t = standard_tables(n)
    raws = [S.fringes(n, a, seed=200 + f) for f in range(2)]
    p0 = db_params(_params(n, a, "cubic", True))
    g = GpuChain(p0, t.curve, t.window, t.phasor)
    plain = [g.frame(r).astype(np.float64) for r in raws]
    g.close()
    for r, out in zip(raws, plain):
        assert db_error(out, O.process_frame(r, p0, t.curve, t.window, t.phasor)) <= DB_TOL
    p = db_params(_params(n, a, "cubic", True, bg_removal=True, bg_weight=w_, bg_offset=off))
    g = GpuChain(p, t.curve, t.window, t.phasor)
    g.ctx.request_bg_recording()
    got = [g.frame(r).astype(np.float64) for r in raws]
    bg = g.ctx.get_bg_profile().astype(np.float64)
    g.close()
    want_bg = plain[0].mean(axis=0)
    print(f"background: profile max |diff| {np.max(np.abs(bg - want_bg)):.2e}")
    assert bg.size == n // 2
    assert np.max(np.abs(bg - want_bg)) <= 1e-4
    for f in range(2):
        want = plain[f] - (np.float32(w_) * bg.astype(np.float32) + np.float32(off))
        e = float(np.max(np.abs(got[f] - want)))
        print(f"background frame {f}: max abs deviation {e:.2e}")
        assert e <= 1e-4


def test_hot_swap_curves_between_frames():
    n, a = 1024, 24
    p = db_params(_params(n, a, "cubic", True))
    t1 = standard_tables(n)
    t2 = standard_tables(n, window_type=O.SINE, center=0.45, fill=0.8, d2=2e-6, d3=1e-9,
                         custom_curve=False)
    raw = S.fringes(n, a, seed=8)
    g = GpuChain(p, t1.curve, t1.window, t1.phasor)
    out1 = g.frame(raw)
    g.ctx.set_resampling_curve(t2.curve)
    g.ctx.set_window_curve(t2.window)
    g.ctx.set_dispersion_curve(t2.phasor)
    out2 = g.frame(raw)
    g.close()
    assert db_error(out1, O.process_frame(raw, p, t1.curve, t1.window, t1.phasor)) <= DB_TOL
    assert db_error(out2, O.process_frame(raw, p, t2.curve, t2.window, t2.phasor)) <= DB_TOL


def test_many_ascans_grid_stride():
    """More A-scans than resident workgroups: every A-scan of a long frame
    matches a per-A-scan oracle on a sample."""
    n, a = 2048, 20000
    p = db_params(_params(n, a, "cubic", True))
    raw = S.fringes(n, a, seed=21)
    t = standard_tables(n)
    g = GpuChain(p, t.curve, t.window, t.phasor)
    gpu = g.frame(raw)
    g.close()
    idx = np.r_[0:8, a // 2:a // 2 + 8, a - 8:a]
    ref = O.process_frame(raw[idx], p, t.curve, t.window, t.phasor)
    assert db_error(gpu[idx], ref) <= DB_TOL
    # every A-scan's peak is at one of its reflector depths (size-independent)
    assert np.all(np.isfinite(gpu))
