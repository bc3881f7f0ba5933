"""Real pairs: without dispersion (and without DC removal) the two A-scans of
a lane are the real and imaginary part of one complex FFT, split after the
transform (X0[k] = (Z[k] + conj Z[N-k]) / 2, X1[k] = (Z[k] - conj Z[N-k]) / 2i).
That is the default launch there.  Each case runs the paired launch and the
one-A-scan-per-FFT launch (pinned with ope_hip_set_launch_variant) and
checks both against the FP64 oracle, and the pair against the single launch.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import chain64 as O
from octproengine_amd import synthetic as S
from tests.helpers import (DB_TOL, LIN_TOL, GpuChain, db_error, db_lin_error, db_params, lin_error, lin_params,
                           standard_tables)

pytestmark = pytest.mark.gpu

METHODS = {"none": None, "linear": O.LINEAR, "cubic": O.CUBIC, "lanczos": O.LANCZOS}


def _params(n, a, method, **kw):
    p = O.ChainParams(signal_length=n, ascans_per_bscan=a)
    p.resampling = METHODS[method] is not None
    p.interpolation = METHODS[method] or 0
    p.windowing = True
    p.dispersion = False
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _passes(n):
    return (int(np.log2(n)) + 2) // 3


def _both(p, raw, t, setup=None):
    """(paired output, one-A-scan-per-FFT output) of the same frame."""
    out = []
    for pinned in (False, True):
        g = GpuChain(p, t.curve, t.window, t.phasor, dtype=raw.dtype)
        if setup:
            setup(g)
        if pinned:
            g.ctx.set_launch_variant(1, _passes(p.signal_length) - 1, 0)
        out.append(g.frame(raw))
        g.close()
    return out


@pytest.mark.parametrize("n", [256, 512, 1024, 2048, 4096])
@pytest.mark.parametrize("method", ["none", "linear", "cubic", "lanczos"])
def test_pairs_log_output(n, method):
    a = 67  # odd: the last pair has one A-scan
    p = db_params(_params(n, a, method))
    raw = S.fringes(n, a, seed=300 + n + len(method))
    t = standard_tables(n)
    pair, single = _both(p, raw, t)
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    e_p, e_s = db_error(pair, ref), db_error(single, ref)
    el = db_lin_error(pair, ref)
    d = db_error(pair, single)
    print(f"N={n} {method}: dB err near peak pair={e_p:.2e} single={e_s:.2e}; pair vs single {d:.2e}; "
          f"pair peak-normalised linear {el:.2e}")
    assert e_p <= DB_TOL and el <= LIN_TOL
    assert e_s <= DB_TOL


@pytest.mark.parametrize("n", [512, 2048])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
def test_pairs_linear_output_and_input_types(n, dtype):
    a = 40
    p = lin_params(_params(n, a, "cubic"))
    raw16 = S.fringes(n, a, seed=7)
    raw = {np.uint8: (raw16 >> 8).astype(np.uint8), np.uint16: raw16,
           np.float32: (raw16 / 65535.0).astype(np.float32)}[dtype]
    t = standard_tables(n)
    pair, single = _both(p, raw, t)
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    e_p, e_s = lin_error(pair, ref), lin_error(single, ref)
    print(f"N={n} {np.dtype(dtype).name}: peak-normalised linear error pair={e_p:.2e} single={e_s:.2e}")
    assert e_p <= LIN_TOL and e_s <= LIN_TOL


@pytest.mark.parametrize("a", [1, 2, 3])
def test_pairs_tiny_frames(a):
    n = 1024
    p = db_params(_params(n, a, "cubic"))
    raw = S.fringes(n, a, seed=11)
    t = standard_tables(n)
    pair, single = _both(p, raw, t)
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    assert np.isfinite(pair).all()
    assert db_error(pair, ref) <= DB_TOL and db_lin_error(pair, ref) <= LIN_TOL


@pytest.mark.parametrize("n", [256, 2048])
def test_pairs_partners_fixed_by_position(n):
    """A-scans 2k and 2k+1 of a launch share one FFT and the last A-scan of
    an odd count pairs with itself, so each output depends only on its own
    pair: the frame split at an even A-scan, or cut to its last A-scan,
    gives the same floats bit for bit."""
    from octproengine_amd import _capi
    a = 2 * 37 + 1
    p = db_params(_params(n, a, "cubic"))
    raw = S.fringes(n, a, seed=19)
    t = standard_tables(n)
    g = GpuChain(p, t.curve, t.window, t.phasor)
    g.ctx.set_sample_offset(_capi.OFFSET_FIXED, 20000.0)
    whole = g.frame(raw)
    head = g.frame(raw[:a - 1].copy())
    tail = g.frame(raw[a - 3:].copy())  # one pair and the odd A-scan
    last = g.frame(raw[a - 1:].copy())
    g.close()
    assert np.array_equal(whole[:a - 1], head)
    assert np.array_equal(whole[a - 3:], tail)
    assert np.array_equal(whole[a - 1:], last)


def test_pairs_fpn_profile_auto_min_max_background_flip():
    """The epilogue steps on split spectra: FPN subtraction (the profile is
    applied to twice the spectrum, scaled by two), per-A-scan auto min/max,
    B-scan flip and background subtraction, against the single launch and
    the oracle."""
    n, a, b = 1024, 16, 4
    t = standard_tables(n)
    raw = S.fringes(n, a * b, seed=13)
    rng = np.random.default_rng(3)
    prof = (rng.normal(size=n) * 2e3).astype(np.float32)
    bg = (rng.normal(size=n // 2) * 0.01).astype(np.float32)

    # FPN profile, dB output
    p = db_params(_params(n, a, "cubic", bscans_per_buffer=b, fpn_removal=True))
    pair, single = _both(p, raw, t, lambda g: g.ctx.set_fpn_profile(prof))
    st = O.ChainState()
    st.fpn_profile = prof
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor, st)
    print(f"FPN: pair {db_error(pair, ref):.2e} dB, single {db_error(single, ref):.2e} dB")
    assert db_error(pair, ref) <= DB_TOL and db_lin_error(pair, ref) <= LIN_TOL

    # B-scan flip and background subtraction (dB map)
    p = db_params(_params(n, a, "cubic", bscans_per_buffer=b, bscan_flip=True, bg_removal=True, bg_weight=0.5,
                          bg_offset=0.25))
    pair, single = _both(p, raw, t, lambda g: g.ctx.set_bg_profile(bg))
    st = O.ChainState()
    st.bg_profile = bg
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor, st)
    print(f"flip + background: pair {db_error(pair, ref):.2e} dB, single {db_error(single, ref):.2e} dB")
    assert db_error(pair, ref) <= DB_TOL and db_error(single, ref) <= DB_TOL

    # auto min/max (grayscale min == max): the rule applied to the pair's own
    # dB output (the extremes sit at noise-floor bins)
    p = _params(n, a, "cubic", bscans_per_buffer=b, log_scaling=True, grayscale_min=0.0, grayscale_max=0.0)
    auto = _both(p, raw, t)[0].astype(np.float64)
    db = _both(db_params(_params(n, a, "cubic", bscans_per_buffer=b)), raw, t)[0].astype(np.float64)
    mn, mx = db.min(axis=1, keepdims=True), db.max(axis=1, keepdims=True)
    e = float(np.max(np.abs(auto - (db - mn) / (mx - mn))))
    print(f"auto min/max on pairs: max deviation from the rule {e:.2e}")
    assert e <= 1e-5



def test_pairs_with_fpn_determination():
    """One-shot FPN determination and subtraction without dispersion, paired
    and pinned to one A-scan per FFT, against the oracle with the GPU's own
    profile.  The profile removes the DC term, so the bar in dB is applied
    within 40 dB of the reflector peak (as for the 2048 x 1024 x 16
    workload), the linear bar on every bin."""
    n, a = 2048, 256
    p = db_params(_params(n, a, "cubic", fpn_removal=True, fpn_bscan_count=1))
    raw = S.fringes(n, a, seed=17)
    t = standard_tables(n)
    outs, profs = [], []
    for pinned in (False, True):
        g = GpuChain(p, t.curve, t.window, t.phasor)
        if pinned:
            g.ctx.set_launch_variant(1, _passes(n) - 1, 0)
        g.ctx.request_fpn()
        outs.append(g.frame(raw))
        profs.append(g.ctx.get_fpn_profile())
        g.close()
    assert profs[0].tobytes() == profs[1].tobytes()  # the statistics pass is the same kernel
    st = O.ChainState()
    st.fpn_profile = profs[0].astype(np.float32)
    pr = _params(n, a, "cubic", fpn_removal=True, fpn_bscan_count=1000)
    ref = O.process_frame(raw, db_params(pr), t.curve, t.window, t.phasor, st)
    e_p, e_s = db_error(outs[0], ref, 40.0), db_error(outs[1], ref, 40.0)
    el = db_lin_error(outs[0], ref)
    print(f"FPN determination: dB err within 40 dB of peak pair={e_p:.2e} single={e_s:.2e}; "
          f"within 60 dB pair={db_error(outs[0], ref):.2e} single={db_error(outs[1], ref):.2e}; "
          f"pair peak-normalised linear {el:.2e}")
    assert e_p <= DB_TOL and e_s <= DB_TOL and el <= LIN_TOL
