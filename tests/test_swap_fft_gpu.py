"""The swap schedule of the chain's FFT (launch variant "small-radix pass =
passes", N = 1024, 2048 and 4096): radices 8, 8, S, 8 (8, 8, 8, 8 at 4096)
with the second exchange made inside each wavefront instead of through LDS
(v_permlane32_swap / v_permlane16_swap, and at 4096 a DPP row rotation by 8
with a select for the third register bit).  Every case runs the swap variant and the
default Stockham variant and checks both against the FP64 oracle, and the
two against each other (same arithmetic up to the FFT's rounding order).
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import chain64 as O
from octproengine_amd import synthetic as S
from tests.helpers import (DB_TOL, LIN_TOL, GpuChain, db_error, db_lin_error, db_params, lin_error, lin_params,
                           standard_tables)

pytestmark = pytest.mark.gpu

PASSES = 4  # N = 1024, 2048 and 4096
METHODS = {"none": None, "linear": O.LINEAR, "cubic": O.CUBIC, "lanczos": O.LANCZOS}


def _params(n, a, method, **kw):
    p = O.ChainParams(signal_length=n, ascans_per_bscan=a)
    p.resampling = METHODS[method] is not None
    p.interpolation = METHODS[method] or 0
    p.windowing = True
    p.dispersion = True
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _both(p, raw, t, setup=None, wg_per_cu=0, ascans_per_lane=1):
    """(swap-schedule output, default-variant output) of the same frame."""
    out = []
    for sched in (PASSES, PASSES - 1):
        g = GpuChain(p, t.curve, t.window, t.phasor, dtype=raw.dtype)
        if setup:
            setup(g)
        g.ctx.set_launch_variant(ascans_per_lane, sched, wg_per_cu)
        out.append(g.frame(raw))
        g.close()
    return out


@pytest.mark.parametrize("n", [1024, 2048, 4096])
@pytest.mark.parametrize("method", ["none", "linear", "cubic", "lanczos"])
def test_swap_log_output(n, method):
    a = 67  # odd: the last workgroup iteration is partly empty
    p = db_params(_params(n, a, method))
    raw = S.fringes(n, a, seed=500 + n + len(method))
    t = standard_tables(n)
    swap, default = _both(p, raw, t)
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    e_s, e_d = db_error(swap, ref), db_error(default, ref)
    el = db_lin_error(swap, ref)
    print(f"N={n} {method}: dB err near peak swap={e_s:.2e} default={e_d:.2e}; "
          f"swap vs default {db_error(swap, default):.2e}; swap peak-normalised linear {el:.2e}")
    assert np.isfinite(swap).all()
    assert e_s <= DB_TOL and el <= LIN_TOL
    assert e_d <= DB_TOL


@pytest.mark.parametrize("n", [1024, 2048, 4096])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.uint32, np.float32])
def test_swap_linear_output_and_input_types(n, dtype):
    a = 40
    p = lin_params(_params(n, a, "cubic"))
    raw16 = S.fringes(n, a, seed=9)
    raw = {np.uint8: (raw16 >> 8).astype(np.uint8), np.uint16: raw16,
           np.uint32: raw16.astype(np.uint32) << 8,
           np.float32: (raw16 / 65535.0).astype(np.float32)}[dtype]
    t = standard_tables(n)
    swap, default = _both(p, raw, t)
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    e_s, e_d = lin_error(swap, ref), lin_error(default, ref)
    print(f"N={n} {np.dtype(dtype).name}: peak-normalised linear error swap={e_s:.2e} default={e_d:.2e}")
    assert e_s <= LIN_TOL and e_d <= LIN_TOL


@pytest.mark.parametrize("n", [1024, 2048, 4096])
@pytest.mark.parametrize("window", [0, 1, 64, 300])
def test_swap_with_dc_removal(n, window):
    a = 33
    p = db_params(_params(n, a, "cubic", dc_removal=True, dc_window=window))
    raw = S.fringes(n, a, seed=41 + window)
    t = standard_tables(n)
    swap, default = _both(p, raw, t)
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    e_s, e_d = db_error(swap, ref), db_error(default, ref)
    print(f"N={n} DC window {window}: dB err near peak swap={e_s:.2e} default={e_d:.2e}")
    assert e_s <= DB_TOL and db_lin_error(swap, ref) <= LIN_TOL
    assert e_d <= DB_TOL


def test_swap_fpn_flip_background_and_auto_min_max():
    n, a, b = 2048, 16, 4
    t = standard_tables(n)
    raw = S.fringes(n, a * b, seed=23)
    rng = np.random.default_rng(4)
    prof = (rng.normal(size=n) * 2e3).astype(np.float32)
    bg = (rng.normal(size=n // 2) * 0.01).astype(np.float32)

    p = db_params(_params(n, a, "cubic", bscans_per_buffer=b, fpn_removal=True, bscan_flip=True,
                          bg_removal=True, bg_weight=0.5, bg_offset=0.25))

    def setup(g):
        g.ctx.set_fpn_profile(prof)
        g.ctx.set_bg_profile(bg)

    swap, default = _both(p, raw, t, setup)
    st = O.ChainState()
    st.fpn_profile = prof
    st.bg_profile = bg
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor, st)
    print(f"FPN + flip + background: swap {db_error(swap, ref):.2e} dB, default {db_error(default, ref):.2e} dB")
    assert db_error(swap, ref) <= DB_TOL and db_error(default, ref) <= DB_TOL

    # per-A-scan auto min/max (grayscale min == max) applied to the swap
    # variant's own dB output
    p = _params(n, a, "cubic", bscans_per_buffer=b, log_scaling=True, grayscale_min=0.0, grayscale_max=0.0)
    auto = _both(p, raw, t)[0].astype(np.float64)
    db = _both(db_params(_params(n, a, "cubic", bscans_per_buffer=b)), raw, t)[0].astype(np.float64)
    mn, mx = db.min(axis=1, keepdims=True), db.max(axis=1, keepdims=True)
    e = float(np.max(np.abs(auto - (db - mn) / (mx - mn))))
    print(f"auto min/max with the swap schedule: max deviation from the rule {e:.2e}")
    assert e <= 1e-5


@pytest.mark.parametrize("wg_per_cu", [1, 2, 3])
def test_swap_every_occupancy(wg_per_cu):
    n, a = 2048, 1000
    p = db_params(_params(n, a, "cubic"))
    raw = S.fringes(n, a, seed=77)
    t = standard_tables(n)
    swap, default = _both(p, raw, t, wg_per_cu=wg_per_cu)
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    assert db_error(swap, ref) <= DB_TOL and db_lin_error(swap, ref) <= LIN_TOL


@pytest.mark.parametrize("n", [1024, 2048, 4096])
@pytest.mark.parametrize("case", ["dispersion", "pairs", "pairs-lanczos", "dc"])
def test_swap_two_ascans_per_lane(n, case):
    """Two A-scans per lane: with dispersion each takes its own FFT (the
    staging shares buffer B behind a barrier at the top of each iteration);
    without dispersion the two are the real and imaginary part of one."""
    a = 67
    kw = {"dispersion": case == "dispersion" or case == "dc"}
    if case == "dc":
        kw.update(dc_removal=True, dc_window=48)
    p = db_params(_params(n, a, "lanczos" if case == "pairs-lanczos" else "cubic", **kw))
    raw = S.fringes(n, a, seed=900 + n + len(case))
    t = standard_tables(n)
    swap, default = _both(p, raw, t, ascans_per_lane=2)
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    e_s, e_d = db_error(swap, ref), db_error(default, ref)
    print(f"N={n} {case}, 2 A-scans per lane: dB err near peak swap={e_s:.2e} default={e_d:.2e}")
    assert np.isfinite(swap).all()
    assert e_s <= DB_TOL and db_lin_error(swap, ref) <= LIN_TOL
    assert e_d <= DB_TOL


def test_swap_not_built_elsewhere():
    from octproengine_amd import _capi
    for n, passes in ((256, 3), (512, 3)):
        p = _params(n, 16, "cubic")
        t = standard_tables(n)
        g = GpuChain(p, t.curve, t.window, t.phasor)
        for as_ in (1, 2):
            with pytest.raises(_capi.HipError) as e:
                g.ctx.set_launch_variant(as_, passes, 0)
            assert e.value.code == _capi.EUNSUP
        g.close()
