"""Launch variants of the chain (A-scans per lane, position of the FFT's
small-radix pass or the swap schedule, workgroups per CU) and the autotuner
that picks among them: every variant meets the parity bar against the FP64
oracle, and the tuned context processes frames exactly as a context pinned
to the chosen variant does."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import chain64 as O
from octproengine_amd import synthetic as S
from tests.helpers import DB_TOL, GpuChain, db_error, db_params, standard_tables

pytestmark = pytest.mark.gpu

PASSES = {256: 3, 512: 3, 1024: 4, 2048: 4, 4096: 4}


def _chain(n, a, method, dc=False, window=32):
    p = db_params(O.ChainParams(signal_length=n, ascans_per_bscan=a, resampling=True, interpolation=method,
                                windowing=True, dispersion=True, dc_removal=dc, dc_window=window))
    t = standard_tables(n)
    return p, t


# linear + DC (window 64, the reference default) runs the banded DC path in
# every variant: two A-scans per lane with the staging sharing buffer B,
# the Stockham placements, and N = 4096 (512 lanes per A-scan)
@pytest.mark.parametrize("n", [256, 1024, 2048, 4096])
@pytest.mark.parametrize("method,dc", [(O.CUBIC, False), (O.CUBIC, True), (O.LANCZOS, True), (O.LINEAR, True)])
def test_every_launch_variant_meets_parity(n, method, dc):
    a = 96
    p, t = _chain(n, a, method, dc, window=64 if method == O.LINEAR else 32)
    raw = S.fringes(n, a, seed=n + method)
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    g = GpuChain(p, t.curve, t.window, t.phasor)
    small = n not in (512, 4096)
    variants = [(1, k, 0) for k in range(PASSES[n] if small else 1)] + [(2, PASSES[n] - 1, 0), (1, PASSES[n] - 1, 1)]
    if n in (1024, 2048, 4096):
        variants.append((1, PASSES[n], 0))  # the swap schedule
        variants.append((2, PASSES[n], 0))  # swap schedule, two A-scans per lane
    for v in variants:
        if not small and v[0] == 1 and v[1] < PASSES[n] - 1:
            continue
        g.ctx.set_launch_variant(*v)
        e = db_error(g.frame(raw), ref)
        print(f"N={n} method={method} dc={dc} variant {v}: dB err near peak {e:.2e}")
        assert e <= DB_TOL, v
    g.close()


def test_variant_validation():
    from octproengine_amd import _capi
    p, t = _chain(2048, 16, O.CUBIC)
    g = GpuChain(p, t.curve, t.window, t.phasor)
    for bad in [(3, 3, 0), (1, 5, 0), (1, -1, 0), (1, 3, -1)]:
        with pytest.raises(_capi.HipError) as e:
            g.ctx.set_launch_variant(*bad)
        assert e.value.code == _capi.EINVAL
    with pytest.raises(_capi.HipError) as e:
        g.ctx.set_launch_variant(2, 0, 0)  # other placements exist for one A-scan per lane only
    assert e.value.code == _capi.EUNSUP
    g.close()


@pytest.mark.parametrize("method", [O.CUBIC, O.LANCZOS])
def test_autotune_choice_is_used(method):
    import torch
    n, a = 2048, 4096
    p, t = _chain(n, a, method, dc=True)
    raw = S.fringes(n, a, seed=5)
    g = GpuChain(p, t.curve, t.window, t.phasor)
    d_in = torch.from_numpy(raw.view(np.int16)).cuda()
    d_out = torch.empty((a, n // 2), dtype=torch.float32, device="cuda")
    choice, us = g.ctx.autotune(d_in.data_ptr(), d_out.data_ptr(), a)
    print(f"method={method}: autotune chose {choice}, {us:.1f} us per {a} A-scans")
    assert choice["ascans_per_lane"] in (1, 2) and 0 <= choice["small_radix_pass"] <= 4
    assert choice["workgroups_per_cu"] >= 1 and us > 0
    tuned = g.frame(raw)
    h = GpuChain(p, t.curve, t.window, t.phasor)
    h.ctx.set_launch_variant(choice["ascans_per_lane"], choice["small_radix_pass"], choice["workgroups_per_cu"])
    assert np.array_equal(tuned, h.frame(raw))
    assert db_error(tuned, O.process_frame(raw, p, t.curve, t.window, t.phasor)) <= DB_TOL
    # the context's own frame slot
    g.ctx.autotune()
    g.close()
    h.close()
