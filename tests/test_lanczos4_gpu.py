"""Lanczos-4 k-linearization (OPE_HIP_INTERP_LANCZOS4, 8 taps n0-3 .. n0+4):
the reference's Lanczos kernel (cpu_kernels.tpp:148-196) at a = 4, an extra
mode of the C ABI beside the reference's a = 8.  Checked against the FP64
oracle's Lanczos-4 over every length, with and without dispersion and DC
removal, every input width, every launch variant, a ragged batch, and the
individual k-linearization operation."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import chain64 as O
from octproengine_amd import synthetic as S
from tests.helpers import (DB_TOL, LIN_TOL, GpuChain, db_error, db_lin_error, db_params, lin_error, lin_params,
                           standard_tables)

pytestmark = pytest.mark.gpu


def _params(n, a, **kw):
    p = O.ChainParams(signal_length=n, ascans_per_bscan=a, bscans_per_buffer=1, resampling=True,
                      interpolation=O.LANCZOS4, windowing=True, dispersion=True)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("n", [256, 512, 1024, 2048, 4096])
@pytest.mark.parametrize("disp,dc", [(True, False), (False, False), (True, True)])
def test_lanczos4_chain(n, disp, dc):
    a = 40
    raw = S.fringes(n, a, seed=n + 3 * disp + dc)
    t = standard_tables(n)
    for make, err, tol in ((db_params, db_error, DB_TOL), (lin_params, lin_error, LIN_TOL)):
        p = make(_params(n, a, dispersion=disp, dc_removal=dc, dc_window=64))
        g = GpuChain(p, t.curve, t.window, t.phasor)
        got = g.frame(raw)
        g.close()
        ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
        e = err(got, ref)
        print(f"[lanczos-4] N={n} disp={disp} dc={dc} {make.__name__}: {e:.2e}")
        assert e <= tol


@pytest.mark.parametrize("dtype", [np.uint8, np.uint32, np.float32])
def test_lanczos4_input_types(dtype):
    n, a = 2048, 24
    raw = S.fringes(n, a, seed=11)
    raw = {np.uint8: lambda r: (r >> 8).astype(np.uint8), np.uint32: lambda r: r.astype(np.uint32) * 3,
           np.float32: lambda r: (r / 65535.0).astype(np.float32)}[dtype](raw)
    t = standard_tables(n)
    p = db_params(_params(n, a))
    g = GpuChain(p, t.curve, t.window, t.phasor, dtype=dtype)
    got = g.frame(raw)
    g.close()
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    assert db_error(got, ref) <= DB_TOL and db_lin_error(got, ref) <= LIN_TOL


@pytest.mark.parametrize("n", [1024, 2048, 4096])
def test_lanczos4_every_launch_variant(n):
    a = 64
    passes = {1024: 4, 2048: 4, 4096: 4}[n]
    raw = S.fringes(n, a, seed=n)
    t = standard_tables(n)
    p = db_params(_params(n, a, dc_removal=True, dc_window=32))
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    g = GpuChain(p, t.curve, t.window, t.phasor)
    variants = [(1, passes - 1, 0), (2, passes - 1, 0), (1, passes - 1, 1)]
    if n != 4096:
        variants += [(1, k, 0) for k in range(passes - 1)] + [(1, passes, 0), (2, passes, 0)]
    for v in variants:
        g.ctx.set_launch_variant(*v)
        e = db_error(g.frame(raw), ref)
        print(f"[lanczos-4] N={n} variant {v}: {e:.2e}")
        assert e <= DB_TOL, v
    g.close()


def test_lanczos4_ragged_and_individual_operation():
    import torch
    from octproengine_amd import _capi
    from tests.helpers import to_device
    chains = {}
    for n in (1024, 2048, 4096):
        p = db_params(_params(n, 64))
        t = standard_tables(n)
        g = GpuChain(p, t.curve, t.window, t.phasor)
        g.ctx.set_sample_offset(_capi.OFFSET_FIXED, 20000.0)
        chains[n] = (g, p, t)
    reqs = []
    for i, (n, a) in enumerate([(1024, 37), (4096, 5), (2048, 100), (1024, 1)]):
        raw = S.fringes(n, a, seed=70 + i)
        reqs.append((n, a, raw, to_device(raw), torch.full((a, n // 2), float("nan"), device="cuda")))
    batch = _capi.RaggedBatch([c[0].ctx for c in chains.values()],
                              [(d.data_ptr(), o.data_ptr(), a, n) for n, a, _, d, o in reqs])
    batch.run(torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    for n, a, raw, _, out in reqs:
        g, p, t = chains[n]
        got = out.cpu().numpy()
        assert np.max(np.abs(got - g.frame(raw))) <= 1e-4
        assert db_error(got, O.process_frame(raw, p, t.curve, t.window, t.phasor)) <= DB_TOL
    batch.close()
    # the individual kLinearization operation with method 3
    g = chains[2048][0]
    rng = np.random.default_rng(3)
    line = rng.normal(0, 1000, 2048).astype(np.float32)
    curve = S.resampling_curve(2048)
    out = np.zeros(2048, np.float32)
    import ctypes
    from tests.helpers import _fp
    rc = _capi.lib.ope_hip_op_klinearization(g.ctx._h, _fp(line), _fp(curve), _capi.LANCZOS4, 2048, 2048,
                                             out.ctypes.data)
    assert rc == 0
    want = O.klin_lanczos(line.astype(np.float64), curve, a=4)
    assert np.max(np.abs(out - want)) <= 1e-4 * np.max(np.abs(want))
    for c in chains.values():
        c[0].close()
