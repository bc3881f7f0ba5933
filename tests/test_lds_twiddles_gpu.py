"""N = 4096 kernels with up to two taps (no resampling, linear) and one
A-scan per lane keep the twiddles of FFT passes 1 and 2 in an LDS table, read
the background constant in the epilogue and pack the tap-window starts, so
that two 512-lane workgroups fit one CU.  Their outputs must meet the parity
bar against the FP64 oracle with every epilogue option, every input width,
one and two workgroups per CU (bit-identical to each other), and through a
ragged batch's row map."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import chain64 as O
from octproengine_amd import synthetic as S
from tests.helpers import DB_TOL, GpuChain, db_error, db_params, standard_tables

pytestmark = pytest.mark.gpu

N = 4096


def _params(a, interp, disp=True, **kw):
    p = O.ChainParams(signal_length=N, ascans_per_bscan=a)
    p.resampling = interp is not None
    p.interpolation = interp or 0
    p.windowing = True
    p.dispersion = disp
    for k, v in kw.items():
        setattr(p, k, v)
    return db_params(p)


@pytest.mark.parametrize("interp", [None, O.LINEAR])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.uint32, np.float32])
def test_input_widths_and_grids(interp, dtype):
    a = 600  # more A-scans than one workgroup per CU holds at once
    p = _params(a, interp)
    t = standard_tables(N)
    raw = S.fringes(N, a, seed=41)
    if dtype == np.uint8:
        raw = (raw >> 8).astype(np.uint8)
    elif dtype == np.uint32:
        raw = raw.astype(np.uint32) * 65537
    elif dtype == np.float32:
        raw = (raw.astype(np.float32) / 65535.0).astype(np.float32)
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    g = GpuChain(p, t.curve, t.window, t.phasor, dtype=dtype)
    outs = []
    for per_cu in (1, 2):
        g.ctx.set_launch_variant(1, 3, per_cu)
        outs.append(g.frame(raw))
    g.close()
    e = db_error(outs[1], ref)
    print(f"[lds twiddles] interp={interp} dtype={np.dtype(dtype)}: {e:.2e} dB; grids equal: "
          f"{np.array_equal(outs[0], outs[1])}")
    assert e <= DB_TOL
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("interp", [None, O.LINEAR])
def test_fpn_profile_and_background(interp):
    """The FPN constant and the background constant read in the epilogue,
    against the oracle and against the arithmetic of the background step
    (cpu_backend.cpp:283-301) on the GPU's own unsubtracted output."""
    a = 64
    t = standard_tables(N)
    rng = np.random.default_rng(7)
    prof = (rng.normal(size=N) * 1e4).astype(np.float32)  # N/2 complex pairs
    raws = [S.fringes(N, a, seed=300 + f) for f in range(2)]
    p0 = _params(a, interp, fpn_removal=True)
    g = GpuChain(p0, t.curve, t.window, t.phasor)
    g.ctx.set_fpn_profile(prof)
    plain = [g.frame(r) for r in raws]
    g.close()
    state = O.ChainState(fpn_profile=prof)
    e = db_error(plain[0], O.process_frame(raws[0], p0, t.curve, t.window, t.phasor, state))
    print(f"[lds twiddles] interp={interp} FPN profile: {e:.2e} dB")
    assert e <= DB_TOL
    w_, off = 0.75, 0.1
    p = _params(a, interp, fpn_removal=True, bg_removal=True, bg_weight=w_, bg_offset=off)
    g = GpuChain(p, t.curve, t.window, t.phasor)
    g.ctx.set_fpn_profile(prof)
    g.ctx.request_bg_recording()
    got = [g.frame(r).astype(np.float64) for r in raws]
    bg = g.ctx.get_bg_profile().astype(np.float64)
    g.close()
    assert np.max(np.abs(bg - plain[0].astype(np.float64).mean(axis=0))) <= 1e-4
    want = plain[1].astype(np.float64) - (np.float32(w_) * bg.astype(np.float32) + np.float32(off))
    dev = float(np.max(np.abs(got[1] - want)))
    print(f"[lds twiddles] interp={interp} background: max abs deviation {dev:.2e}")
    assert dev <= 1e-4


def test_ragged_row_map():
    """Linear requests of 4096 and 1024 samples in one batch: each equal to
    its own frame launch, bit for bit."""
    import torch
    from octproengine_amd import _capi
    from tests.helpers import to_device
    chains = {}
    for n in (1024, 4096):
        p = db_params(O.ChainParams(signal_length=n, ascans_per_bscan=64, resampling=True,
                                    interpolation=O.LINEAR, windowing=True, dispersion=True))
        tt = standard_tables(n)
        g = GpuChain(p, tt.curve, tt.window, tt.phasor)
        g.ctx.set_sample_offset(_capi.OFFSET_FIXED, 20000.0)
        chains[n] = (g, p, tt)
    reqs = []
    for i, (n, a) in enumerate([(4096, 300), (1024, 70), (4096, 1), (4096, 513)]):
        raw = S.fringes(n, a, seed=90 + i)
        reqs.append((n, a, raw, to_device(raw), torch.full((a, n // 2), float("nan"), device="cuda")))
    batch = _capi.RaggedBatch([c[0].ctx for c in chains.values()],
                              [(d.data_ptr(), o.data_ptr(), a, n) for n, a, _, d, o in reqs])
    batch.run(torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    for n, a, raw, _, out in reqs:
        g, p, tt = chains[n]
        got = out.cpu().numpy()
        assert np.array_equal(got, g.frame(raw)), (n, a)
        assert db_error(got, O.process_frame(raw, p, tt.curve, tt.window, tt.phasor)) <= DB_TOL
    batch.close()
    for c in chains.values():
        c[0].close()
