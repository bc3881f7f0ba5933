"""Sample offset (ope_hip_set_sample_offset): the chain runs on x - m0 and
adds m0 * IFFT(S * phasor) back in FP64 through the epilogue constant.

The outputs must be those of the chain on x (parity with the FP64 oracle in
every mode and input type), the automatic offset must be the rounded mean of
the first 32 A-scans, FPN profiles must be the oracle's, and the offset
must follow curve hot-swaps, graphs and the DC-removal switch.  With
FPN removal the offset is what brings weak bins within 60 dB of the peak
inside the 1e-3 dB bar."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import chain64 as O
from octproengine_amd import synthetic as S
from tests.helpers import (DB_TOL, LIN_TOL, GpuChain, db_error, db_lin_error, db_params, standard_tables,
                           to_device)

pytestmark = pytest.mark.gpu


def _params(n, a, interp=O.CUBIC, **kw):
    p = O.ChainParams(signal_length=n, ascans_per_bscan=a, bscans_per_buffer=1, resampling=interp is not None,
                      interpolation=interp or 0, windowing=True, dispersion=True)
    for k, v in kw.items():
        setattr(p, k, v)
    return db_params(p)


def _raw(n, a, dtype, seed):
    raw = S.fringes(n, a, seed=seed)
    if dtype == np.uint8:
        return (raw >> 8).astype(np.uint8)
    if dtype == np.uint32:
        return raw.astype(np.uint32) * np.uint32(40000)
    if dtype == np.float32:
        return (raw / 65535.0).astype(np.float32)
    return raw


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.uint32, np.float32])
def test_auto_offset_is_the_rounded_mean_and_keeps_parity(dtype):
    from octproengine_amd import _capi
    n, a = 2048, 96
    p = _params(n, a)
    t = standard_tables(n)
    raw = _raw(n, a, dtype, 31)
    g = GpuChain(p, t.curve, t.window, t.phasor, dtype=dtype)
    assert g.ctx.get_sample_offset() == (_capi.OFFSET_AUTO, 0.0)  # undetermined
    got = g.frame(raw)
    mode, m0 = g.ctx.get_sample_offset()
    x = O.convert_input(raw[:32], True, False).astype(np.float64)
    want = x.mean() if dtype == np.float32 else np.rint(x.mean())
    print(f"[offset {np.dtype(dtype)}] m0 = {m0!r}, expected {want!r}")
    assert mode == _capi.OFFSET_AUTO
    assert abs(m0 - want) <= 1e-6 * abs(want)
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    e, el = db_error(got, ref), db_lin_error(got, ref)
    print(f"[offset {np.dtype(dtype)}] dB err near peak {e:.2e}, linear {el:.2e}")
    assert e <= DB_TOL and el <= LIN_TOL
    # the offset is kept: a second frame with another level does not move it
    g.frame((raw // 2).astype(dtype) if dtype != np.float32 else raw * 0.5)
    assert g.ctx.get_sample_offset()[1] == m0
    g.close()


@pytest.mark.parametrize("mode,value", [(0, 0.0), (2, 0.0), (2, 20000.0), (2, 65535.0), (2, 7.0)])
def test_every_offset_gives_the_chain_on_x(mode, value):
    """Without FPN removal any offset (even one far from the level) stays
    within the parity bar: the result is the chain on x."""
    n, a = 1024, 128
    p = _params(n, a)
    t = standard_tables(n)
    raw = S.fringes(n, a, seed=32)
    g = GpuChain(p, t.curve, t.window, t.phasor)
    g.ctx.set_sample_offset(mode, value)
    got = g.frame(raw)
    assert g.ctx.get_sample_offset() == (mode, value if mode == 2 else 0.0)
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    e, el = db_error(got, ref), db_lin_error(got, ref)
    print(f"[offset mode {mode} value {value}] dB err {e:.2e}, linear {el:.2e}")
    assert e <= DB_TOL and el <= LIN_TOL
    g.close()


@pytest.mark.parametrize("interp,disp", [(O.CUBIC, True), (O.LINEAR, False), (O.LANCZOS, True), (None, True)])
def test_fpn_removal_within_60db_with_offset(interp, disp):
    """FPN removal takes the DC term away, so weak bins come within 60 dB of
    the reflector peak: with the offset they meet the 1e-3 dB bar there, and
    the profile is the oracle's (the statistics add the offset's term back).
    Without the offset the same bins carry the rounding of the DC term."""
    from octproengine_amd import _capi
    n, a = 2048, 512
    p = _params(n, a, interp, dispersion=disp, fpn_removal=True, fpn_bscan_count=1)
    t = standard_tables(n)
    warm = S.fringes(n, a, seed=33)
    raw = S.fringes(n, a, seed=34)
    profs, outs = [], []
    for mode in (_capi.OFFSET_AUTO, _capi.OFFSET_OFF):
        g = GpuChain(p, t.curve, t.window, t.phasor)
        g.ctx.set_sample_offset(mode)
        g.ctx.request_fpn()
        g.frame(warm)
        profs.append(g.ctx.get_fpn_profile())
        outs.append(g.frame(raw))
        g.close()
    refs = []
    for prof in profs:
        st = O.ChainState()
        st.fpn_profile = prof.astype(np.float32)
        refs.append(O.process_frame(raw, p, t.curve, t.window, t.phasor, st))
    ref = refs[0]
    e_on, e_off = db_error(outs[0], ref), db_error(outs[1], refs[1])
    # the profile against the oracle's determination (bins away from DC,
    # where the reference's own segment choice hinges on float rounding)
    det = O.ChainState()
    det.fpn_requested = True
    O.process_frame(warm, p, t.curve, t.window, t.phasor, det)
    want = det.fpn_profile.astype(np.float64)
    far = np.repeat(np.arange(want.size // 2) >= 16, 2)
    dprof = np.max(np.abs(profs[0][far] - want[far])) / np.max(np.abs(want))
    print(f"[fpn {interp} disp={disp}] dB err within 60 dB: offset {e_on:.2e}, none {e_off:.2e}; "
          f"profile vs oracle {dprof:.2e} of the peak (bins >= 16)")
    assert e_on <= DB_TOL and db_lin_error(outs[0], ref) <= LIN_TOL
    assert dprof <= 1e-4


def test_offset_follows_bitshift_and_hot_swaps():
    from octproengine_amd import _capi
    n, a = 2048, 128
    p = _params(n, a, fpn_removal=True, fpn_bscan_count=1)
    t = standard_tables(n)
    raw = S.fringes(n, a, seed=35)
    g = GpuChain(p, t.curve, t.window, t.phasor)
    g.ctx.request_fpn()
    g.frame(raw)
    m0 = g.ctx.get_sample_offset()[1]
    prof = g.ctx.get_fpn_profile()
    # dispersion curve hot-swap: D is rebuilt for the new phasor
    t2 = standard_tables(n, d2=9e-6, d3=-4e-9)
    g.ctx.set_dispersion_curve(t2.phasor)
    got = g.frame(raw)
    st = O.ChainState()
    st.fpn_profile = prof.astype(np.float32)
    ref = O.process_frame(raw, p, t.curve, t.window, t2.phasor, st)
    assert db_error(got, ref) <= DB_TOL and db_lin_error(got, ref) <= LIN_TOL
    # window hot-swap, then the FPN switch off
    w2 = O.generate_window(n, O.HANN, 0.45, 0.8)
    g.ctx.set_window_curve(w2)
    got = g.frame(raw)
    st = O.ChainState()
    st.fpn_profile = prof.astype(np.float32)
    ref = O.process_frame(raw, p, t.curve, w2, t2.phasor, st)
    assert db_error(got, ref) <= DB_TOL
    from tests.helpers import chain_params
    p_nofpn = _params(n, a)
    g.ctx.set_params(chain_params(p_nofpn))
    got = g.frame(raw)
    ref = O.process_frame(raw, p_nofpn, t.curve, w2, t2.phasor)
    assert db_error(got, ref) <= DB_TOL
    assert g.ctx.get_sample_offset()[1] == m0
    # bitshift: the automatic offset is determined again on shifted samples
    p_shift = _params(n, a, bitshift=True)
    g.ctx.set_params(chain_params(p_shift))
    assert g.ctx.get_sample_offset() == (_capi.OFFSET_AUTO, 0.0)
    got = g.frame(raw)
    m1 = g.ctx.get_sample_offset()[1]
    assert m1 == np.rint((raw[:32] >> 4).astype(np.float64).mean())
    ref = O.process_frame(raw, p_shift, t.curve, w2, t2.phasor)
    assert db_error(got, ref) <= DB_TOL
    g.close()


def test_offset_inactive_with_dc_removal_and_bad_arguments():
    from octproengine_amd import _capi
    n, a = 1024, 64
    p = _params(n, a, dc_removal=True, dc_window=16)
    t = standard_tables(n)
    raw = S.fringes(n, a, seed=36)
    g = GpuChain(p, t.curve, t.window, t.phasor)
    got = g.frame(raw)
    assert g.ctx.get_sample_offset() == (_capi.OFFSET_AUTO, 0.0)  # not used, still undetermined
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    assert db_error(got, ref) <= DB_TOL
    for mode, value in ((3, 0.0), (-1, 0.0), (_capi.OFFSET_FIXED, 0.5), (_capi.OFFSET_FIXED, float("inf"))):
        with pytest.raises(_capi.HipError) as e:
            g.ctx.set_sample_offset(mode, value)
        assert e.value.code == _capi.EINVAL
    g.close()


def test_graph_capture_determines_a_pending_offset():
    """A graph captured while an automatic offset is undetermined determines
    it from its first frame before the capture: replays and later eager
    frames agree bit for bit."""
    import torch
    from octproengine_amd import _capi
    n, a, frames = 2048, 512, 3
    p = _params(n, a)
    t = standard_tables(n)
    g = GpuChain(p, t.curve, t.window, t.phasor)
    g.ctx.set_sample_offset(_capi.OFFSET_OFF)
    g.frame(S.fringes(n, a, seed=39))  # the FPN statistics of the first B-scan
    g.ctx.set_sample_offset(_capi.OFFSET_AUTO)
    raws = [S.fringes(n, a, seed=40 + f) for f in range(frames)]
    d_in = [to_device(r) for r in raws]
    d_out = [torch.full((a, n // 2), float("nan"), dtype=torch.float32, device="cuda") for _ in range(frames)]
    graph = _capi.FrameGraph(g.ctx, [x.data_ptr() for x in d_in], [y.data_ptr() for y in d_out], a)
    m0 = g.ctx.get_sample_offset()[1]
    assert m0 == np.rint(raws[0][:32].astype(np.float64).mean())
    for rep in range(2):
        graph.launch(torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        for r, y in zip(raws, d_out):
            assert np.array_equal(y.cpu().numpy(), g.frame(r)), f"replay {rep}"
    graph.close()
    g.close()


def test_curve_hot_swap_waits_for_frames_on_a_caller_stream():
    """A frame running on a non-default torch stream keeps the tables it
    started with: the setter waits for it before uploading."""
    import torch
    n, a = 2048, 1 << 16
    p = _params(n, 1024)
    t = standard_tables(n)
    g = GpuChain(p, t.curve, t.window, t.phasor)
    d_in = S.fringes_torch(n, a, seed=41)
    out_old = torch.empty((a, n // 2), dtype=torch.float32, device="cuda")
    out_race = torch.empty_like(out_old)
    g.ctx.process_device(d_in.data_ptr(), out_old.data_ptr(), a, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    curve2 = (t.curve + 0.5).astype(np.float32)
    for _ in range(3):
        out_race.fill_(float("nan"))
        torch.cuda.synchronize()
        with torch.cuda.stream(side):
            g.ctx.process_device(d_in.data_ptr(), out_race.data_ptr(), a, side.cuda_stream)
        g.ctx.set_resampling_curve(curve2)  # issued while the frame runs
        torch.cuda.synchronize()
        assert torch.equal(out_race, out_old)
        g.ctx.set_resampling_curve(t.curve)
    g.close()
