"""Steps only the reference's CUDA backend has (its CPU backend ignores the
flags): input bitshift (cuda_kernels.cu:45-70), B-scan flip
(cuda_kernels.cu:553-571) and sinusoidal scan correction
(cuda_kernels.cu:408-437), each against the oracle's restatement of the CUDA
kernel (oracle/chain64.py), whose closed-form checks are in test_oracle.py."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import chain64 as O
from octproengine_amd import synthetic as S
from tests.helpers import DB_TOL, GpuChain, db_error, db_params, standard_tables

pytestmark = pytest.mark.gpu


def _params(n, a, b, **kw):
    p = db_params(O.ChainParams(signal_length=n, ascans_per_bscan=a, bscans_per_buffer=b, resampling=True,
                                interpolation=O.CUBIC, windowing=True, dispersion=True))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.uint32, np.float32])
def test_bitshift(dtype):
    n, a = 1024, 40
    raw16 = S.fringes(n, a, seed=21)
    raw = {np.uint8: (raw16 >> 8).astype(np.uint8), np.uint16: raw16,
           np.uint32: raw16.astype(np.uint32) * 65537,
           np.float32: (raw16 / 65535.0).astype(np.float32)}[dtype]
    p = _params(n, a, 1, bitshift=True)
    t = standard_tables(n)
    g = GpuChain(p, t.curve, t.window, t.phasor, dtype=dtype)
    got = g.frame(raw)
    g.close()
    ref = O.process_frame(raw, p, t.curve, t.window, t.phasor)
    e = db_error(got, ref)
    print(f"bitshift {np.dtype(dtype).name}: dB err near peak {e:.2e}")
    assert e <= DB_TOL
    if dtype != np.float32:
        p.bitshift = False
        unshifted = O.process_frame(raw, p, t.curve, t.window, t.phasor)
        assert np.max(np.abs(unshifted - ref)) > 1.0  # the option changes the data


@pytest.mark.parametrize("bscans", [4, 3, 1])
def test_bscan_flip_is_a_row_permutation(bscans):
    n, a = 512, 24
    raw = S.fringes(n, a * bscans, seed=30 + bscans)
    t = standard_tables(n)
    g0 = GpuChain(_params(n, a, bscans), t.curve, t.window, t.phasor)
    plain = g0.frame(raw)
    g0.close()
    p = _params(n, a, bscans, bscan_flip=True)
    g = GpuChain(p, t.curve, t.window, t.phasor)
    got = g.frame(raw)
    g.close()
    want = O.bscan_flip(plain, a)
    assert np.array_equal(got, want.astype(np.float32))
    assert db_error(got, O.process_frame(raw, p, t.curve, t.window, t.phasor)) <= DB_TOL
    if bscans >= 2:
        assert not np.array_equal(got, plain)


@pytest.mark.parametrize("flip,bg", [(False, False), (True, False), (True, True)])
def test_sinusoidal_correction(flip, bg):
    """The correction interpolates between two A-scans, so near-peak dB
    parity of a corrected A-scan would mix in the other A-scan's noise
    floor.  Instead: the corrected frame equals the oracle's restatement of
    the CUDA kernel applied to the GPU's own uncorrected frame (itself
    checked against FP64 here), then the background steps."""
    n, a, b = 512, 64, 2
    w_, off = 0.5, 0.25
    raws = [S.fringes(n, a * b, seed=40 + f) for f in range(2)]
    t = standard_tables(n)
    p0 = _params(n, a, b, bscan_flip=flip)
    g0 = GpuChain(p0, t.curve, t.window, t.phasor)
    plain = [g0.frame(r).astype(np.float64) for r in raws]
    g0.close()
    for r, out in zip(raws, plain):
        assert db_error(out, O.process_frame(r, p0, t.curve, t.window, t.phasor)) <= DB_TOL
    curve = O.sinusoidal_curve(a)
    corrected = [O.sinusoidal_correction(out, a, curve) for out in plain]
    p = _params(n, a, b, bscan_flip=flip, sinusoidal=True, bg_removal=bg, bg_weight=w_, bg_offset=off)
    g = GpuChain(p, t.curve, t.window, t.phasor)
    if bg:
        g.ctx.request_bg_recording()
    got = [g.frame(r).astype(np.float64) for r in raws]
    if bg:
        prof = g.ctx.get_bg_profile().astype(np.float64)
        assert np.max(np.abs(prof - corrected[0].mean(axis=0))) <= 1e-4
        sub = np.float32(w_) * prof.astype(np.float32) + np.float32(off)
        want = [c - sub for c in corrected]
    else:
        want = corrected
    g.close()
    # the curve is generated on the host with the C library's acosf, the
    # restatement with NumPy's float32 arccos: positions may differ by an
    # ulp, which moves an interpolated value by ulp * |f1 - f0|
    ulp = float(np.spacing(np.float32(a)))
    for f in range(2):
        jump = float(np.max(np.abs(np.diff(plain[f], axis=0))))
        e = float(np.max(np.abs(got[f] - want[f])))
        bound = 1e-4 + 2 * ulp * jump
        print(f"sinusoidal flip={flip} bg={bg} frame {f}: max |GPU - restatement| = {e:.2e} dB "
              f"(bound {bound:.2e})")
        assert e <= bound
