"""Differentiable chain for autofocus (theta = d2, d3, c0..c3).

CPU: the FP64 oracle's analytic gradient (oracle/autofocus64.py) is pinned
against torch float64 autograd of an independent transcription of the
forward, and against central finite differences; a planted sign error in
the oracle's backward must fail those checks.
GPU: the hand-written HIP forward/backward (ope_hip_autofocus_*) against the
oracle and against finite differences (h = 1e-3 relative for d2, d3), and
50 Adam steps through torch.autograd recovering the dispersion of fringes
generated with d2 = 8e-6, d3 = -3e-9."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import autofocus64 as F
from oracle import chain64 as O
from octproengine_amd import synthetic as S

TRUE_D2, TRUE_D3 = 8e-6, -3e-9


def fitted_curve_coeffs(n: int) -> np.ndarray:
    """Least-squares cubic in u = j/(N-1) of the generator's k^-1 curve."""
    r = S.resampling_curve(n, clamp=False).astype(np.float64)
    u = np.arange(n) / (n - 1)
    a = np.vstack([u ** 0, u, u ** 2, u ** 3]).T
    return np.linalg.lstsq(a, r, rcond=None)[0]


def setup(n, a, seed=1, dtype=np.float32):
    raw = S.fringes(n, a, seed=seed, dtype=dtype, d2=TRUE_D2 * (2048 / n) ** 2, d3=TRUE_D3 * (2048 / n) ** 3)
    window = O.generate_window(n, O.HANN, 0.5, 0.95)
    return raw, window, fitted_curve_coeffs(n)


def torch_entropy(x, theta, window, method):
    """Independent float64 transcription of the forward for autograd."""
    import torch
    n = x.shape[-1]
    x = torch.as_tensor(x, dtype=torch.float64)
    u = torch.arange(n, dtype=torch.float64) / (n - 1)
    r = theta[2] + theta[3] * u + theta[4] * u ** 2 + theta[5] * u ** 3
    rc = torch.clamp(r, 7.0, n - 9.0)
    n1 = torch.floor(rc.detach()).long()
    t = rc - n1
    if method == F.LINEAR:
        y = x[:, n1] + (x[:, n1 + 1] - x[:, n1]) * t
    else:
        y0, y1, y2, y3 = x[:, n1 - 1], x[:, n1], x[:, n1 + 1], x[:, n1 + 2]
        ca = -y0 + 3 * (y1 - y2) + y3
        cb = 2 * y0 - 5 * y1 + 4 * y2 - y3
        cc = -y0 + y2
        y = 0.5 * t * (ca * t * t + cb * t + cc) + y1
    d = torch.arange(n, dtype=torch.float64) - n / 2
    phi = theta[0] * d ** 2 + theta[1] * d ** 3
    w = torch.as_tensor(window, dtype=torch.float64)
    s = torch.complex(y * w * torch.cos(phi), y * w * torch.sin(phi))
    X = torch.fft.ifft(s, dim=-1, norm="forward")[:, : n // 2]
    intensity = X.real ** 2 + X.imag ** 2
    total = intensity.sum()
    return torch.log(total) - (intensity * torch.log(intensity)).sum() / total


def fd_steps(theta, method):
    """d2, d3: 1e-3 relative; curve coefficients: a displacement of at most
    1e-3 (cubic) or 1e-4 (linear, whose interpolant has kinks) samples."""
    hc = 1e-3 if method == F.CUBIC else 1e-4
    return np.r_[1e-3 * abs(theta[0]), 1e-3 * abs(theta[1]), hc, hc, hc, hc]


def central_fd(x, theta, window, method, h):
    g = np.zeros(6)
    for i in range(6):
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h[i]
        tm[i] -= h[i]
        g[i] = (F.forward(x, tp, window, method)[0] - F.forward(x, tm, window, method)[0]) / (2 * h[i])
    return g


def away_from_optimum(c):
    return np.r_[0.5 * TRUE_D2, 0.3 * TRUE_D3, c + np.r_[0.5, 3.0, -2.0, 1.0]]


@pytest.mark.parametrize("method", [F.CUBIC, F.LINEAR])
def test_oracle_gradient_matches_torch_autograd(method):
    import torch
    raw, window, c = setup(256, 12)
    x = F.prepare(raw, 32)
    theta = away_from_optimum(c)
    g = F.gradient(x, theta, window, method)
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    loss = torch_entropy(x, th, window, method)
    loss.backward()
    assert abs(float(loss) - F.forward(x, theta, window, method)[0]) < 1e-12
    rel = np.abs(th.grad.numpy() - g) / np.abs(g)
    assert rel.max() < 1e-9, rel


@pytest.mark.parametrize("method", [F.CUBIC, F.LINEAR])
def test_oracle_gradient_matches_finite_differences(method):
    raw, window, c = setup(512, 16)
    x = F.prepare(raw, 32)
    theta = away_from_optimum(c)
    g = F.gradient(x, theta, window, method)
    fd = central_fd(x, theta, window, method, fd_steps(theta, method) / 10)
    rel = np.abs(fd - g) / np.abs(g)
    assert rel.max() < 1e-5, rel


def test_planted_sign_error_in_backward_is_caught(monkeypatch):
    raw, window, c = setup(256, 8)
    x = F.prepare(raw, 32)
    theta = away_from_optimum(c)
    good = F.gradient(x, theta, window, F.CUBIC)
    real_fft = np.fft.fft
    # the backward's forward DFT with the wrong sign (an inverse DFT)
    monkeypatch.setattr(F.np.fft, "fft", lambda a, axis=-1: np.fft.ifft(a, axis=axis) * a.shape[axis])
    bad = F.gradient(x, theta, window, F.CUBIC)
    monkeypatch.setattr(F.np.fft, "fft", real_fft)
    fd = central_fd(x, theta, window, F.CUBIC, fd_steps(theta, F.CUBIC) / 10)
    assert np.max(np.abs(good - fd) / np.abs(fd)) < 1e-5
    assert np.max(np.abs(bad - fd) / np.abs(fd)) > 1e-2


def test_generator_dispersion_is_removed_by_matching_coefficients():
    raw, window, c = setup(2048, 32)
    x = F.prepare(raw, 64)
    at = lambda d2, d3: F.forward(x, np.r_[d2, d3, c], window, F.CUBIC)[0]
    best = at(TRUE_D2, TRUE_D3)
    assert best < at(0.0, 0.0) and best < at(-TRUE_D2, -TRUE_D3)
    assert best < at(1.1 * TRUE_D2, TRUE_D3) and best < at(0.9 * TRUE_D2, TRUE_D3)


# ---------------------------------------------------------------- GPU


def gpu_context(n, method, dtype):
    from octproengine_amd import _capi
    from tests.helpers import np_to_code
    params = _capi.make_params(signal_length=n, ascans_per_bscan=512, bscans_per_buffer=1,
                               input_type=np_to_code(dtype), resampling_enabled=1, interpolation=method,
                               windowing_enabled=1, dc_removal_enabled=1, dc_window_size=64)
    ctx = _capi.Context(params)
    ctx.set_window_curve(O.generate_window(n, O.HANN, 0.5, 0.95))
    return ctx


@pytest.mark.gpu
@pytest.mark.parametrize("method,dtype", [(F.CUBIC, np.float32), (F.LINEAR, np.float32), (F.CUBIC, np.uint16)])
def test_gpu_loss_and_gradient_match_oracle_and_fd(method, dtype):
    from tests.helpers import to_device
    n, a = 2048, 512
    raw, window, c = setup(n, a, seed=2, dtype=dtype)
    x = F.prepare(O.convert_input(raw), 64)
    ctx = gpu_context(n, method, dtype)
    d_in = to_device(raw)
    theta = away_from_optimum(c)
    print(f"[autofocus] N={n} A={a} method={method} dtype={np.dtype(dtype)}", flush=True)
    loss, s, _ = ctx.autofocus_loss(d_in.data_ptr(), a, theta)
    loss64, s64, _, _ = F.forward(x, theta, window, method)
    assert abs(loss - loss64) < 1e-5 * abs(loss64), (loss, loss64)
    assert abs(s - s64) < 1e-5 * s64
    g = ctx.autofocus_grad(d_in.data_ptr(), a, theta, s, loss)
    g64 = F.gradient(x, theta, window, method)
    fd = central_fd(x, theta, window, method, fd_steps(theta, method))
    rel64 = np.abs(g - g64) / np.abs(g64)
    rel_fd = np.abs(g - fd) / np.abs(fd)
    print(f"[autofocus] |grad - FP64| / |FP64| = {np.array2string(rel64, precision=2)}")
    print(f"[autofocus] |grad - FD|   / |FD|   = {np.array2string(rel_fd, precision=2)}")
    assert rel64.max() < 1e-3
    assert rel_fd[:2].max() < 1e-3  # dispersion: h = 1e-3 relative
    assert rel_fd.max() < 2e-3
    ctx.close()


@pytest.mark.gpu
def test_gpu_adam_autofocus_recovers_dispersion():
    import torch
    from octproengine_amd.autofocus import autofocus
    from tests.helpers import to_device
    n, a = 2048, 512
    raw, window, c = setup(n, a, seed=3)
    ctx = gpu_context(n, F.CUBIC, np.float32)
    frames = to_device(raw)
    theta0 = np.r_[0.0, 0.0, c]
    theta, hist = autofocus(ctx, frames, theta0, steps=50, lr=0.1, scale=np.r_[1e-5, 1e-9, 1, 1, 1, 1],
                            free=(True, True, False, False, False, False))
    print(f"[autofocus] loss {hist[0]:.4f} -> {hist[-1]:.4f}; d2 {theta[0]:.4e}, d3 {theta[1]:.4e}")
    assert hist[-1] < hist[0] - 0.5
    assert abs(theta[0] - TRUE_D2) < 0.05 * TRUE_D2
    assert abs(theta[1] - TRUE_D3) < 0.3 * abs(TRUE_D3)
    # the torch.autograd path gives the kernel's gradient
    from octproengine_amd.autofocus import entropy
    th = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    entropy(th, ctx, frames).backward()
    loss, s, _ = ctx.autofocus_loss(frames.data_ptr(), a, theta0)
    assert np.allclose(th.grad.numpy(), ctx.autofocus_grad(frames.data_ptr(), a, theta0, s, loss), rtol=1e-12)
    ctx.close()
