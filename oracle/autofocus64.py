"""FP64 NumPy oracle of the differentiable chain used for autofocus.

TEST INFRASTRUCTURE ONLY: the checker of the HIP autofocus kernels.  Only
``tests/`` and the ``cpu_baseline`` leg of ``bench.py`` may import it.

The reference has no autofocus; the chain steps follow its CPU backend:
  rolling-average DC removal      cpu_kernels.tpp:56-82   (chain64.rolling_average_dc)
  linear / cubic k-linearization  cpu_kernels.tpp:84-146  (same interpolants)
  resampling polynomial           processorconfiguration.cpp:438-459 (u = j/(N-1))
  window, dispersion multiply     cpu_kernels.tpp:198-235
  unnormalised inverse FFT        cpu_kernels.tpp:238-267
Parameters theta = (d2, d3, c0, c1, c2, c3): resampling positions
r[j] = clamp(c0 + c1 u + c2 u^2 + c3 u^3, 7, N-9) and dispersion phase
phi[j] = d2 (j-N/2)^2 + d3 (j-N/2)^3.  Loss: Shannon entropy of the
intensity I = |X[k]|^2, k < N/2, normalised over the whole frame.

Pinning (tests/test_autofocus.py): ``gradient`` is checked against torch
float64 autograd of an independent transcription of ``forward`` and against
central finite differences of ``forward``; ``forward``'s chain steps reuse
the pinned functions of chain64.
"""

from __future__ import annotations

import numpy as np

from oracle import chain64 as O

LINEAR, CUBIC = 0, 1


def positions(n: int, c):
    """Resampling positions, their clamped values and the clamp mask."""
    u = np.arange(n, dtype=np.float64) / (n - 1)
    r = c[0] + u * (c[1] + u * (c[2] + u * c[3]))
    rc = np.clip(r, 7.0, n - 9.0)
    inside = (r > 7.0) & (r < n - 9.0)
    return u, rc, inside


def phase(n: int, d2: float, d3: float):
    d = np.arange(n, dtype=np.float64) - n / 2
    return d, d * d * (d2 + d3 * d)


def interpolate(x: np.ndarray, rc: np.ndarray, method: int):
    """y[j] = interpolant of x at rc[j] and dy/dt (t = fractional part)."""
    n1 = np.floor(rc).astype(np.int64)
    t = rc - n1
    if method == LINEAR:
        x1, x2 = x[..., n1], x[..., n1 + 1]
        return x1 + (x2 - x1) * t, x2 - x1
    y0, y1, y2, y3 = x[..., n1 - 1], x[..., n1], x[..., n1 + 1], x[..., n1 + 2]
    a = -y0 + 3.0 * (y1 - y2) + y3
    b = 2.0 * y0 - 5.0 * y1 + 4.0 * y2 - y3
    c = -y0 + y2
    return 0.5 * t * (a * t * t + b * t + c) + y1, 0.5 * (3.0 * a * t * t + 2.0 * b * t + c)


def prepare(raw: np.ndarray, dc_window: int | None) -> np.ndarray:
    """Input samples as float64, DC-removed when dc_window is not None."""
    x = np.asarray(raw, dtype=np.float64)
    return O.rolling_average_dc(x, dc_window) if dc_window is not None else x


def forward(x: np.ndarray, theta, window: np.ndarray, method: int):
    """x: [A][N] prepared samples.  Returns (loss, S, Q, cache)."""
    n = x.shape[-1]
    d2, d3, c = theta[0], theta[1], theta[2:6]
    u, rc, inside = positions(n, c)
    y, dydt = interpolate(x, rc, method)
    d, phi = phase(n, d2, d3)
    e = np.exp(1j * phi)
    w = np.asarray(window, dtype=np.float64)
    s = y * w * e
    X = O.ifft_unnormalized(s)[..., : n // 2]
    intensity = X.real ** 2 + X.imag ** 2
    S = intensity.sum()
    with np.errstate(divide="ignore", invalid="ignore"):
        Q = np.where(intensity > 0, intensity * np.log(intensity), 0.0).sum()
    loss = np.log(S) - Q / S
    cache = dict(u=u, inside=inside, dydt=dydt, d=d, e=e, w=w, s=s, X=X, I=intensity, S=S, loss=loss)
    return loss, S, Q, cache


def gradient(x: np.ndarray, theta, window: np.ndarray, method: int) -> np.ndarray:
    """d loss / d theta by the analytic backward of the chain."""
    n = x.shape[-1]
    loss, S, _, k = forward(x, theta, window, method)
    intensity, X = k["I"], k["X"]
    with np.errstate(divide="ignore"):
        dI = np.where(intensity > 0, -(np.log(intensity / S) + loss) / S, 0.0)
    gX = np.zeros(x.shape, dtype=np.complex128)
    gX[..., : n // 2] = 2.0 * dI * X
    gs = np.fft.fft(gX, axis=-1)  # sum_k gX[k] exp(-2 pi i jk/N)
    dphi = -np.imag(np.conj(gs) * k["s"])
    dy = np.real(np.conj(gs) * k["w"] * k["e"])
    dr = dy * k["dydt"] * k["inside"]
    d, u = k["d"], k["u"]
    return np.array([
        (dphi * d ** 2).sum(), (dphi * d ** 3).sum(),
        dr.sum(), (dr * u).sum(), (dr * u ** 2).sum(), (dr * u ** 3).sum(),
    ])


def finite_difference(x: np.ndarray, theta, window: np.ndarray, method: int, rel: float = 1e-3,
                      scale=None) -> np.ndarray:
    """Central differences with step rel * max(|theta_i|, scale_i)."""
    theta = np.asarray(theta, dtype=np.float64)
    scale = np.zeros(6) if scale is None else np.asarray(scale, dtype=np.float64)
    g = np.zeros(6)
    for i in range(6):
        h = rel * max(abs(theta[i]), scale[i])
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        g[i] = (forward(x, tp, window, method)[0] - forward(x, tm, window, method)[0]) / (2 * h)
    return g
