"""Seeded synthetic OCT fringes and the matching resampling/dispersion curves.

Fringe model (per A-scan, N samples):
    raw[n] = clip(rint(DC + sum_r a_r cos(2 pi z_r k(n) / N + phi(k(n))) + eps), 0, 65535)
    k(n)   = n + c * (n - N/2)^2            (non-linear k sampling, c = 3e-5)
with per-A-scan reflector depths z_r ~ U(50, 900) * N/2048, amplitudes
a_r ~ U(500, 3000), Gaussian noise eps ~ N(0, sigma^2), DC = 20000.  After
resampling at r[j] = k^-1(j) the fringe of reflector r is a pure tone at
bin z_r.  ``phi`` is an optional dispersion phase,
phi(k) = d2 (k-N/2)^2 + d3 (k-N/2)^3: the chain's dispersion step with the
same coefficients (phase curve ``dispersion_phase(N, d2, d3)``) removes it.
(The chain keeps bins k < N/2 of an inverse FFT, i.e. the exp(-i arg)
half of each cosine, which the compensation multiplies by exp(+i phi).)

numpy (``fringes``) is used for parity-sized data; ``fringes_torch`` makes
the same kind of data directly in GPU memory for throughput runs.
"""

from __future__ import annotations

import math

import numpy as np

K_CURV = 3e-5
DC = 20000.0


def k_of_n(n: int, c: float = K_CURV) -> np.ndarray:
    i = np.arange(n, dtype=np.float64)
    return i + c * (i - n / 2) ** 2


def resampling_curve(n: int, c: float = K_CURV, clamp: bool = True) -> np.ndarray:
    """r[j] = k^-1(j) (float32), clamped to [7, N-9] like generated curves."""
    j = np.arange(n, dtype=np.float64)
    b = 1.0 - c * n
    disc = b * b - 4.0 * c * (c * n * n / 4.0 - j)
    r = (-b + np.sqrt(np.maximum(disc, 0.0))) / (2.0 * c)
    if clamp:
        r = np.clip(r, 7.0, n - 9.0)
    return r.astype(np.float32)


def dispersion_phase(n: int, d2: float, d3: float, k0: float | None = None) -> np.ndarray:
    """phi[j] = d2 (j-k0)^2 + d3 (j-k0)^3 as float32 (k0 = N/2 by default)."""
    if k0 is None:
        k0 = n / 2
    j = np.arange(n, dtype=np.float64) - k0
    return (d2 * j * j + d3 * j * j * j).astype(np.float32)


def fringes(n: int, n_ascans: int, seed: int = 0, reflectors: int = 3, sigma: float = 30.0,
            dtype=np.uint16, d2: float = 0.0, d3: float = 0.0) -> np.ndarray:
    """[n_ascans, n] raw fringes.  float32 output is scaled to [0, 1]."""
    rng = np.random.default_rng(seed)
    kk = k_of_n(n)
    z = rng.uniform(50.0, 900.0, size=(n_ascans, reflectors)) * (n / 2048.0)
    a = rng.uniform(500.0, 3000.0, size=(n_ascans, reflectors))
    phi = 0.0
    if d2 or d3:
        # dispersion acts on the linear-k axis, so it is evaluated at k(n)
        kc = kk - n / 2
        phi = d2 * kc * kc + d3 * kc * kc * kc
    out = np.empty((n_ascans, n), dtype=np.float64)
    chunk = max(1, (1 << 22) // n)
    for s in range(0, n_ascans, chunk):
        e = min(n_ascans, s + chunk)
        arg = 2.0 * math.pi * z[s:e, :, None] * kk[None, None, :] / n + phi
        out[s:e] = DC + np.einsum("ar,arn->an", a[s:e], np.cos(arg))
    out += rng.normal(0.0, sigma, size=out.shape)
    out = np.clip(np.rint(out), 0, 65535)
    if np.dtype(dtype) == np.float32:
        return (out / 65535.0).astype(np.float32)
    return out.astype(dtype)


def fringes_torch(n: int, n_ascans: int, seed: int = 0, reflectors: int = 3, sigma: float = 30.0,
                  device: str = "cuda", dtype=None, chunk: int = 4096, d2: float = 0.0, d3: float = 0.0):
    """Same model as ``fringes`` generated on the GPU with torch (uint16 as
    int16 bit patterns in a torch.int16 tensor, or float32 in [0, 1])."""
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    kk64 = k_of_n(n)
    kc = kk64 - n / 2
    kk = torch.as_tensor(kk64, device=device, dtype=torch.float32)
    phi = torch.as_tensor(d2 * kc * kc + d3 * kc * kc * kc, device=device, dtype=torch.float32)
    want_f32 = dtype == torch.float32
    out = torch.empty((n_ascans, n), device=device,
                      dtype=torch.float32 if want_f32 else torch.int16)
    for s in range(0, n_ascans, chunk):
        e = min(n_ascans, s + chunk)
        m = e - s
        z = torch.rand((m, reflectors, 1), generator=g, device=device) * 850.0 + 50.0
        z = z * (n / 2048.0)
        a = torch.rand((m, reflectors, 1), generator=g, device=device) * 2500.0 + 500.0
        acc = torch.full((m, n), DC, device=device, dtype=torch.float32)
        for r in range(reflectors):
            acc += (a[:, r] * torch.cos((2.0 * math.pi / n) * z[:, r] * kk[None, :] + phi[None, :]))
        acc += torch.randn((m, n), generator=g, device=device) * sigma
        acc = torch.clamp(torch.round(acc), 0, 65535)
        if want_f32:
            out[s:e] = acc / 65535.0
        else:
            out[s:e] = (acc.to(torch.int32) - ((acc >= 32768).to(torch.int32) << 16)).to(torch.int16)
    return out
