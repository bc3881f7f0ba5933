"""The real-pair split used by the chain kernel without dispersion, checked
on the CPU with the kernel's own lane/element indexing.

Lane t (of TPA = N/8) holds Z[t + e*TPA], e = 0..7, after the FFT of
z = y0 + i*y1.  For the kept bins k = t + o*TPA (o < 4) the kernel reads its
partner Z[N - k] at lane TPA - t, element 7 - o (lane 0: its own element
(8 - o) mod 8), and forms 2*X0[k] = Z[k] + conj Z[N-k] and
2*X1[k] = -i (Z[k] - conj Z[N-k]).  A wrong partner index (off by one
lane or element) must fail.
"""

from __future__ import annotations

import numpy as np
import pytest


def _split(Z, n, partner):
    tpa = n // 8
    lanes = Z.reshape(8, tpa).T  # lanes[t, e] = Z[t + e*TPA]
    x0 = np.zeros(n // 2, complex)
    x1 = np.zeros(n // 2, complex)
    for t in range(tpa):
        for o in range(4):
            pt, pe = partner(t, o, tpa)
            z, p = lanes[t, o], lanes[pt, pe]
            k = t + o * tpa
            x0[k] = (z + np.conj(p)) / 2
            x1[k] = -1j * (z - np.conj(p)) / 2
    return x0, x1


def kernel_partner(t, o, tpa):
    return (tpa - t, 7 - o) if t > 0 else (0, (8 - o) & 7)


@pytest.mark.parametrize("n", [256, 512, 1024, 2048, 4096])
def test_pair_split_equals_two_transforms(n):
    rng = np.random.default_rng(n)
    y0, y1 = rng.normal(size=n), rng.normal(size=n)
    Z = np.fft.ifft(y0 + 1j * y1) * n  # unnormalised inverse, as the chain's
    x0, x1 = _split(Z, n, kernel_partner)
    r0 = np.fft.ifft(y0)[: n // 2] * n
    r1 = np.fft.ifft(y1)[: n // 2] * n
    scale = np.abs(Z).max()
    assert np.max(np.abs(x0 - r0)) < 1e-12 * scale
    assert np.max(np.abs(x1 - r1)) < 1e-12 * scale


@pytest.mark.parametrize("wrong", [
    lambda t, o, tpa: ((tpa - t + 1) % tpa, 7 - o) if t > 0 else (0, (8 - o) & 7),
    lambda t, o, tpa: (tpa - t, (8 - o) & 7) if t > 0 else (0, (8 - o) & 7),
    lambda t, o, tpa: (tpa - t, 7 - o) if t > 0 else (0, 7 - o),
])
def test_planted_wrong_partner_is_caught(wrong):
    n = 1024
    rng = np.random.default_rng(1)
    y0, y1 = rng.normal(size=n), rng.normal(size=n)
    Z = np.fft.ifft(y0 + 1j * y1) * n
    x0, _ = _split(Z, n, wrong)
    r0 = np.fft.ifft(y0)[: n // 2] * n
    assert np.max(np.abs(x0 - r0)) > 1e-3 * np.abs(Z).max()
