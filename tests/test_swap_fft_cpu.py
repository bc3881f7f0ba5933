"""The index algebra of the chain's swap-schedule FFT (N = 1024, 2048), on
the CPU with the kernel's own lane fields, LDS address constants and
lane-half swap semantics (chain_kernels.hpp, SwapFft / fft_swap):

* the result equals the unnormalised inverse DFT and lands in the layout
  v[t][e] = X[t + e*N/8] that the epilogue reads;
* both LDS exchanges are one-to-one and fit the padded buffer (PADN slots);
* every 32-lane half of every 8-byte LDS access hits 32 distinct bank
  pairs (64 banks of 4 bytes);
* planted mistakes (a lane-half swap with its outputs exchanged, a wrong
  lane field, a wrong address constant) fail one of these checks.
"""

from __future__ import annotations

import numpy as np
import pytest

PADN = {1024: 1146, 2048: 2298}  # Geo<N>::PADN


def idft(v):
    r = len(v)
    m = np.arange(r)
    return np.exp(2j * np.pi * np.outer(m, m) / r) @ v


def permlane32(a, b):
    """v_permlane32_swap: a keeps lanes 0-31 and takes b's 0-31 as its 32-63;
    b takes a's 32-63 as its 0-31 and keeps its own 32-63."""
    return np.concatenate([a[:32], b[:32]]), np.concatenate([a[32:], b[32:]])


def permlane16(a, b):
    """v_permlane16_swap: odd rows (16 lanes) of a trade with even rows of b."""
    r = lambda x, i: x[16 * i:16 * i + 16]
    return (np.concatenate([r(a, 0), r(b, 0), r(a, 2), r(b, 2)]),
            np.concatenate([r(a, 1), r(b, 1), r(a, 3), r(b, 3)]))


class Swap:
    """SwapFft<N> restated."""

    def __init__(self, n, **plant):
        self.n = n
        self.tpa = n // 8
        self.s = n // 512
        self.rq = 8 // self.s
        self.e3 = 264 if n == 2048 else 136
        self.plant = plant

    def r2_of(self, t):
        if self.plant.get("r2"):
            return (t >> 3) & 7
        return ((t >> 3) & 1) | ((t >> 5) & 6) if self.n == 2048 else ((t >> 3) & 3) | ((t >> 4) & 4)

    def d2_of(self, t):
        return (t >> 4) & 3 if self.n == 2048 else (t >> 5) & 1

    def k1hi_of(self, t):
        return (t >> 3) & 6 if self.n == 2048 else (t >> 3) & 4

    def skew(self, k0):
        a, b, c = (34, 68, 144) if self.n == 2048 else (36, 72, 144)
        if self.plant.get("skew"):
            a = 32
        return a * (k0 & 1) + b * ((k0 >> 1) & 1) + c * (k0 >> 2)

    def c_d1(self, d1):
        return 288 * d1 if self.n == 2048 else 16 * (d1 & 1) + 284 * (d1 >> 1)

    def e1_write_base(self, t):
        return (t & 31) + (288 if self.n == 2048 else 284) * (t >> 5)

    def e1_read_base(self, t):
        return self.r2_of(t) + 8 * self.d2_of(t) + self.skew(t & 7)

    def e3_write_base(self, t):
        return (t & 7) + 8 * self.k1hi_of(t) + self.e3 * self.r2_of(t)

    def e3_write_off(self, e):
        return 8 * (e % self.rq) + 64 * (e // self.rq)


def run(n, x, **plant):
    """fft_swap<N, 1> on one sequence; also returns the LDS accesses."""
    F = Swap(n, **plant)
    tpa, w = F.tpa, (lambda m, p: np.exp(2j * np.pi * p / m))
    t = np.arange(tpa)
    v = np.array([[x[i + tpa * e] for e in range(8)] for i in range(tpa)], complex)
    acc = {"E1 write": [], "E1 read": [], "E3 write": [], "E3 read": []}
    # pass 0 and exchange 1
    v = np.array([idft(row) for row in v])
    v[:, 1:] *= np.array([[w(n, i * e) for e in range(1, 8)] for i in range(tpa)])
    lds = np.full(PADN[n], np.nan + 0j)
    for e in range(8):
        addr = np.array([F.e1_write_base(i) for i in t]) + F.skew(e)
        acc["E1 write"].append(addr)
        lds[addr] = v[:, e]
    for e in range(8):
        addr = np.array([F.e1_read_base(i) for i in t]) + F.c_d1(e)
        acc["E1 read"].append(addr)
        v[:, e] = lds[addr]
    # pass 1
    r1 = np.array([8 * F.d2_of(i) + F.r2_of(i) for i in t])
    v = np.array([idft(row) for row in v])
    v[:, 1:] *= np.array([[w(n, 8 * r1[i] * e) for e in range(1, 8)] for i in t])
    # exchange 2 in each wavefront
    for wave in range(tpa // 64):
        ln = slice(64 * wave, 64 * wave + 64)
        for e in range(4):
            a, b = permlane32(v[ln, e], v[ln, e + 4])
            if plant.get("swap_order"):
                a, b = b, a
            v[ln, e], v[ln, e + 4] = a, b
        if F.s == 4:
            for e in (0, 1, 4, 5):
                v[ln, e], v[ln, e + 2] = permlane16(v[ln, e], v[ln, e + 2])
    # pass 2
    for i in t:
        for q in range(F.rq):
            es = [q + F.rq * m for m in range(F.s)]
            v[i, es] = idft(v[i, es])
    r2 = np.array([F.r2_of(i) for i in t])
    for e in range(F.rq, 8):
        v[:, e] *= np.array([w(n, 64 * r2[i] * (e // F.rq)) for i in t])
    # exchange 3 and pass 3
    lds = np.full(PADN[n], np.nan + 0j)
    for e in range(8):
        addr = np.array([F.e3_write_base(i) for i in t]) + F.e3_write_off(e)
        acc["E3 write"].append(addr)
        lds[addr] = v[:, e]
    for e in range(8):
        addr = t + F.e3 * e
        acc["E3 read"].append(addr)
        v[:, e] = lds[addr]
    v = np.array([idft(row) for row in v])
    X = np.zeros(n, complex)
    for e in range(8):
        X[t + tpa * e] = v[:, e]
    return X, acc


def conflict_cycles(addr):
    """Cycles of one 8-byte access of 64 lanes: two groups of 32 lanes, a
    group costs the largest number of distinct dwords in one bank."""
    tot = 0
    for g in (addr[:32], addr[32:]):
        banks = {}
        for a in g:
            for k in range(2):
                d = 2 * int(a) + k
                banks.setdefault(d % 64, set()).add(d)
        tot += max(len(s) for s in banks.values())
    return tot


@pytest.mark.parametrize("n", [1024, 2048])
def test_swap_schedule_is_the_inverse_dft(n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    X, acc = run(n, x)
    ref = np.fft.ifft(x) * n
    assert np.max(np.abs(X - ref)) < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("n", [1024, 2048])
def test_lds_exchanges_one_to_one_in_bounds_conflict_free(n):
    rng = np.random.default_rng(1)
    _, acc = run(n, rng.normal(size=n) + 0j)
    for name, per_e in acc.items():
        allad = np.concatenate(per_e)
        assert allad.min() >= 0 and allad.max() < PADN[n], name
        assert len(np.unique(allad)) == n, name
        for addr in per_e:
            for wave in range(len(addr) // 64):
                assert conflict_cycles(addr[64 * wave:64 * wave + 64]) == 2, name


@pytest.mark.parametrize("plant", ["swap_order", "r2"])
def test_planted_index_mistakes_change_the_result(plant):
    n = 2048
    rng = np.random.default_rng(2)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    try:
        X, _ = run(n, x, **{plant: True})
    except (IndexError, ValueError):
        return
    ref = np.fft.ifft(x) * n
    assert not np.max(np.abs(X - ref)) < 1e-6 * np.abs(ref).max()


def test_planted_address_constant_conflicts():
    _, acc = run(2048, np.ones(2048, complex), skew=True)
    worst = max(conflict_cycles(a[64 * w:64 * w + 64]) for a in acc["E1 read"] for w in range(4))
    assert worst > 2



# The Lanczos-4 kernels read the twiddles of passes 1 and 2 from a 504-entry
# LDS table (chain_kernels.hpp twl_swap_exponent, fft_swap_small tw1/tw2).
def twl_swap_exponent(n, i, rq):
    if i < 448:
        return (8 * (i % 64) * (1 + i // 64)) & (n - 1)
    j = i - 448
    return (64 * (j % 8) * ((rq + j // 8) // rq)) & (n - 1)


def twl_swap_reads(F, t, plant=None):
    """(entry index, register exponent) of every table read of lane t."""
    r1 = 8 * F.d2_of(t) + F.r2_of(t)
    out = []
    for e in range(1, 8):
        idx = (e - 1) * 64 + (F.r2_of(t) if plant == "r1" else r1)
        out.append((idx, (8 * r1 * e) & (F.n - 1)))
    for e in range(F.rq, 8):
        idx = 448 + (e - F.rq) * 8 + F.r2_of(t)
        out.append((idx, (64 * F.r2_of(t) * (e // F.rq)) & (F.n - 1)))
    return out


@pytest.mark.parametrize("n", [1024, 2048])
def test_lds_twiddle_table_holds_the_register_twiddles(n):
    F = Swap(n)
    for t in range(F.tpa):
        for idx, expo in twl_swap_reads(F, t):
            assert 0 <= idx < 7 * 8 + 7 * 64
            assert twl_swap_exponent(n, idx, F.rq) == expo, (t, idx)
    # within a wavefront the reads of one entry row are broadcasts: at most
    # 8 distinct entries per read instruction
    for wave in range(F.tpa // 64):
        for k in range(7 + 8 - F.rq):
            idxs = {twl_swap_reads(F, t)[k][0] for t in range(64 * wave, 64 * wave + 64)}
            assert len(idxs) <= 8


def test_planted_lds_twiddle_row_mistake_is_caught():
    F = Swap(2048)
    bad = sum(twl_swap_exponent(2048, idx, F.rq) != expo
              for t in range(F.tpa) for idx, expo in twl_swap_reads(F, t, plant="r1"))
    assert bad > 0


# ------------------------------------------------------------------ N = 4096
#
# Radices 8, 8, 8, 8 with n = 512 d0 + 64 d1 + 8 d2 + r2 and
# k = k0 + 8 k1 + 64 k2 + 512 k3 (chain_kernels.hpp, Swap4096 / fft_swap4096):
# exchange 1 through LDS at (t & 63) + 64 k0 + 512 (t >> 6), read at
# t + 512 d1; exchange 2 inside each wavefront, register bits 2, 1, 0 with
# lane bits 5, 4, 3 (lane-half swaps and a row rotation by 8 with a select);
# exchange 3 through LDS at k0 + 8 k1 + 64 k2 + 513 r2, read at t + 513 r2.

PADN4096 = 4381  # Geo<4096>::PADN


def row_ror8(a):
    """DPP row_ror:8 on 64 lanes: lane i of each 16-lane row reads lane
    (i + 8) mod 16 of that row."""
    out = np.empty_like(a)
    for r in range(4):
        row = a[16 * r:16 * r + 16]
        out[16 * r:16 * r + 16] = np.roll(row, -8)
    return out


def swap_bit3(a, b, lanes, plant=False):
    """Register bit 0 (a, b) trades places with lane bit 3: a' = hi ? ror8(b)
    : a, b' = hi ? b : ror8(a), hi = lane bit 3."""
    hi = (lanes & 8) != 0
    if plant:
        hi = ~hi
    pa, pb = row_ror8(a), row_ror8(b)
    return np.where(hi, pb, a), np.where(hi, b, pa)


def run4096(x, e3=513, plant=None):
    n, tpa = 4096, 512
    w = lambda p: np.exp(2j * np.pi * p / n)  # noqa: E731
    t = np.arange(tpa)
    v = np.array([[x[i + tpa * e] for e in range(8)] for i in range(tpa)], complex)
    acc = {"E1 write": [], "E1 read": [], "E3 write": [], "E3 read": []}
    # pass 0: DFT-8 over d0, twiddle w_N^(r0 k0), r0 = t
    v = np.array([idft(row) for row in v])
    v[:, 1:] *= np.array([[w(i * e) for e in range(1, 8)] for i in t])
    lds = np.full(PADN4096, np.nan + 0j)
    for e in range(8):
        addr = (t & 63) + 64 * e + 512 * (t >> 6)
        acc["E1 write"].append(addr)
        lds[addr] = v[:, e]
    for e in range(8):
        addr = t + 512 * e
        acc["E1 read"].append(addr)
        v[:, e] = lds[addr]
    # pass 1: DFT-8 over d1, twiddle w_512^(r1 k1), r1 = t & 63
    v = np.array([idft(row) for row in v])
    r1 = (t & 63) if plant != "r1" else (t & 7)
    v[:, 1:] *= np.array([[w(8 * r1[i] * e) for e in range(1, 8)] for i in t])
    # exchange 2 in each wavefront
    for wave in range(tpa // 64):
        ln = slice(64 * wave, 64 * wave + 64)
        lanes = np.arange(64)
        for e in range(4):
            v[ln, e], v[ln, e + 4] = permlane32(v[ln, e], v[ln, e + 4])
        for e in (0, 1, 4, 5):
            v[ln, e], v[ln, e + 2] = permlane16(v[ln, e], v[ln, e + 2])
        for e in (0, 2, 4, 6):
            v[ln, e], v[ln, e + 1] = swap_bit3(v[ln, e], v[ln, e + 1], lanes, plant == "bit3")
    # pass 2: DFT-8 over d2, twiddle w_64^(r2 k2), r2 = t & 7
    v = np.array([idft(row) for row in v])
    v[:, 1:] *= np.array([[w(64 * (i & 7) * e) for e in range(1, 8)] for i in t])
    # exchange 3 and pass 3
    lds = np.full(PADN4096, np.nan + 0j)
    for e in range(8):
        addr = (t >> 6) + 8 * ((t >> 3) & 7) + 64 * e + e3 * (t & 7)
        acc["E3 write"].append(addr)
        lds[addr] = v[:, e]
    for e in range(8):
        addr = t + e3 * e
        acc["E3 read"].append(addr)
        v[:, e] = lds[addr]
    v = np.array([idft(row) for row in v])
    X = np.zeros(n, complex)
    for e in range(8):
        X[t + tpa * e] = v[:, e]
    return X, acc


def test_swap_schedule_4096_is_the_inverse_dft():
    rng = np.random.default_rng(4096)
    x = rng.normal(size=4096) + 1j * rng.normal(size=4096)
    X, _ = run4096(x)
    ref = np.fft.ifft(x) * 4096
    assert np.max(np.abs(X - ref)) < 1e-12 * np.abs(ref).max()


def test_swap_schedule_4096_exchanges_one_to_one_in_bounds_conflict_free():
    _, acc = run4096(np.ones(4096, complex))
    for name, per_e in acc.items():
        allad = np.concatenate(per_e)
        assert allad.min() >= 0 and allad.max() < PADN4096, name
        assert len(np.unique(allad)) == 4096, name
        for addr in per_e:
            for wave in range(8):
                assert conflict_cycles(addr[64 * wave:64 * wave + 64]) == 2, name


@pytest.mark.parametrize("plant", ["bit3", "r1"])
def test_planted_4096_mistakes_change_the_result(plant):
    rng = np.random.default_rng(3)
    x = rng.normal(size=4096) + 1j * rng.normal(size=4096)
    X, _ = run4096(x, plant=plant)
    ref = np.fft.ifft(x) * 4096
    assert not np.max(np.abs(X - ref)) < 1e-6 * np.abs(ref).max()


def test_planted_4096_even_exchange3_stride_conflicts():
    _, acc = run4096(np.ones(4096, complex), e3=512)
    worst = max(conflict_cycles(a[64 * w:64 * w + 64]) for a in acc["E3 write"] for w in range(8))
    assert worst > 2
