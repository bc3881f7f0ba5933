"""Ragged batches (ope_hip_ragged_*): requests of mixed signal length and
A-scan count in one call, one persistent launch per signal length.  Each
request's output must equal, bit for bit, what the same context gives for
it through ope_hip_process_device (the same chain arithmetic; the launch
only reads and writes its A-scans through a row map), and agree with the
FP64 oracle within the parity bar."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import chain64 as O
from octproengine_amd import synthetic as S
from tests.helpers import DB_TOL, GpuChain, chain_params, db_error, db_params, standard_tables, to_device

pytestmark = pytest.mark.gpu

LENGTHS = (256, 512, 1024, 2048, 4096)


def make_chains(lengths, interp=O.CUBIC, dtype=np.uint16, dc=False):
    """interp None: no resampling."""
    chains = {}
    for n in lengths:
        p = db_params(O.ChainParams(signal_length=n, ascans_per_bscan=64, bscans_per_buffer=1,
                                    resampling=interp is not None, interpolation=interp or 0, windowing=True,
                                    dispersion=True,
                                    dc_removal=dc, dc_window=32))
        t = standard_tables(n)
        # untuned contexts: their frames take the default launch, the one the
        # batch's row-map kernels are built for
        g = GpuChain(p, t.curve, t.window, t.phasor, dtype=dtype)
        # a fixed sample offset, so that the batch (built before any frame)
        # and the per-request frames use the same one
        from octproengine_amd import _capi
        g.ctx.set_sample_offset(_capi.OFFSET_FIXED, 20000.0 if dtype != np.float32 else 0.3)
        chains[n] = (g, p, t)
    return chains


def make_requests(rng, lengths, count, max_ascans, dtype=np.uint16):
    import torch
    reqs = []
    for i in range(count):
        n = int(rng.choice(lengths))
        a = int(rng.integers(0, max_ascans + 1)) if i % 7 else 0  # some empty requests
        raw = S.fringes(n, max(a, 1), seed=1000 + i)[:a]
        if dtype == np.float32:
            raw = (raw.astype(np.float32) / 65535.0).astype(np.float32)
        d_in = to_device(raw.reshape(a, n)) if a else torch.empty((1, n), dtype=torch.int16, device="cuda")
        d_out = torch.full((max(a, 1), n // 2), float("nan"), dtype=torch.float32, device="cuda")
        reqs.append(dict(n=n, a=a, raw=raw, d_in=d_in, d_out=d_out))
    return reqs


def run_batch(chains, reqs, chunk=0):
    import torch
    from octproengine_amd import _capi
    batch = _capi.RaggedBatch([c[0].ctx for c in chains.values()],
                              [(r["d_in"].data_ptr(), r["d_out"].data_ptr(), r["a"], r["n"]) for r in reqs],
                              chunk_ascans=chunk)
    batch.run(torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return batch


# (uint16, DC, no resampling): the banded DC path at N = 1024 and 2048
@pytest.mark.parametrize("dtype,dc,interp", [(np.uint16, False, O.CUBIC), (np.float32, True, O.CUBIC),
                                             (np.uint16, True, None)])
def test_ragged_matches_per_request_processing(dtype, dc, interp):
    rng = np.random.default_rng(5)
    chains = make_chains(LENGTHS, interp=interp, dtype=dtype, dc=dc)
    reqs = make_requests(rng, LENGTHS, 24, 90, dtype)
    print(f"[ragged] {len(reqs)} requests, {sum(r['a'] for r in reqs)} A-scans, dtype {np.dtype(dtype)}, dc={dc}, "
          f"interpolation={interp}", flush=True)
    batch = run_batch(chains, reqs)
    worst_exact, worst_oracle = 0.0, 0.0
    for i, r in enumerate(reqs):
        if r["a"] == 0:
            continue
        got = r["d_out"].cpu().numpy()
        assert np.isfinite(got).all(), f"request {i}: outputs not all written"
        chain, p, t = chains[r["n"]]
        ref = chain.frame(r["raw"])
        worst_exact = max(worst_exact, float(np.max(np.abs(got - ref))))
        if i % 3 == 0:
            o = O.process_frame(r["raw"], p, t.curve, t.window, t.phasor)
            worst_oracle = max(worst_oracle, db_error(got, o))
    print(f"[ragged] max |ragged - per-request| = {worst_exact:.3e} dB, vs oracle {worst_oracle:.2e} dB")
    assert worst_exact == 0.0
    assert worst_oracle <= DB_TOL
    lat = batch.latencies_us()
    nonempty = np.array([r["a"] > 0 for r in reqs])
    assert (lat[nonempty] > 0).all() and (lat[~nonempty] == 0).all()
    batch.close()
    for c in chains.values():
        c[0].close()


def test_ragged_runs_again_without_reset():
    """The counters reset themselves: a second and third run redo the whole
    batch with identical output."""
    import torch
    rng = np.random.default_rng(9)
    chains = make_chains((1024, 2048, 4096))
    reqs = make_requests(rng, (1024, 2048, 4096), 12, 200)
    batch = run_batch(chains, reqs)
    first = [r["d_out"].clone() for r in reqs]
    for _ in range(2):
        for r in reqs:
            r["d_out"].fill_(float("nan"))
        batch.run(torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        for r, f in zip(reqs, first):
            if r["a"]:
                assert torch.equal(r["d_out"], f)
    lat = batch.latencies_us()
    assert lat.max() > 0
    batch.close()
    for c in chains.values():
        c[0].close()


def test_ragged_uses_fpn_profile_and_rejects_bad_batches():
    import torch
    from octproengine_amd import _capi
    chains = make_chains((1024, 2048))
    # give the 2048 context an FPN profile through a normal frame
    c2048, p, t = chains[2048]
    pf = db_params(O.ChainParams(signal_length=2048, ascans_per_bscan=64, bscans_per_buffer=1, resampling=True,
                                 interpolation=O.CUBIC, windowing=True, dispersion=True, fpn_removal=True))
    c2048.ctx.set_params(chain_params(pf))
    c2048.ctx.request_fpn()
    warm = S.fringes(2048, 64, seed=3)
    c2048.frame(warm)
    prof = c2048.ctx.get_fpn_profile()
    assert np.any(prof != 0)
    raw = S.fringes(2048, 50, seed=4)
    ref = c2048.frame(raw)  # uses the profile
    d_in = to_device(raw)
    d_out = torch.empty((50, 1024), dtype=torch.float32, device="cuda")
    b = _capi.RaggedBatch([c[0].ctx for c in chains.values()], [(d_in.data_ptr(), d_out.data_ptr(), 50, 2048)])
    b.run(torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert float(np.max(np.abs(d_out.cpu().numpy() - ref))) == 0.0
    b.close()
    # a length without context
    with pytest.raises(_capi.HipError) as e:
        _capi.RaggedBatch([c[0].ctx for c in chains.values()], [(d_in.data_ptr(), d_out.data_ptr(), 1, 512)])
    assert e.value.code == _capi.EINVAL
    # contexts that disagree on interpolation
    lin = make_chains((512,), interp=O.LINEAR)
    with pytest.raises(_capi.HipError) as e:
        _capi.RaggedBatch([chains[1024][0].ctx, lin[512][0].ctx], [])
    assert e.value.code == _capi.EINVAL
    # a pending FPN determination
    c2048.ctx.request_fpn()
    with pytest.raises(_capi.HipError) as e:
        _capi.RaggedBatch([c2048.ctx], [])
    assert e.value.code == _capi.ESTATE
    # chunk_ascans is unused and must not be negative
    with pytest.raises(_capi.HipError) as e:
        _capi.RaggedBatch([chains[1024][0].ctx], [(d_in.data_ptr(), d_out.data_ptr(), 1, 1024)], chunk_ascans=-1)
    assert e.value.code == _capi.EINVAL
    # misaligned input pointer
    with pytest.raises(_capi.HipError) as e:
        _capi.RaggedBatch([chains[1024][0].ctx], [(d_in.data_ptr() + 2, d_out.data_ptr(), 1, 1024)])
    assert e.value.code == _capi.EINVAL
    for c in list(chains.values()) + list(lin.values()):
        c[0].close()


def test_ragged_empty_batches_and_graph_replay():
    """A batch without requests, or whose requests are all empty, launches
    nothing; a batch of several lengths captured into a HIP graph (one
    launch per length) replays bit-identically to its eager run, with eager
    frames of one of its contexts between the replays."""
    import torch
    from octproengine_amd import _capi
    rng = np.random.default_rng(13)
    chains = make_chains((512, 1024, 4096))
    stream = torch.cuda.current_stream().cuda_stream
    empty = _capi.RaggedBatch([c[0].ctx for c in chains.values()], [])
    empty.run(stream)
    zero_in = torch.empty((1, 1024), dtype=torch.int16, device="cuda")
    zero_out = torch.full((1, 512), float("nan"), device="cuda")
    zeros = _capi.RaggedBatch([c[0].ctx for c in chains.values()],
                              [(zero_in.data_ptr(), zero_out.data_ptr(), 0, 1024)] * 3)
    zeros.run(stream)
    torch.cuda.synchronize()
    assert torch.isnan(zero_out).all()
    assert (zeros.latencies_us() == 0).all()
    empty.close()
    zeros.close()

    reqs = make_requests(rng, (512, 1024, 4096), 9, 120)
    batch = run_batch(chains, reqs)
    eager = [r["d_out"].clone() for r in reqs]
    side = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(side):
        for r in reqs:
            r["d_out"].fill_(float("nan"))
        g.capture_begin()
        batch.run(side.cuda_stream)
        g.capture_end()
    torch.cuda.synchronize()
    c1024 = chains[1024][0]
    warm = S.fringes(1024, 40, seed=77)
    for rep in range(3):
        for r in reqs:
            r["d_out"].fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        c1024.frame(warm)  # eager work of one of the batch's contexts in between
        for i, (r, e) in enumerate(zip(reqs, eager)):
            if r["a"]:
                assert torch.equal(r["d_out"], e), f"replay {rep}, request {i}"
    batch.close()
    for c in chains.values():
        c[0].close()


# The second-largest length runs on the dynamic chunk schedule shared by two
# launches: lengths of up to 64 lanes per A-scan (several A-scans per
# workgroup iteration), real pairs (no dispersion, two A-scans per lane) and
# DC removal take it as well as the 1024/2048-point single FFTs.
@pytest.mark.parametrize("lengths,disp,dc", [((512, 256), True, False), ((4096, 1024), False, False),
                                             ((2048, 512), True, True), ((4096, 2048, 256), False, True)])
def test_ragged_dynamic_schedule_shapes(lengths, disp, dc):
    import torch
    from octproengine_amd import _capi
    rng = np.random.default_rng(31)
    chains = {}
    for n in lengths:
        p = db_params(O.ChainParams(signal_length=n, ascans_per_bscan=64, bscans_per_buffer=1, resampling=True,
                                    interpolation=O.CUBIC, windowing=True, dispersion=disp, dc_removal=dc,
                                    dc_window=32))
        t = standard_tables(n)
        g = GpuChain(p, t.curve, t.window, t.phasor)
        g.ctx.set_sample_offset(_capi.OFFSET_FIXED, 20000.0)
        chains[n] = (g, p, t)
    reqs = make_requests(rng, lengths, 30, 700)
    print(f"[ragged] lengths {lengths} disp={disp} dc={dc}: {sum(r['a'] for r in reqs)} A-scans", flush=True)
    batch = run_batch(chains, reqs)
    for rep in range(2):  # the schedule's counters clear themselves between runs
        if rep:
            for r in reqs:
                r["d_out"].fill_(float("nan"))
            batch.run(torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
        for i, r in enumerate(reqs):
            if r["a"] == 0:
                continue
            got = r["d_out"].cpu().numpy()
            chain, p, t = chains[r["n"]]
            assert np.array_equal(got, chain.frame(r["raw"])), (rep, i, r["n"], r["a"])
            if rep == 0 and i % 5 == 0:
                o = O.process_frame(r["raw"], p, t.curve, t.window, t.phasor)
                assert db_error(got, o) <= DB_TOL
    lat = batch.latencies_us()
    nonempty = np.array([r["a"] > 0 for r in reqs])
    assert (lat[nonempty] > 0).all() and (lat[~nonempty] == 0).all()
    batch.close()
    for c in chains.values():
        c[0].close()
