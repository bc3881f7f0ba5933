"""Graph-captured frames (ope_hip_graph_*): replays of a captured ring of
frames give exactly what process_device gives, with eager work and
synchronisation between the replays; capture is refused outside steady
state."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import chain64 as O
from octproengine_amd import synthetic as S
from tests.helpers import GpuChain, chain_params, db_params, standard_tables, to_device

pytestmark = pytest.mark.gpu


def test_graph_replays_match_eager_frames():
    import torch
    from octproengine_amd import _capi
    n, a, frames = 2048, 1024, 4
    p = db_params(O.ChainParams(signal_length=n, ascans_per_bscan=a, bscans_per_buffer=1, resampling=True,
                                interpolation=O.CUBIC, windowing=True, dispersion=True, fpn_removal=True))
    t = standard_tables(n)
    chain = GpuChain(p, t.curve, t.window, t.phasor)
    chain.ctx.request_fpn()
    chain.frame(S.fringes(n, a, seed=1))  # FPN determined
    chain.frame(S.fringes(n, a, seed=2))  # statistics of the next determination accumulated
    raws = [S.fringes(n, a, seed=10 + f) for f in range(frames)]
    eager = [chain.frame(r) for r in raws]
    d_in = [to_device(r) for r in raws]
    d_out = [torch.full((a, n // 2), float("nan"), dtype=torch.float32, device="cuda") for _ in range(frames)]
    g = _capi.FrameGraph(chain.ctx, [x.data_ptr() for x in d_in], [y.data_ptr() for y in d_out], a)
    stream = torch.cuda.current_stream().cuda_stream
    for rep in range(3):
        for y in d_out:
            y.fill_(float("nan"))
        g.launch(stream)
        torch.cuda.synchronize()
        for y, e in zip(d_out, eager):
            assert np.array_equal(y.cpu().numpy(), e), f"replay {rep}"
        chain.frame(raws[0])  # eager work between replays
    g.close()
    # not steady state: a pending FPN request
    chain.ctx.request_fpn()
    with pytest.raises(_capi.HipError) as e:
        _capi.FrameGraph(chain.ctx, [d_in[0].data_ptr()], [d_out[0].data_ptr()], a)
    assert e.value.code == _capi.ESTATE
    chain.close()


def test_graph_refused_while_fpn_statistics_accumulate():
    from octproengine_amd import _capi
    import torch
    n, a = 1024, 256
    p = db_params(O.ChainParams(signal_length=n, ascans_per_bscan=a, bscans_per_buffer=1, resampling=True,
                                interpolation=O.LINEAR, windowing=True, dispersion=False))
    t = standard_tables(n)
    chain = GpuChain(p, t.curve, t.window, t.phasor)
    x = to_device(S.fringes(n, a, seed=1))
    y = torch.empty((a, n // 2), dtype=torch.float32, device="cuda")
    with pytest.raises(_capi.HipError) as e:
        _capi.FrameGraph(chain.ctx, [x.data_ptr()], [y.data_ptr()], a)
    assert e.value.code == _capi.ESTATE
    chain.frame(S.fringes(n, a, seed=2))
    g = _capi.FrameGraph(chain.ctx, [x.data_ptr()], [y.data_ptr()], a)
    g.launch(torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert np.isfinite(y.cpu().numpy()).all()
    g.close()
    chain.close()
