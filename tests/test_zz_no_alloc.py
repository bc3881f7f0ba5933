"""ope_hip_op_min_variance_mean_device: getMinimumVarianceMean on device
memory with its segment sums in the library's static scratch array (it
allocates nothing).  Checked against the allocating host-pointer entry point
on two calls in a row, the second smaller, and against the scratch limit."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ctx():
    from octproengine_amd import _capi
    return _capi.Context(_capi.make_params(signal_length=1024, ascans_per_bscan=64), 0, 1)


def _rows(rng, height, width):
    return (rng.normal(0, 50, (height, width)) + 1j * rng.normal(0, 50, (height, width))
            + (rng.normal(0, 5, width) + 1j * rng.normal(0, 5, width))).astype(np.complex64)


def test_device_form_matches_the_allocating_form():
    import torch
    from octproengine_amd import _capi
    rng = np.random.default_rng(61)
    with _ctx() as ctx:
        stream = torch.cuda.current_stream().cuda_stream
        for height, width, segments in ((1024, 2048, 8), (96, 300, 8), (64, 7, 16)):
            rows = _rows(rng, height, width)
            want = ctx.min_variance_mean(rows, segments)
            inter = np.stack([rows.real, rows.imag], -1).astype(np.float32)
            d_in = torch.from_numpy(inter).cuda()
            d_out = torch.full((2 * width,), float("nan"), dtype=torch.float32, device="cuda")
            print(f"[no-alloc] height {height} width {width} segments {segments}", flush=True)
            ctx.min_variance_mean_device(d_in.data_ptr(), width, height, segments, d_out.data_ptr(), stream)
            torch.cuda.synchronize()
            got = d_out.cpu().numpy()
            assert np.array_equal(got, want), np.max(np.abs(got - want))
        # beyond the scratch: 2 MiB of FP64 triples holds segments * width <= 87381
        d_in = torch.zeros((16, 8192, 2), dtype=torch.float32, device="cuda")
        d_out = torch.zeros((2 * 8192,), dtype=torch.float32, device="cuda")
        with pytest.raises(_capi.HipError) as e:
            ctx.min_variance_mean_device(d_in.data_ptr(), 8192, 16, 16, d_out.data_ptr(), stream)
        assert e.value.code == _capi.EINVAL
        ctx.min_variance_mean_device(d_in.data_ptr(), 8192, 16, 8, d_out.data_ptr(), stream)  # 65536: fits
        torch.cuda.synchronize()
        assert not d_out.any()
