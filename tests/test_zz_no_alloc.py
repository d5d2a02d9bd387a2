"""swpc_kernel_vmax_noalloc: the allocation-free kernel__vmax.

Through the C ABI, its three maxima (written to a device buffer the test
owns) equal those of swpc_kernel_vmax bit for bit on two calls in a row, the
second on a smaller grid, so stale partials of the first call in the
library's scratch array must not leak into the second.  Both entry points
are also checked against the FP32 oracle's kernel__vmax
(src/swpc_3d/m_kernel.f90:350-374).  A surface beyond the scratch array's
capacity and an output that is not device memory are refused with an error
before anything is launched."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from gpu_helpers import gpu_from_model, small_config

pytestmark = pytest.mark.gpu


def test_noalloc_vmax_equals_vmax_on_two_calls(orc):
    import torch
    d_out = torch.empty(3, dtype=torch.float32, device="cuda:0")
    src2 = [dict(x=-0.9, y=0.8, z=3.1, t0=0.0, tr=0.6, mo=2.5e15, strike=120.0, dip=35.0, rake=-60.0)]
    cases = [small_config(nx=72, ny=64, nz=40, na=8, nt=40),
             small_config(nx=48, ny=48, nz=36, na=8, nt=40, sources=src2)]
    seen = []
    for n, cfg in enumerate(cases):
        m = O.build_model(cfg)
        run = O.OracleRun(m, 32)
        run.run(40)
        g = gpu_from_model(m)
        g.step(1, 40, use_graph=True, block=10)
        want = g.kernel_vmax()
        d_out.fill_(-1.0)
        torch.cuda.synchronize()
        g.kernel_vmax_noalloc(d_out.data_ptr())
        g.synchronize()
        got = d_out.cpu().numpy()
        print(f"call {n}: {cfg.nx}x{cfg.ny}x{cfg.nz} vmax {want} noalloc {got}", flush=True)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (n, got, want)
        assert np.array_equal(want, run.vmax()), (n, want, run.vmax())
        assert np.all(want > 0), want
        seen.append(want)
        g.close()
    # the two calls see different maxima, so stale partials of the first would show in the second
    assert not np.any(seen[0] == seen[1]), seen


def test_noalloc_vmax_refuses_large_surface_and_host_output(orc):
    from openswpc_amd.gpu import SwpcError, SwpcGPU
    # 300 x 8300 cells: 2 x 8288 sampling blocks of 256 cells, beyond the 16384 the scratch holds
    big = SwpcGPU(300, 8300, 4, 1, 0, "pml", 0.5, 0.5, 0.5, 0.02)
    with pytest.raises(SwpcError, match="scratch holds"):
        big.kernel_vmax_noalloc(1 << 20)
    big.close()
    cfg = small_config()
    g = gpu_from_model(O.build_model(cfg))
    host = np.zeros(3, np.float32)
    with pytest.raises(SwpcError, match="not device memory"):
        g.kernel_vmax_noalloc(host.ctypes.data_as(C.c_void_p).value)
    g.close()
