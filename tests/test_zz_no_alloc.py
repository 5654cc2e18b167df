"""lm_adjoint_forward_noalloc, the allocation-free variant of
lm_adjoint_forward, against that entry point through the C ABI.

Two calls in a row on one device, the second with fewer particles and steps
(the static scratch keeps the first call's stage fields, which the second
must overwrite before reading): positions, checkpoints and the outside-grid
count must be bit-identical to lm_adjoint_forward's on the same inputs.  The
first call's 240 steps on the 81x81-node grid need more stage fields than
the 32 MiB scratch holds, so it also runs in several chunks.  A grid too
large for the scratch is rejected on the host.
"""
import ctypes

import numpy as np
import pytest
import torch

from lagrangian_microbes_amd import _lib, adjoint, configs
from lagrangian_microbes_amd._lib import LmGrid, check, ptr, stream_ptr


def _forward(entry, w, f, lon, lat, nsteps, ckpt):
    plan = adjoint.AdjointPlan(f, 0.0, w.dt, nsteps, ckpt)
    u = torch.tensor(f.u, dtype=torch.float32, device="cuda")
    v = torch.tensor(f.v, dtype=torch.float32, device="cuda")
    field = plan.interleave(u, v)
    x = torch.tensor(lon, dtype=torch.float64, device="cuda")
    y = torch.tensor(lat, dtype=torch.float64, device="cuda")
    ck = torch.full((plan.nseg, x.numel(), 2), np.nan, dtype=torch.float64, device="cuda")
    L = _lib.load()
    args = (x.numel(), ptr(x), ptr(y), ptr(field), plan.nlev, ctypes.byref(plan.grid), ptr(plan.lev0), ptr(plan.w),
            plan.lev0_host_ptr(), plan.nsteps, plan.dt, plan.ckpt, ptr(ck))
    if entry == "noalloc":
        check(L.lm_adjoint_forward_noalloc(*args, ptr(plan.oob), stream_ptr()))
    else:
        scratch = torch.empty(adjoint.forward_scratch_bytes(plan), dtype=torch.uint8, device="cuda")
        check(L.lm_adjoint_forward(*args, ptr(scratch), scratch.numel(), ptr(plan.oob), stream_ptr()))
    torch.cuda.synchronize()
    return x.cpu().numpy(), y.cpu().numpy(), ck.cpu().numpy(), int(plan.oob.item())


@pytest.mark.gpu
def test_noalloc_matches_forward_twice(gpu):
    w = configs.adjoint_1m_10day()
    f = w.velocity_field()
    lon, lat = w.particles()
    rng = np.random.default_rng(11)
    # 240 steps, then a smaller second call: 4,999 particles (a ragged block), 37 steps
    for n, steps, ckpt in ((20_000, 240, 16), (4_999, 37, 8)):
        sel = np.sort(rng.choice(lon.size, n, replace=False))
        a = _forward("noalloc", w, f, lon[sel], lat[sel], steps, ckpt)
        b = _forward("forward", w, f, lon[sel], lat[sel], steps, ckpt)
        for p, q in zip(a[:3], b[:3]):
            assert np.array_equal(p, q)
        assert a[3] == b[3] == 0
        assert not np.array_equal(a[0], lon[sel])


def test_noalloc_rejects_large_grid():
    L = _lib.load()
    g = LmGrid(1000, 700, 0.0, 0.0, 4.0, 4.0)  # 700,000 nodes: 3 stage fields exceed 32 MiB
    lev = np.zeros(3, dtype=np.int32)
    rc = L.lm_adjoint_forward_noalloc(10, None, None, None, 2, ctypes.byref(g), None, None,
                                      lev.ctypes.data_as(ctypes.c_void_p), 1, 3600.0, 16, None, None, None)
    assert rc != 0 and b"static scratch" in L.lm_last_error()
    g = LmGrid(1000, 699, 0.0, 0.0, 4.0, 4.0)  # 699,000 nodes fit: rejected only for the NULL arrays
    rc = L.lm_adjoint_forward_noalloc(10, None, None, None, 2, ctypes.byref(g), None, None,
                                      lev.ctypes.data_as(ctypes.c_void_p), 1, 3600.0, 16, None, None, None)
    assert rc != 0 and b"NULL" in L.lm_last_error()
