"""GPU test of vt_project_noalloc: vt_project with every per-call buffer in
the library's fixed device array, called through the C-ABI.

Two calls in a row (the second with fewer rays) must give vt_project's bits,
on a periodic clustered mesh and on the degenerate lattice rays, whose
K_EXPAND rays run in the second (deferred) launch; the call must not lower
the device's free memory; a ray count above the limit is refused.
"""
import ctypes as C
import sys

import numpy as np
import pytest
import torch

from conftest import LATTICE_RAYS, cubic_lattice

pytestmark = pytest.mark.gpu

import vortrace_amd as vt  # noqa: E402
from vortrace_amd import _lib, grid, synthetic  # noqa: E402

MAX_RAYS = 1 << 22


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _fn():
    L = _lib.lib()
    f = L.vt_project_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p,
                  C.c_void_p]
    return f


def _noalloc(pc, s, e, mode):
    """vt_project_noalloc on device tensors; returns (rc, out)."""
    cloud = pc._cloud
    nout = 3 if mode == vt.ReductionMode.VolumeRender else cloud.get_nfields()
    out = torch.full((s.shape[0], nout), -7.0, dtype=torch.float64, device=s.device)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = _fn()(cloud._handle(), C.c_void_p(s.data_ptr()), C.c_void_p(e.data_ptr()),
               s.shape[0], int(mode), C.c_void_p(out.data_ptr()), stream)
    return rc, out


def _regular(pc, s, e, mode):
    return vt.Projection(s, e).project_tensor(pc._cloud, mode)


def _case(name):
    dev = torch.device("cuda")
    if name == "clustered":
        pos, rho = synthetic.clustered_mesh(60_000, seed=71, n_clumps=16)
        pc = vt.ProjectionCloud(pos, rho, boundbox=[0, 100] * 3, periodic=True)
        s, e = grid.generate_projection_grid([0.0, 100.0], 96, [0.0, 100.0], None)
        s, e = s.reshape(-1, 3), e.reshape(-1, 3)
        mode = vt.ReductionMode.Sum
    else:
        pos, f, vol, bb = cubic_lattice()
        pc = vt.ProjectionCloud(pos, np.column_stack([f, 2 * f]), boundbox=bb, vol=vol)
        s = np.array([r[0] for r in LATTICE_RAYS] * 30, float)
        e = np.array([r[1] for r in LATTICE_RAYS] * 30, float)
        mode = vt.ReductionMode.Max
    return pc, torch.from_numpy(s).to(dev), torch.from_numpy(e).to(dev), mode


@pytest.mark.parametrize("name", ["clustered", "lattice"])
def test_noalloc_matches_vt_project(name):
    pc, s, e, mode = _case(name)
    n = s.shape[0]
    _say(f"  no-alloc {name}: {n} rays, then {n // 3}")
    want = _regular(pc, s, e, mode)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    rc, got = _noalloc(pc, s, e, mode)
    _lib.check(rc)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    st = _lib.last_call_stats()
    _say(f"    {st.segments} crossings, {st.deferred_rays} deferred rays; free memory "
         f"{free0 - free1} bytes lower")
    if name == "lattice":
        assert st.deferred_rays > 0
    assert torch.equal(got, want)
    assert free1 >= free0 - (1 << 20)
    # a second, smaller call reuses the same array
    m = n // 3
    s2, e2 = s[m:2 * m].contiguous(), e[m:2 * m].contiguous()
    rc, got2 = _noalloc(pc, s2, e2, mode)
    _lib.check(rc)
    assert torch.equal(got2, _regular(pc, s2, e2, mode))


def test_noalloc_refuses_more_rays_than_it_holds():
    pc, s, e, mode = _case("clustered")
    cloud = pc._cloud
    out = torch.empty((1, 1), dtype=torch.float64, device=s.device)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    # refused before any device access
    rc = _fn()(cloud._handle(), C.c_void_p(s.data_ptr()), C.c_void_p(e.data_ptr()),
               MAX_RAYS + 1, int(mode), C.c_void_p(out.data_ptr()), stream)
    with pytest.raises(ValueError, match="at most 4194304 per call"):
        _lib.check(rc)
    rc, got = _noalloc(pc, s[:0], e[:0], mode)
    assert rc == 0 and got.numel() == 0
