"""The allocation-free solve entry point (``efea_solve_pcg_noalloc``) against
``efea_solve_pcg`` through the C ABI: the same kernels on the same grids, so
the solutions and the solve reports agree bitwise, on two calls in a row (the
second on a smaller system, which reuses the static scratch the first one
left behind), in FP64 and mixed mode; a system beyond the static scratch is
refused on the host with a clear error."""

import ctypes as C

import pytest
import torch

import easyfea_amd as ea
from easyfea_amd import Models, Simulations, _lib
from easyfea_amd.device import ptr, stream_ptr

pytestmark = pytest.mark.gpu

NOALLOC_SCRATCH_BYTES = 16 << 20


def _system(nx, ny, mode):
    """A cantilever's Dirichlet-reduced solve ready for the C ABI: matrix,
    known mask, block-Jacobi inverse, right-hand side, start vector."""
    L, h = 120.0, 13.0
    mesh = ea.quad4_rectangle(L, h, nx, ny)
    simu = Simulations.Elastic(mesh, Models.Elastic.Isotropic(2, 210000.0, 0.3, planeStress=True, thickness=h))
    simu.preconditioner = "bj"
    simu.solver = "hip_fp64" if mode == 64 else "hip_mixed"
    simu.mixed_adaptive = False
    simu.add_dirichlet(mesh.Nodes_Conditions(lambda x, y, z: x == 0), [0, 0], ["x", "y"])
    simu.add_surfLoad(mesh.Nodes_Conditions(lambda x, y, z: x == L), [-800 / h**2], ["y"])
    A = simu._device_K()
    b = simu._neumann_device("elastic")
    x = torch.zeros(A.Ndof, dtype=torch.float64, device="cuda")
    job = simu._make_job(A, b, x, "elastic")
    job.setup(mode)
    return simu, A, job


def _args(A, job, mode, x):
    return (mode, A.Nn, A.dof_n, A.group, ptr(A.node_ptr), ptr(A.node_cols), ptr(A.values),
            ptr(A.values32 if mode == 32 else None), ptr(A.corr32 if mode == 32 else None), ptr(job.known),
            ptr(job.Minv), ptr(job.b), ptr(x))


def _solve_both(A, job, mode):
    lib = _lib.load()
    tol = (1e-12, 1e-9, 0.0, 100000, 32, 0, 0)
    x_ws = job.x.clone()
    nbytes = int(lib.efea_pcg_workspace_bytes(A.Nn, A.dof_n, A.group))
    ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    info_ws = (C.c_double * 4)()
    rc1 = _lib.call("efea_solve_pcg", *_args(A, job, mode, x_ws), ptr(ws), nbytes, *tol, stream_ptr(), info_ws)
    x_na = job.x.clone()
    info_na = (C.c_double * 4)()
    rc2 = _lib.call("efea_solve_pcg_noalloc", *_args(A, job, mode, x_na), *tol, stream_ptr(), info_na)
    torch.cuda.synchronize()
    return rc1, rc2, x_ws, x_na, list(info_ws), list(info_na)


@pytest.mark.parametrize("mode", [64, 32])
def test_noalloc_equals_workspace_solve_twice(mode):
    sizes = [(120, 13), (40, 5)]  # the second call on a smaller system
    for nx, ny in sizes:
        simu, A, job = _system(nx, ny, mode)
        print(f"no-alloc solve: mode {mode}, QUAD4 {nx}x{ny}, {A.Ndof} DOF", flush=True)
        rc1, rc2, x_ws, x_na, info_ws, info_na = _solve_both(A, job, mode)
        assert rc1 == 0 and rc2 == 0, (rc1, rc2, _lib.last_error())
        assert info_ws == info_na and info_na[3] == 1, (info_ws, info_na)
        assert torch.equal(x_ws, x_na), float((x_ws - x_na).abs().max())


def test_noalloc_refuses_a_system_beyond_its_scratch():
    """Checked on the host before any device work (the pointers are never read)."""
    lib = _lib.load()
    dummy = torch.zeros(16, dtype=torch.float64, device="cuda")
    p = ptr(dummy)
    Nnode = 300_000  # 900,000 DOF at dof_n 3: about 36 MB of scratch
    assert int(lib.efea_pcg_workspace_bytes(Nnode, 3, 16)) > NOALLOC_SCRATCH_BYTES
    info = (C.c_double * 4)()
    rc = lib.efea_solve_pcg_noalloc(64, Nnode, 3, 16, p, p, p, None, None, p, p, p, p, 1e-12, 1e-9, 0.0, 10, 32, 0,
                                    0, stream_ptr(), info)
    assert rc == -4
    assert "static scratch" in _lib.last_error()
