"""Solver features that must not change results (GPU tests): the step-1
timing harness computes the solver's own product, chunks launched
asynchronously with systems joining mid-flight stay bitwise equal to solo
solves, predicted state read-backs leave FP64 iterates unchanged, and the
lagged coarse inverse keeps the staggered history within the solver
tolerance of rebuilding it every increment.
"""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from easyfea_amd import _lib, workloads  # noqa: E402
from easyfea_amd.device import BatchSolver, drive, ptr, stream_ptr  # noqa: E402


def _sent_system(cells):
    simu = workloads.sent(cells, solver="hip_fp64")
    workloads.sent_bc(simu, 1e-4)
    simu._assemble_elastic()
    return simu


@pytest.mark.parametrize("two_level", [0, 1])
def test_step1_harness_computes_the_solver_product(two_level):
    simu = _sent_system(40)
    A = simu._Ku
    known_np, _ = simu._dirichlet_state("elastic")
    known = torch.from_numpy(known_np).cuda()
    nagg = A.coarse().nagg if two_level else 0
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    z = torch.rand(A.Ndof, dtype=torch.float64, device="cuda", generator=g)
    q = torch.full_like(z, float("nan"))
    scratch = torch.empty(_lib.load().efea_step1_bench_scratch_bytes(A.Nn, A.dof_n, nagg), dtype=torch.uint8,
                          device="cuda")
    print(f"step1 harness: two_level {two_level}, Nn {A.Nn}, group {A.group}, nagg {nagg}", flush=True)
    _lib.call("efea_step1_bench", two_level, 64, A.Nn, A.dof_n, A.group, ptr(A.node_ptr), ptr(A.node_cols),
              ptr(A.values), None, ptr(known), nagg, ptr(z), ptr(q), ptr(scratch), 3, stream_ptr())
    # the harness's p is z + 1 * 0 and its q = A p: the SpMV's own sums
    ref = A.spmv(z, known=known)
    torch.cuda.synchronize()
    assert torch.equal(q, ref)


def test_batch_join_while_a_chunk_runs_is_bitwise_solo():
    specs = [workloads.ElasticRequestSpec(i, s, workloads.lognormal(
        210000.0, 0.2, workloads.gaussian_field((s, s), 0.1 * s, np.random.default_rng(i)).T.ravel()), 0.0)
        for i, s in enumerate([20, 33, 27])]
    alone = []
    for s in specs:
        u, info = drive(workloads.elastic_request(s, workloads.square_quad4(s.side), "fp64", "two-level"))
        alone.append((u.clone(), info))
    gens, jobs = [], []
    for s in specs:
        gen = workloads.elastic_request(s, workloads.square_quad4(s.side), "fp64", "two-level")
        gens.append(gen)
        jobs.append(next(gen))
    b = BatchSolver(64, 2, jobs[0].A.group, "two-level", capacity=4)
    for j in jobs:
        assert j.A.group == b.group and j.precond == "two-level"
    jobs[0].setup(64)
    b.add(jobs[0])
    done = {}
    assert b.launch(32)
    # the other systems join while the first chunk is on the device
    for j in jobs[1:]:
        j.setup(64)
        b.add(j)
    while len(b):
        for job, info in b.wait():
            done[id(job)] = info
        b.launch(32)
    b.wait()
    b.close()
    for k, (j, gen) in enumerate(zip(jobs, gens)):
        info = done[id(j)]
        try:
            gen.send(info)
        except StopIteration as stop:
            u = stop.value[0]
        u_ref, info_ref = alone[k]
        print(f"side {specs[k].side}: {info.iterations} it (alone {info_ref.iterations})", flush=True)
        assert info.iterations == info_ref.iterations and info.status == 1
        assert torch.equal(u, u_ref)


def test_predicted_polls_leave_fp64_iterates_unchanged():
    mesh = workloads.square_quad4(48)
    spec = workloads.ElasticRequestSpec(7, 48, np.full(48 * 48, 210000.0), 0.0)
    u0, info0 = drive(workloads.elastic_request(spec, mesh, "fp64", "two-level"))
    u0 = u0.clone()
    for expected in (info0.iterations - 20, info0.iterations + 3, 5):
        gen = workloads.elastic_request(spec, mesh, "fp64", "two-level")
        job = next(gen)
        job.expected_iterations = expected
        assert job._polls() == (8, max(1, int(0.95 * expected)))
        try:
            gen.send(job.run())
        except StopIteration as stop:
            u, info = stop.value
        assert info.iterations == info0.iterations and info.rel_residual == info0.rel_residual
        assert torch.equal(u, u0)


def test_lagged_coarse_inverse_stays_within_solver_tolerance():
    out = {}
    for reuse in (False, True):
        simu = workloads.sent(40, solver="hip_fp64")
        simu.coarse_reuse = reuse
        dofs = workloads.sent_top_y_dofs(simu)
        forces, reused = [], []
        for n in range(1, 5):
            workloads.sent_bc(simu, n * 1e-4)
            simu.Solve(1.0, 500, 1)
            simu.Save_Iter()
            forces.append(float(np.sum(simu.Calc_Reaction(dofs))))
            info = simu.last_solve["elastic"]
            assert info.precond == "two-level" and info.status == 1
            reused.append(info.coarse_reused)
        out[reuse] = (np.array(forces), simu._u.clone(), simu._d.clone(), reused)
    f0, u0, d0, r0 = out[False]
    f1, u1, d1, r1 = out[True]
    assert not any(r0) and r1[0] is False and all(r1[1:])
    assert np.max(np.abs(f1 - f0) / np.abs(f0)) < 1e-7
    assert float(torch.linalg.vector_norm(u1 - u0) / torch.linalg.vector_norm(u0)) < 1e-7
    assert float(torch.linalg.vector_norm(d1 - d0)) <= 1e-7 * max(float(torch.linalg.vector_norm(d0)), 1e-30)


def test_batch_launch_and_wait_without_systems():
    h = C.c_void_p()
    assert _lib.load().efea_batch_create(64, 2, 4, 0, 2, C.byref(h)) == 0
    try:
        # nothing active: launch is a no-op, wait returns 0
        assert _lib.load().efea_batch_launch(h, 8, None) == 0
        assert _lib.load().efea_batch_wait(h, None) == 0
    finally:
        _lib.load().efea_batch_destroy(h)


GROUPS = [(1, 2), (1, 4), (1, 8), (1, 16), (2, 4), (2, 8), (2, 16), (3, 8), (3, 16), (3, 32)]


def _simu_for(dof_n):
    import easyfea_amd as ea
    from easyfea_amd import Models, Simulations

    if dof_n == 1:
        mesh = ea.quad4_rectangle(1.0, 1.0, 30, 20)
        simu = Simulations.Thermal(mesh, Models.Thermal(1.3, 1.0))
        simu.add_dirichlet(mesh.Nodes_Conditions(lambda x, y, z: x == 0), [0.0], ["t"])
        simu.add_dirichlet(mesh.Nodes_Conditions(lambda x, y, z: x == 1.0), [2.0], ["t"])
        return simu, lambda: simu._device_KC()[0]
    if dof_n == 2:
        mesh = ea.tri3_rectangle(2.0, 1.0, 30, 15)
        simu = Simulations.Elastic(mesh, Models.Elastic.Isotropic(2, 210000.0, 0.3))
        simu.add_dirichlet(mesh.Nodes_Conditions(lambda x, y, z: x == 0), [0, 0], ["x", "y"])
        simu.add_dirichlet(mesh.Nodes_Conditions(lambda x, y, z: x == 2.0), [1e-3, 0], ["x", "y"])
        return simu, simu._device_K
    mesh = ea.hexa8_box(2.0, 1.0, 1.0, 10, 5, 5)
    simu = Simulations.Elastic(mesh, Models.Elastic.Isotropic(3, 210000.0, 0.3))
    simu.add_dirichlet(mesh.Nodes_Conditions(lambda x, y, z: x == 0), [0, 0, 0], ["x", "y", "z"])
    simu.add_dirichlet(mesh.Nodes_Conditions(lambda x, y, z: x == 2.0), [1e-3, 0, 0], ["x", "y", "z"])
    return simu, simu._device_K


@pytest.mark.parametrize("dof_n,group", GROUPS)
def test_every_lanes_per_node_group_multiplies_and_solves(dof_n, group):
    simu, getK = _simu_for(dof_n)
    simu.solver = "hip_fp64"
    simu.preconditioner = "bj"
    u_default = simu.Solve().copy()
    K = getK()
    K.group = group
    print(f"dof_n {dof_n}, group {group}: Nn {K.Nn}", flush=True)
    x = torch.rand(K.Ndof, dtype=torch.float64, device="cuda")
    y = K.spmv(x).cpu().numpy()
    ref = K.to_scipy() @ x.cpu().numpy()
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref))
    simu._u.zero_()
    u = simu.Solve()
    assert simu.last_solve[simu.problemType].status == 1
    assert np.linalg.norm(u - u_default) <= 1e-7 * np.linalg.norm(u_default)
