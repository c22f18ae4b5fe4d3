"""Continuous batching: every solve gives the same result alone and in the mix
(GPU tests).

Each system in a batch runs on the grid it would get alone, with the same
per-block code, so its iterates, iteration count and final residual must be
bitwise those of the single-system path, whatever else shares the launch and
whenever it joined.  The serving loop's results are also checked against the
CPU oracle (elastic: sparse direct solve within 1e-8).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from easyfea_amd import ElemType, workloads  # noqa: E402
from easyfea_amd.device import BatchExecutor, drive  # noqa: E402
from easyfea_amd.mesh import tri3_notched_square  # noqa: E402
from easyfea_amd.serving import Request, Server, estimate_bytes  # noqa: E402
from oracle import efea_oracle as O  # noqa: E402


@pytest.fixture(autouse=True)
def _force_fp32_krylov(monkeypatch):
    """Small systems sit below the size where hip_mixed switches to FP32
    Krylov products; force that path so that it is exercised."""
    import easyfea_amd.device as dv

    monkeypatch.setattr(dv, "MIXED_MIN_MATRIX_BYTES", 0)


def _elastic_spec(rid, side):
    z = workloads.gaussian_field((side, side), 0.1 * side, np.random.default_rng(rid))
    return workloads.ElasticRequestSpec(rid, side, workloads.lognormal(210000.0, 0.2, z.T.ravel()), 0.0)


def _alone(spec, mode, precond="auto"):
    mesh = workloads.square_quad4(spec.side)
    u, info = drive(workloads.elastic_request(spec, mesh, mode, precond))
    return u.clone(), info


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
@pytest.mark.parametrize("precond,sides", [("bj", [14, 23, 31, 9]), ("two-level", [36, 45, 20, 52])])
def test_batch_bitwise_equals_single_with_staggered_joins(mode, precond, sides):
    specs = [_elastic_spec(i, s) for i, s in enumerate(sides)]
    ref = [_alone(s, mode, precond) for s in specs]
    assert all(info.precond == precond for _, info in ref)
    ex = BatchExecutor(capacity=8, check_every=32)
    gens, results = {}, {}

    def start(spec):
        g = workloads.elastic_request(spec, workloads.square_quad4(spec.side), mode, precond)
        job = next(g)
        job.owner = spec.rid
        gens[spec.rid] = g
        ex.submit(job)

    start(specs[0])
    for _ in range(2):
        for job, info in ex.step():
            _finish(gens, results, job, info)
    start(specs[1])
    start(specs[2])
    for job, info in ex.step():
        _finish(gens, results, job, info)
    start(specs[3])
    while ex.active():
        for job, info in ex.step():
            _finish(gens, results, job, info)
    ex.close()
    for spec, (u_ref, info_ref) in zip(specs, ref):
        u, info = results[spec.rid]
        print(f"side {spec.side} {mode}: {info.iterations} it (alone {info_ref.iterations}), "
              f"replacements {info.replacements}")
        assert info.status == 1 and info.precond == precond
        assert info.iterations == info_ref.iterations
        assert info.replacements == info_ref.replacements
        assert info.rel_residual == info_ref.rel_residual
        assert torch.equal(u, u_ref), f"side {spec.side}: batched result differs from the solo solve"


def _finish(gens, results, job, info):
    try:
        gens[job.owner].send(info)
    except StopIteration as stop:
        results[job.owner] = (stop.value[0].clone(), stop.value[1])


def test_server_elastic_stream_matches_alone_and_oracle():
    sides = [12, 30, 17, 25, 40, 8]
    specs = []
    for i, s in enumerate(sides):
        sp_ = _elastic_spec(100 + i, s)
        sp_.arrival = 0.002 * i
        specs.append(sp_)
    meshes = workloads.MeshCache(workloads.square_quad4)
    reqs = []
    for spec in specs:
        m = meshes.get(spec.side)
        est = estimate_bytes(m.Nn, m.Ne, 4, (2,), (18 * m.Nn * 2,), True)
        reqs.append(Request(spec.rid, spec.arrival,
                            (lambda sp_=spec, m_=m: workloads.elastic_request(sp_, m_, "mixed")), est))
    srv = Server(hbm_budget=3 * max(r.est_bytes for r in reqs), capacity=8, check_every=32)
    srv.run(reqs)
    srv.close()
    assert srv.max_active >= 2, "requests never overlapped on the device"
    for spec, req in zip(specs, reqs):
        assert req.error is None, req.error
        u, info = req.result
        u_ref, info_ref = _alone(spec, "mixed")
        assert torch.equal(u, u_ref)
        assert info.iterations == info_ref.iterations
        # oracle: sparse direct solve of the same system
        mesh = meshes.get(spec.side)
        left = mesh.Nodes_Conditions(lambda x, y, z: x == 0.0)
        right = mesh.Nodes_Conditions(lambda x, y, z: x == 1.0)
        ref = O.ElasticOracle(mesh.connect, mesh.coord, "QUAD4", spec.E, 0.3, True, 1.0,
                              {"SEG2": mesh.dict_groupElem[ElemType.SEG2].connect})
        ref.add_dirichlet(left, [0, 0], ["x", "y"])
        ref.add_surfLoad(right, [1.0], ["x"])
        x = ref.Solve()
        err = np.linalg.norm(u.cpu().numpy() - x) / np.linalg.norm(x)
        assert err < 1e-8, f"side {spec.side}: {err:.2e} vs the direct solve"


def test_server_phasefield_realisations_match_alone():
    specs = workloads.serving_phasefield_specs(n=3, steps=3, ne_lo=800, ne_hi=1800, seed=3)
    meshes = workloads.MeshCache(tri3_notched_square)
    clock = __import__("time").perf_counter
    alone = []
    for spec in specs:
        s2 = workloads.PhaseFieldRequestSpec(spec.rid, spec.cells, spec.Gc, 0.0, spec.steps)
        simu = drive(workloads.phasefield_request(s2, meshes.get(spec.cells), "mixed", clock))
        alone.append((np.array(s2.forces), simu._d.clone(), simu._u.clone()))
    reqs = [Request(s.rid, 0.0, (lambda s_=s: workloads.phasefield_request(s_, meshes.get(s_.cells), "mixed",
                                                                            clock)), 1)
            for s in specs]
    srv = Server(hbm_budget=1 << 40, capacity=8)
    srv.run(reqs)
    srv.close()
    assert srv.max_active >= 2
    for spec, req, (f_ref, d_ref, u_ref) in zip(specs, reqs, alone):
        assert req.error is None, req.error
        simu = req.result
        assert np.array_equal(np.array(spec.forces), f_ref)
        assert torch.equal(simu._d, d_ref) and torch.equal(simu._u, u_ref)
        assert len(spec.step_latency) == spec.steps


def test_executor_queues_jobs_beyond_batch_capacity():
    """A batch of capacity 2 fed 5 solves: the extra jobs wait and join as
    slots free up; every result stays bitwise the solo one."""
    specs = [_elastic_spec(200 + i, s) for i, s in enumerate([12, 18, 9, 15, 21])]
    ref = [_alone(s, "mixed", "bj") for s in specs]
    ex = BatchExecutor(capacity=2, check_every=32)
    gens, results = {}, {}
    for spec in specs:
        g = workloads.elastic_request(spec, workloads.square_quad4(spec.side), "mixed", "bj")
        job = next(g)
        job.owner = spec.rid
        gens[spec.rid] = g
        ex.submit(job)
    assert ex.active() == 5
    while ex.active():
        assert max(len(b) for b in ex.batches.values()) <= 2
        for job, info in ex.step():
            _finish(gens, results, job, info)
    ex.close()
    for spec, (u_ref, info_ref) in zip(specs, ref):
        u, info = results[spec.rid]
        assert info.iterations == info_ref.iterations and torch.equal(u, u_ref)
