"""Host logic and the C ABI (CPU only: no compute call reaches a GPU)."""

import ctypes as C
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

import easyfea_amd as ea
from easyfea_amd import _lib
from easyfea_amd.assembly_map import build_csr_map, build_vec_map
from oracle import efea_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ---------------------------------------------------------------- C ABI
def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "efea.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(efea_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = _lib.load()
    declared = _declared_symbols()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(lib, name), f"{name} missing from libefea_hip.so"
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"
    assert lib.efea_version() == 1


def test_abi_argument_errors_are_reported_without_gpu():
    lib = _lib.load()
    rc = lib.efea_elastic_ke(1, 64, -1, None, None, None, 4, None, 0, 0.3, 1, 1.0, None, None)
    assert rc == -1
    assert "negative size" in _lib.last_error()
    rc = lib.efea_solve_pcg(64, 0, 2, 16, None, None, None, None, None, None, None, None, None, None, 0, 1e-12,
                            1e-9, 1e-3, 10, 1, 0, 0, None, None)
    assert rc == -1
    rc = lib.efea_pf_split_points(4, 7, None, None, 0, 0.3, 0, None, None, None, None, None)
    assert rc == -1 and "null pointer" in _lib.last_error()
    h = C.c_void_p()
    assert lib.efea_batch_create(16, 2, 16, 0, 8, C.byref(h)) == -1 and "mode" in _lib.last_error()
    assert lib.efea_batch_create(64, 2, 12, 0, 8, C.byref(h)) == -1 and "group" in _lib.last_error()
    assert lib.efea_batch_create(64, 2, 16, 1, 0, C.byref(h)) == -1 and "capacity" in _lib.last_error()
    assert h.value is None
    assert lib.efea_batch_run(None, 32, None) == -1
    assert lib.efea_batch_add(None, 10, None, None, None, None, None, None, None, 0, None, None, None, None, 0,
                              None, None, None, 0, 1e-12, 1e-9, 1e-3, 10, 0, None) == -1
    assert lib.efea_coarse_ld(0, 2) == -1
    assert lib.efea_coarse_ld(100, 2) == 320 and lib.efea_coarse_ld(100, 3) == 608 and lib.efea_coarse_ld(5, 1) == 32
    assert lib.efea_solve_pcg2(64, 10, 2, 16, None, None, None, None, None, None, None, 0, None, None, None, None,
                               0, None, None, None, 0, 1e-12, 1e-9, 1e-3, 10, 32, 0, 0, None, None) == -1
    assert lib.efea_split_f32(-1, 2, None, None, None, None, None) == -1
    assert lib.efea_step1_bench(1, 64, 10, 2, 4, None, None, None, None, None, 0, None, None, None, 3, None) == -1
    assert "nagg" in _lib.last_error()
    assert lib.efea_step1_bench_scratch_bytes(100, 2, 16) >= 2 * 100 * 2 * 8
    assert lib.efea_batch_launch(None, 8, None) == -1 and lib.efea_batch_wait(None, None) == -1
    jid = C.c_int(5)
    assert lib.efea_batch_pop(None, C.byref(jid), None) == -1
    assert lib.efea_batch_destroy(None) == 0
    # pipelined block Jacobi: one grid pass of at most efea_pcgp_max_lanes() lanes
    assert lib.efea_pcgp_max_lanes() == 1024 * 256
    fake = C.c_void_p(256)  # never dereferenced: the size checks come first
    assert lib.efea_solve_pcgp(64, 0, 1, 4, *[fake] * 10, 1 << 30, 1e-12, 1e-9, 0, 10, 8, 0, None, fake) == -1
    assert "null pointer or empty" in _lib.last_error()
    assert lib.efea_solve_pcgp(16, 100, 1, 4, *[fake] * 10, 1 << 30, 1e-12, 1e-9, 0, 10, 8, 0, None, fake) == -1
    assert "precision" in _lib.last_error()
    assert lib.efea_solve_pcgp(64, 70000, 1, 4, *[fake] * 10, 1 << 30, 1e-12, 1e-9, 0, 10, 8, 0, None, fake) == -1
    assert "lanes" in _lib.last_error()
    assert lib.efea_solve_pcgp(64, 100, 1, 4, *[fake] * 10, 16, 1e-12, 1e-9, 0, 10, 8, 0, None, fake) == -4
    assert "workspace" in _lib.last_error()
    assert lib.efea_pcgp_workspace_bytes(100, 2) >= 10 * 100 * 2 * 8
    # its deferred form: the same checks, and at least one queued iteration
    assert lib.efea_pcgp_start(64, 0, 1, 4, *[fake] * 10, 1 << 30, 1e-12, 1e-9, 0, 10, 8, 40, None) == -1
    assert "efea_pcgp_start: null pointer or empty" in _lib.last_error()
    assert lib.efea_pcgp_start(64, 100, 1, 4, *[fake] * 10, 1 << 30, 1e-12, 1e-9, 0, 10, 8, 0, None) == -1
    assert "first_check" in _lib.last_error()
    assert lib.efea_pcgp_start(64, 100, 1, 4, *[fake] * 10, 16, 1e-12, 1e-9, 0, 10, 8, 40, None) == -4
    resumed = C.c_int(7)
    assert lib.efea_pcgp_finish(64, 100, 1, 4, *[fake] * 10, 1 << 30, 1e-12, 1e-9, 0, 10, 8, 40, None, None,
                                C.byref(resumed)) == -1
    assert "null info" in _lib.last_error()
    assert lib.efea_pcgp_finish(64, 100, 1, 4, *[fake] * 10, 1 << 30, 1e-12, 1e-9, 0, 10, 8, 0, None, fake,
                                C.byref(resumed)) == -1
    assert "first_check" in _lib.last_error()
    assert lib.efea_pcgp_finish(32, 100, 1, 4, fake, fake, fake, None, *[fake] * 6, 1 << 30, 1e-12, 1e-9, 0, 10,
                                8, 40, None, fake, None) == -1
    assert "precision" in _lib.last_error()
    # FP32 values need their row-sum correction (the iterations read it unconditionally)
    assert lib.efea_solve_pcgp(32, 100, 1, 4, fake, fake, fake, fake, None, *[fake] * 5, 1 << 30, 1e-12, 1e-9, 0,
                               10, 8, 0, None, fake) == -1
    assert "row-sum correction" in _lib.last_error()
    # the two-level pipelined solve's deferred form: its checks, then first_check
    args = [64, 100, 2, 4, *[fake] * 7, 10, fake, fake, fake, fake, 64, 5, fake, fake, 2, fake, fake, fake, fake,
            1 << 30, 1e-12, 1e-9, 10, 8]
    assert lib.efea_pcg2p_start(*args, 0, None) == -1
    assert "efea_pcg2p_start: first_check" in _lib.last_error()
    assert lib.efea_pcg2p_start(*args[:11], 0, *args[12:], 40, None) == -1
    assert "null pointer or empty" in _lib.last_error()
    assert lib.efea_pcg2p_finish(*args, 40, None, None, None) == -1
    assert "null info" in _lib.last_error()
    assert lib.efea_pcg2p_finish(*args, 0, None, fake, None) == -1
    assert "efea_pcg2p_finish: first_check" in _lib.last_error()
    assert lib.efea_solve_pcg2p(32, *args[1:8], None, *args[9:], 40, None, fake) == -1
    assert "row-sum correction" in _lib.last_error()
    # BiCGStab (non-symmetric systems)
    assert lib.efea_solve_bicgstab(0, 2, 4, *[fake] * 8, 1 << 30, 1e-12, 10, 8, None, fake) == -1
    assert "null pointer or empty" in _lib.last_error()
    assert lib.efea_solve_bicgstab(100, 2, 4, *[fake] * 8, 16, 1e-12, 10, 8, None, fake) == -4
    assert "workspace" in _lib.last_error()
    assert lib.efea_bicgstab_workspace_bytes(100, 2) >= 8 * 100 * 2 * 8


def test_workspace_sizes_and_groups():
    lib = _lib.load()
    # TRI3 / QUAD4 (7-9 neighbours): 4 lanes, 2 for scalar patterns from 5e4 nodes; quadratic 2D (up to 25): 16
    assert lib.efea_default_group(2, 9, 10**6) == 4 and lib.efea_default_group(1, 7, 40501) == 4
    assert lib.efea_default_group(1, 9, 10**6) == 2 and lib.efea_default_group(1, 13, 10**6) == 8
    assert lib.efea_default_group(2, 21, 100) == 8 and lib.efea_default_group(1, 25, 100) == 16
    # TETRA4 (15): 8 lanes; HEXA8 (27) and TETRA10 (65): 16
    assert lib.efea_default_group(3, 15, 1000) == 8
    assert lib.efea_default_group(3, 27, 1000) == 16 and lib.efea_default_group(3, 65, 1000) == 16
    small = lib.efea_pcg_workspace_bytes(100, 2, 16)
    big = lib.efea_pcg_workspace_bytes(10000, 2, 16)
    assert big > small >= 4 * 100 * 2 * 8


# ---------------------------------------------------------------- meshes
def test_notched_square_topology():
    n = 20
    m = ea.tri3_notched_square(n)
    assert m.Ne == 2 * n * n and m.Nn == (n + 1) ** 2 + n // 2
    # every element positively oriented, total area 1
    X = m.coord[m.connect][..., :2]
    area = 0.5 * ((X[:, 1, 0] - X[:, 0, 0]) * (X[:, 2, 1] - X[:, 0, 1]) - (X[:, 2, 0] - X[:, 0, 0]) * (X[:, 1, 1] - X[:, 0, 1]))
    assert np.all(area > 0) and abs(area.sum() - 1.0) < 1e-12
    # lower and upper crack faces share no node left of the tip
    y = m.coord[:, 1]
    crack = np.nonzero((np.abs(y - 0.5) < 1e-12) & (m.coord[:, 0] < 0.5 - 1e-12))[0]
    assert crack.size == n  # originals + copies
    below = m.connect[(X[..., 1].mean(1) < 0.5)]
    above = m.connect[(X[..., 1].mean(1) > 0.5)]
    assert not (set(below.ravel()) & set(above.ravel()) & set(crack.tolist()))
    # the tip node is shared
    tip = np.nonzero((np.abs(y - 0.5) < 1e-12) & (np.abs(m.coord[:, 0] - 0.5) < 1e-12))[0]
    assert tip.size == 1 and tip[0] in below and tip[0] in above


def test_rectangle_and_box_boundaries():
    m = ea.quad4_rectangle(120.0, 13.0, 360, 39)
    assert m.Nn == 14440 and m.Ne == 14040
    nL = m.Nodes_Conditions(lambda x, y, z: x == 120.0)
    assert nL.size == 40
    seg = m.dict_groupElem[ea.ElemType.SEG2]
    assert seg.Get_Elements_Nodes(nL).size == 39
    b = ea.hexa8_cube(4)
    assert b.Nn == 125 and b.Ne == 64
    top = b.Nodes_Conditions(lambda x, y, z: z == 1.0)
    assert b.dict_groupElem[ea.ElemType.QUAD4].Get_Elements_Nodes(top).size == 16
    g = O.Geometry("HEXA8", b.connect, b.coord, "rigi", 3)
    assert np.all(g.detJ > 0)


# ---------------------------------------------------------------- maps
@pytest.mark.parametrize("maker,dof_n", [
    (lambda: ea.quad4_rectangle(120.0, 13.0, 36, 9), 2),
    (lambda: ea.tri3_notched_square(16), 2),
    (lambda: ea.tri3_notched_square(16), 1),
    (lambda: ea.hexa8_box(1.0, 2.0, 1.5, 4, 3, 5), 3),
])
def test_scatter_map_matches_scipy_canonical(maker, dof_n):
    m = maker()
    cm = build_csr_map(m.connect, m.Nn, dof_n)
    inv, indices, indptr, nnz = O.csr_map(m.connect, dof_n, m.Nn * dof_n)
    assert np.array_equal(cm.indptr, indptr) and np.array_equal(cm.indices, indices)
    assert cm.indptr.dtype == np.int32 and cm.indices.dtype == np.int32
    slot = np.empty(cm.perm.size, dtype=np.int64)
    slot[cm.perm] = np.repeat(np.arange(cm.nnz), np.diff(cm.seg_ptr))
    assert np.array_equal(slot, inv)
    # within a slot, entries in increasing element-entry order (bincount order)
    starts = cm.seg_ptr[:-1]
    ok = np.ones(cm.perm.size, dtype=bool)
    ok[starts[starts < cm.perm.size]] = False
    assert np.all(np.diff(cm.perm)[ok[1:]] > 0)
    # gather-form scatter of random entries equals bincount exactly
    rng = np.random.default_rng(0)
    vals = rng.standard_normal(cm.perm.size)
    out = np.array([vals[cm.perm[a:b]].sum() if b > a else 0.0 for a, b in zip(cm.seg_ptr[:-1], cm.seg_ptr[1:])])
    seq = np.zeros(cm.nnz)
    for s_ in range(cm.nnz):
        acc = 0.0
        for k in cm.perm[cm.seg_ptr[s_]:cm.seg_ptr[s_ + 1]]:
            acc += vals[k]
        seq[s_] = acc
    assert np.array_equal(seq, np.bincount(inv, weights=vals, minlength=nnz))
    assert np.allclose(out, seq, rtol=1e-14, atol=1e-14)


def test_node_adjacency_matches_pattern():
    m = ea.tri3_notched_square(8)
    cm = build_csr_map(m.connect, m.Nn, 2)
    # row n*2+a holds, per neighbour m, columns 2m, 2m+1
    for n in range(0, m.Nn, 7):
        nb = cm.node_cols[cm.node_ptr[n]:cm.node_ptr[n + 1]]
        for a in range(2):
            r = 2 * n + a
            cols = cm.indices[cm.indptr[r]:cm.indptr[r + 1]]
            assert np.array_equal(cols, (nb[:, None] * 2 + np.arange(2)).ravel())


def test_vector_map():
    m = ea.quad4_rectangle(2.0, 1.0, 5, 3)
    vm = build_vec_map(m.connect, m.Nn, 1)
    rows = m.connect.ravel()
    vals = np.arange(rows.size, dtype=float)
    out = np.array([vals[vm.perm[a:b]].sum() for a, b in zip(vm.seg_ptr[:-1], vm.seg_ptr[1:])])
    assert np.array_equal(out, np.bincount(rows, weights=vals, minlength=m.Nn))


# ---------------------------------------------------------------- models
def test_models_parameters_and_observers():
    mat = ea.Models.Elastic.Isotropic(2, 210000.0, 0.3, planeStress=False, thickness=2.0)
    lam, mu, K = O.lame(210000.0, 0.3, 2, False)
    assert abs(mat.get_lambda() - lam) < 1e-9 and abs(mat.get_mu() - mu) < 1e-9
    assert abs(mat.get_bulk() - K) < 1e-9
    pfm = ea.Models.PhaseField(mat, "Miehe", "AT2", 2.7, 0.01)
    assert abs(pfm.k - 2.7 * 0.01) < 1e-18 and pfm.c_w == 2.0
    flags = []

    class Obs:
        def Need_Update(self, value=True):
            flags.append(value)

    o = Obs()
    pfm._Add_observer(o)
    mat._Add_observer(o)
    mat.E = 200000.0
    pfm.Gc = 3.0
    assert len(flags) >= 2
    with pytest.raises(AssertionError):
        mat.v = 0.5
    with pytest.raises(AssertionError):
        pfm.l0 = -1.0
    with pytest.raises(ValueError):
        ea.Models.PhaseField(mat, "Unknown", "AT2", 1.0, 1.0)


def test_solver_type_names():
    assert [s.value for s in ea.SolverType] == ["hip_fp64", "hip_mixed"]
    from easyfea_amd.Solvers import REFERENCE_ONLY

    assert "scipy" in REFERENCE_ONLY and "pypardiso" in REFERENCE_ONLY


def test_product_path_refuses_to_run_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    mesh = ea.quad4_rectangle(1.0, 1.0, 2, 2)
    mat = ea.Models.Elastic.Isotropic(2)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        ea.Simulations.Elastic(mesh, mat)


# ---------------------------------------------------------------- serving loop
class _FakeJob:
    def __init__(self, chunks):
        self.chunks = chunks
        self.owner = None
        self.request = None


class _FakeExecutor:
    """Completes each job after its number of chunks (no GPU)."""

    def __init__(self):
        self.jobs = []
        self.inflight = []
        self.peak = 0

    def submit(self, job):
        self.jobs.append(job)
        self.peak = max(self.peak, len(self.jobs))

    def active(self):
        return len(self.jobs)

    def launch(self):
        self.inflight = list(self.jobs)
        return bool(self.inflight)

    def collect(self):
        # only the jobs present at launch advanced; jobs submitted while the
        # chunk ran join the next one
        done = []
        for j in self.inflight:
            j.chunks -= 1
            if j.chunks <= 0:
                done.append(j)
        self.jobs = [j for j in self.jobs if j not in done]
        self.inflight = []
        return [(j, "info") for j in done]


def _work(nsolves, chunks, log, rid):
    infos = []
    for _ in range(nsolves):
        infos.append((yield _FakeJob(chunks)))
    log.append(rid)
    return infos


def test_serving_admission_by_budget_and_continuous_batching():
    from easyfea_amd.serving import Request, Server, percentiles

    log = []
    t = [0.0]
    reqs = [Request(i, 0.0, (lambda i=i: _work(2, 1 + i % 3, log, i)), est_bytes=40) for i in range(6)]
    ex = _FakeExecutor()
    srv = Server(hbm_budget=100, executor=ex, clock=lambda: t[0])
    srv.run(reqs)
    assert all(r.error is None and r.result == ["info", "info"] for r in reqs)
    assert ex.peak == 2            # two requests of 40 bytes fit a budget of 100
    assert srv.max_resident <= 100
    assert sorted(log) == list(range(6))
    assert all(r.solves == 2 for r in reqs)
    # a request larger than the budget still runs when nothing else is resident
    big = [Request(0, 0.0, (lambda: _work(1, 1, log, 9)), est_bytes=1000)]
    Server(hbm_budget=100, executor=_FakeExecutor(), clock=lambda: t[0]).run(big)
    assert big[0].result == ["info"]
    p = percentiles([1.0, 2.0, 3.0, None])
    assert p["p50"] == 2.0


def test_serving_failed_request_does_not_stop_others():
    from easyfea_amd.serving import Request, Server

    def bad():
        yield _FakeJob(1)
        raise RuntimeError("solve failed")

    log = []
    reqs = [Request(0, 0.0, bad, 1), Request(1, 0.0, (lambda: _work(1, 2, log, 1)), 1)]
    Server(hbm_budget=10, executor=_FakeExecutor(), clock=lambda: 0.0).run(reqs)
    assert isinstance(reqs[0].error, RuntimeError) and reqs[1].error is None and log == [1]


def test_workload_fields_are_seeded_and_lognormal():
    from easyfea_amd import workloads

    z1 = workloads.gaussian_field((64, 64), 6.4, np.random.default_rng(3))
    z2 = workloads.gaussian_field((64, 64), 6.4, np.random.default_rng(3))
    assert np.array_equal(z1, z2)
    assert abs(z1.mean()) < 1e-12 and abs(z1.std() - 1) < 1e-12
    # neighbouring cells correlated, distant ones much less
    c1 = np.corrcoef(z1[:, :-1].ravel(), z1[:, 1:].ravel())[0, 1]
    assert c1 > 0.8
    E = workloads.lognormal(210000.0, 0.2, np.random.default_rng(0).standard_normal(200000))
    assert abs(E.mean() / 210000.0 - 1) < 3e-3 and abs(E.std() / E.mean() - 0.2) < 3e-3
    specs = workloads.serving_elastic_specs(n=20, qps=50.0)
    assert all(32 <= s.side <= 256 and s.E.shape == (s.side * s.side,) for s in specs)
    arr = np.array([s.arrival for s in specs])
    assert arr[0] == 0 and np.all(np.diff(arr) >= 0)
    pf = workloads.serving_phasefield_specs(n=8, steps=5)
    assert all(100 <= s.cells <= 200 and s.cells % 2 == 0 for s in pf)


def test_tetra_meshes():
    from easyfea_amd.mesh import tetra4_box, tetra10_from_tetra4

    m = tetra4_box(1.0, 2.0, 1.5, 3, 2, 4)
    assert m.Ne == 6 * 3 * 2 * 4 and m.Nn == 4 * 3 * 5
    X = m.coord[m.connect]
    vol = np.einsum("ei,ei->e", np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]), X[:, 3] - X[:, 0]) / 6
    assert np.all(vol > 0) and abs(vol.sum() - 3.0) < 1e-12
    # conforming: every interior triangle face is shared by exactly two tetrahedra
    faces = np.sort(np.concatenate([m.connect[:, [0, 1, 2]], m.connect[:, [0, 1, 3]], m.connect[:, [0, 2, 3]],
                                    m.connect[:, [1, 2, 3]]]), axis=1)
    _, cnt = np.unique(faces, axis=0, return_counts=True)
    assert cnt.max() == 2
    m10 = tetra10_from_tetra4(m)
    c = m10.connect
    for k, (a, b) in enumerate([(0, 1), (1, 2), (0, 2), (0, 3), (2, 3), (1, 3)]):
        assert np.allclose(m10.coord[c[:, 4 + k]], 0.5 * (m10.coord[c[:, a]] + m10.coord[c[:, b]]))
    g = O.Geometry("TETRA10", c, m10.coord, "rigi", 3)
    assert abs(g.wJ.sum() - 3.0) < 1e-12


def test_aggregation_covers_nodes_and_splits_crack_lips():
    from easyfea_amd.coarse import aggregate, coarse_pairs, prolongator

    m = ea.tri3_notched_square(40)
    cm = build_csr_map(m.connect, m.Nn, 2)
    ag = aggregate(m.coord, cm.node_ptr, cm.node_cols, 64)
    assert np.array_equal(np.sort(ag.agg_nodes), np.arange(m.Nn))
    sizes = np.diff(ag.agg_ptr)
    assert sizes.min() >= 16 and sizes.max() <= 64 + 64 // 4 * 4
    # every aggregate is connected in the node graph
    import scipy.sparse as sps
    from scipy.sparse.csgraph import connected_components

    rows = np.repeat(np.arange(m.Nn), np.diff(cm.node_ptr))
    G = sps.csr_matrix((np.ones(rows.size), (rows, cm.node_cols)), shape=(m.Nn, m.Nn))
    for a in range(ag.nagg):
        nodes = ag.agg_nodes[ag.agg_ptr[a]:ag.agg_ptr[a + 1]]
        assert connected_components(G[nodes][:, nodes], directed=False)[0] == 1
    # the two copies of a crack-lip node never share an aggregate
    y, x = m.coord[:, 1], m.coord[:, 0]
    lips = np.nonzero((np.abs(y - 0.5) < 1e-12) & (x < 0.5 - 1e-12))[0]
    pos = {tuple(np.round(m.coord[i, :2], 12)): [] for i in lips}
    for i in lips:
        pos[tuple(np.round(m.coord[i, :2], 12))].append(ag.agg[i])
    assert all(len(v) == 2 and v[0] != v[1] for v in pos.values())
    # rigid-body modes: K P = 0 for an unconstrained elastic operator
    geo = O.Geometry("TRI3", m.connect, m.coord, "rigi", 2)
    K = O.assemble_csr(O.K_elastic(geo, O.C_isotropic(1.0, 0.3, 2, True)), O.csr_map(m.connect, 2, 2 * m.Nn),
                       2 * m.Nn)
    P = prolongator(ag, 2, 2, np.zeros(2 * m.Nn, bool))
    # the x-translation mode summed over all aggregates is a global rigid mode
    t = np.zeros(ag.nagg * 3)
    t[0::3] = 1.0
    assert np.abs(K @ (P @ t)).max() < 1e-12
    cp = coarse_pairs(ag, cm.node_ptr, cm.node_cols)
    assert cp.pair_ptr[-1] == cm.node_cols.size
    keys = cp.blk_I.astype(np.int64) * ag.nagg + cp.blk_J
    assert np.all(np.diff(keys) > 0)


# ---------------------------------------------------------------- time schemes
@pytest.mark.parametrize("algo", ["parabolic", "newmark", "midpoint", "hht", "hht_newmark", "euler_implicit",
                                  "euler_explicit"])
def test_time_scheme_tables_reproduce_the_reference_rhs(algo):
    """The host's expanded (matrix x history vector) form of A and b equals
    the oracle's restatement of _Solver_Apply_Neumann, Rayleigh damping folded in."""
    from easyfea_amd.Solvers import AlgoType
    from easyfea_amd.time_schemes import expand, history_coefs, matrix_coefs

    rng = np.random.default_rng(11)
    n = 30
    Kd = rng.standard_normal((n, n))
    Md = rng.standard_normal((n, n))
    K, M = sp.csr_matrix(Kd @ Kd.T), sp.csr_matrix(Md @ Md.T + n * np.eye(n))
    cM, cK = 0.3, 0.02
    C = (cK * K + cM * M).tocsr()
    F, u, v, a = rng.standard_normal((4, n))
    params = (0.04, 0.6) if algo == "parabolic" else O.hyperbolic_params(0.04, algo, 0.27, 0.55, 0.2)
    ref_b = O.time_rhs(algo, params, K, C, M, F, u, v, a)
    damping = {"K": cK, "M": cM}
    bases = {"K": K, "M": M}
    b = F.copy()
    for name, (ku, kv, ka) in expand(history_coefs(AlgoType(algo), params), damping).items():
        b += bases[name] @ (ku * u + kv * v + ka * a)
    assert np.linalg.norm(b - ref_b) <= 1e-12 * np.linalg.norm(ref_b)
    A = sum(c * bases[nm] for nm, c in expand(dict(zip("KCM", matrix_coefs(AlgoType(algo), params))),
                                               damping).items())
    cK_, cC_, cM_ = O.time_coefs(algo, params)
    ref_A = (cK_ * K + cC_ * C + cM_ * M).toarray()
    assert np.abs(A.toarray() - ref_A).max() <= 1e-12 * np.abs(ref_A).max()
    # thermal form: C a base matrix of its own
    tb = expand(history_coefs(AlgoType(algo), params), {"C": 1.0})
    assert set(tb) <= {"K", "C", "M"}


def test_time_abi_argument_errors_without_gpu():
    lib = _lib.load()
    assert lib.efea_time_update(9, 10, 0.1, 0.25, 0.5, 0.0, None, None, None, None, None, None, None, None) == -1
    assert lib.efea_time_update(2, 10, 0.1, 0.25, 0.5, 0.0, None, None, None, None, None, None, None, None) == -1
    assert "null pointer" in _lib.last_error()
    one = (C.c_double * 1)()
    p = C.cast(one, C.c_void_p)
    assert lib.efea_time_update(2, 1, 0.0, 0.25, 0.5, 0.0, p, p, p, p, p, None, None, None) == -1
    assert "dt" in _lib.last_error()
    assert lib.efea_lincomb(-1, 1.0, None, 0.0, None, 0.0, None, 0.0, None, None, None) == -1
    assert lib.efea_uv_ke(1, 64, -1, None, None, None, 4, None, 0, 1.0, 2, 1.0, None, None) == -1


def test_thermal_model_parameters_and_observers():
    m = ea.Models.Thermal(1.0, 2.0, thickness=0.5)
    assert m.dim == 1 and m.k == 1.0 and m.c == 2.0 and not m.isHeterogeneous
    flags = []

    class Obs:
        def Need_Update(self, value=True):
            flags.append(value)

    m._Add_observer(Obs())
    m.k = np.array([1.0, 2.0])
    m.c = 0.0
    assert flags == [True, True] and m.isHeterogeneous
    with pytest.raises(AssertionError):
        m.k = -1.0


def test_bounded_lsq_abi_argument_errors_without_gpu():
    lib = _lib.load()
    one = (C.c_double * 1)()
    p = C.cast(one, C.c_void_p)
    need = lib.efea_bounded_lsq_workspace_bytes(100)
    assert need >= 6 * 800 + 100 and lib.efea_bounded_lsq_workspace_bytes(0) > 0
    assert lib.efea_solve_bounded_lsq(-1, p, p, p, p, p, p, p, p, need, 1e-10, 1e-9, 10, 10, None, None) == -1
    assert lib.efea_solve_bounded_lsq(100, p, p, p, p, p, None, p, p, need, 1e-10, 1e-9, 10, 10, None, None) == -1
    assert "null pointer" in _lib.last_error()
    assert lib.efea_solve_bounded_lsq(100, p, p, p, p, p, p, p, p, need - 1, 1e-10, 1e-9, 10, 10, None, None) == -4
    assert "workspace" in _lib.last_error()
    assert lib.efea_solve_bounded_lsq(100, p, p, p, p, p, p, p, p, need, 1.5, 1e-9, 10, 10, None, None) == -1
    assert lib.efea_solve_bounded_lsq(100, p, p, p, p, p, p, p, p, need, 1e-10, 1e-9, 10, 0, None, None) == -1
