"""Pins the CPU oracle against closed forms, invariants and the reference's
own known-answer tests (CPU only).

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails it.
"""

import numpy as np
import pytest
import scipy.sparse as sp

import easyfea_amd as ea
from oracle import efea_oracle as O

ELEMS2D = ["TRI3", "QUAD4"]
ELEMS3D = ["TETRA4", "TETRA10", "HEXA8"]
NODES = {
    "SEG2": [[-1], [1]],
    "TRI3": [[0, 0], [1, 0], [0, 1]],
    "QUAD4": [[-1, -1], [1, -1], [1, 1], [-1, 1]],
    "TETRA4": [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]],
    "TETRA10": [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [.5, 0, 0], [.5, .5, 0], [0, .5, 0], [0, 0, .5],
                [0, .5, .5], [.5, 0, .5]],
    "HEXA8": [[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]],
}
MEASURE = {"SEG2": 2.0, "TRI3": 0.5, "QUAD4": 4.0, "TETRA4": 1 / 6, "TETRA10": 1 / 6, "HEXA8": 8.0}


# ------------------------------------------------------------------ a1-a2
@pytest.mark.parametrize("et", list(NODES))
@pytest.mark.parametrize("mt", ["rigi", "mass"])
def test_gauss_weights_sum_to_measure(et, mt):
    _, w = O.gauss(et, mt)
    assert abs(w.sum() - MEASURE[et]) < 1e-14


def test_gauss_exactness():
    # degree-2-per-direction integrands are exact with the 2-point product rules
    p, w = O.gauss("QUAD4", "rigi")
    assert abs(np.sum(w * p[:, 0] ** 2 * p[:, 1] ** 2) - 4 / 9) < 1e-14
    p, w = O.gauss("HEXA8", "rigi")
    assert abs(np.sum(w * (p[:, 0] * p[:, 1] * p[:, 2]) ** 2) - 8 / 27) < 1e-14
    p, w = O.gauss("TRI3", "mass")  # degree 2 on the reference triangle
    assert abs(np.sum(w * p[:, 0] ** 2) - 1 / 12) < 1e-14
    assert abs(np.sum(w * p[:, 0] * p[:, 1]) - 1 / 24) < 1e-14
    p, w = O.gauss("TETRA10", "rigi")
    assert abs(np.sum(w * p[:, 0] ** 2) - 1 / 60) < 1e-14
    p, w = O.gauss("SEG2", "mass")
    assert abs(np.sum(w * p[:, 0] ** 3 + w * p[:, 0] ** 2) - 2 / 3) < 1e-14


@pytest.mark.parametrize("et", list(NODES))
def test_shape_functions_kronecker_partition_and_derivatives(et):
    nodes = np.array(NODES[et], dtype=float)
    N, _ = O.shape_functions(et, nodes)
    assert np.allclose(N, np.eye(len(nodes)), atol=1e-14)  # tests/FEM/groupElem_test.py:67-80
    rng = np.random.default_rng(0)
    pts = rng.random((5, nodes.shape[1])) * 0.3
    N, dN = O.shape_functions(et, pts)
    assert np.allclose(N.sum(1), 1.0, atol=1e-14)
    assert np.allclose(dN.sum(2), 0.0, atol=1e-13)
    h = 1e-6
    for k in range(nodes.shape[1]):
        e = np.zeros(nodes.shape[1])
        e[k] = h
        Np, _ = O.shape_functions(et, pts + e)
        Nm, _ = O.shape_functions(et, pts - e)
        assert np.allclose((Np - Nm) / (2 * h), dN[:, k, :], atol=1e-8)


def test_product_tables_match_oracle():
    from easyfea_amd.quadrature import N_pg, dN_pg, gauss_points

    for et in ["SEG2", "TRI3", "QUAD4", "TETRA4", "TETRA10", "HEXA8"]:
        for mt in ["rigi", "mass"]:
            p, w = O.gauss(et, mt)
            pp, ww = gauss_points(et, mt)
            assert np.allclose(p, pp, atol=1e-16) and np.allclose(w, ww, atol=1e-16)
            N, dN = O.shape_functions(et, p)
            assert np.allclose(N, N_pg(et, mt), atol=1e-15)
            assert np.allclose(dN, dN_pg(et, mt), atol=1e-14)


# ------------------------------------------------------------------ a3-a5
def test_det_inv_against_numpy():
    """tests/FEM/linalg_test.py:62-86."""
    rng = np.random.default_rng(1)
    for d in (2, 3):
        M = rng.standard_normal((20, 4, d, d))
        assert np.allclose(O._det(M), np.linalg.det(M), rtol=1e-12)
        assert np.allclose(O._inv(M), np.linalg.inv(M), rtol=1e-11, atol=1e-12)


def test_area_and_volume_from_wJ():
    """tests/FEM/mesh_test.py:135-149."""
    m = ea.quad4_rectangle(3.0, 2.0, 7, 5)
    for et, mesh, meas in [("QUAD4", m, 6.0), ("TRI3", ea.tri3_rectangle(3.0, 2.0, 7, 5), 6.0),
                           ("HEXA8", ea.hexa8_box(1.0, 2.0, 1.5, 3, 4, 2), 3.0)]:
        for mt in ("rigi", "mass"):
            g = O.Geometry(et, mesh.connect, mesh.coord, mt, mesh.dim)
            assert abs(g.wJ.sum() - meas) < 1e-12 * meas
    g = O.Geometry("SEG2", m.dict_groupElem[ea.ElemType.SEG2].connect, m.coord, "mass", 2)
    assert abs(g.wJ.sum() - 10.0) < 1e-12  # perimeter


@pytest.mark.parametrize("dim,ps", [(2, True), (2, False), (3, False)])
def test_isotropic_C_closed_form(dim, ps):
    """tests/Models/linear_elastic_test.py:92-133."""
    E, v = 210000.0, 0.3
    if dim == 2 and ps:
        Cv = E / (1 - v**2) * np.array([[1, v, 0], [v, 1, 0], [0, 0, (1 - v) / 2]])
    elif dim == 2:
        Cv = E / ((1 + v) * (1 - 2 * v)) * np.array([[1 - v, v, 0], [v, 1 - v, 0], [0, 0, (1 - 2 * v) / 2]])
    else:
        Cv = np.zeros((6, 6))
        Cv[:3, :3] = v
        np.fill_diagonal(Cv[:3, :3], 1 - v)
        Cv[3:, 3:] = np.eye(3) * (1 - 2 * v) / 2
        Cv *= E / ((1 + v) * (1 - 2 * v))
    r2 = np.sqrt(2)
    T = np.array([[1, 1, r2], [1, 1, r2], [r2, r2, 2]]) if dim == 2 else np.block(
        [[np.ones((3, 3)), r2 * np.ones((3, 3))], [r2 * np.ones((3, 3)), 2 * np.ones((3, 3))]])
    c = Cv * T
    C = O.C_isotropic(E, v, dim, ps)
    assert np.linalg.norm(c - C) / np.linalg.norm(c) < 1e-12
    # the product model gives the same tensor
    mat = ea.Models.Elastic.Isotropic(dim, E, v, planeStress=ps)
    assert np.linalg.norm(mat.C - C) / np.linalg.norm(C) < 1e-14


# ------------------------------------------------------------------ a6
def _rigid_modes(coord_e, dim):
    n = coord_e.shape[0]
    modes = []
    for d in range(dim):
        v = np.zeros((n, dim))
        v[:, d] = 1
        modes.append(v.ravel())
    x = coord_e
    if dim == 2:
        modes.append(np.stack([-x[:, 1], x[:, 0]], 1).ravel())
    else:
        for a, b in [(0, 1), (1, 2), (0, 2)]:
            v = np.zeros((n, 3))
            v[:, a], v[:, b] = -x[:, b], x[:, a]
            modes.append(v.ravel())
    return modes


@pytest.mark.parametrize("et", ELEMS2D + ELEMS3D)
def test_element_stiffness_invariants(et):
    rng = np.random.default_rng(2)
    nodes = np.array(NODES[et], dtype=float)
    dim = nodes.shape[1]
    X = np.zeros((len(nodes), 3))
    X[:, :dim] = nodes @ (np.eye(dim) + 0.2 * rng.standard_normal((dim, dim))) + 0.05 * rng.standard_normal(
        (len(nodes), dim))
    connect = np.arange(len(nodes))[None]
    g = O.Geometry(et, connect, X, "rigi", dim)
    K = O.K_elastic(g, O.C_isotropic(210000.0, 0.3, dim, True))[0]
    assert np.allclose(K, K.T, rtol=0, atol=1e-10 * np.abs(K).max())
    for v in _rigid_modes(X[:, :dim], dim):
        assert np.linalg.norm(K @ v) < 1e-10 * np.abs(K).max() * np.linalg.norm(v)
    ev = np.linalg.eigvalsh(K)
    nz = 3 if dim == 2 else 6
    assert np.all(ev[:nz] < 1e-8 * ev[-1]) and np.all(ev[nz:] > 1e-6 * ev[-1])


def test_quad4_unit_square_plane_stress_closed_form():
    """Textbook 8x8 stiffness of the unit square (2x2 Gauss), plane stress."""
    E, v, t = 1.0, 0.25, 1.0
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], dtype=float)
    g = O.Geometry("QUAD4", np.arange(4)[None], X, "rigi", 2)
    K = O.K_elastic(g, O.C_isotropic(E, v, 2, True))[0] * t
    k = np.array([1 / 2 - v / 6, 1 / 8 + v / 8, -1 / 4 - v / 12, -1 / 8 + 3 * v / 8,
                  -1 / 4 + v / 12, -1 / 8 - v / 8, v / 6, 1 / 8 - 3 * v / 8])
    idx = [[0, 1, 2, 3, 4, 5, 6, 7], [1, 0, 7, 6, 5, 4, 3, 2], [2, 7, 0, 5, 6, 3, 4, 1], [3, 6, 5, 0, 7, 2, 1, 4],
           [4, 5, 6, 7, 0, 1, 2, 3], [5, 4, 3, 2, 1, 0, 7, 6], [6, 3, 4, 1, 2, 7, 0, 5], [7, 2, 1, 4, 3, 6, 5, 0]]
    Kref = E * t / (1 - v**2) * k[np.array(idx)]
    assert np.max(np.abs(K - Kref)) < 1e-14


def test_tri3_cst_closed_form():
    """Constant-strain triangle: K = t A B^T D B with B from node coordinates."""
    X = np.array([[0.1, 0.2, 0], [1.3, 0.4, 0], [0.5, 1.7, 0]])
    E, v, t = 210000.0, 0.3, 2.0
    g = O.Geometry("TRI3", np.arange(3)[None], X, "rigi", 2)
    K = O.K_elastic(g, O.C_isotropic(E, v, 2, False))[0] * t
    x, y = X[:, 0], X[:, 1]
    A = 0.5 * ((x[1] - x[0]) * (y[2] - y[0]) - (x[2] - x[0]) * (y[1] - y[0]))
    b = np.array([y[1] - y[2], y[2] - y[0], y[0] - y[1]]) / (2 * A)
    c = np.array([x[2] - x[1], x[0] - x[2], x[1] - x[0]]) / (2 * A)
    B = np.zeros((3, 6))
    B[0, 0::2], B[1, 1::2], B[2, 0::2], B[2, 1::2] = b, c, c, b
    D = E / ((1 + v) * (1 - 2 * v)) * np.array([[1 - v, v, 0], [v, 1 - v, 0], [0, 0, (1 - 2 * v) / 2]])
    Kref = t * A * B.T @ D @ B
    assert np.max(np.abs(K - Kref)) < 1e-9 * np.abs(Kref).max()


@pytest.mark.parametrize("et", ["QUAD4", "HEXA8"])
def test_affine_element_matches_high_order_quadrature(et):
    """On a parallelogram / parallelepiped the 2-point rule is exact: compare
    with a 4-point-per-direction brute force of B^T C B."""
    rng = np.random.default_rng(4)
    nodes = np.array(NODES[et], dtype=float)
    dim = nodes.shape[1]
    Amap = np.eye(dim) + 0.3 * rng.standard_normal((dim, dim))
    X = np.zeros((len(nodes), 3))
    X[:, :dim] = nodes @ Amap.T
    C = O.C_isotropic(1.0, 0.3, dim, False)
    g = O.Geometry(et, np.arange(len(nodes))[None], X, "rigi", dim)
    K = O.K_elastic(g, C)[0]
    xg, wg = np.polynomial.legendre.leggauss(4)
    grids = np.meshgrid(*([xg] * dim), indexing="ij")
    pts = np.stack([q.ravel() for q in grids], 1)
    wts = np.prod(np.stack(np.meshgrid(*([wg] * dim), indexing="ij"), 0).reshape(dim, -1), 0)
    Kb = np.zeros_like(K)
    for p, w in zip(pts, wts):
        _, dN = O.shape_functions(et, p[None])
        F = dN[0] @ X[:, :dim]
        dNx = np.linalg.solve(F, dN[0])
        gg = O.Geometry.__new__(O.Geometry)
        gg.dN_x = dNx[None, None]
        B = O.Geometry.B(gg)[0, 0]
        Kb += w * abs(np.linalg.det(F)) * B.T @ C @ B
    assert np.max(np.abs(K - Kb)) < 1e-12 * np.abs(Kb).max()


def test_patch_test_linear_field_reproduced():
    """Distorted QUAD4 patch: a linear displacement imposed on the boundary is
    reproduced exactly inside (no loads)."""
    m = ea.quad4_rectangle(2.0, 1.0, 6, 4)
    rng = np.random.default_rng(5)
    c = m.coord.copy()
    inner = (c[:, 0] > 0) & (c[:, 0] < 2) & (c[:, 1] > 0) & (c[:, 1] < 1)
    c[inner, :2] += 0.08 * (rng.random((inner.sum(), 2)) - 0.5)
    sim = O.ElasticOracle(m.connect, c, "QUAD4", 1000.0, 0.25, False, 1.0)
    bnd = np.nonzero(~inner)[0]
    ux = lambda x, y, z: 1e-3 * (1 + 2 * x - y)
    uy = lambda x, y, z: 1e-3 * (0.5 * x + 3 * y)
    sim.add_dirichlet(bnd, [ux, uy], ["x", "y"])
    u = sim.Solve()
    exact = np.stack([ux(c[:, 0], c[:, 1], 0), uy(c[:, 0], c[:, 1], 0)], 1).ravel()
    assert np.max(np.abs(u - exact)) < 1e-13


# ------------------------------------------------------------------ a7-a9
def test_mass_diffusion_source_invariants():
    m = ea.tri3_rectangle(2.0, 1.5, 5, 4)
    for et, mesh in [("TRI3", m), ("QUAD4", ea.quad4_rectangle(2.0, 1.5, 5, 4))]:
        gm = O.Geometry(et, mesh.connect, mesh.coord, "mass", 2)
        gr = O.Geometry(et, mesh.connect, mesh.coord, "rigi", 2)
        M = O.UV(gm, 2.5)
        assert abs(M.sum() - 2.5 * 3.0) < 1e-12  # tests/FEM/mesh_test.py:151-166
        D = O.GradU_A_GradV(gr, np.eye(2), 0.7)
        assert np.max(np.abs(D.sum(2))) < 1e-12  # constants in the null space
        F = O.V(gm, 1.5)
        assert abs(F.sum() - 1.5 * 3.0) < 1e-12
        Mv = O.UV(gm, 1.0, dof_n=2)
        assert abs(Mv.sum() - 2 * 3.0) < 1e-12


# ------------------------------------------------------------------ a10-a12
def test_assembly_maps_brute_force():
    """tests/FEM/mesh_test.py:84-133."""
    m = ea.tri3_notched_square(4)
    connect = m.connect
    for dof_n in (1, 2, 3):
        a = O.assembly_e(connect, dof_n)
        brute = np.array([[n * dof_n + d for n in e for d in range(dof_n)] for e in connect])
        assert np.array_equal(a, brute)
        rows, cols = O.rows_cols_e(connect, dof_n)
        nd = a.shape[1]
        br = np.array([[a[e, i] for i in range(nd) for _ in range(nd)] for e in range(len(a))])
        bc = np.array([[a[e, j] for _ in range(nd) for j in range(nd)] for e in range(len(a))])
        assert np.array_equal(rows, br) and np.array_equal(cols, bc)


def test_csr_map_and_bincount_vs_dense_accumulation():
    m = ea.quad4_rectangle(2.0, 1.0, 4, 3)
    g = O.Geometry("QUAD4", m.connect, m.coord, "rigi", 2)
    Ke = O.K_elastic(g, O.C_isotropic(1.0, 0.3, 2, True))
    N = m.Nn * 2
    cmap = O.csr_map(m.connect, 2, N)
    K = O.assemble_csr(Ke, cmap, N)
    dense = np.zeros((N, N))
    touched = np.zeros((N, N), dtype=bool)
    a = O.assembly_e(m.connect, 2)
    for e in range(m.Ne):
        dense[np.ix_(a[e], a[e])] += Ke[e]
        touched[np.ix_(a[e], a[e])] = True
    assert np.max(np.abs(K.toarray() - dense)) < 1e-15 * np.abs(dense).max() * 8
    assert K.has_canonical_format
    ref = sp.csr_matrix(touched.astype(float))  # structural pattern (exact zeros stay stored)
    assert np.array_equal(K.indptr, ref.indptr) and np.array_equal(K.indices, ref.indices)


# ------------------------------------------------------------------ a13-a16
def test_neumann_resultants_and_cantilever():
    L, h = 120.0, 13.0
    m = ea.quad4_rectangle(L, h, 120, 13)
    sim = O.ElasticOracle(m.connect, m.coord, "QUAD4", 210000.0, 0.3, True, h,
                          {"SEG2": m.dict_groupElem[ea.ElemType.SEG2].connect})
    n0 = m.Nodes_Conditions(lambda x, y, z: x == 0)
    nL = m.Nodes_Conditions(lambda x, y, z: x == L)
    sim.add_dirichlet(n0, [0, 0], ["x", "y"])
    sim.add_surfLoad(nL, [-800 / h**2], ["y"])
    b = sim.neumann_vector("elastic", sim.Ndof)
    assert abs(b.sum() + 800.0) < 1e-10  # config 1 resultant
    assert abs(b[0::2].sum()) == 0.0
    u = sim.Solve()
    K = sim.Get_K()
    known, unk = O.known_unknown(sim.dir["elastic"][0], sim.Ndof)
    res = K[unk] @ u - b[unk]
    assert np.linalg.norm(res) < 1e-9 * np.linalg.norm(b)
    # Euler-Bernoulli deflection (examples/LinearizedElasticity/Elas1.py:41-42), shear adds ~1%
    I = h**4 / 12
    uy_an = 800 * L**3 / (3 * 210000.0 * I)
    assert 0.99 < -u[1::2].min() / uy_an < 1.03


# ------------------------------------------------------------------ a19-a23
def _strains(dim, n=300, seed=6):
    rng = np.random.default_rng(seed)
    e = rng.standard_normal((1, n, 3 if dim == 2 else 6)) * 1e-3
    return e


@pytest.mark.parametrize("split", ["Amor", "Miehe"])
@pytest.mark.parametrize("dim,ps", [(2, False), (2, True), (3, False)])
def test_split_invariants(split, dim, ps):
    """tests/Models/phasefield_test.py:118-137: C = cP + cM, sigma and psi add up."""
    eps = _strains(dim)
    cP, cM = O.calc_C(split, eps, 210000.0, 0.3, dim, ps)
    C = O.C_isotropic(210000.0, 0.3, dim, ps)
    assert np.max(np.linalg.norm(C - (cP + cM), axis=(-2, -1)) / np.linalg.norm(C)) < 1e-12
    psiP, psiM = O.calc_psi(split, eps, 210000.0, 0.3, dim, ps)
    psi = 0.5 * np.einsum("...i,ij,...j->...", eps, C, eps)
    assert np.max(np.abs(psi - psiP - psiM) / np.abs(psi)) < 1e-11
    # both parts of these splits are non-negative energies
    assert np.all(psiP >= -1e-12 * np.abs(psi)) and np.all(psiM >= -1e-12 * np.abs(psi))


@pytest.mark.parametrize("dim,ps", [(2, False), (2, True), (3, False)])
def test_miehe_closed_form(dim, ps):
    """Miehe 2010: psi+ = lam/2 <tr e>+^2 + mu sum <e_a>+^2."""
    eps = _strains(dim)
    lam, mu, _ = O.lame(210000.0, 0.3, dim, ps)
    M = O.vec_to_mat(eps)
    ev = np.linalg.eigvalsh(M)
    tr = ev.sum(-1)
    ref = lam / 2 * np.maximum(tr, 0) ** 2 + mu * np.sum(np.maximum(ev, 0) ** 2, -1)
    psiP, _ = O.calc_psi("Miehe", eps, 210000.0, 0.3, dim, ps)
    assert np.max(np.abs(psiP - ref) / np.abs(ref).max()) < 1e-12


@pytest.mark.parametrize("dim,ps", [(2, False), (2, True), (3, False)])
def test_amor_closed_form(dim, ps):
    """Amor 2009: psi+ = K/2 <tr e>+^2 + mu e_dev : e_dev."""
    eps = _strains(dim)
    _, mu, K = O.lame(210000.0, 0.3, dim, ps)
    M = O.vec_to_mat(eps)
    tr = np.trace(M, axis1=-2, axis2=-1)
    dev = M - tr[..., None, None] / dim * np.eye(dim)
    ref = K / 2 * np.maximum(tr, 0) ** 2 + mu * np.sum(dev * dev, axis=(-2, -1))
    psiP, _ = O.calc_psi("Amor", eps, 210000.0, 0.3, dim, ps)
    assert np.max(np.abs(psiP - ref) / np.abs(ref).max()) < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_eigen_decomposition_against_eigh(dim):
    """Models/_phasefield.py:970-1039 verification: eigenvalues vs eigh,
    reconstruction and partition of unity of the projectors."""
    eps = _strains(dim)
    M = O.vec_to_mat(eps)
    if dim == 2:
        val, _, Ms = O.eigen_2d(eps)
    else:
        val, _, Ms = O.eigen_3d(eps)
    ref = np.linalg.eigvalsh(M)
    assert np.max(np.linalg.norm(val - ref, axis=-1) / np.linalg.norm(ref, axis=-1)) < 1e-12
    rec = sum(val[..., a, None, None] * Ms[a] for a in range(dim))
    assert np.max(np.linalg.norm(rec - M, axis=(-2, -1)) / np.linalg.norm(M, axis=(-2, -1))) < 1e-12
    assert np.max(np.abs(sum(Ms) - np.eye(dim))) < 1e-12


@pytest.mark.parametrize("dim", [2, 3])
def test_spectral_projector_orthogonality(dim):
    eps = _strains(dim)
    P, Mm = O.spectral_projectors(eps, dim)
    eP = (P @ eps[..., None])[..., 0]
    eM = (Mm @ eps[..., None])[..., 0]
    nn = np.sum(eps * eps, -1)
    assert np.max(np.abs(np.sum(eP * eM, -1)) / nn) < 1e-12
    assert np.max(np.abs(eP + eM - eps)) < 1e-15


def test_split_special_cases():
    # pure tension: psi+ = psi; pure compression: Miehe psi+ = 0
    for dim in (2, 3):
        n = 3 if dim == 2 else 6
        e = np.zeros((1, 2, n))
        e[0, 0, :dim] = [1e-3, 2e-3, 3e-3][:dim]
        e[0, 1, :dim] = [-1e-3, -2e-3, -3e-3][:dim]
        C = O.C_isotropic(210000.0, 0.3, dim, False)
        psi = 0.5 * np.einsum("...i,ij,...j->...", e, C, e)
        pP, _ = O.calc_psi("Miehe", e, 210000.0, 0.3, dim, False)
        assert abs(pP[0, 0] - psi[0, 0]) < 1e-12 * psi[0, 0]
        assert abs(pP[0, 1]) < 1e-12 * psi[0, 1]


def _pos_neg(t, dim):
    """Positive / negative parts of Kelvin-Mandel tensors from numpy's eigh."""
    w, Q = np.linalg.eigh(O.vec_to_mat(t))
    pos = O.mat_to_vec((Q * np.maximum(w, 0)[..., None, :]) @ np.swapaxes(Q, -2, -1))
    neg = O.mat_to_vec((Q * np.minimum(w, 0)[..., None, :]) @ np.swapaxes(Q, -2, -1))
    return pos, neg


def _q(a, A, b):
    return np.einsum("...i,...ij,...j->...", a, np.broadcast_to(A, a.shape + (a.shape[-1],)), b)


@pytest.mark.parametrize("split", O.SPLITS)
@pytest.mark.parametrize("dim,ps", [(2, False), (2, True), (3, False)])
def test_every_split_adds_up_to_C(split, dim, ps):
    """c+ + c- = C for every split of ``PhaseField.SplitType`` (heterogeneous E)."""
    eps = _strains(dim, 120)
    E = 210000.0 * (1 + 0.2 * np.random.default_rng(1).random((1, 120)))
    cP, cM = O.calc_C(split, eps, E, 0.3, dim, ps)
    C = O._C_iso(E, 0.3, dim, ps)
    assert np.max(np.abs(cP + cM - C)) < 1e-11 * np.abs(C).max()


@pytest.mark.parametrize("dim,ps", [(2, False), (2, True), (3, False)])
def test_split_energies_in_closed_form(dim, ps):
    """psi+ of each split from eigh-based positive parts, an independent
    route: Bourdin 1/2 e.Ce; AnisotStrain 1/2 (e.Ce - e-.Ce-), its NoCross
    1/2 e+.Ce+ and PM/MP 1/2 (e+.Ce+ + e+.Ce-); the stress splits with
    s = Ce and S = C^-1: AnisotStress 1/2 (s.Ss - s-.Ss-), NoCross 1/2
    s+.Ss+, Zhang 1/2 e.s+, Stress 1/2 ((1+nu)/E s+.s+ - c <tr s>+^2) with
    c = nu(1+nu)/E in plane strain, nu/E in plane stress and 3D (the
    compliance of each model); He 1/2 |et+|^2 for et = C^(1/2) e."""
    eps = _strains(dim, 200)
    E, nu = 210000.0, 0.3
    C = O._C_iso(E, nu, dim, ps)
    S = np.linalg.inv(C)
    ep, em = _pos_neg(eps, dim)
    sig = (C @ eps[..., None])[..., 0]
    sp, sm = _pos_neg(sig, dim)
    w, Q = np.linalg.eigh(C)
    et = ((Q * np.sqrt(w)) @ Q.T @ eps[..., None])[..., 0]
    etp, _ = _pos_neg(et, dim)
    trs = sig[..., :dim].sum(-1)
    cst = nu * (1 + nu) / E if (dim == 2 and not ps) else nu / E
    ref = {
        "Bourdin": 0.5 * _q(eps, C, eps),
        "AnisotStrain": 0.5 * (_q(eps, C, eps) - _q(em, C, em)),
        "AnisotStrain_NoCross": 0.5 * _q(ep, C, ep),
        "AnisotStrain_PM": 0.5 * (_q(ep, C, ep) + _q(ep, C, em)),
        "AnisotStrain_MP": 0.5 * (_q(ep, C, ep) + _q(em, C, ep)),
        "AnisotStress": 0.5 * (_q(sig, S, sig) - _q(sm, S, sm)),
        "AnisotStress_NoCross": 0.5 * _q(sp, S, sp),
        "AnisotStress_PM": 0.5 * (_q(sp, S, sp) + _q(sp, S, sm)),
        "AnisotStress_MP": 0.5 * (_q(sp, S, sp) + _q(sm, S, sp)),
        "Zhang": 0.5 * np.sum(eps * sp, -1),
        "Stress": 0.5 * ((1 + nu) / E * np.sum(sp * sp, -1) - cst * np.maximum(trs, 0) ** 2),
        "He": 0.5 * np.sum(etp * etp, -1),
    }
    scale = np.abs(ref["Bourdin"]).max()
    for split, r in ref.items():
        psiP, psiM = O.calc_psi(split, eps, E, nu, dim, ps)
        assert np.max(np.abs(psiP - r)) < 1e-11 * scale, split
        assert np.max(np.abs(psiP + psiM - ref["Bourdin"])) < 1e-11 * scale, split


@pytest.mark.parametrize("dim,ps", [(2, False), (2, True), (3, False)])
def test_spectral_splits_in_pure_tension_and_compression(dim, ps):
    """All principal strains (and stresses) positive: c+ = C, c- = 0; all
    negative: c+ = 0, c- = C -- for every split built on spectral projectors."""
    n = 3 if dim == 2 else 6
    e = np.zeros((1, 2, n))
    e[0, 0, :dim] = [1e-3, 2e-3, 1.5e-3][:dim]
    e[0, 1, :dim] = [-1e-3, -2e-3, -1.5e-3][:dim]
    C = O._C_iso(210000.0, 0.3, dim, ps)
    for split in ("Miehe", "He", "Stress", "Zhang", "AnisotStrain", "AnisotStrain_NoCross", "AnisotStrain_PM",
                  "AnisotStress", "AnisotStress_NoCross", "AnisotStress_MP"):
        cP, cM = O.calc_C(split, e, 210000.0, 0.3, dim, ps)
        assert np.max(np.abs(cP[0, 0] - C)) < 1e-10 * np.abs(C).max(), split
        assert np.max(np.abs(cM[0, 0])) < 1e-10 * np.abs(C).max(), split
        assert np.max(np.abs(cP[0, 1])) < 1e-10 * np.abs(C).max(), split
        assert np.max(np.abs(cM[0, 1] - C)) < 1e-10 * np.abs(C).max(), split


def test_pf_coefficients():
    psi = np.array([[0.0, 1.0, 2.0]])
    assert np.allclose(O.pf_r(psi, 2.7, 0.01, "AT2"), 2 * psi + 270.0)
    assert np.allclose(O.pf_f(psi, 2.7, 0.01, "AT2"), 2 * psi)
    f1 = O.pf_f(psi, 2.7, 0.01, "AT1")
    assert np.allclose(f1, np.maximum(2 * psi - 3 * 2.7 / 0.08, 0))
    assert O.pf_k(2.7, 0.01, "AT2") == 2.7 * 0.01 and abs(O.pf_k(2.7, 0.01, "AT1") - 0.75 * 2.7 * 0.01) < 1e-18
    g = O.pf_g(np.zeros(4), np.array([[0, 1, 2]]), np.full((1, 3), 1 / 3))
    assert np.all(g == 1 + 1e-12)


def test_at2_damage_profile_closed_form():
    """AT2 with d = 1 imposed at x = 0 and no driving force: the damage
    problem -Gc l0 d'' + Gc/l0 d = 0 gives d = exp(-x/l0)."""
    l0, Lx = 0.05, 0.6
    m = ea.quad4_rectangle(Lx, 0.01, 600, 1)
    g_m = O.Geometry("QUAD4", m.connect, m.coord, "mass", 2)
    g_r = O.Geometry("QUAD4", m.connect, m.coord, "rigi", 2)
    Gc = 2.7
    Ke = O.UV(g_m, O.pf_r(np.zeros((m.Ne, 4)), Gc, l0, "AT2")) + O.GradU_A_GradV(g_r, np.eye(2), O.pf_k(Gc, l0, "AT2"))
    K = O.assemble_csr(Ke, O.csr_map(m.connect, 1, m.Nn), m.Nn)
    left = m.Nodes_Conditions(lambda x, y, z: x == 0)
    d = O.solve_r1(K, np.zeros(m.Nn), left, np.ones(left.size), m.Nn)
    x = m.coord[:, 0]
    ref = np.cosh((Lx - x) / l0) / np.cosh(Lx / l0)  # finite strip, zero flux at x = Lx
    assert np.max(np.abs(d - ref)) < 2e-4


def test_history_field_monotone_and_elastic_limit():
    """H never decreases; with d = 0 the first increment's reaction equals
    the linear elastic reaction (cross-path invariant)."""
    m = ea.tri3_notched_square(10)
    pf = O.PhaseFieldOracle(m.connect, m.coord, "TRI3", 210000.0, 0.3, False, 1.0, "Miehe", "AT2", 2.7, 0.05)
    top = m.Nodes_Conditions(lambda x, y, z: y == 1.0)
    bot = m.Nodes_Conditions(lambda x, y, z: y == 0.0)
    dofsY = O.dofs_nodes(top, ["y"], ["x", "y"])
    H_prev = pf.old_psiP.copy()
    for it, dep in enumerate([1e-3, 2e-3, 1e-3]):
        pf.Bc_Init()
        pf.add_dirichlet(top, [0, dep], ["x", "y"])
        pf.add_dirichlet(bot, [0], ["y"])
        pf.Solve(1.0, 10, 1)
        pf.Save_Iter()
        assert np.all(pf.old_psiP >= H_prev)
        H_prev = pf.old_psiP.copy()
    # elastic limit: first increment, d = 0 -> Ku equals the elastic K
    pf2 = O.PhaseFieldOracle(m.connect, m.coord, "TRI3", 210000.0, 0.3, False, 1.0, "Miehe", "AT2", 2.7, 0.05)
    pf2.assemble_elastic()
    el = O.ElasticOracle(m.connect, m.coord, "TRI3", 210000.0, 0.3, False, 1.0)
    Ku, Ke = pf2.Ku.toarray(), el.Get_K().toarray()
    # with u = 0 the strain is zero and Miehe returns cP + cM = C exactly up to g = 1 + 1e-12
    assert np.max(np.abs(Ku - Ke)) < 1e-11 * np.abs(Ke).max()
    el.add_dirichlet(top, [0, 1e-4], ["x", "y"])
    el.add_dirichlet(bot, [0], ["y"])
    ue = el.Solve()
    fe = (el.Get_K()[dofsY] @ ue).sum()
    pf2.add_dirichlet(top, [0, 1e-4], ["x", "y"])
    pf2.add_dirichlet(bot, [0], ["y"])
    pf2.Solve(1.0, 10, 1)
    fp = pf2.Calc_Reaction(dofsY).sum()
    assert abs(fp - fe) < 1e-9 * abs(fe)


# ------------------------------------------------ post-processing (8(f) rank 1)
def _distorted_mesh(et, seed=11):
    makers = {
        "TRI3": lambda: ea.tri3_rectangle(2.0, 1.0, 5, 4),
        "QUAD4": lambda: ea.quad4_rectangle(2.0, 1.0, 5, 4),
        "HEXA8": lambda: ea.hexa8_box(1.0, 1.5, 2.0, 3, 2, 3),
        "TETRA4": lambda: ea.mesh.tetra4_box(1.0, 1.5, 2.0, 2, 2, 3),
        "TETRA10": lambda: ea.mesh.tetra10_from_tetra4(ea.mesh.tetra4_box(1.0, 1.5, 2.0, 2, 2, 2)),
    }
    m = makers[et]()
    c = m.coord.copy()
    rng = np.random.default_rng(seed)
    dim = m.dim
    inner = np.ones(m.Nn, bool)
    for k in range(dim):
        inner &= (c[:, k] > c[:, k].min()) & (c[:, k] < c[:, k].max())
    c[inner, :dim] += 0.05 * (rng.random((inner.sum(), dim)) - 0.5)
    return m, c


@pytest.mark.parametrize("et", ELEMS2D + ELEMS3D)
def test_strain_stress_of_a_linear_field_are_the_closed_forms(et):
    """u = A x + b is reproduced exactly by every element (isoparametric
    patch), so eps = sym(A) at every point and sig = lam tr(eps) I + 2 mu eps;
    the von Mises measures and the energy follow in closed form."""
    m, c = _distorted_mesh(et)
    dim = m.dim
    rng = np.random.default_rng(3)
    A = 1e-3 * rng.standard_normal((dim, dim))
    u = (c[:, :dim] @ A.T + 1e-3).ravel()
    E, nu = 2.0e5 * np.exp(0.1 * rng.standard_normal(m.Ne)), 0.3
    geo = O.Geometry(et, m.connect, c, "rigi", dim)
    eps = O.strain_e_pg(geo, u, m.connect)
    sig = O.stress_e_pg(eps, O.C_isotropic(E, nu, dim, True))
    e_t = 0.5 * (A + A.T)
    lam, mu, _ = O.lame(E, nu, dim, True)
    s_t = lam[:, None, None] * np.trace(e_t) * np.eye(dim) + 2 * mu[:, None, None] * e_t
    names = ["xx", "yy", "xy"] if dim == 2 else ["xx", "yy", "zz", "yz", "xz", "xy"]
    idx = {"x": 0, "y": 1, "z": 2}
    for n in names:
        i, j = idx[n[0]], idx[n[1]]
        assert np.allclose(O.strain_or_stress_result_e(eps, "E" + n), e_t[i, j], rtol=0, atol=1e-15)
        assert np.allclose(O.strain_or_stress_result_e(sig, "S" + n), s_t[:, i, j], rtol=1e-12, atol=1e-12)

    def vm(t):
        d = t - np.trace(t, axis1=-2, axis2=-1)[..., None, None] * np.eye(dim) / 3 if dim == 3 else None
        if dim == 3:
            return np.sqrt(1.5 * np.sum(d * d, axis=(-2, -1)))
        xx, yy, xy = t[..., 0, 0], t[..., 1, 1], t[..., 0, 1]
        return np.sqrt(xx**2 + yy**2 - xx * yy + 3 * xy**2)

    assert np.allclose(O.strain_or_stress_result_e(sig, "Svm"), vm(s_t), rtol=1e-12)
    assert np.allclose(O.strain_or_stress_result_e(eps, "Evm"), vm(e_t), rtol=1e-12)
    full = O.strain_or_stress_result_e(eps, "Strain")
    assert full.shape == (m.Ne, len(names))
    # energy density of the constant field times the element volume
    vol = geo.wJ.sum(1)
    w = O.psi_elas_e(geo, eps, sig, thickness=2.0)
    dens = 0.5 * np.einsum("eij,ij->e", s_t, e_t)
    assert np.allclose(w, dens * vol * (2.0 if dim == 2 else 1.0), rtol=1e-12)


@pytest.mark.parametrize("et", ["QUAD4", "HEXA8", "TETRA10"])
def test_deformation_energy_equals_half_u_K_u(et):
    """Cross-path invariant: sum_e 1/2 int sig.eps = 1/2 u^T K u for any u."""
    m, c = _distorted_mesh(et)
    dim = m.dim
    rng = np.random.default_rng(8)
    u = rng.standard_normal(m.Nn * dim)
    E = 2.0e5 * np.exp(0.2 * rng.standard_normal(m.Ne))
    C = O.C_isotropic(E, 0.3, dim, False)
    geo = O.Geometry(et, m.connect, c, "rigi", dim)
    eps = O.strain_e_pg(geo, u, m.connect)
    W = O.psi_elas_e(geo, eps, O.stress_e_pg(eps, C)).sum()
    K = O.assemble_csr(O.K_elastic(geo, C), O.csr_map(m.connect, dim, m.Nn * dim), m.Nn * dim)
    assert abs(W - 0.5 * u @ (K @ u)) <= 1e-12 * abs(W)


def test_von_mises_special_states():
    """Uniaxial stress gives |s|, pure shear sqrt(3)|t| (2D and 3D), a
    hydrostatic 3D state 0; shear enters the Kelvin-Mandel field times sqrt 2."""
    s, t = 3.0, 2.0
    f2 = np.array([[[s, 0, 0], [0, 0, np.sqrt(2) * t]]])
    assert np.allclose(O.result_in_field(f2, "vm")[0], [abs(s), np.sqrt(3) * t])
    assert np.allclose(O.result_in_field(f2, "xy")[0], [0, t])
    f3 = np.array([[[s, 0, 0, 0, 0, 0], [0, 0, 0, 0, 0, np.sqrt(2) * t], [5, 5, 5, 0, 0, 0]]])
    assert np.allclose(O.result_in_field(f3, "vm")[0], [abs(s), np.sqrt(3) * t, 0])
    assert np.allclose(O.result_in_field(f3, "xy")[0], [0, t, 0])


def test_node_values_averages_surrounding_elements():
    """Two triangles sharing an edge, one unused node: shared nodes take the
    mean, the others their element's value, the unused node 0."""
    connect = np.array([[0, 1, 2], [1, 3, 2]])
    v = np.array([2.0, 6.0])
    out = O.node_values(connect, 5, v)
    assert np.array_equal(out, [2.0, 4.0, 4.0, 6.0, 0.0])
    two = O.node_values(connect, 5, np.stack([v, -v], 1))
    assert np.array_equal(two[:, 1], -out)
    m, _ = _distorted_mesh("HEXA8")
    assert np.allclose(O.node_values(m.connect, m.Nn, np.full(m.Ne, 7.5)), 7.5)


@pytest.mark.parametrize("split", ["Amor", "Miehe"])
def test_damaged_stress_limits(split):
    """d = 0: g = 1 + 1e-12 and c+ + c- = C, so the damaged stress is the
    elastic one; d = 1: only c- eps remains (plus 1e-12 c+ eps), which for a
    pure dilatation (all eigenvalues positive) vanishes for both splits."""
    m = ea.tri3_rectangle(2.0, 1.0, 5, 4)
    rng = np.random.default_rng(4)
    u = 1e-3 * rng.standard_normal(m.Nn * 2)
    E = 2e5 * np.exp(0.1 * rng.standard_normal(m.Ne))
    geo = O.Geometry("TRI3", m.connect, m.coord, "rigi", 2)
    s0 = O.pf_stress_e_pg(geo, u, np.zeros(m.Nn), m.connect, split, E, 0.3, False)
    se = O.stress_e_pg(O.strain_e_pg(geo, u, m.connect), O.C_isotropic(E, 0.3, 2, False))
    assert np.max(np.abs(s0 - se)) <= 2e-12 * np.max(np.abs(se))
    dil = (m.coord[:, :2] * 1e-3).ravel()  # u = 1e-3 x: eps = 1e-3 I
    s1 = O.pf_stress_e_pg(geo, dil, np.ones(m.Nn), m.connect, split, E, 0.3, False)
    s_el = O.stress_e_pg(O.strain_e_pg(geo, dil, m.connect), O.C_isotropic(E, 0.3, 2, False))
    assert np.allclose(s1, 1e-12 * s_el, rtol=1e-9, atol=0)



def _repeated_3d(diags):
    """3D strains diag(d) * 2^-10 with exactly repeated eigenvalues: binary
    fractions, so that the invariants and the Lode angle come out exact."""
    e = np.zeros((len(diags), 1, 6))
    for k, d in enumerate(diags):
        e[k, 0, :3] = np.array(d, dtype=float) * 2.0**-10
    return e



# Exactly repeated principal strains where the reference's 3D decomposition
# is the spectral one: all three equal (g = 0), or the two smallest equal and
# not positive (theta = 0; their half-split projectors then carry no energy).
_REPEATED_SPECTRAL = [(2, 2, 2), (-3, -3, -3), (-1, -1, 2), (0, 0, 4), (-4, -4, 0)]
# ... and where it is not: the two smallest equal and positive (theta = 0),
# or the two largest equal (theta = pi/3).
_REPEATED_REFERENCE_ONLY = [(1, 1, 4), (4, 1, 1), (4, 4, 1), (1, 1, -2), (1, 4, 4)]


def _miehe_closed(eps, dim=3):
    lam, mu, _ = O.lame(210000.0, 0.3, dim, False)
    ev = np.linalg.eigvalsh(O.vec_to_mat(eps))
    return lam / 2 * np.maximum(ev.sum(-1), 0) ** 2 + mu * np.sum(np.maximum(ev, 0) ** 2, -1)


def test_eigen_3d_exactly_repeated_eigenvalues():
    """The case masks of Models/_phasefield.py:832-906 where the general
    formula would divide by zero (g = 0; theta = 0 with the repeated pair
    not positive): eigenvalues, reconstruction, partition of unity and the
    spectral projectors against numpy's eigh, and the Miehe energy against
    its closed form."""
    eps = _repeated_3d(_REPEATED_SPECTRAL)
    M = O.vec_to_mat(eps)
    val, _, Ms = O.eigen_3d(eps)
    ref = np.linalg.eigvalsh(M)
    assert np.array_equal(val, ref)
    rec = sum(val[..., a, None, None] * Ms[a] for a in range(3))
    assert np.max(np.abs(rec - M)) <= 1e-15 * np.max(np.abs(M))
    assert np.max(np.abs(sum(Ms) - np.eye(3))) < 1e-15
    P, Mm = O.spectral_projectors(eps, 3)
    eP = (P @ eps[..., None])[..., 0]
    eM = (Mm @ eps[..., None])[..., 0]
    pos, neg = _pos_neg(eps, 3)
    scale = np.max(np.abs(eps))
    assert np.max(np.abs(eP - pos)) < 1e-15 * scale and np.max(np.abs(eM - neg)) < 1e-15 * scale
    closed = _miehe_closed(eps)
    psiP, psiM = O.calc_psi("Miehe", eps, 210000.0, 0.3, 3, False)
    assert np.max(np.abs(psiP - closed)) <= 1e-13 * np.max(closed)
    psi = 0.5 * np.einsum("...i,ij,...j->...", eps, O.C_isotropic(210000.0, 0.3, 3, False), eps)
    assert np.max(np.abs(psiP + psiM - psi)) <= 1e-13 * np.max(psi)


def test_eigen_3d_repeated_positive_or_two_largest_follow_the_reference():
    """Where the reference's closed form is not the spectral split.  Its
    eigenvalues are exact, but it splits a double eigenspace projector into
    normalised halves (theta = 0: M1 = (I - M3) / sqrt 2, then
    M2 = I - M1 - M3, :892, 942-948), which the 6x6 projector treats as two
    eigenprojectors, and for theta = pi/3 it builds the lone projector from
    I_rg = (I1 - sqrt g) / 3 I (:854, 870) where the eigenvalue shift is
    (I1 + sqrt g) / 3.  With the pair in tension (or the two largest equal)
    the Miehe energy then departs from the closed form.  The oracle keeps the
    reference's formulas (parity is with the reference; its package does not
    import here, gmsh being absent, so these values are pinned by reading
    only): this test records the exact eigenvalues, the partition of unity
    and that the departure is the reference's, not the oracle's rounding."""
    eps = _repeated_3d(_REPEATED_REFERENCE_ONLY)
    M = O.vec_to_mat(eps)
    val, _, Ms = O.eigen_3d(eps)
    assert np.array_equal(val, np.linalg.eigvalsh(M))
    assert np.max(np.abs(sum(Ms) - np.eye(3))) < 1e-15
    psiP, psiM = O.calc_psi("Miehe", eps, 210000.0, 0.3, 3, False)
    closed = _miehe_closed(eps)
    dev = np.abs(psiP - closed)[:, 0] / closed[:, 0]
    assert np.all(dev > 1e-3), dev  # a formula departure, not rounding
    psi = 0.5 * np.einsum("...i,ij,...j->...", eps, O.C_isotropic(210000.0, 0.3, 3, False), eps)
    assert np.max(np.abs(psiP + psiM - psi)) <= 1e-13 * np.max(psi)  # c+ + c- = C still


def test_eigen_3d_repeated_branches_meet_the_general_one():
    """A strain a relative 1e-6 away from one of the spectral repeated cases
    takes the general branch (distinct eigenvalues); its Miehe energy and
    stress differ from the repeated branch's by the perturbation's order:
    those masks join the general formula continuously."""
    eps = _repeated_3d(_REPEATED_SPECTRAL)
    pert = eps.copy()
    d = np.diagonal(O.vec_to_mat(eps)[:, 0], axis1=-2, axis2=-1)
    for k in range(eps.shape[0]):
        # the repeated pair (or the first two of a triple) drift apart
        i, j = [(a, b) for a in range(3) for b in range(a + 1, 3) if d[k, a] == d[k, b]][0]
        h = 1e-6 * np.max(np.abs(eps[k]))
        pert[k, 0, i] += h
        pert[k, 0, j] -= h
    val, _, _ = O.eigen_3d(pert)
    assert np.all(np.diff(val, axis=-1) > 0)  # distinct: the general branch
    C = O.C_isotropic(210000.0, 0.3, 3, False)
    p0, _ = O.calc_psi("Miehe", eps, 210000.0, 0.3, 3, False)
    p1, _ = O.calc_psi("Miehe", pert, 210000.0, 0.3, 3, False)
    psi = 0.5 * np.einsum("...i,ij,...j->...", eps, C, eps)
    assert np.max(np.abs(p1 - p0) / psi) < 1e-5
    c0, _ = O.calc_C("Miehe", eps, 210000.0, 0.3, 3, False)
    c1, _ = O.calc_C("Miehe", pert, 210000.0, 0.3, 3, False)
    s0 = (c0 @ eps[..., None])[..., 0]
    s1 = (c1 @ pert[..., None])[..., 0]
    assert np.max(np.abs(s1 - s0)) < 1e-4 * np.max(np.abs(s0))


def test_eigen_2d_repeated_eigenvalues():
    """2D at equal principal strains (the `v1 != v2` mask of
    Models/_phasefield.py:795-802): the projectors fall back to the axes,
    which are eigenprojectors of a multiple of I, and the split is spectral:
    eigh's positive part, Miehe's closed form, in tension and compression."""
    for ps in (False, True):
        e = np.zeros((4, 1, 3))
        e[:, 0, :2] = np.array([[1, 1], [-1, -1], [0, 0], [3, 3]], dtype=float) * 2.0**-10
        val, _, Ms = O.eigen_2d(e)
        assert np.array_equal(val, np.linalg.eigvalsh(O.vec_to_mat(e)))
        assert np.max(np.abs(sum(Ms) - np.eye(2))) == 0
        P, Mm = O.spectral_projectors(e, 2)
        eP = (P @ e[..., None])[..., 0]
        pos, _ = _pos_neg(e, 2)
        assert np.max(np.abs(eP - pos)) <= 1e-15 * np.max(np.abs(e))
        lam, mu, _ = O.lame(210000.0, 0.3, 2, ps)
        ev = np.linalg.eigvalsh(O.vec_to_mat(e))
        closed = lam / 2 * np.maximum(ev.sum(-1), 0) ** 2 + mu * np.sum(np.maximum(ev, 0) ** 2, -1)
        psiP, _ = O.calc_psi("Miehe", e, 210000.0, 0.3, 2, ps)
        assert np.max(np.abs(psiP - closed)) <= 1e-13 * np.max(closed)
