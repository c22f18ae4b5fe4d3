"""Pins the oracle's staggered phase-field loop (``PhaseFieldOracle.Solve``,
``Save_Iter``, ``Calc_Reaction``) by identities built from independently
pinned operators (CPU only).

For every pass of an increment, with ``H`` the history saved at the start of
the increment, the reference's caching order (``Simulations/_phasefield.py:
248-269, 303-432, 494-573, 613-630``) implies:

* ``d_np1`` solves ``Kd(max(psi+(u_n), H)) d = Fd(max(psi+(u_n), H))``;
* ``u_np1`` solves ``Ku(g(d_np1), split of eps(u_n)) u = f`` at the free DOFs;
* ``Calc_Reaction`` is ``Ku(g(d_last), eps(u_n of the last pass)) u``;
* ``Save_Iter`` sets ``H`` to the last assembled psi+: from the final ``u``
  with options 1 and 2 (their energies assemble ``Kd`` after the solve), from
  the last pass's ``u_n`` with options 0 and 3;
* options 1 and 2 converge on the relative change of an energy whose
  matrices are the cached ones at that moment.

The matrices here are assembled with scipy's COO summation from element
matrices of the pinned operators (``UV``, ``GradU_A_GradV``, ``V``,
``K_elastic``, ``calc_psi``, ``calc_C``), not through the oracle's own
``csr_map``; each identity is checked against a deliberately wrong variant
of the oracle (planted on the CPU) to show that it catches that slip.
"""

import numpy as np
import pytest
import scipy.sparse as sp

import easyfea_amd as ea
from oracle import efea_oracle as O

E, NU, GC, L0 = 210000.0, 0.3, 2.7, 0.05


def _dofs(connect, dof_n):
    return (connect[:, :, None] * dof_n + np.arange(dof_n)).reshape(connect.shape[0], -1)


def _coo(Ke, connect, dof_n, N):
    a = _dofs(connect, dof_n)
    nd = a.shape[1]
    rows = np.repeat(a, nd, axis=1).ravel()
    cols = np.tile(a, (1, nd)).ravel()
    return sp.coo_matrix((Ke.ravel(), (rows, cols)), shape=(N, N)).tocsr()


def _vec(Fe, connect, N):
    return np.bincount(connect.ravel(), weights=Fe.reshape(connect.shape[0], -1).ravel(), minlength=N)


class Checker:
    """Independent assembly of the staggered loop's systems for one mesh."""

    def __init__(self, pf):
        self.pf = pf
        self.c = pf.connect
        self.gm = O.Geometry(pf.elemType, pf.connect, pf.coord, "mass", pf.dim)
        self.gr = O.Geometry(pf.elemType, pf.connect, pf.coord, "rigi", pf.dim)
        self.Bm, self.Br = self.gm.B(), self.gr.B()

    def _eps(self, B, u):
        return np.einsum("epij,ej->epi", B, u[_dofs(self.c, self.pf.dim)])

    def psiP(self, u):
        pf = self.pf
        return O.calc_psi(pf.split, self._eps(self.Bm, u), pf.E, pf.v, pf.dim, pf.planeStress)[0]

    def Kd_Fd(self, psi):
        pf = self.pf
        r = O.pf_r(psi, pf.Gc, pf.l0, pf.regu)
        k = O.pf_k(pf.Gc, pf.l0, pf.regu)
        Ke = O.UV(self.gm, r) + O.GradU_A_GradV(self.gr, np.eye(pf.dim), k)
        Fe = O.V(self.gm, O.pf_f(psi, pf.Gc, pf.l0, pf.regu))
        return _coo(Ke * pf.thickness, self.c, 1, pf.Nn), _vec(Fe * pf.thickness, self.c, pf.Nn)

    def Ku(self, u_eps, d):
        pf = self.pf
        cP, cM = O.calc_C(pf.split, self._eps(self.Br, u_eps), pf.E, pf.v, pf.dim, pf.planeStress)
        g = O.pf_g(d, self.c, self.gr.N_pg)
        Ke = O.K_elastic(self.gr, g[..., None, None] * cP + cM) * pf.thickness
        return _coo(Ke, self.c, pf.dim, pf.Nn * pf.dim)


def _relres(A, x, b, rows=None):
    """Normwise residual scaled by ``|A| |x| + |b|`` (the componentwise
    backward-error scale)."""
    r = A @ x - b
    s = abs(A) @ np.abs(x) + np.abs(b)
    if rows is not None:
        r, s = r[rows], s[rows]
    return np.linalg.norm(r) / max(np.linalg.norm(s), 1e-300)


def check_increment(ch, H, Ku_cached, f, known, known_vals, convOption, tolConv, maxIter, fy_dofs=None):
    """Asserts the identities of one ``Solve`` (+ ``Calc_Reaction`` and
    ``Save_Iter``) of ``ch.pf`` from its trace; returns the ``Ku`` cached
    after the increment.  ``Ku_cached`` is the displacement matrix cached
    when the increment started (None if none was assembled yet)."""
    pf = ch.pf
    Ndof = pf.Nn * pf.dim
    free = np.setdiff1d(np.arange(Ndof), known)
    for k, (u_n, d_n, u_np1, d_np1, conv) in enumerate(pf.trace):
        Kd, Fd = ch.Kd_Fd(np.maximum(ch.psiP(u_n), H))
        assert _relres(Kd, d_np1, Fd) < 1e-10, f"pass {k}: damage does not solve Kd(psi+(u_n) v H) d = Fd"
        Ku = ch.Ku(u_n, d_np1)
        assert _relres(Ku, u_np1, f, free) < 1e-10, f"pass {k}: displacement does not solve Ku(d_np1, u_n) u = f"
        assert np.array_equal(u_np1[known], known_vals)
        if convOption in (1, 2):
            # energies with the matrices cached at each moment
            Kd_n = Kd
            Ku_n = Ku_cached if Ku_cached is not None else ch.Ku(u_n, d_n)

            def energy(u, d, Kd_, Ku_):
                e = 0.0 if np.linalg.norm(d) == 0 else 0.5 * d @ (Kd_ @ d)
                if convOption == 2:
                    e += 0.0 if np.linalg.norm(u) == 0 else 0.5 * u @ (Ku_ @ u)
                    e -= 0.0 if (np.linalg.norm(u) == 0 or np.linalg.norm(f) == 0) else u @ f
                return e

            E_n = energy(u_n, d_n, Kd_n, Ku_n)
            Kd_np1, _ = ch.Kd_Fd(np.maximum(ch.psiP(u_np1), H))
            E_np1 = energy(u_np1, d_np1, Kd_np1, Ku)
            ref = abs(E_n - E_np1) if E_np1 == 0 else abs((E_n - E_np1) / E_np1)
            assert abs(conv - ref) <= 1e-7 * ref + 1e-12, f"pass {k}: convergence energy {conv} vs {ref}"
        elif convOption == 0:
            assert conv == np.max(np.abs(d_np1 - d_n))
        Ku_cached = Ku
        last = k == len(pf.trace) - 1
        if tolConv == 1 or d_np1.max() == 0:
            stop = True
        elif convOption == 3:
            dU = np.abs(u_np1 - u_n)
            dU[u_np1 != 0] /= np.abs(u_np1[u_np1 != 0])
            dD = np.abs(d_np1 - d_n)
            dD[d_np1 != 0] /= np.abs(d_np1[d_np1 != 0])
            stop = dD.sum() <= tolConv and dU.sum() <= 0.999 * tolConv
        else:
            stop = conv <= tolConv
        assert stop == last or (last and k + 1 == maxIter), f"pass {k}: stopping rule"
    u_last_n, _, u_fin, d_fin, _ = pf.trace[-1]
    if fy_dofs is not None:
        ref = Ku_cached[fy_dofs] @ u_fin
        assert np.allclose(pf.Calc_Reaction(fy_dofs), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max()), \
            "Calc_Reaction is not the cached Ku (pre-solve strain, new damage) times u"
    pf.Save_Iter()
    src = u_fin if convOption in (1, 2) else u_last_n
    Hn = np.maximum(ch.psiP(src), H)
    assert np.allclose(pf.old_psiP, Hn, rtol=1e-12, atol=1e-14 * max(Hn.max(), 1e-300)), "history after Save_Iter"
    return Ku_cached


def _mesh():
    m = ea.tri3_notched_square(16)
    top = m.Nodes_Conditions(lambda x, y, z: y == 1.0)
    bot = m.Nodes_Conditions(lambda x, y, z: y == 0.0)
    return m, top, bot


def run_history(cls, split, convOption, tolConv, deps, maxIter=60, neumann=False):
    m, top, bot = _mesh()
    pf = cls(m.connect, m.coord, "TRI3", E, NU, False, 1.0, split, "AT2", GC, L0,
             boundary={"SEG2": m.dict_groupElem[ea.ElemType.SEG2].connect})
    ch = Checker(pf)
    Ndof = m.Nn * 2
    fy = O.dofs_nodes(top, ["y"], ["x", "y"])
    Ku = None
    npass = []
    for dep in deps:
        pf.Bc_Init()
        if neumann:
            # load control: traction on the top edge, bottom clamped in y, one corner in x
            pf.add_dirichlet(bot, [0], ["y"])
            pf.add_dirichlet(bot[:1], [0], ["x"])
            pf.add_surfLoad(top, [dep], ["y"])
        else:
            pf.add_dirichlet(top, [0, dep], ["x", "y"])
            pf.add_dirichlet(bot, [0], ["y"])
        dofs, vals = pf.dir["elastic"]
        known = np.unique(dofs)
        kv = np.bincount(dofs, weights=vals, minlength=Ndof)[known]
        f = pf.neumann_vector("elastic", Ndof)
        H = pf.old_psiP.copy()
        pf.Solve(tolConv, maxIter, convOption)
        npass.append(pf.Niter)
        Ku = check_increment(ch, H, Ku, f, known, kv, convOption, tolConv, maxIter, fy)
    return npass


DEPS = [3e-3, 5e-3, 6e-3, 7e-3]


@pytest.mark.parametrize("convOption", [0, 1, 2, 3])
@pytest.mark.parametrize("split", ["Miehe", "Amor"])
def test_staggered_identities(convOption, split):
    npass = run_history(O.PhaseFieldOracle, split, convOption, 1e-3, DEPS)
    assert max(npass) > 1, "the history must hold multi-pass increments"


def test_staggered_single_pass_identities():
    """tolConv = 1: one pass per increment (the workload of the SENT benchmark)."""
    npass = run_history(O.PhaseFieldOracle, "Miehe", 1, 1.0, DEPS + [8e-3])
    assert npass == [1] * 5


def test_option2_with_external_work():
    """Load control (Neumann traction): option 2's external-work term is live."""
    npass = run_history(O.PhaseFieldOracle, "Amor", 2, 1e-4, [120.0, 200.0, 240.0], neumann=True)
    assert max(npass) > 1


# ---------------------------------------------------------- planted slips
class _PostSolveReaction(O.PhaseFieldOracle):
    """Slip: the reaction from a Ku rebuilt with the post-solve strain."""

    def Calc_Reaction(self, dofs):
        self.updU = False
        return super().Calc_Reaction(dofs)


class _HistoryAtAssembly(O.PhaseFieldOracle):
    """Slip: the history advanced at every damage assembly instead of at Save_Iter."""

    def calc_psiP(self):
        p = super().calc_psiP()
        self.old_psiP = p
        return p


class _KuFromPostSolve(O.PhaseFieldOracle):
    """Slip: Ku assembled with the displacement of the previous pass's end
    even when the cache says it was built earlier (no pre-solve strain)."""

    def _solve(self, problemType):
        if problemType == "elastic":
            self.updU = False
        out = super()._solve(problemType)
        if problemType == "elastic":
            self.updU = False  # rebuilt from the solved u at the next request
        return out


class _NoExternalWork(O.PhaseFieldOracle):
    """Slip: option 2's energy without the external work."""

    def psi_ext(self, fu):
        return 0.0


class _HalfExternalWork(O.PhaseFieldOracle):
    """Slip: a 1/2 on the external work."""

    def psi_ext(self, fu):
        return 0.5 * super().psi_ext(fu)


class _DamageEnergyOnly(O.PhaseFieldOracle):
    """Slip: option 2 computed as option 1 (crack energy only)."""

    def _energy(self, convOption, fu):
        return super()._energy(1, fu)


@pytest.mark.parametrize("cls, convOption, neumann", [
    (_PostSolveReaction, 1, False),
    (_HistoryAtAssembly, 0, False),
    (_HistoryAtAssembly, 3, False),
    (_KuFromPostSolve, 1, False),
    (_NoExternalWork, 2, True),
    (_HalfExternalWork, 2, True),
    (_DamageEnergyOnly, 2, False),
])
def test_planted_slips_are_caught(cls, convOption, neumann):
    deps = [120.0, 200.0, 240.0] if neumann else DEPS
    tol = 1e-4 if neumann else 1e-3
    with pytest.raises(AssertionError):
        run_history(cls, "Amor" if neumann else "Miehe", convOption, tol, deps, neumann=neumann)


def test_history_damage_returns_the_maximum():
    """``HistoryDamage`` (:412-416): the returned damage is the node-wise
    maximum of the damage before and after the call; psi+ is not clamped."""
    m, top, bot = _mesh()
    pf = O.PhaseFieldOracle(m.connect, m.coord, "TRI3", E, NU, False, 1.0, "Miehe", "AT2", GC, L0,
                            solver="HistoryDamage")
    ch = Checker(pf)
    prev = np.zeros(m.Nn)
    # unload: the damage solved one increment later (from the unloaded strain)
    # falls, the returned one keeps the maximum
    for dep in [6e-3, 2e-3, 2e-3]:
        pf.Bc_Init()
        pf.add_dirichlet(top, [0, dep], ["x", "y"])
        pf.add_dirichlet(bot, [0], ["y"])
        u_n = pf.u.copy()
        _, d, _ = pf.Solve(1.0, 10, 1)
        Kd, Fd = ch.Kd_Fd(ch.psiP(u_n))  # no history clamp
        assert _relres(Kd, pf.d, Fd) < 1e-10
        assert np.array_equal(d, np.maximum(prev, pf.d))
        pf.Save_Iter()
        prev = d
    assert np.any(pf.d < prev)
