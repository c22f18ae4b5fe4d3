"""Time schemes as coefficients over the base matrices.

Every scheme of the reference solves, per step, ``A x = b`` with

    A = coefK K + coefC C + coefM M
    b = F + K w_K + C w_C + M w_M,   w_X = sum over s in (u_n, v_n, a_n) of T[X][s] s

(``_Solver_Get_K_C_M_coefs_for_time_scheme``, ``Simulations/_simu.py:1384-1436``;
``_Solver_Apply_Neumann``, ``_simu.py:1735-1860``), then updates the state
from ``x`` (``_Solver_Update_solutions``, ``_simu.py:1552-1640``, the
``efea_time_update`` kernel).  The reference applies each matrix to its
predictor vectors (``ut``, ``vt``) as sparse products; the table below is the
same sum with the predictors expanded into ``u_n, v_n, a_n`` so that the
device forms one history vector per base matrix and one SpMV each.

When ``C`` is itself a combination of base matrices (Rayleigh damping
``C = cK K + cM M``, ``Simulations/_elastic.py:146``), ``expand`` folds it in.
"""

from __future__ import annotations

from .Solvers import AlgoType


def matrix_coefs(algo: AlgoType, params: tuple) -> tuple:
    """(coefK, coefC, coefM) of ``A`` (``_simu.py:1384-1436``).  ``params``:
    ``(dt, alpha)`` for parabolic, ``(dt, beta, gamma, alpha)`` otherwise."""
    if algo == AlgoType.elliptic:
        return 1.0, 0.0, 0.0
    if algo == AlgoType.parabolic:
        dt, alpha = params
        return 1.0, 1 / (alpha * dt), 0.0
    dt, beta, gamma, alpha = params
    if algo == AlgoType.newmark:
        return 1.0, gamma / (beta * dt), 1 / (beta * dt ** 2)
    if algo == AlgoType.hht:
        return 1 - alpha, (1 - alpha) * gamma / (beta * dt), (1 - alpha) / (beta * dt ** 2)
    if algo == AlgoType.midpoint:
        return 0.5, 1 / dt, 2 / dt ** 2
    if algo == AlgoType.hht_newmark:
        return 1 - alpha, gamma / (beta * dt), 1 / (beta * dt ** 2)
    if algo == AlgoType.euler_implicit:
        return 1.0, 1 / dt, 1 / dt ** 2
    if algo == AlgoType.euler_explicit:
        return 0.0, 0.0, 1.0
    raise NotImplementedError(f"Algo {algo} is not implemented here.")


def history_coefs(algo: AlgoType, params: tuple) -> dict:
    """``{"K"|"C"|"M": (on u_n, on v_n, on a_n)}``: the right-hand side's
    history terms of ``_Solver_Apply_Neumann`` (``_simu.py:1773-1856``)."""
    z = (0.0, 0.0, 0.0)
    if algo == AlgoType.elliptic:
        return {"K": z, "C": z, "M": z}
    if algo == AlgoType.parabolic:
        # b += 1/(alpha dt) C (u_n + (1 - alpha) dt v_n)
        dt, alpha = params
        c = 1 / (alpha * dt)
        return {"K": z, "C": (c, c * (1 - alpha) * dt, 0.0), "M": z}
    dt, beta, gamma, alpha = params
    if algo in (AlgoType.newmark, AlgoType.hht_newmark):
        # b += (cC C + cM M) ut - C vt,  ut = u + dt v + dt^2/2 (1-2 beta) a,  vt = v + dt (1-gamma) a
        cC, cM = gamma / (beta * dt), 1 / (beta * dt ** 2)
        p = dt ** 2 / 2 * (1 - 2 * beta)
        C = (cC, cC * dt - 1.0, cC * p - dt * (1 - gamma))
        M = (cM, cM * dt, cM * p)
        # hht_newmark: only K is shifted, b -= alpha K u_n
        K = (-alpha, 0.0, 0.0) if algo == AlgoType.hht_newmark else z
        return {"K": K, "C": C, "M": M}
    if algo == AlgoType.midpoint:
        # b += (2/dt^2 M + 1/dt C - 1/2 K) u_n + 2/dt M v_n
        return {"K": (-0.5, 0.0, 0.0), "C": (1 / dt, 0.0, 0.0), "M": (2 / dt ** 2, 2 / dt, 0.0)}
    if algo == AlgoType.hht:
        cM, cC = 1 / (beta * dt ** 2), gamma / (beta * dt)
        cM2, cC2 = (alpha - 1) / (beta * dt), (alpha - 1) * (gamma / beta) + 1
        cM3, cC3 = (alpha - 1) / (2 * beta) + 1, dt * (alpha - 1) * (gamma / (2 * beta) - 1)
        return {"K": (-alpha, 0.0, 0.0),
                "C": (-(alpha - 1) * cC, -cC2, -cC3),
                "M": (-(alpha - 1) * cM, -cM2, -cM3)}
    if algo == AlgoType.euler_implicit:
        # b += (M/dt^2 + C/dt) u_n + M/dt v_n
        return {"K": z, "C": (1 / dt, 0.0, 0.0), "M": (1 / dt ** 2, 1 / dt, 0.0)}
    if algo == AlgoType.euler_explicit:
        # b = F - K u_n - C v_n  (the unknown is a^n)
        return {"K": (-1.0, 0.0, 0.0), "C": (0.0, -1.0, 0.0), "M": z}
    raise NotImplementedError(f"Algo {algo} is not implemented here.")


def expand(coefs: dict, damping: dict) -> dict:
    """Folds ``C = sum_X damping[X] X`` into the base matrices: ``coefs`` maps
    "K", "C", "M" to a scalar or a tuple of scalars; ``damping`` maps base
    names to their weight in ``C`` (``{"C": 1.0}`` when C is a base matrix of
    its own).  Returns ``{base: scalar or tuple}`` without all-zero entries."""
    out: dict = {}

    def add(name, val):
        if isinstance(val, tuple):
            prev = out.get(name, (0.0,) * len(val))
            out[name] = tuple(p + v for p, v in zip(prev, val))
        else:
            out[name] = out.get(name, 0.0) + val

    for name, val in coefs.items():
        if name == "C":
            for base, w in damping.items():
                add(base, tuple(w * x for x in val) if isinstance(val, tuple) else w * val)
        else:
            add(name, val)
    return {k: v for k, v in out.items() if (any(x != 0 for x in v) if isinstance(v, tuple) else v != 0)}
