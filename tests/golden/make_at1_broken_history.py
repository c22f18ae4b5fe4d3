"""Writes ``at1_broken_history.npz``: the oracle's FP64 history of an AT1
notched square driven until it breaks through (CPU only; about two minutes).

Workload: unit square with a horizontal notch from the left edge to the
centre, TRI3 mesh of 24 x 24 cells, E = 210000 MPa, nu = 0.3, plane strain,
thickness 1, Gc = 2.7 N/mm, l0 = 0.04 mm, Amor split, AT1 with the
BoundConstrain irreversibility (``Simulations/_phasefield.py:148-163``,
``Simulations/Solvers.py:376-383``; the oracle takes the exact minimiser of
the bounded least-squares problem); bottom edge uy = 0, top edge ux = 0 and
uy = n * 1e-3 mm for n = 1..18; one staggered pass per increment
(tolConv = 1, convOption = 1), Save_Iter after each, the reaction summed
over the top edge's y DOFs.  The force peaks at increment 9 and falls
below 3 N once the band of d = 1 crosses the square (increments 17 and 18,
86 and 160 nodes at d = 1): from there the displacement system keeps only
the 1e-12 residual stiffness across the band.  Sparse direct displacement
solves (SuperLU).

Stored: the force curve and u, d after every increment; and, in
``at1_damage_system_13.npz``, the bounded least-squares problem of
increment 13's damage solve (sparse A, b, lb, ub), where BVLS alone stops
8 % above the minimum cost.  Run from the repository root:

    python tests/golden/make_at1_broken_history.py
"""

import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from easyfea_amd.mesh import tri3_notched_square  # noqa: E402
from oracle import efea_oracle as O  # noqa: E402

CELLS, STEPS, DEP = 24, 18, 1e-3
E, NU, GC, L0 = 210000.0, 0.3, 2.7, 0.04
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "at1_broken_history.npz")
OUT13 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "at1_damage_system_13.npz")


def main():
    last = {}
    exact = O.bounded_lsq_exact

    def recorded(A, b, lb, ub):
        last.update(A=A.tocsr(), b=np.array(b), lb=np.array(lb), ub=np.array(ub))
        return exact(A, b, lb, ub)

    O.bounded_lsq_exact = recorded
    mesh = tri3_notched_square(CELLS)
    ref = O.PhaseFieldOracle(mesh.connect, mesh.coord, "TRI3", E, NU, False, 1.0, "Amor", "AT1", GC, L0,
                             solver="BoundConstrain")
    top = mesh.Nodes_Conditions(lambda x, y, z: y == 1.0)
    bot = mesh.Nodes_Conditions(lambda x, y, z: y == 0.0)
    dofsY = O.dofs_nodes(top, ["y"], ["x", "y"])
    forces, us, ds = [], [], []
    t0 = time.perf_counter()
    for it in range(1, STEPS + 1):
        ref.Bc_Init()
        ref.add_dirichlet(top, [0, it * DEP], ["x", "y"])
        ref.add_dirichlet(bot, [0], ["y"])
        u, d, _ = ref.Solve(1.0, 500, 1)
        ref.Save_Iter()
        forces.append(float(np.sum(ref.Calc_Reaction(dofsY))))
        us.append(u.copy())
        ds.append(d.copy())
        if it == 13:
            A = last["A"]
            np.savez_compressed(OUT13, data=A.data, indices=A.indices, indptr=A.indptr, shape=A.shape, b=last["b"],
                                lb=last["lb"], ub=last["ub"])
        print(f"increment {it}: force {forces[-1]:.6f} N, max d {d.max():.4f}, nodes at d = 1: "
              f"{int((d >= 1 - 1e-12).sum())}, {time.perf_counter() - t0:.1f} s", flush=True)
    np.savez_compressed(OUT, forces=np.array(forces), u=np.array(us), d=np.array(ds), cells=CELLS, dep=DEP,
                        l0=L0)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
