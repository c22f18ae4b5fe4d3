"""Time-stepping throughput: Newmark structural dynamics of the 64^3 HEXA8
cube and Crank-Nicolson heat conduction on a 1024^2 QUAD4 square.

``python -m easyfea_amd.dynamics_bench --workload newmark-cube --cells 64``
``python -m easyfea_amd.dynamics_bench --workload thermal-square --cells 1024``

Each ``Solve()`` is one time step on the device: the scheme matrix
``A = coefK K + coefC C + coefM M`` (formed once; kept while dt and the
matrices stay), the right-hand side's history products, the PCG solve and
the state update.  The first steps (assembly of K and M or C, code objects,
workspaces) are untimed; ``steps`` further steps are timed one by one.
Reported: DOF x steps per second over the timed steps, p50/p99 step time,
PCG iterations per step.
"""

from __future__ import annotations

import argparse
import json
import time

import numpy as np


def build(workload: str, cells: int, solver: str):
    from . import workloads

    if workload == "newmark-cube":
        return workloads.newmark_cube(cells, 0, solver), "elastic"
    if workload == "thermal-square":
        return workloads.thermal_square(cells, 0, solver), "thermal"
    raise ValueError(workload)


def run(workload: str, cells: int, steps: int, warmup: int = 2, solver: str = "hip_mixed", precond: str = "auto",
        extrapolate: bool = True) -> dict:
    import torch

    simu, pt = build(workload, cells, solver)
    simu.preconditioner = precond
    simu.extrapolate_guess = extrapolate
    for _ in range(warmup):
        simu.Solve()
    torch.cuda.synchronize()
    lat, its = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        t = time.perf_counter()
        simu.Solve()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t)
        its.append(simu.last_solve[pt].iterations)
    total = time.perf_counter() - t0
    ndof = simu.mesh.Nn * simu.Get_dof_n()
    info = simu.last_solve[pt]
    state = simu._u.detach().cpu().numpy()
    return dict(
        value=ndof * steps / total, unit="DOF*steps/s", ndof=ndof, steps=steps, warmup=warmup,
        ms_per_step=1e3 * total / steps, p50_ms=float(np.percentile(lat, 50) * 1e3),
        p99_ms=float(np.percentile(lat, 99) * 1e3), iterations_mean=float(np.mean(its)),
        iterations_max=int(np.max(its)), mode=info.mode, precond=info.precond, algo=str(simu.algo),
        params=list(simu._time_params()), state_norm=float(np.linalg.norm(state)), solver=solver)


def main(argv=None) -> dict:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", default="newmark-cube", choices=["newmark-cube", "thermal-square"])
    ap.add_argument("--cells", type=int, default=None)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--solver", default="hip_mixed", choices=["hip_fp64", "hip_mixed"])
    ap.add_argument("--precond", default="auto", choices=["auto", "bj", "two-level"])
    ap.add_argument("--no-extrapolate", action="store_true", help="start each solve from the previous state")
    args = ap.parse_args(argv)
    cells = args.cells or (64 if args.workload == "newmark-cube" else 1024)
    out = run(args.workload, cells, args.steps, args.warmup, args.solver, args.precond, not args.no_extrapolate)
    out.update(precond_arg=args.precond, extrapolate=not args.no_extrapolate)
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
