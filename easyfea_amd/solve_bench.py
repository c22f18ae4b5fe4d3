"""Single elastic solves: the 360 x 39 QUAD4 cantilever and the 64^3 HEXA8
cube with lognormal Young's modulus.

``python -m easyfea_amd.solve_bench --workload cantilever``
``python -m easyfea_amd.solve_bench --workload cube --cells 64``

One ``Solve()`` is timed from the device-resident mesh (element matrices,
scatter, boundary conditions, block-Jacobi / coarse set-up, PCG and the
final FP64 residual); the sparsity pattern, scatter map and aggregates are
built once before timing (mesh set-up).  Reported per (precision,
preconditioner): solved DOF/s, p50/p99 over repeats, PCG iterations, the
relative error against the FP64 reference and the CPU baseline.

The cube's error reference is an FP64 two-level PCG solve driven to a
relative residual of 1e-12 (a 3D sparse direct solve of this size is out of
reach on the host); ``bench.py --workload cantilever|cube`` adds the CPU
baseline and, for the cantilever, the error against the host direct solve.
"""

from __future__ import annotations

import argparse
import json
import time

import numpy as np


def _timed_solves(simu, repeats):
    import torch

    lat = []
    for _ in range(repeats):
        simu._u.zero_()
        simu.Need_Update()
        torch.cuda.synchronize()
        t = time.perf_counter()
        simu.Solve()
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t)
    return np.array(lat)


def run(workload: str, cells: int, repeats: int, coarse_target=None, on_run=None) -> dict:
    """Times every (preconditioner, precision) pair; ``on_run(simu, precond,
    solver)`` is called after each pair's timed solves (the simulation still
    holds its matrix and last solve)."""
    import torch

    from . import workloads

    make = (lambda solver: workloads.cantilever(solver)) if workload == "cantilever" else \
        (lambda solver: workloads.cube(cells, 0, solver))
    results = {}
    ref_u = None
    for precond in ("bj", "two-level"):
        for solver in ("hip_fp64", "hip_mixed"):
            t0 = time.perf_counter()
            simu = make(solver)
            simu.preconditioner = precond
            simu.coarse_reuse = False  # every timed solve builds its own coarse inverse
            simu.coarse_target = coarse_target
            dev = simu.dev
            A = dev.csr_map(simu.model.dim)
            if precond == "two-level":
                A.coarse(coarse_target)
            setup_s = time.perf_counter() - t0
            simu.Solve()  # warm-up: code objects, workspaces
            lat = _timed_solves(simu, repeats)
            if on_run is not None:
                on_run(simu, precond, solver)
            info = simu.last_solve["elastic"]
            u = simu._u.clone()
            ndof = simu.mesh.Nn * simu.model.dim
            results[(precond, solver)] = dict(
                value=ndof / lat.min(), p50_ms=float(np.percentile(lat, 50) * 1e3),
                p99_ms=float(np.percentile(lat, 99) * 1e3), iterations=info.iterations,
                rel_residual=info.rel_residual, replacements=info.replacements, mode=info.mode,
                fallback=info.fallback, setup_s=setup_s, u=u, ndof=ndof,
                coarse_unknowns=A.coarse(coarse_target).nagg * (6 if simu.model.dim == 3 else 3)
                if precond == "two-level" else None)
            del simu
            torch.cuda.empty_cache()
    ref_kind, ref_u = None, None
    if workload == "cube":
        simu = make("hip_fp64")
        simu.preconditioner = "two-level"
        simu.tol_anorm = 0.0
        simu.tol_residual = 1e-12
        simu.maxit = 20000
        try:
            simu.Solve()
        except Exception:  # residual floor above 1e-12: keep the last iterate
            pass
        ref_u = simu._u.cpu().numpy()
        info = simu.last_solve["elastic"]
        ref_kind = (f"FP64 two-level PCG driven to relative residual {info.rel_residual:.1e} "
                    f"({info.iterations} iterations)")
        del simu
    out = {"workload": workload, "cells": cells if workload == "cube" else None, "reference": ref_kind, "runs": []}
    for (precond, solver), r in results.items():
        u = r.pop("u").cpu().numpy()
        if ref_u is not None:
            r["u_rel_err_vs_reference"] = float(np.linalg.norm(u - ref_u) / np.linalg.norm(ref_u))
        r["precond"] = precond
        r["solver"] = solver
        r["_u"] = u
        out["runs"].append(r)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", choices=["cantilever", "cube"], default="cantilever")
    ap.add_argument("--cells", type=int, default=64)
    ap.add_argument("--repeats", type=int, default=5)
    ap.add_argument("--coarse-target", type=int, default=None,
                    help="largest aggregate of the coarse space (default: 256 nodes in 2D, 1024 in 3D)")
    args = ap.parse_args(argv)
    out = run(args.workload, args.cells, args.repeats, args.coarse_target)
    for r in out["runs"]:
        r.pop("_u")
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
