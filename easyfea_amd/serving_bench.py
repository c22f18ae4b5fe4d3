"""Serving measurements: a Poisson stream of elastic solves and a batch of
concurrent phase-field realisations through the continuous-batching loop.

``python -m easyfea_amd.serving_bench --workload elastic --requests 512 --qps 200``
``python -m easyfea_amd.serving_bench --workload phasefield --realisations 64 --steps 50``

Each prints one JSON line: solved DOF/s over the whole stream, request (or
per-increment) latency p50/p99, and, with ``--compare-fp64``, the same stream
in plain FP64 with the largest relative deviation of the mixed results.
Meshes are built and their sparsity patterns uploaded before the stream
starts (set-up, reported separately); everything per request - element
matrices, scatter, boundary conditions, solves - is inside the latency.
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import workloads
from .mesh import tri3_notched_square
from .serving import Request, Server, estimate_bytes, percentiles

GIB = 2**30
_MESHES = {}  # one mesh cache per generator, shared by the runs of a process


def _mesh_cache(maker):
    c = _MESHES.get(maker)
    if c is None:
        c = _MESHES[maker] = workloads.MeshCache(maker)
    return c


def _elastic_requests(specs, meshes, mode, max_blocks=0):
    reqs = []
    for spec in specs:
        m = meshes.get(spec.side)
        nnz = m.__dict__.setdefault("_nnz2", None) or 18 * m.Nn * 2
        reqs.append(Request(spec.rid, spec.arrival,
                            (lambda s=spec, m_=m: workloads.elastic_request(s, m_, mode, "auto", max_blocks)),
                            estimate_bytes(m.Nn, m.Ne, 4, (2,), (nnz,), mode == "mixed")))
    return reqs


def _prepare_meshes(meshes, keys, dof_n):
    """Builds each mesh, its device patterns and, where the solves will use
    the two-level preconditioner, its aggregates once (set-up)."""
    from .coarse import default_target
    from .device import DeviceMesh

    t = time.perf_counter()
    for k in sorted(set(keys)):
        m = meshes.get(k)
        for dn in dof_n:
            p = DeviceMesh(m).pattern(dn)
            if dn >= 2 and m.Nn >= 4 * default_target(m.dim):
                p.coarse()
    return time.perf_counter() - t


def run_elastic(args) -> dict:
    import torch

    specs = workloads.serving_elastic_specs(args.requests, args.qps, args.side_lo, args.side_hi, args.seed)
    meshes = _mesh_cache(workloads.square_quad4)
    setup_s = _prepare_meshes(meshes, [s.side for s in specs], (2,))
    # warm-up: the whole stream once (code objects, pinned buffers, and the
    # device allocator's cache of every buffer size the stream uses)
    for mode in ("fp64", "mixed"):
        Server(args.hbm_gib * GIB, capacity=args.capacity).run(_elastic_requests(specs, meshes, mode,
                                                                                 args.max_blocks))
    torch.cuda.synchronize()

    def stream(mode):
        reqs = _elastic_requests(specs, meshes, mode, args.max_blocks)
        srv = Server(args.hbm_gib * GIB, capacity=args.capacity, check_every=args.check_every)
        t0 = time.perf_counter()
        srv.run(reqs)
        wall = time.perf_counter() - t0
        srv.close()
        return reqs, wall, srv

    reqs, wall, srv = stream(args.mode)
    errors = [r.error for r in reqs if r.error is not None]
    ndof = sum(2 * (s.side + 1) ** 2 for s in specs)
    lat = percentiles([r.latency for r in reqs])
    its = [r.result[1].iterations for r in reqs if r.result is not None]
    out = {
        "workload": f"serving stream: {len(specs)} QUAD4 plane-stress squares, side U{{{args.side_lo}..{args.side_hi}}}, "
                    f"correlated lognormal E (CoV 0.2), Poisson arrivals at {args.qps} QPS",
        "mode": args.mode,
        "max_blocks": args.max_blocks,
        "value": ndof / wall,
        "unit": "DOF/s",
        "wall_s": wall,
        "requests": len(specs),
        "failed": len(errors),
        "errors": [repr(e)[:200] for e in errors[:3]],
        "offered_qps": args.qps,
        "achieved_qps": len(specs) / wall,
        "p50_latency_ms": lat["p50"] * 1e3,
        "p99_latency_ms": lat["p99"] * 1e3,
        "max_concurrent_solves": srv.max_active,
        "pcg_iters_mean": float(np.mean(its)) if its else None,
        "setup_s": setup_s,
    }
    if args.compare_fp64 and args.mode == "mixed":
        reqs64, wall64, _ = stream("fp64")
        dev = []
        for a, b in zip(reqs, reqs64):
            if a.result is None or b.result is None:
                continue
            ua, ub = a.result[0], b.result[0]
            dev.append(float(torch.linalg.vector_norm(ua - ub) / torch.linalg.vector_norm(ub)))
        out["fp64"] = {"value": ndof / wall64, "wall_s": wall64,
                       "p50_latency_ms": percentiles([r.latency for r in reqs64])["p50"] * 1e3,
                       "p99_latency_ms": percentiles([r.latency for r in reqs64])["p99"] * 1e3}
        out["u_rel_err_vs_fp64_max"] = max(dev) if dev else None
        out["mixed_speedup"] = wall64 / wall
    out["_specs"] = specs
    out["_meshes"] = meshes
    return out


def run_phasefield(args) -> dict:
    import torch

    specs = workloads.serving_phasefield_specs(args.realisations, args.steps, seed=args.seed)
    meshes = _mesh_cache(tri3_notched_square)
    setup_s = _prepare_meshes(meshes, [s.cells for s in specs], (2, 1))
    clock = time.perf_counter

    def requests(specs_, mode):
        reqs = []
        for s in specs_:
            m = meshes.get(s.cells)
            est = estimate_bytes(m.Nn, m.Ne, 3, (2, 1), (28 * m.Nn, 7 * m.Nn), mode == "mixed")
            est += s.steps * m.Nn * 3 * 8 * 2  # saved iterations (u, d per increment)
            reqs.append(Request(s.rid, 0.0, (lambda s_=s, m_=m: workloads.phasefield_request(s_, m_, mode, clock,
                                                                                             args.max_blocks)),
                                est))
        return reqs

    def fresh(specs_):
        return [workloads.PhaseFieldRequestSpec(s.rid, s.cells, s.Gc, 0.0, min(s.steps, 2)) if warm else
                workloads.PhaseFieldRequestSpec(s.rid, s.cells, s.Gc, 0.0, s.steps) for s in specs_]

    # warm-up: two increments of every realisation in both modes (code objects,
    # pinned buffers, the device allocator's cache)
    warm = True
    for mode in ("fp64", "mixed"):
        Server(args.hbm_gib * GIB, capacity=args.capacity).run(requests(fresh(specs), mode))
    torch.cuda.synchronize()
    warm = False

    def stream(specs_, mode):
        srv = Server(args.hbm_gib * GIB, capacity=args.capacity, check_every=args.check_every)
        reqs = requests(specs_, mode)
        t0 = time.perf_counter()
        srv.run(reqs)
        wall = time.perf_counter() - t0
        srv.close()
        return reqs, wall, srv

    run_specs = fresh(specs)
    reqs, wall, srv = stream(run_specs, args.mode)
    errors = [r.error for r in reqs if r.error is not None]
    ndof = sum((2 + 1) * meshes.get(s.cells).Nn * s.steps for s in run_specs)
    step_lat = percentiles([x for s in run_specs for x in s.step_latency])
    out = {
        "workload": f"{len(specs)} concurrent phase-field realisations, notched square TRI3, Ne U[20k, 80k], "
                    f"lognormal Gc field (CoV 0.15), Amor/AT2, {args.steps} x 1e-4 mm increments",
        "mode": args.mode,
        "max_blocks": args.max_blocks,
        "value": ndof / wall,
        "unit": "DOF/s",
        "wall_s": wall,
        "failed": len(errors),
        "errors": [repr(e)[:200] for e in errors[:3]],
        "staggered_iterations_per_s": sum(len(s.step_latency) for s in run_specs) / wall,
        "p50_step_ms": step_lat["p50"] * 1e3,
        "p99_step_ms": step_lat["p99"] * 1e3,
        "max_concurrent_solves": srv.max_active,
        "setup_s": setup_s,
    }
    if args.compare_fp64 and args.mode == "mixed":
        specs64 = fresh(specs)
        reqs64, wall64, _ = stream(specs64, "fp64")
        fdev = 0.0
        for a, b in zip(run_specs, specs64):
            fa, fb = np.array(a.forces), np.array(b.forces)
            if fa.size and fb.size == fa.size:
                fdev = max(fdev, float(np.max(np.abs(fa - fb) / np.maximum(np.abs(fb), 1e-300))))
        out["fp64"] = {"value": ndof / wall64, "wall_s": wall64}
        out["force_curve_max_rel_dev_vs_fp64"] = fdev
        out["mixed_speedup"] = wall64 / wall
    out["_specs"] = specs
    return out


def main(argv=None, quiet: bool = False):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", choices=["elastic", "phasefield"], default="elastic")
    ap.add_argument("--mode", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--requests", type=int, default=512)
    ap.add_argument("--qps", type=float, default=200.0)
    ap.add_argument("--side-lo", type=int, default=32)
    ap.add_argument("--side-hi", type=int, default=256)
    ap.add_argument("--realisations", type=int, default=64)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--capacity", type=int, default=256)
    ap.add_argument("--check-every", type=int, default=32)
    ap.add_argument("--hbm-gib", type=float, default=200.0)
    ap.add_argument("--max-blocks", type=int, default=None,
                    help="grid cap per system of the solver kernels (0: the library default, 1024); default: "
                         "64 for the phase-field batch (throughput), 0 for the elastic stream (latency)")
    ap.add_argument("--compare-fp64", action="store_true")
    ap.add_argument("--adaptive", action="store_true",
                    help="keep FP64 Krylov products for matrices under 128 MiB (single-solve policy)")
    args = ap.parse_args(argv)
    if args.max_blocks is None:
        args.max_blocks = 64 if args.workload == "phasefield" else 0
    import torch

    from . import device

    assert torch.cuda.is_available(), "serving_bench needs a GPU"
    if not args.adaptive:
        # a batch streams many matrices per iteration: FP32 values pay at any size
        device.MIXED_MIN_MATRIX_BYTES = 0
    out = run_elastic(args) if args.workload == "elastic" else run_phasefield(args)
    if not quiet:
        print(json.dumps({k: v for k, v in out.items() if not k.startswith("_")}), flush=True)
    return out


if __name__ == "__main__":
    sys.exit(0 if main() else 1)
