#!/usr/bin/env python
"""Benchmark: EasyFEA's phase-field fracture load history on one MI355X.

Workload (the headline): the single-edge-notched unit square, TRI3 mesh with
200 x 200 cells (80,000 elements, 40,501 nodes, 81,002 displacement + 40,501
damage DOFs), E = 210000 MPa, nu = 0.3, plane strain, Gc = 2.7 N/mm,
l0 = 0.01 mm, Miehe split, AT2, History irreversibility, staggered scheme
with one pass per increment (tolConv = 1, convOption = 1), 100 tensile
increments of 1e-4 mm on the top edge (bottom edge uy = 0, top edge ux = 0).

A step is one whole load history: a fresh simulation (device maps built
beforehand, outside the timed region) runs the 100 increments - damage
assembly and solve, displacement assembly and solve, history update and the
top-edge reaction per increment.  ``value`` is the solved DOF/s of the mixed
path (FP32 matrix values in every Krylov product, FP64 vectors with FP64
residual replacement; the FP32 copy is forced at this size), over the timed
histories; the plain FP64 device path runs the same histories beside it.
Accuracy: both paths against the oracle's sparse-direct FP64 history
(``tests/golden/sent200_history.npz``: force curve, u and d at increments
1, 25, 50, 75, 100) and against each other.  The CPU baseline is the
numpy/scipy oracle on the first increments of the same history.

Unless ``--no-extra``, the line also carries the measurement lines of the
other workloads: the 360 x 39 cantilever, the 64^3 HEXA8 cube, the elastic
serving stream and the 64 phase-field realisations, each with its roofline
and CPU baseline (``--workload NAME`` prints one of them alone).
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solved DOF/s at ≤1e-8 rel err vs FP64 (mixed vs plain); p50/p99 solve latency"
E, NU, GC, L0 = 210000.0, 0.3, 2.7, 0.01
DEP = 1e-4
INCREMENTS = 100
GOLDEN = os.path.join(ROOT, "tests", "golden", "sent200_history.npz")


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def make_sim(mesh, solver, forced_fp32=False, coarse_reuse=True):
    from easyfea_amd import Models, Simulations

    mat = Models.Elastic.Isotropic(2, E, NU, planeStress=False, thickness=1.0)
    pfm = Models.PhaseField(mat, "Miehe", "AT2", GC, L0)
    simu = Simulations.PhaseField(mesh, pfm)
    simu.solver = solver
    simu.mixed_adaptive = not forced_fp32
    simu.coarse_reuse = coarse_reuse
    # one-time setup outside the timed region: sparsity patterns, scatter maps,
    # coarse aggregates (shared by every simulation on the mesh), tables
    simu.dev.csr_map(2).coarse()
    simu.dev.csr_map(1)
    simu.dev.vec_map(1)
    simu.dev.tables("mass")
    simu.dev.tables("rigi")
    return simu


def run_history(simu, mesh, increments, keep=()):
    """One load history on ``simu`` (fresh).  Returns per-increment host
    latencies (the stream synchronised at the end of each increment by the
    reaction read-back), forces, iteration counts and the kept states."""
    top = mesh.Nodes_Conditions(lambda x, y, z: y == 1.0)
    bot = mesh.Nodes_Conditions(lambda x, y, z: y == 0.0)
    dofsY = simu.Bc_dofs_nodes(top, ["y"])
    lat, forces, its_u, its_d, states = [], [], [], [], {}
    fallbacks = rebuilds = 0
    for it in range(1, increments + 1):
        ts = time.perf_counter()
        simu.Bc_Init()
        simu.add_dirichlet(top, [0, it * DEP], ["x", "y"])
        simu.add_dirichlet(bot, [0], ["y"])
        u, d, _ = simu.Solve(1.0, 500, 1)
        simu.Save_Iter()
        forces.append(float(np.sum(simu.Calc_Reaction(dofsY, "elastic"))))
        lat.append(time.perf_counter() - ts)
        eu, ed = simu.last_solve["elastic"], simu.last_solve["damage"]
        its_u.append(eu.iterations)
        its_d.append(ed.iterations)
        fallbacks += int(eu.fallback) + int(ed.fallback)
        rebuilds += int(eu.precond == "two-level" and not eu.coarse_reused)
        if it in keep:
            states[it] = (u.copy(), d.copy())
    return dict(lat=np.array(lat), forces=np.array(forces), its_u=np.array(its_u), its_d=np.array(its_d),
                states=states, fallbacks=fallbacks, coarse_rebuilds=rebuilds,
                modes=sorted({simu.last_solve["elastic"].mode, simu.last_solve["damage"].mode}))


def timed_histories(mesh, solver, forced_fp32, count, keep, reuse=True):
    """``count`` histories, each on a fresh simulation built (and its maps
    warmed) outside the timed region; returns the summed history time, all
    increment latencies and the last history's record."""
    import torch

    total, lats, last, simu = 0.0, [], None, None
    for _ in range(count):
        del simu
        simu = make_sim(mesh, solver, forced_fp32, reuse)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rec = run_history(simu, mesh, INCREMENTS, keep)
        torch.cuda.synchronize()
        total += time.perf_counter() - t0
        lats.append(rec["lat"])
        last = rec
    return dict(total=total, lat=np.concatenate(lats) if lats else np.zeros(0), last=last, simu=simu)


def cpu_baseline(mesh, steps):
    """The oracle (numpy einsum assembly, bincount scatter, scipy SuperLU)."""
    from oracle import efea_oracle as O

    ref = O.PhaseFieldOracle(mesh.connect, mesh.coord, "TRI3", E, NU, False, 1.0, "Miehe", "AT2", GC, L0)
    top = mesh.Nodes_Conditions(lambda x, y, z: y == 1.0)
    bot = mesh.Nodes_Conditions(lambda x, y, z: y == 0.0)
    dofsY = O.dofs_nodes(top, ["y"], ["x", "y"])
    states, forces = {}, []
    t0 = time.perf_counter()
    for it in range(1, steps + 1):
        ref.Bc_Init()
        ref.add_dirichlet(top, [0, it * DEP], ["x", "y"])
        ref.add_dirichlet(bot, [0], ["y"])
        u, d, _ = ref.Solve(1.0, 500, 1)
        ref.Save_Iter()
        forces.append(float(np.sum(ref.Calc_Reaction(dofsY))))
        states[it] = (u.copy(), d.copy())
    return time.perf_counter() - t0, states, np.array(forces)


def _profiled_traffic(kname, ndof, nnz, precision):
    """HBM bytes per launch of ``kname`` on this system from the committed
    PMC summaries (``profiles/*_traffic.json``), or None."""
    pdir = os.path.join(ROOT, "profiles")
    for fn in sorted(os.listdir(pdir)) if os.path.isdir(pdir) else []:
        if not fn.endswith("_traffic.json"):
            continue
        with open(os.path.join(pdir, fn)) as f:
            t = json.load(f)
        k = t.get("kernel", "").replace("void ", "").replace("efea::", "")
        if t.get("ndof") == ndof and t.get("nnz", nnz) == nnz and t.get("precision") == precision and \
                k.startswith(kname + "<"):
            return t["hbm_bytes_per_launch"]
    return None


def dominant_roofline(simu, problemType, reps):
    """Roofline of the last ``problemType`` solve's dominant kernel: the
    product pass of a pipelined two-level solve, step 1 otherwise."""
    last = simu.last_solve[problemType]
    if last.pipelined and last.precond == "two-level":
        return pp_iter_roofline(simu, problemType, reps)
    return step1_roofline(simu, problemType, reps)


def pp_iter_roofline(simu, problemType, reps):
    """Average duration of the pipelined solve's product pass (k_pp_iter:
    n = A m, the eight vector updates, the partials) over ``reps`` launches
    through ``efea_pp_iter_bench``: the solve's own kernel, chunks, matrix
    (FP32 copy with its correction in a mixed solve), mask and aggregate
    coordinates; HIP events on the stream the kernels run on."""
    import torch

    from easyfea_amd import _lib
    from easyfea_amd.device import ptr, stream_ptr

    A = simu._device_K(problemType) if hasattr(simu, "Detect_Damage") else simu._device_K()
    job = simu.last_job
    info = simu.last_solve[problemType]
    assert job.pipelined and job.cs is not None, "the last solve was not the pipelined one"
    precision = 64 if info.mode == "fp64" else 32
    known = simu._mask_cache[problemType][1]
    cs = job.cs
    pc = cs.pipeline(A.pattern, A.group)
    x = torch.empty(A.Ndof, dtype=torch.float64, device="cuda")
    nbytes = _lib.load().efea_pcg2p_workspace_bytes(A.Nn, A.dof_n, pc.nslot)
    ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    if precision == 32:
        A.make_fp32()
    head = (precision, A.Nn, A.dof_n, A.group, ptr(A.node_ptr), ptr(A.node_cols), ptr(A.values),
            ptr(A.values32 if precision == 32 else None), ptr(A.corr32 if precision == 32 else None), ptr(known),
            cs.nagg, ptr(cs.rel), pc.nchunk, ptr(pc.rows), ptr(pc.slot), pc.maxch, ptr(x), ptr(ws), nbytes)
    _lib.call("efea_pp_iter_bench", *head, 20, stream_ptr())
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    _lib.call("efea_pp_iter_bench", *head, reps, stream_ptr())
    e1.record(s)
    e1.synchronize()
    dt = e0.elapsed_time(e1) / 1e3 / reps  # six fill / state kernels, then reps product passes
    hA = A.host
    nnz, nnz_node, Nn, Ndof, DN = A.nnz, hA.node_cols.size, A.Nn, A.Ndof, A.dof_n
    vb = 4 if precision == 32 else 8
    nm = cs.nm
    # values (+ FP32 row-sum correction) + node columns + per-position node and
    # row range + mask + aggregate coordinates + nine vectors read (m gathered
    # once) + eight written + chunk slots and the partials
    algo_bytes = (nnz * vb + nnz_node * 4 + Nn * (4 + 8) + Ndof + Nn * 24 + Ndof * 8 * 9 + Ndof * 8 * 8
                  + pc.nchunk * 4 + pc.nchunk * 8 * (4 + nm))
    if precision == 32:
        algo_bytes += Nn * DN * DN * 4
    kname = "k_pp_iter"
    traffic = _profiled_traffic(kname, Ndof, nnz, precision)
    peak = 8.0e12
    return {
        "kernel": f"{kname}<{DN},{A.group},{'float' if precision == 32 else 'double'}>",
        "bound": "hbm",
        "achieved": algo_bytes / dt / 1e9,
        "peak": peak / 1e9,
        "unit": "GB/s",
        "frac": algo_bytes / dt / peak,
        "traffic": traffic,
        "bytes_per_launch": algo_bytes,
        "avg_launch_us": dt * 1e6,
        "launches_timed": reps,
        "nnz": nnz,
    }


def step1_roofline(simu, problemType, reps):
    """Average duration of the solve's dominant kernel - step 1: p update
    fused into the node-block SpMV (k_pcg2_step1 in a two-level solve,
    k_pcg_step1 in a block-Jacobi one) - over ``reps`` launches through
    ``efea_step1_bench``: the solve's own kernel, grid, matrix values
    (FP32 copy in a mixed solve), Dirichlet mask and partial count, in the
    state of a running solve; HIP events on the stream the kernels run on."""
    import torch

    from easyfea_amd import _lib
    from easyfea_amd.device import ptr, stream_ptr

    A = simu._device_K(problemType) if hasattr(simu, "Detect_Damage") else simu._device_K()
    info = simu.last_solve[problemType]
    two_level = info.precond == "two-level"
    precision = 64 if info.mode == "fp64" else 32
    known = simu._mask_cache[problemType][1]
    nagg = A.coarse().nagg if two_level else 0
    z = torch.rand(A.Ndof, dtype=torch.float64, device="cuda")
    q = torch.empty_like(z)
    vals = A.make_fp32() if precision == 32 else A.values
    scratch = torch.empty(_lib.load().efea_step1_bench_scratch_bytes(A.Nn, A.dof_n, nagg), dtype=torch.uint8,
                          device="cuda")
    head = (int(two_level), precision, A.Nn, A.dof_n, A.group, ptr(A.node_ptr), ptr(A.node_cols), ptr(vals),
            ptr(A.corr32 if precision == 32 else None), ptr(known), nagg, ptr(z), ptr(q), ptr(scratch))
    _lib.call("efea_step1_bench", *head, 20, stream_ptr())
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    _lib.call("efea_step1_bench", *head, reps, stream_ptr())
    e1.record(s)
    e1.synchronize()
    dt = e0.elapsed_time(e1) / 1e3 / reps  # one state kernel, then reps step-1 launches
    hA = A.host
    nnz, nnz_node, Nn, Ndof = A.nnz, hA.node_cols.size, A.Nn, A.Ndof
    vb = 4 if precision == 32 else 8
    # values (+ FP32 row-sum correction) + node columns + node offsets + mask
    # + z, p_old read + p, q written
    algo_bytes = nnz * vb + nnz_node * 4 + (Nn + 1) * 4 + Ndof * 1 + Ndof * 8 * 4
    if precision == 32:
        algo_bytes += Nn * A.dof_n * A.dof_n * 4
    csr_bytes = nnz * (vb + 4) + (Ndof + 1) * 4 + 2 * Ndof * vb
    kname = "k_pcg2_step1" if two_level else "k_pcg_step1"
    traffic = _profiled_traffic(kname, Ndof, nnz, precision)
    peak = 8.0e12
    return {
        "kernel": f"{kname}<{'float' if precision == 32 else 'double'},{A.dof_n},{A.group}>",
        "bound": "hbm",
        "achieved": algo_bytes / dt / 1e9,
        "peak": peak / 1e9,
        "unit": "GB/s",
        "frac": algo_bytes / dt / peak,
        "traffic": traffic,
        "bytes_per_launch": algo_bytes,
        "csr_formula_bytes": csr_bytes,
        "avg_launch_us": dt * 1e6,
        "launches_timed": reps,
        "nnz": nnz,
    }


def _cpu_cantilever():
    """The oracle's cantilever: assembly + SuperLU direct solve on the host."""
    from easyfea_amd import workloads
    from easyfea_amd._utils import ElemType
    from oracle import efea_oracle as O

    mesh = workloads.cantilever().mesh
    t = time.perf_counter()
    ref = O.ElasticOracle(mesh.connect, mesh.coord, "QUAD4", 210000.0, 0.3, True, 13.0,
                          {"SEG2": mesh.dict_groupElem[ElemType.SEG2].connect})
    ref.add_dirichlet(mesh.Nodes_Conditions(lambda x, y, z: x == 0), [0, 0], ["x", "y"])
    ref.add_surfLoad(mesh.Nodes_Conditions(lambda x, y, z: x == 120.0), [-800.0 / 13.0 ** 2], ["y"])
    u = ref.Solve()
    return u, time.perf_counter() - t


def _cpu_cube(n):
    from easyfea_amd import mesh as M
    from easyfea_amd._utils import ElemType
    from oracle import efea_oracle as O

    m = M.hexa8_cube(n)
    rng = np.random.default_rng(0)
    s = np.sqrt(np.log(1.04))
    E = np.exp(rng.normal(np.log(210000.0) - 0.5 * s * s, s, m.Ne))
    t = time.perf_counter()
    ref = O.ElasticOracle(m.connect, m.coord, "HEXA8", E, 0.3, False, 1.0,
                          {"QUAD4": m.dict_groupElem[ElemType.QUAD4].connect})
    ref.add_dirichlet(m.Nodes_Conditions(lambda x, y, z: z == 0), [0, 0, 0], ["x", "y", "z"])
    ref.add_surfLoad(m.Nodes_Conditions(lambda x, y, z: z == 1.0), [-10.0], ["z"])
    u = ref.Solve()
    return u, time.perf_counter() - t


def bench_single(args, workload) -> dict:
    """One elastic solve of the cantilever (360 x 39 QUAD4) or of the 64^3
    HEXA8 cube: every (precision, preconditioner) pair, DOF/s of the best
    repeat, p50/p99 over the repeats, error against the FP64 reference."""
    from easyfea_amd import solve_bench

    roofs = {}

    def on_run(simu, precond, solver):
        if precond == "two-level" and solver == "hip_mixed":
            roofs["head"] = dominant_roofline(simu, "elastic", args.spmv_reps)

    res = solve_bench.run(workload, 64, 5, on_run=on_run)
    cores = len(os.sched_getaffinity(0))
    cpu = None
    if workload == "cantilever":
        u_ref, t_cpu = _cpu_cantilever()
        res["reference"] = "oracle sparse direct solve (scipy SuperLU)"
        for r in res["runs"]:
            r["u_rel_err_vs_reference"] = float(np.linalg.norm(r["_u"] - u_ref) / np.linalg.norm(u_ref))
        if not args.no_cpu:
            cpu = {"value": u_ref.size / t_cpu, "unit": "DOF/s", "cores": 1, "kind": "port",
                   "sample": f"one full cantilever solve on the host (numpy assembly, SuperLU), {t_cpu:.2f} s"}
    elif not args.no_cpu:
        u16, t_cpu = _cpu_cube(16)
        cpu = {"value": u16.size / t_cpu, "unit": "DOF/s", "cores": 1, "kind": "port",
               "sample": f"oracle direct solve of the 16^3 cube ({u16.size} DOF; the 64^3 one does not fit "
                         f"SuperLU) on the host, {t_cpu:.2f} s"}
    for r in res["runs"]:
        r.pop("_u")
    runs = {(r["precond"], r["solver"]): r for r in res["runs"]}
    head = runs[("two-level", "hip_mixed")]
    plain = runs[("two-level", "hip_fp64")]
    cfg = ("cantilever QUAD4 360x39, plane stress, t=13, -800 N at x=120" if workload == "cantilever"
           else "HEXA8 64^3 unit cube, lognormal E (CoV 0.2, seed 0), clamped z=0, -10 MPa on z=1")
    return {
        "metric": METRIC, "value": head["value"], "unit": "DOF/s", "n_gpus": 1, "steps": 1,
        "warmup": 1, "ms_per_step": 1e3 * head["ndof"] / head["value"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 vectors (f32 matrix copy when mixed)",
        "data": "synthetic mesh generated in-process",
        "config": {"workload": cfg, "ndof": head["ndof"], "solver": "two-level PCG, hip_mixed"},
        "p50_solve_ms": head["p50_ms"], "p99_solve_ms": head["p99_ms"],
        "u_rel_err_vs_reference": head["u_rel_err_vs_reference"], "reference": res["reference"],
        "plain_fp64": {"value": plain["value"], "mixed_speedup": head["value"] / plain["value"]},
        "roofline": {k: roofs["head"][k] for k in ("bound", "achieved", "peak", "unit", "frac", "traffic")},
        "roofline_detail": roofs["head"],
        "runs": res["runs"], "cpu_baseline": cpu,
    }


def _cpu_dynamics(workload, cells, steps):
    """The oracle's time steps on the host (matrices assembled once, then per
    step the reference's right-hand side and a SuperLU solve of the
    Dirichlet-reduced system): (DOF, seconds per step)."""
    from easyfea_amd import mesh as MS
    from easyfea_amd import workloads
    from easyfea_amd._utils import ElemType
    from oracle import efea_oracle as O

    if workload == "newmark-cube":
        m = MS.hexa8_cube(cells)
        E = workloads.cube_E(m.Ne, 0)
        algo, params = "newmark", O.hyperbolic_params(5e-8)
        ref = O.DynamicElasticOracle(m.connect, m.coord, "HEXA8", E, 0.3, False, 1.0,
                                     {"QUAD4": m.dict_groupElem[ElemType.QUAD4].connect}, rho=workloads.RHO_STEEL)
        ref.add_dirichlet(m.Nodes_Conditions(lambda x, y, z: z == 0), [0, 0, 0], ["x", "y", "z"])
        ref.add_surfLoad(m.Nodes_Conditions(lambda x, y, z: z == 1.0), [-10.0], ["z"])
        K, C, M = ref.Get_K_C_M()
        u = np.zeros(ref.Ndof)
        pt, unk = "elastic", ["x", "y", "z"]
    else:
        m = MS.quad4_rectangle(1.0, 1.0, cells, cells)
        algo, params = "parabolic", (1e-4, 0.5)
        ref = O.ThermalOracle(m.connect, m.coord, "QUAD4", workloads.thermal_k(cells, 0), 1.0, 1.0, 1.0,
                              {"SEG2": m.dict_groupElem[ElemType.SEG2].connect})
        ref.add_dirichlet(m.Nodes_Conditions(lambda x, y, z: x == 0), [0.0], ["t"], "thermal")
        ref.add_surfLoad(m.Nodes_Conditions(lambda x, y, z: y == 1.0), [5.0], ["t"], "thermal")
        K, C = ref.Get_K_C()
        M = 0 * C
        u = np.zeros(m.Nn)
        pt, unk = "thermal", ["t"]
    F = ref.neumann_vector(pt, u.size)
    dofs, vals = ref.dir[pt]
    orphan = O.dofs_nodes(ref.orphan, unk, unk)
    v, a = np.zeros_like(u), np.zeros_like(u)
    t = time.perf_counter()
    for _ in range(steps):
        u, v, a = O.time_step(algo, params, K, C, M, F, u, v, a, dofs, vals, orphan)
    return u.size, (time.perf_counter() - t) / steps


def bench_dynamics(args, workload) -> dict:
    """Time stepping (easyfea_amd.dynamics_bench): Newmark on the 64^3 cube or
    Crank-Nicolson on the 1024^2 square, with the oracle's host steps of a
    smaller mesh of the same problem as the CPU baseline."""
    from easyfea_amd import dynamics_bench

    cube = workload == "newmark-cube"
    cells = 64 if cube else 1024
    res = dynamics_bench.run(workload, cells, args.dyn_steps, 2)
    cpu = None
    if not args.no_cpu:
        cpu_cells = 16 if cube else 256
        ndof, s_step = _cpu_dynamics(workload, cpu_cells, 3)
        cpu = {"value": ndof / s_step, "unit": "DOF*steps/s", "cores": 1,
               "kind": "port",
               "sample": f"3 oracle time steps of the {cpu_cells}"
                         f"{'^3 cube' if cube else '^2 square'} ({ndof} DOF; matrices assembled once, "
                         f"SuperLU per step) on the host, {s_step:.2f} s per step"}
    cfg = ("HEXA8 64^3 unit cube, lognormal E (CoV 0.2, seed 0), steel rho, clamped z=0, -10 MPa on z=1 "
           "applied at t=0; Newmark beta 1/4 gamma 1/2, dt 5e-8 s" if cube else
           "QUAD4 1024^2 unit square, lognormal k (mean 1, CoV 0.3), c = rho = 1, t=0 at x=0, flux 5 on y=1; "
           "Crank-Nicolson, dt 1e-4")
    return {
        "metric": "time-step throughput", "value": res["value"], "unit": "DOF*steps/s", "n_gpus": 1,
        "steps": res["steps"], "warmup": res["warmup"], "ms_per_step": res["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 vectors (f32 matrix copy when mixed)", "data": "synthetic mesh generated in-process",
        "config": {"workload": cfg, "ndof": res["ndof"], "solver": f"{res['precond']} PCG, {res['mode']}"},
        "p50_step_ms": res["p50_ms"], "p99_step_ms": res["p99_ms"], "iterations_mean": res["iterations_mean"],
        "iterations_max": res["iterations_max"], "cpu_baseline": cpu,
    }


def bench_serving(args, workload) -> dict:
    """The serving streams (easyfea_amd.serving_bench) with the oracle's
    host solves of a bounded sample as the CPU baseline."""
    from easyfea_amd import serving_bench

    wl = workload[6:]
    argv = ["--workload", wl, "--compare-fp64"]
    out = serving_bench.main(argv, quiet=True)
    saturated = None
    if wl == "elastic":
        # the same 512 requests offered far above capacity: throughput bound
        sat = serving_bench.main(["--workload", "elastic", "--qps", "4000", "--max-blocks", "64"], quiet=True)
        saturated = {k: sat[k] for k in ("value", "wall_s", "achieved_qps", "p50_latency_ms", "p99_latency_ms",
                                         "max_concurrent_solves", "offered_qps", "max_blocks")}
    cpu = None
    if wl == "elastic" and not args.no_cpu:
        from easyfea_amd._utils import ElemType
        from oracle import efea_oracle as O

        specs, meshes = out["_specs"][:8], out["_meshes"]
        t0, ndof = time.perf_counter(), 0
        for spec in specs:
            m = meshes.get(spec.side)
            ref = O.ElasticOracle(m.connect, m.coord, "QUAD4", spec.E, 0.3, True, 1.0,
                                  {"SEG2": m.dict_groupElem[ElemType.SEG2].connect})
            ref.add_dirichlet(m.Nodes_Conditions(lambda x, y, z: x == 0.0), [0, 0], ["x", "y"])
            ref.add_surfLoad(m.Nodes_Conditions(lambda x, y, z: x == 1.0), [1.0], ["x"])
            ref.Solve()
            ndof += 2 * m.Nn
        t = time.perf_counter() - t0
        cpu = {"value": ndof / t, "unit": "DOF/s", "cores": 1, "kind": "port",
               "sample": f"first {len(specs)} requests of the stream solved one after another on the host "
                         f"(single-threaded numpy assembly and SuperLU), {t:.1f} s"}
    if wl == "phasefield" and not args.no_cpu:
        from easyfea_amd import workloads
        from easyfea_amd.mesh import tri3_notched_square
        from oracle import efea_oracle as O

        specs, steps = out["_specs"][:2], 2
        t0, ndof = time.perf_counter(), 0
        for spec in specs:
            m = tri3_notched_square(spec.cells)
            Gc_e = workloads.lognormal(2.7, 0.15, spec.Gc[workloads.cell_of_elements(m, spec.cells)])
            ref = O.PhaseFieldOracle(m.connect, m.coord, "TRI3", 210000.0, 0.3, False, 1.0, "Amor", "AT2", Gc_e,
                                     0.01)
            bottom = m.Nodes_Conditions(lambda x, y, z: y == 0.0)
            top = m.Nodes_Conditions(lambda x, y, z: y == 1.0)
            for n in range(1, steps + 1):
                ref.Bc_Init()
                ref.add_dirichlet(bottom, [0], ["y"])
                ref.add_dirichlet(top, [0, n * spec.du], ["x", "y"])
                ref.Solve(1.0, 500, 1)
                ref.Save_Iter()
            ndof += 3 * m.Nn * steps
        t = time.perf_counter() - t0
        cpu = {"value": ndof / t, "unit": "DOF/s", "cores": 1, "kind": "port",
               "sample": f"first {steps} increments of the first {len(specs)} realisations, one after another on "
                         f"the host (single-threaded numpy assembly, SuperLU), {t:.1f} s"}
    roof = _serving_roofline(wl, out, args.spmv_reps)
    line = {k: v for k, v in out.items() if not k.startswith("_")}
    if saturated is not None:
        line["saturated"] = saturated
    line["roofline"] = {k: roof[k] for k in ("bound", "achieved", "peak", "unit", "frac", "traffic")}
    line["roofline_detail"] = roof
    line.update({"metric": METRIC, "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                 "data": "synthetic meshes and random fields generated in-process", "cpu_baseline": cpu})
    return line


def _serving_roofline(wl, out, reps):
    """Step-1 kernel roofline of the serving stream on its largest system
    (the batched kernels run the same body per system): the largest square
    of the elastic stream, or the largest realisation's displacement matrix
    after its first increment; FP32 values, as the serving batches use."""
    from easyfea_amd import Models, Simulations, workloads
    from easyfea_amd.mesh import tri3_notched_square

    specs = out["_specs"]
    if wl == "elastic":
        spec = max(specs, key=lambda s: s.side)
        mesh = out["_meshes"].get(spec.side)
        mat = Models.Elastic.Isotropic(2, E=spec.E, v=0.3, planeStress=True, thickness=1.0)
        simu = Simulations.Elastic(mesh, mat)
        simu.solver = "hip_mixed"
        simu.mixed_adaptive = False
        left, right = workloads._square_sides(mesh)
        simu.add_dirichlet(left, [0, 0], ["x", "y"])
        simu.add_surfLoad(right, [1.0], ["x"])
        simu.Solve()
        r = step1_roofline(simu, "elastic", reps)
        r["system"] = f"largest request: QUAD4 {spec.side}x{spec.side}"
        return r
    spec = max(specs, key=lambda s: s.cells)
    mesh = tri3_notched_square(spec.cells)
    Gc_e = workloads.lognormal(2.7, 0.15, spec.Gc[workloads.cell_of_elements(mesh, spec.cells)])
    simu = workloads._sent_on(mesh, Gc_e, "mixed")
    simu.mixed_adaptive = False
    workloads.sent_bc(simu, spec.du)
    simu.Solve(1.0, 500, 1)
    r = step1_roofline(simu, "elastic", reps)
    r["system"] = f"largest realisation: TRI3 {spec.cells}x{spec.cells} cells, displacement matrix"
    return r


def bench_sent(args) -> dict:
    """The headline: ``--steps`` timed load histories per path (mixed with
    the FP32 copy forced, then plain FP64), each on a fresh simulation."""
    import easyfea_amd as ea

    mesh = ea.tri3_notched_square(args.cells)
    ndof_u, ndof_d = mesh.Nn * 2, mesh.Nn
    dofs_per_inc = ndof_u + ndof_d  # one staggered pass: one displacement + one damage solve
    keep = (1, 25, 50, 75, 100)
    reuse = not args.no_coarse_reuse
    paths = [("mixed", "hip_mixed", True)] + ([] if args.no_plain else [("fp64", "hip_fp64", False)])
    res = {}
    for name, solver, forced in paths:
        if args.warmup > 0:
            timed_histories(mesh, solver, forced, args.warmup, (), reuse)
        log(f"{name}: warmup done ({args.warmup} histories)")
        res[name] = timed_histories(mesh, solver, forced, args.steps, keep, reuse)
        log(f"{name}: {args.steps} histories in {res[name]['total']:.3f} s")
    mixed = res["mixed"]
    value = args.steps * INCREMENTS * dofs_per_inc / mixed["total"]

    def rel(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    accuracy = {}
    golden = None
    if os.path.exists(GOLDEN) and args.cells == 200:
        g = np.load(GOLDEN)
        golden = {k: g[k] for k in g.files}
    for name, r in res.items():
        last = r["last"]
        if golden is not None:
            dev = np.abs(last["forces"] - golden["forces"]) / np.abs(golden["forces"])
            accuracy[f"{name}_vs_oracle_direct"] = {
                "force_curve_max_rel_dev": float(dev.max()),
                "u_rel_err_max": max(rel(last["states"][k][0], golden[f"u_{k}"]) for k in keep),
                "d_rel_err_max": max(rel(last["states"][k][1], golden[f"d_{k}"]) for k in keep),
                "increments_compared": list(keep),
            }
    if "fp64" in res:
        a, b = mixed["last"], res["fp64"]["last"]
        fr = np.abs(a["forces"] - b["forces"]) / np.maximum(np.abs(b["forces"]), 1e-300)
        accuracy["mixed_vs_fp64"] = {
            "force_curve_max_rel_dev": float(fr.max()),
            "u_rel_err_max": max(rel(a["states"][k][0], b["states"][k][0]) for k in keep),
            "d_rel_err_max": max(rel(a["states"][k][1], b["states"][k][1]) for k in keep),
        }
    if golden is not None:
        accuracy["reference"] = ("oracle FP64 history with sparse direct (SuperLU) solves, all 100 increments "
                                 "(tests/golden/sent200_history.npz)")

    roof = dominant_roofline(mixed["simu"], "elastic", args.spmv_reps)
    roof_fp64 = dominant_roofline(res["fp64"]["simu"], "elastic", args.spmv_reps) if "fp64" in res else None

    cpu = None
    if not args.no_cpu and args.cpu_steps > 0:
        n = min(args.cpu_steps, INCREMENTS)
        t_cpu, st_cpu, f_cpu = cpu_baseline(mesh, n)
        cores = len(os.sched_getaffinity(0))
        cpu = {
            "value": n * dofs_per_inc / t_cpu,
            "unit": "DOF/s",
            "cores": 1,
            "kind": "port",
            "sample": f"first {n} of the {INCREMENTS} increments of the same history on the host (numpy "
                      f"einsum assembly + bincount scatter + scipy SuperLU spsolve, single-threaded "
                      f"factorisation; {cores} cores visible), {t_cpu:.1f} s",
        }
        log(f"cpu baseline: {n} increments in {t_cpu:.1f} s")
    lat_ms = mixed["lat"] * 1e3
    last = mixed["last"]
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["total"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 matrix values in the Krylov products; f64 vectors, residual replacement, eigen and history",
        "data": "synthetic: structured notched-square TRI3 mesh generated in-process",
        "config": {
            "workload": f"phase-field SENT square TRI3 {args.cells}x{args.cells} cells, Miehe/AT2/History, "
                        f"one step = the whole history of {INCREMENTS} x 1e-4 mm increments, staggered one pass",
            "elements": int(mesh.Ne),
            "ndof_u": int(ndof_u),
            "ndof_d": int(ndof_d),
            "solver": "displacement: two-level PCG (nodal block Jacobi + rigid-body aggregation coarse space"
                      + (", coarse inverse kept over increments until the iteration count grows 20%"
                         if reuse else "") + "); damage: block-Jacobi PCG; secant start from the last two "
                      "increments; mixed: FP32 values (row-sum corrected) in every Krylov product, FP64 vectors, "
                      "FP64 residual replacement every 32 iterations",
        },
        "increments_timed": args.steps * INCREMENTS,
        "ms_per_increment": mixed["total"] / (args.steps * INCREMENTS) * 1e3,
        "p50_increment_ms": float(np.percentile(lat_ms, 50)),
        "p99_increment_ms": float(np.percentile(lat_ms, 99)),
        "pcg_iters_u_mean": float(last["its_u"].mean()),
        "pcg_iters_d_mean": float(last["its_d"].mean()),
        "solve_modes": last["modes"],
        "fp64_fallbacks": int(last["fallbacks"]),
        "coarse_rebuilds": int(last["coarse_rebuilds"]),
        "peak_force_N": float(last["forces"].max()),
        "final_force_N": float(last["forces"][-1]),
        "accuracy": accuracy,
        "roofline": {k: roof[k] for k in ("bound", "achieved", "peak", "unit", "frac", "traffic")},
        "roofline_detail": roof,
        "cpu_baseline": cpu,
    }
    if "fp64" in res:
        p = res["fp64"]
        out["plain_fp64"] = {
            "value": args.steps * INCREMENTS * dofs_per_inc / p["total"],
            "ms_per_step": p["total"] / args.steps * 1e3,
            "ms_per_increment": p["total"] / (args.steps * INCREMENTS) * 1e3,
            "p50_increment_ms": float(np.percentile(p["lat"] * 1e3, 50)),
            "p99_increment_ms": float(np.percentile(p["lat"] * 1e3, 99)),
            "pcg_iters_u_mean": float(p["last"]["its_u"].mean()),
            "pcg_iters_d_mean": float(p["last"]["its_d"].mean()),
            "mixed_speedup": p["total"] / mixed["total"],
            "roofline_detail": roof_fp64,
        }
    return out


EXTRAS = ("cantilever", "cube", "serve-elastic", "serve-phasefield")


def run_workload(args, name) -> dict:
    if name in ("cantilever", "cube"):
        return bench_single(args, name)
    if name.startswith("serve-"):
        return bench_serving(args, name)
    if name in ("newmark-cube", "thermal-square"):
        return bench_dynamics(args, name)
    raise ValueError(name)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed load histories per path (sent)")
    ap.add_argument("--warmup", type=int, default=2, help="untimed load histories per path (sent)")
    ap.add_argument("--cells", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="headline line only")
    ap.add_argument("--spmv-reps", type=int, default=500)
    ap.add_argument("--dyn-steps", type=int, default=20, help="time steps of the time-stepping workloads")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="end with one device-to-device copy of this many MiB (HBM counter calibration)")
    ap.add_argument("--no-coarse-reuse", action="store_true",
                    help="rebuild the coarse inverse at every displacement solve")
    ap.add_argument("--workload", default="sent",
                    choices=["sent", "cantilever", "cube", "serve-elastic", "serve-phasefield", "newmark-cube",
                             "thermal-square"],
                    help="sent: the phase-field history above (the headline, with the other workloads' lines "
                         "unless --no-extra); the others print the measurement line of the cantilever / 64^3 "
                         "cube single solves, of the serving streams or of the time-stepping workloads alone")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs on a single MI355X (multi-GPU runs are not part "
              "of it); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.workload != "sent":
        print(json.dumps(run_workload(args, args.workload)), flush=True)
        return
    out = bench_sent(args)
    if not args.no_extra:
        from easyfea_amd import device

        floor = device.MIXED_MIN_MATRIX_BYTES
        out["configs"] = {}
        for name in EXTRAS:
            t = time.perf_counter()
            try:
                out["configs"][name] = run_workload(args, name)
            except Exception as e:  # recorded in the line, never hidden
                out["configs"][name] = {"error": repr(e)[:500]}
            finally:
                device.MIXED_MIN_MATRIX_BYTES = floor  # the serving runs lower it
            log(f"{name}: {time.perf_counter() - t:.1f} s")
    print(json.dumps(out), flush=True)
    if args.copy_mib > 0:
        n = args.copy_mib * 2**20 // 8
        src = torch.rand(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()


if __name__ == "__main__":
    main()
