"""Runs the PCG's dominant kernel on a workload's own matrix, for profiling.

``python -m easyfea_amd.kernel_probe --workload pf200 --reps 200``

Launches ``reps`` back-to-back step-1 kernels (p update fused into the
node-block SpMV) on the displacement matrix of the workload, in FP32 and
FP64 values, then one device copy of a known size (``--copy-mib``), which
calibrates the HBM traffic counters (FETCH_SIZE / WRITE_SIZE) read by
``rocprofv3 --pmc``.  Prints the per-launch algorithmic bytes and the
HIP-event average duration.
"""

import argparse
import json

import numpy as np
import torch

from . import Models, Simulations, _lib
from .device import ptr, stream_ptr
from .mesh import hexa8_cube, quad4_rectangle, tri3_notched_square


def _system(workload: str):
    if workload.startswith("pf"):
        n = int(workload[2:])
        mesh = tri3_notched_square(n)
        mat = Models.Elastic.Isotropic(2, 210000.0, 0.3, planeStress=False, thickness=1.0)
        simu = Simulations.PhaseField(mesh, Models.PhaseField(mat, "Miehe", "AT2", 2.7, 0.01))
        simu._assemble_elastic()
        return simu._Ku
    if workload.startswith("cube"):
        n = int(workload[4:])
        mesh = hexa8_cube(n)
        simu = Simulations.Elastic(mesh, Models.Elastic.Isotropic(3, 210000.0, 0.3))
        return simu.Assembly()
    if workload.startswith("square"):
        n = int(workload[6:])
        mesh = quad4_rectangle(1.0, 1.0, n, n)
        simu = Simulations.Elastic(mesh, Models.Elastic.Isotropic(2, 210000.0, 0.3, True, 1.0))
        return simu.Assembly()
    if workload == "cantilever":
        mesh = quad4_rectangle(120.0, 13.0, 360, 39)
        simu = Simulations.Elastic(mesh, Models.Elastic.Isotropic(2, 210000.0, 0.3, True, 13.0))
        return simu.Assembly()
    raise ValueError(workload)


def algorithmic_bytes(A, value_bytes: int) -> int:
    """Bytes one step-1 launch must move: matrix values, node columns and
    offsets, the known-DOF mask, x (z and p_old) read once, p and q written."""
    nnz_node = A.host.node_cols.size
    return (A.nnz * value_bytes + nnz_node * 4 + (A.Nn + 1) * 4 + A.Ndof * 1 + A.Ndof * 8 * 2
            + A.Ndof * 8 * 2)


def time_step1(A, precision: int, reps: int, group: int = 0) -> float:
    x = torch.rand(A.Ndof, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    vals = A.values if precision == 64 else A.make_fp32()
    head = (precision, A.Nn, A.dof_n, group or A.group, ptr(A.node_ptr), ptr(A.node_cols), ptr(vals), None, ptr(x),
            ptr(y), ptr(A.spmv_scratch))
    _lib.call("efea_spmv_repeat", *head, 10, stream_ptr())
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    _lib.call("efea_spmv_repeat", *head, reps, stream_ptr())
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="pf200")
    ap.add_argument("--reps", type=int, default=200)
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--groups", default="", help="comma-separated lanes-per-node values to time as well")
    args = ap.parse_args()
    torch.cuda.set_device(0)
    A = _system(args.workload)
    out = {"workload": args.workload, "ndof": A.Ndof, "nnz": A.nnz, "group": A.group}
    for prec, vb in ((32, 4), (64, 8)):
        dt = time_step1(A, prec, args.reps)
        b = algorithmic_bytes(A, vb)
        out[f"fp{prec}"] = {"avg_us": dt * 1e6, "bytes": b, "GBps": b / dt / 1e9, "frac_of_8TBps": b / dt / 8e12}
    for g in [int(v) for v in args.groups.split(",") if v]:
        for prec, vb in ((32, 4), (64, 8)):
            dt = time_step1(A, prec, args.reps, g)
            b = algorithmic_bytes(A, vb)
            out[f"fp{prec}_g{g}"] = {"avg_us": dt * 1e6, "GBps": b / dt / 1e9}
    n = args.copy_mib * 2**20 // 8
    src = torch.rand(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    out["copy_bytes_each_way"] = n * 8
    # achievable streaming rate on this device: device-to-device copies
    # (read + write bytes over the event-timed average)
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(20):
        dst.copy_(src)
    e1.record(s)
    e1.synchronize()
    out["copy_GBps"] = 2 * n * 8 / (e0.elapsed_time(e1) / 1e3 / 20) / 1e9
    print(json.dumps(out))


if __name__ == "__main__":
    main()
