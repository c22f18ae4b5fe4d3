"""Lanes-per-node autotuner of the node-group solver kernels.

``python -m easyfea_amd.autotune [--reps 200] [--quick]``

Every Krylov kernel that multiplies by a global matrix gives each node
``G`` lanes (a power of two, 2..32) that stride over the node's neighbour
blocks and combine their partial rows with cross-lane shuffles.  Small
``G`` leaves lanes idle on nodes with few neighbours; large ``G`` costs
more shuffle rounds and fewer nodes per wave.  This tuner times the PCG's
step-1 kernel (p update fused into the SpMV, through ``efea_spmv_repeat``)
for every supported ``G`` on meshes of each element family at several
sizes, in FP64 and FP32 values, and prints one JSON line per case plus the
rule the measurements give, which is what ``efea_default_group`` encodes.
Matrix values are random: the kernel's time does not depend on them.
"""

from __future__ import annotations

import argparse
import json

import numpy as np

CANDIDATES = {1: (2, 4, 8, 16), 2: (4, 8, 16), 3: (8, 16, 32)}


def _cases(quick: bool):
    from .mesh import hexa8_cube, quad4_rectangle, tetra4_box, tetra10_from_tetra4, tri3_notched_square

    sq = (64, 256) if quick else (64, 256, 1024)
    pf = (100,) if quick else (100, 400)
    cu = (16,) if quick else (16, 32, 64)
    te = (12,) if quick else (12, 24, 40)
    for n in pf:
        yield f"TRI3 notched square {n}", (lambda n=n: tri3_notched_square(n)), (2, 1)
    for n in sq:
        yield f"QUAD4 square {n}x{n}", (lambda n=n: quad4_rectangle(1.0, 1.0, n, n)), (2, 1)
    for n in cu:
        yield f"HEXA8 cube {n}^3", (lambda n=n: hexa8_cube(n)), (3,)
    for n in te:
        yield f"TETRA4 box {n}^3", (lambda n=n: tetra4_box(1.0, 1.0, 1.0, n, n, n)), (3,)
    for n in te[:-1] if not quick else te:
        yield f"TETRA10 box {n}^3", (lambda n=n: tetra10_from_tetra4(tetra4_box(1.0, 1.0, 1.0, n, n, n))), (3,)


def time_step1(A, precision: int, reps: int, group: int = 0) -> float:
    """Average duration (s) of ``reps`` back-to-back step-1 kernels (p update
    fused into the node-block SpMV, ``efea_spmv_repeat``) on ``A``'s matrix
    with ``group`` lanes per node, timed with HIP events on the stream the
    kernels run on; FP64 or FP32 values."""
    import torch

    from . import _lib
    from .device import ptr, stream_ptr

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


def tune_pattern(pattern, reps: int) -> dict:
    """Times every candidate lanes-per-node value on ``pattern``."""
    import torch

    from . import _lib
    from .device import DeviceCsr

    A = DeviceCsr(pattern)
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    A.values.copy_(torch.rand(A.nnz, dtype=torch.float64, device="cuda", generator=g))
    cnt = np.diff(pattern.host.node_ptr)
    row = {"Nn": int(A.Nn), "dof_n": int(A.dof_n), "nnz": int(A.nnz), "max_cnt": int(cnt.max()),
           "mean_cnt": float(cnt.mean()),
           "default": int(_lib.load().efea_default_group(A.dof_n, int(cnt.max()), A.Nn))}
    for prec in (64, 32):
        t = {G: time_step1(A, prec, reps, G) * 1e6 for G in CANDIDATES[A.dof_n]}
        row[f"fp{prec}_us"] = {str(k): round(v, 3) for k, v in t.items()}
        row[f"best{prec}"] = int(min(t, key=t.get))
        row[f"default_loss{prec}"] = round(t[row["default"]] / min(t.values()) - 1.0, 4)
    del A
    torch.cuda.empty_cache()
    return row


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--reps", type=int, default=200)
    ap.add_argument("--quick", action="store_true", help="one small size per element family")
    args = ap.parse_args(argv)
    import torch

    from .device import DeviceMesh

    assert torch.cuda.is_available(), "the autotuner times kernels on a GPU"
    torch.cuda.set_device(0)
    rows = []
    for name, make, dofs in _cases(args.quick):
        mesh = make()
        for dn in dofs:
            row = {"case": name, **tune_pattern(DeviceMesh(mesh).pattern(dn), args.reps)}
            rows.append(row)
            print(json.dumps(row), flush=True)
        del mesh
    worst = max(max(r["default_loss64"], r["default_loss32"]) for r in rows)
    print(json.dumps({"summary": "default_group vs fastest", "cases": len(rows),
                      "worst_default_loss": worst}), flush=True)
    return rows


if __name__ == "__main__":
    main()
