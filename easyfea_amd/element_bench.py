"""Element-stiffness kernels on the 64^3 HEXA8 cube (262,144 elements,
lognormal E): the isotropic closed form (``efea_elastic_ke``) against the
dense B^T D B on the FP64 matrix cores (``efea_elastic_ke_dense``, with the
same isotropic D formed in the kernel, and with one 6 x 6 D per element read
from memory).  HIP events over ``reps`` launches each, on the current
stream; prints one JSON line.

    python -m easyfea_amd.element_bench [--n 64] [--reps 10]
"""

from __future__ import annotations

import argparse
import json

import numpy as np


def main(argv=None) -> dict:
    import torch

    from . import _lib, workloads
    from ._utils import MatrixType
    from .device import DeviceMesh, ptr, stream_ptr, to_dev
    from .mesh import hexa8_cube

    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--reps", type=int, default=10)
    args = ap.parse_args(argv)
    mesh = hexa8_cube(args.n)
    dev = DeviceMesh(mesh)
    tab, nPg = dev.tables(MatrixType.rigi)
    E_np = workloads.cube_E(mesh.Ne, 0)
    E = to_dev(E_np)
    nu = 0.3
    lam, mu = E_np * nu / ((1 + nu) * (1 - 2 * nu)), E_np / (2 * (1 + nu))
    D_np = np.zeros((mesh.Ne, 6, 6))
    D_np[:, :3, :3] = lam[:, None, None]
    for i in range(3):
        D_np[:, i, i] += 2 * mu
        D_np[:, 3 + i, 3 + i] = 2 * mu
    D = to_dev(D_np.reshape(-1))
    nd = 24
    Ke = torch.empty(dev.Ne * nd * nd, dtype=torch.float64, device="cuda")
    Ke2 = torch.empty_like(Ke)
    kernels = {
        "closed_form": lambda out: _lib.call("efea_elastic_ke", dev.code, 64, dev.Ne, ptr(dev.connect),
                                             ptr(dev.coord), ptr(tab), nPg, ptr(E), 1, nu, 0, 1.0, ptr(out),
                                             stream_ptr()),
        "matrix_cores_isotropic": lambda out: _lib.call("efea_elastic_ke_dense", dev.code, dev.Ne, ptr(dev.connect),
                                                        ptr(dev.coord), ptr(tab), nPg, None, 0, ptr(E), 1, nu,
                                                        ptr(out), stream_ptr()),
        "matrix_cores_D_per_element": lambda out: _lib.call("efea_elastic_ke_dense", dev.code, dev.Ne,
                                                            ptr(dev.connect), ptr(dev.coord), ptr(tab), nPg, ptr(D),
                                                            1, None, 0, 0.0, ptr(out), stream_ptr()),
    }
    res = {"workload": f"HEXA8 {args.n}^3 cube, {dev.Ne} elements, lognormal E", "reps": args.reps,
           "ke_bytes": int(dev.Ne * nd * nd * 8)}
    kernels["closed_form"](Ke)
    s = torch.cuda.current_stream()
    for name, f in kernels.items():
        f(Ke2)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(args.reps):
            f(Ke2)
        e1.record(s)
        e1.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / args.reps
        diff = float(torch.max(torch.abs(Ke2 - Ke)) / torch.max(torch.abs(Ke)))
        res[name] = {"us": us, "ke_write_TBps": res["ke_bytes"] / us / 1e6, "max_diff_vs_closed_form": diff}
    flops = 2.0 * dev.Ne * 32 * 32 * 6 * nPg  # four 16x16 tiles (24 x 24 padded to 32 x 32), K-dim 6 nPg
    res["matrix_core_flops_per_launch"] = flops
    res["matrix_cores_isotropic"]["TFLOPs"] = flops / res["matrix_cores_isotropic"]["us"] / 1e6
    print(json.dumps(res))
    return res


if __name__ == "__main__":
    main()
