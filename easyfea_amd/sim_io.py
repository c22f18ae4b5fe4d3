"""Simulation files: ``_Simu.Save`` and ``Load_Simu``.

The reference pickles the whole simulation object, its meshes saved beside
it (``EasyFEA/Simulations/_simu.py:3157-3214``, ``Load_Simu`` :3333-3366),
with a ``summary.txt``.  A pickle of this library's objects would hold
device tensors and ctypes handles, so the same content is written as data:

* ``folder/Meshes/mesh0.npz`` -- the mesh (``Mesh.Save``);
* ``folder/<filename>.npz`` -- a JSON header (simulation class, model and
  its parameters, solver, time scheme, density, Rayleigh damping, the
  scalar entries of each saved iteration, the boundary conditions) and one
  array per array entry of each saved iteration (``r<i>/<key>``), per
  array-valued parameter (``p/<name>``), per boundary condition (nodes,
  DOFs, values) and per state field (``s/<name>``: the current
  displacement, speed, acceleration, damage and phase-field history);
* ``folder/summary.txt``.

No pickle is written or read (``allow_pickle=False``).  ``Load_Simu``
rebuilds the simulation on the device with the same results list and
state, so that a run continued after loading matches one that was not
interrupted.
"""

from __future__ import annotations

import json
import os
from datetime import datetime

import numpy as np

from .mesh_io import load_mesh

_STATE = ("_u", "_v", "_a", "_d", "_H")


def _param(value, name: str, arrays: dict):
    """JSON-able parameter; arrays go to ``arrays`` under ``p/<name>``."""
    if isinstance(value, np.ndarray):
        arrays[f"p/{name}"] = value
        return {"array": f"p/{name}"}
    if isinstance(value, (np.floating, np.integer)):
        return value.item()
    return value


def _unparam(value, z):
    if isinstance(value, dict) and "array" in value:
        return np.array(z[value["array"]])
    return value


def _model_header(model, arrays: dict) -> dict:
    from .Models import Isotropic, PhaseField, Thermal

    if isinstance(model, Isotropic):
        return {"kind": "Isotropic", "dim": model.dim, "E": _param(model.E, "E", arrays), "v": model.v,
                "planeStress": bool(model.planeStress), "thickness": model.thickness}
    if isinstance(model, PhaseField):
        return {"kind": "PhaseField", "material": _model_header(model.material, arrays),
                "split": str(model.split.value), "regularization": str(model.regularization.value),
                "Gc": _param(model.Gc, "Gc", arrays), "l0": model.l0, "solver": str(model.solver.value)}
    if isinstance(model, Thermal):
        return {"kind": "Thermal", "k": _param(model.k, "k", arrays), "c": _param(model.c, "c", arrays),
                "thickness": model.thickness}
    raise TypeError(f"no file format for the model {type(model).__name__}")


def _model_from(h: dict, z):
    from .Models import Isotropic, PhaseField, Thermal

    if h["kind"] == "Isotropic":
        return Isotropic(h["dim"], _unparam(h["E"], z), h["v"], h["planeStress"], h["thickness"])
    if h["kind"] == "PhaseField":
        return PhaseField(_model_from(h["material"], z), h["split"], h["regularization"], _unparam(h["Gc"], z),
                          h["l0"], h["solver"])
    if h["kind"] == "Thermal":
        return Thermal(_unparam(h["k"], z), _unparam(h["c"], z), h["thickness"])
    raise ValueError(f"unknown model kind {h['kind']}")


def save_simu(simu, folder: str, filename: str = "simulation", additionalInfos: str = "") -> str:
    """``_Simu.Save``: writes the mesh, the simulation file and the summary;
    returns the simulation file's path."""
    import torch

    os.makedirs(folder, exist_ok=True)
    mesh_path = simu.mesh.Save(os.path.join(folder, "Meshes"), "mesh0")
    arrays: dict = {}
    iters = []
    for i, it in enumerate(simu.results):
        scal = {}
        for key, val in it.items():
            if isinstance(val, torch.Tensor):
                arrays[f"r{i}/{key}"] = val.detach().cpu().numpy()
                scal[key] = {"array": f"r{i}/{key}", "device": True}
            elif isinstance(val, np.ndarray):
                arrays[f"r{i}/{key}"] = val
                scal[key] = {"array": f"r{i}/{key}"}
            elif isinstance(val, (np.floating, np.integer)):
                scal[key] = val.item()
            else:
                scal[key] = val
        iters.append(scal)
    state = []
    for name in _STATE:
        val = simu.__dict__.get(name)
        if isinstance(val, torch.Tensor):
            arrays[f"s/{name}"] = val.detach().cpu().numpy()
            state.append(name)
    bcs = {}
    for store in ("_Bc_Dirichlet", "_Bc_Neumann"):
        rows = []
        for j, bc in enumerate(getattr(simu, store)):
            for f in ("nodes", "dofs", "values"):
                arrays[f"b{store}{j}/{f}"] = getattr(bc, f)
            rows.append({"problemType": bc.problemType, "unknowns": list(bc.unknowns),
                         "description": bc.description, "arrays": f"b{store}{j}"})
        bcs[store] = rows
    header = {
        "class": type(simu).__name__,
        "bcs": bcs,
        "mesh": os.path.relpath(mesh_path, folder),
        "model": _model_header(simu.model, arrays),
        "solver": str(simu.solver.value),
        "algo": str(simu.algo.value),
        "parabolicParams": simu._parabolicParams,
        "hyperbolicParams": simu._hyperbolicParams,
        "Niter": simu.Niter,
        "iterations": iters,
        "state": state,
    }
    if hasattr(simu, "_rho"):
        header["rho"] = _param(simu._rho, "rho", arrays)
    if hasattr(simu, "_rayleigh"):
        header["rayleigh"] = list(simu._rayleigh)
    path = os.path.join(folder, f"{filename}.npz")
    np.savez(path, header=np.array(json.dumps(header)), **arrays)
    summary = f"Simulation completed on: {datetime.now()}\n{type(simu).__name__} on {simu.mesh.Nn} nodes, " \
              f"{simu.mesh.Ne} {simu.mesh.elemType.value} elements; model {header['model']['kind']}; " \
              f"solver {header['solver']}; {len(iters)} saved iterations\n"
    if str(additionalInfos):
        summary += "\nAdditional information\n" + str(additionalInfos) + "\n"
    with open(os.path.join(folder, "summary.txt"), "w", encoding="utf8") as fh:
        fh.write(summary)
    return path


def load_simu(folder: str, filename: str = "simulation"):
    """``Load_Simu``: the simulation saved by ``save_simu`` in ``folder``."""
    import torch

    from . import Simulations
    from .Solvers import AlgoType
    from .device import to_dev

    path = os.path.join(folder, f"{filename}.npz")
    assert os.path.exists(path), f"The file {filename}.npz cannot be found."
    with np.load(path, allow_pickle=False) as z:
        h = json.loads(str(z["header"]))
        mesh = load_mesh(os.path.join(folder, h["mesh"]))
        model = _model_from(h["model"], z)
        cls = {"Elastic": Simulations.Elastic, "Thermal": Simulations.Thermal,
               "PhaseField": Simulations.PhaseField}[h["class"]]
        simu = cls(mesh, model, folder)
        simu.solver = h["solver"]
        if "rho" in h:
            simu.rho = _unparam(h["rho"], z)
        if "rayleigh" in h:
            simu._rayleigh = tuple(h["rayleigh"])
        simu._algo = AlgoType(h["algo"])
        simu._parabolicParams = tuple(h["parabolicParams"]) if h["parabolicParams"] else None
        simu._hyperbolicParams = tuple(h["hyperbolicParams"]) if h["hyperbolicParams"] else None
        results = []
        for scal in h["iterations"]:
            it = {}
            for key, val in scal.items():
                if isinstance(val, dict) and "array" in val:
                    arr = np.array(z[val["array"]])
                    it[key] = to_dev(arr) if val.get("device") else arr
                else:
                    it[key] = val
            results.append(it)
        simu._results = results
        for store, rows in h["bcs"].items():
            setattr(simu, store, [Simulations._Bc(r["problemType"], z[r["arrays"] + "/nodes"], z[r["arrays"] + "/dofs"],
                                                  z[r["arrays"] + "/values"], r["unknowns"], r["description"])
                                  for r in rows])
        simu.Niter = h["Niter"]
        for name in h["state"]:
            setattr(simu, name, to_dev(np.array(z[f"s/{name}"])).to(torch.float64))
    if hasattr(simu, "_updD"):  # phase field: operators are rebuilt from the restored fields
        simu._updD = simu._updU = False
    return simu
