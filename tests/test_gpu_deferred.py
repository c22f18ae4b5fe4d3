"""Deferred pipelined solves (``efea_pcgp_start`` / ``efea_pcgp_finish``
for block Jacobi, ``efea_pcg2p_start`` / ``efea_pcg2p_finish`` for the
two-level solve; GPU tests).  The staggered loop queues each damage solve
that has an expected iteration count, assembles the displacement system
behind it and reads the damage solve's end only before the displacement
solve; it queues the displacement solve likewise, with the damage assembly
and energy of the convergence measure behind it.  The split solve returns
exactly what the whole one does: the same x, bit for bit, and the same info,
whether the queued iterations were enough or the finish had to continue
(``resumed``); a staggered history with deferral matches one without it,
also when every solve is forced to continue (the work queued behind it is
then redone)."""

import ctypes as C

import numpy as np
import pytest
import torch

import easyfea_amd as ea
from easyfea_amd import Models, Simulations, _lib
from easyfea_amd import device as dv

pytestmark = pytest.mark.gpu


def _cracked(n=50, solver="hip_fp64"):
    mesh = ea.tri3_notched_square(n)
    mat = Models.Elastic.Isotropic(2, 210000.0, 0.3, planeStress=False, thickness=1.0)
    simu = Simulations.PhaseField(mesh, Models.PhaseField(mat, "Miehe", "AT2", 2.7, 0.02))
    simu.solver = solver
    top = mesh.Nodes_Conditions(lambda x, y, z: y == 1.0)
    bot = mesh.Nodes_Conditions(lambda x, y, z: y == 0.0)
    return simu, top, bot


def _history(simu, top, bot, deps):
    out = []
    for dep in deps:
        simu.Bc_Init()
        simu.add_dirichlet(top, [0, dep], ["x", "y"])
        simu.add_dirichlet(bot, [0], ["y"])
        u, d, _ = simu.Solve(1.0, 10, 1)
        simu.Save_Iter()
        i, e = simu.last_solve["damage"], simu.last_solve["elastic"]
        out.append((u, d, i.iterations, i.pipelined, e.iterations, float(simu.Calc_Reaction().sum()),
                    e.pipelined and e.precond == "two-level", simu.convIter))
    return out


@pytest.mark.parametrize("solver", ["hip_fp64", "hip_mixed"])
def test_deferred_solve_matches_whole(solver, monkeypatch):
    monkeypatch.setattr(dv, "MIXED_MIN_MATRIX_BYTES", 0)
    simu, top, bot = _cracked(50, solver)
    _history(simu, top, bot, [4e-3, 6e-3, 7e-3])
    Kd = simu._device_K("damage")
    b = simu._Fd.clone()
    x0 = simu._d.clone()

    def job_for(expected):
        x = x0.clone()
        job = simu._make_job(Kd, b, x, "damage")
        job.expected_iterations = expected
        return job, x

    job, x_whole = job_for(40)
    whole = job.run()
    assert whole.pipelined and whole.status == 1, whole
    n = whole.iterations
    print(f"{solver}: whole solve {n} iterations, mode {whole.mode}")
    for expected, want_resumed in ((n, False), (max(1, n // 4), True)):
        job, x = job_for(expected)
        pending = job.start()
        assert pending is not None, "a pipelined block-Jacobi solve with an expected count is deferred"
        info = pending.finish()
        print(f"  expected {expected}: queued {dv.deferred_budget(expected)}, resumed {pending.resumed}, "
              f"{info.iterations} iterations")
        assert pending.resumed == want_resumed
        assert torch.equal(x, x_whole)
        assert (info.iterations, info.status, info.replacements, info.mode) == \
            (whole.iterations, whole.status, whole.replacements, whole.mode)
        assert info.rel_residual == whole.rel_residual


def test_deferred_solve_not_taken_without_expected_count():
    simu, top, bot = _cracked(30)
    _history(simu, top, bot, [4e-3])
    Kd = simu._device_K("damage")
    job = simu._make_job(Kd, simu._Fd.clone(), simu._d.clone(), "damage")
    job.expected_iterations = None
    assert job.start() is None
    job.expected_iterations = 30
    job.pipelined_mode = False  # classic solves are never deferred
    assert job.start() is None


@pytest.mark.parametrize("force_resume", [False, True])
def test_deferred_history_matches_plain(force_resume, monkeypatch):
    if force_resume:
        # two queued iterations: every deferred damage solve continues in its
        # finish and the displacement system is assembled a second time
        monkeypatch.setattr(dv, "deferred_budget", lambda expected, kind="pcgp": 2)
    deps = [2e-3, 4e-3, 5.5e-3, 6.5e-3, 7e-3, 7.5e-3]
    a, top, bot = _cracked(50)
    b, _, _ = _cracked(50)
    b.defer_damage = b.defer_displacement = False
    calls = {"pcgp": [0, 0], "pcg2p": [0, 0]}
    orig = dv.PendingSolve.finish

    def counting(self, raise_on_fail=True):
        info = orig(self, raise_on_fail)
        calls[self.kind][0] += 1
        calls[self.kind][1] += int(self.resumed)
        return info

    monkeypatch.setattr(dv.PendingSolve, "finish", counting)
    ha = _history(a, top, bot, deps)
    hb = _history(b, top, bot, deps)
    print(f"force_resume {force_resume}: deferred finishes and continuations {calls}, damage iterations "
          f"{[r[2] for r in ha]}, displacement {[r[4] for r in ha]}, pipelined {[r[6] for r in ha]}")
    # every damage solve that follows one with iterations (its expected count)
    # is deferred; the first increment's damage is 0 (no iterations); every
    # displacement solve after the first, when it runs the pipelined solve
    deferred = sum(1 for k in range(1, len(deps)) if ha[k - 1][2] > 0)
    assert deferred >= len(deps) - 2 and calls["pcgp"][0] == deferred
    deferred_u = sum(1 for k in range(1, len(deps)) if ha[k][6])
    assert calls["pcg2p"][0] == deferred_u
    if force_resume:
        assert calls["pcgp"][1] == deferred and calls["pcg2p"][1] == deferred_u
    for ra, rb in zip(ha, hb):
        assert ra[2:5] == rb[2:5], (ra[2:5], rb[2:5])
        assert np.array_equal(ra[0], rb[0]) and np.array_equal(ra[1], rb[1])
        assert ra[5] == rb[5] and ra[7] == rb[7]  # reaction and convergence measure


def test_deferred_abi_roundtrip():
    """The C entry points themselves: start then finish on a small damage
    system gives efea_solve_pcgp's x and info."""
    simu, top, bot = _cracked(30)
    _history(simu, top, bot, [4e-3, 6e-3])
    Kd = simu._device_K("damage")
    job = simu._make_job(Kd, simu._Fd.clone(), simu._d.clone(), "damage")
    job.pipelined_mode = True
    job._prepare()
    ws = dv.Workspace()
    x0 = job.x.clone()
    args = job._pcgp_args(ws, 8, 12)
    s = dv.stream_ptr()
    raw = (C.c_double * 4)()
    assert _lib.call("efea_solve_pcgp", *args, s, C.cast(raw, C.c_void_p)) == 0
    x_whole, info_whole = job.x.clone(), list(raw)
    job.x.copy_(x0)
    assert _lib.call("efea_pcgp_start", *args, s) == 0
    raw2 = (C.c_double * 4)()
    resumed = C.c_int(-1)
    assert _lib.call("efea_pcgp_finish", *args, s, C.cast(raw2, C.c_void_p), C.byref(resumed)) == 0
    print(f"whole {info_whole}, split {list(raw2)}, resumed {resumed.value}")
    assert info_whole[0] > 20 and resumed.value == 1  # 12 queued iterations were not enough
    assert torch.equal(job.x, x_whole) and list(raw2) == info_whole


@pytest.mark.parametrize("solver", ["hip_fp64", "hip_mixed"])
def test_deferred_two_level_solve_matches_whole(solver, monkeypatch):
    """The displacement system of a cracked body (two-level pipelined PCG):
    queued with the expected count (the finish finds it done) and with a
    quarter of it (the finish continues, across residual replacements)."""
    monkeypatch.setattr(dv, "MIXED_MIN_MATRIX_BYTES", 0)
    simu, top, bot = _cracked(50, solver)
    simu.defer_displacement = False
    _history(simu, top, bot, [4e-3, 6e-3, 7e-3])
    Ku = simu._device_K("elastic")
    b = simu._neumann_device("elastic")
    x0 = torch.zeros_like(simu._u)  # a full solve (the last displacement is this system's solution)

    def job_for(expected):
        x = x0.clone()
        job = simu._make_job(Ku, b, x, "elastic")
        job.before_solve = None
        job.pipelined_mode = True
        job.expected_iterations = expected
        return job, x

    job, x_whole = job_for(150)
    whole = job.run()
    assert whole.pipelined and whole.precond == "two-level" and whole.status == 1, whole
    n = whole.iterations
    print(f"{solver}: whole solve {n} iterations, mode {whole.mode}, replacements {whole.replacements}")
    for expected, want_resumed in ((n, False), (max(1, n // 4), True)):
        job, x = job_for(expected)
        pending = job.start()
        assert pending is not None and pending.kind == "pcg2p"
        info = pending.finish()
        print(f"  expected {expected}: queued {dv.deferred_budget(expected, 'pcg2p')}, resumed {pending.resumed}")
        assert pending.resumed == want_resumed
        assert torch.equal(x, x_whole)
        assert (info.iterations, info.status, info.replacements, info.mode, info.rel_residual) == \
            (whole.iterations, whole.status, whole.replacements, whole.mode, whole.rel_residual)


@pytest.mark.parametrize("conv_option", [0, 2, 3])
def test_deferred_history_other_measures(conv_option, monkeypatch):
    """tolConv < 1 with several passes per increment: the measures of
    convOption 0, 2 and 3 (host-read in 3) with every deferred solve forced
    to continue give the plain loop's passes, fields and measures."""
    monkeypatch.setattr(dv, "deferred_budget", lambda expected, kind="pcgp": 2)
    a, top, bot = _cracked(30)
    b, _, _ = _cracked(30)
    b.defer_damage = b.defer_displacement = False
    for dep in [4e-3, 6e-3, 7e-3]:
        out = []
        for s in (a, b):
            s.Bc_Init()
            s.add_dirichlet(top, [0, dep], ["x", "y"])
            s.add_dirichlet(bot, [0], ["y"])
            u, d, conv = s.Solve(1e-3, 20, conv_option)
            s.Save_Iter()
            out.append((u, d, conv, s.Niter_staggered, s.convIter))
        print(f"convOption {conv_option}, dep {dep}: passes {out[0][3]} / {out[1][3]}, measure {out[0][4]}")
        assert out[0][2:] == out[1][2:]
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
