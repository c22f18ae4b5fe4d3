"""Continuous-batching request loop for many simulations on one GPU.

A request is a generator of ``SolveJob`` s (a simulation's ``Solve`` written
as ``yield job`` / ``info = ...``, e.g. ``Elastic._solve_gen`` or
``PhaseField._staggered``).  The loop

* admits requests in arrival order while their estimated device memory fits
  the HBM budget (``hbm_budget`` bytes), as soon as they have arrived and
  memory is free;
* starts an admitted request at once: its assembly kernels go onto the
  stream between two chunks of iterations, and its first solve joins the
  running batch;
* advances every active solve ``check_every`` PCG iterations per chunk, all
  systems of one (precision, DOF per node, lanes per node) class in one
  launch per PCG step, each class on its own stream (``efea_batch_launch``);
* while a chunk runs, hands each solve finished in the previous chunk its
  ``SolveInfo``: its request either yields its next solve (host work and
  assembly kernels overlap the chunk; the solve joins at the next chunk) or
  completes (``efea_batch_wait`` collects the chunk).

Latency of a request = completion time - arrival time; the loop records
admission and completion times and the per-request results.  The reference
has no serving layer: each simulation calls ``Solve`` in its own process.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Any, Callable, Generator, Optional

import numpy as np
import torch

from .device import BatchExecutor, SolveJob


@dataclass
class Request:
    rid: int
    arrival: float
    start: Callable[[], Generator]  # builds the simulation, returns its job generator
    est_bytes: int
    admitted: Optional[float] = None
    finished: Optional[float] = None
    result: Any = None
    error: Optional[BaseException] = None
    solves: int = 0
    _gen: Optional[Generator] = field(default=None, repr=False)

    @property
    def latency(self) -> Optional[float]:
        return None if self.finished is None else self.finished - self.arrival


def estimate_bytes(Nn: int, Ne: int, nPe: int, dofs: tuple, nnz: tuple, fp32_copy: bool) -> int:
    """Device memory of one simulation: per system (dof_n, nnz) the FP64 and
    optional FP32 values, the element entries, block-Jacobi inverse, solver
    workspace and vectors; plus the mesh and per-element fields."""
    total = Nn * 3 * 8 + Ne * nPe * 4 + Ne * 8 * 8
    for dn, nz in zip(dofs, nnz):
        ndof = Nn * dn
        nd = nPe * dn
        total += nz * (8 + (4 if fp32_copy else 0))
        total += Ne * nd * nd * 8            # element matrices before the scatter
        total += Nn * dn * dn * 8            # block-Jacobi inverse
        total += ndof * 8 * 12 + 64 * 1024   # workspace vectors, b, x, history, partials
    return int(total)


class Server:
    """Admission by HBM budget + continuous batching of the requests' solves."""

    def __init__(self, hbm_budget: int, capacity: int = 256, check_every: int = 32,
                 clock: Callable[[], float] = time.perf_counter, executor=None):
        self.hbm_budget = int(hbm_budget)
        self.exec = executor if executor is not None else BatchExecutor(capacity=capacity, check_every=check_every)
        self.clock = clock
        self.max_active = 0
        self.max_resident = 0
        self.chunks = 0

    def _advance(self, req: Request, info, now_fn) -> None:
        try:
            job = next(req._gen) if info is None else req._gen.send(info)
        except StopIteration as stop:
            req.result = stop.value
            req.finished = now_fn()
            req._gen = None
            return
        except Exception as exc:  # a failed request does not stop the others
            req.error = exc
            req.finished = now_fn()
            req._gen = None
            return
        job.request = req
        req.solves += 1
        self.exec.submit(job)

    def run(self, requests: list, time_limit: float = 3600.0) -> list:
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        active: list = []
        used = 0
        t0 = self.clock()

        def now() -> float:
            return self.clock() - t0

        done_prev: list = []  # solves finished in the last chunk, handed back during the next one
        while pending or active:
            t = now()
            if t > time_limit:
                raise TimeoutError(f"serving loop exceeded {time_limit} s")
            # admission: arrived requests, in order, while memory allows
            while pending and pending[0].arrival <= t and (used + pending[0].est_bytes <= self.hbm_budget
                                                           or not active):
                req = pending.pop(0)
                req.admitted = now()
                used += req.est_bytes
                try:
                    req._gen = req.start()
                except Exception as exc:
                    req.error, req.finished = exc, now()
                    used -= req.est_bytes
                    continue
                active.append(req)
                self._advance(req, None, now)
            # one chunk of every batch, enqueued without waiting; while it runs
            # the host hands the previous chunk's finished solves back to their
            # requests (their next assembly and solve set-up go onto the
            # current stream, their solves join the batches at the next chunk)
            launched = False
            if self.exec.active():
                self.max_active = max(self.max_active, self.exec.active())
                self.max_resident = max(self.max_resident, used)
                self.chunks += 1
                launched = self.exec.launch()
            for job, info in done_prev:
                self._advance(job.request, info, now)
            done_prev = self.exec.collect() if launched else []
            # retire completed requests
            still = []
            for req in active:
                if req._gen is None:
                    used -= req.est_bytes
                else:
                    still.append(req)
            active = still
            if not active and pending and pending[0].arrival > now():
                wait = pending[0].arrival - now()
                if wait > 0:
                    time.sleep(min(wait, 0.05))
        if torch.cuda.is_available():
            torch.cuda.synchronize()
        return requests

    def close(self) -> None:
        self.exec.close()


def percentiles(values, ps=(50, 99)) -> dict:
    v = np.asarray([x for x in values if x is not None], dtype=float)
    if v.size == 0:
        return {f"p{p}": None for p in ps}
    return {f"p{p}": float(np.percentile(v, p)) for p in ps}
