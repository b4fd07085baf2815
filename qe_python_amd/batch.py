"""Many independent Davidson problems (k-points, cells) on one GPU at once.

The reference solves the k-points of an SCF iteration one after another
(``qepy_pw/pw/scf.py:3340`` loop, ``davidson`` call at ``3510-3537``).  On
one MI355X a single 10k-24k plane-wave problem leaves the device idle while
its host thread factorises the small FP64 reduced matrices, and its own
launches are too small to fill 256 CUs late in a solve when few roots remain
unconverged.  ``davidson_batch`` runs the problems together, in one of two
schedules:

* ``"threads"`` (default): every problem's unchanged ``davidson`` on its own
  HIP stream (default priority) from one of ``workers`` host threads;
  kernels of different problems overlap through the hardware queues, and
  the host side of each solve -- the reduced eigensolve's native calls
  (``qepw_rr_standardize`` / ``qepw_rr_spectral``, which release the GIL),
  its LAPACK on a single-threaded BLAS -- runs while the other threads
  launch.  Three workers measured best on one MI355X with the 24-problem
  uneven workload (1.8x over one by one, 15% over two); more are slower
  with four hardware queues per process and with eight alike, so what
  limits them is the host: the Python parts of the solves take turns
  holding the GIL.
* ``"rounds"``: every problem's Davidson is the generator
  ``davidson_steps``; ``schedule.run_many`` advances all of them in lock-step
  rounds on one host thread and one stream.  In a round, H is applied to
  every requested block of one precision at once -- the FFT phase of all
  problems as one grouped launch per stage (``qepw_apply_local_group``:
  sticks, planes, sticks, each workgroup picking its problem from a prefix
  table) -- and every requested host value (residual norms, reduced
  matrices) is read back after one wait.  Its reduced eigensolves stay in
  the yielding Python steps, so its host side is single-threaded.

Results are identical to solving the problems one by one in both schedules:
every kernel is deterministic (fixed-order reductions, no atomics), the
grouped FFT kernels run each problem's bands through the same code as the
single-problem kernels, and each problem owns its Hamiltonian, workspace and
reduced-problem state.
"""

from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import torch

from . import _native as nat
from . import device as dv
from . import schedule as sched
from .diagonalization import DavidsonResult, davidson, davidson_steps
from .eigensolver import _one_blas_thread


_STREAMS: dict = {}


def _worker_streams(dev: torch.device, count: int) -> list:
    """Long-lived default-priority streams of a device: the caching
    allocator keeps one pool per stream, so reusing the streams keeps
    allocations out of later batches."""
    pool = _STREAMS.setdefault(dev.index, [])
    while len(pool) < count:
        pool.append(torch.cuda.Stream(device=dev))
    return pool[:count]


@dataclass
class BatchProblem:
    """One ``davidson`` call of the batch: the reference's per-k arguments
    (``operator``, ``diagonal``, ``number_of_roots``, ``initial_vectors``,
    ``operator_into``) plus per-problem keyword overrides."""

    operator: object
    diagonal: object
    number_of_roots: int
    initial_vectors: object = None
    operator_into: object = None
    options: dict = field(default_factory=dict)


def davidson_batch(problems, *, workers: int | None = None, device=None,
                   schedule: str = "threads", stats: "sched.RoundStats | None" = None,
                   **common) -> list[DavidsonResult]:
    """Solve ``problems`` (a sequence of ``BatchProblem``) together.

    ``common`` holds ``davidson`` keywords shared by all problems (a
    problem's ``options`` override them).  ``schedule`` is "threads"
    (``workers`` host threads, default 3, each driving one problem at a time
    on its own stream) or "rounds" (all problems in lock-step rounds with
    grouped H applications, on the current stream).  Device tensors
    returned with ``device_output=True`` are complete when this returns.
    Problems must not share an operator object (each owns its device
    workspace).  ``stats`` (a ``schedule.RoundStats``) collects the rounds,
    grouped H applications and reads of the "rounds" schedule."""
    problems = list(problems)
    if schedule not in ("rounds", "threads"):
        raise ValueError("schedule must be 'rounds' or 'threads'")
    if not problems:
        return []
    owners = [getattr(p.operator_into or p.operator, "__self__", None) for p in problems]
    shared = [o for o in owners if o is not None]
    if len({id(o) for o in shared}) != len(shared):
        raise ValueError("davidson_batch: problems must not share an operator")
    dev = dv.require_device(device)
    nat.load()
    if schedule == "rounds":
        return _batch_rounds(problems, dev, common, stats)
    return _batch_threads(problems, dev, workers, common)


def _kwargs(problem, common) -> dict:
    kwargs = dict(common)
    kwargs.update(problem.options)
    return kwargs


def _batch_rounds(problems, dev, common, stats) -> list[DavidsonResult]:
    with torch.cuda.device(dev):
        solves = [davidson_steps(p.operator, p.diagonal, p.number_of_roots,
                                 initial_vectors=p.initial_vectors,
                                 operator_into=p.operator_into, **_kwargs(p, common))
                  for p in problems]
        results = sched.run_many(solves, dev, stats)
        torch.cuda.current_stream(dev).synchronize()
    return results


def _batch_threads(problems, dev, workers, common) -> list[DavidsonResult]:
    workers = max(1, min(len(problems), 3 if workers is None else int(workers)))
    streams = _worker_streams(dev, workers)
    # operands uploaded on the caller's stream must be complete before the
    # worker streams read them
    torch.cuda.synchronize(dev)
    slots = list(range(workers))
    results: list[DavidsonResult | None] = [None] * len(problems)

    def solve(index: int) -> None:
        slot = slots.pop()
        try:
            stream = streams[slot]
            problem = problems[index]
            with torch.cuda.device(dev), torch.cuda.stream(stream):
                results[index] = davidson(problem.operator, problem.diagonal,
                                          problem.number_of_roots,
                                          initial_vectors=problem.initial_vectors,
                                          operator_into=problem.operator_into,
                                          **_kwargs(problem, common))
                stream.synchronize()
        finally:
            slots.append(slot)

    # one BLAS thread for every host factorisation of every worker
    with _one_blas_thread(), ThreadPoolExecutor(max_workers=workers) as pool:
        for future in [pool.submit(solve, i) for i in range(len(problems))]:
            future.result()
    return results
