"""Cooperative scheduling of device Davidson solves.

A solve is written as a generator that hands its host-side waits to a
driver instead of blocking on them: it yields

* ``Read(tensor)`` where it needs a device value on the host (a residual
  norm, a small Gram or Rayleigh matrix for a host factorisation) and
  receives the host array back;
* ``Apply(op, src, dst)`` where it needs H applied to a device block.

``run`` drives one solve: a read is a blocking copy, an application runs at
once -- exactly the sequence of the plain loop.  ``run_many`` drives several
independent solves (k-points, cells) in rounds on one host thread: each
round advances every solve to its next request, applies H to all requested
blocks of one precision in one grouped launch per FFT stage
(``PlaneWaveHamiltonian.device_apply_group``), then reads every requested
value back after a single wait.  Device work of all problems queues back to
back, and the host-side factorisations of one problem no longer wait for
another problem's kernels.
"""

from __future__ import annotations

import threading
from contextlib import contextmanager

import torch

from . import device as dv


class Read:
    """A device tensor the solve needs on the host (sent back as numpy)."""

    __slots__ = ("tensor",)

    def __init__(self, tensor: torch.Tensor):
        self.tensor = tensor


class Apply:
    """dst = H src for (nb, npw) device-order blocks of operator ``op``."""

    __slots__ = ("op", "src", "dst")

    def __init__(self, op, src: torch.Tensor, dst: torch.Tensor):
        self.op = op
        self.src = src
        self.dst = dst


_TLS = threading.local()


def may_block() -> bool:
    """Whether the driver of the running steps waits on each read itself
    (``run``, ``blocking``): steps may then do their host round trips inside
    one native call instead of yielding them (``run_many`` needs the yields
    to interleave its solves)."""
    return getattr(_TLS, "blocking", 0) > 0


@contextmanager
def _blocking_driver():
    _TLS.blocking = getattr(_TLS, "blocking", 0) + 1
    try:
        yield
    finally:
        _TLS.blocking -= 1


@contextmanager
def _yielding_driver():
    prev = getattr(_TLS, "blocking", 0)
    _TLS.blocking = 0
    try:
        yield
    finally:
        _TLS.blocking = prev


def blocking(steps):
    """Drive a generator that only reads (no H applications), with blocking
    copies, on whatever stream is current."""
    with _blocking_driver():
        try:
            request = next(steps)
            while True:
                if type(request) is not Read:
                    raise TypeError("blocking() drives read-only steps")
                request = steps.send(request.tensor.cpu().numpy())
        except StopIteration as stop:
            return stop.value


def _one_blas_thread():
    from .eigensolver import _one_blas_thread as limit
    return limit()


def run(steps, device: torch.device):
    """Drive one solve to its return value, on the current stream.  The
    host LAPACK work of the solve (small factorisations between device
    steps) runs on one BLAS thread: the limit is process-wide, so the driver
    holds it for the whole drive rather than the steps around each call."""
    with dv.fixed_stream(device), _one_blas_thread(), _blocking_driver():
        try:
            request = next(steps)
            while True:
                if type(request) is Read:
                    value = request.tensor.cpu().numpy()
                else:
                    request.op.device_apply(request.src, request.dst)
                    value = None
                request = steps.send(value)
        except StopIteration as stop:
            return stop.value


class RoundStats:
    """What ``run_many`` did: rounds, H applications and how they were
    grouped, device-to-host reads (each round's reads share one wait)."""

    def __init__(self):
        self.rounds = 0
        self.applies = 0            # Apply requests
        self.grouped_calls = 0      # device_apply_group calls (one per precision per round)
        self.reads = 0
        self.waits = 0

    def as_dict(self) -> dict:
        return {"rounds": self.rounds, "applies": self.applies,
                "grouped_calls": self.grouped_calls, "reads": self.reads, "waits": self.waits}


def run_many(solves, device: torch.device, stats: RoundStats | None = None) -> list:
    """Drive independent solves in rounds; returns their results in order.

    Each round: every unfinished solve runs until its next request; the
    ``Apply`` requests are grouped by precision and applied together, the
    ``Read`` requests are copied into pinned host memory behind one event,
    and after a single wait every solve gets its value.  All work goes to the
    current stream, so the device runs one problem's kernels while the host
    prepares another's."""
    from .diagonalization import PlaneWaveHamiltonian

    results = [None] * len(solves)
    pending = {}                    # index -> request
    with dv.fixed_stream(device), _one_blas_thread(), _yielding_driver():
        stream = torch.cuda.current_stream(device)
        for i, steps in enumerate(solves):
            try:
                pending[i] = next(steps)
            except StopIteration as stop:
                results[i] = stop.value
        while pending:
            values = {}
            applies = [(i, r) for i, r in pending.items() if type(r) is Apply]
            groups = {}
            for i, r in applies:
                groups.setdefault(r.src.dtype, []).append(r)
            for group in groups.values():
                if all(isinstance(r.op, PlaneWaveHamiltonian) for r in group):
                    PlaneWaveHamiltonian.device_apply_group(
                        [(r.op, r.src, r.dst) for r in group])
                    if stats is not None:
                        stats.grouped_calls += 1
                else:
                    for r in group:
                        r.op.device_apply(r.src, r.dst)
            for i, _ in applies:
                values[i] = None
            reads = [(i, r) for i, r in pending.items() if type(r) is Read]
            staged = []
            for i, r in reads:
                host = torch.empty(r.tensor.shape, dtype=r.tensor.dtype, pin_memory=True)
                host.copy_(r.tensor, non_blocking=True)
                staged.append((i, host))
            if staged:
                stream.synchronize()
                for i, host in staged:
                    values[i] = host.numpy()
            if stats is not None:
                stats.rounds += 1
                stats.applies += len(applies)
                stats.reads += len(staged)
                stats.waits += 1 if staged else 0
            for i in list(pending):
                try:
                    pending[i] = solves[i].send(values[i])
                except StopIteration as stop:
                    results[i] = stop.value
                    del pending[i]
    return results
