"""Drop-in, device-resident replacement of ``qepy_pw/diagonalization.py``.

Same names, signatures, argument meaning and exception types as the
reference module for the h_psi / Davidson path, so that
``qepy_pw/pw/scf.py:29-40`` can bind them unchanged:

* ``PlaneWaveHamiltonian``  (reference ``diagonalization.py:124-369``)
* ``davidson`` / ``DavidsonResult`` (``diagonalization.py:527-880, 113-121``)
* ``FactorizedProjectorTerm`` / ``PackedProjectorTerm`` / ``ProjectorTerm``
  (``diagonalization.py:21-43``)
* ``_lowest_generalized_eigh`` / ``_matmul_f`` (``diagonalization.py:74-110``)

Design (one GPU):

* The Hamiltonian owns a device FFT plan whose rows are in "device order"
  (plane waves grouped by z-stick), V(r) in plane-major (n3, n2, n1) layout,
  the kinetic diagonal, and all projector terms packed into one device beta
  (npw, P) plus a block-diagonal D.  ``apply``/``apply_into`` accept host
  arrays in the reference's row order (or device tensors) and permute at the
  boundary.
* ``davidson`` keeps ``basis_storage``/``applied_storage`` on the device in
  device order.  h_psi, every subspace product (MFMA GEMMs), the residual,
  preconditioner, norms and normalisation run as HIP kernels.  The reduced
  pencil stays on the device too (``qe_python_amd.eigensolver``): FP64
  standardisation and a spectral-division eigensolve built from GEMMs, with
  LAPACK zhegvd on the host (the routine the reference calls from
  ``_native_fft.pyx:667-691``) for the first solve and as the fallback.
  Convergence flags are host scalars, as in the reference.
* ``precision="mixed"`` runs the Davidson in complex64 storage with FP32
  h_psi and GEMMs and hands its Ritz vectors to an FP64 refinement.
* Operators that are not a device ``PlaneWaveHamiltonian`` are still
  accepted: their blocks are copied to the host for each application.

There is no CPU fallback: without a HIP device or without
``libqepw_hip.so`` every entry point raises.
"""

from __future__ import annotations

import ctypes
import dataclasses
from contextlib import nullcontext
from dataclasses import dataclass
from typing import Callable

import numpy as np
import scipy.linalg as sla
import torch

from . import _native as nat
from . import device as dv
from .eigensolver import DeviceRayleighRitz, _lower_inverse, _one_blas_thread
from . import schedule
from .schedule import Apply, Read, blocking, may_block

# default reduced eigensolve of ``davidson``: "device" (GPU spectral
# division) or "host" (LAPACK zhegvd, the reference's routine)
REDUCED_EIGENSOLVER = "device"
# exit of the complex64 phase of the mixed Davidson: |de| below 1e-5 Ha and
# residual norms below 1e-5, about ten times the FP32 noise floor of both
# (unit roundoff 6e-8 times |H| ~ 15 Ha for a 30 Ry cutoff)
MIXED_SWITCH = (1.0e-5, 1.0e-5)
# host Hermitian eigensolves of the complex64 phase in single precision
# (LAPACK cheevd instead of zheevd)
MIXED_SINGLE_EIGH = True
# a complex64 phase whose largest unconverged residual is within 100x of the
# switch residual and has not halved over this many iterations sits at the
# FP32 floor and hands over (far above the floor, slow progress is the
# ordinary slow tail of the top roots and the cheap phase continues)
MIXED_STALL_WINDOW = 4
MIXED_STALL_SPAN = 100.0
# the complex64 phase locks, at a restart, the converged Ritz pairs at the
# bottom of the spectrum (at least this many, ending at a gap of at least
# MIXED_LOCK_GAP) and continues on the others only
MIXED_LOCK_MIN = 8
MIXED_LOCK_GAP = 1.0e-3     # Ha
# classical Gram-Schmidt passes of the complex64 phase's corrections against
# the basis, before their Cholesky QR (on the medium workload a second pass
# leaves the iteration count at 32 and costs ~4% of the solve)
ORTHO_PASSES = 1
# the complex64 corrections' Cholesky QR as one native call (qepw_cholesky_qr)
# when the driver waits on reads itself; False keeps the Python steps
NATIVE_ORTHO = True
# subspace multiplier of the complex64 phase (None: the caller's
# subspace_multiplier, as the FP64 refinement always uses)
MIXED_COARSE_MULTIPLIER = None
# diagnostics: set to a list to receive one record per Davidson iteration
# (storage dtype, iteration, unconverged roots, max |de|, max residual norm
# or None when residuals are not formed)
ITERATION_TRACE = None

BlockOperator = Callable[[np.ndarray], np.ndarray]
BlockOperatorInto = Callable[[np.ndarray, np.ndarray], np.ndarray]
MaterializedProjectorTerm = tuple[np.ndarray, np.ndarray]


@dataclass(frozen=True)
class FactorizedProjectorTerm:
    """Species-centered beta functions plus atom-dependent phase columns."""

    beta: np.ndarray
    phases: np.ndarray
    coupling: np.ndarray


@dataclass(frozen=True)
class PackedProjectorTerm:
    """QE-style atom-packed beta matrix with a precomputed diagonal."""

    beta: np.ndarray
    coupling: np.ndarray
    diagonal: np.ndarray


ProjectorTerm = MaterializedProjectorTerm | FactorizedProjectorTerm | PackedProjectorTerm


@dataclass(frozen=True)
class DavidsonResult:
    eigenvalues: np.ndarray
    eigenvectors: np.ndarray
    residual_norms: np.ndarray
    iterations: float
    converged: bool
    number_unconverged: int = 0
    hamiltonian_applications: int = 0


class _SerialContext:
    """The reference's serial ``MPIContext`` (``qepy_pw/mpi.py:14-31,129-136``)."""

    rank = 0
    size = 1
    comm = None

    def sum_array(self, value):
        return np.ascontiguousarray(value)


_SERIAL = _SerialContext()


def _matmul_f(left: np.ndarray, right: np.ndarray) -> np.ndarray:
    """Host product into a column-major block (``diagonalization.py:74-80``)."""
    result = np.empty((left.shape[0], right.shape[1]), dtype=np.complex128, order="F")
    np.matmul(left, right, out=result)
    return result


def _lowest_generalized_eigh(hamiltonian, overlap, roots):
    """Lowest ``roots`` pairs of H c = e S c by LAPACK zhegvd(itype 1, 'V', 'L')
    on the host (``_native_fft.pyx:667-691``; ``diagonalization.py:83-110``).
    A non-positive-definite overlap raises ``numpy.linalg.LinAlgError``."""
    h = np.array(hamiltonian, dtype=np.complex128, order="F", copy=True)
    s = np.array(overlap, dtype=np.complex128, order="F", copy=True)
    try:
        values, vectors = sla.eigh(h, s, lower=True, type=1, driver="gvd",
                                   overwrite_a=True, overwrite_b=True, check_finite=False)
    except sla.LinAlgError as exc:  # scipy's LinAlgError subclasses numpy's
        raise np.linalg.LinAlgError(f"zhegvd failed: {exc}") from exc
    return values[:roots].copy(), np.asfortranarray(vectors[:, :roots])


def _measure(timers, label, device=None):
    if timers is None:
        if dv.GEMM_TRACE is not None:
            # GEMMs traced by class: the timer label names the class
            return dv.gemm_class(label)
        return nullcontext()
    return _SyncedMeasure(timers, label, device)


class _SyncedMeasure:
    """Reference timer label around device work, synchronised on exit so the
    wall time covers the kernels (``qepy_pw/timing.py:16-35``)."""

    def __init__(self, timers, label, device):
        self.inner = timers.measure(label)
        self.device = device

    def __enter__(self):
        return self.inner.__enter__()

    def __exit__(self, *exc):
        if self.device is not None:
            torch.cuda.synchronize(self.device)
        return self.inner.__exit__(*exc)


class _PrecisionResources:
    """Device operands of H at one precision."""

    def __init__(self, potential, kinetic, beta, coupling):
        self.potential = potential
        self.kinetic = kinetic
        self.beta = beta
        self.coupling = coupling
        self.nproj = 0 if beta is None else int(beta.shape[0])


class PlaneWaveHamiltonian:
    """Apply scalar norm-conserving H on the GPU without an Npw-square array.

    Arguments follow ``qepy_pw/diagonalization.py:127-137``.  Keyword-only
    extras: ``device`` (HIP device, default the current one) and
    ``precision``: "fp64" or "fp32" arithmetic of ``apply``/``apply_into``,
    or "mixed" (``apply`` in FP64; ``davidson`` on this operator runs its
    mixed complex64 + FP64-refinement form).
    """

    def __init__(self, basis, potential_g, projector_terms=(), local_workspace=None,
                 timers=None, real_potential=None, native_potential_layout=False,
                 potential_average=None, *, device=None, precision: str = "fp64"):
        self.device = dv.require_device(device)
        self.basis = basis
        self.potential_g = np.asarray(potential_g)
        self.local_workspace = local_workspace
        self.timers = timers
        mpi = getattr(local_workspace, "mpi", None)
        self.mpi = mpi if mpi is not None else _SERIAL
        if getattr(self.mpi, "size", 1) > 1:
            raise ValueError("the device Hamiltonian runs on one GPU; a distributed "
                             "LocalPotentialWorkspace is not supported")
        if precision not in ("fp64", "fp32", "mixed"):
            raise ValueError("precision must be 'fp64', 'fp32' or 'mixed'")
        self.precision = precision
        shape = (tuple(local_workspace.shape) if local_workspace is not None
                 else tuple(self.potential_g.shape))
        if len(shape) != 3:
            raise ValueError("the FFT grid shape must be three-dimensional")
        self.shape = shape
        indices = np.asarray(basis.indices)
        self.npw = int(len(indices))
        self.local_rows = np.arange(self.npw, dtype=np.int32)
        self.local_kinetic = np.asarray(basis.kinetic, dtype=np.float64)
        self.native_potential_layout = bool(native_potential_layout)
        self.potential_average = (float(potential_average) if potential_average is not None
                                  else float(np.real(self.potential_g.flat[0])))
        self.projector_terms = tuple(projector_terms)
        shared = getattr(local_workspace, "plan", None)
        if (isinstance(shared, dv.DeviceFFTPlan) and shared.device == self.device
                and shared.shape == shape and shared.npw == self.npw
                and np.array_equal(getattr(local_workspace, "indices", None), indices)):
            # a device LocalPotentialWorkspace of this basis: share its plan
            self.plan = shared
        else:
            self.plan = dv.DeviceFFTPlan(indices, shape, self.device)
        self.workspace = dv.Workspace(self.device)
        perm = self.plan.perm
        if real_potential is not None:
            rp = np.asarray(real_potential, dtype=np.float64)
            # device layout (n3, n2, n1): plane z, row y, x fastest
            if self.native_potential_layout:  # reference native layout (n3, n1, n2)
                plane_major = np.ascontiguousarray(
                    np.transpose(rp.reshape(shape[2], shape[0], shape[1]), (0, 2, 1)))
            else:
                plane_major = np.ascontiguousarray(np.transpose(rp.reshape(shape), (2, 1, 0)))
            self._potential64 = torch.from_numpy(plane_major).to(self.device)
            self._real_potential_host = rp
        else:
            vg = torch.from_numpy(np.ascontiguousarray(self.potential_g, dtype=np.complex128))
            self._potential64 = self.plan.inverse_real(vg.to(self.device), self.workspace)
            self._real_potential_host = None
        kin = torch.from_numpy(np.ascontiguousarray(self.local_kinetic[perm])).to(self.device)
        beta, coupling = self._pack_projectors(perm)
        self._res = {torch.complex128: _PrecisionResources(self._potential64, kin, beta, coupling)}
        self.applications = 0
        # set to a list to record (nb, dtype, start, stop) HIP events of every
        # FFT-phase launch (used by the benchmark's live roofline)
        self.fft_events = None

    # -- set-up -----------------------------------------------------------
    def _pack_projectors(self, perm):
        """All terms as one device beta (P, npw) [column-major npw x P, device
        order] and a block-diagonal D (P, P), the packed form of
        ``pw/scf.py:633-686`` for every term type."""
        blocks, couplings = [], []
        for term in self.projector_terms:
            if isinstance(term, PackedProjectorTerm) or (
                    hasattr(term, "diagonal") and hasattr(term, "beta")):
                # a device-built packed term (qe_python_amd.projectors) stays
                # on the device
                blocks.append(term.beta if isinstance(term.beta, torch.Tensor)
                              else np.asarray(term.beta))
                couplings.append(np.asarray(term.coupling))
            elif isinstance(term, FactorizedProjectorTerm) or hasattr(term, "phases"):
                if isinstance(term.beta, torch.Tensor):
                    # device-built (qe_python_amd.projectors): the atom blocks
                    # are packed on the device, in device row order
                    blocks.append(self._device_phase_blocks(term))
                    couplings.append(np.kron(np.eye(int(term.phases.shape[1])),
                                             np.asarray(term.coupling)))
                    continue
                beta = np.asarray(term.beta)
                for atom in range(term.phases.shape[1]):
                    blocks.append(beta * np.asarray(term.phases)[:, atom, None])
                    couplings.append(np.asarray(term.coupling))
            else:
                beta, coupling = term
                blocks.append(np.asarray(beta))
                couplings.append(np.asarray(coupling))
        blocks = [b for b in blocks if b.shape[1]]
        couplings = [c for c in couplings if c.shape[0]]
        if not blocks:
            return None, None
        for b in blocks:
            if b.shape[0] != self.npw:
                raise ValueError("projector rows must match the plane-wave basis")
        nproj = sum(b.shape[1] for b in blocks)
        d_host = np.zeros((nproj, nproj), dtype=np.complex128)
        off = 0
        for b, c in zip(blocks, couplings):
            k = b.shape[1]
            d_host[off:off + k, off:off + k] = c
            off += k
        if any(isinstance(b, (torch.Tensor, _DeviceOrderBlock)) for b in blocks):
            beta = torch.empty((nproj, self.npw), dtype=torch.complex128, device=self.device)
            off = 0
            for b in blocks:
                k = b.shape[1]
                if isinstance(b, _DeviceOrderBlock):
                    beta[off:off + k] = b.rows
                else:
                    beta[off:off + k] = _upload_block(b, self.plan.perm_device, self.npw,
                                                      self.device, torch.complex128)
                off += k
        else:
            beta_host = np.empty((nproj, self.npw), dtype=np.complex128)
            off = 0
            for b in blocks:
                k = b.shape[1]
                beta_host[off:off + k] = b[perm].T
                off += k
            beta = torch.from_numpy(beta_host).to(self.device)
        # column-major D: torch row j holds column j of D
        coupling = torch.from_numpy(np.ascontiguousarray(d_host.T)).to(self.device)
        return beta, coupling

    def _device_phase_blocks(self, term) -> "_DeviceOrderBlock":
        """(natom*nch, npw) device-order atom blocks beta * phases[:, a] of a
        device-built factorized term (``qepw_phase_blocks``)."""
        beta = term.beta.to(self.device, torch.complex128)
        phases = term.phases.to(self.device, torch.complex128)
        if beta.shape[0] != self.npw or phases.shape[0] != self.npw:
            raise ValueError("projector rows must match the plane-wave basis")
        nch, natom = int(beta.shape[1]), int(phases.shape[1])
        b = beta.T.resolve_conj().contiguous()
        ph = phases.T.resolve_conj().contiguous()
        out = torch.empty((natom * nch, self.npw), dtype=torch.complex128, device=self.device)
        perm = self.plan.perm_device
        rc = nat.load().qepw_phase_blocks(
            nat.C128, b.data_ptr(), dv.ld_of(b), nch, ph.data_ptr(), dv.ld_of(ph), natom,
            None if perm is None else perm.data_ptr(), out.data_ptr(), dv.ld_of(out), self.npw,
            dv.stream_handle(self.device))
        nat.check(rc, "qepw_phase_blocks")
        return _DeviceOrderBlock(out)

    def _resources(self, dtype: torch.dtype) -> _PrecisionResources:
        res = self._res.get(dtype)
        if res is None:
            base = self._res[torch.complex128]
            res = _PrecisionResources(
                base.potential.to(torch.float32), base.kinetic.to(torch.float32),
                None if base.beta is None else base.beta.to(torch.complex64),
                None if base.coupling is None else base.coupling.to(torch.complex64))
            self._res[dtype] = res
        return res

    # -- reference attributes ----------------------------------------------
    @property
    def real_potential(self) -> np.ndarray:
        if self._real_potential_host is None:
            plane = self._potential64.cpu().numpy()  # (n3, n2, n1)
            self._real_potential_host = np.ascontiguousarray(np.transpose(plane, (2, 1, 0)))
        return self._real_potential_host

    @property
    def diagonal(self) -> np.ndarray:
        """Hamiltonian diagonal used by QE's ``usnldiag`` preconditioner
        (``diagonalization.py:205-229``); non-packed terms are accumulated on
        the device (``accumulate_projector_diagonal``)."""
        result = self.local_kinetic + self.potential_average
        acc = None
        for term in self.projector_terms:
            if isinstance(term, PackedProjectorTerm):
                result = result + term.diagonal
                continue
            if isinstance(term, FactorizedProjectorTerm):
                beta, coupling, mult = term.beta, term.coupling, term.phases.shape[1]
            else:
                (beta, coupling), mult = term, 1
            if beta.shape[1] == 0:
                continue
            if acc is None:
                acc = torch.zeros(self.npw, dtype=torch.float64, device=self.device)
            if isinstance(beta, torch.Tensor):
                b = beta.T.to(self.device, torch.complex128).resolve_conj().contiguous()
            else:
                b = torch.from_numpy(np.ascontiguousarray(np.asarray(beta).T,
                                                          dtype=np.complex128)).to(self.device)
            d = torch.from_numpy(np.ascontiguousarray(np.asarray(coupling, np.complex128).T)).to(self.device)
            coupled = torch.empty_like(b)
            dv.gemm(nat.OP_N, self.npw, b.shape[0], b.shape[0], 1.0, b, d, 0.0, coupled,
                    self.workspace)
            dv.accumulate_projector_diagonal(coupled, b, self.npw, mult, acc)
        if acc is not None:
            result = result + acc.cpu().numpy()
        return result

    # -- device-order interface ---------------------------------------------
    def device_apply(self, src: torch.Tensor, dst: torch.Tensor) -> None:
        """dst = H src for (nb, npw) device-order tensors (complex128 or
        complex64); ``dst`` must not alias ``src``."""
        nb = int(src.shape[0])
        if nb == 0:
            return
        res = self._resources(src.dtype)
        with _measure(self.timers, "vloc_psi", self.device):
            if self.fft_events is None:
                self.plan.apply_local(src, dst, res.potential, res.kinetic, self.workspace)
            else:
                # HIP events on the launch stream, recorded by the library
                # right around the FFT phase's kernels (recorded once here
                # first, so that torch has created them)
                start = torch.cuda.Event(enable_timing=True)
                stop = torch.cuda.Event(enable_timing=True)
                start.record()
                stop.record()
                self.plan.apply_local_timed(src, dst, res.potential, res.kinetic,
                                            self.workspace, start, stop)
                self.fft_events.append((nb, src.dtype, start, stop))
        if res.beta is not None:
            self._nonlocal_apply(src, dst, res)
        self.applications += nb

    def _nonlocal_apply(self, src: torch.Tensor, dst: torch.Tensor, res) -> None:
        """dst += sum_terms beta D beta^H src (calbec, add_vuspsi)."""
        nb = int(src.shape[0])
        p = res.nproj
        with _measure(self.timers, "h_psi:calbec", self.device):
            becp = torch.empty((nb, p), dtype=src.dtype, device=self.device)
            dv.gemm(nat.OP_C, p, nb, self.npw, 1.0, res.beta, src, 0.0, becp, self.workspace)
        with _measure(self.timers, "add_vuspsi", self.device):
            coupled = torch.empty_like(becp)
            dv.gemm(nat.OP_N, p, nb, p, 1.0, res.coupling, becp, 0.0, coupled, self.workspace)
            dv.gemm(nat.OP_N, self.npw, nb, p, 1.0, res.beta, coupled, 1.0, dst, self.workspace)

    @staticmethod
    def device_apply_group(requests) -> None:
        """``device_apply`` of several Hamiltonians at once: ``requests`` a
        sequence of (hamiltonian, src, dst) in one precision.  The FFT phase of
        all of them runs as one grouped launch per stage
        (``DeviceFFTPlan.apply_local_group``), then each one's projector
        terms; results equal the one-by-one calls."""
        requests = [r for r in requests if int(r[1].shape[0]) > 0]
        if not requests:
            return
        items = []
        for ham, src, dst in requests:
            res = ham._resources(src.dtype)
            items.append((ham.plan, src, dst, res.potential, res.kinetic, ham.workspace))
        dv.DeviceFFTPlan.apply_local_group(items)
        for ham, src, dst in requests:
            res = ham._resources(src.dtype)
            if res.beta is not None:
                ham._nonlocal_apply(src, dst, res)
            ham.applications += int(src.shape[0])

    def to_device(self, vectors, dtype=torch.complex128) -> torch.Tensor:
        """(npw, nb) reference-order host array or device tensor -> (nb, npw)
        device-order tensor."""
        return _upload_block(vectors, self.plan.perm_device, self.npw, self.device, dtype)

    def from_device(self, block: torch.Tensor) -> np.ndarray:
        """(nb, npw) device-order tensor -> (npw, nb) F-order host array."""
        return _download_block(block, self.plan.perm_device, self.npw)

    # -- reference interface -------------------------------------------------
    def apply(self, coefficients):
        with _measure(self.timers, "h_psi"):
            return self._apply(coefficients)

    def apply_into(self, coefficients, out):
        """Apply H directly into caller-owned storage (``diagonalization.py:237-244``)."""
        with _measure(self.timers, "h_psi"):
            return self._apply(coefficients, out=out)

    def _apply(self, coefficients, out=None):
        dtype = torch.complex64 if self.precision == "fp32" else torch.complex128
        if isinstance(coefficients, torch.Tensor):
            vec = coefficients
            was_vector = vec.dim() == 1
            if was_vector:
                vec = vec[:, None]
            if vec.dim() != 2 or vec.shape[0] != self.npw:
                raise ValueError("coefficients must have shape (nplane,) or (nplane, nvec)")
            src = self.to_device(vec, dtype)
            dst = torch.empty_like(src)
            self.device_apply(src, dst)
            res = torch.empty((dst.shape[0], self.npw), dtype=torch.complex128, device=self.device)
            dv.permute_rows(1, self.plan.perm_device, dst, res, self.npw)
            res = res.T
            if out is not None:
                target = out[:, None] if (was_vector and out.dim() == 1) else out
                target.copy_(res)
                return out
            return res[:, 0] if was_vector else res
        vectors = np.asarray(coefficients)
        was_vector = vectors.ndim == 1
        if was_vector:
            vectors = vectors[:, None]
        if vectors.ndim != 2 or vectors.shape[0] != self.npw:
            raise ValueError("coefficients must have rank-local shape "
                             "(nplane_local,) or (nplane_local, nvec)")
        src = self.to_device(vectors, dtype)
        dst = torch.empty_like(src)
        self.device_apply(src, dst)
        result = self.from_device(dst)
        if out is not None:
            target = np.asarray(out)
            if was_vector and target.ndim == 1:
                target = target[:, None]
            if target.shape != result.shape or target.dtype != np.complex128:
                raise ValueError("local-potential output has the wrong shape or dtype")
            target[...] = result
            return out
        return result[:, 0] if was_vector else result


class _DeviceOrderBlock:
    """Projector columns already on the device in device row order, as a
    (P, npw) tensor; ``shape`` reads like the (npw, P) host blocks'."""

    def __init__(self, rows: torch.Tensor):
        self.rows = rows
        self.shape = (int(rows.shape[1]), int(rows.shape[0]))


def _upload_block(vectors, perm, npw, device, dtype):
    if isinstance(vectors, torch.Tensor):
        host = None
        t = vectors
        if t.dim() != 2 or t.shape[0] != npw:
            raise ValueError("vectors must have shape (npw, nb)")
        src = t.T.to(device)
        if src.dtype not in (torch.complex128, torch.complex64):
            src = src.to(torch.complex128)
        src = src.resolve_conj().resolve_neg().contiguous()
    else:
        host = np.asarray(vectors)
        if host.ndim != 2 or host.shape[0] != npw:
            raise ValueError("vectors must have shape (npw, nb)")
        src = torch.from_numpy(np.ascontiguousarray(host.T, dtype=np.complex128)).to(device)
    out = torch.empty((src.shape[0], npw), dtype=dtype, device=device)
    dv.permute_rows(0, perm, src, out, npw)
    return out


def _download_block(block, perm, npw):
    """Reference-order complex128 host copy (npw, nb) of a device-order
    block, copied into pinned memory (a pageable device-to-host copy runs at
    a fraction of the link's rate); the array keeps its pinned buffer alive,
    and torch's host allocator reuses it once the array is released."""
    ref = torch.empty((block.shape[0], npw), dtype=torch.complex128, device=block.device)
    dv.permute_rows(1, perm, block, ref, npw)
    host = torch.empty(ref.shape, dtype=ref.dtype, pin_memory=True)
    host.copy_(ref, non_blocking=True)
    torch.cuda.current_stream(block.device).synchronize()
    return host.numpy().T


def _device_hamiltonian(operator, operator_into):
    for fn in (operator_into, operator):
        owner = getattr(fn, "__self__", None)
        if isinstance(owner, PlaneWaveHamiltonian):
            return owner
    return None


class _HostOperator:
    """A non-device operator driven from device storage (identity row order)."""

    def __init__(self, operator, operator_into, npw, device):
        self.operator = operator
        self.operator_into = operator_into
        self.npw = npw
        self.device = device

    def device_apply(self, src, dst):
        block = np.ascontiguousarray(src.cpu().numpy().T.astype(np.complex128), dtype=np.complex128)
        block = np.asfortranarray(block)
        out = np.empty_like(block, order="F")
        if self.operator_into is not None:
            self.operator_into(block, out)
        else:
            out[...] = self.operator(block)
        dst.copy_(torch.from_numpy(np.ascontiguousarray(out.T)).to(dst.dtype))


def _check_room(active: int, count: int, maximum_subspace: int) -> None:
    """The reference stores the correction block in its fixed
    ``basis_storage`` (``qepy_pw/diagonalization.py:789-791``), which raises
    ``ValueError`` when an initial basis of more than ``maximum_subspace -
    unconverged`` vectors leaves no room for it; after a restart the
    ``active + unconverged`` test always leaves room."""
    if active + count > maximum_subspace:
        raise ValueError(f"Davidson subspace of {maximum_subspace} vectors has no room for "
                         f"{count} corrections after {active} basis vectors; pass fewer "
                         f"initial vectors or a larger subspace_multiplier")


def _host_dtype(dtype: torch.dtype):
    return np.complex64 if dtype == torch.complex64 else np.complex128


def _cholesky_orthonormalize(block: torch.Tensor, npw: int, ws, condition: float = 1.0e6):
    """(steps) Cholesky QR of a block (the complex64 phase's corrections, or a
    Davidson's initial block in either precision): Gram = L L^H in FP64 on the
    host (one thread), block L^-H on the device.  None when the Gram matrix
    is not safely positive definite at the block's precision (its Cholesky
    fails or the squared condition estimate max(diag L)^2 / min(diag L)^2
    exceeds ``condition``); the caller then drops dependent directions with
    the canonical orthonormalisation."""
    n = int(block.shape[0])
    device = block.device
    if NATIVE_ORTHO and may_block():
        # the same steps in one native call (its host round trip inside)
        lib = nat.load()
        out = torch.empty_like(block)
        buf = ws.get(int(lib.qepw_rr_workspace(max(n, 1))))
        ok = ctypes.c_int(0)
        nat.check(lib.qepw_cholesky_qr(dv.precision_code(block.dtype), block.data_ptr(),
                                       dv.ld_of(block), npw, n, condition, out.data_ptr(),
                                       dv.ld_of(out), ctypes.byref(ok), buf.data_ptr(),
                                       buf.numel(), dv.stream_handle(device)),
                  "qepw_cholesky_qr")
        return out if ok.value else None
    gram_t = torch.empty((n, n), dtype=block.dtype, device=device)
    dv.gemm(nat.OP_C, n, n, npw, 1.0, block, block, 0.0, gram_t, ws)
    gram = (yield Read(gram_t)).T.astype(np.complex128)
    gram = 0.5 * (gram + gram.conj().T)
    with _one_blas_thread():
        try:
            chol = np.linalg.cholesky(gram)
        except np.linalg.LinAlgError:
            return None
        d = np.abs(np.diag(chol))
        if d.min() <= 0.0 or (d.max() / d.min()) ** 2 > condition:
            return None
        inv = _lower_inverse(chol).conj().T   # L^-H
    t_dev = dv.upload(inv.T.astype(_host_dtype(block.dtype)), device)
    out = torch.empty_like(block)
    dv.gemm(nat.OP_N, npw, n, n, 1.0, block, t_dev, 0.0, out, ws)
    return out


def _orthonormalize_device(block: torch.Tensor, npw: int, ws, threshold: float = 1.0e-12):
    """Canonical orthonormalisation (``diagonalization.py:450-472``, serial):
    Gram V^H V on the device, its small Hermitian eigensolve on the host,
    V U s^{-1/2} (columns by decreasing Gram eigenvalue) on the device."""
    return blocking(_orthonormalize_steps(block, npw, ws, threshold))


def _orthonormalize_steps(block: torch.Tensor, npw: int, ws, threshold: float = 1.0e-12):
    """(steps) ``_orthonormalize_device`` as a generator of reads."""
    n = int(block.shape[0])
    device = block.device
    if n == 0:
        return block[:0].clone()
    if n == 1:
        sq = dv.column_squared_norms(block, npw)
        if float((yield Read(sq))[0]) == 0.0:
            return block[:0].clone()
        out = torch.empty_like(block)
        dv.normalize_selected_columns(block, sq, dv.index_tensor([0], device),
                                      out, npw)
        return out
    gram_t = torch.empty((n, n), dtype=block.dtype, device=device)
    dv.gemm(nat.OP_C, n, n, npw, 1.0, block, block, 0.0, gram_t, ws)
    gram = (yield Read(gram_t)).T.astype(np.complex128)
    gram = 0.5 * (gram + gram.conj().T)
    with _one_blas_thread():
        squared, rotation = np.linalg.eigh(gram)
    order = np.argsort(squared)[::-1]
    squared = np.maximum(np.real(squared[order]), 0.0)
    if not len(squared) or squared[0] == 0.0:
        return block[:0].clone()
    rank = int(np.count_nonzero(squared > threshold ** 2 * squared[0]))
    transform = rotation[:, order[:rank]] / np.sqrt(squared[:rank])[None, :]
    t_dev = dv.upload(transform.T.astype(_host_dtype(block.dtype)), device)
    out = torch.empty((rank, npw), dtype=block.dtype, device=device)
    dv.gemm(nat.OP_N, npw, rank, n, 1.0, block, t_dev, 0.0, out, ws)
    return out


def davidson(operator, diagonal, number_of_roots, *, initial_vectors=None, tolerance=1.0e-8,
             max_iterations=40, subspace_multiplier=4, residual_factor=10.0,
             residual_energy_scale=None, occupied_roots=None, initial_is_ritz=False,
             initial_is_orthonormal=False, initial_eigenvalues=None, initial_applied=None,
             mpi=None, global_dimension=None, global_row_indices=None, timers=None,
             operator_into=None, device_output=False, reduced_eigensolver=None,
             precision=None, mixed_switch=None):
    """Lowest eigenpairs by restarted block Davidson (QE ``cegterg``) on the
    GPU; semantics of ``qepy_pw/diagonalization.py:527-880``.

    Extra keywords:

    * ``device_output=True`` returns the eigenvectors as a device tensor
      (npw, nroots) in reference row order instead of a host array (no
      device-to-host copy of the vectors).
    * ``reduced_eigensolver`` selects the reduced eigensolve: "device"
      (spectral division on the GPU, the default,
      ``qe_python_amd.eigensolver``) or "host" (LAPACK zhegvd on the copied
      pencil, as the reference does).
    * ``precision``: "fp64" (complex128 storage; FP64 h_psi and GEMMs) or
      "mixed".  The mixed form runs a complex64 phase -- FP32 FFT, V*psi,
      projections and subspace GEMMs on FP32 MFMA with FP32 accumulation,
      while the Cholesky, the reduced eigensolve and the residual norms stay
      FP64 -- until every root has |de| < ``mixed_switch[0]`` Ha and a
      residual norm < ``mixed_switch[1]`` (or the phase stalls at the FP32
      noise floor), then an FP64 refinement: FP64 h_psi of the complex64
      Ritz vectors, Rayleigh-Ritz and FP64 Davidson iterations to
      ``tolerance``.  ``None`` selects "mixed" when the operator is a
      ``PlaneWaveHamiltonian`` built with ``precision="mixed"``, else "fp64".
      The result's ``iterations`` and ``hamiltonian_applications`` count
      both phases, which share ``max_iterations`` (the complex64 phase
      takes at most ``max_iterations - 1``, the refinement the rest; with
      ``max_iterations < 2`` the FP64 form runs).
    """
    steps = davidson_steps(
        operator, diagonal, number_of_roots, initial_vectors=initial_vectors,
        tolerance=tolerance, max_iterations=max_iterations,
        subspace_multiplier=subspace_multiplier, residual_factor=residual_factor,
        residual_energy_scale=residual_energy_scale, occupied_roots=occupied_roots,
        initial_is_ritz=initial_is_ritz, initial_is_orthonormal=initial_is_orthonormal,
        initial_eigenvalues=initial_eigenvalues, initial_applied=initial_applied, mpi=mpi,
        global_dimension=global_dimension, global_row_indices=global_row_indices,
        timers=timers, operator_into=operator_into, device_output=device_output,
        reduced_eigensolver=reduced_eigensolver, precision=precision,
        mixed_switch=mixed_switch)
    return schedule.run(steps, _solve_device(operator, operator_into))


def _solve_device(operator, operator_into) -> torch.device:
    ham = _device_hamiltonian(operator, operator_into)
    return ham.device if ham is not None else dv.require_device()


def davidson_steps(operator, diagonal, number_of_roots, *, initial_vectors=None, tolerance=1.0e-8,
             max_iterations=40, subspace_multiplier=4, residual_factor=10.0,
             residual_energy_scale=None, occupied_roots=None, initial_is_ritz=False,
             initial_is_orthonormal=False, initial_eigenvalues=None, initial_applied=None,
             mpi=None, global_dimension=None, global_row_indices=None, timers=None,
             operator_into=None, device_output=False, reduced_eigensolver=None,
             precision=None, mixed_switch=None):
    """(steps) ``davidson`` as a generator of ``qe_python_amd.schedule``
    requests; its return value is the ``DavidsonResult``.  Drive it with
    ``schedule.run`` (one solve) or ``schedule.run_many`` (several solves
    in rounds, H applied to all of them together)."""
    if reduced_eigensolver is None:
        reduced_eigensolver = REDUCED_EIGENSOLVER
    if reduced_eigensolver not in ("device", "host"):
        raise ValueError("reduced_eigensolver must be 'device' or 'host'")
    if precision not in (None, "fp64", "mixed"):
        raise ValueError("precision must be 'fp64' or 'mixed'")
    switch = MIXED_SWITCH if mixed_switch is None else tuple(float(x) for x in mixed_switch)
    if len(switch) != 2 or min(switch) <= 0.0:
        raise ValueError("mixed_switch must be two positive numbers (|de| in Ha, residual norm)")
    diagonal = np.asarray(diagonal, dtype=float)
    mpi = mpi if mpi is not None else _SERIAL
    if getattr(mpi, "size", 1) > 1:
        raise ValueError("the device Davidson runs on one GPU; MPI size must be 1")
    local_dimension = len(diagonal)
    dimension = int(global_dimension) if global_dimension is not None else local_dimension
    if not 1 <= number_of_roots <= dimension:
        raise ValueError("number_of_roots must be between 1 and the matrix size")
    if tolerance <= 0.0:
        raise ValueError("Davidson tolerance must be positive")
    if max_iterations < 1:
        raise ValueError("Davidson max_iterations must be at least 1")
    if subspace_multiplier < 2:
        raise ValueError("diago_david_ndim must be at least 2")
    if residual_factor is not None and residual_factor <= 0.0:
        raise ValueError("Davidson residual_factor must be positive or None")
    if residual_energy_scale is not None and residual_energy_scale <= 0.0:
        raise ValueError("Davidson residual_energy_scale must be positive or None")
    if occupied_roots is None:
        occupied_roots = number_of_roots
    if not 0 <= occupied_roots <= number_of_roots:
        raise ValueError("occupied_roots must be between zero and number_of_roots")
    root_tolerances = np.full(number_of_roots, max(5.0 * tolerance, 5.0e-6))
    root_tolerances[:occupied_roots] = tolerance

    if initial_vectors is None:
        rows = (np.arange(local_dimension) if global_row_indices is None
                else np.asarray(global_row_indices, dtype=int))
        trial = np.zeros((local_dimension, number_of_roots), dtype=complex)
        for column in range(number_of_roots):
            matches = np.flatnonzero(rows == column)
            if len(matches):
                trial[matches[0], column] = 1.0
    else:
        trial = initial_vectors
        shape = tuple(trial.shape)
        if len(shape) != 2 or shape[0] != local_dimension or shape[1] < number_of_roots:
            raise ValueError("initial_vectors must have shape "
                             "(matrix size, at least number_of_roots)")

    ham = _device_hamiltonian(operator, operator_into)
    npw = local_dimension
    if ham is not None:
        if ham.npw != npw:
            raise ValueError("diagonal length does not match the Hamiltonian basis")
        device, perm_dev, perm = ham.device, ham.plan.perm_device, ham.plan.perm
        op = ham
        ws = ham.workspace
    else:
        device = dv.require_device()
        perm_dev, perm = None, None
        op = _HostOperator(operator, operator_into, npw, device)
        ws = dv.Workspace(device)
    if precision is None:
        precision = "mixed" if (ham is not None and ham.precision == "mixed") else "fp64"
    diag_dev = torch.from_numpy(np.ascontiguousarray(
        diagonal if perm is None else diagonal[perm])).to(device)
    ctx = _PhaseContext(op, ws, device, npw, diag_dev, timers, number_of_roots, dimension,
                        subspace_multiplier, reduced_eigensolver == "device")
    energy_bounds = root_tolerances
    residual_bounds = None if residual_factor is None else residual_factor * root_tolerances

    def applied_block(active):
        if initial_applied is None:
            return None
        if tuple(initial_applied.shape) != (npw, active):
            raise ValueError("initial_applied must match the initial Ritz block")
        return initial_applied

    if precision == "mixed" and max_iterations < 2:
        # no room for a complex64 iteration and the FP64 refinement under the cap
        precision = "fp64"
    if precision == "fp64":
        dtype = torch.complex128
        start = _upload_block(trial, perm_dev, npw, device, dtype)
        phase = yield from _davidson_phase_body(
            ctx, start, dtype, energy_bounds, residual_bounds, residual_energy_scale,
            root_tolerances, max_iterations, initial_is_ritz, initial_is_orthonormal,
            initial_eigenvalues, applied_block, perm_dev)
        return DavidsonResult(phase.values, _result_vectors(phase.X, perm_dev, npw, device_output),
                              phase.residual_norms, phase.iterations, phase.converged,
                              phase.number_unconverged, phase.happ)

    # mixed: complex64 phase to the switch thresholds ...  Both phases share
    # max_iterations (the SCF sums ``iterations`` over its retries,
    # pw/scf.py:3539-3556): the complex64 phase may use all but one, the
    # FP64 refinement what is left
    de_switch, res_switch = switch
    start = _upload_block(trial, perm_dev, npw, device, torch.complex64)
    coarse_ctx = ctx
    if MIXED_COARSE_MULTIPLIER is not None:
        coarse_ctx = dataclasses.replace(
            ctx, subspace_multiplier=min(ctx.subspace_multiplier, MIXED_COARSE_MULTIPLIER))
    coarse = yield from _davidson_phase_body(
        coarse_ctx, start, torch.complex64, np.maximum(energy_bounds, de_switch),
        np.full(number_of_roots, res_switch), None, np.maximum(root_tolerances, de_switch),
        max_iterations - 1, initial_is_ritz, initial_is_orthonormal, initial_eigenvalues,
        applied_block, perm_dev, stall_window=MIXED_STALL_WINDOW, require_residuals=True,
        coarse=True)
    del start
    # ... then the FP64 refinement from its Ritz vectors; the reduced pencil
    # carries S, so the FP32-orthonormal block needs no re-orthonormalisation
    start = torch.empty(coarse.X.shape, dtype=torch.complex128, device=device)
    dv.copy_columns(coarse.X, start, npw)
    coarse.X = None
    fine = yield from _davidson_phase_body(
        ctx, start, torch.complex128, energy_bounds, residual_bounds, residual_energy_scale,
        root_tolerances, max(1, max_iterations - coarse.iterations), False, True, None,
        lambda active: None, perm_dev, seed_spectrum=coarse.spectrum)
    return DavidsonResult(fine.values, _result_vectors(fine.X, perm_dev, npw, device_output),
                          fine.residual_norms, coarse.iterations + fine.iterations,
                          fine.converged, fine.number_unconverged, coarse.happ + fine.happ)


@dataclass
class _PhaseContext:
    op: object
    ws: dv.Workspace
    device: torch.device
    npw: int
    diag_dev: torch.Tensor
    timers: object
    nroots: int
    dimension: int
    subspace_multiplier: int
    spectral: bool


@dataclass
class _PhaseResult:
    values: np.ndarray
    X: torch.Tensor | None
    residual_norms: np.ndarray
    iterations: int
    converged: bool
    number_unconverged: int
    happ: int
    spectrum: np.ndarray | None = None   # eigenvalue estimates, also above the roots


def _davidson_phase_body(ctx, start, dtype, energy_bounds, residual_bounds, residual_energy_scale,
                         root_tolerances, max_iterations, start_is_ritz, start_is_orthonormal,
                         initial_eigenvalues, applied_block, perm_dev, *, stall_window=None,
                         require_residuals=False, coarse=False, seed_spectrum=None):
    """The restarted Davidson loop (``diagonalization.py:640-880``) at one
    storage dtype, from a device-order start block ``start`` (nb, npw).

    ``coarse``: the complex64 phase of the mixed form.  When its FP32
    overlap loses definiteness (new corrections dependent at FP32
    precision) the phase ends on the previous Ritz pairs instead of raising
    ``LinAlgError``; the FP64 refinement continues from them.  At a restart
    it locks the converged Ritz pairs at the bottom of the spectrum (a run of
    at least ``MIXED_LOCK_MIN`` ending at a gap of ``MIXED_LOCK_GAP``): they
    leave the search space, later corrections are orthogonalised against
    them, and the reduced problem shrinks to the remaining roots.

    A root converges when |de| < ``energy_bounds`` and, when given, its
    residual norm < ``residual_bounds`` and its squared residual norm <
    ``residual_energy_scale * root_tolerances``.  ``stall_window``: stop
    unconverged when the largest unconverged residual norm, already below
    ``MIXED_STALL_SPAN`` times the largest residual bound, has not halved
    over that many iterations."""
    op, ws, device, npw, timers = ctx.op, ctx.ws, ctx.device, ctx.npw, ctx.timers
    number_of_roots = ctx.nroots
    # roots still in the search space (all of them unless the complex64
    # phase has locked some) and their per-root bounds
    gidx = np.arange(number_of_roots)
    eb, rb, rt = energy_bounds, residual_bounds, root_tolerances
    locked = None                 # locked Ritz vectors (device rows)
    lock_vals, lock_res = [], []

    def spectrum():
        """Eigenvalue estimates of the phase, locked ones included."""
        top = rr.theta if rr.theta is not None else values
        return np.sort(np.concatenate(lock_vals + [np.asarray(top, dtype=np.float64)]))

    def merged(values, X, residual_norms):
        """Locked and active Ritz pairs together, by ascending value."""
        if locked is None:
            return values, X, residual_norms
        allv = np.concatenate(lock_vals + [values])
        allr = np.concatenate(lock_res + [residual_norms])
        order = np.argsort(allv, kind="stable")
        # source column j (locked first, then X) lands at position pos[j]
        pos = np.empty(len(order), dtype=np.int32)
        pos[order] = np.arange(len(order), dtype=np.int32)
        nl = int(locked.shape[0])
        Xall = torch.empty((len(order), npw), dtype=dtype, device=device)
        dv.copy_columns(locked, Xall, npw, nl, dst_cols=dv.index_tensor(pos[:nl], device))
        dv.copy_columns(X, Xall, npw, len(order) - nl,
                        dst_cols=dv.index_tensor(pos[nl:], device))
        return allv[order], Xall, allr[order]
    if start_is_ritz or start_is_orthonormal:
        basis = start[:number_of_roots]
    else:
        with _measure(timers, "cegterg:over", device):
            # Cholesky QR (one small factorisation) spans the same space as the
            # canonical orthonormalisation when the block is well conditioned,
            # and the Rayleigh-Ritz depends only on that span; the canonical
            # form (a host eigensolve, dropping dependent directions) takes
            # the rest
            basis = None
            if start.shape[0] > 1:
                basis = yield from _cholesky_orthonormalize(
                    start, npw, ws, condition=1.0e6 if dtype == torch.complex64 else 1.0e8)
            if basis is None:
                basis = yield from _orthonormalize_steps(start, npw, ws)
    if basis.shape[0] < number_of_roots:
        raise ValueError("initial Davidson vectors are linearly dependent")
    maximum_subspace = min(ctx.dimension, max(basis.shape[0], 2 * number_of_roots,
                                              ctx.subspace_multiplier * number_of_roots))
    active = int(basis.shape[0])
    V = torch.empty((maximum_subspace, npw), dtype=dtype, device=device)
    HV = torch.empty_like(V)
    V[:active].copy_(basis)
    del basis, start
    given = applied_block(active)
    if given is None:
        yield Apply(op, V[:active], HV[:active])
        happ = active
    else:
        HV[:active].copy_(_upload_block(given, perm_dev, npw, device, dtype))
        happ = 0
    # the reduced pencil Hs = V^H HV, Ss = V^H V lives on the device
    single = MIXED_SINGLE_EIGH and dtype == torch.complex64
    rr = DeviceRayleighRitz(maximum_subspace, number_of_roots, device, ws,
                            spectral=ctx.spectral, host_single=single)
    diag_dev = ctx.diag_dev

    def update_projected(first):
        with _measure(timers, "cegterg:over", device):
            rr.update(V, HV, first, active, npw)

    update_projected(0)

    def ritz_solve():
        with _measure(timers, "cdiaghg", device):
            roots, coef = yield from rr.solve_steps(active)
            if coef.dtype != dtype:
                narrow = torch.empty(coef.shape, dtype=dtype, device=device)
                dv.copy_columns(coef, narrow, int(coef.shape[1]))
                coef = narrow
        return roots, coef

    def deflate(R):
        """With locked pairs the search runs on the operator projected onto
        their complement (the basis is orthogonal to them): its residuals are
        R - L (L^H R)."""
        if locked is not None:
            nl = int(locked.shape[0])
            proj = torch.empty((R.shape[0], nl), dtype=dtype, device=device)
            dv.gemm(nat.OP_C, nl, R.shape[0], npw, 1.0, locked, R, 0.0, proj, ws)
            dv.gemm(nat.OP_N, npw, R.shape[0], nl, -1.0, locked, proj, 1.0, R, ws)
        return R

    def ritz_block(roots, coef):
        """Ritz vectors X = V C and residuals R = HV C - X diag(e)."""
        with _measure(timers, "cegterg:upda", device):
            X = torch.empty((number_of_roots, npw), dtype=dtype, device=device)
            R = torch.empty_like(X)
            dv.gemm(nat.OP_N, npw, number_of_roots, active, 1.0, V[:active], coef, 0.0, X, ws)
            dv.gemm(nat.OP_N, npw, number_of_roots, active, 1.0, HV[:active], coef, 0.0, R, ws)
            e = dv.upload(np.asarray(roots, dtype=np.float64), device)
            dv.subtract_band_energies(R, X, e, npw)
            deflate(R)
        return X, R

    def selected_residuals(roots, coef, cols):
        """Residuals of the roots ``cols`` only, packed:
        R_u = HV C_u - V (C_u diag(e_u)) -- all the next expansion needs."""
        with _measure(timers, "cegterg:upda", device):
            idx = dv.index_tensor(cols, device)
            e = dv.upload(np.asarray(roots[cols], dtype=np.float64), device)
            # the unconverged roots' coefficients C_u and C_u diag(e_u)
            cu = torch.empty((len(cols), active), dtype=dtype, device=device)
            ce = torch.empty_like(cu)
            dv.copy_columns(coef, cu, active, src_cols=idx)
            dv.copy_columns(coef, ce, active, src_cols=idx, scale=e)
            R = torch.empty((len(cols), npw), dtype=dtype, device=device)
            dv.gemm(nat.OP_N, npw, len(cols), active, 1.0, HV[:active], cu, 0.0, R, ws)
            dv.gemm(nat.OP_N, npw, len(cols), active, -1.0, V[:active], ce, 1.0, R, ws)
            deflate(R)
        return R

    track = (residual_bounds is not None or residual_energy_scale is not None
             or require_residuals)
    coef = None
    if start_is_ritz:
        if initial_eigenvalues is None:
            ip = dv.column_inner_products(V[:number_of_roots], HV[:number_of_roots], npw)
            values = (yield Read(ip))[0::2].copy()
        else:
            values = np.asarray(initial_eigenvalues[:number_of_roots], dtype=float)
        X = V[:number_of_roots].clone()
        R = HV[:number_of_roots].clone()
        dv.subtract_band_energies(R, X, torch.from_numpy(np.ascontiguousarray(values)).to(device),
                                  npw)
        # the start block is its own Ritz block: coefficients = identity
        coef = torch.empty((number_of_roots, number_of_roots), dtype=dtype, device=device)
        dv.set_block(coef, number_of_roots, number_of_roots, 1.0)
    else:
        values, coef = yield from ritz_solve()
        X, R = ritz_block(values, coef)
        if seed_spectrum is not None and rr.theta is not None:
            # estimates from an earlier stage above the roots: the next solve
            # can place its shift instead of diagonalising densely
            known = len(rr.theta)
            rr.seed(np.concatenate([rr.theta, np.asarray(seed_spectrum)[known:]]))

    def residual_norms_of(block):
        sq = yield Read(dv.column_squared_norms(block, npw))
        return np.sqrt(np.maximum(sq, 0.0))

    residual_norms = ((yield from residual_norms_of(R)) if track
                      else np.empty(number_of_roots))
    completed = 0
    number_unconverged = number_of_roots
    unconverged = np.ones(number_of_roots, dtype=bool)
    # residual rows of the unconverged roots for the next expansion: R holds
    # all roots (rows ``cols``) or, without residual tracking, only them
    # (rows 0..len(cols)-1)
    packed = False
    progress = []
    for iteration in range(1, max_iterations + 1):
        completed = iteration
        previous = values
        with _measure(timers, "g_psi", device):
            cols = np.flatnonzero(unconverged).astype(np.int32)
            rows = np.arange(len(cols), dtype=np.int32) if packed else cols
            rows_dev = dv.upload(rows, device)
            e_dev = dv.upload(values[cols], device)
            corrections = torch.empty((len(cols), npw), dtype=dtype, device=device)
            dv.qe_precondition(R, rows_dev, e_dev, diag_dev, corrections, npw)
        if coarse and len(cols):
            # complex64 phase: the block is made orthogonal to the basis
            # (classical Gram-Schmidt, twice) and orthonormal within itself
            # (Cholesky QR, or canonical orthonormalisation dropping
            # directions below 1e-4 of the largest) -- the same span, so the
            # same Rayleigh-Ritz, with an FP32 overlap near I that stays
            # positive definite
            with _measure(timers, "cegterg:orth", device):
                proj = torch.empty((len(cols), active), dtype=dtype, device=device)
                if locked is not None:
                    lproj = torch.empty((len(cols), locked.shape[0]), dtype=dtype, device=device)
                for _ in range(ORTHO_PASSES):
                    dv.gemm(nat.OP_C, active, len(cols), npw, 1.0, V[:active], corrections,
                            0.0, proj, ws)
                    dv.gemm(nat.OP_N, npw, len(cols), active, -1.0, V[:active], proj, 1.0,
                            corrections, ws)
                    if locked is not None:
                        nl = int(locked.shape[0])
                        dv.gemm(nat.OP_C, nl, len(cols), npw, 1.0, locked, corrections, 0.0,
                                lproj, ws)
                        dv.gemm(nat.OP_N, npw, len(cols), nl, -1.0, locked, lproj, 1.0,
                                corrections, ws)
                block = yield from _cholesky_orthonormalize(corrections, npw, ws)
                if block is None:
                    block = yield from _orthonormalize_steps(corrections, npw, ws,
                                                             threshold=1.0e-4)
            if not block.shape[0]:
                break
            _check_room(active, int(block.shape[0]), maximum_subspace)
            old = active
            active += int(block.shape[0])
            V[old:active].copy_(block)
            keep = np.arange(block.shape[0])
            del block
        else:
            # _normalize_columns (diagonalization.py:500-524): no orthogonalisation
            sq_dev = dv.column_squared_norms(corrections, npw)
            sq = yield Read(sq_dev)
            if not len(sq):
                break
            scale = max(float(np.max(sq)), 1.0e-300)
            keep = np.flatnonzero(sq > (1.0e-14) ** 2 * scale).astype(np.int32)
            if not len(keep):
                break
            _check_room(active, len(keep), maximum_subspace)
            old = active
            active += len(keep)
            dv.normalize_selected_columns(corrections, sq_dev, dv.upload(keep, device),
                                          V[old:active], npw)
        del X, R, corrections
        X = None
        yield Apply(op, V[old:active], HV[old:active])
        happ += len(keep)
        update_projected(old)
        coef_prev = coef
        try:
            values, coef = yield from ritz_solve()
        except np.linalg.LinAlgError:
            if not coarse:
                raise
            active, values, coef = old, previous, coef_prev
            packed = False
            break
        if track:
            X, R = ritz_block(values, coef)
            packed = False
            residual_norms = yield from residual_norms_of(R)
        converged = np.abs(values - previous) < eb
        if rb is not None:
            converged &= residual_norms < rb
        if residual_energy_scale is not None:
            converged &= residual_norms ** 2 < residual_energy_scale * rt
        unconverged = ~converged
        number_unconverged = int(np.count_nonzero(unconverged))
        if ITERATION_TRACE is not None:
            ITERATION_TRACE.append((str(dtype), iteration, number_unconverged,
                                    float(np.max(np.abs(values - previous))),
                                    float(np.max(residual_norms)) if track else None))
        if number_unconverged == 0:
            if not track:
                X, R = ritz_block(values, coef)
                residual_norms = yield from residual_norms_of(R)
            est = spectrum()
            values, X, residual_norms = merged(values, X, residual_norms)
            return _PhaseResult(values, X, residual_norms, iteration, True, 0, happ, est)
        if stall_window:
            worst = float(np.max(residual_norms[unconverged]))
            progress.append(worst)
            near_floor = (residual_bounds is not None
                          and worst < MIXED_STALL_SPAN * float(np.max(residual_bounds)))
            if (near_floor and len(progress) > stall_window
                    and min(progress[-stall_window:]) > 0.5 * min(progress[:-stall_window])):
                break
        if active + number_unconverged > maximum_subspace:
            if ITERATION_TRACE is not None:
                ITERATION_TRACE.append(("restart", iteration, int(np.count_nonzero(converged)),
                                        active, maximum_subspace))
            with _measure(timers, "cegterg:last", device):
                if X is None:
                    X, R = ritz_block(values, coef)
                    packed = False
                # lock the longest converged run from the bottom of the active
                # roots that ends at a gap of at least MIXED_LOCK_GAP: the
                # locked set is then a spectral subspace and no near-degenerate
                # cluster is split between locked and active roots
                run = int(np.argmin(converged)) if not converged.all() else len(converged)
                while run > 0 and values[run] - values[run - 1] < MIXED_LOCK_GAP:
                    run -= 1
                done = np.arange(run)
                old_theta = None
                if coarse and len(done) >= MIXED_LOCK_MIN:
                    # lock the run; the search continues on the roots above it
                    # the run and the rest are contiguous column ranges
                    rest = np.arange(run, number_of_roots)
                    nl = 0 if locked is None else int(locked.shape[0])
                    grown = torch.empty((nl + run, npw), dtype=dtype, device=device)
                    if nl:
                        grown[:nl].copy_(locked)
                    grown[nl:].copy_(X[:run])
                    locked = grown
                    lock_vals.append(values[done])
                    lock_res.append(residual_norms[done] if track
                                    else np.zeros(len(done)))
                    X, R = X[run:], R[run:]
                    values, residual_norms = values[rest], residual_norms[rest]
                    gidx = gidx[rest]
                    eb = energy_bounds[gidx]
                    rb = None if residual_bounds is None else residual_bounds[gidx]
                    rt = root_tolerances[gidx]
                    number_of_roots = len(rest)
                    unconverged = unconverged[rest]
                    number_unconverged = int(np.count_nonzero(unconverged))
                    # the search space keeps its size: a space shrunk with the
                    # root count restarts too often for the slow top roots
                    old_theta = rr.theta
                    rr = DeviceRayleighRitz(maximum_subspace, number_of_roots, device, ws,
                                            spectral=ctx.spectral, host_single=single)
                    if ITERATION_TRACE is not None:
                        ITERATION_TRACE.append(("lock", iteration, int(locked.shape[0]),
                                                number_of_roots, maximum_subspace))
                active = number_of_roots
                V[:active].copy_(X)
                e = dv.upload(values, device)
                dv.restart_applied(X, R, e, HV[:active], npw)
                rr.restart(values)
                if len(lock_vals) and rr.theta is None and old_theta is not None:
                    # the fresh pencil after locking: the previous estimates
                    # above the locked run
                    rr.seed(np.asarray(old_theta)[len(done):])
                # the basis is the Ritz block itself
                coef = torch.empty((number_of_roots, number_of_roots), dtype=dtype,
                                   device=device)
                dv.set_block(coef, number_of_roots, number_of_roots, 1.0)
        elif X is None:
            R = selected_residuals(values, coef, np.flatnonzero(unconverged))
            packed = True
    if X is None or packed:
        X, R = ritz_block(values, coef)
    if not track:
        residual_norms = yield from residual_norms_of(R)
    est = spectrum()
    values, X, residual_norms = merged(values, X, residual_norms)
    return _PhaseResult(values, X, residual_norms, completed, False, number_unconverged, happ,
                        est)


def _result_vectors(X, perm_dev, npw, device_output=False):
    if device_output:
        ref = torch.empty((X.shape[0], npw), dtype=torch.complex128, device=X.device)
        dv.permute_rows(1, perm_dev, X, ref, npw)
        return ref.T
    return np.asfortranarray(_download_block(X, perm_dev, npw))


def conjugate_gradient(*args, **kwargs):
    """Band-by-band CG (reference ``diagonalization.py:993-1186``) on the
    device ``h_psi``: ``qe_python_amd.iterative.conjugate_gradient``."""
    from .iterative import conjugate_gradient as solver
    return solver(*args, **kwargs)


def parallel_orbital(*args, **kwargs):
    """ParO (reference ``diagonalization.py:1328-1467``) on the device
    ``h_psi``: ``qe_python_amd.iterative.parallel_orbital``."""
    from .iterative import parallel_orbital as solver
    return solver(*args, **kwargs)


def rmm_diis(*args, **kwargs):
    """RMM-DIIS (reference ``diagonalization.py:1528-1756``) on the device
    ``h_psi``: ``qe_python_amd.iterative.rmm_diis``."""
    from .iterative import rmm_diis as solver
    return solver(*args, **kwargs)
