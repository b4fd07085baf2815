"""Device Rayleigh-Ritz of the Davidson subspace (the reduced eigensolve).

The reference solves the reduced pencil H c = e S c of every Davidson
iteration with LAPACK zhegvd on the host (``generalized_eigh``,
``qepy_pw/_native_fft.pyx:667-691``; ``_lowest_generalized_eigh``,
``qepy_pw/diagonalization.py:83-110``).  Here the pencil stays on the device:

1. ``update`` writes the new rows of H = V^H HV and S = V^H V with MFMA GEMMs
   straight into device matrices and completes them Hermitian
   (``update_projected``, ``diagonalization.py:658-697``).
2. Standardisation W = L^-1 (S = L L^H) is extended incrementally: the old
   block of S never changes between restarts, so only L21 = S21 W11^H and the
   Cholesky of the new diagonal block S22 - L21 L21^H (nnew x nnew, on the
   host) are new; then A = W H W^H gains its new rows.
3. The lowest ``nroots`` eigenpairs of A come from a spectral division:
   sigma is placed above the previous nroots-th Ritz value (the new Ritz
   values can only be lower, Cauchy interlacing), the scaled Newton-Schulz
   iteration X <- (a/2) X (3 I - a^2 X^2) (Chen & Chow, 2014) drives
   X0 = (A - sigma I)/||A - sigma I|| to sign(A - sigma I) with GEMMs only,
   P = (I - X)/2 projects on the m >= nroots eigenvectors below sigma, and a
   Rayleigh-Ritz over range(P Omega) (m x m, on the host) gives the pairs;
   C = W^H U are the coefficients in the Davidson basis.
4. Fallback: at the first Rayleigh-Ritz, when fewer than nroots + 2 Ritz
   values of the previous solve are known, or when the sign iteration has
   not converged (an eigenvalue sits at sigma), the standardised A is
   copied to the host and diagonalised densely (LAPACK heevd, FP64).
   ``spectral=False`` instead copies the pencil to the host and solves it by
   zhegvd exactly as the reference does.

The small host factorisations (Cholesky of the new overlap block, the
eigenproblem of the sampled range) run on one BLAS thread: at these sizes
(<= a few hundred) a multi-threaded BLAS next to torch's own OpenMP pool
spends far longer in thread hand-off than in arithmetic.

Eigenvalues and the S-orthonormality of the coefficients agree with zhegvd to
rounding (tested against the host solve on every iteration of the small and
medium workloads).
"""

from __future__ import annotations

import ctypes

import numpy as np
import scipy.linalg as sla
import torch
from threadpoolctl import ThreadpoolController

from . import _native as nat
from . import device as dv
from .schedule import Read, blocking, may_block

C128 = torch.complex128
_THREADPOOLS = None


def _one_blas_thread():
    """Context limiting the host BLAS/LAPACK to one thread."""
    global _THREADPOOLS
    if _THREADPOOLS is None:
        _THREADPOOLS = ThreadpoolController()
    return _THREADPOOLS.limit(limits=1, user_api="blas")


def _lower_inverse(lower: np.ndarray) -> np.ndarray:
    """Inverse of a nonsingular lower-triangular complex matrix (LAPACK
    ztrtri; its strictly upper part is zero)."""
    inv, info = sla.lapack.ztrtri(np.asarray(lower, dtype=np.complex128), lower=1)
    if info != 0:
        raise np.linalg.LinAlgError(f"ztrtri failed (info {info})")
    return np.tril(inv)


def _cholesky_and_inverse(m: np.ndarray):
    """Lower Cholesky factor of a Hermitian positive-definite host matrix
    and its inverse, on one BLAS thread."""
    with _one_blas_thread():
        lower = np.linalg.cholesky(m)
        return lower, _lower_inverse(lower)


def _scaled_ns_alphas(lower: float, tol: float = 1e-15, limit: int = 60) -> list[float]:
    """Chen-Chow scaling factors for the sign iteration from a lower bound
    ``lower`` of |eigenvalues| of X0 (all |x| <= 1)."""
    alphas = []
    ell = max(min(lower, 1.0), 1e-16)
    while 1.0 - ell > tol and len(alphas) < limit:
        a = np.sqrt(3.0 / (1.0 + ell + ell * ell))
        alphas.append(float(a))
        ell = 0.5 * a * ell * (3.0 - a * a * ell * ell)
    return alphas


_GAUSS: dict = {}

# Run the standardisation and the spectral solve as native calls when the
# driver waits on reads itself (qepw_rr_standardize / qepw_rr_spectral: the
# same steps, their host LAPACK calls inside the library); False keeps the
# Python steps everywhere
NATIVE_RR = True


def _gaussian_columns(size: int, device) -> torch.Tensor:
    """A fixed seeded complex Gaussian (size, size) block on the device,
    drawn and uploaded once per process: the range sample only has to be
    generic, and reusing it keeps the host out of the loop and the solves
    reproducible."""
    key = (str(device), int(size))
    block = _GAUSS.get(key)
    if block is None:
        rng = np.random.default_rng(1234)
        g = (rng.standard_normal((size, size)) + 1j * rng.standard_normal((size, size)))
        block = dv.upload(g / np.sqrt(2.0), device)
        _GAUSS[key] = block
    return block


class DeviceRayleighRitz:
    """The device-resident reduced problem of one Davidson solve."""

    def __init__(self, maxsub: int, nroots: int, device, ws: dv.Workspace, *,
                 spectral: bool = True, extra: int = 16, check: bool = False,
                 host_single: bool = False):
        self.host_single = host_single   # host Hermitian eigensolves in complex64
        self.maxsub = int(maxsub)
        self.nroots = int(nroots)
        self.device = device
        self.ws = ws
        self.spectral = spectral
        self.extra = int(extra)
        self.check = check
        def z():
            m = torch.empty((self.maxsub, self.maxsub), dtype=C128, device=device)
            dv.set_block(m, self.maxsub, self.maxsub)
            return m
        # column-major matrices: tensor[c, r] holds element (r, c)
        self.H, self.S, self.W, self.A = z(), z(), z(), z()
        self.X, self.X2, self.Y = z(), z(), z()
        # read-back buffer of the spectral solve: sig[0:2] and its partial
        # sums, bound[0] and the column sums behind it (written by
        # qepw_sign_check / qepw_shift_scale before any read), then the Gram
        # matrix of the range sample -- one contiguous prefix per read-back
        self._head = 2 * self.maxsub + 4
        self.rb = torch.empty(self._head + 2 * self.maxsub * self.maxsub, dtype=torch.float64,
                              device=device)
        self.sig = self.rb[:self.maxsub + 2]
        self.bound = self.rb[self.maxsub + 2:2 * self.maxsub + 3]
        self.standard = 0            # W and A are valid for [:standard]
        self.theta = None            # previous Ritz values (host, ascending)
        self.sigma = None            # shift of the previous spectral solve
        self.drift = None            # max change of the lowest k Ritz values between solves
        self._last_k = None
        self._bound_prev = None      # ||A - sigma I||_1 of the previous solve (host)
        self._slow_shift = False     # the previous sign iteration needed extra steps
        self.u_prev = None           # previous lowest eigenvectors of A (device, (k, maxsub))
        self.stats = {"spectral": 0, "dense": 0, "host": 0, "ns_iterations": 0,
                      "max_check_error": 0.0, "skipped": 0, "ns_unconverged": 0,
                      "too_few_below": 0, "gram_not_pd": 0}

    # -- views -------------------------------------------------------------
    def block(self, mat: torch.Tensor, r0: int, r1: int, c0: int, c1: int) -> torch.Tensor:
        """Column-major view of rows r0:r1, columns c0:c1."""
        return mat[c0:c1, r0:r1]

    def host_block(self, mat: torch.Tensor, n: int):
        """(steps) The leading n x n block of a column-major device matrix
        as a host array."""
        return (yield Read(mat[:n, :n])).T

    def _gemm(self, op, m, n, k, alpha, a, b, beta, c):
        dv.gemm(op, m, n, k, alpha, a, b, beta, c, self.ws)

    # -- step 1: reduced matrices -----------------------------------------
    def update(self, V: torch.Tensor, HV: torch.Tensor, first: int, active: int, npw: int):
        """New rows first:active of H = V^H HV and S = V^H V.  complex64
        blocks are multiplied on FP32 MFMA with FP32 accumulation and the
        rows widened to the FP64 pencil."""
        nnew = active - first
        for mat, right in ((self.H, HV), (self.S, V)):
            target = self.block(mat, first, active, 0, active)
            if V.dtype == C128:
                self._gemm(nat.OP_C, nnew, active, npw, 1.0, V[first:active], right[:active],
                           0.0, target)
            else:
                rows = torch.empty((active, nnew), dtype=V.dtype, device=self.device)
                self._gemm(nat.OP_C, nnew, active, npw, 1.0, V[first:active], right[:active],
                           0.0, rows)
                dv.copy_columns(rows, target, nnew, active)
            rc = nat.load().qepw_hermitian_complete(mat.data_ptr(), self.maxsub, first, active,
                                                    dv.stream_handle(self.device))
            nat.check(rc, "qepw_hermitian_complete")

    def restart(self, values: np.ndarray):
        """Restart block (diagonalization.py:843-868): H = diag(e), S = I."""
        k = len(values)
        e = dv.upload(np.asarray(values, dtype=np.float64), self.device)
        lib = nat.load()
        st = dv.stream_handle(self.device)
        for mat, d in ((self.H, e.data_ptr()), (self.S, None), (self.W, None),
                       (self.A, e.data_ptr())):
            nat.check(lib.qepw_set_diagonal(mat.data_ptr(), self.maxsub, k, d, st),
                      "qepw_set_diagonal")
        self.standard = k
        # the shift stays valid: the new basis holds the k Ritz vectors, all
        # below it, so by interlacing at least k eigenvalues stay below it.
        # The restart keeps the Ritz values: the next solve's drift is
        # measured against them
        self._last_k = np.asarray(values[:self.nroots], dtype=np.float64).copy()
        # the Ritz vectors are the new basis: the eigenvectors of A = diag
        if self.u_prev is None:
            self.u_prev = torch.empty((self.nroots, self.maxsub), dtype=C128, device=self.device)
        dv.set_block(self.u_prev, self.maxsub, self.nroots, 1.0)

    def seed(self, values: np.ndarray):
        """Ritz-value estimates (ascending) of the pencil to come, e.g. those
        a previous stage found above the roots: they let the first solve after
        a restart place its shift instead of diagonalising densely."""
        self.theta = np.sort(np.asarray(values, dtype=np.float64))
        self.sigma = None

    # -- step 2: standardisation ------------------------------------------
    def _standardize(self, active: int):
        first = self.standard
        if first >= active:
            return
        nnew = active - first
        st = dv.stream_handle(self.device)
        if first == 0:
            m = yield from self.host_block(self.S, active)
        else:
            # Z = L21^H = W11 S12 ; M = S22 - Z^H Z
            Z = torch.empty((nnew, first), dtype=C128, device=self.device)
            self._gemm(nat.OP_N, first, nnew, first, 1.0, self.block(self.W, 0, first, 0, first),
                       self.block(self.S, 0, first, first, active), 0.0, Z)
            M = torch.empty((nnew, nnew), dtype=C128, device=self.device)
            dv.copy_columns(self.block(self.S, first, active, first, active), M, nnew, nnew)
            self._gemm(nat.OP_C, nnew, nnew, first, -1.0, Z, Z, 1.0, M)
            m = (yield Read(M)).T
        m = 0.5 * (m + m.conj().T)
        try:
            _, w22 = _cholesky_and_inverse(m)
        except np.linalg.LinAlgError as exc:
            raise np.linalg.LinAlgError(f"zhegvd failed: overlap not positive definite ({exc})")
        w22_dev = dv.upload(w22.T, self.device)
        dv.copy_columns(w22_dev, self.block(self.W, first, active, first, active), nnew, nnew)
        # the strictly upper part of the new columns stays zero
        if first > 0:
            dv.set_block(self.block(self.W, 0, first, first, active), first, nnew)
        if first > 0:
            # W21 = -W22 (Z^H W11)
            T = torch.empty((first, nnew), dtype=C128, device=self.device)
            self._gemm(nat.OP_C, nnew, first, first, 1.0, Z,
                       self.block(self.W, 0, first, 0, first), 0.0, T)
            self._gemm(nat.OP_N, nnew, first, nnew, -1.0, w22_dev, T, 0.0,
                       self.block(self.W, first, active, 0, first))
        # new rows of A: A[first:, :] = W[first:, :] H W^H
        R = torch.empty((active, nnew), dtype=C128, device=self.device)
        self._gemm(nat.OP_N, nnew, active, active, 1.0,
                   self.block(self.W, first, active, 0, active),
                   self.block(self.H, 0, active, 0, active), 0.0, R)
        self._gemm(nat.OP_N | nat.OP_BC, nnew, active, active, 1.0, R,
                   self.block(self.W, 0, active, 0, active), 0.0,
                   self.block(self.A, first, active, 0, active))
        rc = nat.load().qepw_hermitian_complete(self.A.data_ptr(), self.maxsub, first, active, st)
        nat.check(rc, "qepw_hermitian_complete")
        self.standard = active

    # -- step 3/4: eigenpairs --------------------------------------------
    def solve(self, active: int):
        """Lowest nroots pairs: (values host float64, coefficients device
        (nroots, active) column-major active x nroots), with blocking reads."""
        return blocking(self.solve_steps(active))

    def solve_steps(self, active: int):
        """(steps) ``solve`` as a generator of ``Read`` requests
        (``qe_python_amd.schedule``)."""
        if not self.spectral:
            values, coef, _ = yield from self._host(active)
            self.stats["host"] += 1
            return values, coef
        k = self.nroots
        # under a driver that waits on each read itself, the standardisation
        # and the spectral solve run as one native call each (their host
        # round trips inside the library); under run_many they yield
        native = NATIVE_RR and may_block()
        spectral = self.theta is not None and active > k + 8
        out = None
        if native:
            prepared = self._spectral_native_prepare(active) if spectral else None
            self._standardize_native(active)
            if spectral:
                out = self._spectral_native_run(prepared)
        else:
            yield from self._standardize(active)
            if spectral:
                out = yield from self._spectral(active)
        if not spectral:
            self.stats["skipped"] += 1
        if out is not None:
            self.stats["spectral"] += 1
        else:
            out = yield from self._dense(active)
            self.stats["dense"] += 1
        values, coef = out
        if self.check:
            ref, _, _ = yield from self._host(active)
            err = float(np.max(np.abs(ref - values)))
            self.stats["max_check_error"] = max(self.stats["max_check_error"], err)
        return values, coef

    def _host(self, active):
        """zhegvd of the copied pencil on the host, as the reference does."""
        h = yield from self.host_block(self.H, active)
        s = yield from self.host_block(self.S, active)
        h = np.array(h, dtype=np.complex128, order="F")
        s = np.array(s, dtype=np.complex128, order="F")
        try:
            w, v = sla.eigh(h, s, lower=True, type=1, driver="gvd", overwrite_a=True,
                            overwrite_b=True, check_finite=False)
        except sla.LinAlgError as exc:
            raise np.linalg.LinAlgError(f"zhegvd failed: {exc}") from exc
        k = self.nroots
        coef = torch.from_numpy(np.ascontiguousarray(v[:, :k].T)).to(self.device)
        return w[:k].copy(), coef, w.copy()

    def _dense(self, n):
        """All eigenpairs of the standardised A on the host (LAPACK heevd on
        one BLAS thread, ~2 ms at 150 rows; the first solve of a pencil and
        the fallbacks); C = W^H U on the device."""
        k = self.nroots
        a = yield from self.host_block(self.A, n)
        w, v = self._host_eigh(0.5 * (a + a.conj().T))
        U = dv.upload(np.ascontiguousarray(v[:, :k].T), self.device)
        coef = torch.empty((k, n), dtype=C128, device=self.device)
        self._gemm(nat.OP_C, n, k, n, 1.0, self.block(self.W, 0, n, 0, n), U, 0.0, coef)
        self._keep_vectors(U, n)
        values = np.asarray(w, dtype=np.float64)
        self._note_drift(values)
        self.theta = values
        self.sigma = None
        return values[:k].copy(), coef

    def _host_eigh(self, a: np.ndarray):
        """All eigenpairs of a Hermitian host matrix (LAPACK heevd), in
        complex64 when ``host_single`` (the complex64 phase of the mixed form,
        whose vectors are complex64 anyway), returned in FP64."""
        with _one_blas_thread():
            if not self.host_single:
                return np.linalg.eigh(a)
            w, v = sla.eigh(a.astype(np.complex64), driver="evd", check_finite=False)
        return w.astype(np.float64), v.astype(np.complex128)

    def _can_keep_shift(self) -> bool:
        theta, k = self.theta, self.nroots
        # a shift whose last sign iteration needed the extra steps has an
        # eigenvalue close to it (Ritz values above it drift down towards it
        # as the basis grows): a new one is taken in a gap instead
        return (self.sigma is not None and theta is not None and not self._slow_shift
                and k <= len(theta) <= k + self.extra and theta[-1] < self.sigma
                and self.drift is not None and self.drift < 0.1 * (self.sigma - theta[-1]))

    def _note_drift(self, values):
        k = self.nroots
        low = np.asarray(values[:k], dtype=np.float64)
        if self._last_k is not None and len(self._last_k) == len(low):
            self.drift = float(np.max(np.abs(low - self._last_k)))
        self._last_k = low.copy()

    def _keep_vectors(self, U, n):
        k = self.nroots
        if self.u_prev is None:
            self.u_prev = torch.empty((k, self.maxsub), dtype=C128, device=self.device)
        dv.set_block(self.u_prev, self.maxsub, k)
        dv.copy_columns(U, self.u_prev, n, k)

    def _choose_shift(self):
        """(sigma, lo, hi): the shift of the next spectral solve and the
        eigenvalue bracket around it, from the previous Ritz values; None
        when too few are known."""
        k = self.nroots
        theta = self.theta
        if self._can_keep_shift():
            # keep the previous shift: the subspace only grew since, so by
            # Cauchy interlacing at least the len(theta) >= k eigenvalues
            # found below it last time still lie below it; while the Ritz
            # values still drop fast (drift comparable to the distance to the
            # shift) many more would, so a shift from the gaps is taken then
            sigma = self.sigma
            lo_v, hi_v = theta[-1], 2.0 * sigma - theta[-1]
        elif len(theta) < k + 2:
            # too few values known above the k-th for a gap: one mean spacing
            # of the top values above the last.  The grown A has at least
            # len(theta) >= k eigenvalues at or below theta[-1] (interlacing),
            # so at least k lie below sigma; an eigenvalue close to sigma only
            # slows the iteration, which the convergence check catches
            if len(theta) < k:
                return None
            top = theta[max(0, len(theta) - 5):]
            spacing = (top[-1] - top[0]) / max(1, len(top) - 1)
            if not spacing > 0.0:
                return None
            sigma = theta[-1] + spacing
            lo_v, hi_v = theta[-1], theta[-1] + 2.0 * spacing
        else:
            # a new shift in the widest gap among theta[k-1+margin .. hi-1]:
            # at least k + margin eigenvalues of the grown A lie below it
            hi = min(len(theta), k + self.extra)
            margin = min(4, len(theta) - k - 1)
            lo_i = k - 1 + margin
            gaps = np.diff(theta[lo_i:hi])
            j = int(np.argmax(gaps))
            lo_v, hi_v = theta[lo_i + j], theta[lo_i + j + 1]
            sigma = 0.5 * (lo_v + hi_v)
            spacing = (theta[hi - 1] - theta[k - 1]) / max(1, hi - k)
            if self.drift is not None and self.drift > 3.0 * spacing:
                # the Ritz values still move by several spacings of the
                # window: the old gaps say little about the grown A, whose new
                # eigenvalues crowd in above theta[k-1] (up to one per new
                # basis vector), so the shift goes just above theta[k-1] --
                # still at least k eigenvalues below it, and far fewer for
                # the host eigensolve; a close eigenvalue only costs sign
                # iteration steps
                gap = theta[k] - theta[k - 1]
                lo_v, hi_v = theta[k - 1], theta[k - 1] + 0.1 * gap
                sigma = 0.5 * (lo_v + hi_v)
        return sigma, lo_v, hi_v

    def _standardize_native(self, active: int):
        """``_standardize`` as one native call (qepw_rr_standardize)."""
        first = self.standard
        if first >= active:
            return
        lib = nat.load()
        buf = self.ws.get(int(lib.qepw_rr_workspace(self.maxsub)))
        not_pd = ctypes.c_int(0)
        nat.check(lib.qepw_rr_standardize(self.H.data_ptr(), self.S.data_ptr(),
                                          self.W.data_ptr(), self.A.data_ptr(), self.maxsub,
                                          first, active, ctypes.byref(not_pd), buf.data_ptr(),
                                          buf.numel(), dv.stream_handle(self.device)),
                  "qepw_rr_standardize")
        if not_pd.value:
            raise np.linalg.LinAlgError("zhegvd failed: overlap not positive definite")
        self.standard = active

    def _spectral_native(self, n):
        """``_spectral`` with the sign iteration, the range sample and the
        Rayleigh-Ritz on it as one native call (qepw_rr_spectral)."""
        return self._spectral_native_run(self._spectral_native_prepare(n))

    def _spectral_native_prepare(self, n):
        """The shift and the arguments of the native spectral solve, or None
        when no shift can be placed.  They depend only on the previous Ritz
        values, so they are ready before the standardisation runs and the
        device does not wait for them between the two calls."""
        k = self.nroots
        shift = self._choose_shift()
        if shift is None:
            return None
        sigma, lo_v, hi_v = shift
        lib = nat.load()
        coef = torch.empty((k, n), dtype=C128, device=self.device)
        buf = self.ws.get(int(lib.qepw_rr_workspace(self.maxsub)))
        args = nat.RRSpectralArgs(
            n=n, k=k, maxsub=self.maxsub, count_hint=len(self.theta) + 24,
            host_single=int(self.host_single), sigma=sigma,
            gap=hi_v - lo_v, bound_prev=self._bound_prev or 0.0, A=self.A.data_ptr(),
            W=self.W.data_ptr(), X=self.X.data_ptr(), X2=self.X2.data_ptr(),
            Y=self.Y.data_ptr(), gauss=_gaussian_columns(self.maxsub, self.device).data_ptr(),
            coef=coef.data_ptr(), rb=self.rb.data_ptr(), work=buf.data_ptr(),
            work_bytes=buf.numel(), stream=dv.stream_handle(self.device))
        # buf stays referenced with the arguments that point into it
        return args, coef, sigma, buf

    def _spectral_native_run(self, prepared):
        if prepared is None:
            return None
        args, coef, sigma, _ = prepared
        k = self.nroots
        lib = nat.load()
        # the previous eigenvectors (allocated here, after the standardisation
        # has succeeded, so that a failed one leaves no unset block behind)
        have_prev = self.u_prev is not None
        if not have_prev:
            self.u_prev = torch.empty((k, self.maxsub), dtype=C128, device=self.device)
        args.have_prev = int(have_prev)
        args.u_prev = self.u_prev.data_ptr()
        values = np.empty(self.maxsub, dtype=np.float64)
        res = nat.RRSpectralResult()
        nat.check(lib.qepw_rr_spectral(ctypes.byref(args), values.ctypes.data,
                                       ctypes.byref(res)), "qepw_rr_spectral")
        self.stats["ns_iterations"] += res.ns_steps
        self._bound_prev = res.bound
        self._slow_shift = bool(res.slow_shift)
        if res.status == 4:
            raise np.linalg.LinAlgError("heevd failed on the reduced Rayleigh-Ritz matrix")
        if res.status:
            self.stats[{1: "ns_unconverged", 2: "too_few_below", 3: "gram_not_pd"}[res.status]] += 1
            if not have_prev:
                self.u_prev = None
            return None
        w = values[:res.m_below].copy()
        self._note_drift(w)
        self.theta = w
        self.sigma = sigma
        return w[:k].copy(), coef

    def _spectral(self, n):
        k = self.nroots
        st = dv.stream_handle(self.device)
        lib = nat.load()
        theta = self.theta
        shift = self._choose_shift()
        if shift is None:
            return None
        sigma, lo_v, hi_v = shift
        A = self.block(self.A, 0, n, 0, n)
        X, X2, Y = (self.block(m, 0, n, 0, n) for m in (self.X, self.X2, self.Y))
        nat.check(lib.qepw_shift_scale(A.data_ptr(), self.maxsub, n, sigma, X.data_ptr(),
                                       self.maxsub, self.bound.data_ptr(), st),
                  "qepw_shift_scale")
        # the scaling schedule needs ||A - sigma I||_1 only as an estimate
        # (the device scales X0 by the exact value; the scaled steps map
        # (0, 1] into itself for any schedule, a poor one just converges
        # later): twice the previous solve's value avoids a read-back here
        if self._bound_prev is None:
            self._bound_prev = float((yield Read(self.bound[:1]))[0])
            bound_est = self._bound_prev
        else:
            bound_est = 2.0 * self._bound_prev
        lower = 0.25 * (hi_v - lo_v) / bound_est
        alphas = _scaled_ns_alphas(lower) + [1.0, 1.0]
        buf = self.ws.get(int(lib.qepw_sign_iteration_workspace(n)))
        in_x2 = ctypes.c_int(0)

        def sign_steps(schedule):
            nonlocal X, X2
            # the steps run back to back inside the library, as products of
            # commuting Hermitian matrices whose lower triangles are computed
            # and mirrored; Y is left holding the square of the result
            coeffs = (ctypes.c_double * len(schedule))(*schedule)
            nat.check(lib.qepw_sign_iteration(X.data_ptr(), X2.data_ptr(), Y.data_ptr(),
                                              self.maxsub, n, coeffs, len(schedule),
                                              ctypes.byref(in_x2), buf.data_ptr(), buf.numel(),
                                              st), "qepw_sign_iteration")
            if in_x2.value:
                X, X2 = X2, X
            self.stats["ns_iterations"] += len(schedule)
            nat.check(lib.qepw_sign_check(X.data_ptr(), self.maxsub, Y.data_ptr(), self.maxsub, n,
                                          self.sig.data_ptr(), st), "qepw_sign_check")

        def sampled_range(count):
            """Q0 = (I - X) Omega / 2 for Omega = [previous eigenvectors,
            Gaussian columns] (count columns) and its Gram matrix, with the
            sign check and the exact bound ahead of it in one device buffer
            (a single read-back)."""
            # rows [:count] and columns [:n] are all written below
            omega = torch.empty((count, self.maxsub), dtype=C128, device=self.device)
            start = 0
            if self.u_prev is not None:
                start = min(k, count)
                dv.copy_columns(self.u_prev, omega, n, start)
            if start < count:
                gauss = _gaussian_columns(self.maxsub, self.device)
                dv.copy_columns(gauss[start:count], omega[start:count], n, count - start)
            Om = omega[:, :n]
            Q0 = torch.empty((count, n), dtype=C128, device=self.device)
            dv.copy_columns(Om, Q0, n, count)
            self._gemm(nat.OP_N, n, count, n, -0.5, X, Om, 0.5, Q0)
            h = self._head
            G = torch.view_as_complex(self.rb[h:h + 2 * count * count].view(count, count, 2))
            self._gemm(nat.OP_C, count, count, n, 1.0, Q0, Q0, 0.0, G)
            return Q0, self.rb[:h + 2 * count * count]

        def unpack(rb, count):
            host = yield Read(rb)
            g = host[self._head:].view(np.complex128).reshape(count, count).T
            return (float(host[0]), float(host[1]), float(host[self.maxsub + 2]), g)

        # optimistic path: one read-back carries the sign check, the exact
        # bound and the Gram matrix of a range sample sized for the expected
        # number of eigenvalues below sigma
        sign_steps(alphas)
        count = min(n, len(theta) + 24)
        Q0, rb = sampled_range(count)
        res2, trace, self._bound_prev, g = yield from unpack(rb, count)
        attempt = 0
        self._slow_shift = res2 >= 1e-20 * n
        while res2 >= 1e-20 * n and attempt < 2:
            sign_steps([1.0] * 12)
            res2, trace = (float(v) for v in (yield Read(self.sig[:2])))
            attempt += 1
            g = None
        if res2 >= 1e-20 * n:
            self.stats["ns_unconverged"] += 1
            return None
        m_below = int(round(0.5 * (n - trace)))
        if m_below < k:
            self.stats["too_few_below"] += 1
            return None
        if g is None or m_below > count:
            count = max(count, m_below)
            Q0, rb = sampled_range(count)
            _, _, _, g = yield from unpack(rb, count)
        mp = m_below
        Q0 = Q0[:mp]
        g = g[:mp, :mp]
        g = 0.5 * (g + g.conj().T)
        # Q0 = P Omega has exactly rank m_below = mp (Omega: the previous
        # eigenvectors, then Gaussian columns); its Gram matrix is factored by
        # Cholesky and Q = Q0 L^-H is an orthonormal basis of range(P)
        try:
            _, lg_inv = _cholesky_and_inverse(g)
        except np.linalg.LinAlgError:
            self.stats["gram_not_pd"] += 1
            return None
        t = lg_inv.conj().T
        t_dev = dv.upload(t.T, self.device)
        Q = torch.empty_like(Q0)
        self._gemm(nat.OP_N, n, mp, mp, 1.0, Q0, t_dev, 0.0, Q)
        AQ = torch.empty_like(Q)
        self._gemm(nat.OP_N, n, mp, n, 1.0, A, Q, 0.0, AQ)
        B = torch.empty((mp, mp), dtype=C128, device=self.device)
        self._gemm(nat.OP_C, mp, mp, n, 1.0, Q, AQ, 0.0, B)
        bt = (yield Read(B)).T
        bt = 0.5 * (bt + bt.conj().T)
        w, y = self._host_eigh(bt)
        if len(w) < k:
            return None
        y_dev = dv.upload(y[:, :k].T, self.device)
        U = torch.empty((k, n), dtype=C128, device=self.device)
        self._gemm(nat.OP_N, n, k, mp, 1.0, Q, y_dev, 0.0, U)
        coef = torch.empty((k, n), dtype=C128, device=self.device)
        self._gemm(nat.OP_C, n, k, n, 1.0, self.block(self.W, 0, n, 0, n), U, 0.0, coef)
        self._keep_vectors(U, n)
        self._note_drift(w)
        self.theta = np.asarray(w, dtype=np.float64)
        self.sigma = sigma
        return w[:k].copy(), coef
