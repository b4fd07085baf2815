"""CG, ParO and RMM-DIIS on the device ``h_psi`` (drop-ins for the
reference's ``conjugate_gradient``, ``parallel_orbital`` and ``rmm_diis``,
``qepy_pw/diagonalization.py:993-1186, 1189-1467, 1487-1756``).

The bands live on the device as (nb, npw) blocks in the plan's row order, as
in the device Davidson: H applications go through
``PlaneWaveHamiltonian.device_apply`` (or a host operator driven from device
storage), the block products through ``qepw_gemm``, and every per-column
update, preconditioner and reduction through the HIP kernels of
``libqepw_hip.so`` (``qepw_column_axpby``, ``qepw_qe_precondition``,
``qepw_rmm_precondition``, ``qepw_history_combine``,
``qepw_column_inner_products``, ``qepw_column_diagonal_expectations``).
Only per-band scalars (inner products, line-search coefficients, the
active-set decisions) and the small reduced eigenproblems (ParO's
generalized Rayleigh-Ritz, RMM-DIIS's per-band DIIS and closing rotation,
the initial-subspace rotation) are computed on the host, as the reference
computes them with LAPACK.  The three solvers keep the reference's
control flow step for step, so their iteration and H-application counts
follow the reference's algorithm.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _native as nat
from . import device as dv
from .diagonalization import (
    _SERIAL, DavidsonResult, _HostOperator, _device_hamiltonian, _lowest_generalized_eigh,
    _orthonormalize_device, _result_vectors, _upload_block,
)
from .eigensolver import _one_blas_thread

C128 = torch.complex128


def _guard(x):
    """x where |x| > 1e-30, else +-1e-30 with the sign of x (+ for 0), as the
    reference guards its CG denominators (``diagonalization.py:1248-1252``)."""
    return np.where(np.abs(x) > 1.0e-30, x, np.copysign(1.0e-30, x + (x == 0.0)))


def _root_tolerances(number_of_roots, tolerance, occupied_roots, full_accuracy):
    """``_iterative_root_tolerances`` (``diagonalization.py:883-899``)."""
    if occupied_roots is None or full_accuracy:
        occupied_roots = number_of_roots
    if not 0 <= occupied_roots <= number_of_roots:
        raise ValueError("occupied_roots must be between zero and number_of_roots")
    tol = np.full(number_of_roots, max(5.0 * tolerance, 5.0e-6))
    tol[:occupied_roots] = tolerance
    return tol


class _Blocks:
    """Device operator, workspace and the block algebra of one solve."""

    def __init__(self, operator, operator_into, diagonal, mpi):
        mpi = mpi if mpi is not None else _SERIAL
        if getattr(mpi, "size", 1) > 1:
            raise ValueError("the device solvers run on one GPU; MPI size must be 1")
        diagonal = np.asarray(diagonal, dtype=float)
        self.npw = npw = len(diagonal)
        ham = _device_hamiltonian(operator, operator_into)
        if ham is not None:
            if ham.npw != npw:
                raise ValueError("diagonal length does not match the Hamiltonian basis")
            self.device, self.perm_dev, perm = ham.device, ham.plan.perm_device, ham.plan.perm
            self.op, self.ws = ham, ham.workspace
        else:
            self.device = dv.require_device()
            self.perm_dev, perm = None, None
            self.op = _HostOperator(operator, operator_into, npw, self.device)
            self.ws = dv.Workspace(self.device)
        self.diagonal = diagonal if perm is None else diagonal[perm]
        self.diag = torch.from_numpy(np.ascontiguousarray(self.diagonal)).to(self.device)
        self.happ = 0
        self._dots = torch.empty(64, dtype=torch.float64, device=self.device)

    # -- movement -----------------------------------------------------------
    def upload(self, vectors) -> torch.Tensor:
        return _upload_block(vectors, self.perm_dev, self.npw, self.device, C128)

    def result(self, block, device_output):
        return _result_vectors(block, self.perm_dev, self.npw, device_output)

    def col(self, values) -> torch.Tensor:
        # pinned and asynchronous: a pageable copy would wait for the queue
        return dv.upload(np.array(values, dtype=np.float64).reshape(-1), self.device)

    def index(self, positions) -> torch.Tensor:
        """Host positions -> device int32 map for ``take`` / ``put``."""
        return dv.index_tensor(positions, self.device)

    def take(self, block: torch.Tensor, idx: torch.Tensor) -> torch.Tensor:
        """The columns idx of a (nb, npw) block, gathered into a new block
        (qepw_copy_columns)."""
        count = int(idx.shape[0])
        out = torch.empty((count, block.shape[1]), dtype=block.dtype, device=self.device)
        if count:
            dv.copy_columns(block, out, self.npw, count, src_cols=idx)
        return out

    def put(self, dst: torch.Tensor, idx: torch.Tensor, src: torch.Tensor) -> None:
        """dst[idx[j]] = src[j] (qepw_copy_columns)."""
        if int(idx.shape[0]):
            dv.copy_columns(src, dst, self.npw, int(idx.shape[0]), dst_cols=idx)

    # -- algebra ------------------------------------------------------------
    def apply(self, src: torch.Tensor) -> torch.Tensor:
        dst = torch.empty_like(src)
        if src.shape[0]:
            self.op.device_apply(src, dst)
        self.happ += int(src.shape[0])
        return dst

    def gram_dev(self, a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
        """A^H B as a device (nb, na) tensor (column-major na x nb)."""
        out = torch.empty((b.shape[0], a.shape[0]), dtype=C128, device=self.device)
        dv.gemm(nat.OP_C, a.shape[0], b.shape[0], self.npw, 1.0, a, b, 0.0, out, self.ws)
        return out

    def gram(self, a, b) -> np.ndarray:
        return self.gram_dev(a, b).cpu().numpy().T

    def rotate(self, a: torch.Tensor, coef: np.ndarray) -> torch.Tensor:
        """(nc, npw) block of A @ coef for a host (na, nc) coefficient matrix."""
        out = torch.empty((coef.shape[1], self.npw), dtype=C128, device=self.device)
        t = dv.upload(np.asarray(coef, dtype=np.complex128).T, self.device)
        dv.gemm(nat.OP_N, self.npw, coef.shape[1], a.shape[0], 1.0, a, t, 0.0, out, self.ws)
        return out

    def project_out(self, ref: torch.Tensor, x: torch.Tensor, *others) -> None:
        """x -= ref (ref^H x); each (r, y) in ``others`` gets y -= r (ref^H x)."""
        if not ref.shape[0] or not x.shape[0]:
            return
        # one vector (CG's projections): qepw_gemm runs these one-column
        # products as matrix-vector passes
        p = self.gram_dev(ref, x)
        dv.gemm(nat.OP_N, self.npw, x.shape[0], ref.shape[0], -1.0, ref, p, 1.0, x, self.ws)
        for r, y in others:
            dv.gemm(nat.OP_N, self.npw, x.shape[0], ref.shape[0], -1.0, r, p, 1.0, y, self.ws)

    def dots(self, *pairs) -> list:
        """Per-column <l|r> of each (l, r) pair (complex host arrays), with one
        device-to-host copy for all of them."""
        total = sum(int(l.shape[0]) for l, _ in pairs)
        if 2 * total > self._dots.numel():
            self._dots = torch.empty(2 * total, dtype=torch.float64, device=self.device)
        lib = nat.load()
        off = 0
        for l, r in pairs:
            nb = int(l.shape[0])
            if nb:
                rc = lib.qepw_column_inner_products(
                    nat.C128, l.data_ptr(), dv.ld_of(l), r.data_ptr(), dv.ld_of(r), self.npw, nb,
                    self._dots.data_ptr() + 16 * off, dv.stream_handle(self.device))
                nat.check(rc, "qepw_column_inner_products")
            off += nb
        host = self._dots[:2 * total].cpu().numpy().view(np.complex128)
        out, off = [], 0
        for l, _ in pairs:
            nb = int(l.shape[0])
            out.append(host[off:off + nb].copy())
            off += nb
        return out

    def axpby(self, x, a, y, b) -> None:
        """y = a x + b y per column; a, b: host arrays, scalars or None (1)."""
        n = int(y.shape[0])
        if not n:
            return

        def coef(c):
            if c is None:
                return None
            return self.col(np.broadcast_to(np.asarray(c, dtype=np.float64), (n,)))
        dv.column_axpby(x, coef(a), y, coef(b), self.npw)

    def scale(self, src, s=None, cfac=None, dst=None):
        """dst = s[g] cfac[b] src (s: device per row, cfac: host per column)."""
        dst = torch.empty_like(src) if dst is None else dst
        c = None if cfac is None else self.col(np.broadcast_to(np.asarray(cfac, float),
                                                               (src.shape[0],)))
        rc = nat.load().qepw_scale_rows(
            nat.C128, src.data_ptr(), dv.ld_of(src), None if s is None else s.data_ptr(), 1,
            None if c is None else c.data_ptr(), dst.data_ptr(), dv.ld_of(dst), self.npw,
            src.shape[0], dv.stream_handle(self.device))
        nat.check(rc, "qepw_scale_rows")
        return dst

    def residual_norms(self, hv, v, values) -> np.ndarray:
        r = hv.clone()
        if r.shape[0]:
            dv.subtract_band_energies(r, v, self.col(values), self.npw)
        sq = dv.column_squared_norms(r, self.npw).cpu().numpy()
        return np.sqrt(np.maximum(sq, 0.0))

    # -- shared set-up ------------------------------------------------------
    def initial_subspace(self, n, initial_vectors, *, initial_is_orthonormal,
                         initial_applied=None, initial_eigenvalues=None, initial_is_ritz=False):
        """``_initial_iterative_subspace`` (``diagonalization.py:932-990``)."""
        shape = tuple(getattr(initial_vectors, "shape", ()))
        if len(shape) != 2 or shape[1] < n:
            raise ValueError("initial_vectors must contain at least number_of_roots columns")
        if shape[0] != self.npw:
            raise ValueError("initial_vectors must have one row per plane wave")
        trial = self.upload(initial_vectors)
        if initial_is_orthonormal:
            v = trial[:n].clone()
        else:
            v = _orthonormalize_device(trial, self.npw, self.ws)[:n]
        if v.shape[0] < n:
            raise ValueError("initial vectors are linearly dependent")
        if initial_is_ritz:
            if initial_applied is None or initial_eigenvalues is None:
                raise ValueError("initial Ritz vectors require H-vectors and eigenvalues")
            if tuple(initial_applied.shape) != (self.npw, n):
                raise ValueError("initial_applied must match initial vectors")
            return (v.contiguous(), np.asarray(initial_eigenvalues[:n], dtype=float).copy(),
                    self.upload(initial_applied), 0)
        v = v.contiguous()
        hv = self.apply(v)
        proj = self.gram(v, hv)
        proj = 0.5 * (proj + proj.conj().T)
        with _one_blas_thread():
            w, u = np.linalg.eigh(proj)
        u = u[:, :n]
        return self.rotate(v, u), np.real(w[:n]).copy(), self.rotate(hv, u), n


# ---------------------------------------------------------------------------
# conjugate gradient (QE ccgdiagg)
# ---------------------------------------------------------------------------

def conjugate_gradient(operator, diagonal, number_of_roots, *, initial_vectors,
                       tolerance=1.0e-8, max_iterations=20, occupied_roots=None,
                       full_accuracy=False, initial_is_orthonormal=False, mpi=None,
                       operator_into=None, initial_applied=None, initial_eigenvalues=None,
                       initial_is_ritz=False, device_output=False):
    """Band-by-band CG (``diagonalization.py:993-1186``): the Ry-kinetic
    preconditioner P = 1 + T + sqrt(1 + (T - 1)^2), the P-metric gradient
    orthogonal to psi and to the lower bands, Polak-Ribiere conjugation with
    QE's analytic psi correction and the exact two-vector line minimisation.
    ``device_output=True`` returns the eigenvectors as a device tensor."""
    if tolerance <= 0.0:
        raise ValueError("diagonalization tolerance must be positive")
    if max_iterations < 1:
        raise ValueError("diago_cg_maxiter must be at least 1")
    s = _Blocks(operator, operator_into, diagonal, mpi)
    v, values, hv, _ = s.initial_subspace(
        number_of_roots, initial_vectors, initial_is_orthonormal=initial_is_orthonormal,
        initial_applied=initial_applied, initial_eigenvalues=initial_eigenvalues,
        initial_is_ritz=initial_is_ritz)
    tol = _root_tolerances(number_of_roots, tolerance, occupied_roots, full_accuracy)
    t_ry = 2.0 * s.diagonal
    p_host = 1.0 + t_ry + np.sqrt(1.0 + (t_ry - 1.0) ** 2)
    p = s.col(p_host)
    pinv = s.col(1.0 / p_host)
    unconverged, iteration_sum = 0, 0
    for band in range(number_of_roots):
        psi, hpsi = v[band:band + 1], hv[band:band + 1]
        lower, hlower = v[:band], hv[:band]
        if band:
            s.project_out(lower, psi, (hlower, hpsi))
            (nn,) = s.dots((psi, psi))
            nrm = np.sqrt(max(float(np.real(nn[0])), 1.0e-300))
            s.scale(psi, cfac=1.0 / nrm, dst=psi)
            s.scale(hpsi, cfac=1.0 / nrm, dst=hpsi)
        energy = float(np.real(s.dots((psi, hpsi))[0][0]))
        converged = False
        gg_prev = cgn_prev = sin_prev = 0.0
        cg = pg_prev = None
        for it in range(1, max_iterations + 1):
            g = s.scale(hpsi, s=pinv)
            ppsi = s.scale(psi, s=pinv)
            m0, m1 = s.dots((psi, g), (psi, ppsi))
            e_p = float(np.real(m0[0] / m1[0]))
            s.axpby(ppsi, -e_p, g, None)
            if band:
                s.project_out(lower, g)
            pg = s.scale(g, s=p)
            if it > 1:
                cross, gg = (float(np.real(x[0])) for x in s.dots((g, pg_prev), (g, pg)))
            else:
                gg = float(np.real(s.dots((g, pg))[0][0]))
            pg_prev = pg
            if not np.isfinite(gg) or gg <= 1.0e-30:
                break
            if it == 1:
                cg = g.clone()
            else:
                gamma = (gg - cross) / max(gg_prev, 1.0e-300)
                s.axpby(g, None, cg, gamma)
                s.axpby(psi, -(gamma * cgn_prev * sin_prev), cg, None)
            gg_prev = gg
            hcg = s.apply(cg)
            c2, l0, l1 = s.dots((cg, cg), (psi, hcg), (cg, hcg))
            cg2 = float(np.real(c2[0]))
            if cg2 <= 1.0e-30:
                break
            cgn = np.sqrt(cg2)
            a0 = 2.0 * float(np.real(l0[0])) / cgn
            b0 = float(np.real(l1[0])) / cg2
            den = energy - b0
            theta = 0.5 * np.arctan(a0 / den if den != 0.0 else np.copysign(np.inf, a0))
            co, si = float(np.cos(theta)), float(np.sin(theta))
            cos2, sin2 = co * co - si * si, 2.0 * co * si
            r0 = 0.5 * ((energy - b0) * cos2 + a0 * sin2 + energy + b0)
            r1 = 0.5 * (-(energy - b0) * cos2 - a0 * sin2 + energy + b0)
            if r1 < r0:
                theta += 0.5 * np.pi
                co, si = float(np.cos(theta)), float(np.sin(theta))
            old = energy
            energy = float(min(r0, r1))
            s.axpby(cg, si / cgn, psi, co)
            s.axpby(hcg, si / cgn, hpsi, co)
            cgn_prev, sin_prev = cgn, si
            if abs(energy - old) < tol[band]:
                converged = True
                break
        unconverged += 0 if converged else 1
        iteration_sum += it + 1
        values[band] = energy
    order = np.argsort(values)
    values = values[order]
    idx = s.index(order)
    v, hv = s.take(v, idx), s.take(hv, idx)
    return DavidsonResult(values, s.result(v, device_output), s.residual_norms(hv, v, values),
                          max(1, int(round(iteration_sum / number_of_roots))), unconverged == 0,
                          unconverged, s.happ)


# ---------------------------------------------------------------------------
# ParO (QE paro_k_new with bpcg_k)
# ---------------------------------------------------------------------------

def _paro_bpcg(s: _Blocks, reference, x0, hx0, energies, tolerance, max_iterations=5):
    """``_paro_bpcg`` (``diagonalization.py:1189-1325``): block preconditioned
    CG on (H - e) c = e x0 - H x0 in the complement of ``reference``, with the
    rollback of a step that raises the quadratic functional."""
    count = int(x0.shape[0])
    corr = torch.zeros_like(x0)
    hcorr = torch.zeros_like(x0)
    if count == 0:
        return corr, hcorr
    e_all = np.asarray(energies, dtype=float)
    rhs = hx0.clone()
    s.axpby(x0, e_all, rhs, -1.0)                      # e x0 - H x0
    prec = torch.empty_like(rhs)
    dv.qe_precondition(rhs, None, s.col(e_all), s.diag, prec, s.npw)
    s.project_out(reference, prec)
    g0 = np.real(s.dots((prec, rhs))[0])
    thresholds = np.maximum(0.01 * tolerance, 0.01 * g0)
    dirs = prec.clone()
    f_prev = np.zeros(count)
    active = np.arange(count)
    for it in range(1, max_iterations + 1):
        if not len(active):
            break
        full = len(active) == count
        ia = None if full else s.index(active)

        def take(block):
            return block if full else s.take(block, ia)

        d = take(dirs)
        hd = s.apply(d)
        e = e_all[active]
        dd, dhd = s.dots((d, d), (d, hd))
        gamma = np.real(dhd) - e * np.real(dd)
        alpha = g0[active] / _guard(gamma)
        c_act, hc_act = take(corr), take(hcorr)
        s.axpby(d, alpha, c_act, None)
        s.axpby(hd, alpha, hc_act, None)
        r_act = take(rhs)
        grad = r_act.clone()
        s.axpby(c_act, e, grad, None)
        s.axpby(hc_act, -1.0, grad, None)
        newp = torch.empty_like(grad)
        dv.qe_precondition(grad, None, s.col(e), s.diag, newp, s.npw)
        s.project_out(reference, newp)
        g2, g1, cn, ch, cr = s.dots((take(prec), grad), (newp, grad), (c_act, c_act),
                                    (c_act, hc_act), (c_act, r_act))
        g2, g1 = np.real(g2), np.real(g1)
        fval = 0.5 * (np.real(ch) - e * np.real(cn)) - np.real(cr)
        rose = fval > f_prev[active]
        back = rose & (f_prev[active] < 0.0)
        if np.any(back):
            undo = np.where(back, -alpha, 0.0)
            s.axpby(d, undo, c_act, None)
            s.axpby(hd, undo, hc_act, None)
        if not full:
            s.put(corr, ia, c_act)
            s.put(hcorr, ia, hc_act)
        done = (np.abs(g1) < thresholds[active]) | rose | (it == max_iterations)
        keep = ~done
        if not np.any(keep):
            break
        cont = active[keep]
        beta = (g1[keep] - g2[keep]) / _guard(g0[cont])
        ik = s.index(np.flatnonzero(keep))
        d_keep = s.take(d, ik)
        new_keep = s.take(newp, ik)
        s.axpby(new_keep, None, d_keep, beta)
        ic = s.index(cont)
        s.put(dirs, ic, d_keep)
        s.put(prec, ic, new_keep)
        g0[cont] = g1[keep]
        f_prev[cont] = fval[keep]
        active = cont
    return corr, hcorr


def parallel_orbital(operator, diagonal, number_of_roots, *, initial_vectors,
                     tolerance=1.0e-8, max_iterations=20, occupied_roots=None,
                     full_accuracy=False, initial_is_orthonormal=False, mpi=None,
                     operator_into=None, initial_applied=None, initial_eigenvalues=None,
                     initial_is_ritz=False, device_output=False):
    """QE ``paro_k_new`` (``diagonalization.py:1328-1467``): per outer
    iteration a bpcg correction block for the unconverged trusted roots (and
    the leading extra columns), then the generalized Rayleigh-Ritz of
    [basis, corrections]; converged on Ritz-value changes below
    ``tolerance``.  Like the reference it recomputes H psi and ignores
    ``initial_applied`` / ``initial_eigenvalues`` / ``initial_is_ritz``,
    ``occupied_roots`` and ``full_accuracy``.  For one or two roots the
    reference's first candidate list takes basis columns it has not filled
    (``np.empty`` storage); here they are zero, so such a call ends in a
    ``LinAlgError`` from the singular overlap instead of reading whatever
    memory held."""
    if tolerance <= 0.0 or max_iterations < 1:
        raise ValueError("invalid ParO diagonalization controls")
    del initial_applied, initial_eigenvalues, initial_is_ritz, occupied_roots, full_accuracy
    s = _Blocks(operator, operator_into, diagonal, mpi)
    n = number_of_roots
    v, values, hv, _ = s.initial_subspace(n, initial_vectors,
                                          initial_is_orthonormal=initial_is_orthonormal)
    maxsub = min(s.npw, n + max((n + 1) // 2, 4))
    basis = torch.zeros((maxsub, s.npw), dtype=C128, device=s.device)
    hbasis = torch.zeros_like(basis)
    basis[:n] = v
    hbasis[:n] = hv
    conv = np.zeros(n, dtype=bool)
    for _ in range(max_iterations):
        nconv = int(np.count_nonzero(conv))
        nact = min((maxsub - nconv) // 2, maxsub - n)
        base = max(nconv + nact, n)
        trusted = min(nconv + nact, n)
        if nact <= 0:
            break
        unconv = np.flatnonzero(~conv[:trusted])
        cols = [int(c) for c in unconv]
        energies = list(values[unconv])
        if len(cols) < nact:
            if not len(unconv):
                break
            last = float(values[unconv[-1]])
            for c in range(n, base):
                cols.append(c)
                energies.append(last)
                if len(cols) == nact:
                    break
        cols = cols[:nact]
        ic = s.index(cols)
        corr, hcorr = _paro_bpcg(s, basis[:n], s.take(basis, ic), s.take(hbasis, ic),
                                 np.asarray(energies[:nact], dtype=float), tolerance)
        dim = base + nact
        basis[base:dim] = corr
        hbasis[base:dim] = hcorr
        b, hb = basis[:dim], hbasis[:dim]
        h_small, s_small = s.gram(b, hb), s.gram(b, b)
        with _one_blas_thread():
            w, c = _lowest_generalized_eigh(h_small, s_small, dim)
        new_b, new_hb = s.rotate(b, c), s.rotate(hb, c)
        basis[:dim] = new_b
        hbasis[:dim] = new_hb
        old = values.copy()
        values = np.asarray(w[:n], dtype=float).copy()
        conv[:] = False
        conv[:trusted] = np.abs(values[:trusted] - old[:trusted]) < tolerance
        if np.all(conv):
            break
    v, hv = basis[:n], hbasis[:n]
    nun = int(np.count_nonzero(~conv))
    return DavidsonResult(values, s.result(v, device_output), s.residual_norms(hv, v, values),
                          s.happ / n, nun == 0, nun, s.happ)


# ---------------------------------------------------------------------------
# RMM-DIIS (QE crmmdiagg)
# ---------------------------------------------------------------------------

def _blocked_orthonormalize_pair(s: _Blocks, v, hv, block_size):
    """``_blocked_orthonormalize_pair`` (``diagonalization.py:1487-1525``):
    blocked Gram-Schmidt (two projections per block, then the block's
    canonical orthonormalisation) carrying H psi through the same rotations."""
    n = int(v.shape[0])
    out = torch.empty_like(v)
    hout = torch.empty_like(hv)
    active = 0
    for start in range(0, n, block_size):
        stop = min(start + block_size, n)
        blk, hblk = v[start:stop].clone(), hv[start:stop].clone()
        if active:
            for _ in range(2):
                s.project_out(out[:active], blk, (hout[:active], hblk))
        gram = s.gram(blk, blk)
        gram = 0.5 * (gram + gram.conj().T)
        with _one_blas_thread():
            sq, rot = np.linalg.eigh(gram)
        order = np.argsort(sq)[::-1]
        sq = np.maximum(np.real(sq[order]), 0.0)
        if not len(sq) or sq[0] <= 0.0:
            continue
        keep = sq > 1.0e-24 * sq[0]
        t = rot[:, order][:, keep] / np.sqrt(sq[keep])[None, :]
        r = t.shape[1]
        out[active:active + r] = s.rotate(blk, t)
        hout[active:active + r] = s.rotate(hblk, t)
        active += r
        if active >= n:
            break
    return out[:active], hout[:active]


def rmm_diis(operator, diagonal, number_of_roots, *, initial_vectors, tolerance=1.0e-8,
             max_iterations=20, history_dimension=4, converge=False, gram_schmidt_block=16,
             occupied_roots=None, full_accuracy=False, initial_is_orthonormal=False, mpi=None,
             operator_into=None, initial_applied=None, initial_eigenvalues=None,
             initial_is_ritz=False, device_output=False):
    """QE ``crmmdiagg`` (``diagonalization.py:1528-1756``): ``history_dimension``
    DIIS steps per call (``max_iterations`` is not used, as in the
    reference), each minimising every unconverged band's residual over its
    own history, a Kresse-Furthmueller preconditioned correction and a
    bounded quadratic line search; then blocked Gram-Schmidt and a closing
    Rayleigh-Ritz."""
    del max_iterations
    if history_dimension < 2:
        raise ValueError("diago_rmm_ndim must be at least 2")
    if gram_schmidt_block < 1:
        raise ValueError("diago_gs_nblock must be at least 1")
    s = _Blocks(operator, operator_into, diagonal, mpi)
    n, nd, npw = number_of_roots, history_dimension, s.npw
    v, values, hv, _ = s.initial_subspace(
        n, initial_vectors, initial_is_orthonormal=initial_is_orthonormal,
        initial_applied=initial_applied, initial_eigenvalues=initial_eigenvalues,
        initial_is_ritz=initial_is_ritz)
    tol = _root_tolerances(n, tolerance, occupied_roots, full_accuracy)
    v, hv = v.contiguous(), hv.contiguous()
    vh = torch.empty((nd, n, npw), dtype=C128, device=s.device)
    hh = torch.empty_like(vh)
    eh = np.empty((n, nd))
    res = hv.clone()
    dv.subtract_band_energies(res, v, s.col(values), npw)
    unconv = np.ones(n, dtype=bool)
    t_ry = s.col(2.0 * s.diagonal)
    acc_iter = 0
    for idiis in range(nd):
        active = np.flatnonzero(unconv)
        acc_iter += len(active)
        if not len(active):
            break
        na = len(active)
        ia = s.index(active)
        # history slot idiis of the active bands (gather and scatter by the
        # same map)
        dv.copy_columns(v, vh[idiis], npw, na, src_cols=ia, dst_cols=ia)
        dv.copy_columns(hv, hh[idiis], npw, na, src_cols=ia, dst_cols=ia)
        eh[active, idiis] = values[active]
        cnt = idiis + 1
        if cnt == 1:
            # v, hv already hold the stored history: normalise in place
            va, hva = s.take(v, ia), s.take(hv, ia)
            (nn,) = s.dots((va, va))
            inv = 1.0 / np.sqrt(np.maximum(np.real(nn), 1.0e-300))
            s.scale(va, cfac=inv, dst=va)
            s.scale(hva, cfac=inv, dst=hva)
            ra = hva.clone()
            dv.subtract_band_energies(ra, va, s.col(eh[active, 0]), npw)
        else:
            # the active bands' history, (cnt, na, npw): rows c n + band of
            # the (nd n, npw) history stores
            ih = s.index((np.arange(cnt)[:, None] * n + active[None, :]).reshape(-1))
            hist_v = s.take(vh.view(nd * n, npw), ih).view(cnt, na, npw)
            hist_h = s.take(hh.view(nd * n, npw), ih).view(cnt, na, npw)
            pairs = []
            for i in range(cnt):
                for j in range(cnt):
                    pairs += [(hist_h[i], hist_h[j]), (hist_h[i], hist_v[j]),
                              (hist_v[i], hist_v[j])]
            d = s.dots(*pairs)
            hhm = np.empty((na, cnt, cnt), dtype=complex)
            hpm = np.empty_like(hhm)
            ppm = np.empty_like(hhm)
            k = 0
            for i in range(cnt):
                for j in range(cnt):
                    hhm[:, i, j], hpm[:, i, j], ppm[:, i, j] = d[k], d[k + 1], d[k + 2]
                    k += 3
            coef = np.zeros((na, cnt), dtype=complex)
            with _one_blas_thread():
                for a, band in enumerate(active):
                    e = eh[band, :cnt]
                    hp = hpm[a]
                    rg = (hhm[a] - hp * e[None, :] - e[:, None] * hp.conj().T
                          + e[:, None] * ppm[a] * e[None, :])
                    try:
                        _, c = _lowest_generalized_eigh(rg, ppm[a], 1)
                        coef[a] = c[:, 0]
                    except (np.linalg.LinAlgError, ValueError):
                        coef[a] = 0.0
                        coef[a, -1] = 1.0
            coef_e = coef * eh[active, :cnt]
            cdev = torch.from_numpy(coef).to(s.device)
            cedev = torch.from_numpy(np.ascontiguousarray(coef_e)).to(s.device)
            va = torch.empty((na, npw), dtype=C128, device=s.device)
            hva = torch.empty_like(va)
            ra = torch.empty_like(va)
            dv.history_combine(vh, cnt, ia, cdev, va, npw)
            dv.history_combine(hh, cnt, ia, cdev, hva, npw)
            dv.history_combine(vh, cnt, ia, cedev, ra, npw)
            s.axpby(hva, None, ra, -1.0)                   # H phi c - phi (e c)
        # Kresse-Furthmueller correction and its H
        ekin = np.maximum(dv.column_diagonal_expectations(va, t_ry, npw).cpu().numpy(), 1.0e-14)
        kc = torch.empty_like(va)
        dv.rmm_precondition(ra, None, t_ry, s.col(ekin), kc, npw)
        hkc = s.apply(kc)
        php, khp, khk, psp, ksp, ksk = (np.real(x) for x in s.dots(
            (va, hva), (kc, hva), (kc, hkc), (va, va), (kc, va), (kc, kc)))
        norm_ref = psp + 2.0 * ksp * 0.5 + ksk * 0.25
        e0 = php / psp
        e1 = (php + khp + 0.25 * khk) / norm_ref
        slope = 2.0 * (khp * psp - php * ksp) / (psp * psp)
        curv = (e1 - e0 - 0.5 * slope) / 0.25
        flat = np.abs(curv) > 1.0e-16
        step = np.where(flat, -0.5 * slope / np.where(flat, curv, 1.0),
                        np.where(slope < 0.0, 1.0, 0.05))
        step = np.clip(step, 0.05, 1.0)
        nrm = np.sqrt(psp + 2.0 * ksp * step + ksk * step * step)
        s.axpby(kc, step / nrm, va, 1.0 / nrm)
        s.axpby(hkc, step / nrm, hva, 1.0 / nrm)
        s.put(v, ia, va)
        s.put(hv, ia, hva)
        old = values.copy()
        values = np.real(s.dots((v, hv))[0])
        res = hv.clone()
        dv.subtract_band_energies(res, v, s.col(values), npw)
        unconv &= ~(np.abs(values - old) < tol)
    v, hv = _blocked_orthonormalize_pair(s, v, hv, gram_schmidt_block)
    iterations = acc_iter / max(n, 1)
    if v.shape[0] < n:
        k = int(v.shape[0])
        sq = dv.column_squared_norms(res[:k].contiguous(), npw).cpu().numpy()
        return DavidsonResult(values, s.result(v, device_output), np.sqrt(np.maximum(sq, 0.0)),
                              iterations, False, n - k, s.happ)
    v, hv = v[:n], hv[:n]
    hs = s.gram(v, hv)
    hs = 0.5 * (hs + hs.conj().T)
    with _one_blas_thread():
        w, u = np.linalg.eigh(hs)
    values = np.real(w[:n]).copy()
    u = u[:, :n]
    v, hv = s.rotate(v, u), s.rotate(hv, u)
    return DavidsonResult(values, s.result(v, device_output), s.residual_norms(hv, v, values),
                          iterations, (not converge) or not np.any(unconv),
                          0 if not converge else int(np.count_nonzero(unconv)), s.happ)
