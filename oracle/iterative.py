"""Oracle CG, ParO and RMM-DIIS: serial restatements of the reference's
other iterative eigensolvers (test infrastructure only: the oracle checks the
device solvers, it is never the measured or shipped code).

Follows, as text, ``qepy_pw/diagonalization.py``:

* ``initial_subspace``   -- ``_initial_iterative_subspace`` (932-990): canonical
  orthonormalisation of the trials unless they are orthonormal, then H and a
  Rayleigh-Ritz rotation of the first ``number_of_roots`` columns (or the
  caller's Ritz block as is).
* ``root_tolerances``    -- ``_iterative_root_tolerances`` (883-899).
* ``conjugate_gradient`` -- 993-1186 (QE ``ccgdiagg``): band by band, the
  Ry-kinetic preconditioner P = 1 + T + sqrt(1 + (T - 1)^2), the P-metric
  gradient, Polak-Ribiere conjugation with QE's analytic psi correction,
  and the two-vector line minimisation by the angle theta.
* ``paro_bpcg``          -- ``_paro_bpcg`` (1189-1325, QE ``bpcg_k``): block
  preconditioned CG for the correction equations, step rollback when the
  quadratic functional rises.
* ``parallel_orbital``   -- 1328-1467 (QE ``paro_k_new``): candidate selection,
  correction block, generalized Rayleigh-Ritz of [basis, corrections],
  convergence on Ritz-value changes.
* ``rmm_diis``           -- 1528-1756 (QE ``crmmdiagg``): per-band DIIS on the
  residual Gram of the history, Kresse-Furthmueller kinetic preconditioner
  (``_native_fft.pyx:1346-1408``), bounded quadratic line search, blocked
  Gram-Schmidt carrying H psi (``_blocked_orthonormalize_pair``, 1487-1525)
  and the closing Rayleigh-Ritz.

Pinned by (tests/test_oracle_iterative.py): the reference's own ParO tests
(``tests/unit/core/test_diagonalization.py:8-50``: 18x18 seed 7321 vs
``eigvalsh`` to 1e-9 with residual < 3e-6; more than 2 x roots H
applications for the diagonal 12x12 case), dense ``eigvalsh`` of the small
plane-wave workload for converged CG and ParO runs, the exact-eigenvector
fixed point of all three solvers, RMM-DIIS repeated as the SCF repeats it
converging to ``eigvalsh``, and the argument ``ValueError`` paths.
Iteration and H-application counts beyond the reference's ParO bound are
parity unpinned (the reference's run cannot be made here).
"""

from __future__ import annotations

import numpy as np

from .davidson import OracleResult, lowest_generalized_eigh, orthonormalize, qe_precondition


def root_tolerances(number_of_roots, tolerance, occupied_roots, full_accuracy):
    if occupied_roots is None or full_accuracy:
        occupied_roots = number_of_roots
    if not 0 <= occupied_roots <= number_of_roots:
        raise ValueError("occupied_roots must be between zero and number_of_roots")
    tol = np.full(number_of_roots, max(5.0 * tolerance, 5.0e-6))
    tol[:occupied_roots] = tolerance
    return tol


def _apply(operator, block):
    return np.asarray(operator(np.asfortranarray(block)), dtype=complex)


def initial_subspace(operator, number_of_roots, initial_vectors, *, initial_is_orthonormal,
                     initial_applied=None, initial_eigenvalues=None, initial_is_ritz=False):
    trial = np.asarray(initial_vectors, dtype=complex)
    if trial.ndim != 2 or trial.shape[1] < number_of_roots:
        raise ValueError("initial_vectors must contain at least number_of_roots columns")
    if initial_is_orthonormal:
        v = trial[:, :number_of_roots].copy()
    else:
        v = orthonormalize(trial)[:, :number_of_roots]
    if v.shape[1] < number_of_roots:
        raise ValueError("initial vectors are linearly dependent")
    if initial_is_ritz:
        if initial_applied is None or initial_eigenvalues is None:
            raise ValueError("initial Ritz vectors require H-vectors and eigenvalues")
        hv = np.array(initial_applied, dtype=complex)
        if hv.shape != v.shape:
            raise ValueError("initial_applied must match initial vectors")
        return v, np.asarray(initial_eigenvalues[:number_of_roots], float).copy(), hv, 0
    hv = _apply(operator, v)
    proj = v.conj().T @ hv
    proj = 0.5 * (proj + proj.conj().T)
    w, u = np.linalg.eigh(proj)
    u = u[:, :number_of_roots]
    return v @ u, np.real(w[:number_of_roots]), hv @ u, number_of_roots


def _residual_norms(hv, v, values):
    return np.linalg.norm(hv - v * values[None, :], axis=0)


def conjugate_gradient(operator, diagonal, number_of_roots, *, initial_vectors,
                       tolerance=1.0e-8, max_iterations=20, occupied_roots=None,
                       full_accuracy=False, initial_is_orthonormal=False,
                       initial_applied=None, initial_eigenvalues=None, initial_is_ritz=False):
    if tolerance <= 0.0:
        raise ValueError("diagonalization tolerance must be positive")
    if max_iterations < 1:
        raise ValueError("diago_cg_maxiter must be at least 1")
    v, values, hv, happ = initial_subspace(
        operator, number_of_roots, initial_vectors, initial_is_orthonormal=initial_is_orthonormal,
        initial_applied=initial_applied, initial_eigenvalues=initial_eigenvalues,
        initial_is_ritz=initial_is_ritz)
    tol = root_tolerances(number_of_roots, tolerance, occupied_roots, full_accuracy)
    t_ry = 2.0 * np.asarray(diagonal, dtype=float)
    p = 1.0 + t_ry + np.sqrt(1.0 + (t_ry - 1.0) ** 2)
    unconverged, iteration_sum = 0, 0
    for band in range(number_of_roots):
        psi, hpsi = v[:, band], hv[:, band]        # views: updated in place
        if band:
            c = v[:, :band].conj().T @ psi
            psi -= v[:, :band] @ c
            hpsi -= hv[:, :band] @ c
            nrm = np.sqrt(max(float(np.real(np.vdot(psi, psi))), 1.0e-300))
            psi /= nrm
            hpsi /= nrm
        energy = float(np.real(np.vdot(psi, hpsi)))
        converged = False
        gg_prev = cg_norm_prev = sine_prev = 0.0
        cg = pg_prev = None
        for it in range(1, max_iterations + 1):
            g = hpsi / p
            ppsi = psi / p
            e_p = float(np.real(np.vdot(psi, g) / np.vdot(psi, ppsi)))
            g = g - e_p * ppsi
            if band:
                g = g - v[:, :band] @ (v[:, :band].conj().T @ g)
            if it > 1:
                cross = float(np.real(np.vdot(g, pg_prev)))
            pg_prev = g * p
            gg = float(np.real(np.vdot(g, pg_prev)))
            if not np.isfinite(gg) or gg <= 1.0e-30:
                break
            if it == 1:
                cg = g.copy()
            else:
                gamma = (gg - cross) / max(gg_prev, 1.0e-300)
                cg = gamma * cg + g - (gamma * cg_norm_prev * sine_prev) * psi
            gg_prev = gg
            hcg = _apply(operator, cg[:, None])[:, 0]
            happ += 1
            cg2 = float(np.real(np.vdot(cg, cg)))
            if cg2 <= 1.0e-30:
                break
            cg_norm = np.sqrt(cg2)
            a0 = 2.0 * float(np.real(np.vdot(psi, hcg))) / cg_norm
            b0 = float(np.real(np.vdot(cg, hcg))) / cg2
            den = energy - b0
            theta = 0.5 * np.arctan(a0 / den if den != 0.0 else np.copysign(np.inf, a0))
            co, si = float(np.cos(theta)), float(np.sin(theta))
            cos2, sin2 = co * co - si * si, 2.0 * co * si
            roots = (0.5 * ((energy - b0) * cos2 + a0 * sin2 + energy + b0),
                     0.5 * (-(energy - b0) * cos2 - a0 * sin2 + energy + b0))
            if roots[1] < roots[0]:
                theta += 0.5 * np.pi
                co, si = float(np.cos(theta)), float(np.sin(theta))
            old = energy
            energy = float(min(roots))
            psi *= co
            psi += (si / cg_norm) * cg
            hpsi *= co
            hpsi += (si / cg_norm) * hcg
            cg_norm_prev, sine_prev = cg_norm, si
            if abs(energy - old) < tol[band]:
                converged = True
                break
        unconverged += 0 if converged else 1
        iteration_sum += it + 1
        values[band] = energy
    order = np.argsort(values)
    values, v, hv = values[order], v[:, order], hv[:, order]
    return OracleResult(values, v, _residual_norms(hv, v, values),
                        max(1, int(round(iteration_sum / number_of_roots))), unconverged == 0,
                        unconverged, happ)


def _guard(x):
    """x where |x| > 1e-30, else +-1e-30 (the sign of x, + for 0)."""
    return np.where(np.abs(x) > 1.0e-30, x, np.copysign(1.0e-30, x + (x == 0.0)))


def _col_dot(a, b):
    return np.real(np.einsum("gb,gb->b", a.conj(), b))


def paro_bpcg(operator, diagonal, reference, x0, hx0, energies, tolerance, max_iterations=5):
    count = x0.shape[1]
    corr = np.zeros_like(x0)
    hcorr = np.zeros_like(x0)
    if count == 0:
        return corr, hcorr, 0
    rhs = x0 * energies[None, :] - hx0
    prec = qe_precondition(rhs, energies, diagonal)
    prec = prec - reference @ (reference.conj().T @ prec)
    g0 = _col_dot(prec, rhs)
    thresholds = np.maximum(0.01 * tolerance, 0.01 * g0)
    dirs = prec.copy()
    f_prev = np.zeros(count)
    active = np.arange(count)
    happ = 0
    for it in range(1, max_iterations + 1):
        if not len(active):
            break
        d = dirs[:, active]
        hd = _apply(operator, d)
        happ += len(active)
        e = energies[active]
        gamma = _col_dot(d, hd) - e * _col_dot(d, d)
        alpha = g0[active] / _guard(gamma)
        corr[:, active] += d * alpha[None, :]
        hcorr[:, active] += hd * alpha[None, :]
        grad = rhs[:, active] + corr[:, active] * e[None, :] - hcorr[:, active]
        g2 = _col_dot(prec[:, active], grad)
        newp = qe_precondition(grad, e, diagonal)
        newp = newp - reference @ (reference.conj().T @ newp)
        g1 = _col_dot(newp, grad)
        fval = (0.5 * (_col_dot(corr[:, active], hcorr[:, active])
                       - e * _col_dot(corr[:, active], corr[:, active]))
                - _col_dot(corr[:, active], rhs[:, active]))
        rose = fval > f_prev[active]
        back = rose & (f_prev[active] < 0.0)
        if np.any(back):
            failed = active[back]
            corr[:, failed] -= dirs[:, failed] * alpha[back][None, :]
            hcorr[:, failed] -= hd[:, back] * alpha[back][None, :]
        done = (np.abs(g1) < thresholds[active]) | rose | (it == max_iterations)
        keep = ~done
        if not np.any(keep):
            break
        cont = active[keep]
        beta = (g1[keep] - g2[keep]) / _guard(g0[cont])
        dirs[:, cont] = newp[:, keep] + dirs[:, cont] * beta[None, :]
        prec[:, cont] = newp[:, keep]
        g0[cont] = g1[keep]
        f_prev[cont] = fval[keep]
        active = cont
    return corr, hcorr, happ


def parallel_orbital(operator, diagonal, number_of_roots, *, initial_vectors,
                     tolerance=1.0e-8, max_iterations=20, initial_is_orthonormal=False):
    if tolerance <= 0.0 or max_iterations < 1:
        raise ValueError("invalid ParO diagonalization controls")
    v, values, hv, happ = initial_subspace(operator, number_of_roots, initial_vectors,
                                           initial_is_orthonormal=initial_is_orthonormal)
    n = number_of_roots
    maxsub = min(v.shape[0], n + max((n + 1) // 2, 4))
    basis = np.empty((v.shape[0], maxsub), dtype=complex)
    hbasis = np.empty_like(basis)
    basis[:, :n], hbasis[:, :n] = v, hv
    conv = np.zeros(n, dtype=bool)
    diagonal = np.asarray(diagonal, dtype=float)
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
        energies = np.asarray(energies[:nact], dtype=float)
        corr, hcorr, a = paro_bpcg(operator, diagonal, basis[:, :n], basis[:, cols],
                                   hbasis[:, cols], energies, tolerance)
        happ += a
        dim = base + nact
        basis[:, base:dim], hbasis[:, base:dim] = corr, hcorr
        b, hb = basis[:, :dim], hbasis[:, :dim]
        w, c = lowest_generalized_eigh(b.conj().T @ hb, b.conj().T @ b, dim)
        basis[:, :dim], hbasis[:, :dim] = b @ c, hb @ c
        old = values.copy()
        values = np.asarray(w[:n], dtype=float).copy()
        conv[:] = False
        conv[:trusted] = np.abs(values[:trusted] - old[:trusted]) < tolerance
        if np.all(conv):
            break
    v, hv = basis[:, :n].copy(), hbasis[:, :n].copy()
    nun = int(np.count_nonzero(~conv))
    return OracleResult(values, v, _residual_norms(hv, v, values), happ / n, nun == 0, nun, happ)


def kresse_furthmueller(residuals, kinetic_ry, expectations):
    """K(x) r with x = T / (1.5 <T>), K = -4/(3 <T>) (27 + 18x + 12x^2 + 8x^3)
    / (27 + 18x + 12x^2 + 8x^3 + 16x^4) (``_native_fft.pyx:1346-1408``)."""
    x = kinetic_ry[:, None] / (1.5 * expectations[None, :])
    num = 27.0 + x * (18.0 + x * (12.0 + 8.0 * x))
    den = num + 16.0 * x ** 4
    return (-4.0 / 3.0) * num / (expectations[None, :] * den) * residuals


def blocked_orthonormalize_pair(v, hv, block_size):
    out = np.empty_like(v)
    hout = np.empty_like(hv)
    active = 0
    for start in range(0, v.shape[1], block_size):
        stop = min(start + block_size, v.shape[1])
        blk, hblk = v[:, start:stop].copy(), hv[:, start:stop].copy()
        if active:
            for _ in range(2):
                proj = out[:, :active].conj().T @ blk
                blk -= out[:, :active] @ proj
                hblk -= hout[:, :active] @ proj
        gram = blk.conj().T @ blk
        gram = 0.5 * (gram + gram.conj().T)
        sq, rot = np.linalg.eigh(gram)
        order = np.argsort(sq)[::-1]
        sq = np.maximum(np.real(sq[order]), 0.0)
        if not len(sq) or sq[0] <= 0.0:
            continue
        keep = sq > 1.0e-24 * sq[0]
        t = rot[:, order][:, keep] / np.sqrt(sq[keep])[None, :]
        r = t.shape[1]
        out[:, active:active + r] = blk @ t
        hout[:, active:active + r] = hblk @ t
        active += r
        if active >= v.shape[1]:
            break
    return out[:, :active], hout[:, :active]


def rmm_diis(operator, diagonal, number_of_roots, *, initial_vectors, tolerance=1.0e-8,
             max_iterations=20, history_dimension=4, converge=False, gram_schmidt_block=16,
             occupied_roots=None, full_accuracy=False, initial_is_orthonormal=False,
             initial_applied=None, initial_eigenvalues=None, initial_is_ritz=False):
    del max_iterations                       # crmmdiagg runs history_dimension steps
    if history_dimension < 2:
        raise ValueError("diago_rmm_ndim must be at least 2")
    if gram_schmidt_block < 1:
        raise ValueError("diago_gs_nblock must be at least 1")
    v, values, hv, happ = initial_subspace(
        operator, number_of_roots, initial_vectors, initial_is_orthonormal=initial_is_orthonormal,
        initial_applied=initial_applied, initial_eigenvalues=initial_eigenvalues,
        initial_is_ritz=initial_is_ritz)
    tol = root_tolerances(number_of_roots, tolerance, occupied_roots, full_accuracy)
    npw, n, nd = v.shape[0], number_of_roots, history_dimension
    vh = np.empty((npw, n, nd), dtype=complex)
    hh_hist = np.empty_like(vh)
    eh = np.empty((n, nd))
    res = hv - v * values[None, :]
    unconv = np.ones(n, dtype=bool)
    t_ry = 2.0 * np.asarray(diagonal, dtype=float)
    acc_iter = 0
    for idiis in range(nd):
        active = np.flatnonzero(unconv)
        acc_iter += len(active)
        if not len(active):
            break
        vh[:, active, idiis] = v[:, active]
        hh_hist[:, active, idiis] = hv[:, active]
        eh[active, idiis] = values[active]
        for band in active:
            cnt = idiis + 1
            phi, hphi, e = vh[:, band, :cnt], hh_hist[:, band, :cnt], eh[band, :cnt]
            if cnt == 1:
                nrm = np.sqrt(max(float(np.real(np.vdot(phi[:, 0], phi[:, 0]))), 1.0e-300))
                v[:, band] = phi[:, 0] / nrm
                hv[:, band] = hphi[:, 0] / nrm
                res[:, band] = (hphi[:, 0] - e[0] * phi[:, 0]) / nrm
                continue
            hh = hphi.conj().T @ hphi
            hp = hphi.conj().T @ phi
            pp = phi.conj().T @ phi
            rg = hh - hp * e[None, :] - e[:, None] * hp.conj().T + e[:, None] * pp * e[None, :]
            try:
                _, c = lowest_generalized_eigh(rg, pp, 1)
                c = c[:, 0]
            except (np.linalg.LinAlgError, ValueError):
                c = np.zeros(cnt, dtype=complex)
                c[-1] = 1.0
            v[:, band] = phi @ c
            hv[:, band] = hphi @ c
            res[:, band] = hv[:, band] - phi @ (e * c)
        ekin = np.maximum(np.einsum("g,gb->b", t_ry, np.abs(v[:, active]) ** 2), 1.0e-14)
        k = kresse_furthmueller(res[:, active], t_ry, ekin)
        hk = _apply(operator, k)
        happ += len(active)
        for j, band in enumerate(active):
            psi, hpsi, kp, hkp = v[:, band], hv[:, band], k[:, j], hk[:, j]
            php, khp, khk, psp, ksp, ksk = (float(np.real(np.vdot(a, b))) for a, b in (
                (psi, hpsi), (kp, hpsi), (kp, hkp), (psi, psi), (kp, psi), (kp, kp)))
            norm_ref = psp + ksp + 0.25 * ksk
            e0 = php / psp
            e1 = (php + khp + 0.25 * khk) / norm_ref
            slope = 2.0 * (khp * psp - php * ksp) / (psp * psp)
            curv = (e1 - e0 - 0.5 * slope) / 0.25
            step = -0.5 * slope / curv if abs(curv) > 1.0e-16 else (1.0 if slope < 0.0 else 0.05)
            step = float(np.clip(step, 0.05, 1.0))
            nrm = np.sqrt(psp + 2.0 * ksp * step + ksk * step * step)
            psi *= 1.0 / nrm
            psi += (step / nrm) * kp
            hpsi *= 1.0 / nrm
            hpsi += (step / nrm) * hkp
        old = values.copy()
        values = _col_dot(v, hv)
        res = hv - v * values[None, :]
        unconv &= ~(np.abs(values - old) < tol)
    v, hv = blocked_orthonormalize_pair(v, hv, gram_schmidt_block)
    if v.shape[1] < n:
        return OracleResult(values, v, np.linalg.norm(res[:, :v.shape[1]], axis=0),
                            acc_iter / max(n, 1), False, n - v.shape[1], happ)
    v, hv = v[:, :n], hv[:, :n]
    hs = v.conj().T @ hv
    hs = 0.5 * (hs + hs.conj().T)
    w, u = np.linalg.eigh(hs)
    values = np.real(w[:n])
    v, hv = v @ u[:, :n], hv @ u[:, :n]
    return OracleResult(values, v, _residual_norms(hv, v, values), acc_iter / max(n, 1),
                        (not converge) or not np.any(unconv),
                        0 if not converge else int(np.count_nonzero(unconv)), happ)
