"""Per-k projector assembly on the device (QE ``init_us_2``).

The SCF rebuilds every k-point's nonlocal projectors each iteration
(``qepy_pw/pw/scf.py:3383-3395``): per species, its radial tables are
interpolated at |k+G| and multiplied by the real harmonics and (-i)^l
(``UPFPseudo.projector_basis``, ``qepy_pw/upf.py:483-541``, with
``_qe_cubic_interpolate`` ``upf.py:880-912`` and
``_qe_real_spherical_harmonics`` ``upf.py:722-772``), then every atom's
block gets its structure phase exp(-i (k+G).tau) and the blocks are packed
(``pw/scf.py:657-686``).  Here both steps run on the device
(``qepw_projector_basis``, ``qepw_build_projectors``), so the packed
projectors never cross PCIe (the medium workload's are 98 MB per k-point).

The radial tables themselves (spherical-Bessel transforms of the UPF
projectors on the uniform grid i * dq) stay host set-up, computed once per
species as the reference does (``upf.py:270-326``).
"""

from __future__ import annotations

import numpy as np
import torch

from . import _native as nat
from . import device as dv
from .diagonalization import FactorizedProjectorTerm, PackedProjectorTerm


def _device_rows(array, device, dtype=torch.float64) -> torch.Tensor:
    if isinstance(array, torch.Tensor):
        return array.to(device, dtype).contiguous()
    return torch.from_numpy(np.ascontiguousarray(array, dtype=np.float64)).to(device, dtype)


def projector_basis(gk, table, dq: float, channels, device=None) -> torch.Tensor:
    """Species projectors beta (npw, nch) on the device, complex128, rows in
    the order of ``gk`` (npw, 3) Cartesian k+G (bohr^-1).  ``table`` (nq,
    nradial): radial transforms on the grid i * dq; ``channels``: [(radial
    index, l <= 3, QE ylmr2 index m), ...].  The returned tensor is the
    transpose of a column-major (nch, npw) buffer, i.e. laid out as the
    reference's F-ordered ``beta_matrix``."""
    dev = dv.require_device(device)
    q = _device_rows(gk, dev)
    if q.dim() != 2 or q.shape[1] != 3:
        raise ValueError("gk must have shape (npw, 3)")
    tab = np.asarray(table.cpu() if isinstance(table, torch.Tensor) else table, dtype=np.float64)
    if tab.ndim == 1:
        tab = tab[:, None]
    nq, nrad = tab.shape
    chans = np.asarray(channels, dtype=np.int32).reshape(-1, 3)
    if len(chans) and (np.any(chans[:, 0] < 0) or np.any(chans[:, 0] >= nrad)
                       or np.any(chans[:, 1] < 0) or np.any(chans[:, 1] > 3)
                       or np.any(chans[:, 2] < 0) or np.any(chans[:, 2] > 2 * chans[:, 1])):
        raise ValueError("channels must be (radial index, l <= 3, 0 <= m <= 2l)")
    npw = int(q.shape[0])
    # the table must cover every |k+G| (the kernel's interpolation reads four
    # points from floor(|q| / dq)): checked on the host copy of the vectors
    host_q = (gk.detach().to("cpu", torch.float64).numpy() if isinstance(gk, torch.Tensor)
              else np.asarray(gk, dtype=np.float64))
    qmax = float(np.sqrt(np.max(np.sum(host_q * host_q, axis=1)))) if npw else 0.0
    if npw and int(np.floor(qmax / dq)) + 3 >= nq:
        raise ValueError(f"the radial table (q up to {(nq - 1) * dq:.4g}) does not cover "
                         f"|k+G| = {qmax:.4g} with the four-point interpolation")
    out = torch.empty((len(chans), npw), dtype=torch.complex128, device=dev)
    if npw and len(chans):
        # column-major table: torch row p holds radial transform p
        tab_dev = torch.from_numpy(np.ascontiguousarray(tab.T)).to(dev)
        ch = torch.from_numpy(np.ascontiguousarray(chans.T)).to(dev)
        rc = nat.load().qepw_projector_basis(
            q.data_ptr(), npw, tab_dev.data_ptr(), nq, nq, float(dq), ch[0].data_ptr(),
            ch[1].data_ptr(), ch[2].data_ptr(), len(chans), out.data_ptr(), npw,
            dv.stream_handle(dev))
        nat.check(rc, "qepw_projector_basis")
    return out.T


def expanded_coupling(channels, dij_ry) -> np.ndarray:
    """The species' (nch, nch) coupling in Hartree for ``channels`` [(radial
    index, l, m), ...]: D_ij = dij_ry[p_i, p_j] / 2 where the channels share
    l and m, else 0 (``UPFPseudo.projector_basis``, ``upf.py:542-555``)."""
    chans = np.asarray(channels, dtype=np.int64).reshape(-1, 3)
    dij = np.asarray(dij_ry, dtype=np.float64)
    same = (chans[:, None, 1] == chans[None, :, 1]) & (chans[:, None, 2] == chans[None, :, 2])
    return np.where(same, 0.5 * dij[np.ix_(chans[:, 0], chans[:, 0])], 0.0)


def packed_projector_term(beta_species: torch.Tensor, couplings, positions, gk,
                          device=None) -> PackedProjectorTerm:
    """The atom-packed ``PackedProjectorTerm`` of one species on the device:
    beta[:, a*nch + i] = beta_species[:, i] exp(-i (k+G).tau_a) (a device
    tensor, rows in the order of ``gk``), the block-diagonal coupling of the
    per-atom ``couplings`` (each (nch, nch), Ha) and the preconditioner's
    diagonal sum_a sum_ij beta_gi D_a,ij conj(beta_gj) (host, as the
    reference's packed term carries it)."""
    dev = dv.require_device(device)
    bs = beta_species.to(dev, torch.complex128)
    npw, nch = int(bs.shape[0]), int(bs.shape[1])
    tau = _device_rows(positions, dev)
    natom = int(tau.shape[0])
    couplings = [np.asarray(c, dtype=np.complex128) for c in couplings]
    if len(couplings) != natom or any(c.shape != (nch, nch) for c in couplings):
        raise ValueError("one (nch, nch) coupling per atom is required")
    q = _device_rows(gk, dev)
    src = bs.T.contiguous()                     # column-major (npw x nch)
    packed = torch.empty((natom * nch, npw), dtype=torch.complex128, device=dev)
    if npw and natom and nch:
        rc = nat.load().qepw_build_projectors(
            nat.C128, src.data_ptr(), dv.ld_of(src), nch, q.data_ptr(), tau.data_ptr(), natom,
            packed.data_ptr(), dv.ld_of(packed), npw, dv.stream_handle(dev))
        nat.check(rc, "qepw_build_projectors")
    coupling = np.zeros((natom * nch, natom * nch), dtype=np.complex128)
    for a, c in enumerate(couplings):
        coupling[a * nch:(a + 1) * nch, a * nch:(a + 1) * nch] = c
    # the phases cancel in the diagonal: beta_s (sum_a D_a) beta_s^H
    diag = torch.zeros(npw, dtype=torch.float64, device=dev)
    if npw and nch and natom:
        dsum = torch.from_numpy(np.ascontiguousarray(sum(couplings).T)).to(dev)
        coupled = torch.empty_like(src)
        ws = dv.Workspace(dev)
        dv.gemm(nat.OP_N, npw, nch, nch, 1.0, src, dsum, 0.0, coupled, ws)
        dv.accumulate_projector_diagonal(coupled, src, npw, 1.0, diag)
    return PackedProjectorTerm(packed.T, coupling, diag.cpu().numpy())


def factorized_projector_term(beta_species: torch.Tensor, coupling, positions, gk,
                              device=None) -> FactorizedProjectorTerm:
    """The reference's ``FactorizedProjectorTerm`` (species beta, atom phase
    columns, one shared coupling; ``pw/scf.py:692-707``, used when a species
    has more than 64 packed channels) with device tensors: beta (npw, nch) as
    given, phases[g, a] = exp(-i (k+G)_g . tau_a) computed on the device
    (``qepw_build_projectors`` of a unit column), rows in the order of
    ``gk``.  ``PlaneWaveHamiltonian`` packs its atom blocks on the device
    (``qepw_phase_blocks``)."""
    dev = dv.require_device(device)
    bs = beta_species.to(dev, torch.complex128)
    npw = int(bs.shape[0])
    coupling = np.asarray(coupling, dtype=np.complex128)
    if coupling.shape != (bs.shape[1], bs.shape[1]):
        raise ValueError("coupling must be (nch, nch)")
    tau = _device_rows(positions, dev)
    natom = int(tau.shape[0])
    q = _device_rows(gk, dev)
    ones = torch.ones((1, npw), dtype=torch.complex128, device=dev)
    phases = torch.empty((natom, npw), dtype=torch.complex128, device=dev)
    if npw and natom:
        rc = nat.load().qepw_build_projectors(
            nat.C128, ones.data_ptr(), npw, 1, q.data_ptr(), tau.data_ptr(), natom,
            phases.data_ptr(), npw, npw, dv.stream_handle(dev))
        nat.check(rc, "qepw_build_projectors")
    return FactorizedProjectorTerm(bs, phases.T, coupling)
