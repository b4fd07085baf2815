"""Hand-written backward of the band energy on the GPU.

E(psi, V, tau) = sum_n f_n <psi_n | H | psi_n>,  H = T + V_loc(r) + sum_a B_a D_a B_a^H,
B_a = beta(k+G) exp(-i (k+G).tau_a)

with the gradients

* dE/dV(r)    = (1/N) sum_n f_n |psi_n(r)|^2   (psi(r) the unnormalised
  backward FFT; the density of ``_native_fft.pyx:2592-2648`` divided by N),
* dE/dpsi*    = f_n H psi_n   (returned as torch's 2 dE/dpsi*),
* dE/dtau_a,c = 2 sum_n f_n Re sum_{i in a} conj(dO_c[i,n]) (D O)[i,n],
  O = B^H psi, dO_c = i B^H ((k+G)_c psi) -- the force branch of
  ``_nonlocal_derivatives`` (``qepy_pw/pw/scf.py:1521-1631``) with the
  projectors recomputed from tau on the device instead of storing dB/dtau.

Everything runs through ``libqepw_hip.so`` (``qepw_apply_local``,
``qepw_gemm``, ``qepw_accumulate_density``, ``qepw_build_projectors``,
``qepw_scale_rows``, ``qepw_projector_forces``); ``band_energy`` wraps the
forward and backward in a ``torch.autograd.Function``.  Both complex128 and
complex64 storage are supported (FP64 accumulation of every reduction).
"""

from __future__ import annotations

import numpy as np
import torch

from . import _native as nat
from . import device as dv


class BandEnergyModel:
    """Device operands of E for one k-point: FFT plan, kinetic, k+G vectors,
    the species projectors beta (npw, nch) and the per-atom couplings D_a."""

    def __init__(self, basis, shape, beta_species, couplings, *, device=None):
        self.device = dv.require_device(device)
        self.shape = tuple(int(s) for s in shape)
        self.plan = dv.DeviceFFTPlan(np.asarray(basis.indices), self.shape, self.device)
        self.npw = self.plan.npw
        perm = self.plan.perm
        self.workspace = dv.Workspace(self.device)
        self.kinetic = torch.from_numpy(np.ascontiguousarray(
            np.asarray(basis.kinetic, float)[perm])).to(self.device)
        self.gk = torch.from_numpy(np.ascontiguousarray(
            np.asarray(basis.vectors, float)[perm])).to(self.device)
        beta = np.asarray(beta_species, dtype=np.complex128)
        self.nch = int(beta.shape[1])
        self.beta = torch.from_numpy(np.ascontiguousarray(beta[perm].T)).to(self.device)
        couplings = [np.asarray(c, dtype=float) for c in couplings]
        self.natom = len(couplings)
        nproj = self.natom * self.nch
        d_host = np.zeros((nproj, nproj), dtype=np.complex128)
        for a, c in enumerate(couplings):
            d_host[a * self.nch:(a + 1) * self.nch, a * self.nch:(a + 1) * self.nch] = c
        self.coupling = torch.from_numpy(np.ascontiguousarray(d_host.T)).to(self.device)
        self.offsets = torch.arange(0, nproj + 1, self.nch, dtype=torch.int32, device=self.device)
        self._cast = {}

    def operands(self, dtype):
        if dtype == torch.complex128:
            return self.beta, self.coupling, self.kinetic
        if dtype not in self._cast:
            self._cast[dtype] = (self.beta.to(dtype), self.coupling.to(dtype),
                                 self.kinetic.to(torch.float32))
        return self._cast[dtype]

    def projectors(self, tau: torch.Tensor, dtype) -> torch.Tensor:
        """(P, npw) device-order packed projector B(tau)."""
        beta, _, _ = self.operands(dtype)
        out = torch.empty((self.natom * self.nch, self.npw), dtype=dtype, device=self.device)
        rc = nat.load().qepw_build_projectors(
            dv.precision_code(dtype), beta.data_ptr(), dv.ld_of(beta), self.nch,
            self.gk.data_ptr(), tau.data_ptr(), self.natom, out.data_ptr(), dv.ld_of(out),
            self.npw, dv.stream_handle(self.device))
        nat.check(rc, "qepw_build_projectors")
        return out

    def apply(self, psi, potential, proj):
        """H psi, O = B^H psi and D O for (nb, npw) device-order psi."""
        dtype = psi.dtype
        _, coupling, kinetic = self.operands(dtype)
        nb = psi.shape[0]
        p = proj.shape[0]
        hpsi = torch.empty_like(psi)
        self.plan.apply_local(psi, hpsi, potential, kinetic, self.workspace)
        overlap = torch.empty((nb, p), dtype=dtype, device=self.device)
        dv.gemm(nat.OP_C, p, nb, self.npw, 1.0, proj, psi, 0.0, overlap, self.workspace)
        coupled = torch.empty_like(overlap)
        dv.gemm(nat.OP_N, p, nb, p, 1.0, coupling, overlap, 0.0, coupled, self.workspace)
        dv.gemm(nat.OP_N, self.npw, nb, p, 1.0, proj, coupled, 1.0, hpsi, self.workspace)
        return hpsi, overlap, coupled

    def scale_rows(self, src, dst, s=None, sstride=0, cfac=None):
        rc = nat.load().qepw_scale_rows(
            dv.precision_code(src.dtype), src.data_ptr(), dv.ld_of(src),
            None if s is None else s.data_ptr(), sstride,
            None if cfac is None else cfac.data_ptr(), dst.data_ptr(), dv.ld_of(dst),
            self.npw, src.shape[0], dv.stream_handle(self.device))
        nat.check(rc, "qepw_scale_rows")

    def density(self, psi, weights, scale):
        """scale * sum_b weights[b] |psi_b(r)|^2, (n3, n2, n1) float64."""
        lib = nat.load()
        prec = dv.precision_code(psi.dtype)
        nb = psi.shape[0]
        n1, n2, n3 = self.shape
        out = torch.empty((n3, n2, n1), dtype=torch.float64, device=self.device)
        need = lib.qepw_accumulate_density_workspace(self.plan.handle, prec, nb)
        buf = self.workspace.get(need)
        rc = lib.qepw_accumulate_density(self.plan.handle, prec, psi.data_ptr(), dv.ld_of(psi),
                                         nb, weights.data_ptr(), float(scale), out.data_ptr(),
                                         buf.data_ptr(), buf.numel(), dv.stream_handle(self.device))
        nat.check(rc, "qepw_accumulate_density")
        return out

    def position_gradient(self, psi, proj, coupled, weights):
        """dE/dtau (natom, 3) float64."""
        nb = psi.shape[0]
        p = proj.shape[0]
        scaled = torch.empty_like(psi)
        d_overlap = torch.empty((3, nb, p), dtype=psi.dtype, device=self.device)
        for c in range(3):
            self.scale_rows(psi, scaled, s=self.gk[:, c], sstride=3)
            dv.gemm(nat.OP_C, p, nb, self.npw, 1j, proj, scaled, 0.0, d_overlap[c],
                    self.workspace)
        out = torch.empty((self.natom, 3), dtype=torch.float64, device=self.device)
        rc = nat.load().qepw_projector_forces(
            dv.precision_code(psi.dtype), d_overlap.data_ptr(), p, nb * p,
            coupled.data_ptr(), dv.ld_of(coupled), weights.data_ptr(), nb,
            self.offsets.data_ptr(), self.natom, out.data_ptr(), dv.stream_handle(self.device))
        nat.check(rc, "qepw_projector_forces")
        return out


def _to_plane_major(potential: torch.Tensor, dtype) -> torch.Tensor:
    # (n1, n2, n3) -> (n3, n2, n1): a layout copy, no arithmetic
    return potential.detach().permute(2, 1, 0).contiguous().to(dtype)


class _BandEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, psi, potential, positions, model, occupations, dtype):
        device = model.device
        real = torch.float64 if dtype == torch.complex128 else torch.float32
        psi_ref = psi.detach().to(device)
        psi_dev = torch.empty((psi_ref.shape[1], model.npw), dtype=dtype, device=device)
        src = psi_ref.T.contiguous()
        if src.dtype != dtype and src.dtype not in (torch.complex128, torch.complex64):
            src = src.to(torch.complex128)
        dv.permute_rows(0, model.plan.perm_device, src, psi_dev, model.npw)
        v_dev = _to_plane_major(potential.to(device), real)
        tau = positions.detach().to(device=device, dtype=torch.float64).contiguous()
        proj = model.projectors(tau, dtype)
        hpsi, overlap, coupled = model.apply(psi_dev, v_dev, proj)
        f = np.asarray(occupations, dtype=np.float64).reshape(-1)
        ip = dv.column_inner_products(psi_dev, hpsi, model.npw)
        energy = float(np.dot(f, ip.cpu().numpy()[0::2]))
        ctx.model = model
        ctx.save_for_backward(psi_dev, hpsi, proj, coupled)
        ctx.occupations = f
        ctx.psi_dtype = psi.dtype
        return torch.tensor(energy, dtype=torch.float64, device=device)

    @staticmethod
    def backward(ctx, grad_energy):
        model = ctx.model
        psi_dev, hpsi, proj, coupled = ctx.saved_tensors
        g = float(grad_energy)
        # the per-band weights f_n dE are host scalars, uploaded once
        wf_host = ctx.occupations * g
        wf = dv.upload(wf_host, model.device)
        grad_psi = grad_v = grad_tau = None
        if ctx.needs_input_grad[0]:
            scaled = torch.empty_like(hpsi)
            model.scale_rows(hpsi, scaled, cfac=dv.upload(2.0 * wf_host, model.device))
            ref = torch.empty_like(scaled)
            dv.permute_rows(1, model.plan.perm_device, scaled, ref, model.npw)
            grad_psi = ref.T.to(ctx.psi_dtype)
        if ctx.needs_input_grad[1]:
            n = float(np.prod(model.shape))
            dens = model.density(psi_dev, wf, 1.0 / n)   # (n3, n2, n1)
            grad_v = dens.permute(2, 1, 0).contiguous()
        if ctx.needs_input_grad[2]:
            grad_tau = model.position_gradient(psi_dev, proj, coupled, wf)
        return grad_psi, grad_v, grad_tau, None, None, None


def band_energy(model: BandEnergyModel, psi: torch.Tensor, potential: torch.Tensor,
                positions: torch.Tensor, occupations, precision: str = "fp64") -> torch.Tensor:
    """E = sum_n f_n <psi_n|H|psi_n> (Hartree) as a differentiable scalar.

    psi: (npw, nb) complex, reference row order; potential: (n1, n2, n3)
    real V(r) in Hartree; positions: (natom, 3) Cartesian bohr; occupations
    f_n.  precision "fp64" (complex128) or "fp32" (complex64 storage and
    arithmetic, FP64 reductions)."""
    if precision not in ("fp64", "fp32"):
        raise ValueError("precision must be 'fp64' or 'fp32'")
    dtype = torch.complex128 if precision == "fp64" else torch.complex64
    if psi.dim() != 2 or psi.shape[0] != model.npw:
        raise ValueError("psi must have shape (npw, nb)")
    if tuple(potential.shape) != model.shape:
        raise ValueError("potential must have the FFT grid shape")
    if tuple(positions.shape) != (model.natom, 3):
        raise ValueError("positions must have shape (natom, 3)")
    if len(occupations) != psi.shape[1]:
        raise ValueError("one occupation per band is required")
    return _BandEnergy.apply(psi, potential, positions, model, occupations, dtype)
