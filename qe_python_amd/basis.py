"""Plane-wave basis and FFT-grid construction (host side, set-up only).

These routines build the inputs of the hot path: the QE-ordered G list of a
k-point, its kinetic diagonal and the FFT grid shape.  They follow the
reference's ``qepy_pw/basis.py`` so that a basis built here has the same row
order, and therefore the same Davidson storage layout, as the reference's:

* ``next_fast_len``       -> ``basis.py:27-38``  (``_next_fast_len_real``)
* ``qe_energy_order``     -> ``basis.py:1588-1616`` (``_qe_energy_order``)
* ``make_bases``          -> ``basis.py:1483-1577`` (with ``_basis_bound``
  1456-1470 and ``_miller_grid`` 1473-1478)
* ``fft_shape``           -> ``basis.py:1619-1682``
* ``potential_matrix``    -> ``basis.py:1685-1688``
* ``LocalPotentialWorkspace`` -> ``basis.py:457-555, 1210-1282`` (the device FFT
  plan of one k-point and the SCF's density accumulation)

Apart from ``LocalPotentialWorkspace`` (which holds device state and imports
the device modules when it is built), nothing here touches the GPU or runs per
Hamiltonian application.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def next_fast_len(target: int) -> int:
    """Smallest 2,3,5-smooth integer >= target (``basis.py:27-38``)."""
    candidate = max(1, int(target))
    while True:
        rest = candidate
        for factor in (2, 3, 5):
            while rest % factor == 0:
                rest //= factor
        if rest == 1:
            return candidate
        candidate += 1


def qe_energy_order(values: np.ndarray, tolerance: float = 1.0e-8) -> np.ndarray:
    """QE ``hpsort_eps`` order: sort by value, runs closer than ``tolerance``
    keep their incoming (index) order (``basis.py:1588-1616``)."""
    array = np.asarray(values, dtype=float)
    order = np.argsort(array, kind="stable")
    if len(order) < 2 or np.all(np.diff(array[order]) >= tolerance):
        return order
    start = 0
    n = len(order)
    while start < n:
        stop = start + 1
        ref = array[order[start]]
        while stop < n and abs(array[order[stop]] - ref) < tolerance:
            stop += 1
        if stop - start > 1:
            order[start:stop] = np.sort(order[start:stop])
        start = stop
    return order


@dataclass(frozen=True)
class PlaneWaveBasis:
    """Materialised per-k basis: Miller indices, Cartesian k+G, kinetic.

    Mirrors the attributes of the reference ``PlaneWaveBasis``
    (``basis.py:106-199``) that the hot path reads: ``indices`` (npw, 3)
    int32, ``vectors`` (npw, 3) bohr^-1 and ``kinetic`` = 0.5|k+G|^2 (Ha).
    """

    indices: np.ndarray
    vectors: np.ndarray
    kinetic: np.ndarray
    k_crystal: np.ndarray | None = None

    def __len__(self) -> int:
        return len(self.indices)

    def materialize(self) -> "PlaneWaveBasis":
        return self


def _basis_bound(reciprocal, k_crystal, ecut_ha, minimum_stretch) -> int:
    return int(np.ceil((np.sqrt(2.0 * ecut_ha)
                        + np.linalg.norm(k_crystal @ reciprocal))
                       / minimum_stretch)) + 1


def _miller_cube(bound: int) -> np.ndarray:
    side = 2 * bound + 1
    grid = np.indices((side, side, side), dtype=np.int32) - bound
    return np.moveaxis(grid, 0, -1).reshape(-1, 3)


def make_bases(reciprocal: np.ndarray, k_crystal: np.ndarray,
               ecut_ry: float) -> list[PlaneWaveBasis]:
    """Bases of several k-points, rows in QE order (``basis.py:1483-1577``).

    The global G list is ordered by |G|^2 (ties within 1e-8 keep Miller-cube
    order); each k keeps the G with 0.5|k+G|^2 <= ecut_ha + 1e-12, ordered by
    |k+G|^2 with ties keeping global order.
    """
    reciprocal = np.asarray(reciprocal, dtype=float)
    kpoints = np.atleast_2d(np.asarray(k_crystal, dtype=float))
    if kpoints.shape[1] != 3 or not len(kpoints):
        raise ValueError("k_crystal must have shape (3,) or (nk, 3)")
    ecut_ha = 0.5 * float(ecut_ry)
    stretch = float(np.linalg.svd(reciprocal, compute_uv=False)[-1])
    bounds = [_basis_bound(reciprocal, k, ecut_ha, stretch) for k in kpoints]
    cube = _miller_cube(max(bounds))
    gvec = cube @ reciprocal
    cube = cube[qe_energy_order(np.einsum("ij,ij->i", gvec, gvec))]
    out = []
    for k in kpoints:
        vectors = (cube + k) @ reciprocal
        kinetic = 0.5 * np.einsum("ij,ij->i", vectors, vectors)
        keep = np.flatnonzero(kinetic <= ecut_ha + 1.0e-12)
        if not len(keep):
            raise ValueError("ecutwfc produces an empty plane-wave basis")
        keep = keep[qe_energy_order(kinetic[keep])]
        vec = vectors[keep]
        out.append(PlaneWaveBasis(
            np.ascontiguousarray(cube[keep], dtype=np.int32),
            np.ascontiguousarray(vec),
            np.ascontiguousarray(0.5 * np.einsum("ij,ij->i", vec, vec)),
            np.asarray(k, dtype=float).copy(),
        ))
    return out


def make_basis(reciprocal, k_crystal, ecut_ry) -> PlaneWaveBasis:
    return make_bases(reciprocal, k_crystal, ecut_ry)[0]


def fft_shape(bases, reciprocal=None, ecutrho_ry=None,
              fft_factors=(1, 1, 1)) -> tuple[int, int, int]:
    """FFT grid from the charge-density sphere (``basis.py:1619-1682``)."""
    if not bases:
        raise ValueError("at least one plane-wave basis is required")
    if (reciprocal is None) != (ecutrho_ry is None):
        raise ValueError("reciprocal and ecutrho_ry must be supplied together")
    if reciprocal is not None:
        rho = make_basis(np.asarray(reciprocal, float), np.zeros(3), float(ecutrho_ry))
        largest = np.max(np.abs(rho.indices), axis=0)
        shape = []
        for value, factor in zip(largest, fft_factors):
            order = int(2 * value + 1)
            while True:
                cand = next_fast_len(order)
                if cand % factor == 0:
                    shape.append(cand)
                    break
                order = cand + 1
        return tuple(shape)
    span = np.max(np.stack([np.max(b.indices, 0) - np.min(b.indices, 0)
                            for b in bases]), axis=0)
    shape = []
    for value in span:
        order = next_fast_len(int(2 * value + 2))
        if order % 2:
            order = next_fast_len(order + 1)
        shape.append(order)
    return tuple(shape)


def potential_matrix(vg: np.ndarray, indices: np.ndarray) -> np.ndarray:
    """Dense <G|V|G'> = V(G-G') on the FFT grid (``basis.py:1685-1688``)."""
    delta = indices[:, None, :] - indices[None, :, :]
    slots = tuple(delta[:, :, a] % vg.shape[a] for a in range(3))
    return vg[slots]


def linear_slots(indices: np.ndarray, shape) -> np.ndarray:
    """C-order FFT slot of each G (``basis.py:515-523``)."""
    idx = np.asarray(indices)
    return np.ravel_multi_index(tuple(idx[:, a] % shape[a] for a in range(3)),
                                tuple(shape)).astype(np.int64)


class LocalPotentialWorkspace:
    """Device counterpart of the reference's ``LocalPotentialWorkspace``
    (``qepy_pw/basis.py:457-555``, serial branch) for one k-point basis: the
    sphere-pruned FFT plan on the GPU, shared by a ``PlaneWaveHamiltonian``
    built with ``local_workspace=`` this object, and the occupied-state
    density of the SCF's ``sum_band`` (``accumulate_density``,
    ``basis.py:1210-1282``; called from ``pw/scf.py:1382-1451``).

    Only the single-process form exists here (one GPU): an ``mpi`` of more
    than one rank is rejected, as the device Hamiltonian does.
    """

    def __init__(self, indices, shape, *, mpi=None, timers=None, device=None, **options):
        from . import device as dv
        idx = indices.indices if hasattr(indices, "indices") else indices
        idx = np.asarray(idx)
        if idx.ndim != 2 or idx.shape[1] != 3:
            raise ValueError("indices must have shape (nplane, 3)")
        self.shape = tuple(int(s) for s in shape)
        if len(self.shape) != 3:
            raise ValueError("the FFT grid shape must be three-dimensional")
        if mpi is not None and getattr(mpi, "size", 1) > 1:
            raise ValueError("the device workspace runs on one GPU; a distributed "
                             "LocalPotentialWorkspace is not supported")
        self.mpi = mpi
        self.timers = timers
        self.local_slab = slice(0, self.shape[2])
        self.device = dv.require_device(device)
        self.indices = idx
        self.plan = dv.DeviceFFTPlan(idx, self.shape, self.device)
        self.workspace = dv.Workspace(self.device)
        self.npw = self.plan.npw

    def accumulate_density(self, density, coefficients, band_weights) -> None:
        """density += sum_b band_weights[b] |psi_b(r)|^2, psi_b(r) the
        unnormalised backward FFT of coefficients[:, b] (reference row order).
        ``density``: (n1, n2, n3) float64, a numpy array or a device tensor,
        updated in place; ``coefficients``: (npw, nb) complex, host or
        device."""
        import torch

        from . import _native as nat
        from . import device as dv
        weights = np.asarray(band_weights, dtype=float)
        nb = len(weights)
        if getattr(coefficients, "ndim", 0) != 2 or coefficients.shape[1] != nb:
            raise ValueError("density coefficients and band weights disagree")
        if tuple(density.shape) != self.shape:
            raise ValueError("density accumulator has the wrong local shape")
        if coefficients.shape[0] != self.npw:
            raise ValueError("density coefficients do not match the basis")
        if nb == 0:
            return
        context = self.timers.measure("fftw") if self.timers is not None else None
        if context is not None:
            context.__enter__()
        try:
            from .diagonalization import _upload_block
            # device order, one band per row of a (nb, npw) tensor
            psi = _upload_block(coefficients, self.plan.perm_device, self.npw, self.device,
                                torch.complex128)
            w_dev = dv.upload(weights, self.device)
            n1, n2, n3 = self.shape
            out = torch.empty((n3, n2, n1), dtype=torch.float64, device=self.device)
            self._accumulate_device(psi, w_dev, out)
            grid = out.permute(2, 1, 0)
            if isinstance(density, torch.Tensor):
                density += grid.to(density.device, density.dtype)
            else:
                density += grid.cpu().numpy()
        finally:
            if context is not None:
                context.__exit__(None, None, None)

    def _accumulate_device(self, psi, weights, out) -> None:
        """out (n3, n2, n1) = sum_b weights[b] |psi_b(r)|^2 for device-order
        bands psi (nb, npw) and device weights (nb), on the current stream."""
        from . import _native as nat
        from . import device as dv
        lib = nat.load()
        prec = dv.precision_code(psi.dtype)
        nb = int(psi.shape[0])
        need = lib.qepw_accumulate_density_workspace(self.plan.handle, prec, nb)
        buf = self.workspace.get(need)
        rc = lib.qepw_accumulate_density(self.plan.handle, prec, psi.data_ptr(), dv.ld_of(psi),
                                         nb, weights.data_ptr(), 1.0, out.data_ptr(),
                                         buf.data_ptr(), buf.numel(),
                                         dv.stream_handle(self.device))
        nat.check(rc, "qepw_accumulate_density")
