"""Writes the golden fixtures of the small workload (run from the repo root:
``python tests/golden/make_golden.py``).

small_hpsi.npz:
  indices, real_potential   -- inputs (pin the generator: basis order, V(r))
  trial4                    -- first 4 trial bands
  hpsi                      -- oracle H psi of trial4
  dense_eigenvalues         -- lowest 24 eigenvalues (Ha) of the dense
                               brute-force Hamiltonian (numpy eigvalsh)
  kat_*                     -- the reference's V_loc known answer
                               (tests/unit/core/test_basis.py:76-89)
"""

import sys
from pathlib import Path
from types import SimpleNamespace

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle.dense import dense_hamiltonian, potential_matrix  # noqa: E402
from oracle.hpsi import OracleHamiltonian  # noqa: E402
from qe_python_amd.synthetic import make_system  # noqa: E402


def main():
    s = make_system("small")
    beta, coup, diag = s.packed_term()
    term = SimpleNamespace(beta=beta, coupling=coup, diagonal=diag)
    ham = OracleHamiltonian(s.basis.indices, s.basis.kinetic, s.shape, s.real_potential,
                            (term,), potential_average=float(np.real(s.potential_g.flat[0])))
    dense = dense_hamiltonian(s.basis.indices, s.basis.kinetic, s.real_potential, (term,))
    evals = np.linalg.eigvalsh(dense)[:24]
    shape = (6, 6, 6)
    grid = np.indices(shape).sum(axis=0)
    vr = 0.3 + 0.2 * np.cos(2.0 * np.pi * grid / shape[0])
    vg = np.fft.fftn(vr) / np.prod(shape)
    idx = np.asarray([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0]], dtype=np.int32)
    c = np.asarray([1.0 + 0.2j, -0.3j, 0.4, -0.1 + 0.5j])
    np.savez_compressed(
        Path(__file__).parent / "small_hpsi.npz",
        indices=s.basis.indices, real_potential=s.real_potential,
        trial4=s.trial[:, :4], hpsi=ham.apply(s.trial[:, :4]), dense_eigenvalues=evals,
        kat_potential_g=vg, kat_indices=idx, kat_coefficients=c,
        kat_expected=potential_matrix(vg, idx) @ c)
    print("lowest dense eigenvalues (Ry):", 2 * evals[:10])


if __name__ == "__main__" and len(sys.argv) == 1:
    main()


def medium():
    """medium_davidson.npz: oracle Davidson of the medium workload (128 roots,
    tolerance 0.5e-10 Ha, residual_factor=None), ~1 minute on 8 cores."""
    from oracle.davidson import davidson
    s = make_system("medium")
    beta, coup, diag = s.packed_term()
    term = SimpleNamespace(beta=beta, coupling=coup, diagonal=diag)
    ham = OracleHamiltonian(s.basis.indices, s.basis.kinetic, s.shape, s.real_potential,
                            (term,), potential_average=float(np.real(s.potential_g.flat[0])))
    res = davidson(ham.apply, ham.diagonal, 128, initial_vectors=s.trial, tolerance=0.5e-10,
                   residual_factor=None, operator_into=ham.apply_into)
    np.savez_compressed(Path(__file__).parent / "medium_davidson.npz",
                        eigenvalues=res.eigenvalues, residual_norms=res.residual_norms,
                        iterations=res.iterations, converged=res.converged,
                        hamiltonian_applications=res.hamiltonian_applications)
    print("medium:", res.iterations, res.converged, res.hamiltonian_applications)


if __name__ == "__main__" and "--medium" in sys.argv:
    medium()


BATCH_PROBLEMS = [(cell, k) for cell in ("cell54", "cell64", "cell72") for k in (0, 1)]


def batch():
    """batch_davidson.npz: oracle Davidson (tolerance 0.5e-10 Ha,
    residual_factor=None, up to 200 iterations) of the first two seeded k-points of each cell of
    the uneven batched workload, keys ``<cell>_k<i>``; a few minutes on 8
    cores."""
    from oracle.davidson import davidson
    from qe_python_amd.synthetic import batch_kpoints
    out = {}
    for cell, k in BATCH_PROBLEMS:
        s = make_system(cell, kpoint=batch_kpoints(cell)[k])
        beta, coup, diag = s.packed_term()
        term = SimpleNamespace(beta=beta, coupling=coup, diagonal=diag)
        ham = OracleHamiltonian(s.basis.indices, s.basis.kinetic, s.shape, s.real_potential,
                                (term,), potential_average=float(np.real(s.potential_g.flat[0])))
        res = davidson(ham.apply, ham.diagonal, s.nroots, initial_vectors=s.trial,
                       tolerance=0.5e-10, residual_factor=None, operator_into=ham.apply_into,
                       max_iterations=200)
        assert res.converged, (cell, k)
        out[f"{cell}_k{k}"] = res.eigenvalues
        out[f"{cell}_k{k}_npw"] = np.asarray(s.npw)
        print(cell, k, s.npw, res.iterations, res.converged, flush=True)
    np.savez_compressed(Path(__file__).parent / "batch_davidson.npz", **out)


if __name__ == "__main__" and "--batch" in sys.argv:
    batch()
