"""The Davidson golden fixtures carry their own error bounds.

``tests/golden/make_golden.py`` solves the medium workload and the 24
problems of the uneven batch far past the 1e-10 Ry bar (|de| < 5e-13 Ha,
residual norms < 1e-9) and stores, beside the eigenvalues, bounds that do
not rest on the Davidson's stop test: with R = H X - X diag(w) recomputed by
the oracle h_psi on orthonormal Ritz vectors X, each eigenvalue is within
||R||_2 of an eigenvalue of H and within ||R||_2^2 / gap of its own.  These
checks hold the fixtures to the bars the GPU tests compare against:
5e-11 Ha (1e-10 Ry) for the FP64 form and 5e-9 Ha (1e-8 Ry) for the mixed
form and the batch.
"""

from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).parent / "golden"


def test_medium_golden_is_bounded_below_the_fp64_bar():
    d = np.load(GOLDEN / "medium_davidson.npz")
    w = d["eigenvalues"]
    assert w.shape == (128,) and np.all(np.diff(w) >= 0.0)
    assert float(d["bound_quadratic"]) < 0.5e-10 * 1e-3
    assert float(d["bound_linear"]) < 1e-9
    assert np.all(d["residual_norms"] < 1e-9)
    assert float(d["orthonormality"]) < 1e-12
    # the gap above the block, with the next Ritz values, is what makes the
    # quadratic bound meaningful
    assert float(d["gap"]) > 1e-4 and d["above"][0] > w[-1]


def test_batch_golden_covers_all_24_problems():
    d = np.load(GOLDEN / "batch_davidson.npz")
    roots = {"cell54": 64, "cell64": 96, "cell72": 128}
    npw = {"cell54": (9000, 11500), "cell64": (14500, 17500), "cell72": (22500, 25500)}
    for cell, n in roots.items():
        for k in range(8):
            key = f"{cell}_k{k}"
            assert d[key].shape == (n,), key
            lo, hi = npw[cell]
            assert lo < int(d[key + "_npw"]) < hi, key
            assert float(d[key + "_bound"]) < 0.5e-10 * 1e-3, key
            assert float(d[key + "_bound_linear"]) < 1e-9, key
