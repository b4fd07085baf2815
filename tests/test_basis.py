"""Basis / grid construction against the reference's own tests and sizes.

Mirrors ``tests/unit/core/test_basis.py:18-89`` of the reference (fast FFT
lengths, basis order and cutoff, grid factors, V_loc known answer) plus the
plane-wave / stick counts of the synthetic workloads.
"""

import numpy as np
import pytest

from qe_python_amd.basis import (fft_shape, linear_slots, make_bases, next_fast_len,
                                 potential_matrix, qe_energy_order)
from qe_python_amd.synthetic import make_system


@pytest.mark.parametrize(("target", "expected"),
                         [(0, 1), (1, 1), (7, 8), (11, 12), (17, 18), (31, 32)])
def test_fast_lengths_are_235_smooth(target, expected):
    assert next_fast_len(target) == expected


def test_bases_sorted_by_kinetic_within_cutoff_and_unique():
    reciprocal = 2.0 * np.pi / 5.0 * np.eye(3)
    bases = make_bases(reciprocal, np.asarray([[0.0, 0.0, 0.0], [0.25, 0.0, 0.0]]), 10.0)
    for b in bases:
        assert np.all(np.diff(b.kinetic) >= -1.0e-12)
        assert np.max(b.kinetic) <= 5.0 + 1.0e-12
        assert len(np.unique(b.indices, axis=0)) == len(b)


def test_fft_shape_respects_factors():
    reciprocal = 2.0 * np.pi / 6.0 * np.eye(3)
    bases = make_bases(reciprocal, np.zeros(3), 8.0)
    shape = fft_shape(bases, reciprocal, ecutrho_ry=32.0, fft_factors=(2, 3, 1))
    assert shape[0] % 2 == 0 and shape[1] % 3 == 0
    assert all(next_fast_len(s) == s for s in shape)


def test_energy_order_keeps_index_order_inside_ties():
    values = np.array([1.0, 0.5, 1.0 + 1e-10, 0.5 - 1e-10, 2.0])
    order = qe_energy_order(values)
    assert list(order) == [1, 3, 0, 2, 4]


@pytest.mark.parametrize("name,npw,grid,sticks", [
    ("small", 2945, (36, 36, 36), 241),
    ("medium", 23847, (72, 72, 72), 997),
    ("cell54", 10059, (54, 54, 54), 561),
    ("cell64", 15967, (64, 64, 64), 769),
])
def test_workload_sizes(name, npw, grid, sticks):
    s = make_system(name)
    assert s.npw == npw
    assert s.shape == grid
    m = s.basis.indices
    xy = np.unique((m[:, 0] % grid[0]) * grid[1] + (m[:, 1] % grid[1]))
    assert len(xy) == sticks
    # G sphere fits the grid without aliasing: slots are unique
    assert len(np.unique(linear_slots(m, grid))) == npw


def test_potential_matrix_is_convolution():
    rng = np.random.default_rng(3)
    vg = rng.standard_normal((6, 6, 6)) + 1j * rng.standard_normal((6, 6, 6))
    idx = np.array([[0, 0, 0], [1, 0, 0], [0, -1, 2]])
    mat = potential_matrix(vg, idx)
    assert mat[1, 2] == vg[(1 - 0) % 6, (0 + 1) % 6, (0 - 2) % 6]


def test_gemm_wrapper_guards_raise_before_the_library():
    """The GEMM wrapper's operand checks (undersized views, 3-D tensors,
    non-unit inner strides, lazy conj views, mixed dtypes) raise ValueError
    on the host before any device call."""
    import torch

    from qe_python_amd import _native as nat
    from qe_python_amd import device as dv
    z = lambda *s, dt=torch.complex128: torch.zeros(s, dtype=dt)  # noqa: E731
    with pytest.raises(ValueError, match="smaller"):
        dv.gemm(nat.OP_N, 10, 4, 3, 1.0, z(3, 9), z(4, 3), 0.0, z(4, 10), None)
    with pytest.raises(ValueError, match="smaller"):
        dv.gemm(nat.OP_C, 10, 4, 30, 1.0, z(10, 30), z(4, 30), 0.0, z(3, 10), None)
    with pytest.raises(ValueError, match="unit inner stride"):
        dv.gemm(nat.OP_N, 4, 4, 4, 1.0, z(2, 4, 4), z(4, 4), 0.0, z(4, 4), None)
    with pytest.raises(ValueError, match="unit inner stride"):
        dv.gemm(nat.OP_N, 4, 4, 4, 1.0, z(4, 8)[:, ::2], z(4, 4), 0.0, z(4, 4), None)
    # one-column operands with strided elements: the kernels would read the
    # column contiguously
    with pytest.raises(ValueError, match="unit inner stride"):
        dv.gemm(nat.OP_N, 4, 1, 4, 1.0, z(4, 4), z(1, 8)[:, ::2], 0.0, z(1, 4), None)
    with pytest.raises(ValueError, match="unit inner stride"):
        dv.gemm(nat.OP_C, 1, 1, 4, 1.0, z(4, 4).T[:1], z(1, 4), 0.0, z(1, 1), None)
    with pytest.raises(ValueError, match="lazy conj"):
        dv.gemm(nat.OP_N, 4, 4, 4, 1.0, z(4, 4).conj(), z(4, 4), 0.0, z(4, 4), None)
    with pytest.raises(ValueError, match="one complex dtype"):
        dv.gemm(nat.OP_N, 4, 4, 4, 1.0, z(4, 4, dt=torch.complex64), z(4, 4), 0.0, z(4, 4), None)
    assert dv.ld_of(z(5, 7)) == 7 and dv.ld_of(z(1, 7)) == 7 and dv.ld_of(z(5, 7)[:, :3]) == 7
