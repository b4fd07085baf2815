"""Binding the device path into the reference's SCF module.

``qepy_pw/pw/scf.py:29-40`` imports its eigensolver names from
``qepy_pw.diagonalization`` by value, so the SCF keeps whatever objects were
bound when it was imported.  ``bind`` rebinds them in the SCF module's own
namespace -- the mechanism the reference's tests use to patch
``qepy_pw.pw.scf.davidson`` (``tests/integration/test_qe_reference.py:195-216``)
-- after which the SCF builds device Hamiltonians
(``pw/scf.py:3396-3411``), rotates the starting subspace through them
(``pw/scf.py:923-953``) and calls the device ``davidson`` with its retry
loop (``pw/scf.py:3497-3556``) without a change to its own code.
"""

from __future__ import annotations

from . import diagonalization as _dev

_ABSENT = object()

# the names pw/scf.py:29-40 imports, in its order
SCF_NAMES = ("FactorizedProjectorTerm", "PackedProjectorTerm", "PlaneWaveHamiltonian",
             "ProjectorTerm", "_lowest_generalized_eigh", "_matmul_f", "conjugate_gradient",
             "davidson", "parallel_orbital", "rmm_diis")


def bind(scf_module, precision: str = "fp64"):
    """Rebind every name of ``SCF_NAMES`` in ``scf_module`` to the device
    implementation and return the previous bindings (a dict, for
    ``unbind``).  ``precision`` is the ``PlaneWaveHamiltonian`` precision
    the SCF's operators get: "fp64", or "mixed" (``apply`` stays FP64;
    every Davidson call on such an operator runs the mixed form)."""
    if precision not in ("fp64", "mixed"):
        raise ValueError("precision must be 'fp64' or 'mixed'")
    previous = {name: getattr(scf_module, name, _ABSENT) for name in SCF_NAMES}
    for name in SCF_NAMES:
        setattr(scf_module, name, getattr(_dev, name))
    if precision != "fp64":
        base = _dev.PlaneWaveHamiltonian

        class PlaneWaveHamiltonian(base):
            """``PlaneWaveHamiltonian`` whose Davidson runs the mixed form."""

            def __init__(self, *args, **kwargs):
                kwargs.setdefault("precision", precision)
                super().__init__(*args, **kwargs)

        setattr(scf_module, "PlaneWaveHamiltonian", PlaneWaveHamiltonian)
    return previous


def unbind(scf_module, previous: dict) -> None:
    """Restore the bindings ``bind`` returned."""
    for name, value in previous.items():
        if value is _ABSENT:
            if hasattr(scf_module, name):
                delattr(scf_module, name)
        else:
            setattr(scf_module, name, value)
