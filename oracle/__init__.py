"""CPU oracle for the plane-wave h_psi + Davidson path -- TEST INFRASTRUCTURE.

This package restates, in NumPy/SciPy, the reference's serial algorithms for
the hot path (sjhong6230/QE_python, qepy-pw 1.0.2).  It exists only to check
the device implementation and to time the CPU baseline:

* only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
  leg may import it, and only as the checker / the timed CPU baseline;
* the product package ``qe_python_amd`` never imports it.

Modules and the reference code each follows:

* ``oracle.hpsi``     -- ``PlaneWaveHamiltonian._apply`` / ``diagonal``
  (``qepy_pw/diagonalization.py:205-369``), the serial V_loc sandwich
  (``basis.py:1354-1455`` -> ``_native_fft.pyx:2338-2455``),
  ``prepare_potential`` (``basis.py:777-792``, ``_native_fft.pyx:2188-2202``)
  and density accumulation (``_native_fft.pyx:2592-2648``).
* ``oracle.davidson`` -- ``davidson`` and its helpers
  (``diagonalization.py:46-110, 372-880``) with the elementwise kernels of
  ``_native_fft.pyx:852-899, 1153-1291, 1495-1530``.
* ``oracle.dense``    -- dense brute-force H (``basis.py:1685-1688`` plus
  T and B D B^H) used to pin ``oracle.hpsi`` and ``oracle.davidson``.
* ``oracle.forces``   -- the force branch of ``_nonlocal_derivatives``
  (``pw/scf.py:1521-1631``) and dE/dV = rho/N.

Pinning status of every function is listed in the module docstrings; the
checks live in ``tests/test_oracle_*.py``.  The reference's own native
extension cannot be built here (it needs mpi4py at build time and FFTW, both
absent), so there is no ``oracle/_ref`` build.
"""
