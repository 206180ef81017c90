"""CPU oracle for pyIID's scattering path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``pyiid_amd``) imports,
links or executes anything in this package.  It is used by ``tests/``, by
``__graft_entry__.smoke()`` (as the checker) and by the ``cpu_baseline`` leg of
``bench.py``; never as the thing measured or shipped.

Contents
--------
``oracle.scatter``
    NumPy FP64 restatement of the reference's CPU algorithms: experiment grids,
    form-factor broadcasting, flat pair indexing, normalisation, Debye F(Q),
    the reduced-structure-factor gradient (with the reference's -1/2
    convention), the FFT F(Q)->G(r) transform and its closed dense-matrix
    form, Rw / chi^2 and their gradients, the Calc1D energy/forces and the
    leapfrog step.  Each function cites the reference file:line it follows.
``oracle.native``
    ctypes loader for ``debye_oracle.c``: the same Debye sums in C (FP64, and a
    float32-per-pair restatement of the reference 'flat' CPU processor used as
    the timed CPU baseline), multithreaded with pthreads.

Every function is pinned by a check that is not a retyping of its formula
(closed forms, invariants, finite differences, an independent FFT path, the
reference's own known answers and fixtures); the pins live in
``tests/test_oracle_*.py``.  The one part left "parity unpinned" is the
per-element form-factor values for Z other than carbon (the reference takes
them from ``xraylib``, which is absent); single-element results do not depend
on them.
"""
