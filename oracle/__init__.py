"""FP64 CPU oracle for the HPRMAT dense complex LU+solve hot path.

TEST INFRASTRUCTURE ONLY.  This package is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only to compute expected values or to time the reference CPU
path.  The product (``hprmat_amd``) never imports it and never falls back to
it.

Each module restates one part of the reference in plain NumPy/SciPy, written
in the loop-by-loop style of the Fortran it follows (citations are relative to
the jinleiphys/HPRMAT source tree):

* ``lagrange``  - ``rmat_ini``: Gauss-Legendre mesh, T+L (kinetic + Bloch),
                  Bloch blocks and boundary values ``q1``/``q2``
                  (src/rmatrix_hp.F90:61-159, LEGZO :741-783).
* ``channels``  - eta = 0 open-channel functions O = G + iF, rho*O' and the
                  closed-channel Whittaker log-derivative
                  (src/rmatrix_hp.F90:224-244, src/special_functions.f).
* ``rmatrix``   - assembly of C = H - E + Bloch, R = q2^T X, R -> S
                  (src/rmatrix_hp.F90:246-368).
* ``solve``     - type-1 ZGESV path, type-2 mixed path and the GPU
                  refinement semantics (src/rmat_solvers.F90:98-296,
                  src/cusolver_interface.cu:321-443, 737-795).
* ``synth``     - seeded synthetic Lagrange-mesh systems of the benchmark
                  workloads (potential recipe in this module's docstring).

Pinning (what fixes each function independently of the formulas restated
here) is recorded in each module's header and exercised in ``tests/``.
"""
