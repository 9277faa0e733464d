"""CPU oracle for the petgem CSEM forward hot path -- TEST INFRASTRUCTURE ONLY.

This package is a plain FP64 numpy/scipy restatement of the reference
algorithm (ocastilloreyes/petgem 1.0).  It exists to *check* the HIP path:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker.  The product package
``petgem_amd`` never imports, links or executes anything in here, and fails
loudly when its HIP library is missing instead of falling back to this code.

Modules and the reference code each one restates (paths relative to the
reference repository root):

* ``quadrature`` -- ``petgem/hvfem.py:1055-1127,1603-1610`` (symmetric tet
  rules of order 2/4/6 and their mapping to the unit tetrahedron).
* ``basis`` -- ``petgem/hvfem.py:319-1052`` (``shape3DETet`` and helpers:
  hierarchical Nedelec H(curl) shape functions and curls, p <= 3).
* ``topology`` -- ``petgem/mesh.py:18-125,149-321``, ``petgem/vectors.py:15-132``,
  ``petgem/hvfem.py:15-220`` (edge/face numbering, face owners, DOF
  connectivity, orientations, boundary DOFs).
* ``fem`` -- ``petgem/hvfem.py:101-316,2303-2553``, ``petgem/solver.py:156-316,552-590``,
  ``petgem/postprocessing.py:479-616`` (element matrices, CSR assembly,
  Dirichlet BC, point-dipole source, receiver interpolation, direct solve).
* ``krylov`` -- COCG / GMRES(m) restatements used as CPU baselines.

Pinning status of each function (how a plausible mistake would be caught) is
recorded in the module docstrings and in the tests under ``tests/``.
"""
