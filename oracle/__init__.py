"""CPU oracle for the SIMPLE/PISO inner loop of JohnCSu/WarpCFD.

TEST INFRASTRUCTURE ONLY.  This package is a plain numpy restatement of the
reference algorithm, written by reading the reference files as text.  It is
the checker for the HIP path: only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import it.  The product package
(``warpcfd_amd``) never imports it and has no CPU fallback.

Modules and what pins them (the tests named here live under ``tests/``):

* ``oracle.mesh``    -- preprocess/mesh.py + FV/mesh_structs.py:141-251 +
  preprocess/grid.py:165-194.  Pinned by closed forms on uniform grids
  (areas, unit normals, psi = 1/2, volume sums), by closed-cell sum(n A) = 0,
  by a brute-force pair loop written like mesh.py:263-281 on tiny meshes
  (test_oracle_mesh.py) and by permutation invariance.
* ``oracle.fv``      -- FV/model.py, FV/interpolation_Schemes, FV/kernels,
  FV/terms, FV/Implicit_Schemes, FV/kernels/equation_kernels.py.  Pinned by
  exactness for linear fields, the 7-point Laplacian stencil, hand-computed
  upwind rows, the Patankar relaxation closed form and the identity row
  (test_oracle_fv.py).
* ``oracle.krylov``  -- our restatement of warp.optim.linear.cg/bicgstab
  (third party, not vendored; the reference pins ``warp==1.0.4`` in
  requirements.txt:9 while its code needs warp-lang >= 1.4).  Pinned by the
  true residual ||b - A x|| and by scipy.sparse.linalg.spsolve on the same
  systems (test_oracle_krylov.py).  Iterate-level behaviour (stopping
  test, check interval) is *parity unpinned*: no reference test or vector
  fixes it; only converged fields are compared.
* ``oracle.simple``  -- FV/implicit_Solvers/incompressible.py:72-172.
  Pinned loosely by the Ghia et al. (1982) centreline table that ships with
  the reference (examples/LDC/u_velocity_results.txt,
  examples/LDC/v_velocity_results.csv) and by FP32-vs-FP64 agreement.

Cell centroids and volumes: the reference takes them from VTK
(mesh.py:32-33), which is absent.  For hexahedra VTK's cell centre is the
trilinear parametric centre, i.e. the mean of the 8 vertices, which is what
we use.  The volume is computed with the divergence theorem over the same
fan-triangulated faces the reference uses for areas; for planar-faced hexes
this equals VTK's value, for jittered (warped) hexes it is *parity unpinned*.
"""
