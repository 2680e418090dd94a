"""examples/LDC/LDC.py:1-63 run through the reference's own import path
(``warp_cfd``) with only its Warp and pyvista lines removed (``import warp``,
``wp.init()``, ``float_dtype=wp.float32`` -> ``np.float32``): a 16 x 16
cavity, 20 SIMPLE iterations via ``solver.run``, results read with
``model.cell_values.numpy()`` and ``struct_member_to_array``, against the
oracle on the same mesh (velocity rel-L2 <= 1e-4)."""
import numpy as np
import pytest

from cases import oracle_case, rel_l2

pytestmark = pytest.mark.gpu


def test_ldc_example_through_warp_cfd(gpu):
    from warp_cfd.FV.model import FVM
    from warp_cfd.FV.implicit_Solvers import IncompressibleSolver
    from warp_cfd.preprocess import create_2D_grid, Mesh, define_boundary_walls
    from oracle.simple import OracleSolver
    n = 16
    Re = 100
    G, nu = 1, 1 / Re
    dz = 0.1
    pv_mesh = create_2D_grid((0, 0, 0), n, n, 1, 1, dz=dz, element_type='hex', display_mesh=False)
    m = Mesh(pv_mesh)
    define_boundary_walls(m)
    model = FVM(m, output_variables=['u', 'v', 'w', 'p'], float_dtype=np.float32)
    model.material.create_incompressible_newtonian_fluid(viscosity=nu, density=1.)
    model.set_reference_pressure(0, 0.)
    model.boundary.no_slip_wall('-X')
    model.boundary.no_slip_wall('+X')
    model.boundary.no_slip_wall('-Y')
    model.boundary.velocity_BC('+Y', u=1, v=0, w=0)
    model.boundary.slip_wall('+Z')
    model.boundary.slip_wall('-Z')
    model.finalize()
    solver = IncompressibleSolver(model, 0.7, 0.3, correction=False)
    solver.set_iterations(20, 1, 1)
    solver.run(10)
    velocity = model.cell_values.numpy()[:, 0:3].reshape(-1, 3)
    p = model.cell_values.numpy()[:, 3]
    centroids = model.struct_member_to_array('centroid', 'cells')
    assert isinstance(velocity, np.ndarray) and velocity.shape == (n * n, 3)

    of = oracle_case(pv_mesh.points, next(iter(pv_mesh.cells_dict.values())), 'cavity')
    os_ = OracleSolver(of, 0.7, 0.3, nu=nu)
    for _ in range(20):
        os_.outer_iteration()
    err = rel_l2(velocity, of.cell_values[:, :3])
    print(f'LDC through warp_cfd: velocity rel-L2 vs oracle {err:.2e}')
    assert err <= 1e-4
    assert rel_l2(p, of.cell_values[:, 3]) <= 1e-4
    np.testing.assert_allclose(centroids, of.centroids, atol=1e-6)
