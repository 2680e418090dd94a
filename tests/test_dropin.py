"""The reference's Python surface, without a GPU: its import lines
(examples/*/*.py, warp_cfd/test) resolve through the ``warp_cfd`` package to
the same objects as ``warpcfd_amd``; create_2D_grid gives Mesh the grid it
reads; Material and Boundary keep the reference's calls and error types."""
import importlib

import numpy as np
import pytest


def test_reference_import_lines_resolve_to_the_package():
    from warp_cfd.FV.model import FVM
    from warp_cfd.FV.implicit_Solvers import IncompressibleSolver
    from warp_cfd.preprocess import create_2D_grid, Mesh, define_boundary_walls
    from warp_cfd.preprocess.mesh import Mesh as Mesh2
    from warp_cfd.preprocess.grid import create_2D_grid as grid2
    from warp_cfd.FV.terms import ConvectionTerm, DiffusionTerm, GradTerm, Equation, DdtTerm
    import warpcfd_amd.FV.model as real_model
    import warpcfd_amd.preprocess as real_pre
    import warpcfd_amd.FV.terms as real_terms
    assert FVM is real_model.FVM and Mesh is Mesh2 is real_pre.Mesh
    assert create_2D_grid is grid2 is real_pre.create_2D_grid
    assert define_boundary_walls is real_pre.define_boundary_walls
    assert ConvectionTerm is real_terms.ConvectionTerm and Equation is real_terms.Equation
    assert importlib.import_module('warp_cfd.FV.implicit_Solvers').IncompressibleSolver \
        is IncompressibleSolver
    with pytest.raises(ImportError):
        importlib.import_module('warp_cfd.no_such_module')


@pytest.mark.parametrize('element_type,n_cells,n_nodes', [('hex', 48, 8), ('wedge', 96, 6),
                                                          ('tet', 288, 4)])
def test_create_2D_grid_feeds_mesh(element_type, n_cells, n_nodes, tmp_path):
    from warp_cfd.preprocess import Mesh, create_2D_grid, define_boundary_walls, read_msh
    g = create_2D_grid((0.5, -1.0, 0.0), 8, 6, 2.0, 1.5, dz=0.1, element_type=element_type,
                       save=str(tmp_path / 'grid'))
    m = Mesh(g)
    define_boundary_walls(m)
    assert m.cells.shape == (n_cells, n_nodes)
    assert sorted(m.groups) == ['+X', '+Y', '+Z', '-X', '-Y', '-Z']
    lo, hi = m.nodes.min(axis=0), m.nodes.max(axis=0)
    np.testing.assert_allclose(lo, [0.5, -1.0, 0.0], atol=1e-6)
    np.testing.assert_allclose(hi, [2.5, 0.5, 0.1], atol=1e-6)
    np.testing.assert_allclose(m.cell_volumes.sum(), 2.0 * 1.5 * 0.1, rtol=1e-5)
    back = read_msh(str(tmp_path / 'grid.msh'))
    np.testing.assert_array_equal(back.cells, m.cells)
    np.testing.assert_allclose(back.points, g.points)
    with pytest.raises(NotImplementedError):
        create_2D_grid((0, 0, 0), 4, 4, 1, 1, element_type='wedge', unstructured_wedge=True)
    with pytest.raises(ValueError):
        create_2D_grid((0, 0, 0), 4, 4, 1, 1, element_type='quad')


def test_material_calls_and_errors():
    from warp_cfd.material import Material
    mat = Material()
    with pytest.raises(ValueError, match='density is unset'):
        mat.density
    mat.create_incompressible_newtonian_fluid(viscosity=0.02, density=2.0)
    assert mat.viscosity == 0.02 and mat.density == 2.0 and mat.kinematic_viscosity == 0.01
    mat.set_density(4, incompressible=False)
    mat.set_viscosity(0.5, newtonian=True)
    assert mat.density == 4.0 and mat.viscosity == 0.5
    for bad in (0.0, -1.0, float('nan'), float('inf')):
        with pytest.raises(ValueError, match='positive finite'):
            mat.set_viscosity(bad)


def test_boundary_errors_keep_the_reference_types():
    from warpcfd_amd.FV.boundary import Boundary
    from warpcfd_amd.FV.field import Field
    fields = {o: Field(o, i) for i, o in enumerate('uvwp')}
    b = Boundary(fields, {'inlet': np.array([7, 8])}, boundary_ids=[7, 8, 9])
    with pytest.raises(AssertionError, match='not on the mesh boundary'):
        b.fixed_BC([3, 7], u=1.0)
    with pytest.raises(AssertionError, match="no face group 'outlet'"):
        b.fixed_BC('outlet', u=1.0)
    with pytest.raises(AssertionError, match='no field given'):
        b.fixed_BC('inlet')
    with pytest.raises(ValueError, match='face_ids must be'):
        b.fixed_BC(7.5, u=1.0)
    b.velocity_BC('inlet', 1.0, 0.0, 0.0)
    with pytest.raises(ValueError, match='lack a condition'):
        b.finalize()
    b.slip_wall(9)
    b.finalize()
