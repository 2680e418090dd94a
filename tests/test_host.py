"""Host-side behaviour of the API mirror (no GPU needed): boundary
bookkeeping and the errors the reference raises."""
import numpy as np
import pytest

from warpcfd_amd.preprocess import Mesh, cavity_mesh, define_boundary_walls
from warpcfd_amd.FV import FVM
from warpcfd_amd import _lib


def _model():
    nodes, cells = cavity_mesh(6)
    m = Mesh(nodes, cells)
    define_boundary_walls(m)
    return FVM(m, device='cpu')


def test_boundary_helpers_set_types_like_the_reference():
    model = _model()
    b = model.boundary
    b.no_slip_wall('-X'); b.no_slip_wall('+X'); b.no_slip_wall('-Y')
    b.velocity_BC('+Y', u=1.0, v=0.0, w=0.0)
    b.slip_wall('+Z'); b.slip_wall('-Z')
    g = model.mesh.groups
    rows = np.isin(b.boundary_ids, g['+Y'].ids)
    assert np.all(b.boundary_type[rows] == [1, 1, 1, 2])
    assert np.all(b.boundary_values[rows] == [1, 0, 0, 0])
    rows = np.isin(b.boundary_ids, g['+Z'].ids)
    assert np.all(b.boundary_type[rows] == [2, 2, 2, 2])       # T12
    b.finalize()
    assert b.uniform_velocity_types()


def test_unassigned_boundary_raises():
    model = _model()
    model.boundary.no_slip_wall('-X')
    with pytest.raises(ValueError):
        model.boundary.finalize()


def test_pressure_bc_on_internal_face_rejected():
    model = _model()
    internal = model.face_properties.internal_face_ids[:3]
    with pytest.raises(AssertionError):
        model.boundary.pressure_BC(internal, p=0.0)


def test_device_path_has_no_cpu_fallback():
    model = _model()
    b = model.boundary
    for w in ('-X', '+X', '-Y', '+Y', '+Z', '-Z'):
        b.no_slip_wall(w)
    with pytest.raises(_lib.WcfdError):
        model.finalize()


def test_solver_needs_finalized_model():
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    with pytest.raises(ValueError):
        IncompressibleSolver(_model(), 0.7, 0.3)
