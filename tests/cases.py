"""Shared set-ups for the tests: the same mesh and boundary conditions given
to the oracle (CPU restatement) and to the device path."""
import numpy as np

from warpcfd_amd.preprocess.grid import structured_hex, cavity_mesh


def apply_bcs(boundary, kind):
    if kind == 'cavity':
        for w in ('-X', '+X', '-Y'):
            boundary.no_slip_wall(w)
        boundary.velocity_BC('+Y', u=1.0, v=0.0, w=0.0)
        boundary.slip_wall('+Z'); boundary.slip_wall('-Z')
    elif kind == 'closed_cavity':
        # every wall no-slip: on cells without z-symmetry (tetrahedra) the
        # slip walls above let mass through, since slip_wall gives w a zero
        # normal gradient too (boundary.py:101-105)
        for w in ('-X', '+X', '-Y', '-Z', '+Z'):
            boundary.no_slip_wall(w)
        boundary.velocity_BC('+Y', u=1.0, v=0.0, w=0.0)
    elif kind == 'channel':
        boundary.velocity_BC('-X', u=1.0, v=0.0, w=0.0)
        boundary.pressure_BC('+X', p=0.0)
        for w in ('-Y', '+Y', '-Z', '+Z'):
            boundary.no_slip_wall(w)
    elif kind == 'channel2d':
        boundary.velocity_BC('-X', u=1.0, v=0.0, w=0.0)
        boundary.pressure_BC('+X', p=0.0)
        boundary.no_slip_wall('-Y'); boundary.no_slip_wall('+Y')
        boundary.slip_wall('-Z'); boundary.slip_wall('+Z')
    else:
        raise ValueError(kind)


def oracle_case(nodes, cells, kind, dtype=np.float32, ref_cell=0, geom_dtype=None):
    """``geom_dtype``: the mesh geometry's type (default: float32 with float32
    fields, float64 with float64 fields); float32 geometry with float64
    fields is the reference's float64 model (mesh.py:22-33)."""
    from oracle.mesh import OracleMesh
    from oracle.fv import OracleFVM
    if geom_dtype is None:
        geom_dtype = np.float32 if dtype == np.float32 else np.float64
    om = OracleMesh(nodes, cells, geom_dtype=geom_dtype)
    om.define_boundary_walls()
    of = OracleFVM(om, dtype)
    if ref_cell is not None:
        of.set_reference_pressure(ref_cell, 0.0)
    apply_bcs(of.boundary, kind)
    of.boundary.finalize()
    return of


def device_case(nodes, cells, kind, ref_cell=0, nu=0.01, renumber=True, float_dtype=np.float32):
    from warpcfd_amd.preprocess import Mesh, define_boundary_walls
    from warpcfd_amd.FV import FVM
    m = Mesh(nodes, cells)
    define_boundary_walls(m)
    model = FVM(m, renumber=renumber, float_dtype=float_dtype)
    model.material.create_incompressible_newtonian_fluid(viscosity=nu, density=1.0)
    if ref_cell is not None:
        model.set_reference_pressure(ref_cell, 0.0)
    apply_bcs(model.boundary, kind)
    model.finalize()
    return model


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def jittered_box(n=(9, 7, 5), jitter=0.1, seed=1, lengths=(1.0, 0.8, 0.6)):
    return structured_hex(n, lengths, jitter=jitter, seed=seed)
