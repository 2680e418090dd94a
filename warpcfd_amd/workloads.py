"""Case builders for the standard workloads.

* ``cavity_model(n, re)``: lid-driven cavity on the unit square, n x n x 1
  hexahedra (dz = 0.1), lid u = 1 on +Y, no-slip -X/+X/-Y, slip +-Z, nu = 1/Re,
  reference pressure 0 in cell 0 (examples/LDC/LDC.py:24-45).
* ``channel_model(...)``: box channel, uniform inlet u = 1 on -X, p = 0 on +X,
  no-slip walls, interior vertices jittered (seeded).
* ``cylinder_model(...)``: transient flow past a cylinder -- a box with a
  voxel-carved cylinder, inflow u = 1 on -X, p = 0 on +X, slip walls on
  +-Y and +-Z, no-slip cylinder, nu = U D / Re.
* ``batch64(seed)``: the 64-cavity batch -- N = round(exp(U(ln 32, ln 256))),
  Re ~ U(100, 1000), one seeded generator.

Models are returned un-finalized (``FVM`` objects with boundary conditions,
material and reference pressure set), ready for ``FVM.finalize`` (one case on
its own) or for ``batch.CaseSpec`` (a case in an arena).
"""
import numpy as np

from .FV import FVM
from .preprocess import (Mesh, box_and_body_walls, cavity_mesh, channel_mesh, cylinder_mesh,
                         define_boundary_walls, structured_hex)


def cavity_bcs(b):
    for w in ('-X', '+X', '-Y'):
        b.no_slip_wall(w)
    b.velocity_BC('+Y', u=1.0, v=0.0, w=0.0)
    b.slip_wall('+Z')
    b.slip_wall('-Z')


def channel_bcs(b):
    b.velocity_BC('-X', u=1.0, v=0.0, w=0.0)
    b.pressure_BC('+X', p=0.0)
    for w in ('-Y', '+Y', '-Z', '+Z'):
        b.no_slip_wall(w)


def cavity_model(n, re=100.0, device='cuda', float_dtype=np.float32):
    nodes, cells = cavity_mesh(int(n))
    m = Mesh(nodes, cells)
    define_boundary_walls(m)
    model = FVM(m, device=device, float_dtype=float_dtype)
    model.material.create_incompressible_newtonian_fluid(viscosity=1.0 / float(re), density=1.0)
    model.set_reference_pressure(0, 0.0)
    cavity_bcs(model.boundary)
    return model


def channel_model(counts=(200, 40, 40), lengths=(10.0, 2.0, 2.0), re=200.0, jitter=0.1, seed=0,
                  device='cuda', float_dtype=np.float32):
    """Re on the channel height: nu = U H / Re with U = 1, H = lengths[1]."""
    nodes, cells = channel_mesh(*counts, lengths=lengths, jitter=jitter, seed=seed)
    m = Mesh(nodes, cells)
    define_boundary_walls(m)
    model = FVM(m, device=device, float_dtype=float_dtype)
    model.material.create_incompressible_newtonian_fluid(viscosity=lengths[1] / float(re),
                                                         density=1.0)
    channel_bcs(model.boundary)
    return model


def batch64_draws(seed=0, n_cases=64, n_range=(32, 256), re_range=(100.0, 1000.0)):
    """(N, Re) of every case of the batch, from one seeded generator."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(n_range[0]), np.log(n_range[1])
    N = np.round(np.exp(rng.uniform(lo, hi, n_cases))).astype(int)
    Re = rng.uniform(re_range[0], re_range[1], n_cases)
    return [(int(n), float(r)) for n, r in zip(N, Re)]


def cylinder_bcs(b):
    b.velocity_BC('-X', u=1.0, v=0.0, w=0.0)
    b.pressure_BC('+X', p=0.0)
    for w in ('-Y', '+Y', '-Z', '+Z'):
        b.slip_wall(w)
    b.no_slip_wall('cylinder')


def cylinder_model(h=0.05, lengths=(20.0, 8.0, 1.0), centre=(5.0, 4.0), radius=0.5, re=100.0,
                   jitter=0.1, seed=0, device='cuda'):
    """Re on the diameter: nu = U D / Re with U = 1, D = 2 radius."""
    nodes, cells = cylinder_mesh(h, lengths, centre, radius, jitter=jitter, seed=seed)
    m = Mesh(nodes, cells)
    box_and_body_walls(m, lengths)
    model = FVM(m, device=device)
    model.material.create_incompressible_newtonian_fluid(viscosity=2.0 * radius / float(re),
                                                         density=1.0)
    cylinder_bcs(model.boundary)
    return model
