"""Transient PISO (incompressible.py:91-156 with steady_state=False, two
inner correctors, no relaxation) past a voxel-carved cylinder: the GPU path
against the oracle step by step, with and without the float64 pressure
refinement."""
import numpy as np
import pytest

from cases import rel_l2

pytestmark = pytest.mark.gpu

GEOM = dict(h=0.125, lengths=(6.0, 4.0, 0.5), centre=(2.0, 2.0), radius=0.5)
DT = 0.5 * GEOM['h'] / 1.0          # CFL 0.5 on the inflow velocity
STEPS = 5


def _oracle(nodes, cells, walls, p_refine):
    from oracle.mesh import OracleMesh
    from oracle.fv import OracleFVM
    from oracle.simple import OracleSolver
    from warpcfd_amd.workloads import cylinder_bcs
    om = OracleMesh(nodes, cells)
    om.groups = {k: np.asarray(v) for k, v in walls.items()}
    of = OracleFVM(om, np.float32)
    cylinder_bcs(of.boundary)
    of.boundary.finalize()
    s = OracleSolver(of, None, None, steady_state=False, nu=0.01, implicit_outflow=True,
                     p_refine=p_refine)
    s.set_iterations(1, 2, 1)
    s.set_time_stepping(STEPS, DT)
    return of, s


@pytest.mark.parametrize('p_refine', [0, 1])
def test_piso_cylinder_vs_oracle(gpu, p_refine):
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    from warpcfd_amd.preprocess import Mesh, box_and_body_walls, cylinder_mesh
    from warpcfd_amd.FV import FVM
    from warpcfd_amd.workloads import cylinder_bcs
    nodes, cells = cylinder_mesh(GEOM['h'], GEOM['lengths'], GEOM['centre'], GEOM['radius'],
                                 jitter=0.1, seed=5)
    m = Mesh(nodes, cells)
    walls = box_and_body_walls(m, GEOM['lengths'])
    model = FVM(m)
    model.material.create_incompressible_newtonian_fluid(viscosity=0.01, density=1.0)
    cylinder_bcs(model.boundary)
    model.finalize()
    sv = IncompressibleSolver(model, None, None, steady_state=False, implicit_outflow=True)
    sv.set_iterations(1, 2, 1)
    sv.set_time_stepping(STEPS, DT)
    sv.p_refine = p_refine
    of, os_ = _oracle(nodes, cells, walls, p_refine)
    for step in range(STEPS):
        d = sv.outer_iteration(dt=DT)
        os_.outer_iteration(DT)
        got = model.cell_values.cpu().numpy()
        ref = of.cell_values
        err = [rel_l2(got[:, q], ref[:, q]) for q in (0, 1, 3)]
        print(f'step {step}: rel-L2 u, v, p = {err}, p iterations {d["p_iters"]}')
        assert err[0] <= 1e-4 and err[1] <= 1e-3 and err[2] <= 1e-3, (step, err)
    # the flow has developed round the cylinder
    assert np.abs(got[:, 1]).max() > 0.05


@pytest.mark.parametrize('p_refine', [0, 1])
def test_piso_cylinder_tight_krylov_vs_oracle(gpu, p_refine):
    """The same five PISO steps with every Krylov solve converged (atol 1e-9,
    up to 4000 iterations) on the device and in the float32 oracle, plus a
    float64 oracle run (atol 1e-12) as the exact answer.  u and v agree with
    the float32 oracle within 1e-4 at every step.  p does not get closer with
    tight solves: the pressure equation's condition number (about 2e3 here)
    amplifies the float32 rounding of its assembled coefficients, and the
    float32 oracle itself sits 6e-4 (first step from rest) to 5e-5 from the
    float64 p.  So the device's p is held to that floor: with the float64
    refinement (p_refine 1), no further from the float64 p than the float32
    oracle is, or within 1e-4, at every step (measured: 2.4e-4 at the first
    step, 3e-5 to 4e-5 after, against the oracle's 4.4e-4 and 2e-5 to
    9e-5); without it the device and the float32 oracle are two independent
    samples of the float32 floor (1.2e-4 against 5e-5 at step 1), held to
    twice that floor."""
    from oracle.fv import OracleFVM
    from oracle.mesh import OracleMesh
    from oracle.simple import OracleSolver
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    from warpcfd_amd.preprocess import Mesh, box_and_body_walls, cylinder_mesh
    from warpcfd_amd.FV import FVM
    from warpcfd_amd.workloads import cylinder_bcs
    nodes, cells = cylinder_mesh(GEOM['h'], GEOM['lengths'], GEOM['centre'], GEOM['radius'],
                                 jitter=0.1, seed=5)
    m = Mesh(nodes, cells)
    walls = box_and_body_walls(m, GEOM['lengths'])
    model = FVM(m)
    model.material.create_incompressible_newtonian_fluid(viscosity=0.01, density=1.0)
    cylinder_bcs(model.boundary)
    model.finalize()
    sv = IncompressibleSolver(model, None, None, steady_state=False, implicit_outflow=True)
    sv.set_iterations(1, 2, 1)
    sv.set_time_stepping(STEPS, DT)
    sv.p_refine = p_refine
    sv.tol, sv.max_iter = 1e-9, 4000
    of, os_ = _oracle(nodes, cells, walls, p_refine)
    os_.atol, os_.maxiter = 1e-9, 4000
    om64 = OracleMesh(np.asarray(nodes, np.float32).astype(np.float64), cells,
                      geom_dtype=np.float64)
    om64.groups = {k: np.asarray(v) for k, v in walls.items()}
    of64 = OracleFVM(om64, np.float64)
    cylinder_bcs(of64.boundary)
    of64.boundary.finalize()
    os64 = OracleSolver(of64, None, None, steady_state=False, nu=0.01, implicit_outflow=True,
                        atol=1e-12, maxiter=4000)
    os64.set_iterations(1, 2, 1)
    for step in range(STEPS):
        d = sv.outer_iteration(dt=DT)
        os_.outer_iteration(DT)
        os64.outer_iteration(DT)
        got = model.cell_values.cpu().numpy()
        err = [rel_l2(got[:, q], of.cell_values[:, q]) for q in (0, 1, 3)]
        p_dev = rel_l2(got[:, 3], of64.cell_values[:, 3])
        p_f32 = rel_l2(of.cell_values[:, 3], of64.cell_values[:, 3])
        print(f'tight, p_refine {p_refine}, step {step}: vs float32 oracle u, v, p = {err}; '
              f'p vs float64: device {p_dev:.2e}, float32 oracle {p_f32:.2e}; '
              f'p iterations {d["p_iters"]}')
        assert err[0] <= 1e-4 and err[1] <= 1e-4, (step, err)
        floor = max(1e-4, p_f32) * (1.0 if p_refine else 2.0)
        assert p_dev <= floor, (step, p_dev, p_f32)
