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
