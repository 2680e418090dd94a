"""GPU runs against the committed oracle fixtures (tests/golden)."""
import os

import numpy as np
import pytest

from cases import device_case, rel_l2

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def test_cavity24_fixed_point_vs_fp64_oracle(gpu):
    """700 iterations with tight Krylov tolerances reach the fixed point: the
    float32 GPU fields agree with the float64 oracle's fixed point to 1e-5
    (the float32 oracle's own distance is ~5e-7)."""
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    from warpcfd_amd.preprocess.grid import cavity_mesh
    z = np.load(os.path.join(G, 'cavity24_converged.npz'))
    nodes, cells = cavity_mesh(24)
    model = device_case(nodes, cells, 'cavity')
    sv = IncompressibleSolver(model, 0.7, 0.3)
    sv.tol = 1e-9
    for _ in range(700):
        d = sv.outer_iteration()
    assert d['continuity'] < 1e-7
    got = model.cell_values.cpu().numpy()
    err = [rel_l2(got[:, q], z['f64'][:, q]) for q in (0, 1, 3)]
    print('rel-L2 vs float64 fixed point (u, v, p):', err)
    assert max(err) <= 1e-5


@pytest.mark.parametrize('rtol0', [0.0, 0.1])
def test_cavity24_monitor_stop_vs_fp64_oracle(gpu, rtol0):
    """Stopped by the convergence monitor (continuity <= 1e-6 and estimated
    velocity error <= 5e-5): within 1e-4 of the float64 fixed point, for the
    reference Krylov setting and for the warm-start-relative one."""
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    from warpcfd_amd.preprocess.grid import cavity_mesh
    z = np.load(os.path.join(G, 'cavity24_converged.npz'))
    nodes, cells = cavity_mesh(24)
    model = device_case(nodes, cells, 'cavity')
    sv = IncompressibleSolver(model, 0.7, 0.3)
    if rtol0:
        sv.tol, sv.rtol, sv.rtol0 = 0.0, 0.0, rtol0
    else:
        sv.tol = 1e-9
    it, done, mon = sv.run_to_convergence(1e-6, 5e-5, 1e-7, max_outer=2000)
    assert done
    got = model.cell_values.cpu().numpy()
    err = [rel_l2(got[:, q], z['f64'][:, q]) for q in (0, 1)]
    print(f'stopped after {it} iterations, monitor {mon.estimate}, rel-L2 (u, v) {err}')
    assert max(err) <= 1e-4


def test_cavity64_500_iterations_reference_settings(gpu):
    """The 64x64 cavity, 500 iterations, reference solver settings
    (examples/LDC/LDC.py set-up).  With atol = 1e-5 each pressure solve
    fixes p only to that tolerance: the smooth error modes of the Laplacian
    (eigenvalues ~1e-6 here) are left unresolved, so two correct
    implementations drift apart in p (tens of percent) while the velocity,
    which sees only grad p, stays within ~1e-2.  Converged fields are compared
    at 1e-4 in the test above; here velocity <= 1e-2 and both runs sit equally
    close to the Ghia data."""
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    from warpcfd_amd.preprocess.grid import cavity_mesh
    z = np.load(os.path.join(G, 'cavity64_ref500.npz'))
    nodes, cells = cavity_mesh(64)
    model = device_case(nodes, cells, 'cavity')
    sv = IncompressibleSolver(model, 0.7, 0.3)
    for _ in range(500):
        sv.outer_iteration()
    got = model.cell_values.cpu().numpy()
    ref = z['cell_values']
    print('cavity64 rel-L2 u,v,p:', [rel_l2(got[:, q], ref[:, q]) for q in (0, 1, 3)])
    assert rel_l2(got[:, :2], ref[:, :2]) <= 1e-2
    from test_golden import _centrelines
    _, ug, vg = _centrelines(got, 64)
    _, ur, vr = _centrelines(ref, 64)
    assert abs(ug.min() - ur.min()) < 2e-3 and abs(vg.max() - vr.max()) < 2e-3


def test_cavity64_fixed_point_vs_fp64_oracle(gpu):
    """The 64 x 64 cavity (the LDC.py set-up) run to its fixed point with tight
    Krylov tolerances: u, v and p within 1e-4 of the float64 oracle run to
    continuity 1e-10 (committed fixture cavity64_converged.npz)."""
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    from warpcfd_amd.preprocess.grid import cavity_mesh
    z = np.load(os.path.join(G, 'cavity64_converged.npz'))
    nodes, cells = cavity_mesh(64)
    model = device_case(nodes, cells, 'cavity')
    sv = IncompressibleSolver(model, 0.7, 0.3)
    sv.tol = 1e-9
    sv.max_iter = 2000
    for it in range(4000):
        d = sv.outer_iteration()
        if it >= int(z['iters64']) and d['continuity'] < 1e-7:
            break
    got = model.cell_values.cpu().numpy()
    err = [rel_l2(got[:, q], z['f64'][:, q]) for q in (0, 1, 3)]
    print(f'cavity64: {it + 1} iterations, continuity {d["continuity"]:.2e}, '
          f'rel-L2 vs float64 fixed point (u, v, p): {err}')
    assert d['continuity'] < 1e-6
    assert max(err) <= 1e-4
