"""The double-field build of the device path (libwcfd64.so, behind
FVM(float_dtype=np.float64) as the reference's float_dtype=wp.float64,
FV/model.py:54-55) against the oracle in float64.

Both sides take the mesh geometry in float32, as the reference does for a
float64 model (preprocess/mesh.py:22-33); the device layout also rounds the
derived geometry (skew vector, |d|, A n) to float32 before casting, where
the oracle forms it in float64, so operators agree to ~1e-7 rather than to
double rounding: tolerance 1e-6 per operator.  Whole runs: the 24 x 24
cavity's fixed point within 3e-6 of the float64 oracle fixture (whose
geometry is float64: its float32 rounding moves p by ~1e-6), transient PISO within 1e-4 every step in u, v and p with
converged Krylov solves on both sides (the bound float32 cannot meet for p,
tests/test_gpu_transient.py), and a float64 batch arena against the
fixture.  The float32 path's distance from the float64 build is printed."""
import os

import numpy as np
import pytest

from cases import device_case, jittered_box, oracle_case, rel_l2

pytestmark = pytest.mark.gpu
F64 = np.float64
TOL = 1e-6
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def _pair64(kind='channel', n=(9, 7, 5), jitter=0.1, seed=1):
    nodes, cells = jittered_box(n, jitter, seed)
    of = oracle_case(nodes, cells, kind, dtype=F64, geom_dtype=np.float32)
    model = device_case(nodes, cells, kind, float_dtype=F64)
    return of, model


def _set_state(model, of, seed=0):
    import torch
    rng = np.random.default_rng(seed)
    C = of.m.num_cells
    cv = rng.standard_normal((C, 4))
    cg = 0.5 * rng.standard_normal((C, 4, 3))
    of.cell_values[:] = cv
    of.cell_gradients[:] = cg
    dev = model._cv.device
    model._cv.copy_(model.dm.cells_to_device(torch.from_numpy(cv.T.copy()).to(dev)))
    model._cg.copy_(model.dm.cells_to_device(torch.from_numpy(cg.reshape(C, 12).T.copy()).to(dev)))
    model._uvw_grad_zero = False


def test_fp64_operators_vs_oracle(gpu):
    import torch
    of, model = _pair64()
    assert model.lib.wcfd_real_bytes() == 8 and model._cv.dtype == torch.float64
    _set_state(model, of, seed=2)
    for x in (of, model):
        x.set_boundary_conditions(); x.face_interpolation(); x.calculate_mass_flux()
    assert rel_l2(model.face_values.numpy(), of.face_values) <= TOL
    assert rel_l2(model.mass_fluxes.numpy(), of.mass_fluxes) <= TOL
    of.calculate_gradients(); model.calculate_gradients()
    assert rel_l2(model.cell_gradients.numpy(), of.cell_gradients) <= TOL
    assert rel_l2(model.divFlux().numpy(), of.divFlux()) <= TOL


def test_fp64_assembly_vs_oracle(gpu):
    """The composable terms and Equation in float64: momentum system
    (convection - diffusion, -grad p, relaxed) against the oracle's."""
    from oracle.fv import OracleEquation, convection_weights, diffusion_weights, grad_term_weights
    from warpcfd_amd.FV.terms import ConvectionTerm, DiffusionTerm, Equation, GradTerm
    of, model = _pair64(seed=4)
    _set_state(model, of, seed=4)
    for x in (of, model):
        x.set_boundary_conditions(); x.face_interpolation(); x.calculate_mass_flux()
    of.calculate_gradients(); model.calculate_gradients()
    uvw = ['u', 'v', 'w']
    conv, diff, grad = ConvectionTerm(model, uvw), DiffusionTerm(model, uvw), GradTerm(model, 'p')
    conv(model); diff(model, viscosity=0.013); grad(model)
    eq = Equation(model, fields=uvw)
    assert eq.values.dtype.itemsize == 8 and eq.tol == 1e-6     # equation.py:35-39 (float64)
    eq.form_system([conv, -diff], explicit_terms=-1.0 * grad, fvm=model)
    eq.relax(0.7, model)
    oe = OracleEquation(of, [0, 1, 2])
    oe.form_system([('face', convection_weights(of), 1.0),
                    ('face', diffusion_weights(of, 0.013, (0, 1, 2)), -1.0)],
                   explicit=[(grad_term_weights(of), -1.0)])
    oe.relax(0.7)
    assert rel_l2(eq.matrix.toarray(), oe.matrix().toarray()) <= TOL
    assert rel_l2(eq.rhs.numpy(), oe.rhs) <= TOL


def _cavity24(float_dtype, iters=700):
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    from warpcfd_amd.preprocess.grid import cavity_mesh
    nodes, cells = cavity_mesh(24)
    model = device_case(nodes, cells, 'cavity', float_dtype=float_dtype)
    sv = IncompressibleSolver(model, 0.7, 0.3)
    sv.tol = 1e-13 if float_dtype == F64 else 1e-9
    for _ in range(iters):
        d = sv.outer_iteration()
    return model.cell_values.numpy(), d


def test_fp64_cavity24_fixed_point(gpu):
    z = np.load(os.path.join(G, 'cavity24_converged.npz'))
    got64, d = _cavity24(F64)
    assert d['continuity'] < 1e-10
    err = [rel_l2(got64[:, q], z['f64'][:, q]) for q in (0, 1, 3)]
    got32, _ = _cavity24(np.float32)
    err32 = [rel_l2(got32[:, q], got64[:, q]) for q in (0, 1, 3)]
    print(f'cavity24: float64 build vs float64 oracle fixture (u, v, p) {err}; '
          f'float32 path vs float64 build {err32}')
    assert max(err) <= 3 * TOL
    assert max(err32) <= 1e-5


def test_fp64_piso_cylinder_vs_oracle(gpu):
    """Five PISO steps past the small voxel-carved cylinder, both sides in
    float64 with converged Krylov solves: u, v and p within 1e-4 at every
    step (measured far below)."""
    from oracle.fv import OracleFVM
    from oracle.mesh import OracleMesh
    from oracle.simple import OracleSolver
    from test_gpu_transient import DT, GEOM, STEPS
    from warpcfd_amd.FV import FVM
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    from warpcfd_amd.preprocess import Mesh, box_and_body_walls, cylinder_mesh
    from warpcfd_amd.workloads import cylinder_bcs
    nodes, cells = cylinder_mesh(GEOM['h'], GEOM['lengths'], GEOM['centre'], GEOM['radius'],
                                 jitter=0.1, seed=5)
    m = Mesh(nodes, cells)
    walls = box_and_body_walls(m, GEOM['lengths'])
    model = FVM(m, float_dtype=F64)
    model.material.create_incompressible_newtonian_fluid(viscosity=0.01, density=1.0)
    cylinder_bcs(model.boundary)
    model.finalize()
    sv = IncompressibleSolver(model, None, None, steady_state=False, implicit_outflow=True)
    sv.set_iterations(1, 2, 1)
    sv.tol, sv.max_iter = 1e-12, 4000
    om = OracleMesh(nodes, cells, geom_dtype=np.float32)
    om.groups = {k: np.asarray(v) for k, v in walls.items()}
    of = OracleFVM(om, F64)
    cylinder_bcs(of.boundary)
    of.boundary.finalize()
    os_ = OracleSolver(of, None, None, steady_state=False, nu=0.01, implicit_outflow=True,
                       atol=1e-12, maxiter=4000)
    os_.set_iterations(1, 2, 1)
    for step in range(STEPS):
        sv.outer_iteration(dt=DT)
        os_.outer_iteration(DT)
        got = model.cell_values.numpy()
        err = [rel_l2(got[:, q], of.cell_values[:, q]) for q in (0, 1, 3)]
        print(f'float64 PISO step {step}: rel-L2 u, v, p = {err}')
        assert max(err) <= 1e-4, (step, err)


def test_fp64_arena_case_fixed_point(gpu):
    """A case in a float64 batch arena (beside another) reaches the float64
    oracle's fixed point of the 24 x 24 cavity within 3e-6."""
    import torch
    from test_gpu_batch import cavity_spec
    from warpcfd_amd.batch import CaseArena
    z = np.load(os.path.join(G, 'cavity24_converged.npz'))
    specs = [cavity_spec(24, 100.0, stop=0.0, max_outer=700), cavity_spec(12, 400.0, stop=0.0,
                                                                          max_outer=700)]
    arena = CaseArena.for_cases(specs, chunk=128, block=128, atol=1e-13, tol=1e-13, maxiter=4000,
                                dtype=torch.float64)
    assert arena.lib.wcfd_real_bytes() == 8 and arena.cv.dtype == torch.float64
    arena.run(max_steps=701)
    got = arena.cell_values(0)
    err = [rel_l2(got[:, q], z['f64'][:, q]) for q in (0, 1, 3)]
    print(f'float64 arena cavity24 vs float64 oracle fixture (u, v, p): {err}')
    assert max(err) <= 3 * TOL
