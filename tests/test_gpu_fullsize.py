"""Parity at the full size of the benchmark channel: 10 x 2 x 2 box,
200 x 40 x 40 hexahedra with jittered interior vertices (seed 0), 320 000
cells, 942 400 internal and 35 200 boundary faces, inlet u = 1 on -X,
pressure outlet on +X, no-slip walls.

At this size every launch spans hundreds to thousands of workgroups, so the
XCD-aware workgroup numbering, the grid-stride tails and the int16
neighbour offsets of the renumbered mesh (bandwidth ~1 600 cells) run at the
shapes the benchmark uses; the smaller parity tests do not reach them.
* each operator after a seeded random state -- boundary values, face
  interpolation with skew correction, mass flux, Green-Gauss gradients,
  divergence, the momentum and pressure matrices and right-hand sides --
  within rel-L2 1e-5 of the oracle (float32, different summation order);
* two SIMPLE iterations with the reference solver settings (Krylov atol
  1e-5, maxiter 500, check every 10): velocity and pressure within 1e-4
  (measured 2.8e-5 and 1.6e-5) and the continuity residual within 1e-4
  relative (measured 5e-6); the device run repeated is bitwise identical.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from cases import device_case, oracle_case, rel_l2
from test_gpu_parity import KERNEL_TOL, _dev_matrix_ref, _set_state

pytestmark = pytest.mark.gpu

DIMS, LENGTHS = (200, 40, 40), (10.0, 2.0, 2.0)


@pytest.fixture(scope='module')
def channel():
    from warpcfd_amd.preprocess import channel_mesh
    nodes, cells = channel_mesh(*DIMS, LENGTHS, jitter=0.1, seed=0)
    assert len(cells) == 320000
    return nodes, cells


@pytest.fixture(scope='module')
def pair(channel, gpu):
    nodes, cells = channel
    of = oracle_case(nodes, cells, 'channel', ref_cell=None)
    model = device_case(nodes, cells, 'channel', ref_cell=None)
    assert of.m.num_faces == 942400 + 35200
    return of, model


def _sparse_rel(A, B):
    return spla.norm(A - B) / spla.norm(B)


def test_face_and_cell_operators_full_size(pair):
    of, model = pair
    _set_state(model, of, seed=11)
    for x in (of, model):
        x.set_boundary_conditions(); x.face_interpolation(); x.calculate_mass_flux()
    assert rel_l2(model.face_values.cpu().numpy(), of.face_values) <= KERNEL_TOL
    assert rel_l2(model.mass_fluxes.cpu().numpy(), of.mass_fluxes) <= KERNEL_TOL
    of.calculate_gradients(); model.calculate_gradients()
    assert rel_l2(model.cell_gradients.cpu().numpy(), of.cell_gradients) <= KERNEL_TOL
    assert rel_l2(model.divFlux().cpu().numpy(), of.divFlux()) <= KERNEL_TOL


def test_assembly_full_size(pair):
    import torch
    from oracle.fv import OracleEquation, convection_weights, diffusion_weights, grad_term_weights
    from warpcfd_amd.FV.device import ptr, stream_ptr
    of, model = pair
    _set_state(model, of, seed=12)
    for x in (of, model):
        x.set_boundary_conditions(); x.face_interpolation(); x.calculate_mass_flux()
    of.calculate_gradients()
    C = of.m.num_cells
    dev = model._cv.device
    # momentum: upwind convection, Laplacian, -grad p, Patankar relaxation, implicit outflow
    nu, alpha = 0.01, 0.7
    eq = OracleEquation(of, [0, 1, 2])
    eq.form_system([('face', convection_weights(of, implicit_outflow=True), 1.0),
                    ('face', diffusion_weights(of, nu, (0, 1, 2)), -1.0)],
                   explicit=[(grad_term_weights(of), -1.0)])
    eq.relax(alpha)
    vals = torch.zeros(7, C, dtype=torch.float32, device=dev)
    rhs = torch.zeros(3, C, dtype=torch.float32, device=dev)
    assert model.lib.wcfd_assemble_momentum(
        model.dm.desc_ref, model.bc, ptr(model._cv), ptr(model._fv), ptr(model._flux),
        ptr(model._cg), 1, nu, 1.0, alpha, 0.0, 2, ptr(vals), ptr(rhs), stream_ptr()) == 0
    torch.cuda.synchronize()
    A_ref = eq.matrix().tocsr()[0::3, 0::3]
    assert _sparse_rel(_dev_matrix_ref(model, vals), A_ref) <= KERNEL_TOL
    assert rel_l2(model.dm.cells_to_reference(rhs).cpu().numpy(),
                  eq.rhs.reshape(C, 3).T) <= KERNEL_TOL
    assert rel_l2(model.cell_gradients.cpu().numpy(), of.cell_gradients) <= KERNEL_TOL
    # pressure: Laplacian of a seeded positive face field = divergence of the flux
    rua_f = (0.1 + np.random.default_rng(13).random(of.m.num_faces)).astype(np.float32)
    eq = OracleEquation(of, [3])
    eq.form_system([('face', diffusion_weights(of, rua_f, (3,)), 1.0)])
    eq.add_RHS(of.divFlux())
    pv = torch.zeros(7, C, dtype=torch.float32, device=dev)
    prhs = torch.zeros(C, dtype=torch.float32, device=dev)
    rua_dev = model.dm.faces_to_device(torch.from_numpy(rua_f).to(dev)).contiguous()
    assert model.lib.wcfd_assemble_pressure(
        model.dm.desc_ref, model.bc, ptr(model._fv), ptr(model._flux), ptr(rua_dev),
        ptr(model._cg), 0, 0, model.reference_cell_device, 0.0, ptr(pv), ptr(prhs), None,
        stream_ptr()) == 0
    torch.cuda.synchronize()
    assert _sparse_rel(_dev_matrix_ref(model, pv), eq.matrix().tocsr()) <= KERNEL_TOL
    assert rel_l2(model.dm.cells_to_reference(prhs).cpu().numpy(), eq.rhs) <= KERNEL_TOL


def test_simple_full_size_vs_oracle(channel, gpu):
    from oracle.simple import OracleSolver
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    nodes, cells = channel
    of = oracle_case(nodes, cells, 'channel', ref_cell=None)
    os_ = OracleSolver(of, 0.7, 0.3, nu=0.01, implicit_outflow=True)
    runs = []
    for _ in range(2):
        model = device_case(nodes, cells, 'channel', ref_cell=None)
        sv = IncompressibleSolver(model, 0.7, 0.3, implicit_outflow=True)
        for _ in range(2):
            sv.outer_iteration()
        runs.append((model.cell_values.cpu().numpy(), sv.history[-1]['continuity']))
        del sv, model
    for _ in range(2):
        os_.outer_iteration()
    got, cont = runs[0]
    assert np.array_equal(runs[1][0], got)
    err_u = rel_l2(got[:, :3], of.cell_values[:, :3])
    err_p = rel_l2(got[:, 3], of.cell_values[:, 3])
    ref_cont = os_.history[-1]['continuity']
    print(f'320k channel, 2 SIMPLE iterations: velocity rel-L2 {err_u:.2e}, p {err_p:.2e}, '
          f'continuity {cont:.6e} vs {ref_cont:.6e}')
    assert err_u <= 1e-4 and err_p <= 1e-4
    assert abs(cont - ref_cont) <= 1e-4 * ref_cont
