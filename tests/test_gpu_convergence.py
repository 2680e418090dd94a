"""The device convergence report (wcfd_convergence_report, behind
FVM.check_convergence / IncompressibleSolver.check_convergence) against the
oracle's restatement of check_convergence (model.py:278-315):
* on the device's own state after five SIMPLE iterations -- the oracle
  function applied to the same matrix, right-hand side, velocity, rUA grad p,
  pressure and fluxes, converted to reference order: every norm within 1e-9
  relative (float64 sums in another order);
* against an independent oracle run of the same case (reference Krylov
  settings, so the runs differ by the CG stops): every norm within 1e-4
  (measured 1e-6 to 2e-6), the same verdicts;
on a jittered hexahedral channel and a tetrahedral closed cavity.  Also the
reference's run(): a check every steps_per_check steps and after the last,
logged in model.Convergence; and the per-case report of a batch arena
(wcfd_batch_convergence_report)."""
import numpy as np
import pytest
import scipy.sparse as sp

from cases import device_case, jittered_box, oracle_case, rel_l2
from test_gpu_parity import _dev_matrix_ref

pytestmark = pytest.mark.gpu


def _cases():
    from warpcfd_amd.preprocess.grid import cavity_mesh, split_hexes
    nodes, cells = jittered_box((12, 6, 5), 0.1, 13, lengths=(2.0, 1.0, 0.8))
    yield 'hex channel', nodes, cells, 'channel', None, True
    nodes, hexes = cavity_mesh(8)
    yield 'tetra cavity', nodes, split_hexes(hexes, 'tetra'), 'closed_cavity', 0, False


def _close(a, b, tol):
    return abs(a - b) <= tol * max(abs(b), 1e-30)


@pytest.mark.parametrize('case', list(_cases()), ids=lambda c: c[0])
def test_report_vs_oracle(gpu, case):
    from oracle.simple import OracleSolver, convergence_report, convergence_verdicts
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    name, nodes, cells, kind, ref_cell, outflow = case
    model = device_case(nodes, cells, kind, ref_cell=ref_cell)
    sv = IncompressibleSolver(model, 0.7, 0.3, implicit_outflow=outflow)
    of = oracle_case(nodes, cells, kind, ref_cell=ref_cell)
    os_ = OracleSolver(of, 0.7, 0.3, nu=0.01, implicit_outflow=outflow)
    for _ in range(5):
        sv.outer_iteration()
        os_.outer_iteration()
    got = model.convergence_norms(sv.mom_values, None, sv.mom_rhs, (sv.rUA, model._cg[9:12]),
                                  sv.p_x)
    # the oracle function on the device's own state, in reference order
    dm = model.dm
    A = sp.kron(_dev_matrix_ref(model, sv.mom_values), sp.identity(3), format='csr')
    u = model.cell_values.cpu().numpy()[:, :3].reshape(-1)
    b = dm.cells_to_reference(sv.mom_rhs).cpu().numpy().T.reshape(-1)
    rua = dm.cells_to_reference(sv.rUA).cpu().numpy()
    grad_p = model.cell_gradients.cpu().numpy()[:, 3]
    corr = (rua[:, None] * grad_p).astype(np.float32).reshape(-1)
    p = dm.cells_to_reference(sv.p_x).cpu().numpy()
    of_same = oracle_case(nodes, cells, kind, ref_cell=ref_cell)
    of_same.mass_fluxes[:] = model.mass_fluxes.cpu().numpy()
    same = convergence_report(A, u, b, of_same.divFlux(), corr, p)
    for k in range(3):
        assert _close(got['r'][k], same['r'][k], 1e-9), (k, got, same)
        assert _close(got['b'][k], same['b'][k], 1e-9), (k, got, same)
    for key in ('vel_corr', 'p', 'div'):
        assert _close(got[key], same[key], 1e-9), (key, got, same)
    # an independent oracle run of the same case
    ref = os_.convergence_report()
    print(f'{name}: device {got}\n{name}: oracle {ref}')
    for k in range(3):
        if ref['b'][k] > 1e-6:
            assert _close(got['r'][k], ref['r'][k], 1e-4), (k, got, ref)
            assert _close(got['b'][k], ref['b'][k], 1e-4), (k, got, ref)
    for key in ('vel_corr', 'p', 'div'):
        assert _close(got[key], ref[key], 1e-4), (key, got, ref)
    verdict_ref, conv_ref = convergence_verdicts(ref)
    assert sv.check_convergence(verbose=False) == conv_ref
    logged = model.Convergence.log_list[-1]
    for n, (_, ok) in verdict_ref.items():
        assert logged[n][1] == ok, n
    assert rel_l2(model.cell_values.cpu().numpy(), of.cell_values) <= 1e-3


def test_run_checks_like_the_reference(gpu):
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    from warpcfd_amd.preprocess.grid import cavity_mesh
    nodes, cells = cavity_mesh(12)
    model = device_case(nodes, cells, 'cavity')
    sv = IncompressibleSolver(model, 0.7, 0.3)
    sv.set_iterations(25, 1, 1)
    sv.run(steps_per_check=10)
    # steps 0, 10, 20 and the final check
    assert len(model.Convergence.log_list) == 4
    last = model.Convergence.log_list[-1]
    assert set(last) == {'momentum_x', 'momentum_y', 'momentum_z', 'continuity',
                         'local_continuity', 'velocity_correction', 'pressure_correction'}
    assert last['continuity'][0] == last['local_continuity'][0]
    assert all(np.isfinite(v[0]) for v in last.values())


def _case_inputs(arena, k):
    """Case k's momentum matrix (3C x 3C), u, b, rUA grad p, p and
    sum |div F| from the arena's buffers, in the case's reference order."""
    sl = arena.slots[k]
    spec, c0 = sl['spec'], sl['c0']
    L, C = spec.layout, spec.n_cells
    cols = slice(c0, c0 + C)
    # the ELL rows as COO (boundary slots add their zero coefficient to the
    # diagonal), renumbered to the reference order
    vals = arena.mom_values[:, cols].cpu().numpy()
    nbr = arena.slot_nbr[:, cols].cpu().numpy().astype(np.int64) - c0
    n2o = np.argsort(L.old2new)
    rows = np.tile(np.arange(C), 7)
    cix = np.concatenate([np.arange(C)] + [nbr[s] for s in range(6)])
    A = sp.csr_matrix((vals.reshape(-1), (n2o[rows], n2o[cix])), shape=(C, C))
    ref = lambda t: t[..., L.old2new]
    u = arena.cell_values(k)[:, :3].reshape(-1)
    b = ref(arena.mom_rhs[:, cols].cpu().numpy()).T.reshape(-1)
    rua = ref(arena.rua[cols].cpu().numpy())
    gp = ref(arena.cg[9:12, cols].cpu().numpy()).T
    corr = (rua[:, None] * gp).astype(np.float32).reshape(-1)
    p = ref(arena.p_x[cols].cpu().numpy())
    # divergence over the arena's slot tables (local-face order, float32)
    flux = arena.flux.cpu().numpy()
    sf = arena.slot_face[:, cols].cpu().numpy()
    side = arena.slot_side[:, cols].cpu().numpy()
    div = np.zeros(C, np.float32)
    for s in range(6):
        f = flux[sf[s]]
        div = np.where(side[s] == 0, div + f, div - f)
    return sp.kron(A, sp.identity(3), format='csr'), u, b, div, corr, p


def test_batch_report_per_case(gpu):
    """Per-case reports of a batch arena: a case's report is bitwise the same
    alone and third in a mix (other offsets and neighbours), and its norms
    equal the oracle function on the case's own arena state within 1e-9."""
    from oracle.simple import convergence_report
    from test_gpu_batch import KW, cavity_spec, channel_spec
    from warpcfd_amd.batch import CaseArena
    x = cavity_spec(16, 400.0)
    alone = CaseArena.for_cases([x], **KW)
    mix = CaseArena.for_cases([cavity_spec(12, 800.0), channel_spec(), x], **KW)
    for _ in range(6):
        alone.step()
        mix.step()
    ra, rm = alone.convergence_report(), mix.convergence_report()
    assert ra[0] == rm[2]
    for k in range(3):
        got = rm[k]
        same = convergence_report(*_case_inputs(mix, k))
        for j in range(3):
            assert _close(got['r'][j], same['r'][j], 1e-9), (k, j, got, same)
            assert _close(got['b'][j], same['b'][j], 1e-9), (k, j, got, same)
        for key in ('vel_corr', 'p', 'div'):
            assert _close(got[key], same[key], 1e-9), (k, key, got, same)


def test_check_convergence_takes_the_reference_call(gpu):
    """model.check_convergence(vel_equation.matrix, HbyA, vel_equation.rhs,
    div_u, vel_correction.flatten(), p_new) -- incompressible.py:163, with the
    reference-order objects of the composable API (a scipy matrix on rows
    3c + k, interleaved (3C,) vectors, a (C,) pressure) -- gives the oracle
    report of those same arrays; so do the Equation object itself and the
    (C, 3) forms.  log=True sets model.res = |p_new| as the reference does."""
    import torch
    from oracle.simple import convergence_report, convergence_verdicts
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    from warpcfd_amd.FV.terms import ConvectionTerm, DiffusionTerm, Equation, GradTerm
    nodes, cells = jittered_box((8, 6, 4), 0.1, 21, lengths=(1.0, 0.8, 0.5))
    model = device_case(nodes, cells, 'cavity')
    sv = IncompressibleSolver(model, 0.7, 0.3)
    for _ in range(4):
        sv.outer_iteration()
    # the momentum system of the current state, term by term (incompressible.py:95-112)
    model.set_boundary_conditions()
    model.face_interpolation()
    model.calculate_gradients()
    model.calculate_mass_flux()
    uvw = ['u', 'v', 'w']
    conv, diff, grad = ConvectionTerm(model, uvw, 'upwind'), DiffusionTerm(model, uvw), \
        GradTerm(model, 'p')
    conv(model)
    diff(model, viscosity=0.01)
    grad(model)
    vel_equation = Equation(model, fields=uvw)
    vel_equation.form_system([conv, -diff], explicit_terms=-1.0 * grad, fvm=model)
    vel_equation.relax(0.7, model)
    C = model.num_cells
    HbyA = model.cell_values[:, :3].reshape(-1).contiguous()          # interleaved (3C,)
    rua_ref = model.dm.cells_to_reference(sv.rUA)
    vel_correction = model.get_gradient('p', coeff=rua_ref)             # (C, 3)
    p_new = model.dm.cells_to_reference(sv.p_x).contiguous()            # (C,)
    div_u = torch.zeros(C, dtype=torch.float32, device=HbyA.device)
    A = vel_equation.matrix
    assert A.shape == (3 * C, 3 * C)
    expect = convergence_report(A, HbyA.cpu().numpy(), vel_equation.rhs.cpu().numpy(),
                                model.divFlux().cpu().numpy(),
                                vel_correction.cpu().numpy().reshape(-1), p_new.cpu().numpy())
    forms = [(A, HbyA, vel_equation.rhs, vel_correction.flatten()),
             (vel_equation, HbyA.reshape(C, 3), vel_equation.rhs.reshape(C, 3).cpu().numpy(),
              vel_correction)]
    for mat, u, b, corr in forms:
        got = model.convergence_norms(mat, u, b, corr, p_new)
        for k in range(3):
            assert _close(got['r'][k], expect['r'][k], 1e-9), (k, got, expect)
            assert _close(got['b'][k], expect['b'][k], 1e-9), (k, got, expect)
        for key in ('vel_corr', 'p', 'div'):
            assert _close(got[key], expect[key], 1e-9), (key, got, expect)
    verdicts, conv_ok = convergence_verdicts(expect)
    ok = model.check_convergence(A, HbyA, vel_equation.rhs, div_u, vel_correction.flatten(), p_new,
                                 verbose=False)
    assert ok == conv_ok
    logged = model.Convergence.log_list[-1]
    for name, (_, v) in verdicts.items():
        assert logged[name][1] == v, name
    np.testing.assert_array_equal(model.res, np.abs(p_new.cpu().numpy()))


def test_true_verdicts_at_the_fixed_point(gpu):
    """From the float32 fixed point of the 24 x 24 cavity (committed fixture),
    three more SIMPLE iterations on the device and in the oracle: the device
    report's verdicts equal the oracle's criterion by criterion, and the
    momentum criteria pass (True) on both."""
    import os
    from oracle.simple import OracleSolver, convergence_verdicts
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    from warpcfd_amd.preprocess.grid import cavity_mesh
    fix = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden',
                               'cavity24_converged.npz'))
    nodes, cells = cavity_mesh(24)
    model = device_case(nodes, cells, 'cavity')
    model.set_initial_conditions(fix['f32'])
    sv = IncompressibleSolver(model, 0.7, 0.3)
    of = oracle_case(nodes, cells, 'cavity')
    of.cell_values[:] = fix['f32']
    os_ = OracleSolver(of, 0.7, 0.3, nu=0.01)
    for _ in range(3):
        sv.outer_iteration()
        os_.outer_iteration()
    verdicts, conv_ok = convergence_verdicts(os_.convergence_report())
    ok = sv.check_convergence(verbose=False)
    logged = model.Convergence.log_list[-1]
    print({n: (logged[n], v) for n, v in verdicts.items()})
    for name, (_, v) in verdicts.items():
        assert logged[name][1] == v, (name, logged[name], verdicts[name])
    for name in ('momentum_x', 'momentum_y'):
        assert verdicts[name][1] and logged[name][1], name
    assert ok == conv_ok


def test_batch_report_of_a_finished_case(gpu):
    """A case that the device monitor has stopped (poll() reports it
    finished, its workgroups dropped from the stepping launches) is still
    reported from its final state until it is released: the norms equal the
    oracle function on that state within 1e-9."""
    from oracle.simple import convergence_report
    from test_gpu_batch import KW, cavity_spec
    from warpcfd_amd.batch import CaseArena
    quick = cavity_spec(8, 100.0, max_outer=20)         # stopped by its iteration cap
    slow = cavity_spec(12, 400.0)
    arena = CaseArena.for_cases([quick, slow], **KW)
    done = []
    for _ in range(40):
        arena.step()
        done += arena.poll()
        if 0 in done:
            break
    assert 0 in done and 1 not in done
    rep = arena.convergence_report()
    assert rep[0] is not None and rep[1] is not None
    same = convergence_report(*_case_inputs(arena, 0))
    for j in range(3):
        assert _close(rep[0]['r'][j], same['r'][j], 1e-9), (j, rep[0], same)
        assert _close(rep[0]['b'][j], same['b'][j], 1e-9), (j, rep[0], same)
    for key in ('vel_corr', 'p', 'div'):
        assert _close(rep[0][key], same[key], 1e-9), (key, rep[0], same)
