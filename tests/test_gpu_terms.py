"""The composable term / equation API (warpcfd_amd.FV.terms) on the GPU.

* every term's weights against the oracle's statement of the reference's
  term (convection upwind / central / upwindLinear, diffusion central and
  corrected with scalar or per-face viscosity, ddt, grad), on a jittered
  channel with random fields;
* Equation.form_system + relax + replace_row against the oracle's Equation,
  and solve_Axb against scipy's direct solve of the oracle's matrix;
* the composable momentum system against the fused assembly kernel.
Tolerance: rel-L2 <= 1e-5 per array (float32), 1e-4 for solved fields.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from cases import rel_l2
from test_gpu_parity import _pair, _set_state

pytestmark = pytest.mark.gpu

UVW = ['u', 'v', 'w']


def _state(seed=3, n=(12, 9, 7)):
    of, model = _pair('channel', n=n, jitter=0.1, seed=seed)
    _set_state(model, of, seed)
    of.set_boundary_conditions(); of.face_interpolation(); of.calculate_mass_flux()
    model.set_boundary_conditions()
    model.face_interpolation()
    model.calculate_mass_flux()
    return of, model


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize('scheme', ['upwind', 'central difference', 'upwindLinear'])
def test_convection_weights_vs_oracle(gpu, scheme):
    from oracle.fv import convection_weights
    from warpcfd_amd.FV.terms import ConvectionTerm
    of, model = _state()
    assert rel_l2(_np(model.mass_fluxes), of.mass_fluxes) <= 1e-5
    t = ConvectionTerm(model, UVW, scheme)
    t(model)
    got = _np(t.weights)
    want = convection_weights(of, (0, 1, 2), scheme=scheme)
    for j in range(3):
        assert rel_l2(got[..., j], want[..., j]) <= 1e-5, (scheme, j)


@pytest.mark.parametrize('correction', [False, True])
def test_diffusion_weights_vs_oracle(gpu, correction):
    import torch
    from oracle.fv import diffusion_weights
    from warpcfd_amd.FV.terms import DiffusionTerm
    of, model = _state(seed=4)
    t = DiffusionTerm(model, UVW, correction=correction)
    t(model, 0.013)
    want = diffusion_weights(of, np.float32(0.013), (0, 1, 2), correction=correction)
    got = _np(t.weights)
    for j in range(3):
        assert rel_l2(got[..., j], want[..., j]) <= 1e-5, (correction, j)
    # per-face viscosity (the pressure equation's rUA_f), reference face order
    rng = np.random.default_rng(5)
    nu_f = (0.5 + rng.random(model.num_faces)).astype(np.float32)
    tp = DiffusionTerm(model, 'p', correction=correction)
    tp(model, torch.from_numpy(nu_f))
    want = diffusion_weights(of, nu_f, (3,), correction=correction)
    got = _np(tp.weights)
    for j in range(3):
        assert rel_l2(got[..., j], want[..., j]) <= 1e-5, ('per face', correction, j)


def test_ddt_and_grad_weights_vs_oracle(gpu):
    from oracle.fv import ddt_weights, grad_term_weights
    from warpcfd_amd.FV.terms import DdtTerm, GradTerm
    of, model = _state(seed=6)
    d = DdtTerm(model, UVW)
    d(model, 0.02)
    assert rel_l2(_np(d.weights), ddt_weights(of, 0.02, (0, 1, 2))) <= 1e-6
    g = GradTerm(model, 'p')
    g(model)
    assert rel_l2(_np(g.weights), grad_term_weights(of, 3)) <= 1e-6


def _momentum_pair(of, model, scheme='upwind', dt=None):
    from oracle.fv import (OracleEquation, convection_weights, ddt_weights, diffusion_weights,
                           grad_term_weights)
    from warpcfd_amd.FV.terms import ConvectionTerm, DdtTerm, DiffusionTerm, Equation, GradTerm
    nu = np.float32(0.02)
    conv = ConvectionTerm(model, UVW, scheme)
    diff = DiffusionTerm(model, UVW)
    grad = GradTerm(model, 'p')
    conv(model); diff(model, float(nu)); grad(model)
    implicit = [conv, -diff]
    o_impl = [('face', convection_weights(of, (0, 1, 2), scheme=scheme), 1.0),
              ('face', diffusion_weights(of, nu, (0, 1, 2)), -1.0)]
    if dt is not None:
        ddt = DdtTerm(model, UVW)
        ddt(model, dt)
        implicit = [ddt] + implicit
        o_impl = [('cell', ddt_weights(of, dt, (0, 1, 2)), 1.0)] + o_impl
    eq = Equation(model, UVW)
    eq.form_system(implicit, [-grad])
    oe = OracleEquation(of, [0, 1, 2])
    oe.form_system(o_impl, [(grad_term_weights(of, 3), -1.0)])
    return eq, oe


@pytest.mark.parametrize('scheme,dt', [('upwind', None), ('central difference', 0.05),
                                       ('upwindLinear', None)])
def test_momentum_equation_vs_oracle_and_solve(gpu, scheme, dt):
    of, model = _state(seed=7)
    eq, oe = _momentum_pair(of, model, scheme, dt)
    A_got = eq.matrix.toarray()
    A_want = oe.matrix().toarray()
    assert rel_l2(A_got, A_want) <= 1e-5
    assert rel_l2(_np(eq.rhs), oe.rhs) <= 1e-5
    assert rel_l2(_np(eq.diagonal), oe.diagonal) <= 1e-5
    eq.relax(0.7)
    oe.relax(0.7)
    assert rel_l2(eq.matrix.toarray(), oe.matrix().toarray()) <= 1e-5
    assert rel_l2(_np(eq.rhs), oe.rhs) <= 1e-5
    # solve: three right-hand sides on the shared matrix, against scipy
    eq.tol = 1e-6
    eq.max_iter = 2000
    (iters, res, thr), x = eq.solve_Axb()
    assert iters > 0 and res <= thr
    xs = spla.spsolve(oe.matrix().astype(np.float64).tocsc(), oe.rhs.astype(np.float64))
    assert rel_l2(_np(x), xs) <= 1e-4


def test_pressure_equation_vs_oracle_and_solve(gpu):
    """DiffusionTerm('p') with per-face rUA_f, the divergence added with
    add_RHS, reference row 0 replaced, solved by CG."""
    import torch
    from oracle.fv import OracleEquation, diffusion_weights
    from warpcfd_amd.FV.terms import DiffusionTerm, Equation, linear
    of, model = _state(seed=8)
    rng = np.random.default_rng(9)
    rua_f = (0.5 + rng.random(model.num_faces)).astype(np.float32)
    div = of.divFlux().astype(np.float32)
    d = DiffusionTerm(model, 'p')
    d(model, torch.from_numpy(rua_f))
    eq = Equation(model, 'p', solver=linear.cg)
    eq.form_system([d])
    eq.add_RHS(torch.from_numpy(div))
    eq.replace_row(0, 0.0)
    oe = OracleEquation(of, [3])
    oe.form_system([('face', diffusion_weights(of, rua_f, (3,)), 1.0)])
    oe.add_RHS(div)
    oe.replace_row(0, 0.0)
    assert rel_l2(eq.matrix.toarray(), oe.matrix().toarray()) <= 1e-5
    assert rel_l2(_np(eq.rhs), oe.rhs) <= 1e-5
    eq.tol = 1e-6
    eq.max_iter = 4000
    (iters, res, thr), x = eq.solve_Axb()
    assert iters > 0 and res <= thr
    xs = spla.spsolve(oe.matrix().astype(np.float64).tocsc(), oe.rhs.astype(np.float64))
    assert rel_l2(_np(x), xs) <= 1e-4
    # warm start from a given vector, updated in place
    x0 = torch.zeros(model.num_cells, dtype=torch.float32, device=model._cv.device)
    _, x1 = eq.solve_Axb(x0)
    assert x1 is x0 and rel_l2(_np(x0), xs) <= 1e-4


def test_composable_momentum_equals_fused_assembly(gpu):
    """conv - diff with -grad p, relaxed: the same matrix and right-hand
    sides as the fused wcfd_assemble_momentum kernel."""
    import torch
    from warpcfd_amd.FV.device import ptr, stream_ptr
    of, model = _state(seed=11)
    eq, _ = _momentum_pair(of, model)
    eq.relax(0.7)
    C = model.num_cells
    dev = model._cv.device
    values = torch.zeros(7, C, dtype=torch.float32, device=dev)
    rhs = torch.zeros(3, C, dtype=torch.float32, device=dev)
    code = model.lib.wcfd_assemble_momentum(model.dm.desc_ref, model.bc, ptr(model._cv),
                                            ptr(model._fv), ptr(model._flux), ptr(model._cg), 0,
                                            0.02, 1.0, 0.7, 0.0, 0, ptr(values), ptr(rhs),
                                            stream_ptr())
    assert code == 0
    for o in range(3):
        assert rel_l2(_np(eq.values[o]), _np(values)) <= 1e-6, o
    assert rel_l2(_np(eq._rhs), _np(rhs)) <= 1e-6


def test_term_scale_and_layout(gpu):
    from warpcfd_amd.FV.terms import DiffusionTerm
    _, model = _state(seed=12, n=(5, 4, 3))
    d = DiffusionTerm(model, UVW)
    d(model, 0.5)
    assert d.scale == 1.0 and (-d).scale == -1.0 and (2 * d).scale == 2.0 and d.scale == 1.0
    assert (-d)._w is d._w                       # shallow copies share the weights
    assert tuple(d.weights.shape) == (model.num_cells, 6, 3, 3)
    with pytest.raises(ValueError):
        DiffusionTerm(model, ['q'])
