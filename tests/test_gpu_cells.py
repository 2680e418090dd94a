"""Tetrahedron and wedge meshes on the device path, against the oracle.

Jittered boxes split into Kuhn tetrahedra or prisms.  The kernels keep six
slots per cell; the slots a tetrahedron (4 faces) or wedge (5) does not
have point at the layout's null face, so these tests check that it adds
nothing anywhere:
* face values and fluxes, Green-Gauss gradients and divergence;
* the momentum matrix and right-hand sides (with and without the corrected
  Laplacian), and the pressure matrix, in the reference's CSR pattern;
* composable term weights (the reference's (C, faces per cell, O, 3)
  layout) and the Equation built from them;
* SIMPLE iterations on a channel, and a closed cavity converged to a
  continuity residual below 1e-7 (u, v, p within 1e-4);
* transient PISO step by step;
* the Krylov kernels' slot count (4 or 5) against a padded six-slot read;
* hexahedral, tetrahedral and wedge cases in one batch arena.
Tolerances: rel-L2 <= 1e-5 per kernel output (float32), 1e-4 for fields.
"""
import numpy as np
import pytest

from cases import device_case, oracle_case, rel_l2
from test_gpu_parity import KERNEL_TOL, _dev_matrix_ref, _set_state

pytestmark = pytest.mark.gpu

TYPES = ['tetra', 'wedge']


def _mesh(cell_type, n=(5, 4, 3), jitter=0.1, seed=1, lengths=(1.0, 0.8, 0.6)):
    from warpcfd_amd.preprocess.grid import structured_cells
    return structured_cells(n, lengths, jitter=jitter, seed=seed, cell_type=cell_type)


def _pair(cell_type, kind='channel', seed=1, ref_cell=0, **kw):
    nodes, cells = _mesh(cell_type, seed=seed, **kw)
    return (oracle_case(nodes, cells, kind, ref_cell=ref_cell),
            device_case(nodes, cells, kind, ref_cell=ref_cell))


def _prepare(of, model, seed):
    _set_state(model, of, seed=seed)
    for x in (of, model):
        x.set_boundary_conditions(); x.face_interpolation(); x.calculate_mass_flux()


@pytest.mark.parametrize('cell_type', TYPES)
def test_face_and_cell_operators(gpu, cell_type):
    of, model = _pair(cell_type)
    assert model.layout.null_face == model.layout.n_internal - 1
    _prepare(of, model, 2)
    assert rel_l2(model.face_values.cpu().numpy(), of.face_values) <= KERNEL_TOL
    assert rel_l2(model.mass_fluxes.cpu().numpy(), of.mass_fluxes) <= KERNEL_TOL
    of.calculate_gradients(); model.calculate_gradients()
    assert rel_l2(model.cell_gradients.cpu().numpy(), of.cell_gradients) <= KERNEL_TOL
    assert rel_l2(model.divFlux().cpu().numpy(), of.divFlux()) <= KERNEL_TOL
    # the null face carries nothing
    f = model.layout.null_face
    assert float(model._flux[f]) == 0.0


@pytest.mark.parametrize('cell_type,correction', [('tetra', False), ('tetra', True),
                                                  ('wedge', False), ('wedge', True)])
def test_momentum_assembly(gpu, cell_type, correction):
    import torch
    from oracle.fv import OracleEquation, convection_weights, diffusion_weights, grad_term_weights
    from warpcfd_amd.FV.device import ptr, stream_ptr
    of, model = _pair(cell_type, seed=3)
    _prepare(of, model, 3)
    of.calculate_gradients()
    nu, alpha = 0.013, 0.7
    eq = OracleEquation(of, [0, 1, 2])
    eq.form_system([('face', convection_weights(of), 1.0),
                    ('face', diffusion_weights(of, nu, (0, 1, 2), correction), -1.0)],
                   explicit=[(grad_term_weights(of), -1.0)])
    eq.relax(alpha)
    C = of.m.num_cells
    dev = model._cv.device
    vals = torch.zeros(7, C, dtype=torch.float32, device=dev)
    rhs = torch.zeros(3, C, dtype=torch.float32, device=dev)
    assert model.lib.wcfd_assemble_momentum(model.dm.desc_ref, model.bc, ptr(model._cv), ptr(model._fv),
                                            ptr(model._flux), ptr(model._cg), 1, nu, 1.0, alpha, 0.0,
                                            int(correction), ptr(vals), ptr(rhs), stream_ptr()) == 0
    fc = of.m.faces_per_cell
    assert torch.all(vals[1 + fc:] == 0)                      # padded slots
    A_ref = eq.matrix().tocsr()[0::3, 0::3].toarray()
    assert rel_l2(_dev_matrix_ref(model, vals).toarray(), A_ref) <= KERNEL_TOL
    assert rel_l2(model.dm.cells_to_reference(rhs).cpu().numpy(), eq.rhs.reshape(C, 3).T) <= KERNEL_TOL


@pytest.mark.parametrize('cell_type', TYPES)
def test_pressure_assembly(gpu, cell_type):
    import torch
    from oracle.fv import OracleEquation, diffusion_weights
    from warpcfd_amd.FV.device import ptr, stream_ptr
    of, model = _pair(cell_type, kind='cavity', seed=4)
    _prepare(of, model, 4)
    rng = np.random.default_rng(5)
    rua_f = (0.1 + rng.random(of.m.num_faces)).astype(np.float32)
    eq = OracleEquation(of, [3])
    eq.form_system([('face', diffusion_weights(of, rua_f, (3,)), 1.0)])
    eq.add_RHS(of.divFlux())
    eq.replace_row(0, 0.0)
    dev = model._cv.device
    C = of.m.num_cells
    vals = torch.zeros(7, C, dtype=torch.float32, device=dev)
    rhs = torch.zeros(C, dtype=torch.float32, device=dev)
    rua_dev = model.dm.faces_to_device(torch.from_numpy(rua_f).to(dev)).contiguous()
    assert model.lib.wcfd_assemble_pressure(model.dm.desc_ref, model.bc, ptr(model._fv),
                                            ptr(model._flux), ptr(rua_dev), ptr(model._cg), 0, 0,
                                            model.reference_cell_device, 0.0, ptr(vals), ptr(rhs),
                                            None, stream_ptr()) == 0
    assert rel_l2(_dev_matrix_ref(model, vals).toarray(), eq.matrix().toarray()) <= KERNEL_TOL
    assert rel_l2(model.dm.cells_to_reference(rhs).cpu().numpy(), eq.rhs) <= KERNEL_TOL


@pytest.mark.parametrize('cell_type', TYPES)
def test_composable_terms(gpu, cell_type):
    from oracle.fv import (OracleEquation, convection_weights, diffusion_weights,
                           grad_term_weights)
    from warpcfd_amd.FV.terms import ConvectionTerm, DiffusionTerm, Equation, GradTerm
    of, model = _pair(cell_type, seed=6)
    _prepare(of, model, 6)
    of.calculate_gradients(); model.calculate_gradients()
    uvw = ['u', 'v', 'w']
    conv = ConvectionTerm(model, uvw, 'upwindLinear')
    diff = DiffusionTerm(model, uvw, correction=True)
    grad = GradTerm(model, 'p')
    conv(model); diff(model, 0.02); grad(model)
    fc = of.m.faces_per_cell
    got = conv.weights.cpu().numpy()
    assert got.shape == (of.m.num_cells, fc, 3, 3)
    want = convection_weights(of, (0, 1, 2), scheme='upwindLinear')
    for j in range(3):
        assert rel_l2(got[..., j], want[..., j]) <= KERNEL_TOL, j
    wd = diffusion_weights(of, np.float32(0.02), (0, 1, 2), correction=True)
    assert rel_l2(diff.weights.cpu().numpy(), wd) <= KERNEL_TOL
    eq = Equation(model, uvw)
    eq.form_system([conv, -diff], [-grad])
    oe = OracleEquation(of, [0, 1, 2])
    oe.form_system([('face', want, 1.0), ('face', wd, -1.0)], [(grad_term_weights(of, 3), -1.0)])
    assert rel_l2(eq.matrix.toarray(), oe.matrix().toarray()) <= KERNEL_TOL
    assert rel_l2(eq.rhs.cpu().numpy(), oe.rhs) <= KERNEL_TOL


@pytest.mark.parametrize('cell_type', TYPES)
def test_simple_channel_vs_oracle(gpu, cell_type):
    from oracle.simple import OracleSolver
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    nodes, cells = _mesh(cell_type, n=(8, 4, 3), seed=7, lengths=(2.0, 1.0, 0.8))
    of = oracle_case(nodes, cells, 'channel', ref_cell=None)
    model = device_case(nodes, cells, 'channel', ref_cell=None)
    os_ = OracleSolver(of, 0.7, 0.3, nu=0.01, implicit_outflow=True)
    sv = IncompressibleSolver(model, 0.7, 0.3, implicit_outflow=True)
    for _ in range(15):
        os_.outer_iteration()
        sv.outer_iteration()
    got = model.cell_values.cpu().numpy()
    assert rel_l2(got[:, :3], of.cell_values[:, :3]) <= 1e-4
    assert rel_l2(got[:, 3], of.cell_values[:, 3]) <= 1e-4


@pytest.mark.parametrize('cell_type', TYPES)
def test_cavity_converged_vs_oracle(gpu, cell_type):
    from oracle.simple import OracleSolver
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    from warpcfd_amd.preprocess.grid import split_hexes, cavity_mesh
    nodes, hexes = cavity_mesh(12)
    cells = split_hexes(hexes, cell_type)
    # every wall no-slip: with the slip walls of the 2-D cavity, tetrahedra
    # give w != 0 and the zero-gradient w lets mass through them, which
    # keeps the continuity residual at ~2e-3 in the reference's algorithm
    of = oracle_case(nodes, cells, 'closed_cavity')
    model = device_case(nodes, cells, 'closed_cavity')
    os_ = OracleSolver(of, 0.7, 0.3, nu=0.01, atol=1e-9)
    sv = IncompressibleSolver(model, 0.7, 0.3)
    sv.tol = 1e-9
    for _ in range(500):
        if os_.outer_iteration()['continuity'] < 1e-7:
            break
    for _ in range(500):
        if sv.outer_iteration()['continuity'] < 1e-7:
            break
    assert os_.history[-1]['continuity'] < 1e-7
    assert sv.history[-1]['continuity'] < 1e-7
    got = model.cell_values.cpu().numpy()
    for q in (0, 1, 3):
        assert rel_l2(got[:, q], of.cell_values[:, q]) <= 1e-4, q


@pytest.mark.parametrize('cell_type', TYPES)
def test_krylov_reads_only_used_slots(gpu, cell_type):
    """The descriptor's n_slots (4 or 5) makes the Krylov kernels skip the
    null slots; the padded six-slot read (n_slots 0) adds only zero products,
    so both give the same iterates and iteration counts.  Both match scipy."""
    import ctypes
    import scipy.sparse.linalg as spla
    import torch
    from test_gpu_parity import _csr_to_ell, _spd_system
    from warpcfd_amd._lib import SolveInfo
    from warpcfd_amd.FV.device import ptr, stream_ptr
    _, model = _pair(cell_type, seed=8)
    L = model.layout
    fc = 4 if cell_type == 'tetra' else 5
    assert model.dm.desc.n_slots == fc
    A = _spd_system(model, 9)
    n = L.num_cells
    dev = model._cv.device
    ell = torch.from_numpy(_csr_to_ell(L, A.data)).to(dev)
    assert torch.all(ell[1 + fc:] == 0)
    rng = np.random.default_rng(10)
    b = torch.from_numpy(rng.standard_normal(n).astype(np.float32)).to(dev)
    b3 = torch.from_numpy(rng.standard_normal((3, n)).astype(np.float32)).to(dev)
    lib = model.lib
    ws = torch.zeros(lib.wcfd_krylov_workspace_bytes(n, 3), dtype=torch.uint8, device=dev)
    out = {}
    for ns in (fc, 0):
        model.dm.desc.n_slots = ns
        for variant in (0, 1):
            x = torch.zeros(n, dtype=torch.float32, device=dev)
            info = SolveInfo()
            assert lib.wcfd_cg_ell(model.dm.desc_ref, ptr(ell), variant, ptr(b), ptr(x), 1e-7, 1e-7,
                                   2000, 10, ptr(ws), ctypes.byref(info), stream_ptr()) == 0
            out[(ns, variant)] = (x.cpu().numpy(), info.iterations)
        x3 = torch.zeros(3, n, dtype=torch.float32, device=dev)
        info = SolveInfo()
        assert lib.wcfd_bicgstab_ell(model.dm.desc_ref, ptr(-ell), 3, ptr(b3), ptr(x3), 1e-7, 1e-7,
                                     2000, 10, ptr(ws), ctypes.byref(info), stream_ptr()) == 0
        out[(ns, 'bicg')] = (x3.cpu().numpy(), info.iterations)
    model.dm.desc.n_slots = fc
    for key in (0, 1, 'bicg'):
        (xa, ia), (xb, ib) = out[(fc, key)], out[(0, key)]
        assert ia == ib and np.array_equal(xa, xb), key
    xs = spla.spsolve(A.astype(np.float64).tocsc(), b.cpu().numpy().astype(np.float64))
    assert rel_l2(out[(fc, 1)][0], xs) <= 1e-5


def _closed_cavity_spec(cell_type, n=12, re=100.0):
    from cases import apply_bcs
    from warpcfd_amd.FV import FVM
    from warpcfd_amd.batch import CaseSpec
    from warpcfd_amd.preprocess import Mesh, cavity_mesh, define_boundary_walls
    from warpcfd_amd.preprocess.grid import split_hexes
    nodes, cells = cavity_mesh(n)
    if cell_type != 'hexahedron':
        cells = split_hexes(cells, cell_type)
    m = Mesh(nodes, cells)
    define_boundary_walls(m)
    model = FVM(m)
    model.material.create_incompressible_newtonian_fluid(viscosity=1.0 / re, density=1.0)
    model.set_reference_pressure(0, 0.0)
    apply_bcs(model.boundary, 'closed_cavity')
    return CaseSpec(model, stop=1e-6, max_outer=3000, name=f'{cell_type}{n}'), nodes, cells


def test_batch_of_mixed_cell_types(gpu):
    """A hexahedral, a tetrahedral and a wedge cavity in one arena (six
    slots for all; the null faces of the tetrahedra and wedges are ordinary
    internal faces of their cases): every case stops by the monitor, the
    tetrahedral case is bitwise the same alone and in the mix, and each
    case's velocity is within 1e-4 of the FP64 oracle's fixed point."""
    from oracle.simple import OracleSolver
    from test_gpu_batch import KW, _run
    from warpcfd_amd.batch import CaseArena
    built = [_closed_cavity_spec(ct) for ct in ('hexahedron', 'tetra', 'wedge')]
    specs = [b[0] for b in built]
    mix = CaseArena.for_cases(specs, **KW)
    _run(mix)
    alone = CaseArena.for_cases([_closed_cavity_spec('tetra')[0]], **KW)
    _run(alone)
    assert alone.iterations(0) == mix.iterations(1)
    assert np.array_equal(alone.cell_values(0), mix.cell_values(1))
    for k, (_, nodes, cells) in enumerate(built):
        assert mix.converged(k), (k, mix.continuity(k))
        of = oracle_case(np.asarray(nodes, np.float32).astype(np.float64), cells, 'closed_cavity',
                         dtype=np.float64, ref_cell=0)
        s = OracleSolver(of, 0.7, 0.3, nu=0.01, atol=1e-12)
        for _ in range(600):
            if s.outer_iteration()['continuity'] < 1e-10:
                break
        got = mix.cell_values(k)
        assert rel_l2(got[:, :3], of.cell_values[:, :3]) <= 1e-4, k


@pytest.mark.parametrize('cell_type', TYPES)
def test_piso_vs_oracle(gpu, cell_type):
    """Transient PISO (ddt in the momentum matrix, two correctors, no
    relaxation) on a jittered channel of tetrahedra or wedges, step by step
    against the oracle: u within 1e-4, v and p within 1e-3 (each step's
    pressure solves stop at an absolute tolerance, as in the hexahedral
    cylinder test)."""
    from oracle.simple import OracleSolver
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    nodes, cells = _mesh(cell_type, n=(8, 4, 3), seed=9, lengths=(2.0, 1.0, 0.8))
    of = oracle_case(nodes, cells, 'channel', ref_cell=None)
    model = device_case(nodes, cells, 'channel', ref_cell=None)
    dt = 0.05
    os_ = OracleSolver(of, None, None, steady_state=False, nu=0.01, implicit_outflow=True)
    os_.set_iterations(1, 2, 1)
    sv = IncompressibleSolver(model, None, None, steady_state=False, implicit_outflow=True)
    sv.set_iterations(1, 2, 1)
    for step in range(5):
        sv.outer_iteration(dt=dt)
        os_.outer_iteration(dt)
        got = model.cell_values.cpu().numpy()
        err = [rel_l2(got[:, q], of.cell_values[:, q]) for q in (0, 1, 3)]
        assert err[0] <= 1e-4 and err[1] <= 1e-3 and err[2] <= 1e-3, (step, err)
