"""GPU parity of every hot-path kernel against the oracle (CPU restatement of
the reference), through the C ABI.  Tolerance: rel-L2 <= 1e-5 per kernel
(float32 arithmetic, different summation order), <= 1e-4 for whole SIMPLE
runs."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from cases import apply_bcs, device_case, jittered_box, oracle_case, rel_l2

pytestmark = pytest.mark.gpu

KERNEL_TOL = 1e-5


def _torch():
    import torch
    return torch


def _set_state(model, of, seed=0):
    """Random cell values and gradients, identical on both sides."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    C = of.m.num_cells
    cv = rng.standard_normal((C, 4)).astype(np.float32)
    cg = (0.5 * rng.standard_normal((C, 4, 3))).astype(np.float32)
    of.cell_values[:] = cv
    of.cell_gradients[:] = cg
    dev = model._cv.device
    model._cv.copy_(model.dm.cells_to_device(torch.from_numpy(cv.T.copy()).to(dev)))
    model._cg.copy_(model.dm.cells_to_device(torch.from_numpy(cg.reshape(C, 12).T.copy()).to(dev)))
    model._uvw_grad_zero = False


def _pair(kind='channel', n=(9, 7, 5), jitter=0.1, seed=1):
    nodes, cells = jittered_box(n, jitter, seed)
    of = oracle_case(nodes, cells, kind)
    model = device_case(nodes, cells, kind)
    return of, model


def _dev_matrix_ref(model, values):
    """Device ELL matrix (device order) -> scipy CSR in reference cell order,
    converted on the GPU by wcfd_ell_to_csr and checked against the host
    conversion."""
    torch = _torch()
    from warpcfd_amd.FV.device import ptr, stream_ptr
    L = model.layout
    rp = L.row_ptr.astype(np.int64)
    rows = np.repeat(np.arange(L.num_cells), np.diff(rp))
    csr = torch.zeros(L.nnz, dtype=torch.float32, device=values.device)
    assert model.lib.wcfd_ell_to_csr(model.dm.desc_ref, ptr(values), ptr(csr), stream_ptr()) == 0
    v = csr.cpu().numpy()
    assert np.array_equal(v, L.ell_to_csr(values.cpu().numpy()))
    r = L.new2old[rows]
    c = L.new2old[L.col]
    return sp.csr_matrix((v, (r, c)), shape=(L.num_cells, L.num_cells))


def test_face_interpolation_and_flux(gpu):
    of, model = _pair()
    _set_state(model, of)
    of.set_boundary_conditions(); of.face_interpolation(); of.calculate_mass_flux()
    model.set_boundary_conditions(); model.face_interpolation(); model.calculate_mass_flux()
    assert rel_l2(model.face_values.cpu().numpy(), of.face_values) <= KERNEL_TOL
    assert rel_l2(model.mass_fluxes.cpu().numpy(), of.mass_fluxes) <= KERNEL_TOL


def test_face_interpolation_upwind(gpu):
    """face_interpolation(upwind=True) (interpolation_functions.py:18-28)
    against the oracle: random cell values and gradients (so the skew term
    is live on the jittered mesh), the same mass fluxes on both sides, of
    both signs and with some exact zeros (which take the neighbour)."""
    torch = _torch()
    of, model = _pair()
    _set_state(model, of, seed=4)
    rng = np.random.default_rng(8)
    flux = rng.standard_normal(of.mass_fluxes.shape).astype(np.float32)
    flux[::5] = 0.0
    of.mass_fluxes[:] = flux
    dev = model._cv.device
    model._flux.copy_(model.dm.faces_to_device(torch.from_numpy(flux).to(dev)))
    of.set_boundary_conditions(); of.face_interpolation(upwind=True)
    model.set_boundary_conditions(); model.face_interpolation(upwind=True)
    got = model.face_values.cpu().numpy()
    assert rel_l2(got, of.face_values) <= KERNEL_TOL
    # the linear scheme gives other face values on the same state
    model.face_interpolation()
    assert not np.array_equal(model.face_values.cpu().numpy(), got)


def test_init_global_arrays_zeroes_the_fields(gpu):
    """init_global_arrays (model.py:155-176), which finalize calls, gives
    zeroed cell values, gradients, face values and mass fluxes again after
    a state has been set and the face pass run."""
    of, model = _pair()
    _set_state(model, of, seed=5)
    model.set_boundary_conditions(); model.face_interpolation(); model.calculate_mass_flux()
    assert np.abs(model.cell_values.cpu().numpy()).max() > 0
    model.init_global_arrays()
    for arr in (model.cell_values, model.cell_gradients, model.face_values, model.mass_fluxes):
        a = arr.cpu().numpy()
        assert a.size and not a.any()


def test_green_gauss_and_divergence(gpu):
    of, model = _pair()
    _set_state(model, of, seed=3)
    for x in (of, model):
        x.set_boundary_conditions(); x.face_interpolation(); x.calculate_mass_flux()
    of.calculate_gradients(); model.calculate_gradients()
    assert rel_l2(model.cell_gradients.cpu().numpy(), of.cell_gradients) <= KERNEL_TOL
    assert rel_l2(model.divFlux().cpu().numpy(), of.divFlux()) <= KERNEL_TOL
    # calculate_gradients('p') zeroes the other gradients (model.py:237)
    of.calculate_gradients((3,)); model.calculate_gradients('p')
    g = model.cell_gradients.cpu().numpy()
    assert np.all(g[:, :3] == 0)
    assert rel_l2(g, of.cell_gradients) <= KERNEL_TOL


@pytest.mark.parametrize('correction,outflow', [(False, False), (True, False), (False, True)])
def test_momentum_assembly(gpu, correction, outflow):
    from oracle.fv import OracleEquation, convection_weights, diffusion_weights, grad_term_weights
    torch = _torch()
    of, model = _pair(seed=5)
    _set_state(model, of, seed=5)
    for x in (of, model):
        x.set_boundary_conditions(); x.face_interpolation(); x.calculate_mass_flux()
    of.calculate_gradients()
    nu, alpha = 0.013, 0.7
    eq = OracleEquation(of, [0, 1, 2])
    wc = convection_weights(of, implicit_outflow=outflow)
    wd = diffusion_weights(of, nu, (0, 1, 2), correction)
    eq.form_system([('face', wc, 1.0), ('face', wd, -1.0)], explicit=[(grad_term_weights(of), -1.0)])
    eq.relax(alpha)
    A_ref = eq.matrix().tocsr()[0::3, 0::3].toarray()
    C = of.m.num_cells
    dev = model._cv.device
    vals = torch.zeros(7, C, dtype=torch.float32, device=dev)
    rhs = torch.zeros(3, C, dtype=torch.float32, device=dev)
    from warpcfd_amd.FV.device import ptr, stream_ptr
    code = model.lib.wcfd_assemble_momentum(model.dm.desc_ref, model.bc, ptr(model._cv), ptr(model._fv),
                                            ptr(model._flux), ptr(model._cg), 1, nu, 1.0, alpha, 0.0,
                                            int(correction) | (2 if outflow else 0), ptr(vals), ptr(rhs),
                                            stream_ptr())
    assert code == 0
    torch.cuda.synchronize()
    A_dev = _dev_matrix_ref(model, vals).toarray()
    assert rel_l2(A_dev, A_ref) <= KERNEL_TOL
    rhs_ref = eq.rhs.reshape(C, 3).T
    assert rel_l2(model.dm.cells_to_reference(rhs).cpu().numpy(), rhs_ref) <= KERNEL_TOL
    # the kernel's gradients equal calculate_gradients()
    assert rel_l2(model.cell_gradients.cpu().numpy(), of.cell_gradients) <= KERNEL_TOL


def test_transient_momentum_assembly(gpu):
    from oracle.fv import OracleEquation, convection_weights, diffusion_weights, grad_term_weights, ddt_weights
    torch = _torch()
    of, model = _pair(seed=6)
    _set_state(model, of, seed=6)
    for x in (of, model):
        x.set_boundary_conditions(); x.face_interpolation(); x.calculate_mass_flux()
    of.calculate_gradients()
    nu, dt = 0.01, 0.02
    eq = OracleEquation(of, [0, 1, 2])
    eq.form_system([('cell', ddt_weights(of, dt), 1.0), ('face', convection_weights(of), 1.0),
                    ('face', diffusion_weights(of, nu, (0, 1, 2)), -1.0)],
                   explicit=[(grad_term_weights(of), -1.0)])
    C = of.m.num_cells
    dev = model._cv.device
    vals = torch.zeros(7, C, dtype=torch.float32, device=dev)
    rhs = torch.zeros(3, C, dtype=torch.float32, device=dev)
    from warpcfd_amd.FV.device import ptr, stream_ptr
    code = model.lib.wcfd_assemble_momentum(model.dm.desc_ref, model.bc, ptr(model._cv), ptr(model._fv),
                                            ptr(model._flux), ptr(model._cg), 1, nu, 1.0, 0.0, dt, 0,
                                            ptr(vals), ptr(rhs), stream_ptr())
    assert code == 0
    A_ref = eq.matrix().tocsr()[0::3, 0::3].toarray()
    assert rel_l2(_dev_matrix_ref(model, vals).toarray(), A_ref) <= KERNEL_TOL
    assert rel_l2(model.dm.cells_to_reference(rhs).cpu().numpy(), eq.rhs.reshape(C, 3).T) <= KERNEL_TOL


@pytest.mark.parametrize('kind,ref_cell', [('channel', None), ('cavity', 0)])
def test_pressure_assembly(gpu, kind, ref_cell):
    from oracle.fv import OracleEquation, diffusion_weights
    torch = _torch()
    nodes, cells = jittered_box((8, 6, 4), 0.1, 7)
    of = oracle_case(nodes, cells, kind, ref_cell=ref_cell)
    model = device_case(nodes, cells, kind, ref_cell=ref_cell)
    _set_state(model, of, seed=7)
    for x in (of, model):
        x.set_boundary_conditions(); x.face_interpolation(); x.calculate_mass_flux()
    rng = np.random.default_rng(8)
    rua_f = (0.1 + rng.random(of.m.num_faces)).astype(np.float32)
    div = of.divFlux()
    eq = OracleEquation(of, [3])
    eq.form_system([('face', diffusion_weights(of, rua_f, (3,)), 1.0)])
    eq.add_RHS(div)
    if ref_cell is not None:
        eq.replace_row(ref_cell, 0.0)
    dev = model._cv.device
    C = of.m.num_cells
    vals = torch.zeros(7, C, dtype=torch.float32, device=dev)
    rhs = torch.zeros(C, dtype=torch.float32, device=dev)
    rua_dev = model.dm.faces_to_device(torch.from_numpy(rua_f).to(dev)).contiguous()
    from warpcfd_amd.FV.device import ptr, stream_ptr
    code = model.lib.wcfd_assemble_pressure(model.dm.desc_ref, model.bc, ptr(model._fv), ptr(model._flux),
                                            ptr(rua_dev), ptr(model._cg), 0, 0, model.reference_cell_device,
                                            0.0, ptr(vals), ptr(rhs), None, stream_ptr())
    assert code == 0
    assert rel_l2(_dev_matrix_ref(model, vals).toarray(), eq.matrix().toarray()) <= KERNEL_TOL
    assert rel_l2(model.dm.cells_to_reference(rhs).cpu().numpy(), eq.rhs) <= KERNEL_TOL


def _spd_system(model, seed):
    """A pressure-like system on the device pattern: symmetric, diagonally
    dominant, negative definite (the sign the reference's Laplacian has)."""
    L = model.layout
    rng = np.random.default_rng(seed)
    rp = L.row_ptr.astype(np.int64)
    rows = np.repeat(np.arange(L.num_cells), np.diff(rp))
    off = rows != L.col
    w = np.zeros(L.nnz)
    lo = off & (rows < L.col)
    wv = 0.1 + rng.random(lo.sum())
    A = sp.csr_matrix((np.r_[wv, wv], (np.r_[rows[lo], L.col[lo]], np.r_[L.col[lo], rows[lo]])),
                      shape=(L.num_cells, L.num_cells))
    d = -(np.asarray(A.sum(axis=1)).ravel() + 0.05)
    A = (A + sp.diags(d)).tocsr()
    A.sort_indices()
    return A.astype(np.float32)


def _csr_to_ell(L, data):
    ell = np.zeros((7, L.num_cells), np.float32)
    ell[0] = data[L.diag_pos]
    for k in range(6):
        ok = L.slot_pos[k] >= 0
        ell[k + 1][ok] = data[L.slot_pos[k][ok]]
    return ell


def test_ell_solvers_match_csr(gpu):
    """The device path's slot-major matrices give the same SpMV and solves as
    the CSR entry points (int16 neighbour offsets)."""
    torch = _torch()
    from warpcfd_amd._lib import SolveInfo
    from warpcfd_amd.FV.device import ptr, stream_ptr
    nodes, cells = jittered_box((14, 10, 8), 0.1, 31)
    model = device_case(nodes, cells, 'channel')
    L = model.layout
    assert L.slot_dnb is not None
    A = _spd_system(model, 7)
    n = L.num_cells
    dev = model._cv.device
    lib = model.lib
    ell = torch.from_numpy(_csr_to_ell(L, A.data)).to(dev)
    csr = torch.from_numpy(A.data).to(dev)
    rng = np.random.default_rng(8)
    x = torch.from_numpy(rng.standard_normal((3, n)).astype(np.float32)).to(dev)
    y1 = torch.zeros(3, n, dtype=torch.float32, device=dev)
    y2 = torch.zeros(3, n, dtype=torch.float32, device=dev)
    assert lib.wcfd_spmv_ell(model.dm.desc_ref, ptr(ell), 3, ptr(x), ptr(y1), stream_ptr()) == 0
    assert lib.wcfd_spmv(n, ptr(model.dm.row_ptr), ptr(model.dm.col), ptr(csr), 3, ptr(x), ptr(y2),
                         stream_ptr()) == 0
    assert rel_l2(y1.cpu().numpy(), y2.cpu().numpy()) <= 1e-6
    b = torch.from_numpy(rng.standard_normal(n).astype(np.float32)).to(dev)
    ws = torch.zeros(lib.wcfd_krylov_workspace_bytes(n, 3), dtype=torch.uint8, device=dev)
    xa = torch.zeros(n, dtype=torch.float32, device=dev)
    xb = torch.zeros(n, dtype=torch.float32, device=dev)
    ia, ib = SolveInfo(), SolveInfo()
    for variant in (0, 1):
        xa.zero_(); xb.zero_()
        assert lib.wcfd_cg_ell(model.dm.desc_ref, ptr(ell), variant, ptr(b), ptr(xa), 1e-6, 1e-6, 1000, 10,
                               ptr(ws), ctypes.byref(ia), stream_ptr()) == 0
        assert lib.wcfd_cg(n, ptr(model.dm.row_ptr), ptr(model.dm.col), ptr(csr), ptr(model.dm.diag_pos),
                           variant, ptr(b), ptr(xb), 1e-6, 1e-6, 1000, 10, ptr(ws), ctypes.byref(ib),
                           stream_ptr()) == 0
        assert ia.iterations == ib.iterations and ia.converged
        assert rel_l2(xa.cpu().numpy(), xb.cpu().numpy()) <= 1e-5
    b3 = torch.from_numpy(rng.standard_normal((3, n)).astype(np.float32)).to(dev)
    xa3 = torch.zeros(3, n, dtype=torch.float32, device=dev)
    negell = -ell
    assert lib.wcfd_bicgstab_ell(model.dm.desc_ref, ptr(negell), 3, ptr(b3), ptr(xa3), 1e-6, 1e-6, 1000, 10,
                                 ptr(ws), ctypes.byref(ia), stream_ptr()) == 0
    xs = np.stack([spla.spsolve((-A).astype(np.float64).tocsc(), b3[q].cpu().numpy().astype(np.float64))
                   for q in range(3)])
    assert ia.converged and rel_l2(xa3.cpu().numpy(), xs) <= 1e-5


@pytest.mark.parametrize('variant', [0, 1])
def test_spmv_and_cg(gpu, variant):
    torch = _torch()
    from warpcfd_amd._lib import SolveInfo
    from warpcfd_amd.FV.device import ptr, stream_ptr
    nodes, cells = jittered_box((10, 8, 6), 0.1, 9)
    model = device_case(nodes, cells, 'channel')
    A = _spd_system(model, 1)
    L = model.layout
    assert np.array_equal(A.indptr, L.row_ptr) and np.array_equal(A.indices, L.col)
    dev = model._cv.device
    n = L.num_cells
    vals = torch.from_numpy(A.data).to(dev)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((3, n)).astype(np.float32)
    y = torch.zeros(3, n, dtype=torch.float32, device=dev)
    lib = model.lib
    assert lib.wcfd_spmv(n, ptr(model.dm.row_ptr), ptr(model.dm.col), ptr(vals), 3,
                         ptr(torch.from_numpy(x).to(dev)), ptr(y), stream_ptr()) == 0
    ref = np.stack([A @ x[q] for q in range(3)])
    assert rel_l2(y.cpu().numpy(), ref) <= KERNEL_TOL
    b = rng.standard_normal(n).astype(np.float32)
    xs = spla.spsolve(A.astype(np.float64).tocsc(), b.astype(np.float64))
    xd = torch.zeros(n, dtype=torch.float32, device=dev)
    ws = torch.zeros(lib.wcfd_krylov_workspace_bytes(n, 3), dtype=torch.uint8, device=dev)
    info = SolveInfo()
    code = lib.wcfd_cg(n, ptr(model.dm.row_ptr), ptr(model.dm.col), ptr(vals), ptr(model.dm.diag_pos),
                       variant, ptr(torch.from_numpy(b).to(dev)), ptr(xd), 1e-7, 1e-7, 2000, 10, ptr(ws),
                       ctypes.byref(info), stream_ptr())
    assert code == 0 and info.converged
    assert rel_l2(xd.cpu().numpy(), xs) <= 1e-5
    # iteration count and residual follow the oracle's statement of the solver
    from oracle import krylov
    xo = np.zeros(n, np.float32)
    solve = krylov.cg if variant == 0 else krylov.cg_single_reduction
    it, res = solve(A, b, xo, (1.0 / A.diagonal()).astype(np.float32), 1e-7, maxiter=2000)
    assert it == info.iterations
    assert rel_l2(xd.cpu().numpy(), xo) <= 1e-5


def test_bicgstab_three_vectors(gpu):
    torch = _torch()
    from warpcfd_amd._lib import SolveInfo
    from warpcfd_amd.FV.device import ptr, stream_ptr
    nodes, cells = jittered_box((10, 8, 6), 0.1, 11)
    model = device_case(nodes, cells, 'channel')
    L = model.layout
    rng = np.random.default_rng(4)
    S = _spd_system(model, 3)
    # add an upwind-like non-symmetric part on the same pattern
    N = S.copy()
    N.data = N.data + np.where(N.indices != np.repeat(np.arange(L.num_cells), np.diff(N.indptr)),
                               -0.05 * rng.random(N.nnz), 0.0).astype(np.float32)
    A = (-N).tocsr()
    A.sort_indices()
    n = L.num_cells
    dev = model._cv.device
    vals = torch.from_numpy(A.data.astype(np.float32)).to(dev)
    b = rng.standard_normal((3, n)).astype(np.float32)
    x = torch.zeros(3, n, dtype=torch.float32, device=dev)
    lib = model.lib
    ws = torch.zeros(lib.wcfd_krylov_workspace_bytes(n, 3), dtype=torch.uint8, device=dev)
    info = SolveInfo()
    code = lib.wcfd_bicgstab(n, ptr(model.dm.row_ptr), ptr(model.dm.col), ptr(vals), ptr(model.dm.diag_pos),
                             3, ptr(torch.from_numpy(b).to(dev)), ptr(x), 1e-7, 1e-7, 2000, 10, ptr(ws),
                             ctypes.byref(info), stream_ptr())
    assert code == 0 and info.converged
    Ad = A.astype(np.float64).tocsc()
    xs = np.stack([spla.spsolve(Ad, b[q].astype(np.float64)) for q in range(3)])
    assert rel_l2(x.cpu().numpy(), xs) <= 1e-5


def _run_pair(kind, nodes, cells, iters, ref_cell=0, u_relax=0.7, p_relax=0.3, tol=None,
              implicit_outflow=False, cg_variant=0):
    from oracle.simple import OracleSolver
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    of = oracle_case(nodes, cells, kind, ref_cell=ref_cell)
    model = device_case(nodes, cells, kind, ref_cell=ref_cell)
    os_ = OracleSolver(of, u_relax, p_relax, nu=0.01, atol=tol, implicit_outflow=implicit_outflow,
                       cg_variant=cg_variant)
    sv = IncompressibleSolver(model, u_relax, p_relax, implicit_outflow=implicit_outflow)
    sv.cg_variant = cg_variant
    if tol is not None:
        sv.tol = tol
    for _ in range(iters):
        os_.outer_iteration()
        sv.outer_iteration()
    return of, model, os_, sv


def test_simple_cavity_vs_oracle(gpu):
    """30 SIMPLE iterations with the reference's solver settings (atol 1e-5,
    maxiter 500, check every 10).  Velocity agrees to 1e-4.  Pressure only to
    1e-3: each CG stops at ||r|| <= 1e-5, which fixes p only to that
    tolerance, and a one-ulp difference can move a stop by one check interval
    (10 iterations).  Converged fields are compared in the next test."""
    from warpcfd_amd.preprocess.grid import cavity_mesh
    nodes, cells = cavity_mesh(24)
    of, model, os_, sv = _run_pair('cavity', nodes, cells, 30)
    got = model.cell_values.cpu().numpy()
    assert rel_l2(got[:, :2], of.cell_values[:, :2]) <= 1e-4
    assert rel_l2(got[:, 3], of.cell_values[:, 3]) <= 1e-3
    assert abs(sv.history[-1]['continuity'] - os_.history[-1]['continuity']) <= \
        0.05 * os_.history[-1]['continuity']


def test_simple_cavity_converged_vs_oracle(gpu):
    """Converged fields (continuity residual < 1e-7, Krylov atol 1e-9): u, v,
    p within 1e-4 of the oracle."""
    from oracle.simple import OracleSolver
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    from warpcfd_amd.preprocess.grid import cavity_mesh
    nodes, cells = cavity_mesh(24)
    of = oracle_case(nodes, cells, 'cavity')
    model = device_case(nodes, cells, 'cavity')
    os_ = OracleSolver(of, 0.7, 0.3, nu=0.01, atol=1e-9)
    sv = IncompressibleSolver(model, 0.7, 0.3)
    sv.tol = 1e-9
    for _ in range(400):
        if os_.outer_iteration()['continuity'] < 1e-7:
            break
    for _ in range(400):
        if sv.outer_iteration()['continuity'] < 1e-7:
            break
    assert sv.history[-1]['continuity'] < 1e-7
    got = model.cell_values.cpu().numpy()
    for q in (0, 1, 3):
        assert rel_l2(got[:, q], of.cell_values[:, q]) <= 1e-4, q


@pytest.mark.parametrize('cg_variant', [0, 1])
def test_simple_jittered_channel_vs_oracle(gpu, cg_variant):
    nodes, cells = jittered_box((12, 6, 5), 0.1, 13, lengths=(2.0, 1.0, 0.8))
    of, model, os_, sv = _run_pair('channel', nodes, cells, 20, ref_cell=None, implicit_outflow=True,
                                   cg_variant=cg_variant)
    got = model.cell_values.cpu().numpy()
    assert rel_l2(got[:, :3], of.cell_values[:, :3]) <= 1e-4
    assert rel_l2(got[:, 3], of.cell_values[:, 3]) <= 1e-4


def test_simple_is_deterministic(gpu):
    from warpcfd_amd.preprocess.grid import cavity_mesh
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    nodes, cells = cavity_mesh(20)
    outs = []
    for _ in range(2):
        model = device_case(nodes, cells, 'cavity')
        sv = IncompressibleSolver(model, 0.7, 0.3)
        for _ in range(10):
            sv.outer_iteration()
        outs.append(model.cell_values.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])


def test_cg_after_bicgstab_in_one_workspace(gpu):
    """Regression: CG and the 3-vector BiCGSTAB share the SIMPLE workspace;
    the solver state must not overlap either solver's vectors."""
    torch = _torch()
    from warpcfd_amd._lib import SolveInfo
    from warpcfd_amd.FV.device import ptr, stream_ptr
    nodes, cells = jittered_box((24, 16, 12), 0.1, 21)
    model = device_case(nodes, cells, 'channel')
    A = _spd_system(model, 5)
    n = model.layout.num_cells
    dev = model._cv.device
    lib = model.lib
    vals = torch.from_numpy(A.data).to(dev)
    negA = torch.from_numpy((-A).data.astype(np.float32)).to(dev)
    rng = np.random.default_rng(6)
    b3 = torch.from_numpy(rng.standard_normal((3, n)).astype(np.float32)).to(dev)
    b1 = rng.standard_normal(n).astype(np.float32)
    ws = torch.zeros(lib.wcfd_simple_workspace_bytes(n, model.num_faces), dtype=torch.uint8, device=dev)
    args = (ptr(model.dm.row_ptr), ptr(model.dm.col))
    for _ in range(2):
        x3 = torch.zeros(3, n, dtype=torch.float32, device=dev)
        info = SolveInfo()
        assert lib.wcfd_bicgstab(n, *args, ptr(negA), ptr(model.dm.diag_pos), 3, ptr(b3), ptr(x3), 1e-6, 1e-6,
                                 500, 10, ptr(ws), ctypes.byref(info), stream_ptr()) == 0
        x1 = torch.zeros(n, dtype=torch.float32, device=dev)
        info = SolveInfo()
        assert lib.wcfd_cg(n, *args, ptr(vals), ptr(model.dm.diag_pos), 1, ptr(torch.from_numpy(b1).to(dev)),
                           ptr(x1), 1e-6, 1e-6, 1000, 10, ptr(ws), ctypes.byref(info), stream_ptr()) == 0
        assert info.converged and info.iterations > 0
        xs = spla.spsolve(A.astype(np.float64).tocsc(), b1.astype(np.float64))
        assert rel_l2(x1.cpu().numpy(), xs) <= 1e-4
