"""Pins oracle.fv (restatement of the reference's FV operators) with closed
forms that a plausible mistake would fail: linear-field exactness of the
skew-corrected interpolation and of Green-Gauss, the reference's inverted
interpolation weight on a stretched grid (T1), a divergence-free uniform
flow, hand-computed upwind and Laplacian rows, Patankar relaxation and the
identity row."""
import numpy as np
import pytest

from oracle.fv import (OracleEquation, OracleFVM, convection_weights, diffusion_weights,
                       grad_term_weights)
from oracle.mesh import OracleMesh
from warpcfd_amd.preprocess.grid import structured_hex


def _fvm(nodes, cells, bc='dirichlet_all', dtype=np.float64):
    m = OracleMesh(nodes, cells, geom_dtype=dtype)
    m.define_boundary_walls()
    f = OracleFVM(m, dtype)
    return f


def _linear(f, a, b0):
    x = f.centroids
    for o in range(4):
        f.cell_values[:, o] = x @ a[o] + b0[o]
        f.cell_gradients[:, o] = a[o]
    bnd = f.boundary
    xf = f.face_centroid[bnd.boundary_ids]
    for o in range(4):
        bnd.values[:, o] = xf @ a[o] + b0[o]
        bnd.types[:, o] = 1


@pytest.mark.parametrize('jitter', [0.0, 0.12])
def test_linear_field_face_values_exact(jitter):
    nodes, cells = structured_hex((5, 4, 3), (1.0, 0.8, 0.6), jitter=jitter, seed=3)
    f = _fvm(nodes, cells)
    rng = np.random.default_rng(0)
    a = rng.standard_normal((4, 3)); b0 = rng.standard_normal(4)
    _linear(f, a, b0)
    f.set_boundary_conditions()
    f.face_interpolation()
    exact = f.face_centroid @ a.T + b0
    np.testing.assert_allclose(f.face_values, exact, atol=1e-12)
    if jitter > 0:
        # without the skew term the jittered faces are not exact
        f.cell_gradients[:] = 0
        f.face_interpolation()
        assert np.abs(f.face_values - exact).max() > 1e-4


def test_green_gauss_linear_exact_uniform():
    nodes, cells = structured_hex((5, 4, 3), (1.0, 0.8, 0.6))
    f = _fvm(nodes, cells)
    rng = np.random.default_rng(1)
    a = rng.standard_normal((4, 3)); b0 = rng.standard_normal(4)
    _linear(f, a, b0)
    f.set_boundary_conditions(); f.face_interpolation()
    f.calculate_gradients()
    np.testing.assert_allclose(f.cell_gradients, np.broadcast_to(a, f.cell_gradients.shape), atol=1e-10)
    f.calculate_gradients((3,))
    assert np.all(f.cell_gradients[:, :3] == 0)          # T3: the others are zeroed
    np.testing.assert_allclose(f.cell_gradients[:, 3], np.broadcast_to(a[3], (f.m.num_cells, 3)), atol=1e-10)


def test_inverted_interpolation_weight_stretched():
    """T1: psi = |x_f - x_O| / |x_O - x_N| weights the OWNER, so on a stretched
    grid phi_f = (h1 phi_O + h2 phi_N) / (h1 + h2), not the textbook weights."""
    xs = np.array([0.0, 1.0, 4.0])           # cells of width 1 and 3
    ys = np.array([0.0, 1.0]); zs = np.array([0.0, 1.0])
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing='ij')
    nodes = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    from warpcfd_amd.preprocess.grid import structured_hex as sh
    _, cells = sh((2, 1, 1))
    f = _fvm(nodes, cells)
    f.cell_values[:, 0] = [2.0, 10.0]
    f.face_interpolation((0,))
    fi = f.m.internal_face_ids[0]
    assert f.m.adjacent_cells[fi, 0] == 0
    h1, h2 = 1.0, 3.0
    assert f.face_values[fi, 0] == pytest.approx((h1 * 2.0 + h2 * 10.0) / (h1 + h2))


def test_uniform_flow_is_divergence_free():
    nodes, cells = structured_hex((5, 4, 3), (1.0, 1.0, 1.0), jitter=0.12, seed=5)
    f = _fvm(nodes, cells)
    u = np.array([0.3, -1.2, 0.7])
    f.cell_values[:, :3] = u
    bnd = f.boundary
    bnd.values[:, :3] = u; bnd.types[:, :3] = 1; bnd.types[:, 3] = 2
    f.set_boundary_conditions(); f.face_interpolation(); f.calculate_mass_flux()
    div = f.divFlux()
    assert np.abs(div).max() < 1e-12
    assert np.abs(f.mass_fluxes).max() > 1e-3


def _line(n):
    nodes, cells = structured_hex((n, 1, 1), (float(n), 1.0, 1.0))
    return _fvm(nodes, cells)


def test_upwind_rows_closed_form():
    f = _line(3)
    f.cell_values[:, 0] = 1.0
    bnd = f.boundary
    bnd.types[:, :3] = 1
    bnd.values[:, 0] = 1.0
    bnd.types[:, 3] = 2
    f.set_boundary_conditions(); f.face_interpolation(); f.calculate_mass_flux()
    eq = OracleEquation(f, [0, 1, 2])
    eq.form_system([('face', convection_weights(f), 1.0)])
    A = eq.matrix().toarray()[0::3, 0::3]
    Fv = 1.0            # u A with A = 1
    # boundary faces only reach b (convection.py:63-67): the outflow face adds
    # nothing to the last diagonal
    expect = np.array([[Fv, 0, 0], [-Fv, Fv, 0], [0, -Fv, 0]])
    np.testing.assert_allclose(A, expect, atol=1e-12)
    # inflow at x=0 brings F_b phi_b = -1 * 1 to b with a minus sign: b_0 = +1;
    # outflow at x=3: b_2 = -(1 * 1)
    np.testing.assert_allclose(eq.rhs[0::3], [1.0, 0.0, -1.0], atol=1e-12)
    # the implicit-outflow variant moves the Neumann outflow face to the diagonal
    out = np.isin(bnd.boundary_ids, f.m.groups['+X'])
    bnd.types[out, :3] = 2
    bnd.values[out, :3] = 0.0
    f.set_boundary_conditions(); f.face_interpolation(); f.calculate_mass_flux()
    eq.form_system([('face', convection_weights(f, implicit_outflow=True), 1.0)])
    np.testing.assert_allclose(eq.matrix().toarray()[0::3, 0::3],
                               [[Fv, 0, 0], [-Fv, Fv, 0], [0, -Fv, Fv]], atol=1e-12)
    np.testing.assert_allclose(eq.rhs[0::3], [1.0, 0.0, 0.0], atol=1e-12)


def _uniform_flow(n=(6, 5, 4), u=(0.7, -0.4, 0.25)):
    """Uniform grid, uniform velocity u on every face (fluxes of both signs)."""
    nodes, cells = structured_hex(n, (1.2, 1.0, 0.8))
    f = _fvm(nodes, cells)
    u = np.asarray(u, np.float64)
    f.cell_values[:, :3] = u
    bnd = f.boundary
    bnd.values[:, :3] = u; bnd.types[:, :3] = 1; bnd.types[:, 3] = 2
    f.set_boundary_conditions(); f.face_interpolation(); f.calculate_mass_flux()
    return f, u


def test_central_convection_of_quadratic_field_closed_form():
    """Central weights (psi, 1 - psi) on a uniform grid in uniform flow: for
    phi = x^2 the face averages are x_f^2 + h^2/4, and the opposite faces'
    h^2/4 cancel, so an interior row of the assembled operator is exactly
    V u_x 2 x_c.  Upwind misses it by V u_x h, so the closed form tells the
    schemes apart."""
    f, u = _uniform_flow()
    x = f.centroids[:, 0]
    phi = x * x
    eq = OracleEquation(f, [0])
    eq.form_system([('face', convection_weights(f, outputs=(0,), scheme='central difference'),
                     1.0)])
    interior = np.all(f.m.neighbors >= 0, axis=1)
    assert interior.sum() > 10
    exact = f.volumes * u[0] * 2.0 * x
    np.testing.assert_allclose((eq.matrix() @ phi)[interior], exact[interior], rtol=1e-10,
                               atol=1e-12)
    eq.form_system([('face', convection_weights(f, outputs=(0,)), 1.0)])
    h = 1.2 / 6
    np.testing.assert_allclose((eq.matrix() @ phi)[interior],
                               (exact - f.volumes * u[0] * h)[interior], rtol=1e-10, atol=1e-12)


def test_upwind_linear_reconstructs_from_the_upwind_cell():
    """upwindLinear: on every internal face, w0 phi_O + w1 phi_N + w2 (the
    explicit part; the skew term vanishes on a uniform grid) equals
    F (phi_up + grad(phi)_up . (x_f - x_up)), up = the upwind cell, for
    fluxes of both signs.  phi = x^2 with its exact gradient 2x at each
    centroid, so a gradient taken from the wrong cell fails."""
    f, u = _uniform_flow()
    x = f.centroids
    f.cell_values[:, 0] = x[:, 0] ** 2
    f.cell_gradients[:, 0] = 0.0
    f.cell_gradients[:, 0, 0] = 2.0 * x[:, 0]
    w = convection_weights(f, outputs=(0,), scheme='upwindLinear')
    m = f.m
    fi = m.internal_face_ids
    own, nb = m.adjacent_cells[fi, 0], m.adjacent_cells[fi, 1]
    k0, k1 = m.cell_face_index[fi, 0], m.cell_face_index[fi, 1]
    F = f.mass_fluxes[fi]
    assert (F > 0).any() and (F < 0).any()
    upc = np.where(F > 0, own, nb)
    xf = f.face_centroid[fi]
    expect = F * (x[upc, 0] ** 2 + 2.0 * x[upc, 0] * (xf[:, 0] - x[upc, 0]))
    phi = f.cell_values[:, 0]
    got_own = w[own, k0, 0, 0] * phi[own] + w[own, k0, 0, 1] * phi[nb] + w[own, k0, 0, 2]
    np.testing.assert_allclose(got_own, expect, rtol=1e-12, atol=1e-12)
    # the neighbour's row carries the same face with the flux negated
    got_nb = w[nb, k1, 0, 0] * phi[nb] + w[nb, k1, 0, 1] * phi[own] + w[nb, k1, 0, 2]
    np.testing.assert_allclose(got_nb, -expect, rtol=1e-12, atol=1e-12)


def test_upwind_tie_goes_to_neighbour():
    """T6: F == 0 puts the implicit weight on the neighbour column."""
    f = _line(2)
    f.mass_fluxes[:] = 0.0
    w = convection_weights(f)
    fi = f.m.internal_face_ids[0]
    o, k0 = f.m.adjacent_cells[fi, 0], f.m.cell_face_index[fi, 0]
    assert w[o, k0, 0, 0] == 0.0 and w[o, k0, 0, 1] == 0.0  # 0 * F either way
    f.mass_fluxes[fi] = -2.0
    w = convection_weights(f)
    assert w[o, k0, 0, 1] == -2.0 and w[o, k0, 0, 0] == 0.0


def test_laplacian_rows_closed_form():
    n = 4
    f = _line(n)
    nu = 0.37
    bnd = f.boundary
    bnd.types[:] = 2                         # all Neumann ...
    xm = bnd.boundary_ids[np.isin(bnd.boundary_ids, f.m.groups['-X'])]
    rows = np.isin(bnd.boundary_ids, f.m.groups['-X'])
    bnd.types[rows, 0] = 1                   # ... but u Dirichlet on -X
    bnd.values[rows, 0] = 2.0
    f.set_boundary_conditions()
    eq = OracleEquation(f, [0])
    eq.form_system([('face', diffusion_weights(f, nu, (0,)), -1.0)])
    A = eq.matrix().toarray()
    D = nu * 1.0 / 1.0
    expect = np.diag([D + 2 * D, 2 * D, 2 * D, D]) - D * (np.eye(n, k=1) + np.eye(n, k=-1))
    np.testing.assert_allclose(A, expect, atol=1e-12)
    np.testing.assert_allclose(eq.rhs, [2 * D * 2.0, 0, 0, 0], atol=1e-12)
    del xm


def test_relax_and_replace_row():
    nodes, cells = structured_hex((3, 3, 1), (1.0, 1.0, 0.1))
    f = _fvm(nodes, cells)
    rng = np.random.default_rng(2)
    f.cell_values[:] = rng.standard_normal(f.cell_values.shape)
    f.boundary.types[:, :3] = 1; f.boundary.types[:, 3] = 2
    f.set_boundary_conditions(); f.face_interpolation(); f.calculate_mass_flux()
    eq = OracleEquation(f, [0, 1, 2])
    eq.form_system([('face', diffusion_weights(f, 0.1, (0, 1, 2)), -1.0)],
                   explicit=[(grad_term_weights(f), -1.0)])
    d0 = eq.diagonal.copy(); b0 = eq.rhs.copy()
    alpha = 0.7
    eq.relax(alpha)
    np.testing.assert_allclose(eq.diagonal, d0 / alpha, rtol=1e-14)
    phi = f.cell_values[:, :3].reshape(-1)
    np.testing.assert_allclose(eq.rhs, b0 + (1 - alpha) / alpha * d0 * phi, rtol=1e-12)
    peq = OracleEquation(f, [3])
    peq.form_system([('face', diffusion_weights(f, np.ones(f.m.num_faces), (3,)), 1.0)])
    peq.replace_row(0, 0.25)
    A = peq.matrix().toarray()
    assert A[0, 0] == 1.0 and np.all(A[0, 1:] == 0)
    assert np.any(A[1:, 0] != 0)                  # T7: the column stays, A is non-symmetric
    assert peq.rhs[0] == 0.25


@pytest.mark.parametrize('jitter', [0.0, 0.12])
def test_upwind_face_values_take_the_upwind_cell(jitter):
    """interpolation_functions.py:18-28: with zero gradients (no skew term) an
    internal face takes its owner's value where its mass flux is > 0 and its
    neighbour's where it is < 0 or exactly 0; boundary faces are untouched by
    the scheme.  Random cell values, fluxes of every sign and some zeros."""
    nodes, cells = structured_hex((5, 4, 3), (1.0, 0.8, 0.6), jitter=jitter, seed=4)
    f = _fvm(nodes, cells)
    rng = np.random.default_rng(5)
    f.cell_values[:] = rng.standard_normal(f.cell_values.shape)
    f.cell_gradients[:] = 0
    flux = rng.standard_normal(f.mass_fluxes.shape)
    flux[::7] = 0.0
    f.mass_fluxes[:] = flux
    f.face_interpolation(upwind=True)
    m = f.m
    fi = m.internal_face_ids
    own, nb = m.adjacent_cells[fi, 0], m.adjacent_cells[fi, 1]
    for o in range(4):
        want = np.where(flux[fi] > 0, f.cell_values[own, o], f.cell_values[nb, o])
        np.testing.assert_array_equal(f.face_values[fi, o], want)
    assert np.any(flux[fi] == 0) and np.any(flux[fi] < 0) and np.any(flux[fi] > 0)


def test_upwind_and_linear_share_the_skew_term():
    """On a jittered grid with non-zero gradients, upwind minus linear face
    values equal upwind minus linear interpolants of the cell values alone:
    the skew correction (skewness_functions.py:7-20) is the same term in
    both schemes."""
    nodes, cells = structured_hex((5, 4, 3), (1.0, 0.8, 0.6), jitter=0.12, seed=6)
    f = _fvm(nodes, cells)
    rng = np.random.default_rng(7)
    f.cell_values[:] = rng.standard_normal(f.cell_values.shape)
    f.cell_gradients[:] = rng.standard_normal(f.cell_gradients.shape)
    f.mass_fluxes[:] = rng.standard_normal(f.mass_fluxes.shape)
    f.face_interpolation(upwind=True)
    up = f.face_values.copy()
    f.face_interpolation()
    lin = f.face_values.copy()
    m = f.m
    fi = m.internal_face_ids
    own, nb = m.adjacent_cells[fi, 0], m.adjacent_cells[fi, 1]
    psi, psi2 = f.psi[fi], f.psi2[fi]
    for o in range(4):
        cv = f.cell_values[:, o]
        d_val = np.where(f.mass_fluxes[fi] > 0, cv[own], cv[nb]) - (cv[own] * psi + psi2 * cv[nb])
        np.testing.assert_allclose(up[fi, o] - lin[fi, o], d_val, atol=1e-12)
    skew = lin[fi, 0] - (f.cell_values[own, 0] * psi + psi2 * f.cell_values[nb, 0])
    assert np.abs(skew).max() > 1e-3          # the skew term is really there
