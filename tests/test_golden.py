"""Committed fixtures (tests/golden/make_golden.py): mesh tables regression,
the oracle's cavity against the Ghia et al. Re=100 centreline data that ships
with the reference, and float32-vs-float64 agreement of converged fields."""
import json
import os

import numpy as np

from oracle.mesh import OracleMesh
from warpcfd_amd.preprocess.mesh import Mesh

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def test_mesh_tables_fixture():
    z = np.load(os.path.join(G, 'mesh_3x3x2_jitter.npz'))
    m = OracleMesh(z['nodes'], z['cells'])
    pm = Mesh(z['nodes'], z['cells'])
    for k in ('adjacent_cells', 'cell_face_index', 'neighbors', 'face_side', 'nnz_per_cell',
              'face_offset_index', 'cell_faces'):
        assert np.array_equal(getattr(m, k), z[k]), k
    assert np.array_equal(pm.face_properties.adjacent_cells, z['adjacent_cells'])
    r, c = pm.csr_pattern(3)
    assert np.array_equal(r, z['csr3_rowptr']) and np.array_equal(c, z['csr3_cols'])
    r, c = pm.csr_pattern(1)
    assert np.array_equal(r, z['csr1_rowptr']) and np.array_equal(c, z['csr1_cols'])
    area, psi = pm.face_struct_arrays()
    assert np.array_equal(psi, z['psi']) and np.array_equal(area, z['area'])


def _centrelines(cv, n):
    u = cv[:, 0].reshape(n, n)      # [j (y), i (x)]
    v = cv[:, 1].reshape(n, n)
    h = 1.0 / n
    yc = (np.arange(n) + 0.5) * h
    u_mid = 0.5 * (u[:, n // 2 - 1] + u[:, n // 2])     # x = 0.5
    v_mid = 0.5 * (v[n // 2 - 1, :] + v[n // 2, :])     # y = 0.5
    return yc, u_mid, v_mid


def test_cavity_against_ghia():
    z = np.load(os.path.join(G, 'cavity64_ref500.npz'))
    gh = json.load(open(os.path.join(G, 'ghia_re100.json')))
    yc, u_mid, v_mid = _centrelines(z['cell_values'], 64)
    gu = np.array(gh['u_on_x_0.5']); gv = np.array(gh['v_on_y_0.5'])
    inner = (gu[:, 0] > yc[0]) & (gu[:, 0] < yc[-1])
    du = np.abs(np.interp(gu[inner, 0], yc, u_mid) - gu[inner, 1]).max()
    inner = (gv[:, 0] > yc[0]) & (gv[:, 0] < yc[-1])
    dv = np.abs(np.interp(gv[inner, 0], yc, v_mid) - gv[inner, 1]).max()
    # first-order upwind on 64x64 after 500 iterations: within a few percent of
    # the lid speed (the reference's own 41x41 plot is ~0.02 off at the extrema)
    assert du < 0.04 and dv < 0.04, (du, dv)
    assert u_mid.min() < -0.18 and v_mid.max() > 0.15 and v_mid.min() < -0.21


def test_float32_vs_float64_converged():
    z = np.load(os.path.join(G, 'cavity24_converged.npz'))
    a, b = z['f32'], z['f64']
    for q in (0, 1, 3):
        assert np.linalg.norm(a[:, q] - b[:, q]) / np.linalg.norm(b[:, q]) < 1e-4
