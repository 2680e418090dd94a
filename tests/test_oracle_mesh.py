"""Pins oracle.mesh (restatement of preprocess/mesh.py + mesh_structs.py) and
checks the product preprocess against it, bit-exact for every integer table.

Pins: uniform-grid closed forms, closed-cell sum(A n) = 0, a brute-force pair
loop written like mesh.py:263-281, the CSR closed form, and relabelling
invariance (a plausible orientation or owner mistake changes these)."""
import numpy as np
import pytest

from oracle.mesh import OracleMesh
from warpcfd_amd.preprocess.grid import structured_hex, cavity_mesh
from warpcfd_amd.preprocess.mesh import Mesh
from warpcfd_amd.preprocess.layout import DeviceLayout

INT_TABLES = ['adjacent_cells', 'cell_face_index', 'neighbors', 'face_side', 'nnz_per_cell',
              'face_offset_index', 'is_boundary', 'unique_faces', 'cell_faces']


def test_uniform_grid_closed_forms():
    n = (4, 3, 2)
    L = (2.0, 1.5, 1.0)
    nodes, cells = structured_hex(n, L)
    m = OracleMesh(nodes, cells)
    h = np.array(L) / np.array(n)
    C = np.prod(n)
    assert m.num_cells == C
    Fi = (n[0] - 1) * n[1] * n[2] + n[0] * (n[1] - 1) * n[2] + n[0] * n[1] * (n[2] - 1)
    Fb = 2 * (n[0] * n[1] + n[1] * n[2] + n[0] * n[2])
    assert len(m.internal_face_ids) == Fi and len(m.boundary_face_ids) == Fb
    np.testing.assert_allclose(m.volumes, np.prod(h), rtol=1e-6)
    np.testing.assert_allclose(m.volumes.astype(np.float64).sum(), np.prod(L), rtol=1e-6)
    # hex local faces: 0 bottom (-z), 1 top (+z), 2 front (-y), 3 right (+x), 4 back (+y), 5 left (-x)
    expect_n = np.array([[0, 0, -1], [0, 0, 1], [0, -1, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0]], float)
    expect_a = np.array([h[0] * h[1], h[0] * h[1], h[0] * h[2], h[1] * h[2], h[0] * h[2], h[1] * h[2]])
    np.testing.assert_allclose(m.normal, np.broadcast_to(expect_n, m.normal.shape), atol=1e-6)
    np.testing.assert_allclose(m.area, np.broadcast_to(expect_a, m.area.shape), rtol=1e-6)
    psi = m.norm_distance[m.internal_face_ids, 0]
    np.testing.assert_allclose(psi, 0.5, rtol=1e-6)
    # owner = lower cell id; distance = x_N - x_O is +h along one axis
    fi = m.internal_face_ids
    assert np.all(m.adjacent_cells[fi, 0] < m.adjacent_cells[fi, 1])
    d = m.cell_distance[fi]
    np.testing.assert_allclose(np.sort(np.abs(d), axis=1)[:, :2], 0, atol=1e-6)
    assert np.all(d.max(axis=1) > 0)
    # walls
    m.define_boundary_walls()
    sizes = {k: len(v) for k, v in m.groups.items()}
    assert sizes == {'+X': n[1] * n[2], '-X': n[1] * n[2], '+Y': n[0] * n[2], '-Y': n[0] * n[2],
                     '+Z': n[0] * n[1], '-Z': n[0] * n[1]}


def test_closed_cells_and_face_ids_jittered():
    nodes, cells = structured_hex((5, 4, 3), (1.0, 1.0, 1.0), jitter=0.15, seed=4)
    m = OracleMesh(nodes, cells, geom_dtype=np.float64)
    # sum over the 6 faces of A n vanishes for every closed cell
    s = (m.area[:, :, None] * m.normal).sum(axis=1)
    assert np.abs(s).max() < 1e-12
    # face ids are the lexicographic order of the sorted node tuples
    u = m.unique_faces
    assert np.all(np.diff(u[:, 0]) >= 0)
    keys = [tuple(r) for r in u]
    assert keys == sorted(keys)
    # every internal face: the owner's normal points towards the neighbour
    fi = m.internal_face_ids
    n_o = m.normal[m.adjacent_cells[fi, 0], m.cell_face_index[fi, 0]]
    assert np.all((n_o * m.cell_distance[fi]).sum(axis=1) > 0)


@pytest.mark.parametrize('shape,jitter', [((1, 1, 1), 0.0), ((2, 2, 1), 0.0), ((3, 3, 1), 0.1),
                                          ((3, 2, 2), 0.12)])
def test_vectorised_pairs_equal_reference_loop(shape, jitter):
    nodes, cells = structured_hex(shape, (1.0, 1.0, 1.0), jitter=jitter, seed=2)
    a = OracleMesh(nodes, cells, pair_loop=True)
    b = OracleMesh(nodes, cells)
    for name in INT_TABLES + ['distance', 'norm_distance']:
        assert np.array_equal(getattr(a, name), getattr(b, name)), name


def test_csr_pattern_closed_form():
    nodes, cells = structured_hex((4, 3, 2), (1.0, 1.0, 1.0), jitter=0.1, seed=1)
    m = OracleMesh(nodes, cells)
    for O in (1, 3):
        offs, cols, rows = m.csr_pattern(O)
        C = m.num_cells
        counts = np.repeat(1 + (m.neighbors >= 0).sum(axis=1), O)
        assert np.array_equal(np.diff(offs), counts)
        for r in range(C * O):
            c, k = divmod(r, O)
            nb = m.neighbors[c][m.neighbors[c] >= 0]
            expect = sorted(O * np.concatenate([[c], nb]) + k)
            assert list(cols[offs[r]:offs[r + 1]]) == expect


def test_relabelling_invariance():
    """Shuffling the cell order permutes the tables consistently."""
    nodes, cells = structured_hex((4, 3, 2), (1.0, 1.0, 1.0), jitter=0.1, seed=6)
    rng = np.random.default_rng(0)
    perm = rng.permutation(cells.shape[0])
    a = OracleMesh(nodes, cells)
    b = OracleMesh(nodes, cells[perm])
    assert np.array_equal(a.volumes[perm], b.volumes)
    assert np.array_equal(a.area[perm], b.area)
    assert np.array_equal(a.normal[perm], b.normal)
    inv = np.argsort(perm)
    nb_a = a.neighbors[perm]
    nb_b = np.where(b.neighbors >= 0, perm[np.maximum(b.neighbors, 0)], -1)
    assert np.array_equal(nb_a, nb_b)
    assert sorted(a.nnz_per_cell[perm]) == sorted(b.nnz_per_cell)
    del inv


@pytest.mark.parametrize('jitter', [0.0, 0.1])
def test_product_mesh_matches_oracle(jitter):
    nodes, cells = structured_hex((6, 5, 3), (1.0, 0.7, 0.5), jitter=jitter, seed=8)
    om = OracleMesh(nodes, cells, pair_loop=True)
    pm = Mesh(nodes, cells)
    fp, cp = pm.face_properties, pm.cell_properties
    prod = {'adjacent_cells': fp.adjacent_cells, 'cell_face_index': fp.cell_face_index,
            'neighbors': cp.neighbors, 'face_side': cp.face_side, 'nnz_per_cell': cp.nnz_per_cell,
            'face_offset_index': cp.face_offset_index, 'is_boundary': fp.is_boundary,
            'unique_faces': fp.unique_faces, 'cell_faces': cp.faces}
    for k, v in prod.items():
        assert np.array_equal(getattr(om, k), v), k
    for k, v in {'distance': fp.distance, 'normal': cp.normal, 'area': cp.area,
                 'c2f': cp.cell_centroid_to_face_centroid, 'face_centroid': fp.centroid,
                 'centroids': cp.centroids, 'volumes': cp.volumes}.items():
        assert np.array_equal(getattr(om, k), v), k
    area, psi = pm.face_struct_arrays()
    assert np.array_equal(om.norm_distance[:, 0], psi)
    assert np.array_equal(om.face_area, area)
    for O in (1, 3):
        r1, c1, _ = om.csr_pattern(O)
        r2, c2 = pm.csr_pattern(O)
        assert np.array_equal(r1, r2) and np.array_equal(c1, c2)


@pytest.mark.parametrize('renumber', [False, True])
def test_device_layout_consistency(renumber):
    nodes, cells = structured_hex((6, 5, 4), (1.0, 1.0, 1.0), jitter=0.1, seed=9)
    pm = Mesh(nodes, cells)
    L = DeviceLayout(pm, renumber=renumber)
    fp, cp = pm.face_properties, pm.cell_properties
    C = cp.num_cells
    # permutations are inverse of each other
    assert np.array_equal(L.old2new[L.new2old], np.arange(C))
    assert np.array_equal(L.face_old2new[L.face_new2old], np.arange(fp.num_faces))
    # internal faces first, owners non-decreasing, owner role kept from the reference
    fo = L.face_new2old
    assert np.all(fp.is_boundary[fo[:L.n_internal]] == 0) and np.all(fp.is_boundary[fo[L.n_internal:]] == 1)
    assert np.all(np.diff(L.owner[:L.n_internal]) >= 0)
    assert np.array_equal(L.new2old[L.owner], fp.adjacent_cells[fo, 0])
    assert np.array_equal(L.new2old[L.neigh], fp.adjacent_cells[fo[:L.n_internal], 1])
    # slots: the face in slot k of device cell c is reference face cells[c_ref, k]
    for k in range(6):
        assert np.array_equal(fo[L.slot_face[k]], cp.faces[L.new2old, k])
    # CSR positions point at (c, neighbour) and (c, c)
    rows = np.repeat(np.arange(C), np.diff(L.row_ptr))
    assert np.array_equal(L.col[L.diag_pos], np.arange(C))
    for k in range(6):
        ok = L.slot_pos[k] >= 0
        assert np.array_equal(rows[L.slot_pos[k][ok]], np.arange(C)[ok])
        assert np.array_equal(L.col[L.slot_pos[k][ok]], L.slot_nbr[k][ok])
    # ELL neighbour tables: boundary slots point at the row itself, int16
    # offsets equal the int32 neighbours, ELL <-> CSR conversion is lossless
    for k in range(6):
        bnd = L.slot_pos[k] < 0
        assert np.array_equal(L.slot_nbr[k][bnd], np.arange(C)[bnd])
    assert L.slot_dnb is not None
    assert np.array_equal(L.slot_dnb.astype(np.int64) + np.arange(C), L.slot_nbr)
    rng = np.random.default_rng(3)
    csr = rng.standard_normal(L.nnz).astype(np.float32)
    ell = np.zeros((7, C), np.float32)
    ell[0] = csr[L.diag_pos]
    for k in range(6):
        ok = L.slot_pos[k] >= 0
        ell[k + 1][ok] = csr[L.slot_pos[k][ok]]
    assert np.array_equal(L.ell_to_csr(ell), csr)
    # rows sorted and a symmetric pattern
    for c in range(C):
        seg = L.col[L.row_ptr[c]:L.row_ptr[c + 1]]
        assert np.all(np.diff(seg) > 0)
    if renumber:
        # RCM narrows the band compared with the lexicographic order
        bw = np.abs(L.col - rows).max()
        L0 = DeviceLayout(pm, renumber=False)
        rows0 = np.repeat(np.arange(C), np.diff(L0.row_ptr))
        assert bw <= np.abs(L0.col - rows0).max()


def test_cavity_groups():
    from warpcfd_amd.preprocess.grid import define_boundary_walls
    nodes, cells = cavity_mesh(8)
    pm = Mesh(nodes, cells)
    walls = define_boundary_walls(pm)
    assert {k: len(v) for k, v in walls.items()} == {'+X': 8, '-X': 8, '+Y': 8, '-Y': 8, '+Z': 64, '-Z': 64}
    om = OracleMesh(nodes, cells)
    for k, v in om.define_boundary_walls().items():
        assert np.array_equal(v, walls[k])


def test_product_connectivity_matches_oracle_full_channel():
    """The native face matching at the full 200 x 40 x 40 channel (320 000
    cells, 977 600 faces): face numbering, owner/neighbour pairs and the
    per-cell tables bit-exact against the oracle's np.unique restatement."""
    from warpcfd_amd.preprocess.grid import channel_mesh
    nodes, cells = channel_mesh(200, 40, 40)
    om = OracleMesh(nodes, cells)
    pm = Mesh(nodes, cells)
    fp, cp = pm.face_properties, pm.cell_properties
    assert fp.num_faces == 977600
    for k, v in {'adjacent_cells': fp.adjacent_cells, 'cell_face_index': fp.cell_face_index,
                 'neighbors': cp.neighbors, 'face_side': cp.face_side,
                 'is_boundary': fp.is_boundary, 'unique_faces': fp.unique_faces,
                 'cell_faces': cp.faces}.items():
        assert np.array_equal(getattr(om, k), v), k
