"""Host mesh topology and geometry in the reference's order.

Mirrors ``warp_cfd.preprocess.Mesh`` (preprocess/mesh.py:17-302): face ids in
lexicographic order of sorted node tuples (mesh.py:115-121), per face set
(a wedge's triangles first, then its quadrilaterals, mesh.py:101-104,
142-157), owner = lower cell id (mesh.py:205,280), fan-triangulated face
areas and normals per cell-local face (mesh.py:161-183), face centroid =
node mean (mesh.py:128), cell->face-centroid vectors (mesh.py:140),
neighbour / face_side / nnz / face_offset_index tables (mesh.py:259-300).
The Python pair loop of mesh.py:263-281 is replaced by a sort over face
ids, which gives identical integer tables.

Cells are hexahedra (8 nodes, 6 faces), tetrahedra (4 nodes, 4 faces) or
wedges (6 nodes, 5 faces), in VTK node order, one type per mesh
(mesh.py:20).  pyvista is not available, so the mesh is built from raw
arrays: ``Mesh(nodes, cells)``.  An object with ``.points`` and
``.cells_dict`` (a pyvista grid) is accepted as well.
"""
import numpy as np

VTK_TETRA, VTK_HEXAHEDRON, VTK_WEDGE = 10, 12, 13

#: local faces of a VTK hexahedron, outward by the right-hand rule (mesh.py:69-76)
HEX_FACE_NODES = np.array([[3, 2, 1, 0], [4, 5, 6, 7], [0, 1, 5, 4],
                           [1, 2, 6, 5], [3, 7, 6, 2], [3, 0, 4, 7]], dtype=np.int32)
#: tetrahedron (mesh.py:78-83)
TET_FACE_NODES = np.array([[2, 1, 0], [0, 1, 3], [3, 2, 0], [1, 2, 3]], dtype=np.int32)
#: wedge: two triangles, then three quadrilaterals (mesh.py:85-91)
WEDGE_TRI_NODES = np.array([[0, 2, 1], [3, 4, 5]], dtype=np.int32)
WEDGE_QUAD_NODES = np.array([[0, 3, 5, 2], [1, 2, 5, 4], [0, 1, 4, 3]], dtype=np.int32)

#: VTK type -> (name, nodes per cell, face sets in local face order)
CELL_TYPES = {
    VTK_HEXAHEDRON: ('HEXAHEDRON', 8, (HEX_FACE_NODES,)),
    VTK_TETRA: ('TETRA', 4, (TET_FACE_NODES,)),
    VTK_WEDGE: ('WEDGE', 6, (WEDGE_TRI_NODES, WEDGE_QUAD_NODES)),
}
_BY_NODES = {v[1]: k for k, v in CELL_TYPES.items()}


def cell_type_of(cells):
    """VTK type of a (C, n) cell array: n = 8, 4 or 6."""
    n = np.asarray(cells).shape[-1]
    if n not in _BY_NODES:
        raise ValueError('only Tetra and Hexahedron and Wedge elements are supported '
                         '(cells of 4, 8 or 6 nodes)')
    return _BY_NODES[n]


def local_faces(cell_type):
    """Local faces of a cell type in face order, each a node list."""
    return [list(r) for t in CELL_TYPES[cell_type][2] for r in t]
class group:
    """preprocess/mesh.py:8-15."""

    def __init__(self, name, ids, type):
        self.name = name
        self.ids = ids
        self.type = type


class _Props:
    pass


def match_faces(local):
    """Unique faces of one face set and the slots holding each, by
    wcfd_match_faces (host code in libwcfd.so).  local: (C, Fs, Ns) node
    ids.  Returns (unique (K, Ns) ascending node ids in lexicographic row
    order, face id per slot (C, Fs), slots (K, 2) c * Fs + k ascending, -1
    in the second for a boundary face): the numbering of np.unique over the
    sorted rows (mesh.py:108-148) and the pairs of mesh.py:263-288."""
    from .. import _lib
    lib = _lib.load()
    local = np.ascontiguousarray(local, dtype=np.int32)
    C, Fs, Ns = local.shape
    cap = C * Fs
    face_of = np.empty((C, Fs), np.int32)
    uniq = np.empty((cap, Ns), np.int32)
    slots = np.empty((cap, 2), np.int32)
    k = lib.wcfd_match_faces(local.ctypes.data, C, Fs, Ns, face_of.ctypes.data,
                             uniq.ctypes.data, slots.ctypes.data, cap)
    if k < 0:
        if -k == _lib.WCFD_EMESH:
            raise AssertionError('a face is shared by more than two cells')
        _lib.check(-k, 'match_faces')
    return uniq[:k].copy(), face_of, slots[:k].astype(np.int64)


def _fan(coords):
    v0 = coords[..., 0, :]
    parts = [0.5 * np.cross(coords[..., i, :] - v0, coords[..., i + 1, :] - v0)
             for i in range(1, coords.shape[-2] - 1)]
    area = sum(np.linalg.norm(p, axis=-1) for p in parts)
    return area, sum(parts) / area[..., None]


def _volume(x, sets):
    """Divergence theorem over the fan triangles of the faces, float64.
    (Stand-in for VTK compute_cell_sizes, mesh.py:33; exact for tetrahedra
    and planar faces.)"""
    vol = np.zeros(x.shape[0])
    for t in sets:
        fc = x[:, t]
        v0 = fc[:, :, 0]
        for i in range(1, t.shape[1] - 1):
            cr = np.cross(fc[:, :, i] - v0, fc[:, :, i + 1] - v0)
            vol += np.einsum('cfk,cfk->c', cr, v0 + fc[:, :, i] + fc[:, :, i + 1]) / 18.0
    return np.abs(vol)


class Mesh:
    def __init__(self, mesh, cells=None, dtype=np.float32):
        if cells is None:
            pts = np.asarray(mesh.points)
            cd = mesh.cells_dict
            if len(cd) != 1:
                raise AssertionError('Meshes can only contain a single cell type for now, Got multiple')
            ctype, cells = next(iter(cd.items()))
            if int(ctype) not in CELL_TYPES:
                raise AssertionError('only Tetra and Hexahedron and Wedge elements are supported')
            nodes = pts
        else:
            nodes = mesh
        self.float_dtype = np.float32
        self.int_dtype = np.int32
        self.dtype = dtype
        self.nodes = np.asarray(nodes, dtype=np.float32)            # mesh.py:26 (T13)
        self.cells = np.asarray(cells, dtype=np.int32)               # mesh.py:31
        if self.cells.ndim != 2:
            raise ValueError('cells must be a (C, n) array of VTK cells')
        ct = cell_type_of(self.cells)
        self.cellType = (CELL_TYPES[ct][0], ct)
        self._sets = CELL_TYPES[ct][2]
        x = self.nodes.astype(np.float64)[self.cells]
        # VTK cell centre = the parametric centre = the vertex mean for all three types
        self.cell_centroids = x.mean(axis=1).astype(np.float32)
        self.cell_volumes = _volume(x, self._sets).astype(np.float32)
        self.groups = {}
        self.gridType = 'Unstructured'
        self.face_properties, self.cell_properties = self.get_mesh_properties()

    @property
    def dimension(self):
        assert self.nodes.shape[-1] == 3, 'Only 3 Dimensional Meshes are supported'
        return 3

    def add_group(self, name, id, group_type):                      # mesh.py:53-57
        assert group_type in {'face', 'cell', 'point', 'edge'}
        assert name not in self.groups, 'name for group already exists!'
        self.groups[name] = group(name, np.asarray(id, np.int32), group_type)

    def add_groups(self, dic, group_type):                          # mesh.py:60-62
        for name, ids in dic.items():
            self.add_group(name, ids, group_type)

    @property
    def boundary_face_ids(self):
        return self.face_properties.boundary_face_ids

    def boundary_normals(self):
        fp, cp = self.face_properties, self.cell_properties
        b = fp.boundary_face_ids
        return cp.normal[fp.adjacent_cells[b, 0], fp.cell_face_index[b, 0]]

    @property
    def faces_per_cell(self):
        return sum(t.shape[0] for t in self._sets)

    def get_mesh_properties(self):
        C = self.cells.shape[0]
        N = max(t.shape[1] for t in self._sets)
        Fc = self.faces_per_cell
        parts, shift, col, pairs = [], 0, 0, []
        for t in self._sets:                                          # mesh.py:108-148
            local = self.cells[:, t]                                  # (C,Fs,Ns)
            Fs = t.shape[0]
            uq, cf, slots = match_faces(local)
            # the face's slots as flat indices of cell_faces (C, Fc)
            pairs.append(np.where(slots >= 0, (slots // Fs) * Fc + col + slots % Fs, -1))
            col += Fs
            a, n = _fan(self.nodes[local])
            fc = self.nodes[uq].mean(axis=1)
            c2f_s = fc[cf] - self.cell_centroids[:, None, :]
            nn = np.full(len(uq), t.shape[1], np.int32)
            if t.shape[1] < N:                                        # tri faces padded with -1
                uq = np.hstack([uq, -np.ones((len(uq), N - t.shape[1]), uq.dtype)])
            parts.append((uq, cf + shift, a, n, fc, c2f_s, nn))
            shift += len(uq)
        uniq = np.concatenate([p[0] for p in parts], axis=0)
        cell_faces = np.concatenate([p[1] for p in parts], axis=1).astype(np.int32)
        area = np.concatenate([p[2] for p in parts], axis=1)
        normal = np.concatenate([p[3] for p in parts], axis=1)
        fcen = np.concatenate([p[4] for p in parts], axis=0)
        c2f = np.concatenate([p[5] for p in parts], axis=1)
        num_nodes = np.concatenate([p[6] for p in parts])
        K = uniq.shape[0]

        # the two (cell, local face) slots of every shared face
        pairs = np.concatenate(pairs, axis=0)
        fi = np.nonzero(pairs[:, 1] >= 0)[0].astype(np.int32)
        a, b = pairs[fi, 0], pairs[fi, 1]                  # slot a has the lower cell id
        ca, ka, cb, kb = a // Fc, a % Fc, b // Fc, b % Fc

        neighbors = np.full((C, Fc), -1, np.int32)
        neighbors[ca, ka] = cb
        neighbors[cb, kb] = ca
        face_side = np.zeros((C, Fc), np.int32)
        face_side[cb, kb] = 1
        adjacent = np.full((K, 2), -1, np.int32)
        cfi = np.full((K, 2), -1, np.int32)
        adjacent[fi, 0], adjacent[fi, 1] = ca, cb
        cfi[fi, 0], cfi[fi, 1] = ka, kb
        is_b = np.ones(K, np.uint8)
        is_b[fi] = 0
        bf = np.nonzero(pairs[:, 1] < 0)[0]
        sb = pairs[bf, 0]
        adjacent[bf, 0] = sb // Fc
        cfi[bf, 0] = sb % Fc
        dist = np.full((K, 3), -1.0, np.float32)
        dist[fi] = self.cell_centroids[cb] - self.cell_centroids[ca]
        mask = neighbors != -1

        fp = _Props()
        fp.num_faces = K
        fp.unique_faces = uniq.astype(np.int32)
        fp.centroid = fcen.astype(np.float32)
        fp.num_nodes = num_nodes
        fp.adjacent_cells = adjacent
        fp.cell_face_index = cfi
        fp.distance = dist
        fp.is_boundary = is_b
        ids = np.arange(K, dtype=np.int32)
        fp.boundary_face_ids = ids[is_b == 1]
        fp.internal_face_ids = ids[is_b == 0]
        fp.num_boundary_faces = len(fp.boundary_face_ids)
        fp.dimension = 3

        cp = _Props()
        cp.num_cells = C
        cp.faces_per_cell = Fc
        cp.nodes_per_face = N
        cp.dimension = 3
        cp.centroids = self.cell_centroids
        cp.volumes = self.cell_volumes
        cp.nodes = self.cells
        cp.faces = cell_faces
        cp.normal = normal.astype(np.float32)
        cp.area = area.astype(np.float32)
        cp.cell_centroid_to_face_centroid = c2f.astype(np.float32)
        cp.neighbors = neighbors
        cp.face_side = face_side
        cp.nnz_per_cell = (1 + mask.sum(axis=1)).astype(np.int32)
        cp.face_offset_index = (np.cumsum(mask, axis=1) * mask).astype(np.int32)
        return fp, cp

    # ------------------------------------------------------------ face structs
    def face_struct_arrays(self):
        """Face-struct members of mesh_structs.py:190-210: owner-side area and
        the interpolation weight psi = |x_f - x_O| / |x_O - x_N| (T1)."""
        fp, cp = self.face_properties, self.cell_properties
        own, k0 = fp.adjacent_cells[:, 0], fp.cell_face_index[:, 0]
        area = cp.area[own, k0]
        psi = np.zeros(fp.num_faces, np.float32)
        fi = fp.internal_face_ids
        o, n = own[fi], fp.adjacent_cells[fi, 1]
        v = cp.cell_centroid_to_face_centroid[o, k0[fi]]
        dd = cp.centroids[o] - cp.centroids[n]
        psi[fi] = np.sqrt((v * v).sum(axis=1, dtype=np.float32)) / np.sqrt((dd * dd).sum(axis=1, dtype=np.float32))
        return area, psi

    def csr_pattern(self, num_outputs=1):
        """The CSR pattern the reference builds (equation_kernels.py:10-38 then
        bsr_from_triplets): row O*c+k holds columns sorted({O*j+k : j in {c} U N(c)}).
        Returns (row_ptr, columns) int32."""
        cp = self.cell_properties
        C, O = cp.num_cells, num_outputs
        nb = np.sort(np.where(cp.neighbors >= 0, cp.neighbors, np.iinfo(np.int32).max), axis=1)
        cols_cell = np.concatenate([np.arange(C, dtype=np.int64)[:, None], nb], axis=1)
        cols_cell = np.sort(cols_cell, axis=1)                     # self + neighbours, sorted
        valid = cols_cell != np.iinfo(np.int32).max
        counts = valid.sum(axis=1)
        rows_cols = [np.where(valid, O * cols_cell + k, -1) for k in range(O)]
        stacked = np.stack(rows_cols, axis=1).reshape(C * O, cols_cell.shape[1])
        row_ptr = np.zeros(C * O + 1, np.int64)
        row_ptr[1:] = np.cumsum(np.repeat(counts, O))
        cols = stacked[stacked >= 0]
        return row_ptr.astype(np.int32), cols.astype(np.int32)
