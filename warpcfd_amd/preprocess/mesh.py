"""Host mesh topology and geometry in the reference's order.

Mirrors ``warp_cfd.preprocess.Mesh`` (preprocess/mesh.py:17-302): face ids in
lexicographic order of sorted node tuples (mesh.py:115-121), owner = lower
cell id (mesh.py:205,280), fan-triangulated face areas and normals per
cell-local face (mesh.py:161-183), face centroid = node mean (mesh.py:128),
cell->face-centroid vectors (mesh.py:140), neighbour / face_side / nnz /
face_offset_index tables (mesh.py:259-300).  The Python pair loop of
mesh.py:263-281 is replaced by a sort over face ids, which gives identical
integer tables.

pyvista is not available, so the mesh is built from raw arrays:
``Mesh(nodes, cells)`` with ``cells`` (C,8) VTK hexahedra.  An object with
``.points`` and ``.cells_dict`` (a pyvista grid) is accepted as well.
"""
import numpy as np

VTK_HEXAHEDRON = 12

#: local faces of a VTK hexahedron, outward by the right-hand rule (mesh.py:69-76)
HEX_FACE_NODES = np.array([[3, 2, 1, 0], [4, 5, 6, 7], [0, 1, 5, 4],
                           [1, 2, 6, 5], [3, 7, 6, 2], [3, 0, 4, 7]], dtype=np.int32)


class group:
    """preprocess/mesh.py:8-15."""

    def __init__(self, name, ids, type):
        self.name = name
        self.ids = ids
        self.type = type


class _Props:
    pass


def _fan(coords):
    v0 = coords[..., 0, :]
    parts = [0.5 * np.cross(coords[..., i, :] - v0, coords[..., i + 1, :] - v0)
             for i in range(1, coords.shape[-2] - 1)]
    area = sum(np.linalg.norm(p, axis=-1) for p in parts)
    return area, sum(parts) / area[..., None]


def _hex_volume(x):
    """Divergence theorem over the fan triangles of the 6 faces, float64.
    (Stand-in for VTK compute_cell_sizes, mesh.py:33.)"""
    fc = x[:, HEX_FACE_NODES]
    v0 = fc[:, :, 0]
    vol = np.zeros(x.shape[0])
    for i in (1, 2):
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
            if int(ctype) != VTK_HEXAHEDRON:
                raise ValueError('only hexahedral meshes are supported')
            nodes = pts
        else:
            nodes = mesh
        self.float_dtype = np.float32
        self.int_dtype = np.int32
        self.dtype = dtype
        self.cellType = ('HEXAHEDRON', VTK_HEXAHEDRON)
        self.nodes = np.asarray(nodes, dtype=np.float32)            # mesh.py:26 (T13)
        self.cells = np.asarray(cells, dtype=np.int32)               # mesh.py:31
        if self.cells.ndim != 2 or self.cells.shape[1] != 8:
            raise ValueError('cells must be a (C,8) array of VTK hexahedra')
        x = self.nodes.astype(np.float64)[self.cells]
        self.cell_centroids = x.mean(axis=1).astype(np.float32)      # VTK cell centre (hex)
        self.cell_volumes = _hex_volume(x).astype(np.float32)
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

    def get_mesh_properties(self):
        C = self.cells.shape[0]
        local = self.cells[:, HEX_FACE_NODES]                        # (C,6,4)
        keyed = np.sort(local, axis=2).reshape(-1, 4)
        uniq, inv = np.unique(keyed, axis=0, return_inverse=True)
        cell_faces = inv.reshape(C, 6).astype(np.int32)
        area, normal = _fan(self.nodes[local])
        fcen = self.nodes[uniq].mean(axis=1)
        c2f = fcen[cell_faces] - self.cell_centroids[:, None, :]
        K = uniq.shape[0]

        # pair the two (cell, local face) slots of every shared face
        flat = cell_faces.ravel()
        order = np.lexsort((np.arange(flat.size), flat))
        counts = np.bincount(flat, minlength=K)
        if counts.max() > 2:
            raise AssertionError('a face is shared by more than two cells')
        first = np.concatenate([[0], np.cumsum(counts)[:-1]])
        fi = np.nonzero(counts == 2)[0].astype(np.int32)
        a, b = order[first[fi]], order[first[fi] + 1]      # slot a has the lower cell id
        ca, ka, cb, kb = a // 6, a % 6, b // 6, b % 6

        neighbors = np.full((C, 6), -1, np.int32)
        neighbors[ca, ka] = cb
        neighbors[cb, kb] = ca
        face_side = np.zeros((C, 6), np.int32)
        face_side[cb, kb] = 1
        adjacent = np.full((K, 2), -1, np.int32)
        cfi = np.full((K, 2), -1, np.int32)
        adjacent[fi, 0], adjacent[fi, 1] = ca, cb
        cfi[fi, 0], cfi[fi, 1] = ka, kb
        is_b = np.ones(K, np.uint8)
        is_b[fi] = 0
        bf = np.nonzero(counts == 1)[0]
        sb = order[first[bf]]
        adjacent[bf, 0] = sb // 6
        cfi[bf, 0] = sb % 6
        dist = np.full((K, 3), -1.0, np.float32)
        dist[fi] = self.cell_centroids[cb] - self.cell_centroids[ca]
        mask = neighbors != -1

        fp = _Props()
        fp.num_faces = K
        fp.unique_faces = uniq.astype(np.int32)
        fp.centroid = fcen.astype(np.float32)
        fp.num_nodes = np.full(K, 4, np.int32)
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
        cp.faces_per_cell = 6
        cp.nodes_per_face = 4
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
        stacked = np.stack(rows_cols, axis=1).reshape(C * O, 7)
        row_ptr = np.zeros(C * O + 1, np.int64)
        row_ptr[1:] = np.cumsum(np.repeat(counts, O))
        cols = stacked[stacked >= 0]
        return row_ptr.astype(np.int32), cols.astype(np.int32)
