"""Oracle restatement of the reference mesh preprocessing (test infrastructure).

Follows preprocess/mesh.py (Mesh.__init__ 18-43, get_faces 64-159,
calculate_face_normal_and_area 161-183, get_neighbors 186-205,
get_mesh_properties 220-302), FV/mesh_structs.py:init_structs 141-251 (psi,
face area) and preprocess/grid.py:define_boundary_walls 165-194.

Hexahedra, tetrahedra and wedges follow get_faces' three branches (the
local face tables of mesh.py:69-91; a wedge's two triangles form the first
face set and its three quadrilaterals the second, mesh.py:101-104).  The
reference's own struct selection rejects its tet and wedge meshes
(mesh_structs.py:112-132 asks a tetrahedron for 3 faces and a wedge for
3-node faces); the arrays restated here are those the rest of the reference
computes for them, which is what a corrected selection (4 faces of 3 nodes;
5 faces, padded to 4 nodes) would hand its kernels.
"""
import numpy as np

# Local face table of a VTK hexahedron, mesh.py:69-76.
HEX_FACES = np.array([[3, 2, 1, 0],
                      [4, 5, 6, 7],
                      [0, 1, 5, 4],
                      [1, 2, 6, 5],
                      [3, 7, 6, 2],
                      [3, 0, 4, 7]], dtype=np.int32)
# VTK tetrahedron, mesh.py:78-83.
TET_FACES = np.array([[2, 1, 0], [0, 1, 3], [3, 2, 0], [1, 2, 3]], dtype=np.int32)
# VTK wedge, mesh.py:85-91 (-1 pads the two triangles).
WEDGE_FACES = np.array([[0, 2, 1, -1], [3, 4, 5, -1], [0, 3, 5, 2], [1, 2, 5, 4],
                        [0, 1, 4, 3]], dtype=np.int32)

#: nodes per cell -> (reference cell type name, VTK type, face sets in local order)
CELL_TYPES = {
    8: ('HEXAHEDRON', 12, [HEX_FACES]),
    4: ('TETRA', 10, [TET_FACES]),
    6: ('WEDGE', 13, [WEDGE_FACES[0:2, :-1], WEDGE_FACES[2:, :]]),
}


def face_sets(cells):
    """The face sets of get_faces (mesh.py:101-104) for a (C, n) cell array."""
    n = np.asarray(cells).shape[-1]
    if n not in CELL_TYPES:
        raise ValueError('only tetrahedra (4 nodes), wedges (6) and hexahedra (8) are supported')
    return CELL_TYPES[n][2]


def fan_area_normal(face_coords):
    """mesh.py:161-183: fan triangulation from node 0.

    area = sum_i |1/2 (v_i - v0) x (v_{i+1} - v0)|, normal = sum_i (1/2 cross) / area
    (not unit length for a warped face).  Arithmetic stays in the dtype of
    ``face_coords``, as numpy does for the reference's float32 nodes.
    """
    v0 = face_coords[..., 0, :]
    areas = []
    normals = []
    for i in range(1, face_coords.shape[-2] - 1):
        vi = face_coords[..., i, :]
        vii = face_coords[..., i + 1, :]
        nv = 1 / 2 * np.cross(vi - v0, vii - v0)
        areas.append(np.linalg.norm(nv, axis=-1))
        normals.append(nv)
    area = sum(areas)
    normal = sum(normals) / area[..., np.newaxis]
    return area, normal


def cell_volume(nodes, cells):
    """Cell volume by the divergence theorem over the fan triangles of the
    cell's faces (V = 1/3 sum_tri (x_c . n_tri A_tri)), evaluated in float64.
    Stand-in for VTK compute_cell_sizes (mesh.py:33), which is absent; exact
    for a tetrahedron and for cells with planar faces."""
    x = nodes.astype(np.float64)[cells]               # (C,n,3)
    vol = np.zeros(cells.shape[0])
    for table in face_sets(cells):
        fc = x[:, table]                               # (C,Fs,Ns,3)
        v0 = fc[:, :, 0, :]
        for i in range(1, table.shape[1] - 1):
            a = fc[:, :, i, :] - v0
            b = fc[:, :, i + 1, :] - v0
            cr = np.cross(a, b)                         # 2 A n for the triangle
            tri_c = (v0 + fc[:, :, i, :] + fc[:, :, i + 1, :]) / 3.0
            vol += (cr * tri_c).sum(axis=(1, 2)) / 6.0
    return np.abs(vol)


def hex_volume(nodes, cells):
    return cell_volume(nodes, cells)


class OracleMesh:
    """Reference-order mesh arrays (hexahedra, tetrahedra or wedges).

    Attribute names follow cells_data / faces_data (mesh.py:306-510) and the
    Cell/Face structs (mesh_structs.py:5-64).
    """

    def __init__(self, nodes, cells, geom_dtype=np.float32, pair_loop=False):
        gd = np.dtype(geom_dtype)
        self.geom_dtype = gd
        # mesh.py:26,31 -- the reference always casts to float32/int32 (T13).
        self.nodes = np.asarray(nodes, dtype=gd)
        self.cells = np.asarray(cells, dtype=np.int32)
        C = self.cells.shape[0]
        self.num_cells = C
        self.cell_type = CELL_TYPES[self.cells.shape[1]][0] if self.cells.shape[1] in CELL_TYPES \
            else None
        sets = face_sets(self.cells)
        # mesh.py:32 (VTK cell_centers: the parametric centre, which is the
        # vertex mean for a hexahedron, a tetrahedron and a wedge).
        self.centroids = self.nodes.astype(np.float64)[self.cells].mean(axis=1).astype(gd)
        # mesh.py:33 (VTK compute_cell_sizes, abs) -> divergence theorem.
        self.volumes = cell_volume(self.nodes, self.cells).astype(gd)

        # ---- get_faces, mesh.py:99-159: one face set, or two for a wedge ----
        N = max(t.shape[1] for t in sets)
        parts, shift = [], 0
        for table in sets:
            faces = self.cells[:, table]                          # (C,Fs,Ns)
            ns = table.shape[1]
            faces_reshaped = np.sort(faces, axis=2).reshape(-1, ns)
            uf, inv = np.unique(faces_reshaped, axis=0, return_inverse=True)
            cfi = inv.reshape(C, table.shape[0]).astype(np.int32)
            area, normal = fan_area_normal(self.nodes[faces])      # (C,Fs), (C,Fs,3)
            fcen = self.nodes[uf].mean(axis=1)                     # mesh.py:128
            to_face = fcen[cfi] - self.centroids[:, np.newaxis, :]  # mesh.py:140
            if ns < N:                                             # mesh.py:131-135
                uf = np.hstack([uf, -np.ones((uf.shape[0], N - ns), uf.dtype)])
            parts.append((uf, cfi + shift, area, normal, np.full(len(uf), ns, np.int32), fcen,
                          to_face))
            shift += len(uf)
        unique_faces = np.concatenate([p[0] for p in parts], axis=0)
        cell_face_ids = np.concatenate([p[1] for p in parts], axis=1).astype(np.int32)

        self.unique_faces = unique_faces.astype(np.int32)
        self.cell_faces = cell_face_ids
        self.area = np.concatenate([p[2] for p in parts], axis=1).astype(gd)
        self.normal = np.concatenate([p[3] for p in parts], axis=1).astype(gd)
        self.face_num_nodes = np.concatenate([p[4] for p in parts])
        self.face_centroid = np.concatenate([p[5] for p in parts], axis=0).astype(gd)
        self.c2f = np.concatenate([p[6] for p in parts], axis=1).astype(gd)
        K = unique_faces.shape[0]
        self.num_faces = K
        Fc = cell_face_ids.shape[1]
        self.faces_per_cell = Fc
        self.nodes_per_face = N

        # ---- get_mesh_properties, mesh.py:220-302 ----
        self.adjacent_cells = -np.ones((K, 2), np.int32)
        self.cell_face_index = -np.ones((K, 2), np.int32)
        self.distance = -np.ones((K, 3), gd)
        self.is_boundary = np.ones(K, np.uint8)
        self.neighbors = -np.ones((C, Fc), np.int32)
        self.face_side = np.zeros((C, Fc), np.int32)
        self.nnz_per_cell = np.ones(C, np.int32)

        if pair_loop:
            self._pairs_loop()
        else:
            self._pairs_vectorised()

        # boundary faces: owner = the only cell (mesh.py:285-288)
        bc, bk = np.nonzero(self.neighbors == -1)
        bf = self.cell_faces[bc, bk]
        self.adjacent_cells[bf, 0] = bc
        self.cell_face_index[bf, 0] = bk

        ids = np.arange(K, dtype=np.int32)
        self.boundary_face_ids = ids[self.is_boundary == 1]
        self.internal_face_ids = ids[self.is_boundary == 0]
        self.num_boundary_faces = len(self.boundary_face_ids)
        mask = self.neighbors != -1
        self.face_offset_index = (np.cumsum(mask, axis=1) * mask).astype(np.int32)  # mesh.py:299-300

        # ---- init_structs, mesh_structs.py:190-210 ----
        own = self.adjacent_cells[:, 0]
        k0 = self.cell_face_index[:, 0]
        self.face_area = self.area[own, k0]                      # owner-side area (200)
        self.norm_distance = np.zeros((K, 2), gd)
        self.cell_distance = np.zeros((K, 3), gd)
        fi = self.internal_face_ids
        o = own[fi]
        n = self.adjacent_cells[fi, 1]
        num = np.sqrt((self.c2f[o, k0[fi]] ** 2).sum(axis=1, dtype=gd))
        den = np.sqrt(((self.centroids[o] - self.centroids[n]) ** 2).sum(axis=1, dtype=gd))
        psi = (num / den).astype(gd)                               # T1: |x_f-x_O|/|x_O-x_N|
        self.norm_distance[fi, 0] = psi
        self.norm_distance[fi, 1] = gd.type(1.0) - psi
        self.cell_distance[fi] = self.distance[fi]
        self.groups = {}

    # mesh.py:263-281, written as the reference does (small meshes only).
    def _pairs_loop(self):
        cf = self.cell_faces
        K = self.num_faces
        face_cells = [[] for _ in range(K)]
        for c in range(self.num_cells):
            for k in range(self.faces_per_cell):
                face_cells[cf[c, k]].append(c)
        pairs = sorted({tuple(sorted(v)) for v in face_cells if len(v) == 2})
        for i, j in pairs:
            inter, ii, jj = np.intersect1d(cf[i], cf[j], return_indices=True)
            assert ii.shape[0] == 1
            f = inter[0]
            self.neighbors[i, ii[0]] = j
            self.neighbors[j, jj[0]] = i
            self.face_side[i, ii[0]] = 0
            self.face_side[j, jj[0]] = 1
            self.nnz_per_cell[i] += 1
            self.nnz_per_cell[j] += 1
            self.distance[f] = self.centroids[j] - self.centroids[i]
            self.is_boundary[f] = 0
            self.adjacent_cells[f] = [i, j]
            self.cell_face_index[f] = [ii[0], jj[0]]

    # Same result as the loop: every internal face is shared by exactly two
    # cells; owner = lower cell id (T14), face_side 0 for it, 1 for the other.
    def _pairs_vectorised(self):
        cf = self.cell_faces.reshape(-1)
        C = self.num_cells
        order = np.argsort(cf, kind='stable')          # (cell,k) slots grouped by face
        fs = cf[order]
        counts = np.bincount(fs, minlength=self.num_faces)
        if counts.max() > 2:
            raise ValueError('a face is shared by more than two cells')
        start = np.concatenate([[0], np.cumsum(counts)[:-1]])
        internal = np.nonzero(counts == 2)[0]
        s0 = order[start[internal]]
        s1 = order[start[internal] + 1]
        Fc = self.faces_per_cell
        c0, k0 = s0 // Fc, s0 % Fc
        c1, k1 = s1 // Fc, s1 % Fc
        swap = c0 > c1
        ci = np.where(swap, c1, c0); ki = np.where(swap, k1, k0)
        cj = np.where(swap, c0, c1); kj = np.where(swap, k0, k1)
        self.neighbors[ci, ki] = cj
        self.neighbors[cj, kj] = ci
        self.face_side[ci, ki] = 0
        self.face_side[cj, kj] = 1
        self.nnz_per_cell += np.bincount(np.concatenate([ci, cj]), minlength=C).astype(np.int32)
        self.distance[internal] = self.centroids[cj] - self.centroids[ci]
        self.is_boundary[internal] = 0
        self.adjacent_cells[internal, 0] = ci
        self.adjacent_cells[internal, 1] = cj
        self.cell_face_index[internal, 0] = ki
        self.cell_face_index[internal, 1] = kj

    # ---- grid.py:165-194 ----
    def define_boundary_walls(self):
        b = self.boundary_face_ids
        k = self.cell_face_index[b, 0]
        c = self.adjacent_cells[b, 0]
        n = self.normal[c, k]
        tol = 1e-3
        walls = {}
        for name, e in (("+X", [1, 0, 0]), ("-X", [-1, 0, 0]), ("+Y", [0, 1, 0]),
                        ("-Y", [0, -1, 0]), ("+Z", [0, 0, 1]), ("-Z", [0, 0, -1])):
            walls[name] = b[np.abs(n @ np.array(e, dtype=n.dtype) - 1) < tol]
        for name, ids in walls.items():
            assert name not in self.groups
            self.groups[name] = ids
        return walls

    # ---- COO_Arrays + calculate_BSR_matrix_indices + bsr_from_triplets ----
    def csr_pattern(self, num_outputs):
        """utils.py:7-28, equation_kernels.py:10-38, equation.py:53-59.

        Triplets are written exactly as the kernel does, then sorted and merged
        the way bsr_from_triplets does (row-major, columns ascending).
        """
        C = self.num_cells
        O = num_outputs
        nnz_pc = self.nnz_per_cell
        nnz = int(O * nnz_pc.sum())
        offsets = np.zeros(C, np.int64)
        offsets[1:] = np.cumsum(O * nnz_pc)[:-1]
        rows = np.zeros(nnz, np.int64)
        cols = np.zeros(nnz, np.int64)
        for o in range(O):
            tot = offsets + nnz_pc * o
            rows[tot] = O * np.arange(C) + o
            cols[tot] = rows[tot]
            for k in range(self.faces_per_cell):
                f = self.cell_faces[:, k]
                internal = self.is_boundary[f] == 0
                cc = np.nonzero(internal)[0]
                t = offsets[cc] + nnz_pc[cc] * o + self.face_offset_index[cc, k]
                rows[t] = O * cc + o
                cols[t] = O * self.neighbors[cc, k] + o
        key = rows * (C * O) + cols
        order = np.argsort(key, kind='stable')
        rows, cols = rows[order], cols[order]
        if np.any(np.diff(rows * (C * O) + cols) == 0):
            raise AssertionError('duplicate triplet')
        offs = np.zeros(C * O + 1, np.int64)
        np.add.at(offs, rows + 1, 1)
        offs = np.cumsum(offs)
        return offs.astype(np.int32), cols.astype(np.int32), rows.astype(np.int32)
