"""Device layout: RCM-renumbered cells, face-ordered structure-of-arrays.

The reference keeps per-cell and per-face AoS structs (mesh_structs.py:5-64,
~350 B per hex cell, ~80 B per face) and every kernel loads whole structs.
Here the same information becomes flat arrays in the order the HIP kernels
stream them:

* cells are renumbered by reverse Cuthill-McKee, so the neighbours of a cell
  sit close to it in memory (the gathers of the face and SpMV kernels then
  stay inside a narrow band that the L2 holds);
* internal faces come first, sorted by (new owner, new neighbour), then the
  boundary faces sorted by owner: consecutive lanes of a face kernel read
  consecutive owners;
* the owner / neighbour role of every face is the reference's (lower
  *reference* id owns the face, mesh.py:280), kept after renumbering, because
  the interpolation weight psi (mesh_structs.py:206) and the upwind tie rule
  (convectionScheme.py:42-47) depend on which side owns the face;
* cell->face "slots" (k = 0..5, the reference's local face order) are stored
  slot-major (``slot_face[k*C + c]``) for coalesced per-cell loops;
* tetrahedra (4 faces) and wedges (5) fill the remaining slots with one
  *null face*: an internal face placed after the real internal faces, with
  area 0, flux 0, unit distance and normal, and no neighbour (slot_pos -1,
  slot_nbr = the cell).  Every slot term of the kernels is a multiple of a
  face's area or flux, so the null face adds exactly 0 and the kernels keep
  a fixed six-slot loop for every cell type;
* the C x C CSR pattern (columns sorted) is shared by the pressure matrix and
  the three momentum components; ``slot_pos`` / ``diag_pos`` give the CSR
  index each assembly thread writes, so assembly never searches
  (the reference searches, equation_kernels.py:43-49, 77).
"""
import numpy as np

NSLOT = 6


def rcm_order(cell_neighbors):
    """Reverse Cuthill-McKee order of the cell adjacency graph (new -> old)."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    C = cell_neighbors.shape[0]
    r, k = np.nonzero(cell_neighbors >= 0)
    c = cell_neighbors[r, k]
    g = sp.csr_matrix((np.ones(r.size, np.int8), (r, c)), shape=(C, C))
    return np.asarray(reverse_cuthill_mckee(g, symmetric_mode=True), dtype=np.int64)


class DeviceLayout:
    """Host (numpy) image of every array the device path reads.

    ``new2old`` / ``old2new`` map cells, ``face_new2old`` / ``face_old2new``
    map faces between reference order and device order.  ``bnd_ref_row``
    gives, for each device boundary face, the row of the reference's
    boundary arrays (ordered like ``boundary_face_ids``).
    """

    def __init__(self, mesh, renumber=True):
        fp, cp = mesh.face_properties, mesh.cell_properties
        C, K = cp.num_cells, fp.num_faces
        Fc = cp.faces_per_cell
        if Fc > NSLOT:
            raise ValueError(f'cells with {Fc} faces: at most {NSLOT} are supported')
        null = 1 if Fc < NSLOT else 0
        self.faces_per_cell = Fc
        # device face count (with the null face) and the reference's
        self.num_cells, self.num_faces, self.num_ref_faces = C, K + null, K
        new2old = rcm_order(cp.neighbors) if renumber else np.arange(C, dtype=np.int64)
        old2new = np.empty(C, np.int64)
        old2new[new2old] = np.arange(C)
        self.new2old, self.old2new = new2old, old2new

        own_ref = fp.adjacent_cells[:, 0].astype(np.int64)
        nb_ref = fp.adjacent_cells[:, 1].astype(np.int64)
        k0 = fp.cell_face_index[:, 0].astype(np.int64)
        k1 = fp.cell_face_index[:, 1].astype(np.int64)
        internal = fp.is_boundary == 0
        own_new = old2new[own_ref]
        nb_new = np.where(internal, old2new[np.maximum(nb_ref, 0)], -1)
        fi_ref = np.nonzero(internal)[0]
        fb_ref = np.nonzero(~internal)[0]
        fi_sorted = fi_ref[np.lexsort((nb_new[fi_ref], own_new[fi_ref]))]
        fb_sorted = fb_ref[np.lexsort((k0[fb_ref], own_new[fb_ref]))]
        # the null face (reference id K, no reference counterpart) follows
        # the internal faces
        face_new2old = np.concatenate([fi_sorted, np.full(null, K, np.int64), fb_sorted])
        face_old2new = np.empty(K, np.int64)
        real = face_new2old < K
        face_old2new[face_new2old[real]] = np.nonzero(real)[0]
        self.face_new2old, self.face_old2new = face_new2old, face_old2new
        self.null_face = fi_sorted.size if null else -1
        Fi = fi_sorted.size + null
        Fb = fb_sorted.size
        self.n_internal, self.n_boundary = Fi, Fb
        # device boundary face j (0..Fb-1) <- reference boundary row
        bref_row = np.empty(K, np.int64)
        bref_row[fp.boundary_face_ids] = np.arange(fp.boundary_face_ids.size)
        self.bnd_ref_row = bref_row[fb_sorted]

        f32 = np.float32
        area, psi = mesh.face_struct_arrays()
        c2f = cp.cell_centroid_to_face_centroid
        cen = cp.centroids
        fo = face_new2old[real]
        nf = len(fo)
        ins = lambda a, v: np.insert(a, Fi - null, v, axis=-1) if null else a   # null face row
        app = lambda a, v: np.concatenate([a, np.asarray(v, a.dtype).reshape(a.shape[:-1] + (1,))],
                                          axis=-1) if null else a
        self.owner = ins(own_new[fo], C - 1).astype(np.int32)
        self.neigh = app(nb_new[fi_sorted], C - 1).astype(np.int32)
        self.area = ins(area[fo], 0.0).astype(f32)
        n_own = cp.normal[own_ref[fo], k0[fo]]
        unit = np.array([[1.0], [0.0], [0.0]], f32)
        self.normal = ins(np.ascontiguousarray(n_own.T), unit[:, 0]).astype(f32)  # (3,F)
        psi_i = psi[fi_sorted]
        self.psi = app(psi_i.astype(f32), 0.5)
        oi, ni = own_ref[fi_sorted], nb_ref[fi_sorted]
        rk_dash = cen[oi] * psi_i[:, None] + cen[ni] * (f32(1.0) - psi_i)[:, None]
        self.skew = app(np.ascontiguousarray((fp.centroid[fi_sorted] - rk_dash).T).astype(f32),
                        np.zeros((3, 1), f32))                             # (3,Fi)
        d = fp.distance[fi_sorted]
        dmag = np.sqrt((d * d).sum(axis=1, dtype=f32))
        self.dist = app(np.ascontiguousarray(d.T).astype(f32), unit)      # (3,Fi) O->N
        cb = c2f[own_ref[fb_sorted], k0[fb_sorted]]
        bdist = np.sqrt((cb * cb).sum(axis=1, dtype=f32))
        self.bdist = bdist.astype(f32)                                       # |x_f - x_P|
        # |d| for every face: internal |x_N - x_O|, boundary |x_f - x_P|
        self.dmag = np.concatenate([app(dmag.astype(f32), 1.0), bdist]).astype(f32)
        self.volume = cp.volumes[new2old].astype(f32)
        self.centroid = np.ascontiguousarray(cen[new2old].T).astype(f32)     # (3,C)
        assert self.owner.size == nf + null == self.num_faces

        # slot tables (slot-major), padded to NSLOT with the null face
        pad = NSLOT - Fc
        padc = lambda a, v: np.concatenate([a, np.full(a.shape[:1] + (pad,) + a.shape[2:], v,
                                                       a.dtype)], axis=1) if pad else a
        cf_new = padc(face_old2new[cp.faces[new2old]], self.null_face)     # (C,6)
        self.slot_face = np.ascontiguousarray(cf_new.T).astype(np.int32)     # (6,C)
        self.slot_side = np.ascontiguousarray(padc(cp.face_side[new2old], 0).T).astype(np.int8)
        sn = cp.normal[new2old] * area[cp.faces[new2old]][:, :, None]       # A_f n_{c,k}
        sn = padc(sn, 0.0)
        self.slot_sn = np.ascontiguousarray(np.transpose(sn, (2, 1, 0))).astype(f32)  # (3,6,C)
        nbrs = padc(cp.neighbors[new2old], -1)
        nb_slot_new = np.where(nbrs >= 0, old2new[np.maximum(nbrs, 0)], -1)    # (C,6)

        # CSR over new ids, columns sorted; positions of slots and diagonals
        big = np.iinfo(np.int64).max
        cols7 = np.concatenate([np.arange(C)[:, None], np.where(nb_slot_new >= 0, nb_slot_new, big)], axis=1)
        srt = np.sort(cols7, axis=1)
        valid = srt != big
        cnt = valid.sum(axis=1)
        row_ptr = np.zeros(C + 1, np.int64)
        row_ptr[1:] = np.cumsum(cnt)
        self.row_ptr = row_ptr.astype(np.int32)
        self.col = srt[valid].astype(np.int32)
        self.nnz = int(row_ptr[-1])
        # rank of each slot's column within the sorted row
        rank = (srt[:, None, :] < np.where(nb_slot_new >= 0, nb_slot_new, -1)[:, :, None]).sum(axis=2)
        slot_pos = np.where(nb_slot_new >= 0, row_ptr[:-1, None] + rank, -1)
        self.slot_pos = np.ascontiguousarray(slot_pos.T).astype(np.int32)    # (6,C)
        self.diag_pos = (row_ptr[:-1] + (srt < np.arange(C)[:, None]).sum(axis=1)).astype(np.int32)

        # neighbour cell per slot (new ids); boundary slots point at the cell
        # itself (their matrix coefficient is 0), so SpMV needs no branch
        own_id = np.arange(C)[:, None]
        nb_self = np.where(nb_slot_new >= 0, nb_slot_new, own_id)
        self.slot_nbr = np.ascontiguousarray(nb_self.T).astype(np.int32)
        delta = nb_self - own_id
        self.band = int(np.abs(delta).max()) if delta.size else 0
        # int16 neighbour offsets when the RCM band allows it (halves index bytes)
        self.slot_dnb = (np.ascontiguousarray(delta.T).astype(np.int16)
                         if self.band <= np.iinfo(np.int16).max else None)
        # owner-side local index of each face and the neighbour-side one (a
        # padded slot for the null face)
        self.face_k0 = ins(k0[fo], NSLOT - 1).astype(np.int8)
        self.face_k1 = app(k1[fi_sorted], NSLOT - 1).astype(np.int8)

    ARENA_FIELDS = ('volume', 'slot_face', 'slot_side', 'slot_sn', 'slot_nbr', 'slot_dnb',
                    'owner', 'neigh', 'psi', 'skew', 'area', 'normal', 'dist', 'dmag', 'old2new',
                    'num_cells', 'num_faces', 'n_internal', 'n_boundary', 'nnz', 'band',
                    'null_face')

    def arena_view(self):
        """The arrays a case needs inside a batch arena (no CSR pattern, no
        face maps): about half the bytes, for shipping prepared cases
        between processes."""
        import types
        return types.SimpleNamespace(**{k: getattr(self, k) for k in self.ARENA_FIELDS})

    def ell_to_csr(self, ell):
        """Host conversion of ELL values [7][C] (diagonal, 6 slots) to values
        on (row_ptr, col)."""
        ell = np.asarray(ell).reshape(7, self.num_cells)
        out = np.zeros(self.nnz, ell.dtype)
        out[self.diag_pos] = ell[0]
        for k in range(6):
            ok = self.slot_pos[k] >= 0
            out[self.slot_pos[k][ok]] = ell[k + 1][ok]
        return out

    # ---------------------------------------------------------------- maps
    def cells_to_device(self, a):
        """Reference cell order -> device order (axis 0)."""
        return np.asarray(a)[self.new2old]

    def cells_to_reference(self, a):
        return np.asarray(a)[self.old2new]

    def faces_to_device(self, a):
        """Reference face order -> device order (axis 0); the null face gets 0."""
        a = np.asarray(a)
        if self.null_face >= 0:
            a = np.concatenate([a, np.zeros((1,) + a.shape[1:], a.dtype)], axis=0)
        return a[self.face_new2old]

    def faces_to_reference(self, a):
        return np.asarray(a)[self.face_old2new]
