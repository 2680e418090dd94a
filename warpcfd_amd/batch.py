"""Many independent steady cases on one GPU: a device arena and batched SIMPLE.

No counterpart in the reference, which solves one case per ``FVM`` object
(incompressible.py:95-156 is the per-case step repeated here for every case).

A ``CaseArena`` holds, at a fixed capacity, the mesh arrays, fields, matrices
and boundary conditions of many cases side by side in one ``wcfd_mesh``:
every case takes a contiguous range of cells, of internal faces and of
boundary faces (first fit), with its cell and face indices offset into the
arena.  No face couples two cases, so the arena's matrices are block
diagonal.  One ``wcfd_batch_simple_iteration`` call advances every active
case by one SIMPLE iteration: face and cell kernels run over the whole arena
and skip cells and faces whose case is inactive or absent; the Krylov and
continuity reductions give each case its own workgroups, partials, scalars
and stopping test.  A case's arithmetic therefore does not depend on which
other cases share the arena or where it sits in it (its results are bitwise
the same alone or in a mix).  The device clears a case's active flag when
its continuity residual reaches the case's stop value; the host only polls
the flags.

Cases are admitted and released while others run (``admit`` / ``release``),
which is what the serving loop in ``warpcfd_amd.serve`` builds on.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib
from .FV.device import ptr, require_gpu, stream_ptr
from .preprocess.layout import DeviceLayout


class CaseSpec:
    """One steady case prepared on the host for an arena.

    ``model`` is a ``warpcfd_amd.FV.FVM`` whose boundary conditions, material
    and reference pressure are set (it need not be finalized; no device
    memory of its own is used).  The case is converged when its continuity
    residual is <= ``stop`` and the convergence monitor's estimate of the
    velocity error is <= ``stop_u`` (or the per-iteration velocity change
    sits at the float32 floor, <= ``floor_u``); ``max_outer`` caps its SIMPLE
    iterations.
    """

    def __init__(self, model, stop=1e-6, max_outer=20000, name=None, layout=None, stop_u=5e-5,
                 floor_u=1e-7):
        b = model.boundary
        b.finalize()
        if not b.uniform_velocity_types():
            raise NotImplementedError('u, v and w must share the boundary-condition type on every '
                                      'face')
        L = layout if layout is not None else DeviceLayout(model.mesh, renumber=model._renumber)
        if L.slot_dnb is None:
            raise ValueError('the case band does not fit int16 neighbour offsets')
        self.bc_type, self.bc_value = b.host_arrays(L, np.float64)   # cast on upload
        self.layout = L.arena_view()
        rho = float(model.material.density)
        self.density = rho
        self.nu = float(model.material.viscosity) / rho
        rc = model.reference_pressure_cell_id
        self.ref_cell = -1 if rc is None else int(L.old2new[rc])
        self.ref_value = float(model.reference_pressure)
        self.stop = float(stop)
        self.stop_u = float(stop_u)
        self.floor_u = float(floor_u)
        self.max_outer = int(max_outer)
        self.name = name
        self.n_cells = L.num_cells
        self.n_internal = L.n_internal
        self.n_boundary = L.n_boundary

    def device_bytes(self):
        """Arena bytes the case occupies (mesh, fields, matrices, BCs)."""
        C, Fi, Fb = self.n_cells, self.n_internal, self.n_boundary
        return CaseArena.bytes_for(C, Fi, Fb)


ITEMS_PER_BLOCK = 256   # items per workgroup of the face and cell kernels


def item_blocks(ranges, per=ITEMS_PER_BLOCK):
    """Workgroup item ranges of a compacted launch: every [a, b) of
    ``ranges`` cut into pieces of at most ``per`` items, as a (2, G) int32
    array of starts and ends."""
    ranges = [(int(a), int(b)) for a, b in ranges if b > a]
    if not ranges:
        return np.zeros((2, 0), np.int32)
    a = np.array([r[0] for r in ranges], np.int64)
    b = np.array([r[1] for r in ranges], np.int64)
    n = (b - a + per - 1) // per
    first = np.repeat(np.cumsum(n) - n, n)
    j = np.arange(int(n.sum()), dtype=np.int64) - first
    st = np.repeat(a, n) + per * j
    en = np.minimum(st + per, np.repeat(b, n))
    return np.stack([st, en]).astype(np.int32)


class _Ranges:
    """First-fit allocator of contiguous index ranges in [0, cap)."""

    def __init__(self, cap):
        self.cap = int(cap)
        self.free = [(0, self.cap)]          # sorted, disjoint, coalesced

    def alloc(self, n):
        if n == 0:
            return 0
        for i, (a, m) in enumerate(self.free):
            if m >= n:
                if m == n:
                    self.free.pop(i)
                else:
                    self.free[i] = (a + n, m - n)
                return a
        return None

    def release(self, a, n):
        if n == 0:
            return
        self.free.append((a, n))
        self.free.sort()
        out = []
        for s, m in self.free:
            if out and out[-1][0] + out[-1][1] == s:
                out[-1] = (out[-1][0], out[-1][1] + m)
            else:
                out.append((s, m))
        self.free = out

    def largest(self):
        return max((m for _, m in self.free), default=0)


class CaseArena:
    """Fixed-capacity device arena of up to ``max_cases`` cases.

    Capacities are in cells, internal faces and boundary faces.  ``chunk``
    rows per workgroup of ``block`` threads define every case's reduction
    workgroups (keep them fixed to compare results across arenas bitwise).
    Solver settings are shared by all cases: relaxation factors, Krylov
    tolerances, density and scheme flags.
    """

    # bytes per cell / internal face / boundary face held by the arena
    CELL_BYTES = (4 + 6 * 4 + 6 + 3 * 6 * 4 + 6 * 4 + 6 * 2 + 4 + 4 * 4 + 12 * 4 + 3 * 4 + 4 + 4 +
                  7 * 4 + 3 * 4 + 7 * 4 + 4)
    FACE_BYTES = 4 + 4 + 3 * 4 + 4 + 4 + 4 * 4 + 4 + 4
    INTERNAL_BYTES = 4 + 4 + 3 * 4 + 3 * 4
    BOUNDARY_BYTES = 4 + 4 * 4

    @classmethod
    def bytes_for(cls, C, Fi, Fb):
        # fields and matrices + mesh + the Krylov/workspace vectors (27 floats per cell)
        return (C * (cls.CELL_BYTES + 27 * 4) + (Fi + Fb) * (cls.FACE_BYTES + 4) +
                Fi * cls.INTERNAL_BYTES + Fb * cls.BOUNDARY_BYTES)

    def __init__(self, cell_cap, internal_cap, boundary_cap, max_cases, device='cuda',
                 chunk=256, block=256, density=1.0, alpha_u=0.7, alpha_p=0.3, atol=0.0,
                 tol=1e-6, rtol0=0.0, maxiter=2000, check_every=10, check_every_u=0, p_refine=0,
                 maxiter_u=0,
                 cg_variant=_lib.CG_SINGLE_REDUCTION, implicit_outflow=True, correction=False,
                 dtype=torch.float32):
        """``dtype``: field type, torch.float32 (libwcfd.so) or torch.float64
        (libwcfd64.so, every field, matrix and geometry array in double)."""
        self.device = require_gpu(device)
        self.dtype = dtype
        self.lib = _lib.for_dtype(dtype)
        Cc, Fic, Fbc, K = int(cell_cap), int(internal_cap), int(boundary_cap), int(max_cases)
        if min(Cc, Fic, Fbc, K) <= 0:
            raise ValueError('capacities must be positive')
        if block % 64 or not 64 <= block <= 512 or chunk <= 0:
            raise ValueError('block must be a multiple of 64 in [64, 512], chunk > 0')
        self.Cc, self.Fic, self.Fbc, self.K = Cc, Fic, Fbc, K
        self.F = Fic + Fbc
        self.chunk, self.block = int(chunk), int(block)
        self.Gmax = (Cc + chunk - 1) // chunk + K
        per = ITEMS_PER_BLOCK
        self.item_cap = {'cells': (Cc + per - 1) // per + K,
                         'faces': (Fic + per - 1) // per + (Fbc + per - 1) // per + 2 * K,
                         'bfaces': (Fbc + per - 1) // per + K}
        dev = self.device
        f32, i32 = dtype, torch.int32
        z = lambda *s, dt=f32: torch.zeros(*s, dtype=dt, device=dev)
        F = self.F
        # mesh
        self.volume = z(Cc)
        self.slot_face = z(6, Cc, dt=i32)
        self.slot_side = z(6, Cc, dt=torch.int8)
        self.slot_sn = z(3, 6, Cc)
        self.slot_nbr = z(6, Cc, dt=i32)
        self.slot_dnb = z(6, Cc, dt=torch.int16)
        self.owner = z(F, dt=i32)
        self.area = z(F)
        self.normal = z(3, F)
        self.dmag = torch.ones(F, dtype=f32, device=dev)
        self.neigh = z(Fic, dt=i32)
        self.psi = z(Fic)
        self.skew = z(3, Fic)
        self.dist = z(3, Fic)
        self.bc_type = torch.full((4, Fbc), 2, dtype=torch.uint8, device=dev)
        self.bc_value = z(4, Fbc)
        self.cell_case = torch.full((Cc,), -1, dtype=i32, device=dev)
        self.face_case = torch.full((F,), -1, dtype=i32, device=dev)
        # state
        self.cv = z(4, Cc)
        self.cg = z(12, Cc)
        self.fv = z(4, F)
        self.flux = z(F)
        self.hbya = z(3, Cc)
        self.p_x = z(Cc)
        self.rua = z(Cc)
        self.rua_face = z(F)
        self.mom_values = z(7, Cc)
        self.mom_rhs = z(3, Cc)
        self.p_values = z(7, Cc)
        self.p_rhs = z(Cc)
        self.resid = torch.zeros(K, 4, dtype=torch.float64, device=dev)
        # per case
        self.cell_off = z(K + 1, dt=i32)
        self.cell_end = z(K, dt=i32)
        self.blk_off = z(K + 1, dt=i32)
        self.blk_case = z(self.Gmax, dt=i32)
        self.active = z(K, dt=torch.uint8)
        self.nu = z(K)
        self.ref_cell = torch.full((K,), -1, dtype=i32, device=dev)
        self.ref_value = z(K)
        self.stop = torch.zeros(K, 3, dtype=torch.float64, device=dev)
        self.monitor = torch.zeros(K, _lib.MONITOR_LEN, dtype=torch.float64, device=dev)
        self.outer_iters = z(K, dt=i32)
        # compacted face / cell launch ranges ([2][n] starts, ends)
        self.items = {name: z(2 * cap, dt=i32) for name, cap in self.item_cap.items()}
        self.n_items = {name: 0 for name in self.item_cap}
        self.workspace = torch.zeros(
            self.lib.wcfd_batch_simple_workspace_bytes(Cc, F, K, self.Gmax), dtype=torch.uint8,
            device=dev)
        # host bookkeeping
        self.cells = _Ranges(Cc)
        self.ifaces = _Ranges(Fic)
        self.bfaces = _Ranges(Fbc)
        self.slots = [None] * K           # slot -> dict(spec, c0, i0, b0)
        self.h_cell_off = np.zeros(K + 1, np.int32)
        self.h_cell_end = np.zeros(K, np.int32)
        self.h_active = np.zeros(K, np.uint8)
        self.h_steps = np.zeros(K, np.int64)   # iterations issued while the host saw the case active
        self.n_blocks = 0
        self.max_case_blocks = 0
        # settings
        self.density = float(density)
        self.alpha_u, self.alpha_p = alpha_u, alpha_p
        self.atol, self.tol, self.rtol0 = atol, tol, rtol0
        self.maxiter, self.check_every, self.check_every_u = maxiter, check_every, check_every_u
        self.p_refine = p_refine
        self.maxiter_u = maxiter_u
        self.cg_variant = cg_variant
        self.scheme_flags = (_lib.SCHEME_IMPLICIT_OUTFLOW if implicit_outflow else 0) | \
            (_lib.SCHEME_CORRECTED if correction else 0)
        self._build_descriptors()

    # ----------------------------------------------------------- capacity
    @classmethod
    def for_cases(cls, specs, max_cases=None, **kw):
        """An arena exactly large enough for ``specs`` (all admitted)."""
        C = sum(s.n_cells for s in specs)
        Fi = sum(s.n_internal for s in specs)
        Fb = sum(s.n_boundary for s in specs)
        kw.setdefault('density', specs[0].density)
        a = cls(max(C, 1), max(Fi, 1), max(Fb, 1), max_cases or len(specs), **kw)
        for s in specs:
            if a.admit(s) is None:
                raise RuntimeError('arena sizing error')
        return a

    @classmethod
    def for_budget(cls, hbm_bytes, max_cases, cells_per_face=(3.0, 2.1), **kw):
        """An arena whose capacity fits ``hbm_bytes`` of device memory, with
        internal / boundary face capacities in the given ratio to cells
        (3-D hex meshes have about 3 internal faces per cell; one-cell-thick
        2-D meshes about 2 boundary faces per cell)."""
        ri, rb = cells_per_face
        per_cell = cls.bytes_for(1, ri, rb)
        C = int(hbm_bytes // per_cell)
        return cls(C, int(C * ri) + 1, int(C * rb) + 1, max_cases, **kw)

    def fits(self, spec):
        return (spec.n_cells <= self.cells.largest() and spec.n_internal <= self.ifaces.largest()
                and spec.n_boundary <= self.bfaces.largest() and None in self.slots)

    # --------------------------------------------------------- admission
    def admit(self, spec):
        """Place ``spec`` in the arena with zero initial fields; returns its
        slot, or None if no slot or no contiguous space is free."""
        if abs(spec.density - self.density) > 0:
            raise ValueError('every case of an arena shares the density')
        if None not in self.slots:
            return None
        C, Fi, Fb = spec.n_cells, spec.n_internal, spec.n_boundary
        c0 = self.cells.alloc(C)
        if c0 is None:
            return None
        i0 = self.ifaces.alloc(Fi)
        if i0 is None:
            self.cells.release(c0, C)
            return None
        b0 = self.bfaces.alloc(Fb)
        if b0 is None:
            self.cells.release(c0, C)
            self.ifaces.release(i0, Fi)
            return None
        k = self.slots.index(None)
        self.slots[k] = {'spec': spec, 'c0': c0, 'i0': i0, 'b0': b0}
        self._upload(k)
        self.h_cell_off[k] = c0
        self.h_cell_end[k] = c0 + C
        self.h_active[k] = 1
        self.h_steps[k] = 0
        self._build_descriptors()
        return k

    def release(self, k):
        """Free slot ``k`` (its cells and faces become free space)."""
        sl = self.slots[k]
        if sl is None:
            return
        s = sl['spec']
        c0, i0, b0 = sl['c0'], sl['i0'], sl['b0']
        self.cell_case[c0:c0 + s.n_cells] = -1
        self.face_case[i0:i0 + s.n_internal] = -1
        self.face_case[self.Fic + b0:self.Fic + b0 + s.n_boundary] = -1
        self.active[k] = 0
        self.cells.release(c0, s.n_cells)
        self.ifaces.release(i0, s.n_internal)
        self.bfaces.release(b0, s.n_boundary)
        self.slots[k] = None
        self.h_active[k] = 0
        self.h_cell_off[k] = 0
        self.h_cell_end[k] = 0
        self._build_descriptors()

    def _upload(self, k):
        sl = self.slots[k]
        s, c0, i0, b0 = sl['spec'], sl['c0'], sl['i0'], sl['b0']
        L = s.layout
        C, Fi, Fb = s.n_cells, s.n_internal, s.n_boundary
        dev = self.device
        t = lambda a, dt=None: torch.from_numpy(np.ascontiguousarray(a if dt is None else
                                                                     np.asarray(a).astype(dt))).to(dev)
        cs = slice(c0, c0 + C)
        fi = slice(i0, i0 + Fi)
        gb = slice(self.Fic + b0, self.Fic + b0 + Fb)    # boundary faces as arena faces
        lb = slice(b0, b0 + Fb)                          # columns of the BC arrays
        # local face f -> arena face: internal f -> i0 + f, boundary Fi + j -> Fic + b0 + j
        sf = L.slot_face.astype(np.int64)
        sf = np.where(sf < Fi, sf + i0, sf - Fi + self.Fic + b0).astype(np.int32)
        self.volume[cs] = t(L.volume)
        self.slot_face[:, cs] = t(sf)
        self.slot_side[:, cs] = t(L.slot_side)
        self.slot_sn[:, :, cs] = t(L.slot_sn)
        self.slot_nbr[:, cs] = t(L.slot_nbr.astype(np.int64) + c0, np.int32)
        self.slot_dnb[:, cs] = t(L.slot_dnb)
        own = L.owner.astype(np.int64) + c0
        self.owner[fi] = t(own[:Fi], np.int32)
        self.owner[gb] = t(own[Fi:], np.int32)
        for dst, src in ((self.area, L.area), (self.dmag, L.dmag)):
            dst[fi] = t(src[:Fi])
            dst[gb] = t(src[Fi:])
        self.normal[:, fi] = t(L.normal[:, :Fi])
        self.normal[:, gb] = t(L.normal[:, Fi:])
        if Fi:
            self.neigh[fi] = t(L.neigh.astype(np.int64) + c0, np.int32)
            self.psi[fi] = t(L.psi)
            self.skew[:, fi] = t(L.skew)
            self.dist[:, fi] = t(L.dist)
        if Fb:
            self.bc_type[:, lb] = t(s.bc_type)
            self.bc_value[:, lb] = t(s.bc_value)
        # zero state of the case's rows
        for a in (self.cv, self.cg, self.hbya, self.mom_values, self.mom_rhs, self.p_values):
            a[:, cs] = 0
        for a in (self.p_x, self.rua, self.p_rhs):
            a[cs] = 0
        for a in (self.fv,):
            a[:, fi] = 0
            a[:, gb] = 0
        for a in (self.flux, self.rua_face):
            a[fi] = 0
            a[gb] = 0
        self.cell_case[cs] = k
        self.face_case[fi] = k
        self.face_case[gb] = k
        self.nu[k] = s.nu
        self.ref_cell[k] = (c0 + s.ref_cell) if s.ref_cell >= 0 else -1
        self.ref_value[k] = s.ref_value
        self.stop[k] = torch.tensor([s.stop, s.stop_u, s.floor_u], dtype=torch.float64)
        self.monitor[k] = 0
        self.outer_iters[k] = 0
        self.resid[k] = 0
        self.active[k] = 1

    def _build_descriptors(self):
        """Workgroup ranges of the resident, active cases and the C-ABI
        structs.  Cases that have finished get no workgroups, so the
        row-parallel kernels launch only the workgroups of cases still
        iterating; a case's own partition (and so its arithmetic) does not
        change when others drop out."""
        K, ch = self.K, self.chunk
        nb = np.zeros(K, np.int64)
        for k, sl in enumerate(self.slots):
            if sl is not None and self.h_active[k]:
                nb[k] = (sl['spec'].n_cells + ch - 1) // ch
        blk_off = np.zeros(K + 1, np.int32)
        blk_off[1:] = np.cumsum(nb)
        G = int(blk_off[-1])
        blk_case = np.repeat(np.arange(K, dtype=np.int32), nb)
        self.n_blocks = G
        self.max_case_blocks = int(nb.max()) if K else 0
        cell_off = self.h_cell_off.copy()
        cell_off[K] = self.Cc
        dev = self.device
        self.cell_off.copy_(torch.from_numpy(cell_off).to(dev))
        self.cell_end.copy_(torch.from_numpy(self.h_cell_end).to(dev))
        self.blk_off.copy_(torch.from_numpy(blk_off).to(dev))
        if G:
            self.blk_case[:G] = torch.from_numpy(blk_case).to(dev)
        # face and cell kernels: the active cases' cells, internal faces
        # (numbered first) and boundary faces (after all internal faces)
        live = [sl for k, sl in enumerate(self.slots) if sl is not None and self.h_active[k]]
        cr = [(sl['c0'], sl['c0'] + sl['spec'].n_cells) for sl in live]
        ir = [(sl['i0'], sl['i0'] + sl['spec'].n_internal) for sl in live]
        br = [(sl['b0'], sl['b0'] + sl['spec'].n_boundary) for sl in live]
        lists = {'cells': item_blocks(cr), 'bfaces': item_blocks(br),
                 'faces': item_blocks(ir + [(self.Fic + a, self.Fic + b) for a, b in br])}
        for name, blk in lists.items():
            n = blk.shape[1]
            assert n <= self.item_cap[name]
            self.n_items[name] = n
            if n:
                both = np.zeros(2 * n, np.int32)
                both[:n], both[n:] = blk[0], blk[1]
                self.items[name][:2 * n] = torch.from_numpy(both).to(dev)
        m = _lib.WcfdMesh()
        m.n_cells, m.n_faces, m.n_internal, m.n_boundary = self.Cc, self.F, self.Fic, self.Fbc
        m.nnz, m.n_slots = 0, 0           # 0: six slots, right for every cell type
        for name in ('owner', 'neigh', 'psi', 'skew', 'area', 'normal', 'dist', 'dmag', 'volume',
                     'slot_face', 'slot_side', 'slot_sn', 'slot_nbr', 'slot_dnb'):
            setattr(m, name, getattr(self, name).data_ptr())
        m.slot_pos = m.diag_pos = m.row_ptr = m.col = None
        self.desc = m
        b = _lib.WcfdBatch()
        b.n_cases, b.n_blocks, b.chunk, b.block = K, G, self.chunk, self.block
        b.max_case_blocks = self.max_case_blocks
        b.n_cell_blk, b.n_face_blk, b.n_bface_blk = (self.n_items[k] for k in
                                                     ('cells', 'faces', 'bfaces'))
        b.cell_blk, b.face_blk, b.bface_blk = (self.items[k].data_ptr() for k in
                                               ('cells', 'faces', 'bfaces'))
        for name in ('cell_off', 'cell_end', 'blk_off', 'blk_case', 'cell_case', 'face_case',
                     'active', 'nu', 'ref_cell', 'ref_value', 'stop', 'monitor', 'outer_iters'):
            setattr(b, name, getattr(self, name).data_ptr())
        self.batch = b
        self.bc = _lib.WcfdBC(self.bc_type.data_ptr(), self.bc_value.data_ptr())
        st = _lib.SimpleState()
        for name, tt in (('cell_values', self.cv), ('cell_grad', self.cg),
                         ('face_values', self.fv), ('mass_flux', self.flux),
                         ('hbya', self.hbya), ('p_x', self.p_x), ('rua', self.rua),
                         ('rua_face', self.rua_face), ('mom_values', self.mom_values),
                         ('mom_rhs', self.mom_rhs), ('p_values', self.p_values),
                         ('p_rhs', self.p_rhs), ('resid', self.resid)):
            setattr(st, name, tt.data_ptr())
        self.state = st

    # ----------------------------------------------------------- stepping
    def params(self):
        p = _lib.SimpleParams()
        p.nu = 0.0
        p.density = self.density
        p.alpha_u = float(self.alpha_u)
        p.alpha_p = float(self.alpha_p)
        p.dt = 0.0
        p.scheme_flags = self.scheme_flags
        p.ref_cell = -1
        p.ref_value = 0.0
        p.n_inner = p.n_ortho = 1
        p.atol, p.tol, p.rtol0 = float(self.atol), float(self.tol), float(self.rtol0)
        p.maxiter, p.check_every = int(self.maxiter), int(self.check_every)
        p.check_every_u = int(self.check_every_u)
        p.p_refine = int(self.p_refine)
        p.maxiter_u = int(self.maxiter_u)
        p.first_iteration = 0
        p.cg_variant = int(self.cg_variant)
        return p

    def launch_key(self):
        """Grid sizes of every launch of a batched iteration; a captured graph
        stays valid while they are unchanged (the ranges themselves are read
        from device memory at run time)."""
        return (self.n_blocks, self.max_case_blocks, self.n_items['cells'],
                self.n_items['faces'], self.n_items['bfaces'])

    def step(self, info=False, stream=None, budget=None):
        """One SIMPLE iteration of every active case.  With ``info`` returns
        per-slot (momentum iterations, pressure iterations, continuity).
        ``budget`` = (groups_u, groups_p) runs the Krylov solves with that
        fixed launch budget instead of polling the host (the mode a captured
        graph replays)."""
        if self.n_blocks == 0:
            return None
        prm = self.params()
        if budget is not None:
            prm.graph_groups_u, prm.graph_groups_p = (int(b) for b in budget)
            info = False
        arr = (_lib.SimpleInfo * self.K)() if info else None
        code = self.lib.wcfd_batch_simple_iteration(
            ctypes.byref(self.desc), ctypes.byref(self.batch), self.bc, ctypes.byref(prm),
            ctypes.byref(self.state), ptr(self.workspace), arr, stream_ptr(stream))
        _lib.check(code, 'batch simple iteration')
        self.h_steps += self.h_active.astype(np.int64)
        if info:
            return [(arr[k].momentum.iterations, arr[k].pressure.iterations, arr[k].continuity)
                    for k in range(self.K)]
        return None

    # ---------------------------------------------------------- HIP graph
    def capture(self, budget):
        """Capture one batched SIMPLE iteration with the fixed Krylov launch
        budget ``(groups_u, groups_p)`` as a HIP graph (torch's stream
        capture; the library launches only kernels in this mode, with no
        memset, copy or host read).  Kernels read the batch descriptors
        through pointers, so the graph stays valid while cases finish; it
        must be captured again after an admission or release changes the
        number of workgroups.  Call ``step(budget=...)`` once before the
        first capture, so every kernel is loaded outside the capture."""
        side = torch.cuda.Stream(device=self.device)
        side.wait_stream(torch.cuda.current_stream())
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=side):
            self.step(stream=side, budget=budget)
        torch.cuda.current_stream().wait_stream(side)
        self._graph = {'graph': g, 'blocks': self.launch_key(), 'budget': tuple(budget)}
        return g

    def replay(self):
        """One batched SIMPLE iteration by replaying the captured graph."""
        gr = getattr(self, '_graph', None)
        if gr is None or gr['blocks'] != self.launch_key():
            raise RuntimeError('no graph for the current arena contents: call capture() again')
        gr['graph'].replay()
        self.h_steps += self.h_active.astype(np.int64)

    def budget_for(self, res, margin=1.5):
        """Launch budget (groups_u, groups_p) covering ``margin`` times the
        most Krylov iterations any resident case needed in ``res`` (the
        return value of ``step(info=True)``)."""
        cu = self.check_every_u or self.check_every
        cp = self.check_every
        act = [k for k, sl in enumerate(self.slots) if sl is not None]
        mu = max([1] + [res[k][0] for k in act])
        mp = max([1] + [res[k][1] for k in act])
        return (max(1, math.ceil(margin * mu / cu)), max(1, math.ceil(margin * mp / cp)))

    def poll(self):
        """Synchronise and read the active flags (one small copy); cases past
        their iteration cap are deactivated here.  Returns the slots that
        are resident and no longer active.  The device stops cases by
        itself, so polling less often only delays this bookkeeping."""
        act = self.active.cpu().numpy()
        capped = []
        for k, sl in enumerate(self.slots):
            # an active case has run exactly h_steps[k] iterations
            if sl is not None and act[k] and self.h_steps[k] >= sl['spec'].max_outer:
                capped.append(k)
        for k in capped:
            self.active[k] = 0
            act[k] = 0
        finished = bool(np.any((self.h_active != 0) & (act == 0)))
        self.h_active[:] = act
        if finished:
            self._build_descriptors()      # drop the finished cases' workgroups
        return [k for k, sl in enumerate(self.slots) if sl is not None and not act[k]]

    def _report_batch(self):
        """A copy of the batch descriptor whose workgroup ranges cover every
        resident case, finished ones included (the stepping descriptor drops
        those), so a case can be reported after poll() has stopped it.  A
        case gets the same ceil(C / chunk) workgroups in both descriptors, so
        its report does not depend on which descriptor it came from."""
        K, ch = self.K, self.chunk
        nb = np.zeros(K, np.int64)
        for k, sl in enumerate(self.slots):
            if sl is not None:
                nb[k] = (sl['spec'].n_cells + ch - 1) // ch
        blk_off = np.zeros(K + 1, np.int32)
        blk_off[1:] = np.cumsum(nb)
        G = int(blk_off[-1])
        dev = self.device
        if getattr(self, '_rep_blk_off', None) is None:
            self._rep_blk_off = torch.zeros(K + 1, dtype=torch.int32, device=dev)
            self._rep_blk_case = torch.zeros(self.Gmax, dtype=torch.int32, device=dev)
        self._rep_blk_off.copy_(torch.from_numpy(blk_off).to(dev))
        if G:
            self._rep_blk_case[:G] = torch.from_numpy(
                np.repeat(np.arange(K, dtype=np.int32), nb)).to(dev)
        b = _lib.WcfdBatch.from_buffer_copy(self.batch)
        b.n_blocks = G
        b.max_case_blocks = int(nb.max()) if K else 0
        b.blk_off, b.blk_case = self._rep_blk_off.data_ptr(), self._rep_blk_case.data_ptr()
        return b

    def convergence_report(self, stream=None):
        """Per-case norms of check_convergence (model.py:278-315) from one
        device pass (``wcfd_batch_convergence_report``, float64): for each
        slot None (empty) or a dict as ``FVM.convergence_norms`` returns --
        ``r``, ``b`` (||A u_k - b_k||_2, ||b_k||_2), ``vel_corr`` (max |rUA
        grad p|), ``p`` (max |p|, the pressure solve's result), ``div`` (sum
        |div F| of the case's mass fluxes).  Every resident case is reported,
        including one that has finished but is not yet released (its final
        state).  A case's report depends only on the case."""
        bt = self._report_batch()
        if bt.n_blocks == 0:
            return [None] * self.K
        need = self.lib.wcfd_batch_convergence_workspace_bytes(bt.n_blocks)
        if getattr(self, '_conv_ws', None) is None or self._conv_ws.numel() < need:
            self._conv_ws = torch.zeros(need, dtype=torch.uint8, device=self.device)
            self._conv_out = torch.zeros(9 * self.K, dtype=torch.float64, device=self.device)
        if stream is not None:
            stream.wait_stream(torch.cuda.current_stream())   # buffers written there
        code = self.lib.wcfd_batch_convergence_report(
            ctypes.byref(self.desc), ctypes.byref(bt), ptr(self.mom_values), ptr(self.cv),
            ptr(self.mom_rhs), ptr(self.rua), ptr(self.cg[9:12]), ptr(self.p_x), ptr(self.flux),
            ptr(self._conv_ws), ptr(self._conv_out), stream_ptr(stream))
        _lib.check(code, 'batch convergence report')
        if stream is not None:
            # the copy below runs on the current stream: order it after the report
            torch.cuda.current_stream().wait_stream(stream)
        o = self._conv_out.cpu().numpy().reshape(self.K, 9)
        rep = []
        for k, sl in enumerate(self.slots):
            if sl is None:
                rep.append(None)
                continue
            rep.append({'r': [float(x) for x in o[k, 0:3]], 'b': [float(x) for x in o[k, 3:6]],
                        'vel_corr': float(o[k, 6]), 'p': float(o[k, 7]), 'div': float(o[k, 8])})
        return rep

    def continuity(self, k):
        r = self.resid[k].cpu().numpy()
        return float(r[0] / r[1]) if r[1] > 0 else 0.0

    def monitor_record(self, k):
        """(continuity, velocity change, mean change, rate, error estimate)
        of slot ``k`` after its last iteration."""
        return tuple(float(v) for v in self.monitor[k, 1:6].cpu().numpy())

    def run(self, max_steps=100000, poll_every=1, budget=None, graph=False, recheck=100,
            margin=1.5):
        """Iterate until no resident case is active; returns the number of
        batch iterations.

        ``budget`` = (groups_u, groups_p): every iteration after the first
        replays a HIP graph captured with that fixed Krylov launch budget.
        ``graph=True``: the budget follows the solves -- every ``recheck``
        iterations one polled iteration measures the Krylov iterations the
        cases need, and the graph is captured again when ``margin`` times
        that no longer fits the budget or needs less than half of it.  A
        solve that needs more than the budget stops at it (like maxiter)."""
        n = 0
        cur = None
        while n < max_steps:
            if budget is not None:
                if n == 0:
                    self.step(budget=budget)
                    self.capture(budget)
                elif self._graph['blocks'] != self.launch_key():
                    self.capture(budget)           # workgroups changed: capture again
                    self.replay()
                else:
                    self.replay()
            elif graph:
                stale = cur is not None and self._graph['blocks'] != self.launch_key()
                if cur is None or n % recheck == 0 or stale:
                    need = self.budget_for(self.step(info=True), margin)
                    if (stale or cur is None or need[0] > cur[0] or need[1] > cur[1]
                            or 2 * need[0] < cur[0] or 2 * need[1] < cur[1]):
                        cur = need
                        self.capture(cur)
                else:
                    self.replay()
            else:
                self.step()
            n += 1
            if n % poll_every == 0:
                self.poll()
                if not any(self.h_active[k] for k, sl in enumerate(self.slots) if sl is not None):
                    break
        return n

    # ------------------------------------------------------------ results
    def cell_values(self, k):
        """(C, 4) numpy array (the arena's field type) of case ``k`` in its
        reference cell order."""
        sl = self.slots[k]
        s, c0 = sl['spec'], sl['c0']
        dev = self.cv[:, c0:c0 + s.n_cells].cpu().numpy()
        return np.ascontiguousarray(dev[:, s.layout.old2new].T)

    def converged(self, k):
        """The monitor's verdict for slot ``k`` (the rule the device applied)."""
        s = self.slots[k]['spec']
        c, _, dbar, _, E = self.monitor_record(k)
        return c <= s.stop and (E <= s.stop_u or dbar <= s.floor_u)

    def iterations(self, k):
        return int(self.outer_iters[k].item())


def run_workers(work, n):
    """Run ``work(j, abort)`` for j in [0, n) on n threads and wait for all of
    them.  The first exception sets ``abort`` (a threading.Event the workers
    check between steps, so the others stop early) and is raised here."""
    import threading
    abort = threading.Event()
    errors = []

    def body(j):
        try:
            work(j, abort)
        except BaseException as e:    # re-raised on the calling thread
            errors.append(e)
            abort.set()

    threads = [threading.Thread(target=body, args=(j,)) for j in range(n)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise errors[0]


def split_groups(sizes, n_groups):
    """Case indices in ``n_groups`` groups of about equal total size: the
    cases in decreasing size (ties by index) dealt out back and forth
    (groups 0, 1, ..., n-1, n-1, ..., 1, 0, 0, 1, ...), so every group gets
    one of the largest cases.  Groups left empty are dropped."""
    n = max(1, int(n_groups))
    order = sorted(range(len(sizes)), key=lambda k: (-sizes[k], k))
    groups = [[] for _ in range(n)]
    for i, k in enumerate(order):
        r, j = divmod(i, n)
        groups[j if r % 2 == 0 else n - 1 - j].append(k)
    return [g for g in groups if g]


class SplitBatch:
    """Steady cases split over several arenas, each stepped by its own host
    thread on its own stream, so that the arenas' kernels run side by side.

    One arena advances its cases in lockstep: each batched SIMPLE iteration
    waits for the slowest Krylov solve among them, and once few cases remain
    every launch costs its fixed overhead rather than HBM time.  Over several
    arenas, a group whose solves end early starts its next iteration while
    the others still iterate, and the launches of different groups overlap
    on the GPU.  A case's arithmetic depends only on its own rows and on the
    workgroup shape (``chunk``, ``block``, the same for every arena), so its
    results are bitwise those of one arena holding every case.

    ``specs``: CaseSpec list; ``n_arenas``: groups (``split_groups`` by cell
    count); ``arena_kw``: CaseArena keywords.  Results are addressed by the
    index of the case in ``specs``.  The library keeps one pinned polling
    buffer and its events per host thread, so each worker thread owns one
    such set.  The arenas overlap only on streams of their own hardware
    queues: with HIP's default of four queues per process, four arenas
    measured slower (two of their streams shared a queue) than with eight
    (``GPU_MAX_HW_QUEUES``, read when HIP starts).
    """

    def __init__(self, specs, n_arenas=4, **arena_kw):
        self.groups = split_groups([s.n_cells for s in specs], n_arenas)
        self.arenas = [CaseArena.for_cases([specs[k] for k in g], **arena_kw)
                       for g in self.groups]
        self.where = {k: (j, i) for j, g in enumerate(self.groups) for i, k in enumerate(g)}
        self.streams = [torch.cuda.Stream(device=a.device) for a in self.arenas]
        self.done_at = {}
        self.steps = [0] * len(self.arenas)

    def run(self, max_steps=100000, poll_every=4):
        """Step every arena on its own thread and stream until none of its
        cases is active (at most ``max_steps`` batched iterations each).
        ``done_at[k]``: seconds from the start to the poll that saw case k
        stop.  Returns the batched iterations of each arena."""
        import time
        torch.cuda.synchronize()          # arenas were filled on the default stream
        t0 = time.perf_counter()

        def work(j, abort):
            a, g = self.arenas[j], self.groups[j]
            with torch.cuda.stream(self.streams[j]):
                n = 0
                try:
                    while n < max_steps and not abort.is_set():
                        a.step()
                        n += 1
                        if n % poll_every == 0 or n == max_steps:
                            now = time.perf_counter() - t0
                            for i in a.poll():
                                self.done_at.setdefault(g[i], now)
                            if not any(a.h_active[i] for i in range(len(g))):
                                break
                finally:
                    self.steps[j] = n
                    self.streams[j].synchronize()

        run_workers(work, len(self.arenas))
        torch.cuda.synchronize()
        return list(self.steps)

    def _at(self, k):
        j, i = self.where[k]
        return self.arenas[j], i

    def cell_values(self, k):
        a, i = self._at(k)
        return a.cell_values(i)

    def converged(self, k):
        a, i = self._at(k)
        return a.converged(i)

    def iterations(self, k):
        a, i = self._at(k)
        return a.iterations(i)

    def continuity(self, k):
        a, i = self._at(k)
        return a.continuity(i)
