"""Oracle restatement of the reference finite-volume operators (test infrastructure).

Every array keeps the reference layout: cell_values (C,4), cell_gradients
(C,4,3), face_values (F,4), face_gradients (F,4), mass_fluxes (F), term
weights (C,Fc,O,3) with Fc faces per cell (6 hexahedron, 5 wedge, 4
tetrahedron).  Loops that the reference runs serially (Warp-CPU launches
in tid order, atomics accumulated in that order) are vectorised over the
outer index and kept sequential over the inner one, so each accumulation
happens in the reference's order.
"""
import numpy as np

DIRICHLET = 1
NEUMANN = 2


class Boundary:
    """FV/boundary/boundary.py:5-121."""

    def __init__(self, boundary_ids, groups, num_outputs=4, dtype=np.float32):
        self.boundary_ids = np.asarray(boundary_ids, np.int32)
        self.groups = groups
        self.values = np.zeros((len(boundary_ids), num_outputs), dtype)
        self.types = np.zeros((len(boundary_ids), num_outputs), np.uint8)
        self.fields = {'u': 0, 'v': 1, 'w': 2, 'p': 3}

    def _rows(self, face_ids):
        if isinstance(face_ids, str):
            face_ids = self.groups[face_ids]
        face_ids = np.atleast_1d(np.asarray(face_ids))
        assert np.all(np.isin(face_ids, self.boundary_ids))
        return np.where(np.isin(self.boundary_ids, face_ids))[0]

    def set_value(self, face_ids, btype, **kw):
        rows = self._rows(face_ids)
        for name, val in kw.items():
            j = self.fields[name]
            self.values[rows, j] = val
            self.types[rows, j] = btype

    def fixed_BC(self, f, **kw): self.set_value(f, DIRICHLET, **kw)
    def gradient_BC(self, f, **kw): self.set_value(f, NEUMANN, **kw)

    def velocity_BC(self, f, u, v, w):                 # boundary.py:77-82
        self.fixed_BC(f, u=u, v=v, w=w); self.gradient_BC(f, p=0.)

    def pressure_BC(self, f, p):                       # boundary.py:84-89
        self.fixed_BC(f, p=p); self.gradient_BC(f, u=0., v=0., w=0.)

    def no_slip_wall(self, f):                         # boundary.py:94-99
        self.fixed_BC(f, u=0., v=0., w=0.); self.gradient_BC(f, p=0.)

    def slip_wall(self, f):                            # boundary.py:101-105 (T12)
        self.gradient_BC(f, u=0., v=0., w=0., p=0.)

    def finalize(self):                                # boundary.py:108-121
        if np.any(self.types < 1):
            raise ValueError('unassigned boundary faces')
        if np.any(self.types > 3):
            raise ValueError('bad boundary type')


class OracleFVM:
    """FV/model.py:51-384 restated over an oracle.mesh.OracleMesh."""

    def __init__(self, mesh, dtype=np.float32):
        self.m = mesh
        self.dtype = np.dtype(dtype)
        dt = self.dtype
        C, F = mesh.num_cells, mesh.num_faces
        self.boundary = Boundary(mesh.boundary_face_ids, mesh.groups, 4, dt)
        self.cell_values = np.zeros((C, 4), dt)
        self.cell_gradients = np.zeros((C, 4, 3), dt)
        self.face_values = np.zeros((F, 4), dt)
        self.face_gradients = np.zeros((F, 4), dt)
        self.mass_fluxes = np.zeros(F, dt)
        self.reference_pressure_cell_id = None
        self.reference_pressure = 0.0
        # geometry in the model dtype (to_NVD_warp, mesh.py:369-404, 471-510)
        g = lambda a: np.asarray(a, dt)
        self.centroids = g(mesh.centroids)
        self.volumes = g(mesh.volumes)
        self.normal = g(mesh.normal)
        self.c2f = g(mesh.c2f)
        self.face_area = g(mesh.face_area)
        self.face_centroid = g(mesh.face_centroid)
        self.psi = g(mesh.norm_distance[:, 0])
        self.psi2 = g(mesh.norm_distance[:, 1])
        self.cell_distance = g(mesh.cell_distance)
        self.material_viscosity = None
        self.material_density = None

    def set_reference_pressure(self, cell_id, value=0.0):   # model.py:151-153
        self.reference_pressure_cell_id = cell_id
        self.reference_pressure = float(value)

    # ---------------------------------------------------------------- model ops
    def set_boundary_conditions(self):                # conditions.py:6-19
        b = self.boundary
        for j in range(4):
            d = b.types[:, j] == DIRICHLET
            nmn = b.types[:, j] == NEUMANN
            self.face_values[b.boundary_ids[d], j] = b.values[d, j]
            self.face_gradients[b.boundary_ids[nmn], j] = b.values[nmn, j]

    def face_interpolation(self, outputs=(0, 1, 2, 3), upwind=False):
        """model.py:210-229; interpolation_kernels.py:5-54 (linear + skew).
        ``upwind``: interpolation_functions.py:18-28, the owner's value where
        the face's mass flux is > 0, else the neighbour's (a zero flux takes
        the neighbour), plus the same skew correction."""
        m = self.m
        fi = m.internal_face_ids
        own = m.adjacent_cells[fi, 0]
        nb = m.adjacent_cells[fi, 1]
        psi = self.psi[fi][:, None]
        psi2 = self.psi2[fi][:, None]
        cv = self.cell_values
        cg = self.cell_gradients
        # skewness_functions.py:7-20
        rk_dash = self.centroids[own] * psi + self.centroids[nb] * psi2
        rk = self.face_centroid[fi] - rk_dash
        up = self.mass_fluxes[fi] > 0
        for o in outputs:
            if upwind:
                val = np.where(up, cv[own, o], cv[nb, o])
            else:
                val = cv[own, o] * psi[:, 0] + psi2[:, 0] * cv[nb, o]
            gi = cg[own, o] * psi + psi2 * cg[nb, o]
            skew = gi[:, 0] * rk[:, 0] + gi[:, 1] * rk[:, 1] + gi[:, 2] * rk[:, 2]
            self.face_values[fi, o] = val + skew
        # boundary, Neumann only (interpolation_kernels.py:33-54)
        b = self.boundary
        bid = b.boundary_ids
        own_b = m.adjacent_cells[bid, 0]
        kb = m.cell_face_index[bid, 0]
        c2f = self.c2f[own_b, kb]
        dist = np.sqrt(c2f[:, 0] * c2f[:, 0] + c2f[:, 1] * c2f[:, 1] + c2f[:, 2] * c2f[:, 2])
        for o in outputs:
            sel = b.types[:, o] == NEUMANN
            self.face_values[bid[sel], o] = dist[sel] * self.face_gradients[bid[sel], o] + cv[own_b[sel], o]

    def calculate_gradients(self, outputs=(0, 1, 2, 3)):
        """model.py:232-239; model_kernels.py:72-90 (zeroes ALL gradients, T3)."""
        self.cell_gradients[:] = 0
        m = self.m
        for o in outputs:
            acc = np.zeros((m.num_cells, 3), self.dtype)
            for k in range(m.faces_per_cell):
                f = m.cell_faces[:, k]
                s = self.face_values[f, o] * self.face_area[f]
                acc = acc + (s[:, None] * self.normal[:, k]) / self.volumes[:, None]
            self.cell_gradients[:, o] = acc

    def calculate_mass_flux(self):
        """model.py:264-266; model_kernels.py:46-60 (owner normal, owner area)."""
        m = self.m
        n = self.normal[m.adjacent_cells[:, 0], m.cell_face_index[:, 0]]
        fv = self.face_values
        dot = fv[:, 0] * n[:, 0] + fv[:, 1] * n[:, 1] + fv[:, 2] * n[:, 2]
        self.mass_fluxes[:] = dot * self.face_area

    def divFlux(self):
        """model.py:269-275; model_kernels.py:27-34."""
        m = self.m
        acc = np.zeros(m.num_cells, self.dtype)
        for k in range(m.faces_per_cell):
            fl = self.mass_fluxes[m.cell_faces[:, k]]
            acc = acc + np.where(m.face_side[:, k] == 0, fl, -fl)
        return acc

    def get_gradient(self, o, coeff=1.0):
        """model.py:242-254; model_kernels.py:8-14."""
        c = np.asarray(coeff, self.dtype)
        if c.ndim == 0:
            return (c * self.cell_gradients[:, o]).astype(self.dtype)
        return (c[:, None] * self.cell_gradients[:, o]).astype(self.dtype)

    def relax(self, new_value, alpha, o):             # model.py:381-384
        a = self.dtype.type(alpha)
        self.cell_values[:, o] = a * new_value + (self.dtype.type(1.0) - a) * self.cell_values[:, o]

    def replace_cell_values(self, outputs, value):    # model_kernels.py:104-109
        outs = list(outputs)
        self.cell_values[:, outs] = value.reshape(-1, len(outs))


# ------------------------------------------------------------------ terms
def convection_weights(fvm, outputs=(0, 1, 2), implicit_outflow=False, scheme='upwind'):
    """convection.py:44-88 with convectionScheme.py: 'upwind' (28-52),
    'central difference' (3-26: owner psi, neighbour 1 - psi) or
    'upwindLinear' (54-88: upwind weights plus the explicit
    grad(phi)_upwind . r of the global field, r = the owner's
    centroid-to-face vector of local face j[0] when the owner is upwind and,
    as the reference reads it, of local face j[1] -- the neighbour's index --
    when the neighbour is).

    Internal explicit (skew) weights are computed like the reference
    (convection.py:76-78, local output index) but never reach b (T2).

    ``implicit_outflow`` (not in the reference): on a boundary face with a
    Neumann condition and outward flux F_b > 0 the face value
    phi_b = phi_P + |x_f - x_P| g is split into an implicit F_b phi_P
    (diagonal) and an explicit F_b |x_f - x_P| g, instead of the reference's
    fully lagged F_b phi_b.  The two agree at convergence."""
    m = fvm.m
    dt = fvm.dtype
    O = len(outputs)
    w = np.zeros((m.num_cells, m.faces_per_cell, O, 3), dt)
    F = fvm.mass_fluxes
    b = m.boundary_face_ids
    ob, kb = m.adjacent_cells[b, 0], m.cell_face_index[b, 0]
    fi = m.internal_face_ids
    own, nb = m.adjacent_cells[fi, 0], m.adjacent_cells[fi, 1]
    k0, k1 = m.cell_face_index[fi, 0], m.cell_face_index[fi, 1]
    Fi = F[fi]
    up = Fi > 0                                         # T6: tie -> neighbour
    psi = fvm.psi[fi][:, None]; psi2 = fvm.psi2[fi][:, None]
    if scheme == 'central difference':
        w0, w1 = psi[:, 0].copy(), psi2[:, 0].copy()
    elif scheme in ('upwind', 'upwindLinear'):
        w0 = np.where(up, dt.type(1), dt.type(0))
        w1 = np.where(up, dt.type(0), dt.type(1))
    else:
        raise ValueError(scheme)
    if scheme == 'upwindLinear':
        upc = np.where(up, own, nb)
        r_up = np.where(up[:, None], fvm.c2f[own, k0], fvm.c2f[own, k1])
    rk = fvm.face_centroid[fi] - (fvm.centroids[own] * psi + fvm.centroids[nb] * psi2)
    bnd = fvm.boundary
    c2f = fvm.c2f[ob, kb]
    dmag_b = np.sqrt(c2f[:, 0] * c2f[:, 0] + c2f[:, 1] * c2f[:, 1] + c2f[:, 2] * c2f[:, 2])
    for lo, go in enumerate(outputs):
        w[ob, kb, lo, 0] = 0
        w[ob, kb, lo, 1] = 0
        w[ob, kb, lo, 2] = F[b] * fvm.face_values[b, go]
        if implicit_outflow:
            out = (bnd.types[:, go] == NEUMANN) & (F[b] > 0)
            w[ob[out], kb[out], lo, 0] = F[b][out]
            w[ob[out], kb[out], lo, 2] = F[b][out] * (dmag_b[out] * fvm.face_gradients[b[out], go])
        gi = fvm.cell_gradients[own, lo] * psi + psi2 * fvm.cell_gradients[nb, lo]
        skew = (gi * rk).sum(axis=1, dtype=dt)
        if scheme == 'upwindLinear':
            gu = fvm.cell_gradients[upc, go]
            lin = gu[:, 0] * r_up[:, 0] + gu[:, 1] * r_up[:, 1] + gu[:, 2] * r_up[:, 2]
            skew = lin + skew
        w[own, k0, lo, 0] = w0 * Fi
        w[own, k0, lo, 1] = w1 * Fi
        w[own, k0, lo, 2] = skew * Fi
        mf = -Fi
        w[nb, k1, lo, 0] = w1 * mf
        w[nb, k1, lo, 1] = w0 * mf
        w[nb, k1, lo, 2] = skew * mf
    return w


def diffusion_weights(fvm, viscosity, outputs, correction=False):
    """diffusion.py:40-49, 113-152 (BC), 155-197 (internal);
    laplacianScheme.py:14-37 (central) and 39-66 (corrected)."""
    m = fvm.m
    dt = fvm.dtype
    O = len(outputs)
    w = np.zeros((m.num_cells, m.faces_per_cell, O, 3), dt)
    nu_arr = np.atleast_1d(np.asarray(viscosity, dt))
    nu_f = (lambda f: nu_arr[0]) if nu_arr.shape[0] == 1 else (lambda f: nu_arr[f])
    bnd = fvm.boundary
    b = bnd.boundary_ids
    ob, kb = m.adjacent_cells[b, 0], m.cell_face_index[b, 0]
    A_b = fvm.face_area[b]
    c2f = fvm.c2f[ob, kb]
    dmag = np.sqrt(c2f[:, 0] * c2f[:, 0] + c2f[:, 1] * c2f[:, 1] + c2f[:, 2] * c2f[:, 2])
    nu_b = nu_f(b)
    for lo, go in enumerate(outputs):
        nmn = bnd.types[:, go] == NEUMANN
        weight = nu_b * A_b / dmag
        w[ob, kb, lo, 0] = np.where(nmn, dt.type(0), -weight)
        w[ob, kb, lo, 1] = 0
        w[ob, kb, lo, 2] = np.where(nmn, nu_b * fvm.face_gradients[b, go] * A_b,
                                    weight * fvm.face_values[b, go])
    fi = m.internal_face_ids
    own, nb = m.adjacent_cells[fi, 0], m.adjacent_cells[fi, 1]
    k0, k1 = m.cell_face_index[fi, 0], m.cell_face_index[fi, 1]
    weight = nu_f(fi) * fvm.face_area[fi]
    d = fvm.cell_distance[fi]
    dmag = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    if not correction:
        f0 = dt.type(-1.0) / dmag
        f1 = dt.type(1.0) / dmag
        f2 = [np.zeros_like(dmag)] * O
    else:
        n_f = fvm.normal[own, k0]
        nd = (n_f * d).sum(axis=1, dtype=dt)
        n1 = d * (dt.type(1.0) / nd)[:, None]
        n1m = np.sqrt((n1 * n1).sum(axis=1, dtype=dt))
        n2 = n_f - n1
        f0 = dt.type(-1.0) / dmag * n1m
        f1 = dt.type(1.0) / dmag * n1m
        psi = fvm.psi[fi][:, None]; psi2 = fvm.psi2[fi][:, None]
        f2 = []
        for go in outputs:
            gf = psi * fvm.cell_gradients[own, go] + psi2 * fvm.cell_gradients[nb, go]
            f2.append((gf * n2).sum(axis=1, dtype=dt))
    for lo in range(O):
        w[own, k0, lo, 0] = f0 * weight
        w[own, k0, lo, 1] = f1 * weight
        w[own, k0, lo, 2] = f2[lo] * weight
        w[nb, k1, lo, 0] = -f1 * weight
        w[nb, k1, lo, 1] = -f0 * weight
        w[nb, k1, lo, 2] = -f2[lo] * weight
    return w


def grad_term_weights(fvm, o=3):
    """grad.py:7-28: w[3c+d] = (grad p)_c[d] V_c."""
    return (fvm.cell_gradients[:, o] * fvm.volumes[:, None]).reshape(-1).astype(fvm.dtype)


def ddt_weights(fvm, dt_step, outputs=(0, 1, 2)):
    """ddt.py:7-33 (backwards_Euler): w[c,o] = [V/dt, -phi V/dt]."""
    d = fvm.dtype
    vdt = fvm.volumes / d.type(dt_step)
    w = np.zeros((fvm.m.num_cells, len(outputs), 2), d)
    for lo, go in enumerate(outputs):
        w[:, lo, 0] = vdt
        w[:, lo, 1] = -fvm.cell_values[:, go] * vdt
    return w


# ---------------------------------------------------------------- equation
class OracleEquation:
    """FV/terms/equation.py:10-204 over the reference's interleaved CSR."""

    def __init__(self, fvm, outputs):
        self.fvm = fvm
        self.outputs = list(outputs)
        self.O = len(outputs)
        self.offsets, self.columns, self.rows = fvm.m.csr_pattern(self.O)
        self.n = fvm.m.num_cells * self.O
        self.values = np.zeros(len(self.columns), fvm.dtype)
        self.rhs = np.zeros(self.n, fvm.dtype)
        self.x = np.zeros(self.n, fvm.dtype)
        # position of (cell c, local face k, output o) and of each diagonal
        m = fvm.m
        C = m.num_cells
        self.diag_pos = np.zeros(self.n, np.int64)
        rr = self.rows
        dpos = np.nonzero(rr == self.columns)[0]
        self.diag_pos[rr[dpos]] = dpos
        self.slot_pos = -np.ones((C, m.faces_per_cell, self.O), np.int64)
        # rows ascending, columns sorted inside a row -> (row, col) keys sorted
        keys = self.rows.astype(np.int64) * self.n + self.columns
        for o in range(self.O):
            for k in range(m.faces_per_cell):
                nbr = m.neighbors[:, k]
                cc = np.nonzero(nbr >= 0)[0]
                row = self.O * cc + o
                col = self.O * nbr[cc].astype(np.int64) + o
                p = np.searchsorted(keys, row.astype(np.int64) * self.n + col)
                assert np.all(keys[p] == row.astype(np.int64) * self.n + col)
                self.slot_pos[cc, k, o] = p

    def reset(self):                                   # equation.py:202-204
        self.rhs[:] = 0
        self.values[:] = 0

    def add_face_term(self, w, scale):
        """calculate_BSR_values_kernel (equation_kernels.py:52-79) and
        calculate_RHS_values_kernel (84-101): diagonal = scale * sum over the
        local faces of w[...,0] (boundary slots included); off-diagonal =
        scale * w[...,1]; b -= scale * w[...,2] for boundary faces only (T2)."""
        d = self.fvm.dtype
        s = d.type(scale)
        m = self.fvm.m
        C = m.num_cells
        for o in range(self.O):
            dp = self.diag_pos[self.O * np.arange(C) + o]
            acc = self.values[dp]
            for k in range(m.faces_per_cell):
                acc = acc + s * w[:, k, o, 0]
            self.values[dp] = acc
            for k in range(m.faces_per_cell):
                sp = self.slot_pos[:, k, o]
                ok = sp >= 0
                self.values[sp[ok]] = self.values[sp[ok]] + s * w[ok, k, o, 1]
        b = m.boundary_face_ids
        cb, kb = m.adjacent_cells[b, 0], m.cell_face_index[b, 0]
        for o in range(self.O):
            row = cb * self.O + o
            # one boundary face per (row, slot): no two faces collide on the same
            # slot, but a row may receive several boundary faces -> ordered add
            np.add.at(self.rhs, row, -s * w[cb, kb, o, 2])

    def add_cell_term(self, w, scale):
        """add_Implicit_cell_based_term_kernel (equation_kernels.py:104-122)."""
        d = self.fvm.dtype
        s = d.type(scale)
        C = self.fvm.m.num_cells
        for o in range(self.O):
            rows = self.O * np.arange(C) + o
            dp = self.diag_pos[rows]
            self.values[dp] = self.values[dp] + s * w[:, o, 0]
            self.rhs[rows] = self.rhs[rows] + (-s * w[:, o, 1])

    def form_system(self, implicit, explicit=()):
        """equation.py:99-124: implicit = [(kind, weights, scale)]."""
        self.reset()
        for kind, w, scale in implicit:
            if kind == 'cell':
                self.add_cell_term(w, scale)
            else:
                self.add_face_term(w, scale)
        for w, scale in explicit:
            self.rhs = self.rhs + w * self.fvm.dtype.type(scale)

    def add_RHS(self, arr, scale=1.0):                  # equation.py:169-170
        self.rhs = self.rhs + arr * self.fvm.dtype.type(scale)

    def relax(self, alpha):
        """implicit_relaxation_kernel (equation_kernels.py:128-152)."""
        d = self.fvm.dtype
        a = d.type(alpha)
        C = self.fvm.m.num_cells
        for o in range(self.O):
            rows = self.O * np.arange(C) + o
            dp = self.diag_pos[rows]
            cv = self.fvm.cell_values[:, self.outputs[o]]
            self.rhs[rows] = self.rhs[rows] + (d.type(1.0) - a) / a * self.values[dp] * cv
            self.values[dp] = self.values[dp] / a

    def replace_row(self, row_id, value):
        """replace_row_kernel (equation_kernels.py:155-175).  The reference
        reads rhs_values[row_id] from a length-1 array (T7): only row 0 is
        defined; the intended value is used for any row here."""
        s, e = self.offsets[row_id], self.offsets[row_id + 1]
        cols = self.columns[s:e]
        self.values[s:e] = np.where(cols == row_id, 1.0, 0.0).astype(self.fvm.dtype)
        self.rhs[row_id] = self.fvm.dtype.type(value)

    @property
    def diagonal(self):                                # equation.py:80-83
        return self.values[self.diag_pos].copy()

    def matrix(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.values, self.columns, self.offsets), shape=(self.n, self.n))
