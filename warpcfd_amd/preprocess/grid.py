"""Synthetic mesh generators and boundary wall tagging.

The reference meshes with gmsh (preprocess/grid.py:8-159), which is absent;
these generators produce the same kind of input -- raw ``(nodes, cells)``
arrays of VTK hexahedra, tetrahedra or wedges -- for the cavity, channel and
cylinder workloads.  ``define_boundary_walls`` follows
preprocess/grid.py:165-194.
"""
import numpy as np

# a lattice cube in VTK hexahedron order: node k at these (x, y, z) offsets
_CUBE = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                  [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.float64)
# Kuhn subdivision: one tetrahedron per order of the axes along the path
# from node 0 to node 6; conforming between lattice neighbours
_KUHN = [[0, 1, 2, 6], [0, 1, 5, 6], [0, 3, 2, 6], [0, 3, 7, 6], [0, 4, 5, 6], [0, 4, 7, 6]]
# two prisms on the (0, 2) diagonal of the bottom and top faces
_PRISMS = [[0, 1, 2, 4, 5, 6], [0, 2, 3, 4, 6, 7]]


def _oriented(split, swap):
    """Orders each sub-cell so that its first three nodes (p1 - p0) x (p2 - p0)
    point at the rest of the cell -- the orientation for which the local
    face tables (preprocess/mesh.py:69-91) give outward normals."""
    out = []
    for t in split:
        t = list(t)
        p = _CUBE[t]
        n = np.cross(p[1] - p[0], p[2] - p[0])
        if np.dot(n, p[len(t) // 2 if len(t) == 6 else 3] - p[0]) < 0:
            for a, b in swap:
                t[a], t[b] = t[b], t[a]
        out.append(t)
    return np.array(out, np.int64)


def split_hexes(cells, cell_type):
    """Split lattice hexahedra (VTK order, as ``structured_hex`` makes them)
    into ``'tetra'`` (6 per hexahedron) or ``'wedge'`` (2 per hexahedron)
    cells, the sub-cells of hexahedron i at rows i*k .. i*k + k - 1."""
    cells = np.asarray(cells)
    if cell_type == 'tetra':
        split = _oriented(_KUHN, [(1, 2)])
    elif cell_type == 'wedge':
        split = _oriented(_PRISMS, [(1, 2), (4, 5)])
    else:
        raise ValueError("cell_type must be 'tetra' or 'wedge'")
    return cells[:, split].reshape(-1, split.shape[1]).astype(np.int32)


def structured_cells(counts, lengths=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), jitter=0.0,
                     seed=0, cell_type='hexahedron'):
    """``structured_hex``'s box as hexahedra, tetrahedra or wedges."""
    nodes, cells = structured_hex(counts, lengths, origin, jitter, seed)
    if cell_type == 'hexahedron':
        return nodes, cells
    return nodes, split_hexes(cells, cell_type)


def structured_hex(counts, lengths=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0),
                   jitter=0.0, seed=0):
    """Box of nx*ny*nz hexahedra, nodes on a lattice, cells lexicographic
    (x fastest), VTK node order (0-3 bottom counter-clockwise, 4-7 top).

    ``jitter`` moves every *interior* node by U(-jitter*h, +jitter*h) per axis
    (h = the cell size along that axis) from ``numpy.random.default_rng(seed)``;
    boundary nodes stay on the box so boundary normals stay axis-aligned.
    Returns (nodes (Nn,3) float64, cells (C,8) int32).
    """
    nx, ny, nz = (int(c) for c in counts)
    Lx, Ly, Lz = lengths
    ox, oy, oz = origin
    xs = ox + np.linspace(0.0, Lx, nx + 1)
    ys = oy + np.linspace(0.0, Ly, ny + 1)
    zs = oz + np.linspace(0.0, Lz, nz + 1)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing='ij')
    nodes = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        h = np.array([Lx / nx, Ly / ny, Lz / nz])
        d = rng.uniform(-jitter, jitter, size=nodes.shape) * h
        ii = np.arange(nx + 1)
        jj = np.arange(ny + 1)
        kk = np.arange(nz + 1)
        K, J, I = np.meshgrid(kk, jj, ii, indexing='ij')
        interior = ((I > 0) & (I < nx) & (J > 0) & (J < ny) & (K > 0) & (K < nz)).ravel()
        nodes[interior] += d[interior]
    i = np.arange(nx)
    j = np.arange(ny)
    k = np.arange(nz)
    K, J, I = np.meshgrid(k, j, i, indexing='ij')
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    sx, sy = 1, nx + 1
    sz = (nx + 1) * (ny + 1)
    n0 = I * sx + J * sy + K * sz
    cells = np.stack([n0, n0 + sx, n0 + sx + sy, n0 + sy,
                      n0 + sz, n0 + sx + sz, n0 + sx + sy + sz, n0 + sy + sz], axis=1)
    return nodes, cells.astype(np.int32)


class UnstructuredGrid:
    """The two members of a pyvista ``UnstructuredGrid`` that ``Mesh`` reads:
    ``points`` (N, 3) and ``cells_dict`` {VTK cell type: (C, n) node ids}."""

    def __init__(self, points, cells, cell_type):
        self.points = np.asarray(points, np.float64)
        self.cells_dict = {int(cell_type): np.asarray(cells, np.int32)}

    @property
    def n_cells(self):
        return next(iter(self.cells_dict.values())).shape[0]


_ELEMENTS = {'hex': (12, None), 'wedge': (13, 'wedge'), 'tet': (10, 'tetra')}


def create_2D_grid(origin, nx, ny, dx, dy, dz=0.1, *, element_type='hex',
                   unstructured_wedge=False, display_mesh=False, save=None,
                   gmsh_version2=False):
    """preprocess/grid.py:8-159 without gmsh: an nx x ny x 1 box of size
    dx x dy x dz at ``origin`` as hexahedra, or each hexahedron split into 2
    wedges or 6 tetrahedra, returned as an ``UnstructuredGrid`` for ``Mesh``.
    ``save`` writes ``{save}.msh`` (format 2.2 if ``gmsh_version2``, else
    4.1, gmsh's default).  ``display_mesh`` opened gmsh's viewer and is
    ignored.  Cell and node numbering is lexicographic (x fastest) rather than
    gmsh's; ``unstructured_wedge`` needs gmsh's mesher and is refused -- read
    such a mesh from its ``.msh`` file with ``read_msh`` instead."""
    if element_type not in _ELEMENTS:
        raise ValueError(f'element_type must be one of {sorted(_ELEMENTS)}, got {element_type!r}')
    if unstructured_wedge:
        raise NotImplementedError('unstructured wedge meshing needs gmsh; mesh with gmsh and '
                                  'load the .msh file with warp_cfd.preprocess.read_msh')
    vtk, split = _ELEMENTS[element_type]
    nodes, cells = structured_hex((nx, ny, 1), (dx, dy, dz), origin)
    if split:
        cells = split_hexes(cells, split)
    if save is not None:
        from .io import write_msh
        write_msh(f'{save}.msh', nodes, cells, version=2.2 if gmsh_version2 else 4.1)
    return UnstructuredGrid(nodes, cells, vtk)


def cavity_mesh(n, dz=0.1):
    """Unit-square n x n x 1 cavity, dz thick (examples/LDC/LDC.py:24-29)."""
    return structured_hex((n, n, 1), (1.0, 1.0, dz))


def channel_mesh(nx=200, ny=40, nz=40, lengths=(10.0, 2.0, 2.0), jitter=0.1, seed=0):
    """Box channel with +-jitter*h interior-node jitter."""
    return structured_hex((nx, ny, nz), lengths, jitter=jitter, seed=seed)


def cylinder_mesh(h=0.05, lengths=(20.0, 8.0, 1.0), centre=(5.0, 4.0), radius=0.5,
                  jitter=0.1, seed=0):
    """Voxel-carved cylinder in a box: cells whose centroid lies inside the
    circle are removed; unused nodes are dropped and the rest renumbered.
    Interior nodes away from the carved surface are jittered."""
    nx, ny, nz = (int(round(L / h)) for L in lengths)
    nodes, cells = structured_hex((nx, ny, nz), lengths, jitter=0.0)
    cen = nodes[cells].mean(axis=1)
    keep = (cen[:, 0] - centre[0]) ** 2 + (cen[:, 1] - centre[1]) ** 2 > radius ** 2
    cells = cells[keep]
    used = np.zeros(nodes.shape[0], bool)
    used[cells.ravel()] = True
    # nodes touching a removed cell lie on the carved surface: keep them fixed
    removed = np.zeros(nodes.shape[0], bool)
    full_cells = structured_hex((nx, ny, nz), lengths, jitter=0.0)[1]
    removed[full_cells[~keep].ravel()] = True
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        hv = np.array([lengths[0] / nx, lengths[1] / ny, lengths[2] / nz])
        d = rng.uniform(-jitter, jitter, size=nodes.shape) * hv
        eps = 1e-9
        lo = np.array([0.0, 0.0, 0.0]); hi = np.array(lengths, float)
        interior = np.all((nodes > lo + eps) & (nodes < hi - eps), axis=1) & ~removed
        nodes = nodes + d * interior[:, None]
    remap = -np.ones(nodes.shape[0], np.int64)
    remap[used] = np.arange(used.sum())
    return nodes[used], remap[cells].astype(np.int32)


def define_boundary_walls(mesh):
    """preprocess/grid.py:165-194: group boundary faces by their owner-side
    normal, |n.e - 1| < 1e-3 for e = +-x, +-y, +-z."""
    b = mesh.boundary_face_ids
    n = mesh.boundary_normals()
    walls = {}
    for name, e in (("+X", [1, 0, 0]), ("-X", [-1, 0, 0]), ("+Y", [0, 1, 0]),
                    ("-Y", [0, -1, 0]), ("+Z", [0, 0, 1]), ("-Z", [0, 0, -1])):
        walls[name] = b[np.abs(n @ np.array(e, dtype=n.dtype) - 1) < 1e-3]
    mesh.add_groups(walls, 'face')
    return walls


def box_and_body_walls(mesh, lengths, origin=(0.0, 0.0, 0.0), name='cylinder', tol=1e-6):
    """Boundary groups of a box with a carved body: faces whose centroid lies
    on a box plane go to '-X', '+X', '-Y', '+Y', '-Z', '+Z'; all other
    boundary faces (the carved surface) to ``name``.  Not in the reference:
    define_boundary_walls (grid.py:165-194) classifies by normal, which puts
    the axis-aligned faces of a voxel-carved body into the box groups."""
    b = mesh.boundary_face_ids
    c = np.asarray(mesh.face_properties.centroid, np.float64)[b]
    lo = np.asarray(origin, np.float64)
    hi = lo + np.asarray(lengths, np.float64)
    walls = {}
    on_box = np.zeros(b.size, bool)
    for ax, (n_lo, n_hi) in enumerate((('-X', '+X'), ('-Y', '+Y'), ('-Z', '+Z'))):
        at_lo = np.abs(c[:, ax] - lo[ax]) < tol * max(1.0, abs(hi[ax]))
        at_hi = np.abs(c[:, ax] - hi[ax]) < tol * max(1.0, abs(hi[ax]))
        walls[n_lo] = b[at_lo & ~on_box]
        on_box |= at_lo
        walls[n_hi] = b[at_hi & ~on_box]
        on_box |= at_hi
    walls[name] = b[~on_box]
    mesh.add_groups(walls, 'face')
    return walls
