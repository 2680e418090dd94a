"""Mesh files in and out: gmsh .msh (format 2.2 and 4.1, ASCII) and VTK
.vtu (XML UnstructuredGrid: ascii, inline binary, appended; zlib-compressed or
not).

The reference builds meshes with gmsh and pyvista (preprocess/grid.py:8-159,
mesh.py:17-43), which are absent here.  ``read_msh`` and ``read_vtu``
return a ``MeshData`` -- points, the cells of one type (hexahedra,
tetrahedra or wedges, the types the reference's Mesh takes, mesh.py:20,37)
in file order, named groups -- that ``Mesh`` accepts as it accepts a pyvista
grid; node ids become 0-based positions in file order, as
``create_2D_grid`` makes them (grid.py:153-154).  ``apply_groups`` turns
physical groups into the mesh's groups: boundary triangles and
quadrilaterals into face groups (matched by their node sets), volume
elements into cell groups.  The writers produce files that gmsh, meshio and
ParaView read, and ``write_vtu`` is also the result export (cell fields
beside the mesh, as the reference's examples attach them to the pyvista
grid, examples/LDC/LDC.py:61-65).
"""
import base64
import re
import struct
import zlib
import xml.etree.ElementTree as ET

import numpy as np

VTK_TETRA, VTK_HEXAHEDRON, VTK_WEDGE = 10, 12, 13
GMSH_TRI, GMSH_QUAD, GMSH_TET, GMSH_HEX, GMSH_PRISM = 2, 3, 4, 5, 6
# gmsh and VTK list the nodes of these three volume elements in the same order
_GMSH_TO_VTK = {GMSH_TET: VTK_TETRA, GMSH_HEX: VTK_HEXAHEDRON, GMSH_PRISM: VTK_WEDGE}
_VTK_TO_GMSH = {v: k for k, v in _GMSH_TO_VTK.items()}
_NODES = {VTK_TETRA: 4, VTK_HEXAHEDRON: 8, VTK_WEDGE: 6}
_TYPE_OF_NODES = {v: k for k, v in _NODES.items()}
_FACE_TYPES = (GMSH_TRI, GMSH_QUAD)


def _vtk_type(cells):
    n = np.asarray(cells).shape[-1]
    if n not in _TYPE_OF_NODES:
        raise ValueError('cells must be tetrahedra (4 nodes), wedges (6) or hexahedra (8)')
    return _TYPE_OF_NODES[n]


def _pad4(faces):
    """Face node lists as (Q, 4), triangles padded with -1 (the reference
    pads a wedge's triangles so, mesh.py:131-135)."""
    f = np.asarray(faces, np.int64)
    f = f.reshape(-1, f.shape[-1]) if f.size else f.reshape(0, 4)
    if f.shape[1] < 4:
        f = np.hstack([f, -np.ones((f.shape[0], 4 - f.shape[1]), np.int64)])
    return f


class MeshData:
    """Points (N, 3), cells (C, n) int32 of one VTK type (10 tetrahedron,
    12 hexahedron, 13 wedge; VTK node order), cell data {name: (C, ...)}
    and groups {name: (dim, array)}: dim 3 -> cell indices, dim 2 -> face
    node ids (Q, 4), triangles padded with -1."""

    def __init__(self, points, cells, cell_data=None, groups=None, cell_type=None):
        self.points = np.asarray(points, np.float64)
        cells = np.asarray(cells, np.int32)
        self.cell_type = int(cell_type) if cell_type is not None else _vtk_type(cells)
        self.cells = cells.reshape(-1, _NODES[self.cell_type])
        self.cell_data = dict(cell_data or {})
        self.groups = dict(groups or {})

    @property
    def hexes(self):
        if self.cell_type != VTK_HEXAHEDRON:
            raise AttributeError('the mesh holds no hexahedra')
        return self.cells

    @property
    def cells_dict(self):
        return {self.cell_type: self.cells}


# ------------------------------------------------------------------ gmsh
def _sections(text):
    out = {}
    for m in re.finditer(r'^\$(\w+)\s*$(.*?)^\$End\1\s*$', text, re.S | re.M):
        out.setdefault(m.group(1), []).append(m.group(2).strip('\n'))
    return out


def _parse(s, dtype, count=None):
    """Whitespace-separated numbers of s, parsed in C (np.fromstring's text
    mode); a token that is not a number ends the parse, which the count
    check reports."""
    a = np.fromstring(s, dtype=dtype, sep=' ')
    if count is not None and a.size != count:
        raise ValueError(f'expected {count} numbers, read {a.size}')
    return a


def _ints(s):
    return _parse(s, np.int64)


def read_msh(path):
    """gmsh .msh, format 2.2 or 4.1, ASCII.  The volume elements -- all
    hexahedra (type 5), tetrahedra (4) or prisms (6) -- become the cells;
    physical groups of triangles (2), quadrilaterals (3) and volume elements
    become named groups (the physical name, else 'physical_<tag>')."""
    with open(path, 'rb') as f:
        raw = f.read()
    text = raw.decode('latin-1')
    sec = _sections(text)
    if 'MeshFormat' not in sec:
        raise ValueError(f'{path}: no $MeshFormat section')
    fmt = sec['MeshFormat'][0].split()
    version, ftype = float(fmt[0]), int(fmt[1])
    if ftype != 0:
        raise NotImplementedError('binary .msh files are not read; save with Mesh.Binary = 0')
    names = {}
    for block in sec.get('PhysicalNames', []):
        lines = block.splitlines()[1:]
        for ln in lines:
            dim, tag, name = ln.split(maxsplit=2)
            names[(int(dim), int(tag))] = name.strip().strip('"')
    if version < 3:
        return _read_msh2(sec, names)
    if 4 <= version < 5:
        return _read_msh4(sec, names)
    raise NotImplementedError(f'.msh format {version}')


def _finish(tags, coords, blocks, names):
    """blocks: list of (element type, physical tag or None, node tags (E, n))."""
    order = np.argsort(tags, kind='stable')
    st = tags[order]

    def index(t):
        p = np.searchsorted(st, t)
        if np.any(p >= st.size) or np.any(st[np.minimum(p, st.size - 1)] != t):
            raise ValueError('an element references a node that is not defined')
        return order[p]
    vols, vol_phys, faces = {}, {}, {}
    for etype, phys, nodes in blocks:
        ph = (np.full(len(nodes), -1 if phys is None else phys, np.int64)
              if np.isscalar(phys) or phys is None else phys)
        if etype in _GMSH_TO_VTK:
            vols.setdefault(etype, []).append(index(nodes))
            vol_phys.setdefault(etype, []).append(ph)
        elif etype in _FACE_TYPES:
            for t in np.unique(ph[ph > 0]):
                faces.setdefault(int(t), []).append(_pad4(index(nodes[ph == t])))
    if not vols:
        raise ValueError('no volume elements (gmsh types 4, 5, 6) in the file')
    if len(vols) > 1:
        raise ValueError('Meshes can only contain a single cell type for now, Got multiple')
    (etype, parts), = vols.items()
    cells = np.concatenate(parts).astype(np.int32)
    hp = np.concatenate(vol_phys[etype])
    groups = {}
    for tag in sorted(faces):
        groups[names.get((2, tag), f'physical_{tag}')] = (2, np.concatenate(faces[tag]).astype(np.int32))
    for tag in np.unique(hp[hp > 0]).tolist():
        groups[names.get((3, tag), f'physical_{tag}')] = (3, np.nonzero(hp == tag)[0].astype(np.int32))
    return MeshData(coords, cells, groups=groups, cell_type=_GMSH_TO_VTK[etype])


def _read_msh2(sec, names):
    lines = sec['Nodes'][0].splitlines()
    n = int(lines[0])
    a = _parse(' '.join(lines[1:n + 1]), np.float64, 4 * n).reshape(n, 4)
    tags, coords = a[:, 0].astype(np.int64), a[:, 1:]
    lines = sec['Elements'][0].splitlines()
    rows = lines[1:int(lines[0]) + 1]
    # element lines of one type and tag count have one length: parse each
    # length class as one array, keeping file order within a type
    width = np.array([ln.count(' ') for ln in rows])
    blocks = []
    keys = []
    for wd in np.unique(width):
        sel = np.nonzero(width == wd)[0]
        arr = _parse(' '.join(rows[i] for i in sel), np.int64).reshape(sel.size, -1)
        for et in np.unique(arr[:, 1]):
            m = arr[:, 1] == et
            ntags = arr[m, 2]
            if np.any(ntags != ntags[0]):
                raise ValueError('mixed tag counts within one element type')
            nt = int(ntags[0])
            phys = arr[m, 3] if nt > 0 else np.full(int(m.sum()), -1, np.int64)
            keys.append((int(et), sel[m]))
            blocks.append((int(et), phys, arr[m, 3 + nt:]))
    # restore file order within each element type
    out = []
    for et in sorted({b[0] for b in blocks}):
        parts = [(k[1], b) for k, b in zip(keys, blocks) if b[0] == et]
        pos = np.concatenate([p for p, _ in parts])
        nodes = np.concatenate([b[2] for _, b in parts])
        phys = np.concatenate([b[1] for _, b in parts])
        o = np.argsort(pos, kind='stable')
        out.append((et, phys[o], nodes[o]))
    return _finish(tags, coords, out, names)


def _read_msh4(sec, names):
    # entity tag -> physical tag, per dimension
    ent_phys = {}
    if 'Entities' in sec:
        lines = sec['Entities'][0].splitlines()
        counts = [int(x) for x in lines[0].split()]
        row = 1
        for dim, cnt in enumerate(counts):
            for _ in range(cnt):
                v = lines[row].split()
                row += 1
                tag = int(v[0])
                nph_at = 4 if dim == 0 else 7
                nph = int(v[nph_at])
                phys = [int(x) for x in v[nph_at + 1:nph_at + 1 + nph]]
                ent_phys[(dim, tag)] = phys[0] if phys else None
    lines = sec['Nodes'][0].splitlines()
    n_blocks = int(lines[0].split()[0])
    row = 1
    tags, coords = [], []
    for _ in range(n_blocks):
        ed, et, para, nn = (int(x) for x in lines[row].split())
        row += 1
        tags.append(_parse(' '.join(lines[row:row + nn]), np.int64, nn))
        row += nn
        xyz = _parse(' '.join(lines[row:row + nn]), np.float64).reshape(nn, -1)[:, :3]
        coords.append(xyz)
        row += nn
    tags = np.concatenate(tags)
    coords = np.concatenate(coords) if coords else np.zeros((0, 3))
    blocks = []
    lines = sec['Elements'][0].splitlines()
    n_blocks = int(lines[0].split()[0])
    row = 1
    for _ in range(n_blocks):
        ed, et, etype, ne = (int(x) for x in lines[row].split())
        row += 1
        if ne:
            block = _parse(' '.join(lines[row:row + ne]), np.int64).reshape(ne, -1)
            blocks.append((etype, ent_phys.get((ed, et)), block[:, 1:]))
        row += ne
    return _finish(tags, coords, blocks, names)


def _face_rows(q):
    """(gmsh type, node lists) pairs of a (Q, 4) face array: triangles
    (padded with -1) and quadrilaterals."""
    tri = q[:, 3] < 0
    return [(t, q[m][:, :n]) for t, m, n in ((GMSH_TRI, tri, 3), (GMSH_QUAD, ~tri, 4)) if m.any()]


def write_msh(path, nodes, cells, face_groups=None, version=2.2, cell_groups=None):
    """ASCII .msh of hexahedra, tetrahedra or wedges (prisms), with physical
    groups: ``face_groups`` {name: (Q, 3 or 4) node ids of boundary faces,
    triangles padded with -1} and ``cell_groups`` {name: cell indices}.
    Node ids are written 1-based in array order."""
    nodes = np.asarray(nodes, np.float64)
    cells = np.asarray(cells, np.int64)
    gtype = _VTK_TO_GMSH[_vtk_type(cells)]
    face_groups = face_groups or {}
    cell_groups = cell_groups or {}
    phys = []                                   # (dim, tag, name, element node arrays)
    tag = 1
    for name, q in face_groups.items():
        phys.append((2, tag, name, _pad4(q)))
        tag += 1
    hex_tag = np.zeros(len(cells), np.int64)
    for name, ids in cell_groups.items():
        hex_tag[np.asarray(ids, np.int64)] = tag
        phys.append((3, tag, name, None))
        tag += 1
    out = []
    w = out.append
    if version < 3:
        w('$MeshFormat\n2.2 0 8\n$EndMeshFormat\n')
    else:
        w('$MeshFormat\n4.1 0 8\n$EndMeshFormat\n')
    if phys:
        w(f'$PhysicalNames\n{len(phys)}\n')
        for dim, t, name, _ in phys:
            w(f'{dim} {t} "{name}"\n')
        w('$EndPhysicalNames\n')
    N = len(nodes)
    if version < 3:
        w(f'$Nodes\n{N}\n')
        w(''.join(f'{i + 1} {x!r} {y!r} {z!r}\n' for i, (x, y, z) in enumerate(nodes.tolist())))
        w('$EndNodes\n')
        rows = []
        eid = 1
        for dim, t, name, q in phys:
            if q is None:
                continue
            for ft, fq in _face_rows(q):
                for e in fq:
                    rows.append(f'{eid} {ft} 2 {t} {t} ' + ' '.join(str(v + 1) for v in e))
                    eid += 1
        for i, e in enumerate(cells):
            t = int(hex_tag[i])
            rows.append(f'{eid} {gtype} 2 {t} {max(t, 1)} ' + ' '.join(str(v + 1) for v in e))
            eid += 1
        w(f'$Elements\n{len(rows)}\n' + '\n'.join(rows) + '\n$EndElements\n')
    else:
        # one entity per physical group, plus a volume entity per hexahedron group
        # and one for hexahedra in none
        surf = [(t, q) for dim, t, name, q in phys if q is not None]
        vols = [t for dim, t, name, q in phys if q is None]
        free = np.any(hex_tag == 0)
        vol_entities = [(i + 1, t) for i, t in enumerate(vols)]
        if free:
            vol_entities.append((len(vols) + 1, 0))
        w('$Entities\n')
        w(f'0 0 {len(surf)} {len(vol_entities)}\n')
        for i, (t, q) in enumerate(surf):
            w(f'{i + 1} 0 0 0 0 0 0 1 {t} 0\n')
        for et, t in vol_entities:
            w(f'{et} 0 0 0 0 0 0 ' + (f'1 {t}' if t else '0') + ' 0\n')
        w('$EndEntities\n')
        w(f'$Nodes\n1 {N} 1 {N}\n3 1 0 {N}\n')
        w(''.join(f'{i + 1}\n' for i in range(N)))
        w(''.join(f'{x!r} {y!r} {z!r}\n' for x, y, z in nodes.tolist()))
        w('$EndNodes\n')
        blocks = []
        eid = 1
        for i, (t, q) in enumerate(surf):
            for ft, fq in _face_rows(q):          # one block per element type
                rows = []
                for e in fq:
                    rows.append(f'{eid} ' + ' '.join(str(v + 1) for v in e))
                    eid += 1
                blocks.append(f'2 {i + 1} {ft} {len(rows)}\n' + '\n'.join(rows))
        for et, t in vol_entities:
            sel = np.nonzero(hex_tag == t)[0]
            rows = []
            for c in sel:
                rows.append(f'{eid} ' + ' '.join(str(v + 1) for v in cells[c]))
                eid += 1
            blocks.append(f'3 {et} {gtype} {len(rows)}\n' + '\n'.join(rows))
        total = eid - 1
        w(f'$Elements\n{len(blocks)} {total} 1 {total}\n' + '\n'.join(blocks) + '\n$EndElements\n')
    with open(path, 'w') as f:
        f.write(''.join(out))


def boundary_faces(mesh, face_ids):
    """Node ids (Q, 4) of faces, in the owner cell's local face node order
    (a closed cycle, as a gmsh or VTK polygon needs), triangles padded with
    -1."""
    from .mesh import local_faces
    fp = mesh.face_properties
    f = np.asarray(face_ids, np.int64)
    own, k0 = fp.adjacent_cells[f, 0], fp.cell_face_index[f, 0]
    table = _pad4([r + [-1] * (4 - len(r)) for r in local_faces(mesh.cellType[1])])
    loc = table[k0]
    out = np.where(loc >= 0, mesh.cells[own[:, None], np.maximum(loc, 0)], -1)
    return out.astype(np.int32)


boundary_quads = boundary_faces


def face_groups_of(mesh):
    """{name: boundary faces} of every face group of ``mesh``."""
    return {name: boundary_faces(mesh, g.ids) for name, g in mesh.groups.items()
            if g.type == 'face'}


def apply_groups(mesh, data):
    """Add the physical groups of ``data`` (a MeshData) to ``mesh`` (a Mesh
    built from it): face groups (each triangle or quadrilateral must be a
    face of the mesh) and cell groups."""
    fp = mesh.face_properties
    keys = np.sort(_pad4(fp.unique_faces), axis=1)
    view = np.ascontiguousarray(keys).view([('', keys.dtype)] * 4).ravel()
    order = np.argsort(view)
    for name, (dim, ids) in data.groups.items():
        if dim == 3:
            mesh.add_group(name, ids, 'cell')
            continue
        q = np.sort(_pad4(ids).astype(keys.dtype), axis=1)
        qv = np.ascontiguousarray(q).view([('', keys.dtype)] * 4).ravel()
        p = np.searchsorted(view[order], qv)
        ok = (p < view.size)
        ok[ok] = view[order][p[ok]] == qv[ok]
        if not np.all(ok):
            raise ValueError(f'group {name!r}: {int((~ok).sum())} faces are no mesh face')
        mesh.add_group(name, order[p].astype(np.int32), 'face')
    return mesh


# ------------------------------------------------------------------- VTK
_VTK_TYPES = {'Float32': np.float32, 'Float64': np.float64, 'Int8': np.int8, 'UInt8': np.uint8,
              'Int16': np.int16, 'UInt16': np.uint16, 'Int32': np.int32, 'UInt32': np.uint32,
              'Int64': np.int64, 'UInt64': np.uint64}
_NP_TO_VTK = {np.dtype(v): k for k, v in _VTK_TYPES.items()}


def _decode_binary(b64, dtype, header, compressed):
    """Inline binary DataArray: base64 of a size header and the data.
    Uncompressed: one size word, encoded with the data or on its own;
    zlib-compressed: the block table (count, block size, last size,
    compressed sizes) encoded on its own, then the compressed blocks."""
    hdt = np.dtype(header)
    hs = hdt.itemsize
    if not compressed:
        blob = base64.b64decode(b64)
        n = int(np.frombuffer(blob[:hs], hdt)[0])
        if len(blob) - hs == n or (len(blob) - hs > n and len(blob) - hs - n < 3):
            return np.frombuffer(blob[hs:hs + n], dtype).copy()
        hchars = ((hs + 2) // 3) * 4
        n = int(np.frombuffer(base64.b64decode(b64[:hchars])[:hs], hdt)[0])
        return np.frombuffer(base64.b64decode(b64[hchars:])[:n], dtype).copy()
    nb = int(np.frombuffer(base64.b64decode(b64[:((hs + 2) // 3) * 4])[:hs], hdt)[0])
    hbytes = (3 + nb) * hs
    enc_hlen = ((hbytes + 2) // 3) * 4
    head = np.frombuffer(base64.b64decode(b64[:enc_hlen])[:hbytes], hdt)
    data = base64.b64decode(b64[enc_hlen:])
    out, pos = [], 0
    for size in head[3:3 + nb]:
        out.append(zlib.decompress(data[pos:pos + int(size)]))
        pos += int(size)
    return np.frombuffer(b''.join(out), dtype).copy()


def _decode_raw(buf, off, dtype, header, compressed):
    hdt = np.dtype(header)
    hs = hdt.itemsize
    if not compressed:
        n = int(np.frombuffer(buf[off:off + hs], hdt)[0])
        return np.frombuffer(buf[off + hs:off + hs + n], dtype).copy()
    nb = int(np.frombuffer(buf[off:off + hs], hdt)[0])
    head = np.frombuffer(buf[off:off + (3 + nb) * hs], hdt)
    pos = off + (3 + nb) * hs
    out = []
    for s in head[3:3 + nb]:
        out.append(zlib.decompress(buf[pos:pos + int(s)]))
        pos += int(s)
    return np.frombuffer(b''.join(out), dtype).copy()


def read_vtu(path):
    """VTK XML UnstructuredGrid of one cell type -- hexahedra (12),
    tetrahedra (10) or wedges (13): points, cells, and every CellData
    array."""
    with open(path, 'rb') as f:
        raw = f.read()
    appended = None
    m = re.search(rb'<AppendedData[^>]*>\s*_', raw)
    if m:
        enc = re.search(rb'encoding="(\w+)"', raw[m.start():m.end()])
        start = m.end()
        end = raw.rfind(b'</AppendedData>')
        body = raw[start:end]
        appended = (body, enc.group(1).decode() if enc else 'raw')
        raw = raw[:m.start()] + b'<AppendedData encoding="x"/></VTKFile>'
    root = ET.fromstring(raw)
    if root.get('type') != 'UnstructuredGrid':
        raise ValueError(f'{path}: not a VTK UnstructuredGrid file')
    header = {'UInt32': np.uint32, 'UInt64': np.uint64}[root.get('header_type', 'UInt32')]
    compressed = root.get('compressor') is not None
    if compressed and 'ZLib' not in root.get('compressor'):
        raise NotImplementedError(f'compressor {root.get("compressor")}')

    def read(da):
        dt = _VTK_TYPES[da.get('type')]
        fmt = da.get('format', 'ascii')
        if fmt == 'ascii':
            a = np.array((da.text or '').split(), dtype=np.float64 if dt in (np.float32, np.float64)
                         else np.int64).astype(dt)
        elif fmt == 'binary':
            a = _decode_binary((da.text or '').strip().encode(), dt, header, compressed)
        elif fmt == 'appended':
            body, enc = appended
            off = int(da.get('offset'))
            if enc == 'base64':
                raise NotImplementedError('base64-encoded appended data')
            a = _decode_raw(body, off, dt, header, compressed)
        else:
            raise ValueError(f'DataArray format {fmt}')
        nc = int(da.get('NumberOfComponents', '1'))
        return a.reshape(-1, nc) if nc > 1 else a

    piece = root.find('UnstructuredGrid/Piece')
    pts = read(piece.find('Points/DataArray'))
    arrays = {da.get('Name'): da for da in piece.find('Cells').findall('DataArray')}
    conn = read(arrays['connectivity']).astype(np.int64)
    offs = read(arrays['offsets']).astype(np.int64)
    types = read(arrays['types']).astype(np.int64)
    kinds = np.unique(types)
    if kinds.size > 1:
        raise ValueError('Meshes can only contain a single cell type for now, Got multiple')
    if kinds.size == 0 or int(kinds[0]) not in _NODES:
        raise ValueError('only tetrahedra (VTK type 10), hexahedra (12) and wedges (13) '
                         'are supported')
    ctype = int(kinds[0])
    nn = _NODES[ctype]
    if not np.array_equal(offs, nn * np.arange(1, len(types) + 1)):
        raise ValueError('inconsistent cell offsets')
    cell_data = {}
    cd = piece.find('CellData')
    if cd is not None:
        for da in cd.findall('DataArray'):
            cell_data[da.get('Name')] = read(da)
    return MeshData(pts, conn.reshape(-1, nn), cell_data=cell_data, cell_type=ctype)


def _encode_binary(a, compress):
    b = np.ascontiguousarray(a).tobytes()
    if not compress:
        return base64.b64encode(struct.pack('<Q', len(b)) + b).decode()
    block = 1 << 15
    chunks = [zlib.compress(b[i:i + block]) for i in range(0, max(len(b), 1), block)] or [b'']
    last = len(b) - block * (len(chunks) - 1) if b else 0
    head = struct.pack(f'<{3 + len(chunks)}Q', len(chunks), block, last, *[len(c) for c in chunks])
    return base64.b64encode(head).decode() + base64.b64encode(b''.join(chunks)).decode()


def write_vtu(path, nodes, cells, cell_data=None, binary=True, compress=True):
    """VTK XML UnstructuredGrid of hexahedra, tetrahedra or wedges (by the
    nodes per cell) with CellData arrays (one value or one vector per
    cell).  binary: inline base64 (UInt64 headers), zlib-compressed when
    ``compress``; else ascii."""
    nodes = np.asarray(nodes, np.float64).reshape(-1, 3)
    cells = np.asarray(cells, np.int64)
    ctype = _vtk_type(cells)
    nn = _NODES[ctype]
    C = len(cells)
    cell_data = cell_data or {}
    fmt = 'binary' if binary else 'ascii'

    def da(name, a, ncomp=None):
        a = np.asarray(a)
        nc = ncomp or (a.shape[1] if a.ndim > 1 else 1)
        attrs = f'type="{_NP_TO_VTK[a.dtype]}" Name="{name}" format="{fmt}"'
        if nc > 1:
            attrs += f' NumberOfComponents="{nc}"'
        body = _encode_binary(a, compress) if binary else ' '.join(
            repr(x) if isinstance(x, float) else str(x) for x in a.reshape(-1).tolist())
        return f'<DataArray {attrs}>\n{body}\n</DataArray>\n'

    comp = ' compressor="vtkZLibDataCompressor"' if (binary and compress) else ''
    out = [f'<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="1.0" '
           f'byte_order="LittleEndian" header_type="UInt64"{comp}>\n<UnstructuredGrid>\n'
           f'<Piece NumberOfPoints="{len(nodes)}" NumberOfCells="{C}">\n']
    out.append('<Points>\n' + da('Points', nodes, 3) + '</Points>\n')
    out.append('<Cells>\n' + da('connectivity', cells.reshape(-1))
               + da('offsets', nn * np.arange(1, C + 1, dtype=np.int64))
               + da('types', np.full(C, ctype, np.uint8)) + '</Cells>\n')
    if cell_data:
        out.append('<CellData>\n')
        for name, a in cell_data.items():
            a = np.asarray(a)
            if a.shape[0] != C:
                raise ValueError(f'cell data {name!r}: {a.shape[0]} values for {C} cells')
            if a.dtype == np.float16 or a.dtype.kind not in 'fiu':
                a = a.astype(np.float64)
            out.append(da(name, a))
        out.append('</CellData>\n')
    out.append('</Piece>\n</UnstructuredGrid>\n</VTKFile>\n')
    with open(path, 'w') as f:
        f.write(''.join(out))
