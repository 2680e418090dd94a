"""Result export: cell fields to a .vtu file, and sampling along a line.

The reference's examples attach the solution to the pyvista grid and sample
it along the cavity's centrelines with ``pv.Line(...).sample(grid)``, then
compare with Ghia et al. (examples/LDC/LDC.py:53-103); pyvista is absent
here.  ``export_vtu`` writes the mesh with u, v, w, p and the velocity
vector as cell data (ParaView reads it); ``sample_line`` returns the value
of the cell that contains each of n evenly spaced points of a segment, as a
probe of cell data does (points on a face shared by two cells take the
lower cell id; points outside the mesh give NaN).
"""
import numpy as np

from .preprocess.io import write_vtu


def export_vtu(model, path, extra=None, binary=True, compress=True):
    """Write ``model``'s mesh (input node and cell order) with its cell
    values: u, v, w, p, 'velocity' (3 components) and any ``extra`` {name:
    per-cell array in input cell order}."""
    cv = model.cell_values
    cv = cv.cpu().numpy() if hasattr(cv, 'cpu') else np.asarray(cv)
    data = {name: cv[:, i].astype(np.float32) for i, name in enumerate(model.output_variables)}
    data['velocity'] = cv[:, :3].astype(np.float32)
    for k, v in (extra or {}).items():
        data[k] = v.cpu().numpy() if hasattr(v, 'cpu') else np.asarray(v)
    write_vtu(path, model.mesh.nodes, model.mesh.cells, data, binary=binary, compress=compress)


def locate_cells(mesh, points, k=8, tol=1e-9):
    """Index of the cell containing each point (-1 outside): the k nearest
    centroids are tested against their cells' six face planes (outward
    normals of the cell, mesh.py's fan normals), lowest id first."""
    from scipy.spatial import cKDTree
    cp, fp = mesh.cell_properties, mesh.face_properties
    cen = np.asarray(cp.centroids, np.float64)
    pts = np.atleast_2d(np.asarray(points, np.float64))
    k = min(k, cen.shape[0])
    _, cand = cKDTree(cen).query(pts, k=k)
    cand = np.sort(cand.reshape(len(pts), k), axis=1)
    nrm = np.asarray(cp.normal, np.float64)                    # (C, 6, 3)
    fcen = np.asarray(fp.centroid, np.float64)[cp.faces]       # (C, 6, 3)
    scale = np.cbrt(np.abs(np.asarray(cp.volumes, np.float64)))
    out = np.full(len(pts), -1, np.int64)
    for j in range(k):
        c = cand[:, j]
        s = np.einsum('pkd,pkd->pk', pts[:, None, :] - fcen[c], nrm[c])
        inside = np.all(s <= tol * scale[c][:, None], axis=1) & (out < 0)
        out[inside] = c[inside]
    return out


def sample_line(mesh, values, start, end, n_points):
    """Values of the cells containing n_points evenly spaced points from
    ``start`` to ``end``.  values: (C,) or (C, k) per input cell, or a dict
    of them.  Returns (points (n, 3), sampled values, cell ids)."""
    t = np.linspace(0.0, 1.0, n_points)
    a, b = np.asarray(start, np.float64), np.asarray(end, np.float64)
    pts = a[None, :] + t[:, None] * (b - a)[None, :]
    cells = locate_cells(mesh, pts)
    def take(v):
        v = np.asarray(v.cpu().numpy() if hasattr(v, 'cpu') else v, np.float64)
        out = v[np.maximum(cells, 0)]
        out[cells < 0] = np.nan
        return out
    if isinstance(values, dict):
        return pts, {k: take(v) for k, v in values.items()}, cells
    return pts, take(values), cells
