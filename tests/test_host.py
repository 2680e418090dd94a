"""Host-side behaviour of the API mirror (no GPU needed): boundary
bookkeeping and the errors the reference raises."""
import numpy as np
import pytest

from warpcfd_amd.preprocess import Mesh, cavity_mesh, define_boundary_walls
from warpcfd_amd.FV import FVM
from warpcfd_amd import _lib


def _model():
    nodes, cells = cavity_mesh(6)
    m = Mesh(nodes, cells)
    define_boundary_walls(m)
    return FVM(m, device='cpu')


def test_boundary_helpers_set_types_like_the_reference():
    model = _model()
    b = model.boundary
    b.no_slip_wall('-X'); b.no_slip_wall('+X'); b.no_slip_wall('-Y')
    b.velocity_BC('+Y', u=1.0, v=0.0, w=0.0)
    b.slip_wall('+Z'); b.slip_wall('-Z')
    g = model.mesh.groups
    rows = np.isin(b.boundary_ids, g['+Y'].ids)
    assert np.all(b.boundary_type[rows] == [1, 1, 1, 2])
    assert np.all(b.boundary_values[rows] == [1, 0, 0, 0])
    rows = np.isin(b.boundary_ids, g['+Z'].ids)
    assert np.all(b.boundary_type[rows] == [2, 2, 2, 2])       # T12
    b.finalize()
    assert b.uniform_velocity_types()


def test_unassigned_boundary_raises():
    model = _model()
    model.boundary.no_slip_wall('-X')
    with pytest.raises(ValueError):
        model.boundary.finalize()


def test_pressure_bc_on_internal_face_rejected():
    model = _model()
    internal = model.face_properties.internal_face_ids[:3]
    with pytest.raises(AssertionError):
        model.boundary.pressure_BC(internal, p=0.0)


def test_device_path_has_no_cpu_fallback():
    model = _model()
    b = model.boundary
    for w in ('-X', '+X', '-Y', '+Y', '+Z', '-Z'):
        b.no_slip_wall(w)
    with pytest.raises(_lib.WcfdError):
        model.finalize()


def test_solver_needs_finalized_model():
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    with pytest.raises(ValueError):
        IncompressibleSolver(_model(), 0.7, 0.3)


def _faces_in_slots(rng, n_faces, nodes_per_face, n_slots):
    """n_faces distinct node sets, each put in one or two random slots of an
    (n_slots, nodes_per_face) table with its nodes in a random order; the
    remaining slots get faces of their own."""
    seen, faces = set(), []
    while len(faces) < n_faces + n_slots:
        f = tuple(sorted(rng.choice(60, nodes_per_face, replace=False).tolist()))
        if f not in seen:
            seen.add(f)
            faces.append(f)
    table = np.empty((n_slots, nodes_per_face), np.int64)
    slots = rng.permutation(n_slots)
    j = 0
    for i in range(n_faces):
        for _ in range(1 + (i % 2)):
            if j < n_slots:
                table[slots[j]] = rng.permutation(faces[i])
                j += 1
    for i in range(j, n_slots):
        table[slots[i]] = rng.permutation(faces[n_faces + i])
    return table


@pytest.mark.parametrize('nodes_per_face', [3, 4])
def test_match_faces_numbers_like_np_unique(nodes_per_face):
    """wcfd_match_faces against np.unique over the sorted node ids (the
    reference's numbering, mesh.py:108-148) and a stable sort of the slots
    by face id (the owner/neighbour pairs, mesh.py:263-288)."""
    from warpcfd_amd.preprocess.mesh import match_faces
    rng = np.random.default_rng(nodes_per_face)
    C, Fs = 40, 6
    local = _faces_in_slots(rng, 150, nodes_per_face, C * Fs).reshape(C, Fs, nodes_per_face)
    uq, face_of, slots = match_faces(local)
    ref_uq, inv = np.unique(np.sort(local, axis=2).reshape(-1, nodes_per_face), axis=0,
                            return_inverse=True)
    assert np.array_equal(uq, ref_uq)
    assert np.array_equal(face_of.ravel(), inv.ravel())
    order = np.argsort(inv.ravel(), kind='stable')
    counts = np.bincount(inv.ravel())
    first = np.concatenate([[0], np.cumsum(counts)[:-1]])
    assert np.array_equal(slots[:, 0], order[first])
    two = counts == 2
    assert 0 < two.sum() < len(counts)
    assert np.array_equal(slots[two, 1], order[first[two] + 1])
    assert (slots[~two, 1] == -1).all()


def test_match_faces_errors():
    from warpcfd_amd.preprocess.mesh import match_faces
    quad = np.array([0, 1, 2, 3])
    three = np.stack([quad, quad[::-1], np.roll(quad, 1)])[:, None, :]   # one face in 3 slots
    with pytest.raises(AssertionError, match='more than two'):
        match_faces(three)
    with pytest.raises(_lib.WcfdError, match='invalid argument'):
        match_faces(-np.ones((2, 1, 4), np.int64))
    lib = _lib.load()
    local = np.arange(8, dtype=np.int32).reshape(2, 1, 4)
    out = np.empty(16, np.int32)
    assert lib.wcfd_match_faces(local.ctypes.data, 2, 1, 4, out.ctypes.data, out.ctypes.data,
                                out.ctypes.data, 1) == -_lib.WCFD_ECAPACITY
