"""Device-resident mesh: the DeviceLayout arrays as torch tensors plus the
C-ABI descriptor (``wcfd_mesh``) that points at them."""
import ctypes

import numpy as np
import torch

from .. import _lib


def require_gpu(device):
    dev = torch.device(device)
    if dev.type != 'cuda' or not torch.cuda.is_available():
        raise _lib.WcfdError('the warpcfd_amd device path needs a ROCm GPU (torch.cuda); '
                             'there is no CPU fallback')
    return dev


class HostReadable(torch.Tensor):
    """A device tensor whose ``numpy()`` first copies to the host, as the
    reference's warp arrays do: its scripts read results with
    ``model.cell_values.numpy()`` (examples/LDC/LDC.py:54-55).  Everything
    else is an ordinary ``torch.Tensor``."""

    def numpy(self, *, force=False):
        return self.detach().cpu().as_subclass(torch.Tensor).numpy()


def host_readable(t):
    return t.as_subclass(HostReadable)


def stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class DeviceMesh:
    """Uploads a ``DeviceLayout`` and builds its ``wcfd_mesh`` descriptor.
    ``dtype`` (torch.float32 or torch.float64) is the field type: the
    geometry is uploaded in it and ``lib`` is the library built for it.  The
    layout's geometry is float32 either way (the reference computes the mesh
    in float32 numpy whatever its model's dtype, preprocess/mesh.py:22-33), so
    a float64 model sees the same geometry values as a float32 one."""

    def __init__(self, layout, device='cuda', dtype=torch.float32):
        self.device = require_gpu(device)
        self.layout = L = layout
        self.dtype = dtype
        self.lib = _lib.for_dtype(dtype)
        up = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a).astype(dt)).to(self.device)
        i32, i8 = np.int32, np.int8
        f32 = np.float64 if dtype == torch.float64 else np.float32      # the field type
        self.owner = up(L.owner, i32)
        self.neigh = up(L.neigh if L.neigh.size else np.zeros(1), i32)
        self.psi = up(L.psi if L.psi.size else np.zeros(1), f32)
        self.skew = up(L.skew if L.skew.size else np.zeros((3, 1)), f32)
        self.area = up(L.area, f32)
        self.normal = up(L.normal, f32)
        self.dist = up(L.dist if L.dist.size else np.zeros((3, 1)), f32)
        self.dmag = up(L.dmag, f32)
        self.volume = up(L.volume, f32)
        self.slot_face = up(L.slot_face, i32)
        self.slot_side = up(L.slot_side, i8)
        self.slot_sn = up(L.slot_sn, f32)
        self.slot_pos = up(L.slot_pos, i32)
        self.diag_pos = up(L.diag_pos, i32)
        self.row_ptr = up(L.row_ptr, i32)
        self.col = up(L.col, i32)
        self.slot_nbr = up(L.slot_nbr, i32)
        self.slot_dnb = up(L.slot_dnb, np.int16) if L.slot_dnb is not None else None
        self.old2new = torch.from_numpy(L.old2new).to(self.device)
        self.new2old = torch.from_numpy(L.new2old).to(self.device)
        self.face_old2new = torch.from_numpy(L.face_old2new).to(self.device)
        self.face_new2old = torch.from_numpy(L.face_new2old).to(self.device)
        self.n_cells, self.n_faces = L.num_cells, L.num_faces
        self.n_internal, self.n_boundary, self.nnz = L.n_internal, L.n_boundary, L.nnz
        d = _lib.WcfdMesh()
        d.n_cells, d.n_faces, d.n_internal = L.num_cells, L.num_faces, L.n_internal
        d.n_boundary, d.nnz, d.n_slots = L.n_boundary, L.nnz, L.faces_per_cell
        for name in ('owner', 'neigh', 'psi', 'skew', 'area', 'normal', 'dist', 'dmag', 'volume',
                     'slot_face', 'slot_side', 'slot_sn', 'slot_pos', 'diag_pos', 'row_ptr', 'col',
                     'slot_nbr'):
            setattr(d, name, getattr(self, name).data_ptr())
        d.slot_dnb = self.slot_dnb.data_ptr() if self.slot_dnb is not None else None
        self.desc = d
        self.desc_ref = ctypes.byref(d)

    # reference order <-> device order
    def cells_to_reference(self, t):
        """(..., C) device order -> (..., C) reference order."""
        return t.index_select(-1, self.old2new)

    def cells_to_device(self, t):
        return t.index_select(-1, self.new2old)

    def faces_to_reference(self, t):
        return t.index_select(-1, self.face_old2new)

    def faces_to_device(self, t):
        """(..., K) reference order -> (..., F) device order; the null face
        of a tetrahedron or wedge mesh gets 0."""
        if self.layout.null_face >= 0:
            t = torch.cat([t, t.new_zeros(t.shape[:-1] + (1,))], dim=-1)
        return t.index_select(-1, self.face_new2old)
