"""ConvectionTerm (warp_cfd/FV/terms/convection.py:8-88).

``interpolation``: 'upwind' (convectionScheme.py:28-52), 'central difference'
(3-26) or 'upwindLinear' (54-88); each internal face also gets the skew
correction as explicit part, which, as in the reference, never reaches the
right-hand side (T2).  A 'custom' interpolation (a Warp function) has no
counterpart here.
"""
import numpy as np
import torch

from ... import _lib
from ..device import ptr, stream_ptr
from .terms import Term

SCHEMES = {'upwind': _lib.CONV_UPWIND, 'central difference': _lib.CONV_CENTRAL,
           'upwindLinear': _lib.CONV_UPWIND_LINEAR}


def face_c2f(fv):
    """(2, 3, Fi) per internal face in device order: x_f - x_owner, and the
    owner's centroid-to-face vector at the neighbour's local face index (what
    upwindLinear reads when the neighbour is upwind, convectionScheme.py:82).
    Built on the host once per model."""
    if getattr(fv, '_face_c2f', None) is None:
        L = fv.layout
        fp, cp = fv.mesh.face_properties, fv.mesh.cell_properties
        fi_ref = L.face_new2old[:L.n_internal]
        real = fi_ref < fp.num_faces                 # not the null face
        own = fp.adjacent_cells[fi_ref[real], 0]
        k0, k1 = fp.cell_face_index[fi_ref[real], 0], fp.cell_face_index[fi_ref[real], 1]
        c2f = cp.cell_centroid_to_face_centroid
        a = np.zeros((2, 3, L.n_internal), fv.float_dtype)
        a[:, :, real] = np.stack([c2f[own, k0].T, c2f[own, k1].T])
        if a.size == 0:
            a = np.zeros((2, 3, 1), fv.float_dtype)
        fv._face_c2f = torch.from_numpy(np.ascontiguousarray(a)).to(fv._cv.device)
    return fv._face_c2f


class ConvectionTerm(Term):
    def __init__(self, fv, fields, interpolation='upwind', custom_interpolation=None):
        super().__init__(fv, fields, implicit=True, need_global_index=True, cell_based=False)
        self.correct_face_interpolation = True
        self.interpolation = interpolation
        allowed = list(SCHEMES) + ['custom']
        assert interpolation in allowed, \
            'interpolation schems supported: upwind, central difference, upwindLinear or custom'
        if interpolation == 'custom':
            raise NotImplementedError('custom interpolation functions are Warp functions; '
                                      'use upwind, central difference or upwindLinear')
        self.scheme = SCHEMES[interpolation]
        self.implicit_outflow = False

    def calculate_weights(self, fv, implicit_outflow=None, **kwargs):
        """convection.py:33-41, from the model's cell gradients, face values
        and mass fluxes.  implicit_outflow (not in the reference): implicit
        upwind flux on Neumann outflow faces."""
        io = self.implicit_outflow if implicit_outflow is None else implicit_outflow
        g, gp = self._outputs()
        c2f = face_c2f(fv) if self.scheme == _lib.CONV_UPWIND_LINEAR else None
        code = fv.lib.wcfd_convection_weights(
            fv.dm.desc_ref, fv.bc, ptr(fv._cg), ptr(fv._fv), ptr(fv._flux),
            ptr(c2f) if c2f is not None else None, self.scheme,
            _lib.SCHEME_IMPLICIT_OUTFLOW if io else 0, gp, self.num_outputs, ptr(self._w),
            stream_ptr())
        _lib.check(code, 'convection weights')
