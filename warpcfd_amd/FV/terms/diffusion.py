"""DiffusionTerm (warp_cfd/FV/terms/diffusion.py:8-197).

Central-difference Laplacian (laplacianScheme.py:14-37), or with
``correction=True`` the over-relaxed non-orthogonal form (39-66), whose
explicit part of internal faces never reaches the right-hand side (T2).
Boundary faces follow each output's condition (diffusion.py:113-152); the
``von_neumann=`` option puts nu * value * A_f on every boundary face of the
single field, and so does ``dirchlet=``, which the reference builds with the
same kernel (diffusion.py:33).
"""
import numpy as np
import torch

from ... import _lib
from ..device import ptr, stream_ptr
from .terms import Term


class DiffusionTerm(Term):
    def __init__(self, fv, fields, *, correction=False, custom_interpolation=None,
                 von_neumann=None, dirchlet=None, need_global_index=True):
        super().__init__(fv, fields, implicit=True, need_global_index=need_global_index,
                         cell_based=False)
        if correction:
            assert need_global_index, 'Fields Must be defined in the FVM model to use corrections'
        self.correction = bool(correction)
        assert custom_interpolation is None, 'Custom laplacian functions not available yet'
        assert not ((von_neumann is not None) and (dirchlet is not None)), \
            'von neumann and dirichlet settings cannot both be not None'
        self.bc_mode, self.bc_value = 0, 0.0
        for v in (von_neumann, dirchlet):
            if v is not None:
                assert len(self.fields) == 1, \
                    'Von neumann or dirichlet option only valid if len of field == 1'
                self.bc_mode, self.bc_value = 1, float(v)

    def calculate_weights(self, fv, viscosity, **kwargs):
        """diffusion.py:40-49.  viscosity: a float, a length-1 array, or one
        value per face in the reference's face order."""
        nu_faces = None
        nu = 0.0
        if isinstance(viscosity, (float, int, np.floating)):
            nu = float(viscosity)
        else:
            v = torch.as_tensor(np.asarray(viscosity) if not torch.is_tensor(viscosity)
                                else viscosity, dtype=fv.tdtype).reshape(-1)
            if v.numel() == 1:
                nu = float(v.item())
            elif v.numel() == fv.num_faces:
                nu_faces = fv.dm.faces_to_device(v.to(fv._cv.device)).contiguous()
            else:
                raise ValueError('viscosity: a float or one value per face')
        g, gp = self._outputs()
        code = fv.lib.wcfd_diffusion_weights(
            fv.dm.desc_ref, fv.bc, ptr(fv._cg), ptr(fv._fv),
            ptr(nu_faces) if nu_faces is not None else None, nu, int(self.correction),
            self.bc_mode, self.bc_value, gp, self.num_outputs, ptr(self._w), stream_ptr())
        _lib.check(code, 'diffusion weights')
