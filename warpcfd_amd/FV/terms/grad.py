"""GradTerm (warp_cfd/FV/terms/grad.py:7-28): the explicit term
(grad phi)_c V_c of one field, one row per cell and direction."""
from ... import _lib
from ..device import ptr, stream_ptr
from .terms import Term


class GradTerm(Term):
    def __init__(self, fv, field):
        assert isinstance(field, str), 'grad Term can only be in terms of 1 field string name only'
        super().__init__(fv, field, implicit=False, need_global_index=True, cell_based=False)
        # one row per direction: the explicit weights have fv.dimension outputs
        import torch
        self._w = torch.zeros((fv.dimension, fv.num_cells), dtype=fv.tdtype,
                              device=fv._cv.device)

    def calculate_weights(self, fv, **kwargs):
        code = fv.lib.wcfd_grad_weights(fv.dm.desc_ref, ptr(fv._cg), int(self.fields[0].index),
                                        ptr(self._w), stream_ptr())
        _lib.check(code, 'grad weights')
