"""DdtTerm (warp_cfd/FV/terms/ddt.py:7-33): backward Euler, a cell-based
implicit term with weights (V/dt, -phi V/dt) per output."""
from ... import _lib
from ..device import ptr, stream_ptr
from .terms import Term


class DdtTerm(Term):
    def __init__(self, fv, fields, time_scheme='backwardEuler'):
        super().__init__(fv, fields, implicit=True, need_global_index=True, cell_based=True)
        if time_scheme != 'backwardEuler':
            raise NotImplementedError('Only backwardEuler valid option now')
        self.time_scheme = time_scheme

    def calculate_weights(self, fv, dt, **kwargs):
        g, gp = self._outputs()
        code = fv.lib.wcfd_ddt_weights(fv.dm.desc_ref, ptr(fv._cv), float(dt), gp,
                                       self.num_outputs, ptr(self._w), stream_ptr())
        _lib.check(code, 'ddt weights')
