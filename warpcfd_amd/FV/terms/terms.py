"""Term: base of the composable finite-volume terms (warp_cfd/FV/terms/terms.py:8-120).

A term holds its weights on the device, computed by a HIP kernel of
libwcfd.so from the model's current state (``term(model, ...)``), and a
scale (+1, -1 or any float through unary minus and multiplication), as the
reference's terms do.  ``Equation.form_system`` accumulates them.

Device layouts (device cell order):
* face-based implicit terms: ``_w`` (O, 3, 6, C) -- output, part (owner
  coefficient, neighbour coefficient, explicit), local face, cell;
* cell-based implicit terms: ``_w`` (O, 2, C);
* explicit terms: ``_w`` (O, C).
``weights`` returns the reference's layout and cell order: (C, 6, O, 3),
(C, O, 2), or the interleaved (C * O,) vector.
"""
import copy
import ctypes

import numpy as np
import torch

from ..field import Field


class Term:
    def __init__(self, fv, fields, implicit, need_global_index, cell_based):
        if not fv.finalized:
            raise ValueError('Please call method model.finalize() first')
        self.fields = []
        if need_global_index:
            if isinstance(fields, str):
                fields = [fields]
            for name in fields:
                if name not in fv.output_variables:
                    raise ValueError(f'Field {name} was not found in original model only variables '
                                     f'{fv.output_variables} are available')
                self.fields.append(fv.fields[name])
            self.global_output_indices = np.array([f.index for f in self.fields], np.int32)
        else:
            assert isinstance(fields, str), \
                'Only one field variable can be defined if global index is False'
            self.fields.append(Field(fields, index=-1))
            # terms.py:37: the kernels then read field 0's data and conditions
            self.global_output_indices = np.zeros(1, np.int32)
        self.need_global_index = need_global_index
        self.local_field_index = {f.name: i for i, f in enumerate(self.fields)}
        self.num_outputs = len(self.fields)
        self.set_implicit(implicit)
        self.cell_based = cell_based
        self.float_dtype = fv.float_dtype
        self.int_dtype = np.int32
        self._scale = 1.0
        self._fv = fv
        C = fv.num_cells
        dev = fv._cv.device
        O = self.num_outputs
        if implicit:
            shape = (O, 2, C) if cell_based else (O, 3, 6, C)
        else:
            shape = (O, C)
        self._w = torch.zeros(shape, dtype=fv.tdtype, device=dev)

    # ------------------------------------------------------------- scale
    def set_implicit(self, implicit):
        self._implicit = bool(implicit)

    def set_scale(self, value):
        self._scale = float(np.float32(value))

    @property
    def scale(self):
        return self._scale

    @property
    def implicit(self):
        """Whether the term has implicit components."""
        return self._implicit

    def __pos__(self):
        self.set_scale(1.0)
        return self

    def __neg__(self):
        t = copy.copy(self)
        t.set_scale(-1.0)
        return t

    def __mul__(self, value):
        if isinstance(value, (float, int)):
            t = copy.copy(self)
            t.set_scale(value)
            return t
        return NotImplemented

    def __rmul__(self, other):
        return self.__mul__(other)

    def __call__(self, fv, *args, **kwargs):
        self.calculate_weights(fv, *args, **kwargs)

    def calculate_weights(self, fv, *args, **kwargs):
        pass

    # ----------------------------------------------------- reference views
    @property
    def weights(self):
        """The weights in the reference's layout and cell order (a copy)."""
        dm = self._fv.dm
        w = self._w
        if not self.implicit:
            return dm.cells_to_reference(w).T.reshape(-1).contiguous()
        if self.cell_based:                              # (O, 2, C) -> (C, O, 2)
            return dm.cells_to_reference(w).permute(2, 0, 1).contiguous()
        w = dm.cells_to_reference(w).permute(3, 2, 0, 1)                  # (C, 6, O, 3)
        return w[:, :self._fv.faces_per_cell].contiguous()   # the mesh's faces per cell

    def _outputs(self):
        """(array, pointer) of the global field indices; keep the array alive
        for the call."""
        return int32_array(self.global_output_indices)


def int32_array(values):
    g = np.ascontiguousarray(values, np.int32)
    return g, g.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
