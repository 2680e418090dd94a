"""Boundary conditions per boundary face and field.

Mirrors warp_cfd/FV/boundary/boundary.py:5-121: values and types are kept on
the host in the reference's order (rows follow ``boundary_face_ids``) until
``finalize``; ``device_arrays`` then reorders them into the device layout,
[4][Fb] row-major, for the kernels.
"""
import numpy as np
import torch

DIRICHLET = 1
NEUMANN = 2


class Boundary:
    def __init__(self, fields, groups, boundary_ids, float_type=np.float32):
        self.fields = fields
        self.num_outputs = len(fields)
        self.groups = groups
        self.boundary_ids = np.asarray(boundary_ids, np.int32)
        # float64 on the host (exact for the Python floats given); cast to the
        # model's field type when uploaded
        self.boundary_values = np.zeros((len(self.boundary_ids), self.num_outputs), np.float64)
        self.boundary_type = np.zeros(self.boundary_values.shape, np.uint8)
        self.float_dtype = float_type
        self._boundary_types = {'dirichlet': DIRICHLET, 'vonNeumann': NEUMANN}
        self.finalized = False

    def set_value(self, face_ids, boundary_type='dirchlet', **output_vars):
        if self.finalized:
            raise RuntimeError('boundary conditions are fixed once the model is finalized')
        face_ids = self.get_face_id_array(face_ids)
        inside = ~np.isin(face_ids, self.boundary_ids)
        assert not inside.any(), \
            f'faces {face_ids[inside][:10].tolist()} are not on the mesh boundary'
        assert output_vars, 'no field given: name at least one of u, v, w, p as a keyword'
        rows = np.where(np.isin(self.boundary_ids, face_ids))[0]
        for name, value in output_vars.items():
            j = self.fields[name].index
            self.boundary_values[rows, j] = value
            self.boundary_type[rows, j] = self._boundary_types[boundary_type]

    def get_face_id_array(self, face_ids):
        if isinstance(face_ids, str):
            assert face_ids in self.groups, \
                f'no face group {face_ids!r}; the mesh has {sorted(self.groups)}'
            g = self.groups[face_ids]
            return np.asarray(getattr(g, 'ids', g))
        if isinstance(face_ids, (int, np.integer, list, tuple)):
            return np.atleast_1d(np.asarray(face_ids, np.int32))
        if isinstance(face_ids, np.ndarray):
            return face_ids
        raise ValueError('face_ids must be a group name, an int, a list or tuple of ints or an '
                         f'integer array, not {type(face_ids).__name__}')

    def gradient_BC(self, face_ids, **output_vars):
        self.set_value(face_ids, boundary_type='vonNeumann', **output_vars)

    def fixed_BC(self, face_ids, **output_vars):
        self.set_value(face_ids, boundary_type='dirichlet', **output_vars)

    def velocity_BC(self, face_ids, u, v, w):
        self.fixed_BC(face_ids, u=u, v=v, w=w)
        self.gradient_BC(face_ids, p=0.)

    def pressure_BC(self, face_ids, p, backflow=None):
        self.fixed_BC(face_ids, p=p)
        self.gradient_BC(face_ids, u=0., v=0., w=0.)

    def mass_flow_BC(self, face_ids, mass_flow, inlet=True):
        raise ValueError('not yet implemented')

    def no_slip_wall(self, face_ids):
        self.fixed_BC(face_ids, u=0., v=0., w=0.)
        self.gradient_BC(face_ids, p=0.)

    def slip_wall(self, face_ids):
        self.gradient_BC(face_ids, u=0., v=0., w=0., p=0.)

    def finalize(self):
        """boundary.py:108-121 checks; the arrays stay on the host."""
        if np.any(self.boundary_type < 1):
            bad = self.boundary_ids[(self.boundary_type < 1).any(axis=1)]
            raise ValueError(f'{len(bad)} boundary faces lack a condition for some field, '
                             f'e.g. faces {bad[:10].tolist()}')
        if np.any(self.boundary_type > 3):
            raise ValueError('a boundary type is neither 1 (Dirichlet) nor 2 (Neumann)')
        self.finalized = True

    def uniform_velocity_types(self):
        """True if u, v and w share the BC type on every face, which lets the
        three momentum components share one matrix."""
        t = self.boundary_type
        return bool(np.all(t[:, 0] == t[:, 1]) and np.all(t[:, 0] == t[:, 2]))

    def host_arrays(self, layout, dtype=np.float32):
        """BC type (uint8) and value (``dtype``) as [4][Fb] in device face
        order."""
        rows = layout.bnd_ref_row
        t = np.ascontiguousarray(self.boundary_type[rows].T).astype(np.uint8)
        v = np.ascontiguousarray(self.boundary_values[rows].T.astype(dtype))
        return t, v

    def device_arrays(self, layout, device, dtype=torch.float32):
        np_dt = np.float64 if dtype == torch.float64 else np.float32
        t, v = self.host_arrays(layout, np_dt)
        if t.size == 0:
            t = np.zeros((4, 1), np.uint8)
            v = np.zeros((4, 1), np_dt)
        return torch.from_numpy(t).to(device), torch.from_numpy(v).to(device)
