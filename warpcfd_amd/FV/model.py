"""FVM: the finite-volume model, device-resident.

Mirrors warp_cfd/FV/model.py:51-384.  The state lives on the GPU in the
device layout (structure-of-arrays, RCM-renumbered cells, face-ordered);
every operator is a HIP kernel of libwcfd.so reached through the C ABI.
Properties such as ``cell_values`` return tensors in the *reference* order
(cell i = the i-th cell of the input mesh), as the reference's arrays are.
"""
import ctypes

import numpy as np
import torch

from .. import _lib
from ..material import Material
from ..preprocess.layout import DeviceLayout
from .boundary import Boundary
from .convergence import Convergence
from .device import DeviceMesh, host_readable, ptr, require_gpu, stream_ptr
from .field import Field


class FVM:
    def __init__(self, mesh, output_variables=None, float_dtype=np.float32, int_dtype=np.int32,
                 device='cuda', renumber=True):
        # model.py:54-55: float32 (libwcfd.so) or float64 (libwcfd64.so); the
        # reductions are float64 either way
        if float_dtype in (torch.float32, torch.float64):
            float_dtype = np.float32 if float_dtype == torch.float32 else np.float64
        if np.dtype(float_dtype) not in (np.float32, np.float64):
            raise ValueError(f'float_dtype must be float32 or float64, got {float_dtype}')
        self.material = Material()
        self.dimension = mesh.dimension
        self.float_dtype = np.dtype(float_dtype).type
        self.tdtype = torch.float64 if self.float_dtype == np.float64 else torch.float32
        self.int_dtype = np.int32
        self.mesh = mesh
        self.skew_correction = True
        self.face_properties = mesh.face_properties
        self.cell_properties = mesh.cell_properties
        self.input_variables = ['x', 'y', 'z', 't']
        self.output_variables = output_variables or ['u', 'v', 'w', 'p']
        if list(self.output_variables) != ['u', 'v', 'w', 'p']:
            raise NotImplementedError("output_variables must be ['u','v','w','p']")
        self.fields = {o: Field(o, i) for i, o in enumerate(self.output_variables)}
        self.boundary = Boundary(self.fields, mesh.groups,
                                 boundary_ids=self.face_properties.boundary_face_ids)
        self.vars_mapping = {v: i for i, v in enumerate(self.output_variables)}
        self.num_outputs = 4
        self.num_nodes = mesh.nodes.shape[0]
        self.num_cells = mesh.cells.shape[0]
        self.num_faces = self.face_properties.num_faces
        self.nodes_per_cell = mesh.cells.shape[1]
        self.faces_per_cell = self.cell_properties.faces_per_cell
        self.nodes_per_face = self.cell_properties.nodes_per_face
        self.reference_pressure_cell_id = None
        self.reference_pressure = 0.0
        self.finalized = False
        self._device = device
        self._renumber = renumber
        # u,v,w gradients are known to be zero (after calculate_gradients('p'))
        self._uvw_grad_zero = True
        self.Convergence = Convergence()
        self._conv_ws = None

    # ------------------------------------------------------------ set-up
    def finalize(self, layout=None):
        """model.py:121-139: upload the mesh, check and upload the BCs,
        allocate the field arrays.  ``layout``: a ``DeviceLayout`` of this
        mesh built earlier (not in the reference; saves renumbering twice)."""
        if self.finalized:
            return
        dev = require_gpu(self._device)
        self.boundary.finalize()
        self.layout = layout if layout is not None else DeviceLayout(self.mesh,
                                                                     renumber=self._renumber)
        self.dm = DeviceMesh(self.layout, dev, dtype=self.tdtype)
        self.bc_type, self.bc_value = self.boundary.device_arrays(self.layout, dev, self.tdtype)
        self.bc = _lib.WcfdBC(self.bc_type.data_ptr(), self.bc_value.data_ptr())
        self.lib = self.dm.lib
        self.init_global_arrays()
        self.finalized = True

    def init_global_arrays(self):
        """model.py:155-176: zeroed cell values and gradients, face values
        and mass fluxes, and the reduction workspace (field-major, device
        cell and face order; face gradients are computed on demand)."""
        dev = self.dm.device
        C, F = self.num_cells, self.dm.n_faces        # device faces (with a null face)
        z = lambda *s: torch.zeros(*s, dtype=self.tdtype, device=dev)
        self._cv = z(4, C)
        self._cg = z(12, C)
        self._fv = z(4, F)
        self._flux = z(F)
        self._reduce_ws = torch.zeros(self.lib.wcfd_reduce_workspace_bytes(F), dtype=torch.uint8,
                                      device=dev)

    def _need(self):
        if not self.finalized:
            raise ValueError('Please call method model.finalize() first')

    def set_initial_conditions(self, IC, output_indices=None):
        """model.py:143-147 (IC in reference cell order)."""
        self._need()
        ic = torch.as_tensor(np.asarray(IC) if not torch.is_tensor(IC) else IC,
                             dtype=self.tdtype, device=self._cv.device)
        if output_indices is None:
            output_indices = list(range(ic.shape[-1]))
        for o in self._indices(output_indices):
            self._cv[o] = self.dm.cells_to_device(ic[:, o].contiguous())
        self._uvw_grad_zero = bool(torch.all(self._cg[:9] == 0))

    def set_reference_pressure(self, cell_id, value=0.):
        self.reference_pressure_cell_id = cell_id
        self.reference_pressure = float(value)

    @property
    def reference_cell_device(self):
        if self.reference_pressure_cell_id is None:
            return -1
        return int(self.layout.old2new[self.reference_pressure_cell_id])

    def _indices(self, output_indices):
        if output_indices is None:
            return [0, 1, 2, 3]
        if isinstance(output_indices, (int, np.integer)):
            return [int(output_indices)]
        if isinstance(output_indices, str):
            assert output_indices in self.output_variables
            return [self.fields[output_indices].index]
        out = []
        for o in output_indices:
            out.append(self.fields[o].index if isinstance(o, str) else int(o))
        return out

    def get_output_indices(self, output_indices):
        return np.asarray(self._indices(output_indices), np.int32)

    @staticmethod
    def _mask(idx):
        m = 0
        for o in idx:
            m |= 1 << o
        return m

    def _s(self):
        return stream_ptr()

    # ------------------------------------------------ reference-order views
    @property
    def cell_values(self):
        """(C,4) tensor of the model's float type in reference cell order (a copy)."""
        self._need()
        return host_readable(self.dm.cells_to_reference(self._cv).T.contiguous())

    @property
    def cell_gradients(self):
        """(C,4,3) in reference cell order (a copy)."""
        self._need()
        return host_readable(
            self.dm.cells_to_reference(self._cg).T.reshape(self.num_cells, 4, 3).contiguous())

    @property
    def face_values(self):
        self._need()
        return host_readable(self.dm.faces_to_reference(self._fv).T.contiguous())

    @property
    def mass_fluxes(self):
        self._need()
        return host_readable(self.dm.faces_to_reference(self._flux).contiguous())

    @property
    def face_gradients(self):
        """(F,4): the Neumann boundary gradients (the only entries the
        reference ever writes, conditions.py:17-18)."""
        fg = np.zeros((self.num_faces, 4), self.float_dtype)
        b = self.boundary
        n = b.boundary_type == 2
        rows, cols = np.nonzero(n)
        fg[b.boundary_ids[rows], cols] = b.boundary_values[rows, cols]
        return torch.from_numpy(fg)

    # ----------------------------------------------------------- operators
    def set_boundary_conditions(self):
        """model.py:206-208."""
        self._need()
        _lib.check(self.lib.wcfd_set_boundary_conditions(self.dm.desc_ref, self.bc, ptr(self._fv),
                                                         15, self._s()), 'set_boundary_conditions')

    def face_interpolation(self, output_indices=None, upwind=False):
        """model.py:210-229: linear + skew correction, or with ``upwind`` the
        upwind cell's value by the sign of the model's mass fluxes
        (interpolation_functions.py:18-28) + the same skew correction."""
        self._need()
        idx = self._indices(output_indices)
        mask = self._mask(idx)
        skew = mask if self.skew_correction else 0
        if self._uvw_grad_zero:
            skew &= 8
        if upwind:
            _lib.check(self.lib.wcfd_face_interpolation_upwind(
                self.dm.desc_ref, self.bc, ptr(self._cv), ptr(self._cg), ptr(self._flux),
                ptr(self._fv), mask, skew, self._s()), 'face_interpolation(upwind=True)')
            return
        _lib.check(self.lib.wcfd_face_interpolation(self.dm.desc_ref, self.bc, ptr(self._cv),
                                                    ptr(self._cg), ptr(self._fv), mask, skew, None,
                                                    None, None, self._s()), 'face_interpolation')

    def calculate_gradients(self, output_indices=None):
        """model.py:232-239 (zeroes every gradient first, as the reference)."""
        self._need()
        idx = self._indices(output_indices)
        mask = self._mask(idx)
        _lib.check(self.lib.wcfd_green_gauss(self.dm.desc_ref, ptr(self._fv), ptr(self._cg), mask, 1,
                                             self._s()), 'calculate_gradients')
        self._uvw_grad_zero = (mask & 7) == 0

    def calculate_mass_flux(self):
        self._need()
        _lib.check(self.lib.wcfd_mass_flux(self.dm.desc_ref, ptr(self._fv), ptr(self._flux),
                                           self._s()), 'calculate_mass_flux')

    def divFlux(self, arr=None):
        """model.py:269-275; returns a (C,) tensor in reference order."""
        self._need()
        d = torch.empty(self.num_cells, dtype=self.tdtype, device=self._cv.device)
        _lib.check(self.lib.wcfd_divergence(self.dm.desc_ref, ptr(self._flux), ptr(d), self._s()),
                   'divFlux')
        out = self.dm.cells_to_reference(d)
        if arr is not None:
            arr.copy_(out)
            return arr
        return host_readable(out)

    # ------------------------------------------ check_convergence arguments
    def _ell_of(self, matrix):
        """Momentum matrix as device ELL values [7][C]: a (7, C) device
        tensor as is; an ``Equation`` (its shared plane); or a scipy matrix
        on the reference's rows (C x C, or 3C x 3C interleaved as
        ``Equation.matrix`` returns it, whose three component blocks must be
        equal -- the matrix the reference's bsr_mv applies)."""
        from .terms.equation import Equation
        C = self.num_cells
        if torch.is_tensor(matrix):
            return matrix
        if isinstance(matrix, Equation):
            if not matrix._shared_matrix():
                raise NotImplementedError('check_convergence: the velocity equation\'s three '
                                          'outputs must share one matrix')
            return matrix.values[0].contiguous()
        import scipy.sparse as sp
        if not sp.issparse(matrix):
            raise TypeError(f'vel_matrix: expected an Equation, its .matrix or a device tensor, '
                            f'got {type(matrix).__name__}')
        A = matrix.tocsr()
        n = A.shape[0]
        if n == 3 * C:
            Au = A[0::3, 0::3].tocsr()
            # rows 3c + k hold the u block for every k, and nothing couples components
            if (A - sp.kron(Au, sp.identity(3), format='csr')).count_nonzero():
                raise NotImplementedError('check_convergence: the u, v and w blocks of the '
                                          'velocity matrix must be equal and uncoupled')
            A = Au
        elif n != C:
            raise ValueError(f'vel_matrix has {n} rows; expected {C} or {3 * C}')
        L = self.layout
        old = L.new2old
        vals = np.zeros((7, C), self.float_dtype)
        vals[0] = np.asarray(A[old, old]).ravel()
        for k in range(6):
            nb = L.slot_nbr[k]
            inner = nb != np.arange(C)
            vals[k + 1, inner] = np.asarray(A[old[inner], old[nb[inner]]]).ravel()
        return torch.from_numpy(vals).to(self._cv.device)

    def _cells_of(self, t, rows):
        """A cell field as a contiguous device tensor of the model's type in device
        order: (rows, C) or (C,) device-order tensors as they are; the
        reference's forms, (C * rows,) interleaved or (C, rows), reordered."""
        if t is None:
            return None
        C, dev = self.num_cells, self._cv.device
        t = torch.as_tensor(np.asarray(t) if not torch.is_tensor(t) else t).to(dev)
        shape = tuple(t.shape)
        if rows == 1 and shape == (C,):
            return t.to(self.tdtype).contiguous()
        if rows > 1 and shape == (rows, C):
            return t.to(self.tdtype).contiguous()
        if rows > 1 and shape in ((C * rows,), (C, rows)):
            ref = t.to(self.tdtype).reshape(C, rows).T.contiguous()
            return self.dm.cells_to_device(ref).contiguous()
        raise ValueError(f'expected {(rows, C) if rows > 1 else (C,)} in device order or '
                         f'{(C * rows,)} / {(C, rows)} in reference order, got {shape}')

    def convergence_norms(self, vel_matrix=None, velocity=None, b=None, velocity_correction=None,
                          p_correction=None):
        """The norms of check_convergence from one device report
        (``wcfd_convergence_report``, float64), as a dict of floats: ``r``
        and ``b`` (3 each: ||A u_k - b_k||_2, ||b_k||_2), ``vel_corr``
        (max |c|), ``p`` (max |p|), ``div`` (sum |div F| of the current mass
        fluxes).  ``vel_matrix``: see ``_ell_of``.  ``velocity`` (None: the
        current u, v, w), ``b`` and ``velocity_correction``: [3][C] device
        order, or the reference's interleaved (3C,) / (C, 3) forms;
        ``velocity_correction`` may also be a pair (rUA [C], grad p [3][C])
        in device order.  ``p_correction``: (C,), either order (only its
        maximum is used).  None skips a group (its norms are NaN)."""
        self._need()
        dev = self._cv.device
        if self._conv_ws is None:
            n = self.lib.wcfd_convergence_workspace_bytes()
            self._conv_ws = torch.zeros(n, dtype=torch.uint8, device=dev)
            self._conv_out = torch.zeros(9, dtype=torch.float64, device=dev)
        vals = self._ell_of(vel_matrix) if vel_matrix is not None else None
        if vals is not None:
            vals = self._cells_of(vals, 7)
            u = self._cells_of(self._cv[:3] if velocity is None else velocity, 3)
            bb = self._cells_of(b, 3)
        else:
            u = bb = None
        rua = corr = None
        if isinstance(velocity_correction, (tuple, list)):
            rua, corr = self._cells_of(velocity_correction[0], 1), \
                self._cells_of(velocity_correction[1], 3)
        else:
            corr = self._cells_of(velocity_correction, 3)
        p = self._cells_of(p_correction, 1)
        _lib.check(self.lib.wcfd_convergence_report(
            self.dm.desc_ref, ptr(vals), ptr(u), ptr(bb), ptr(rua), ptr(corr), ptr(p),
            ptr(self._flux), ptr(self._conv_ws), ptr(self._conv_out), self._s()),
            'convergence report')
        o = self._conv_out.cpu().numpy()
        nan = float('nan')
        return {'r': [float(x) for x in o[0:3]] if vals is not None else [nan] * 3,
                'b': [float(x) for x in o[3:6]] if vals is not None else [nan] * 3,
                'vel_corr': float(o[6]) if corr is not None else nan,
                'p': float(o[7]) if p is not None else nan,
                'div': float(o[8])}

    def check_convergence(self, vel_matrix, velocity, b, div_u=None, velocity_correction=None,
                          p_correction=None, log=True, verbose=True):
        """model.py:278-315: momentum_x/y/z, continuity, local_continuity,
        velocity_correction and pressure_correction of ``self.Convergence``
        from one device report.  Takes the reference's call as it stands
        (incompressible.py:163: ``vel_equation.matrix, HbyA,
        vel_equation.rhs, div_u, vel_correction.flatten(), p_new``) as well as
        device-order buffers; see ``convergence_norms``.  As the reference,
        continuity is taken from the current mass fluxes (``div_u`` is only a
        buffer there, recomputed by divFlux), a missing correction reports NaN
        and leaves its criterion's verdict as it was, the dict is printed and,
        with ``log``, logged and ``self.res`` set to |p_correction| (a host
        array), and a non-finite momentum_x residual raises ValueError.
        Returns has_converged()."""
        n = self.convergence_norms(vel_matrix, velocity, b, velocity_correction, p_correction)
        cv = self.Convergence
        nan = float('nan')
        conv = {
            'momentum_x': cv.momentum_x.check_value(n['r'][0], n['b'][0]),
            'momentum_y': cv.momentum_y.check_value(n['r'][1], n['b'][1]),
            'momentum_z': cv.momentum_z.check_value(n['r'][2], n['b'][2]),
            'continuity': cv.continuity.check_value(n['div']),
            'local_continuity': cv.local_continuity.check_value(n['div']),
            'velocity_correction': (cv.velocity_correction.check_value(n['vel_corr'])
                                    if velocity_correction is not None else nan),
            'pressure_correction': (cv.pressure_correction.check_value(n['p'])
                                    if p_correction is not None else nan),
        }
        if verbose:
            print(conv)
        if log:
            cv.log(conv)
            if p_correction is not None:
                pc = p_correction.detach().cpu().numpy() if torch.is_tensor(p_correction) \
                    else np.asarray(p_correction)
                self.res = np.abs(pc)
        if np.isnan(conv['momentum_x'][0]):
            raise ValueError('momentum_x residual is not finite')
        return cv.has_converged()

    def get_gradient(self, output_index, coeff=1.):
        """model.py:242-254: coeff * grad(phi), (C,3) in reference order."""
        self._need()
        o = self._indices(output_index)[0]
        g = self.cell_gradients[:, o]
        if isinstance(coeff, (float, int)):
            return (torch.tensor(coeff, dtype=self.tdtype, device=g.device) * g)
        c = torch.as_tensor(coeff, dtype=self.tdtype, device=g.device).reshape(-1, 1)
        return c * g

    def relax(self, new_value, alpha, output_index):
        """model.py:256-262 (new_value in reference order)."""
        self._need()
        o = self._indices(output_index)[0]
        nv = self.dm.cells_to_device(torch.as_tensor(new_value, dtype=self.tdtype,
                                                     device=self._cv.device))
        a = torch.tensor(alpha, dtype=self.tdtype)
        self._cv[o] = a.item() * nv + (1.0 - a).item() * self._cv[o]

    def replace_cell_values(self, output_indices, value):
        """model.py:318-325: value is interleaved (C*n,) in reference order."""
        self._need()
        idx = self._indices(output_indices)
        v = torch.as_tensor(value, dtype=self.tdtype, device=self._cv.device).reshape(-1, len(idx))
        for k, o in enumerate(idx):
            self._cv[o] = self.dm.cells_to_device(v[:, k].contiguous())

    def update_cell_values(self, field_idx, value, scale=1.):
        """model.py:329-349."""
        self._need()
        idx = self._indices(field_idx)
        v = torch.as_tensor(value, dtype=self.tdtype, device=self._cv.device)
        for k, o in enumerate(idx):
            if v.numel() == 1:
                self._cv[o] += scale * v.reshape(())
            else:
                vv = v.reshape(self.num_cells, -1)[:, k] if v.ndim > 1 else v
                self._cv[o] += scale * self.dm.cells_to_device(vv.contiguous())

    def struct_member_to_array(self, member='centroid', struct='cells'):
        """model.py:177-190 for the members the examples read."""
        cp, fp = self.cell_properties, self.face_properties
        table = {('centroid', 'cells'): cp.centroids, ('volume', 'cells'): cp.volumes,
                 ('centroid', 'faces'): fp.centroid}
        if (member, struct) not in table:
            raise ValueError(f'{member} of {struct} is not exported')
        return np.asarray(table[(member, struct)])
