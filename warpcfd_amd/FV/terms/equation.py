"""Equation (warp_cfd/FV/terms/equation.py:10-204) on the device.

The system of an equation with O outputs is held as one ELL plane per
output, ``values`` (O, 7, C) -- diagonal, then the six local-face
coefficients -- and ``rhs`` (O, C), device cell order: the reference's
BSR matrix with rows O*c + o (equation_kernels.py:10-38), split by output.
``form_system`` accumulates face-based terms, cell-based terms and explicit
terms with HIP kernels; ``relax`` and ``replace_row`` edit the system in
place; ``solve_Axb`` runs the library's Jacobi-preconditioned BiCGSTAB or CG.

Views in the reference's layout (interleaved rows O*c + o, reference cell
order): ``rhs``, ``x``, ``diagonal``; ``matrix`` is a scipy CSR matrix and
``dense`` a numpy array, both for inspection on small meshes.

Solving: one output, either solver; three outputs sharing one matrix (the
planes are compared) use the three-right-hand-side BiCGSTAB, whose scalars
span all three components as the reference's 3C x 3C solve does.  Other
combinations are rejected.
"""
import ctypes
import types

import numpy as np
import torch

from ... import _lib
from ..device import host_readable, ptr, stream_ptr
from ..field import Field
from .terms import Term, int32_array

# stand-in for warp.optim.linear: the solvers an Equation can be given
linear = types.SimpleNamespace(bicgstab='bicgstab', cg='cg')


def _solver_name(solver):
    if isinstance(solver, str):
        name = solver
    else:
        name = getattr(solver, '__name__', str(solver))
    if name not in ('bicgstab', 'cg'):
        raise ValueError(f'solver must be linear.bicgstab or linear.cg, not {solver!r}')
    return name


class Equation:
    def __init__(self, fvm, fields, solver=linear.bicgstab):
        if not fvm.finalized:
            raise ValueError('Please call method model.finalize() first')
        if isinstance(fields, str):
            fields = [fields]
        need_globals = all(f in fvm.output_variables for f in fields)
        if need_globals:
            self.fields = [fvm.fields[f] for f in fields]
            self.global_output_indices = np.array([f.index for f in self.fields], np.int32)
        else:
            self.fields = [Field(f) for f in fields]
            self.global_output_indices = None
        self.num_outputs = len(fields)
        self.sparse_rows = fvm.num_cells * self.num_outputs
        self.linear_solver = _solver_name(solver)
        self.max_iter = 500
        self.tol = 1.e-5 if fvm.tdtype == torch.float32 else 1.e-6   # equation.py:35-39
        self.check_every = 10
        self.initial = True
        self.float_dtype = fvm.float_dtype
        self.int_dtype = np.int32
        self.fvm = fvm
        self.initial_gradient = False
        C, O = fvm.num_cells, self.num_outputs
        dev = fvm._cv.device
        self.values = torch.zeros((O, 7, C), dtype=fvm.tdtype, device=dev)
        self._rhs = torch.zeros((O, C), dtype=fvm.tdtype, device=dev)
        self._x = torch.zeros((O, C), dtype=fvm.tdtype, device=dev)
        self._ws = None
        self.last_info = None

    # --------------------------------------------------- reference layout
    def _to_ref(self, t):
        """(O, C) device -> (C * O,) interleaved, reference cell order."""
        return host_readable(self.fvm.dm.cells_to_reference(t).T.reshape(-1).contiguous())

    def _to_dev(self, v):
        """(C * O,) interleaved, reference cell order -> (O, C) device."""
        C, O = self.fvm.num_cells, self.num_outputs
        v = torch.as_tensor(np.asarray(v) if not torch.is_tensor(v) else v,
                            dtype=self.fvm.tdtype).to(self.values.device).reshape(C, O)
        return self.fvm.dm.cells_to_device(v.T.contiguous()).contiguous()

    @property
    def rhs(self):
        return self._to_ref(self._rhs)

    @property
    def x(self):
        return self._to_ref(self._x)

    @property
    def output_gradient(self):
        if not hasattr(self, '_output_gradient'):
            self._output_gradient = torch.zeros((self.fvm.num_cells, self.fvm.dimension),
                                                dtype=self.fvm.tdtype, device=self.values.device)
        return self._output_gradient

    @property
    def diagonal(self):
        """The matrix diagonal, (C * O,) in the reference's row order."""
        return self._to_ref(self.values[:, 0, :].contiguous())

    @property
    def matrix(self):
        """The system as a scipy CSR matrix on the reference's rows O*c + o
        (cell numbering of the input mesh)."""
        import scipy.sparse as sp
        L = self.fvm.layout
        O, C = self.num_outputs, self.fvm.num_cells
        v = self.values.cpu().numpy()
        rows, cols, vals = [], [], []
        old = L.new2old
        for o in range(O):
            rows.append(O * old + o)
            cols.append(O * old + o)
            vals.append(v[o, 0])
            for k in range(6):
                nb = L.slot_nbr[k]
                inner = nb != np.arange(C)
                rows.append(O * old[inner] + o)
                cols.append(O * old[nb[inner]] + o)
                vals.append(v[o, k + 1][inner])
        return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                             shape=(O * C, O * C))

    @property
    def dense(self):
        """Dense matrix (O C x O C); for small meshes only."""
        return self.matrix.toarray()

    # ------------------------------------------------------------ system
    def reset(self):
        self._rhs.zero_()
        self.values.zero_()

    def form_system(self, implicit_terms=None, explicit_terms=None, *, fvm=None):
        """equation.py:99-124: implicit terms into the matrix (face-based:
        diagonal and neighbour coefficients, boundary explicit parts into
        the right-hand side; cell-based: diagonal and right-hand side), then
        explicit terms added to the right-hand side."""
        self.reset()
        fvm = self.fvm if fvm is None else fvm
        lib = fvm.lib
        if implicit_terms is not None:
            if isinstance(implicit_terms, Term):
                implicit_terms = [implicit_terms]
            for term in implicit_terms:
                assert term.implicit
                self._check_outputs(term)
                fn = (lib.wcfd_equation_add_cell_term if term.cell_based
                      else lib.wcfd_equation_add_face_term)
                _lib.check(fn(fvm.dm.desc_ref, ptr(term._w), self.num_outputs, term.scale,
                              ptr(self.values), ptr(self._rhs), stream_ptr()), 'form_system')
        if explicit_terms is not None:
            if isinstance(explicit_terms, Term):
                explicit_terms = [explicit_terms]
            for term in explicit_terms:
                assert not term.implicit
                if term._w.shape != self._rhs.shape:
                    raise ValueError(f'explicit term with {term._w.shape[0]} rows per cell in an '
                                     f'equation of {self.num_outputs} outputs')
                self._add_rhs_dev(term._w, term.scale)

    def _check_outputs(self, term):
        if term.num_outputs != self.num_outputs:
            raise ValueError(f'term of {term.num_outputs} outputs in an equation of '
                             f'{self.num_outputs}')

    def _add_rhs_dev(self, arr, scale):
        n = self._rhs.numel()
        _lib.check(self.fvm.lib.wcfd_equation_add_rhs(n, ptr(arr), float(np.float32(scale)),
                                                      ptr(self._rhs), stream_ptr()), 'add_RHS')

    def add_RHS(self, arr, scale=1.):
        """equation.py:169-170; arr in the reference's row order."""
        self._add_rhs_dev(self._to_dev(arr), scale)

    def relax(self, relaxation_factor, fvm=None):
        """equation.py:173-181: b += (1 - a)/a A_ii phi, A_ii /= a, phi the
        outputs' current cell values."""
        fvm = self.fvm if fvm is None else fvm
        assert self.global_output_indices is not None, \
            'Relaxation is only available for fields defined in fvm'
        g, gp = int32_array(self.global_output_indices)
        _lib.check(fvm.lib.wcfd_equation_relax(fvm.dm.desc_ref, ptr(fvm._cv), gp, self.num_outputs,
                                               float(relaxation_factor), ptr(self.values),
                                               ptr(self._rhs), stream_ptr()), 'relax')

    def replace_row(self, row_id, rhs_value):
        """equation.py:184-193: identity rows.  row_id: a row O*c + o of the
        reference's numbering, or an array of them; rhs_value: a float, or
        one value per row."""
        rows = np.atleast_1d(np.asarray(row_id.cpu() if torch.is_tensor(row_id) else row_id,
                                        np.int64))
        vals = np.broadcast_to(np.asarray(rhs_value.cpu() if torch.is_tensor(rhs_value)
                                          else rhs_value, np.float32), rows.shape)
        O = self.num_outputs
        L = self.fvm.layout
        for r, v in zip(rows, vals):
            if not 0 <= r < self.sparse_rows:
                raise ValueError(f'row {r} outside the {self.sparse_rows} rows')
            c = int(L.old2new[r // O])
            _lib.check(self.fvm.lib.wcfd_equation_replace_row(
                self.fvm.dm.desc_ref, O, int(r % O), c, float(v), ptr(self.values), ptr(self._rhs),
                stream_ptr()), 'replace_row')

    # ------------------------------------------------------------- solve
    def _shared_matrix(self):
        v = self.values
        return all(torch.equal(v[o], v[0]) for o in range(1, self.num_outputs))

    def solve_Axb(self, x=None):
        """equation.py:195-200: Jacobi-preconditioned solve with
        atol = tol = self.tol, at most self.max_iter iterations, warm-started
        from x.  x: None (the equation's own vector) or a vector in the
        reference's row order, updated in place.  Returns ((iterations,
        residual, threshold), x)."""
        fvm = self.fvm
        lib = fvm.lib
        O, C = self.num_outputs, fvm.num_cells
        if x is not None:
            self._x.copy_(self._to_dev(x))
        if self._ws is None:
            self._ws = torch.zeros(lib.wcfd_krylov_workspace_bytes(C, 3), dtype=torch.uint8,
                                   device=self.values.device)
        info = _lib.SolveInfo()
        if O == 1:
            if self.linear_solver == 'cg':
                code = lib.wcfd_cg_ell(fvm.dm.desc_ref, ptr(self.values), _lib.CG_STANDARD,
                                       ptr(self._rhs), ptr(self._x), self.tol, self.tol,
                                       self.max_iter, self.check_every, ptr(self._ws),
                                       ctypes.byref(info), stream_ptr())
            else:
                code = lib.wcfd_bicgstab_ell(fvm.dm.desc_ref, ptr(self.values), 1, ptr(self._rhs),
                                             ptr(self._x), self.tol, self.tol, self.max_iter,
                                             self.check_every, ptr(self._ws), ctypes.byref(info),
                                             stream_ptr())
        elif O == 3 and self.linear_solver == 'bicgstab' and self._shared_matrix():
            code = lib.wcfd_bicgstab_ell(fvm.dm.desc_ref, ptr(self.values[0]), 3, ptr(self._rhs),
                                         ptr(self._x), self.tol, self.tol, self.max_iter,
                                         self.check_every, ptr(self._ws), ctypes.byref(info),
                                         stream_ptr())
        else:
            raise NotImplementedError(
                'solve_Axb: one output (BiCGSTAB or CG), or three outputs sharing one matrix '
                '(BiCGSTAB); this equation has ' + str(O) + ' outputs with '
                + ('the same' if O > 1 and self._shared_matrix() else 'different')
                + ' matrices and solver ' + self.linear_solver)
        _lib.check(code, 'solve_Axb')
        self.last_info = info
        result = (info.iterations, info.residual, info.threshold)
        if x is None:
            return result, self.x
        xr = self._to_ref(self._x)
        if torch.is_tensor(x):
            x.copy_(xr.to(x.device).reshape(x.shape))
            return result, x
        return result, xr
