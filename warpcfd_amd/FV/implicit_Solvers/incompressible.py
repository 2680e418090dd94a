"""IncompressibleSolver: steady SIMPLE / transient PISO-like loop on the GPU.

Mirrors warp_cfd/FV/implicit_Solvers/incompressible.py:10-172.  One outer
iteration is a single C-ABI call (``wcfd_simple_iteration``) that launches
the fused kernels and the two Krylov solves; the state (fields, matrices,
Krylov warm starts) never leaves the device.
"""
import ctypes

import numpy as np
import torch

from ... import _lib
from ..device import ptr, stream_ptr


class Monitor:
    """Host side of the convergence monitor of a steady run: the same rule
    (``wcfd_monitor_update`` in libwcfd.so) that the batched SIMPLE step runs
    on the device -- continuity <= stop_c and an a-posteriori estimate of the
    velocity error from the decay rate of the per-iteration velocity change
    <= stop_u (or that change at the float32 floor, <= floor_u)."""

    def __init__(self, stop_c=1e-6, stop_u=5e-5, floor_u=1e-7):
        self.stop = (float(stop_c), float(stop_u), float(floor_u))
        self.record = (ctypes.c_double * _lib.MONITOR_LEN)()
        self.lib = _lib.load()

    def update(self, continuity, velocity_change):
        r = self.lib.wcfd_monitor_update(self.record, float(continuity), float(velocity_change),
                                         *self.stop)
        if r < 0:
            _lib.check(-r, 'monitor')
        return r == 1

    @property
    def estimate(self):
        """(continuity, velocity change, mean change, rate, error estimate)."""
        return tuple(self.record[i] for i in range(1, 6))


class IncompressibleSolver:
    def __init__(self, model, u_relaxation_factor=None, p_relaxation_factor=None, correction=False,
                 steady_state=True, implicit_outflow=False):
        """``implicit_outflow`` is not in the reference: it treats the
        convective flux through Neumann outflow faces implicitly (upwind,
        F_b on the diagonal) instead of lagging F_b phi_b on the right-hand
        side.  The fixed point is unchanged; it keeps outlet rows diagonally
        dominant, without which the reference scheme diverges on
        inlet/outlet flows at cell Peclet numbers of a few (the channel)."""
        if not model.finalized:
            raise ValueError('Please call method model.finalize() before passing model in to solver')
        if not model.boundary.uniform_velocity_types():
            raise NotImplementedError('u, v and w must share the boundary-condition type on every '
                                      'face (true for every helper of Boundary)')
        self.model = model
        self.float_dtype = model.float_dtype
        self.steady_state = steady_state
        self.correction = bool(correction)
        self.implicit_outflow = bool(implicit_outflow)
        self.p_relaxation_factor = p_relaxation_factor
        self.u_relaxation_factor = u_relaxation_factor
        self.material = model.material
        self.lib = model.lib
        C, F, nnz = model.num_cells, model.dm.n_faces, model.dm.nnz
        dev = model._cv.device
        z = lambda *s: torch.zeros(*s, dtype=model.tdtype, device=dev)
        self.HbyA = z(3, C)
        self.p_x = z(C)
        self.rUA = z(C)
        self.rUA_faces = z(F)
        self.mom_values = z(7, C)      # slot-major (ELL) matrices
        self.mom_rhs = z(3, C)
        self.p_values = z(7, C)
        self.p_rhs = z(C)
        self.resid = torch.zeros(4, dtype=torch.float64, device=dev)
        self.workspace = torch.zeros(self.lib.wcfd_simple_workspace_bytes(C, F), dtype=torch.uint8,
                                     device=dev)
        self.max_iter = 500
        # equation.py:35-39: atol 1e-5 for float32 models, 1e-6 for float64
        self.tol = 1e-5 if model.tdtype == torch.float32 else 1e-6
        self.rtol = None         # None: relative tolerance = atol, as the reference's call
        # not in the reference: also stop at rtol0 * ||r_0|| (the warm-start
        # residual); 0 = off
        self.rtol0 = 0.0
        self.check_every = 10
        # momentum solve check interval; None: check_every (the reference's
        # single setting)
        self.check_every_u = None
        # not in the reference: extra pressure-CG sweeps from a float64
        # residual (mixed-precision iterative refinement); 0 = off
        self.p_refine = 0
        # not in the reference: iteration cap of the momentum solve (None:
        # max_iter, as the reference)
        self.max_iter_u = None
        # pressure CG: WCFD_CG_STANDARD (the reference's two-reduction form) or
        # WCFD_CG_SINGLE_REDUCTION (same iterates in exact arithmetic, one
        # fused kernel per iteration)
        self.cg_variant = _lib.CG_STANDARD
        self.set_time_stepping(1, 1.)
        self.NUM_ORTHOGONAL_CORRECTORS = 1
        self.NUM_OUTER_LOOPS = 1
        self.NUM_INNER_LOOPS = 1
        self.history = []
        self._state = _lib.SimpleState()
        for name, t in (('cell_values', model._cv), ('cell_grad', model._cg),
                        ('face_values', model._fv), ('mass_flux', model._flux),
                        ('hbya', self.HbyA), ('p_x', self.p_x), ('rua', self.rUA),
                        ('rua_face', self.rUA_faces), ('mom_values', self.mom_values),
                        ('mom_rhs', self.mom_rhs), ('p_values', self.p_values),
                        ('p_rhs', self.p_rhs), ('resid', self.resid)):
            setattr(self._state, name, t.data_ptr())

    def set_time_stepping(self, MAX_TIME_STEPS, dt, automatic=False, CFL_Max=1.):
        if automatic:
            raise NotImplementedError('automatic time stepping not yet implemented')
        self.MAX_TIME_STEPS = MAX_TIME_STEPS
        self.dt = dt

    def set_iterations(self, NUM_OUTER_LOOPS, NUM_INNER_LOOPS, NUM_ORTHOGONAL_CORRECTORS):
        self.NUM_OUTER_LOOPS = NUM_OUTER_LOOPS
        self.NUM_INNER_LOOPS = NUM_INNER_LOOPS
        self.NUM_ORTHOGONAL_CORRECTORS = NUM_ORTHOGONAL_CORRECTORS

    def params(self, dt=None):
        m = self.model
        p = _lib.SimpleParams()
        rho = float(self.material.density)
        p.nu = float(self.material.viscosity) / rho
        p.density = rho
        p.alpha_u = float(self.u_relaxation_factor) if self.u_relaxation_factor is not None else 0.0
        p.alpha_p = float(self.p_relaxation_factor) if self.p_relaxation_factor is not None else 0.0
        p.dt = 0.0 if self.steady_state else float(self.dt if dt is None else dt)
        p.scheme_flags = (_lib.SCHEME_CORRECTED if self.correction else 0) | \
            (_lib.SCHEME_IMPLICIT_OUTFLOW if self.implicit_outflow else 0)
        p.ref_cell = m.reference_cell_device
        p.ref_value = float(m.reference_pressure)
        p.n_inner = int(self.NUM_INNER_LOOPS)
        p.n_ortho = int(self.NUM_ORTHOGONAL_CORRECTORS)
        p.atol = float(self.tol)
        p.tol = float(self.tol if self.rtol is None else self.rtol)
        p.rtol0 = float(self.rtol0)
        p.maxiter = int(self.max_iter)
        p.check_every = int(self.check_every)
        p.check_every_u = int(self.check_every_u or 0)
        p.p_refine = int(self.p_refine)
        p.maxiter_u = int(self.max_iter_u or 0)
        p.first_iteration = 0 if m._uvw_grad_zero else 1
        p.cg_variant = int(self.cg_variant)
        return p

    def outer_iteration(self, dt=None, stream=None):
        """One pass of incompressible.py:95-156; returns a dict of diagnostics."""
        m = self.model
        prm = self.params(dt)
        info = _lib.SimpleInfo()
        code = self.lib.wcfd_simple_iteration(m.dm.desc_ref, m.bc, ctypes.byref(prm),
                                              ctypes.byref(self._state), ptr(self.workspace),
                                              ctypes.byref(info), stream_ptr(stream))
        _lib.check(code, 'simple iteration')
        m._uvw_grad_zero = True
        d = {'u_iters': info.momentum.iterations, 'u_residual': info.momentum.residual,
             'p_iters': info.pressure.iterations, 'p_residual': info.pressure.residual,
             'continuity': info.continuity, 'velocity_change': info.velocity_change}
        self.history.append(d)
        if not np.isfinite(info.momentum.residual):
            raise ValueError('momentum residual is not finite')
        return d

    def run_to_convergence(self, stop_c=1e-6, stop_u=5e-5, floor_u=1e-7, max_outer=20000):
        """Steady run until the convergence monitor (``Monitor``: continuity
        <= stop_c and estimated velocity error <= stop_u) reports converged;
        not in the reference, whose run() has no stop test.  Returns
        (iterations, converged, monitor)."""
        mon = Monitor(stop_c, stop_u, floor_u)
        it = 0
        done = False
        while it < max_outer and not done:
            d = self.outer_iteration()
            it += 1
            done = mon.update(d['continuity'], d['velocity_change'])
        return it, done, mon

    def check_convergence(self, log=True, verbose=True):
        """The check of incompressible.py:158-163 and 169-172,
        model.check_convergence(vel_equation.matrix, HbyA, vel_equation.rhs,
        div_u, vel_correction, p_new), on the device state of the last outer
        iteration: the relaxed momentum matrix and right-hand side, the
        corrected velocity, rUA grad p and the pressure solve's result."""
        m = self.model
        return m.check_convergence(self.mom_values, None, self.mom_rhs, None,
                                   (self.rUA, m._cg[9:12]), m.dm.cells_to_reference(self.p_x),
                                   log=log, verbose=verbose)

    def run(self, steps_per_check=10, callback=None, check=True):
        """incompressible.py:72-172 (no early stop, as the reference): every
        steps_per_check steps, and once more after the last, the reference's
        convergence check (printed and logged in model.Convergence) unless
        ``check`` is False; ``callback(steps, diagnostics)`` (not in the
        reference) on the same steps."""
        n_t = self.MAX_TIME_STEPS
        for nt in range(n_t):
            for it in range(self.NUM_OUTER_LOOPS):
                d = self.outer_iteration()
                steps = it if self.steady_state else nt
                if steps % steps_per_check == 0:
                    if callback is not None:
                        callback(steps, d)
                    if check:
                        self.check_convergence()
        if check:
            self.check_convergence()
        return self.history
