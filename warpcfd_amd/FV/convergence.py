"""Convergence criteria of an outer iteration (Convergence.py:7-106).

``Criteria`` and ``Convergence`` keep the reference's names, tolerances and
verdict rules: continuity and local_continuity are absolute 1-norms (eps
1e-4), momentum_x/y/z relative 2-norms ||Ax - b|| / max(||b||, 1e-6) (eps
1e-3), pressure_correction and velocity_correction absolute inf-norms (eps
1e-4); ``has_converged`` needs every criterion in ``criterias`` (not
local_continuity) to have passed its last check.

``FVM.check_convergence`` gets the norms of all of them from one device
report (``wcfd_convergence_report``, float64) and hands each to
``Criteria.check_value``; ``Criteria.check`` keeps the reference's generic
host form for arrays a caller holds (numpy norms in float64).
"""
import numpy as np


class Criteria:
    """Convergence.py:7-48."""

    def __init__(self, name, eps=1e-4, tol='relative', ord=1., is_flux=False):
        assert tol == 'relative' or tol == 'abs'
        self.name = name
        self.eps = eps
        self.tol = tol
        self.ord = ord
        self.is_converged = False
        self.non_zero = 1e-6

    @staticmethod
    def _host(a):
        if hasattr(a, 'detach'):
            a = a.detach().cpu().numpy()
        return np.asarray(a, dtype=np.float64)

    def residual(self, arr, rhs=0., ord=None):
        ord = self.ord if ord is None else ord
        return float(np.linalg.norm(self._host(arr) - self._host(rhs), ord))

    def relative_residual(self, arr, rhs, ord=None):
        ord = self.ord if ord is None else ord
        res = self.residual(arr, rhs, ord)
        return res / max(float(np.linalg.norm(np.atleast_1d(self._host(rhs)), ord)),
                         self.non_zero)

    def check_value(self, res, rhs_norm=None):
        """The verdict on a norm computed elsewhere: ``res`` is ||arr - rhs||
        in this criterion's order; for a relative criterion ``rhs_norm`` is
        ||rhs|| and the residual is res / max(rhs_norm, 1e-6)."""
        res = float(res)
        if self.tol == 'relative':
            res = res / max(float(rhs_norm), self.non_zero)
        self.is_converged = bool(res < self.eps)
        return res, self.is_converged

    def check(self, arr, rhs=0., ord=None):
        arr = self._host(arr)
        rhs = np.atleast_1d(self._host(rhs))
        if self.tol == 'abs':
            res = self.residual(arr, rhs, ord)
        else:
            res = self.relative_residual(arr, rhs, ord)
        self.is_converged = bool(res < self.eps)
        return res, self.is_converged


class Convergence:
    """Convergence.py:51-97."""

    def __init__(self):
        self.continuity = Criteria('continuity', eps=1e-4, tol='abs')
        self.local_continuity = Criteria('local_continuity', eps=1e-4, tol='abs')
        self.momentum_x = Criteria('momentum_x', ord=2., eps=1e-3)
        self.momentum_y = Criteria('momentum_y', ord=2., eps=1e-3)
        self.momentum_z = Criteria('momentum_z', ord=2., eps=1e-3)
        self.pressure_correction = Criteria('pressure_correction', tol='abs', eps=1e-4,
                                            ord=np.inf)
        self.velocity_correction = Criteria('velocity_correction', tol='abs', eps=1e-4,
                                            ord=np.inf)
        self.criterias = [self.continuity, self.momentum_x, self.momentum_y, self.momentum_z,
                          self.pressure_correction, self.velocity_correction]
        self.log_list = []

    def check(self, criteria, arr, rhs=0., ord=None):
        return getattr(self, criteria).check(arr, rhs, ord)

    def has_converged(self):
        return all(c.is_converged for c in self.criterias)

    def log(self, step):
        self.log_list.append(step)

    def local_continuity_error(self, div_u, local_err=None):
        """|div u| per cell (Convergence.py:86-97)."""
        out = div_u.abs()
        if local_err is not None:
            local_err.copy_(out)
            return local_err
        return out

    def plot_residuals(self):
        from matplotlib import pyplot as plt
        x = np.arange(len(self.log_list))
        for c in self.criterias:
            plt.plot(x, [step[c.name][0] for step in self.log_list], label=c.name)
        plt.yscale('log')
        plt.legend()
        plt.show()
