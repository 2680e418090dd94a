"""Fluid properties of a model: constant density and dynamic viscosity.

Serves the calls that the solver and the reference's examples make on
``FVM.material`` (warp_cfd/material/mat.py:1-28; examples/LDC/LDC.py:36;
incompressible.py:87-89): ``create_incompressible_newtonian_fluid``, the
``density`` and ``viscosity`` properties, ``set_density`` and
``set_viscosity``.  The SIMPLE/PISO path only ever sees a constant-density
Newtonian fluid, so both properties are stored as positive finite floats and
checked when they are set rather than when a kernel first divides by them.
"""
import math


class Material:
    __slots__ = ('_values',)

    def __init__(self):
        self._values = {}

    @staticmethod
    def _checked(name, value):
        v = float(value)
        if not (math.isfinite(v) and v > 0.0):
            raise ValueError(f'fluid {name} must be a positive finite number, got {value!r}')
        return v

    def _read(self, name):
        if name not in self._values:
            raise ValueError(f'fluid {name} is unset: call create_incompressible_newtonian_fluid '
                             f'or set_{name} before using the model')
        return self._values[name]

    @property
    def density(self):
        return self._read('density')

    @property
    def viscosity(self):
        """Dynamic viscosity mu; the solver uses mu / density."""
        return self._read('viscosity')

    @property
    def kinematic_viscosity(self):
        return self.viscosity / self.density

    def set_density(self, density, incompressible=True):
        """``incompressible`` is accepted for the reference's signature; the
        density is constant on this path either way."""
        self._values['density'] = self._checked('density', density)

    def set_viscosity(self, viscosity, newtonian=True):
        """``newtonian`` is accepted for the reference's signature; the
        viscosity is constant on this path either way."""
        self._values['viscosity'] = self._checked('viscosity', viscosity)

    def create_incompressible_newtonian_fluid(self, viscosity, density):
        self.set_viscosity(viscosity)
        self.set_density(density)
