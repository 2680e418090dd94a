"""Composable finite-volume terms and equations (warp_cfd/FV/terms/__init__.py)."""
from .convection import ConvectionTerm
from .ddt import DdtTerm
from .diffusion import DiffusionTerm
from .equation import Equation, linear
from .grad import GradTerm
from .terms import Term

__all__ = ['ConvectionTerm', 'DiffusionTerm', 'GradTerm', 'DdtTerm', 'Equation', 'Term', 'linear']
