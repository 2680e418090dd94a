"""Finite-volume model, operators and solvers (device path)."""
from .model import FVM
