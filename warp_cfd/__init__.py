"""The reference's import path, ``warp_cfd``, for this package.

Reference scripts import ``warp_cfd.FV.model.FVM``,
``warp_cfd.FV.implicit_Solvers.IncompressibleSolver`` and
``warp_cfd.preprocess.{Mesh, create_2D_grid, define_boundary_walls}``
(examples/LDC/LDC.py:1-7).  Every module ``warp_cfd.X`` is the module
``warpcfd_amd.X`` itself -- the same module object, so classes compare
equal whichever name imported them -- resolved lazily by a finder on
``sys.meta_path``.  Nothing here runs the reference or Warp.
"""
import importlib
import importlib.abc
import importlib.util
import sys

import warpcfd_amd

_ALIAS, _REAL = __name__, warpcfd_amd.__name__


class _AliasFinder(importlib.abc.MetaPathFinder, importlib.abc.Loader):
    def find_spec(self, name, path=None, target=None):
        if not name.startswith(_ALIAS + '.'):
            return None
        real = _REAL + name[len(_ALIAS):]
        spec = importlib.util.find_spec(real)
        if spec is None:
            return None
        return importlib.util.spec_from_loader(
            name, self, is_package=spec.submodule_search_locations is not None)

    def create_module(self, spec):
        # the real module, already executed; the import system keeps its
        # __name__ and __spec__ and files it under the alias as well
        return importlib.import_module(_REAL + spec.name[len(_ALIAS):])

    def exec_module(self, module):
        pass


if not any(isinstance(f, _AliasFinder) for f in sys.meta_path):
    sys.meta_path.insert(0, _AliasFinder())

__path__ = []            # a package, so that warp_cfd.X submodule imports reach the finder
