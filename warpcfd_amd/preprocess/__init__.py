from .mesh import Mesh
from .grid import (define_boundary_walls, structured_hex, cavity_mesh, channel_mesh, cylinder_mesh,
                   box_and_body_walls, create_2D_grid, UnstructuredGrid)
from .io import read_msh, read_vtu, write_msh, write_vtu
from .layout import DeviceLayout
