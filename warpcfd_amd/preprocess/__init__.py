from .mesh import Mesh
from .grid import (define_boundary_walls, structured_hex, cavity_mesh, channel_mesh, cylinder_mesh,
                   box_and_body_walls)
from .layout import DeviceLayout
