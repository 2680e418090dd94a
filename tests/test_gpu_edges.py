"""Edge cases of the device path against the oracle: meshes so small that
most or all faces are boundary faces and a reduction launch has a single,
partly empty workgroup -- one hexahedron, a row of two, a column of three,
and the smallest tetrahedral and wedge boxes.  Five SIMPLE iterations with
the reference solver settings; fields within rel-L2 1e-4 of the oracle
(absolute where the oracle's field is zero)."""
import numpy as np
import pytest

from cases import device_case, jittered_box, oracle_case, rel_l2

pytestmark = pytest.mark.gpu

CASES = [
    ('one hexahedron, channel', (1, 1, 1), 'hexahedron', 'channel'),
    ('two hexahedra, channel', (2, 1, 1), 'hexahedron', 'channel'),
    ('three hexahedra, cavity', (1, 3, 1), 'hexahedron', 'cavity'),
    ('six tetrahedra, closed cavity', (1, 1, 1), 'tetra', 'closed_cavity'),
    ('twelve tetrahedra, closed cavity', (2, 1, 1), 'tetra', 'closed_cavity'),
    ('two wedges, cavity', (1, 1, 1), 'wedge', 'cavity'),
]


@pytest.mark.parametrize('case', CASES, ids=lambda c: c[0])
def test_tiny_meshes_vs_oracle(gpu, case):
    from oracle.simple import OracleSolver
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    from warpcfd_amd.preprocess.grid import split_hexes
    name, shape, cell_type, kind = case
    nodes, cells = jittered_box(shape, 0.0, 0)
    if cell_type != 'hexahedron':
        cells = split_hexes(cells, cell_type)
    ref_cell = None if kind == 'channel' else 0
    outflow = kind == 'channel'
    of = oracle_case(nodes, cells, kind, ref_cell=ref_cell)
    model = device_case(nodes, cells, kind, ref_cell=ref_cell)
    os_ = OracleSolver(of, 0.7, 0.3, nu=0.01, implicit_outflow=outflow)
    sv = IncompressibleSolver(model, 0.7, 0.3, implicit_outflow=outflow)
    for _ in range(5):
        os_.outer_iteration()
        sv.outer_iteration()
    got = model.cell_values.cpu().numpy()
    assert np.all(np.isfinite(got))
    for q in range(4):
        ref = of.cell_values[:, q]
        err = rel_l2(got[:, q], ref)
        assert err <= 1e-4, (name, q, err, got[:, q], ref)
