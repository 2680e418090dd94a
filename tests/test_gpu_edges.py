"""Edge cases of the device path against the oracle: meshes so small that
most or all faces are boundary faces and a reduction launch has a single,
partly empty workgroup -- one hexahedron, a row of two, a column of three,
and the smallest tetrahedral and wedge boxes.  Five SIMPLE iterations with
the reference solver settings; fields within rel-L2 1e-4 of the oracle
(absolute where the oracle's field is zero).  Also the int32 neighbour
indices a mesh too wide for int16 offsets keeps."""
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


def test_int32_neighbour_path_vs_oracle(gpu):
    """Cells in a shuffled order and no renumbering: the neighbour band
    exceeds the int16 range, so the device keeps int32 neighbour indices
    (the other ELL instantiation of every Krylov kernel).  Three SIMPLE
    iterations of a 36 000-cell jittered channel against the oracle on the
    same shuffled mesh: velocity within 1e-4, pressure within 1e-3 (a CG
    stop at ||r|| <= 1e-5 fixes p only to that tolerance)."""
    from oracle.simple import OracleSolver
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    nodes, cells = jittered_box((40, 30, 30), 0.1, 4, lengths=(4.0, 1.0, 1.0))
    cells = cells[np.random.default_rng(4).permutation(len(cells))]
    model = device_case(nodes, cells, 'channel', ref_cell=None, renumber=False)
    assert model.layout.slot_dnb is None and model.dm.slot_dnb is None
    of = oracle_case(nodes, cells, 'channel', ref_cell=None)
    os_ = OracleSolver(of, 0.7, 0.3, nu=0.01, implicit_outflow=True)
    sv = IncompressibleSolver(model, 0.7, 0.3, implicit_outflow=True)
    for _ in range(3):
        os_.outer_iteration()
        sv.outer_iteration()
    got = model.cell_values.cpu().numpy()
    err_u = rel_l2(got[:, :3], of.cell_values[:, :3])
    err_p = rel_l2(got[:, 3], of.cell_values[:, 3])
    print(f'int32 neighbour path: velocity rel-L2 {err_u:.2e}, p {err_p:.2e}')
    assert err_u <= 1e-4 and err_p <= 1e-3
