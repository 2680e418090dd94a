"""A case read from a mesh file runs like the same case built from arrays and
like the oracle on the file's points and cells, and its exported fields read
back unchanged (GPU)."""
import numpy as np
import pytest

from cases import apply_bcs

pytestmark = pytest.mark.gpu


def _solve(mesh_src, iters=15):
    from warpcfd_amd.FV import FVM
    from warpcfd_amd.FV.implicit_Solvers import IncompressibleSolver
    from warpcfd_amd.preprocess import Mesh, define_boundary_walls
    m = Mesh(mesh_src) if not isinstance(mesh_src, tuple) else Mesh(*mesh_src)
    define_boundary_walls(m)
    model = FVM(m)
    model.material.create_incompressible_newtonian_fluid(viscosity=0.01, density=1.0)
    model.set_reference_pressure(0, 0.0)
    apply_bcs(model.boundary, 'cavity')
    model.finalize()
    sv = IncompressibleSolver(model, 0.7, 0.3)
    for _ in range(iters):
        sv.outer_iteration()
    return model


@pytest.mark.parametrize('fmt', ['msh', 'vtu'])
def test_case_from_file_runs_bitwise_and_exports(gpu, tmp_path, fmt):
    from warpcfd_amd.postprocess import export_vtu, sample_line
    from warpcfd_amd.preprocess.grid import cavity_mesh
    from warpcfd_amd.preprocess.io import read_msh, read_vtu, write_msh, write_vtu
    nodes, cells = cavity_mesh(20)
    p = str(tmp_path / f'cavity.{fmt}')
    if fmt == 'msh':
        write_msh(p, nodes, cells, version=4.1)
        d = read_msh(p)
    else:
        write_vtu(p, nodes, cells)
        d = read_vtu(p)
    a = _solve((nodes, cells))
    b = _solve(d)
    va, vb = a.cell_values.cpu().numpy(), b.cell_values.cpu().numpy()
    assert np.array_equal(va, vb)
    # the oracle (the reference's algorithm, float64 numpy) on what the file holds
    from oracle.simple import OracleSolver
    from cases import oracle_case, rel_l2
    of = oracle_case(d.points, d.cells, 'cavity')
    os_ = OracleSolver(of, 0.7, 0.3, nu=0.01)
    for _ in range(15):
        os_.outer_iteration()
    err_u = rel_l2(vb[:, :3], of.cell_values[:, :3])
    err_p = rel_l2(vb[:, 3], of.cell_values[:, 3])
    print(f'{fmt}: 15 SIMPLE iterations, velocity rel-L2 vs oracle {err_u:.2e}, p {err_p:.2e}')
    assert err_u <= 1e-4 and err_p <= 1e-4
    out = str(tmp_path / 'result.vtu')
    export_vtu(b, out)
    r = read_vtu(out)
    for i, name in enumerate('uvwp'):
        np.testing.assert_array_equal(r.cell_data[name], vb[:, i])
    np.testing.assert_array_equal(r.cell_data['velocity'], vb[:, :3])
    # the vertical centreline of the cavity, sampled from the exported fields
    _, s, ids = sample_line(b.mesh, r.cell_data['u'], (0.5 + 1e-3, 0.0, 0.05),
                            (0.5 + 1e-3, 1.0, 0.05), 20)
    assert np.all(ids >= 0) and np.isfinite(s).all() and s.max() > 0.0
