"""Why the channel, cylinder and served workloads use the implicit convective
outflow (WCFD_SCHEME_IMPLICIT_OUTFLOW) instead of the reference's lagged
boundary convection (convection.py:63-67).

On the benchmark channel's geometry (box 10 x 2 x 2, inlet u = 1, outlet
p = 0, no-slip walls, nu = 0.01, interior vertices jittered +-10 % h from
seed 0) at 40 x 8 x 8 cells, SIMPLE with the reference's scheme overflows
to NaN within a few outer iterations, while the implicit outflow, the same
scheme with the outlet's convective flux moved to the matrix diagonal,
converges: the outlet rows regain diagonal dominance.  Both runs are the
oracle (the CPU restatement) with the reference's settings otherwise.
"""
import numpy as np

from cases import oracle_case
from oracle.simple import OracleSolver
from warpcfd_amd.preprocess.grid import structured_hex


def _run(implicit_outflow, iters):
    nodes, cells = structured_hex((40, 8, 8), (10.0, 2.0, 2.0), jitter=0.1, seed=0)
    f = oracle_case(nodes, cells, 'channel', ref_cell=None)
    s = OracleSolver(f, 0.7, 0.3, nu=0.01, implicit_outflow=implicit_outflow)
    cont = []
    with np.errstate(all='ignore'):
        for i in range(iters):
            cont.append(s.outer_iteration()['continuity'])
            if not np.isfinite(f.cell_values).all():
                return i + 1, cont, f
    return None, cont, f


def test_reference_outflow_diverges_on_the_channel():
    blew_up_at, _, _ = _run(False, 15)
    assert blew_up_at is not None and blew_up_at <= 12


def test_implicit_outflow_converges_on_the_same_channel():
    blew_up_at, cont, f = _run(True, 40)
    assert blew_up_at is None
    assert cont[-1] < 5e-3 * cont[0]                      # 8.7e-2 -> 2.0e-4
    assert all(b < a for a, b in zip(cont[8::8], cont[16::8]))   # decreasing
    u = f.cell_values[:, 0]
    assert 1.0 < u.max() < 2.0                             # developing profile
