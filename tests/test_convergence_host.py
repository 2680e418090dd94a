"""The convergence report of check_convergence (model.py:278-315) and the
criteria of Convergence.py:7-106: the oracle's restatement pinned by closed
forms, and the product's Criteria / Convergence mirror against the
reference's rules.  CPU only."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle.simple import convergence_report, convergence_verdicts


def _closed_form_inputs(C=7):
    rng = np.random.default_rng(0)
    u = np.empty(3 * C)
    for k in range(3):
        u[k::3] = k + 1.0                    # distinct per component: a slicing error shows
    vel_corr = np.zeros(3 * C)
    vel_corr[4] = -0.25
    vel_corr[5] = 0.125
    p = rng.standard_normal(C)
    p[3] = -9.0
    div = np.array([1.0, -2.0, 3.0, 0.0, 0.5, -0.5, 0.0])[:C]
    return u, vel_corr, p, div


def test_report_closed_forms():
    C = 7
    u, vel_corr, p, div = _closed_form_inputs(C)
    # A = I, b = 0: ||(A u - b)_k|| = (k + 1) sqrt(C), ||b_k|| = 0
    rep = convergence_report(sp.identity(3 * C, format='csr'), u, np.zeros(3 * C), div,
                             vel_corr, p)
    for k in range(3):
        assert rep['r'][k] == pytest.approx((k + 1) * np.sqrt(C), rel=1e-14)
        assert rep['b'][k] == 0.0
    assert rep['vel_corr'] == 0.25 and rep['p'] == 9.0 and rep['div'] == 7.0
    # A = 2 I, b = u: residual u_k, ||b_k|| = (k + 1) sqrt(C); relative residual 1
    rep = convergence_report(2.0 * sp.identity(3 * C, format='csr'), u, u, div, vel_corr, p)
    for k in range(3):
        assert rep['r'][k] == pytest.approx((k + 1) * np.sqrt(C), rel=1e-14)
        assert rep['b'][k] == pytest.approx((k + 1) * np.sqrt(C), rel=1e-14)
    out, conv = convergence_verdicts(rep)
    assert out['momentum_y'][0] == pytest.approx(1.0) and not conv
    # a matrix coupling x to y only (rows 3c, columns 3c+1): the y values land in r_x
    A = sp.csr_matrix((np.ones(C), (3 * np.arange(C), 3 * np.arange(C) + 1)), shape=(3 * C, 3 * C))
    rep = convergence_report(A, u, np.zeros(3 * C), div, vel_corr, p)
    assert rep['r'] == pytest.approx([2.0 * np.sqrt(C), 0.0, 0.0])


def test_verdict_rules():
    rep = {'r': [1e-9, 2e-9, 0.0], 'b': [1.0, 1.0, 1e-9], 'vel_corr': 5e-5, 'p': 5e-5,
           'div': 5e-5}
    out, conv = convergence_verdicts(rep)
    assert conv
    assert out['momentum_z'] == (0.0, True)          # 0 / max(1e-9, 1e-6)
    rep['div'] = 2e-4                                 # continuity fails
    assert not convergence_verdicts(rep)[1]
    rep['div'] = 5e-5
    rep['r'][2] = 2e-9                                # 2e-9 / 1e-6 = 2e-3 > 1e-3
    assert not convergence_verdicts(rep)[1]


def test_reported_divergence_is_the_pre_correction_flux():
    """The reference passes div_u of the HbyA flux (incompressible.py:131)
    and check_convergence recomputes it from the same mass fluxes: the
    report's continuity equals the L1 norm taken inside the iteration, and
    differs from the corrected flux's."""
    from cases import jittered_box, oracle_case
    from oracle.simple import OracleSolver
    nodes, cells = jittered_box((6, 5, 3), 0.1, 3)
    of = oracle_case(nodes, cells, 'cavity')
    s = OracleSolver(of, 0.7, 0.3, nu=0.01)
    for _ in range(4):
        info = s.outer_iteration()
    rep = s.convergence_report()
    assert rep['div'] == pytest.approx(info['div_hbya_l1'], rel=1e-12)
    assert rep['p'] == pytest.approx(np.abs(s.p_new).max())


def test_criteria_mirror():
    from warpcfd_amd.FV.convergence import Convergence, Criteria
    c = Criteria('m', ord=2., eps=1e-3)
    arr, rhs = np.array([3.0, 4.0]), np.array([0.0, 0.0])
    assert c.check(arr, rhs) == (pytest.approx(5.0 / 1e-6), False)   # ||rhs|| below non_zero
    assert c.check(np.array([1.0, 1.0]), np.array([1.0, 1.0 + 1e-4]))[1]
    assert c.check_value(1e-4, 1.0) == (pytest.approx(1e-4), True)
    a = Criteria('a', tol='abs', eps=1e-4, ord=np.inf)
    assert a.check(np.array([-3e-5, 2e-5])) == (pytest.approx(3e-5), True)
    assert a.check(np.array([-3e-5, 2e-4]))[1] is False
    assert a.check_value(float('nan'))[1] is False
    cv = Convergence()
    assert [x.name for x in cv.criterias] == ['continuity', 'momentum_x', 'momentum_y',
                                              'momentum_z', 'pressure_correction',
                                              'velocity_correction']
    for x in cv.criterias:
        x.is_converged = True
    cv.local_continuity.is_converged = False          # not part of the verdict
    assert cv.has_converged()
    cv.velocity_correction.is_converged = False
    assert not cv.has_converged()


def test_criteria_mirror_agrees_with_oracle_verdicts():
    from warpcfd_amd.FV.convergence import Convergence
    rng = np.random.default_rng(5)
    for _ in range(20):
        rep = {'r': list(10.0 ** rng.uniform(-9, -2, 3)), 'b': list(10.0 ** rng.uniform(-8, 1, 3)),
               'vel_corr': 10.0 ** rng.uniform(-6, -2), 'p': 10.0 ** rng.uniform(-6, -2),
               'div': 10.0 ** rng.uniform(-6, -2)}
        ref, ref_conv = convergence_verdicts(rep)
        cv = Convergence()
        got = {n: getattr(cv, n).check_value(rep['r'][k], rep['b'][k])
               for k, n in enumerate(('momentum_x', 'momentum_y', 'momentum_z'))}
        got['continuity'] = cv.continuity.check_value(rep['div'])
        got['velocity_correction'] = cv.velocity_correction.check_value(rep['vel_corr'])
        got['pressure_correction'] = cv.pressure_correction.check_value(rep['p'])
        for n, v in got.items():
            assert v[0] == pytest.approx(ref[n][0]) and v[1] == ref[n][1], n
        assert cv.has_converged() == ref_conv
