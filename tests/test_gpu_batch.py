"""Batched SIMPLE in a device arena (warpcfd_amd.batch) on the GPU.

* a case converged inside a batch matches the float64 oracle fixture;
* a case's result is bitwise the same alone and in a mix, at another place
  in the arena, and when other cases are admitted and released around it
  while it runs (continuous batching);
* the device-side convergence monitor stops every case within 1e-4 of the
  float64 oracle's fixed point.
"""
import os

import numpy as np
import pytest

from cases import apply_bcs, rel_l2

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def cavity_spec(n, re=100.0, stop=1e-6, max_outer=3000):
    from warpcfd_amd.FV import FVM
    from warpcfd_amd.batch import CaseSpec
    from warpcfd_amd.preprocess import Mesh, cavity_mesh, define_boundary_walls
    nodes, cells = cavity_mesh(n)
    m = Mesh(nodes, cells)
    define_boundary_walls(m)
    model = FVM(m)
    model.material.create_incompressible_newtonian_fluid(viscosity=1.0 / re, density=1.0)
    model.set_reference_pressure(0, 0.0)
    apply_bcs(model.boundary, 'cavity')
    return CaseSpec(model, stop=stop, max_outer=max_outer, name=f'cavity{n}_re{int(re)}')


def channel_spec(stop=1e-6, max_outer=3000):
    from warpcfd_amd.FV import FVM
    from warpcfd_amd.batch import CaseSpec
    from warpcfd_amd.preprocess import Mesh, define_boundary_walls, structured_hex
    nodes, cells = structured_hex((24, 6, 6), (4.0, 1.0, 1.0), jitter=0.1, seed=3)
    m = Mesh(nodes, cells)
    define_boundary_walls(m)
    model = FVM(m)
    model.material.create_incompressible_newtonian_fluid(viscosity=0.02, density=1.0)
    apply_bcs(model.boundary, 'channel')
    return CaseSpec(model, stop=stop, max_outer=max_outer, name='channel')


KW = dict(chunk=128, block=128, atol=1e-9, tol=1e-9, maxiter=2000)


def _run(arena, max_steps=3000):
    n = arena.run(max_steps=max_steps)
    assert n < max_steps
    return n


def test_batch_case_matches_fp64_oracle(gpu):
    """24x24 cavity (Re 100) converged inside a batch of three: every case
    stopped by the device monitor, and the 24x24 case within 1e-4 of the
    float64 oracle's fixed point."""
    from warpcfd_amd.batch import CaseArena
    z = np.load(os.path.join(G, 'cavity24_converged.npz'))
    specs = [cavity_spec(16, 400.0), cavity_spec(24, 100.0), cavity_spec(20, 1000.0)]
    a = CaseArena.for_cases(specs, **KW)
    _run(a)
    for k in range(3):
        assert a.converged(k), (k, a.continuity(k))
        assert a.iterations(k) > 10
    got = a.cell_values(1)
    for q in (0, 1):
        assert rel_l2(got[:, q], z['f64'][:, q]) <= 1e-4, q


def test_alone_vs_mix_bitwise(gpu):
    """The same case alone, and third in a mix with a cavity and a channel
    (other offsets, other neighbours): bitwise equal fields and equal
    iteration counts."""
    from warpcfd_amd.batch import CaseArena
    x = cavity_spec(20, 400.0, stop=1e-6)
    alone = CaseArena.for_cases([x], **KW)
    _run(alone)
    mix = CaseArena.for_cases([cavity_spec(12, 800.0), channel_spec(), x], **KW)
    _run(mix)
    assert alone.iterations(0) == mix.iterations(2)
    assert np.array_equal(alone.cell_values(0), mix.cell_values(2))
    assert alone.converged(0) and mix.converged(2)


def test_continuous_admission_bitwise(gpu):
    """Cases admitted and released while the watched case runs (it starts
    beside one case; after 5 iterations that case leaves and two others take
    free space, one before it in the arena): the watched case's result is
    bitwise its result alone."""
    from warpcfd_amd.batch import CaseArena
    x = cavity_spec(16, 250.0, stop=1e-6)
    alone = CaseArena.for_cases([x], **KW)
    _run(alone)
    y, w1, w2 = cavity_spec(12, 100.0), cavity_spec(8, 300.0), cavity_spec(10, 600.0)
    cap = y.n_cells + x.n_cells + w2.n_cells
    fi = y.n_internal + x.n_internal + w2.n_internal
    fb = y.n_boundary + x.n_boundary + w2.n_boundary
    a = CaseArena(cap, fi, fb, 4, **KW)
    ky = a.admit(y)
    kx = a.admit(x)
    for _ in range(5):
        a.step()
    a.release(ky)
    k1 = a.admit(w1)     # first fit: the space y left, before x
    k2 = a.admit(w2)
    assert None not in (k1, k2)
    assert a.slots[k1]['c0'] < a.slots[kx]['c0']
    _run(a)
    assert a.iterations(kx) == alone.iterations(0)
    assert np.array_equal(a.cell_values(kx), alone.cell_values(0))
    for k in (k1, k2):
        assert a.converged(k)


def test_stop_policy_velocity_error_vs_fp64(gpu):
    """The batch policy (Krylov stop at 0.1 ||r_0||, device monitor stop at
    continuity 1e-6 and estimated velocity error 5e-5) leaves the 24x24 Re 100
    cavity within 1e-4 of the float64 oracle's fixed point."""
    from warpcfd_amd.batch import CaseArena
    z = np.load(os.path.join(G, 'cavity24_converged.npz'))
    a = CaseArena.for_cases([cavity_spec(24, 100.0, stop=1e-6)], chunk=256, block=256, atol=0.0,
                            tol=0.0, rtol0=0.1, maxiter=2000)
    _run(a)
    assert a.converged(0)
    got = a.cell_values(0)
    err = [rel_l2(got[:, q], z['f64'][:, q]) for q in (0, 1, 3)]
    print('rel-L2 vs float64 oracle (u, v, p):', err)
    assert max(err[:2]) <= 1e-4
