"""HIP-graph replay of a batched SIMPLE iteration (CaseArena.capture):
every replay is checked against the same iteration run eagerly, with the
program's own eager work and synchronisations between replays; with a
launch budget that covers every solve, the graph reproduces the polled
(adaptive) path bitwise."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

KW = dict(chunk=128, block=128, atol=0.0, tol=0.0, rtol0=0.1, maxiter=2000, check_every_u=2)


def _arena():
    from warpcfd_amd.batch import CaseArena, CaseSpec
    from warpcfd_amd.workloads import cavity_model
    specs = [CaseSpec(cavity_model(n, re), stop=0.0) for n, re in ((16, 200.0), (12, 700.0),
                                                                  (20, 400.0))]
    return CaseArena.for_cases(specs, **KW)


def test_graph_replays_match_eager_fixed_budget(gpu):
    budget = (3, 4)
    eager, graph = _arena(), _arena()
    eager.step(budget=budget)
    graph.step(budget=budget)
    graph.capture(budget)
    junk = torch.zeros(1 << 20, device='cuda')
    for i in range(8):
        eager.step(budget=budget)
        graph.replay()
        junk.add_(1.0)                      # eager work between replays
        torch.cuda.synchronize()
        for k in range(3):
            assert np.array_equal(eager.cell_values(k), graph.cell_values(k)), (i, k)


def test_graph_with_sufficient_budget_equals_polled_path(gpu):
    polled, graph = _arena(), _arena()
    for _ in range(3):
        polled.step()
    budget = (400, 200)                     # more groups than any solve needs: no truncation
    graph.step(budget=budget)
    graph.capture(budget)
    graph.replay()
    graph.replay()
    torch.cuda.synchronize()
    for k in range(3):
        assert np.array_equal(polled.cell_values(k), graph.cell_values(k)), k
