"""The preconditioner apply replayed from captured HIP graphs against the same
kernels launched one by one from the host: results must be bit-identical,
on every replay (several replays with eager work and synchronisations between
them), for one problem and for a batch (where the replays also carry the
per-problem skip of stopped problems).
"""
import numpy as np
import pytest

from hierarchicaldd_amd import domain
from tests.gpu_common import random_rhs, same_bits

pytestmark = pytest.mark.gpu


def _tol(dom):
    return 1e-6 * float(np.abs(dom.b).max())


def test_apply_replays_match_eager():
    from hierarchicaldd_amd import hdd
    dom, cfg = domain.config0()
    g = hdd.DDSolver(dom, cfg)
    rs = [random_rhs(g.level(0)[1], s) for s in range(3)]
    g.set_graph(False)
    eager = [g.apply_preconditioner(r) for r in rs]
    g.set_graph(True)
    for rep in range(3):
        for r, ze in zip(rs, eager):
            zg = g.apply_preconditioner(r)
            assert same_bits(zg, ze), f"replay {rep} differs"
        g.set_graph(False)
        assert same_bits(g.apply_preconditioner(rs[0]), eager[0])  # eager work between replays
        g.set_graph(True)
    g.close()


def test_pcg_graph_matches_eager():
    from hierarchicaldd_amd import hdd
    dom, cfg = domain.small_box(48, 16, 3, seed=7)
    g = hdd.DDSolver(dom, cfg)
    g.set_graph(False)
    it_e, h_e, _ = g.solve(_tol(dom), 300)
    x_e = g.get(hdd.CH_X)
    g.set_graph(True)
    for _ in range(2):
        it_g, h_g, _ = g.solve(_tol(dom), 300)
        assert it_g == it_e and np.array_equal(h_g, h_e)
        assert same_bits(g.get(hdd.CH_X), x_e)
    g.close()


def test_batch_pcg_graph_matches_eager():
    from hierarchicaldd_amd import hdd
    probs = [domain.small_box(32, 16, 3, seed=1), domain.small_box(48, 16, 3, seed=2), domain.config0()]
    bt = hdd.DDBatch(probs)
    tol = [_tol(d) for d, _ in probs]
    bt.set_graph(False)
    eager = bt.solve_all(tol, 300)
    xe = [bt.x(p) for p in range(len(probs))]
    bt.set_graph(True)
    for _ in range(2):
        res = bt.solve_all(tol, 300)
        for p in range(len(probs)):
            assert res[p][0] == eager[p][0] and np.array_equal(res[p][1], eager[p][1])
            assert same_bits(bt.x(p), xe[p])
    bt.close()
