"""Batched DD-PCG: several independent problems of different sizes in one context.

Every V-cycle stage, interface solve and PCG vector operation is one launch for
all problems of the batch; each problem keeps its own PCG scalars and stopping
test. Bar: each problem's result in the batch is the result it gets alone,
within float rounding of the interface solves (whose launch shapes depend on
the batch):
  * iteration count within +-1 of the same problem solved alone and of the oracle;
  * residual history (max|r| per iteration) within 1e-3 relative or 1e-7 of
    the initial max|r| of the solve alone (common prefix): in the last
    iterations max|r| is ~1e-6 of its start, where float rounding of the
    interface solves moves it by up to a few percent;
  * solution within 1e-4 relative (max-norm) of the solve alone;
and two copies of one problem in a batch agree to 1e-6.
Edge cases in the batch: a zero right-hand side (stops at iteration 0 with
x = 0), a problem with one subdomain (no interface dofs), sizes that are not a
multiple of sub_size.
"""
import numpy as np
import pytest

from hierarchicaldd_amd import domain
from tests.gpu_common import rel_err

pytestmark = pytest.mark.gpu


def _zero_rhs():
    d, cfg = domain.small_box(40, 16, 3, seed=3)
    d.b[:] = 0
    return d, cfg


CASES = [
    ("box32_sub16", lambda: domain.small_box(32, 16, 3, seed=1)),
    ("box48_sub16", lambda: domain.small_box(48, 16, 3, seed=2)),
    ("sphere64_sub16", domain.config0),
    ("box40_zero_rhs", _zero_rhs),
    ("box32_sub32_no_interface", lambda: domain.small_box(32, 32, 3, seed=4)),
]


def _tol(dom):
    m = float(np.abs(dom.b).max())
    return 1e-6 * m if m > 0 else 1e-6


@pytest.fixture(scope="module")
def runs():
    from hierarchicaldd_amd import hdd
    probs = [f() for _, f in CASES]
    alone = []
    for (name, _), (dom, cfg) in zip(CASES, probs):
        print(f"[alone {name}] pages={dom.n_pages}", flush=True)
        g = hdd.DDSolver(dom, cfg)
        it, h, secs = g.solve(_tol(dom), 300)
        alone.append((it, h, g.get(hdd.CH_X), secs))
        g.close()
    print(f"[batch] {len(probs)} problems", flush=True)
    bt = hdd.DDBatch(probs)
    res = bt.solve_all([_tol(d) for d, _ in probs], 300)
    xs = [bt.x(p) for p in range(len(probs))]
    offs = [bt.problem_page_offsets(p) for p in range(len(probs))]
    bt.close()
    return probs, alone, res, xs, offs


def test_batch_layout(runs):
    probs, _, _, xs, offs = runs
    for (dom, _), x, off in zip(probs, xs, offs):
        assert x.shape == (dom.n_pages, 256)
        assert np.array_equal(np.sort(off), np.sort(dom.page_offsets))


@pytest.mark.parametrize("k", range(len(CASES)))
def test_batch_matches_alone(runs, k):
    probs, alone, res, xs, _ = runs
    name = CASES[k][0]
    it_a, h_a, x_a, _ = alone[k]
    it_b, h_b, _ = res[k]
    x_b = xs[k]
    n = min(len(h_a), len(h_b))
    diff = np.abs(h_b[:n] - h_a[:n])
    hrel = float((diff / np.maximum(h_a[:n], 1e-30)).max()) if n else 0.0
    ok = diff <= np.maximum(1e-3 * h_a[:n], 1e-7 * h_a[0]) if n else np.array([True])
    print(f"[{name}] alone {it_a} iterations, batch {it_b}; history rel diff {hrel:.2e}, x rel diff {rel_err(x_b, x_a):.2e}")
    assert it_b >= 0 and abs(it_b - it_a) <= 1
    assert ok.all()
    if np.abs(x_a).max() == 0:
        assert np.abs(x_b).max() == 0
    else:
        assert rel_err(x_b, x_a) < 1e-4
    assert h_b[-1] <= _tol(probs[k][0])


def test_zero_rhs_stops_at_once(runs):
    _, alone, res, xs, _ = runs
    k = [c[0] for c in CASES].index("box40_zero_rhs")
    assert res[k][0] == 0 and alone[k][0] == 0
    assert not np.any(xs[k])


def test_batch_against_oracle(runs, oracle_lib):
    probs, _, res, _, _ = runs
    for (name, _), (dom, cfg), (it_b, _, _) in zip(CASES, probs, res):
        o = oracle_lib.Oracle(dom, cfg)
        it_o, _, _ = o.solve(_tol(dom), 300)
        print(f"[{name}] oracle {it_o} iterations, batch {it_b}")
        assert abs(it_b - it_o) <= 1


def test_two_copies_agree():
    from hierarchicaldd_amd import hdd
    a = domain.small_box(32, 16, 3, seed=5)
    b = domain.config0()
    bt = hdd.DDBatch([a, b, a])
    tol = [_tol(a[0]), _tol(b[0]), _tol(a[0])]
    res = bt.solve_all(tol, 300)
    x0, x2 = bt.x(0), bt.x(2)
    bt.close()
    print(f"[copies] iterations {res[0][0]} / {res[2][0]}, x rel diff {rel_err(x2, x0):.2e}")
    assert res[0][0] == res[2][0]
    assert rel_err(x2, x0) < 1e-6


def test_contexts_solved_at_once_match_alone():
    """hdd_batches_pcg_solve: two single-problem contexts and a batch iterate at once, each on its
    own stream; every problem gives the iterations and the bits of its own context's solve."""
    from hierarchicaldd_amd import hdd
    from tests.gpu_common import same_bits
    p1, p2 = domain.small_box(32, 16, 3, seed=51), domain.config0()
    p3, p4 = domain.small_box(40, 16, 3, seed=52), domain.small_box(48, 16, 3, seed=53)
    ctxs = [hdd.DDSolver(*p1), hdd.DDSolver(*p2), hdd.DDBatch([p3, p4])]
    tol = [_tol(p[0]) for p in (p1, p2, p3, p4)]
    try:
        alone = [ctxs[0].solve(tol[0], 300)[0], ctxs[1].solve(tol[1], 300)[0]] + [r[0] for r in ctxs[2].solve_all(tol[2:], 300)]
        xa = [ctxs[0].get(hdd.CH_X), ctxs[1].get(hdd.CH_X), ctxs[2].x(0), ctxs[2].x(1)]
        res = hdd.solve_concurrently(ctxs, tol, 300)
        xc = [ctxs[0].get(hdd.CH_X), ctxs[1].get(hdd.CH_X), ctxs[2].x(0), ctxs[2].x(1)]
    finally:
        for g in ctxs:
            g.close()
    print(f"[at once] iterations alone {alone}, at once {[r[0] for r in res]}", flush=True)
    assert [r[0] for r in res] == alone
    assert all(r[2] for r in res)
    assert all(same_bits(a, b) for a, b in zip(xa, xc))
