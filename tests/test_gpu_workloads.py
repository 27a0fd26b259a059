"""The batched and served workloads' own generators at sizes the oracle covers.

* Batched uneven instances (domain.batch_problems: edge 32 round(exp(U)), 1-4
  sphere obstacles, a delta right-hand side at a random Active cell, sub_size
  32, vlevels 4), here six problems with edges 64-128: each problem solved in
  one batched context gives the iteration count of its solve alone (+-1) and
  its x within 1e-4 of alone; the edge-64 problems also match the oracle's
  iteration count (+-1) (the oracle's literal 200-sweep sigma bottom takes
  minutes per larger problem).
* The served request stream (domain.served_request: edge log-uniform, up to
  40% blocked by spheres, N(0,1) right-hand side), here four requests with
  edges 32-96 submitted to the serving loop with staggered arrivals: each
  result converges and equals the request solved alone (+-1 iteration, x
  within 1e-4).
"""
import time

import numpy as np
import pytest

from hierarchicaldd_amd import domain
from tests.gpu_common import rel_err

pytestmark = pytest.mark.gpu


def _tol(dom):
    return 1e-6 * float(np.abs(dom.b).max())


def test_batch_problems_workload(oracle_lib):
    from hierarchicaldd_amd import hdd
    probs = domain.batch_problems(6, seed=2, nmax=128)
    print(f"[batch workload] edges {[d.meta['n'] for d, _ in probs]}", flush=True)
    alone = []
    for dom, cfg in probs:
        g = hdd.DDSolver(dom, cfg)
        it, _, _ = g.solve(_tol(dom), 300)
        alone.append((it, g.get(hdd.CH_X)))
        g.close()
    bt = hdd.DDBatch(probs)
    try:
        res = bt.solve_all([_tol(d) for d, _ in probs], 300)
        xb = [bt.x(p) for p in range(len(probs))]
    finally:
        bt.close()
    for p, ((dom, cfg), (it_a, x_a), (it_b, _, _)) in enumerate(zip(probs, alone, res)):
        it_o = oracle_lib.Oracle(dom, cfg).solve(_tol(dom), 300)[0] if dom.meta["n"] <= 64 else None
        d = rel_err(xb[p], x_a)
        print(f"[batch workload {p}] edge {dom.meta['n']}: oracle {it_o}, alone {it_a}, batched {it_b}, x rel diff {d:.2e}", flush=True)
        assert it_a > 0 and it_b > 0
        assert abs(it_b - it_a) <= 1
        assert it_o is None or (it_o > 0 and abs(it_a - it_o) <= 1)
        assert d < 1e-4
    assert sum(d.meta["n"] <= 64 for d, _ in probs) >= 2


def test_served_request_stream():
    from hierarchicaldd_amd import hdd
    rng = np.random.default_rng(5)
    reqs = [domain.served_request(rng, nmin=32, nmax=96) for _ in range(4)]
    print(f"[served stream] edges {[d.meta['n'] for d, _ in reqs]}", flush=True)
    alone = []
    for dom, cfg in reqs:
        g = hdd.DDSolver(dom, cfg)
        it, _, _ = g.solve(_tol(dom), 300)
        alone.append((it, g.get(hdd.CH_X)))
        g.close()
    srv = hdd.Server(setup_threads=2, groups=2, max_batch=2)
    try:
        ids = []
        for dom, cfg in reqs:
            ids.append(srv.submit(dom, cfg, _tol(dom), 300))
            time.sleep(0.05)
        out = [srv.wait(i) for i in ids]
    finally:
        srv.close()
    for k, ((dom, _), (st, x, rep), (it_a, x_a)) in enumerate(zip(reqs, out, alone)):
        d = rel_err(x, x_a)
        print(f"[served stream {k}] edge {dom.meta['n']}: alone {it_a}, served {rep['iterations']} (unit of {rep['batch_size']}, "
              f"{rep['concurrent']} units on the device), x rel diff {d:.2e}", flush=True)
        assert st == 0 and rep["final_norm"] <= _tol(dom)
        assert abs(rep["iterations"] - it_a) <= 1
        assert d < 1e-4
