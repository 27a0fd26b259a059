"""Serving loop: requests of different sizes submitted one after another, set up
on host workers, admitted to the GPU as units that iterate side by side
(continuous batching: a unit admitted while others are mid-solve joins them
at their next iteration).

Bar: each request's result in the mix is the result it gets solved alone
(tolerances as for batched solves: iteration count within +-1, solution
within 1e-4 relative in max-norm; a unit of one request gives the same bits
as the solve alone), and every result meets its own tolerance. Also: a
request that joins while another is mid-solve, cancellation of a running and
of a queued request, a request larger than the HBM budget rejected with an
error, and requests with different V-cycle schedules solved in separate units.
"""
import time

import numpy as np
import pytest

from hierarchicaldd_amd import domain
from tests.gpu_common import rel_err

pytestmark = pytest.mark.gpu


def _tol(dom):
    return 1e-6 * float(np.abs(dom.b).max())


def _requests():
    return [domain.small_box(32, 16, 3, seed=11), domain.small_box(48, 16, 3, seed=12), domain.config0(),
            domain.small_box(40, 16, 3, seed=13), domain.small_box(32, 32, 3, seed=14), domain.small_box(56, 16, 3, seed=15)]


def test_server_results_match_alone():
    from hierarchicaldd_amd import hdd
    reqs = _requests()
    alone = []
    for dom, cfg in reqs:
        g = hdd.DDSolver(dom, cfg)
        it, _, _ = g.solve(_tol(dom), 300)
        alone.append((it, g.get(hdd.CH_X)))
        g.close()
    srv = hdd.Server(setup_threads=2, groups=2, max_batch=3)
    ids = []
    for dom, cfg in reqs:
        ids.append(srv.submit(dom, cfg, _tol(dom), 300))
        time.sleep(0.02)
    out = [srv.wait(i) for i in ids]
    srv.close()
    for k, ((dom, _), (st, x, rep), (it_a, x_a)) in enumerate(zip(reqs, out, alone)):
        # the server returns x in the submitted page order; the solve alone in Morton order (same for these domains)
        print(f"[request {k}] pages={dom.n_pages} alone {it_a} served {rep['iterations']} (group of {rep['batch_size']}), "
              f"latency {rep['latency_s']:.3f}s setup {rep['setup_s']:.3f}s queue {rep['queue_s']:.3f}s solve {rep['solve_s']:.3f}s, "
              f"x rel diff {rel_err(x, x_a):.2e}", flush=True)
        assert st == 0
        assert abs(rep["iterations"] - it_a) <= 1
        assert rel_err(x, x_a) < 1e-4
        assert rep["final_norm"] <= _tol(dom)
        assert rep["latency_s"] >= rep["solve_s"] >= 0


def test_server_input_page_order():
    """x comes back in the page order of the submitted desc, whatever that order is."""
    from hierarchicaldd_amd import hdd
    dom, cfg = domain.small_box(32, 16, 3, seed=21)
    perm = np.random.default_rng(0).permutation(dom.n_pages)
    shuffled = domain.SPGridDomain(dom.size, dom.page_offsets[perm].copy(), dom.flags[perm].copy(), dom.b[perm].copy())
    srv = hdd.Server(setup_threads=1, groups=1, max_batch=2)
    a = srv.submit(dom, cfg, _tol(dom), 300)
    b = srv.submit(shuffled, cfg, _tol(dom), 300)
    _, xa, _ = srv.wait(a)
    _, xb, _ = srv.wait(b)
    srv.close()
    assert rel_err(xb, xa[perm]) < 1e-4


def test_server_rejects_oversized():
    from hierarchicaldd_amd import hdd
    dom, cfg = domain.config0()
    srv = hdd.Server(setup_threads=1, groups=1, max_batch=2, hbm_budget=1e6)
    rid = srv.submit(dom, cfg, _tol(dom), 300)
    with pytest.raises(hdd.HddError, match="HBM budget"):
        srv.wait(rid)
    srv.close()


def test_server_mixed_schedules():
    from hierarchicaldd_amd import hdd
    a = domain.small_box(32, 16, 3, seed=31)
    b = domain.small_box(32, 16, 2, seed=32)
    srv = hdd.Server(setup_threads=2, groups=1, max_batch=4)
    ia = srv.submit(*a, _tol(a[0]), 300)
    ib = srv.submit(*b, _tol(b[0]), 300)
    sa, _, ra = srv.wait(ia)
    sb, _, rb = srv.wait(ib)
    srv.close()
    assert sa == 0 and sb == 0
    assert ra["batch_size"] == 1 and rb["batch_size"] == 1


def test_server_continuous_batching_and_cancel():
    """Two requests that never converge (a negative tolerance: the recursive residual of a
    float PCG keeps falling and would pass any tolerance >= 0) hold both device slots; probe
    requests submitted meanwhile queue, and one cancelled while queued never runs.
    Then one long request is cancelled mid-flight: the probes join the other one
    mid-solve (concurrent units >= 2) and each gives the bits of its solve alone."""
    from hierarchicaldd_amd import hdd
    from tests.gpu_common import same_bits
    long_dom, long_cfg = domain.small_box(48, 16, 3, seed=41)
    probes = [domain.small_box(32, 16, 3, seed=42 + k) for k in range(3)]
    alone = []
    for dom, cfg in probes:
        g = hdd.DDSolver(dom, cfg)
        it, _, _ = g.solve(_tol(dom), 300)
        alone.append((it, g.get(hdd.CH_X)))
        g.close()
    srv = hdd.Server(setup_threads=2, groups=2, max_batch=1)
    try:
        long_a = srv.submit(long_dom, long_cfg, -1.0, 1000000)
        long_b = srv.submit(long_dom, long_cfg, -1.0, 1000000)
        time.sleep(2.0)  # both long requests set up and iterating
        queued = srv.submit(*probes[0], _tol(probes[0][0]), 300)
        assert srv.cancel(queued)
        st_q, x_q, rep_q = srv.wait(queued)
        print(f"[queued request] status {st_q} iterations {rep_q['iterations']} concurrent {rep_q['concurrent']}", flush=True)
        assert st_q == 5 and rep_q["cancelled"] == 1
        assert rep_q["solve_s"] == 0.0 and not x_q.any()
        # free one slot mid-flight: the probes are admitted one unit each beside long_b
        assert srv.cancel(long_a)
        st_a, _, rep_a = srv.wait(long_a)
        print(f"[long request a] status {st_a} iterations {rep_a['iterations']} solve {rep_a['solve_s']:.3f}s", flush=True)
        assert st_a == 5 and rep_a["cancelled"] == 1 and rep_a["iterations"] < 0
        ids = [srv.submit(dom, cfg, _tol(dom), 300) for dom, cfg in probes]
        out = [srv.wait(i) for i in ids]
        assert srv.cancel(long_b)
        st_b, _, rep_b = srv.wait(long_b)
        assert st_b == 5 and rep_b["iterations"] < 0
    finally:
        # requests that never converge must not outlive the test (close() finishes every request)
        srv.cancel(long_a)
        srv.cancel(long_b)
        srv.close()
    joined = 0
    for k, ((dom, _), (st, x, rep), (it_a, x_a)) in enumerate(zip(probes, out, alone)):
        bits = same_bits(x, x_a)
        print(f"[probe {k}] alone {it_a} served {rep['iterations']}, concurrent units {rep['concurrent']}, "
              f"bit-identical {bits}, x rel diff {rel_err(x, x_a):.2e}", flush=True)
        assert st == 0 and rep["iterations"] == it_a
        assert bits
        joined += rep["concurrent"] >= 2
    # long_b was mid-solve (it never converges) while every probe ran
    assert joined == len(probes)


def test_served_request_bits_on_a_worker_share():
    """A request set up by one of four set-up workers (each on a quarter of the host's
    threads) gives the bits of its solve alone (set up on all of them): the set-up does
    not depend on the thread count, down to the bottom spectrum's block Lanczos sums and
    which factorisation fronts run on the GPU."""
    from hierarchicaldd_amd import hdd
    from tests.gpu_common import same_bits
    dom, cfg = domain.sphere_case(96, 32, 4, delta=(32, 32, 32))
    g = hdd.DDSolver(dom, cfg)
    try:
        it_a, _, _ = g.solve(_tol(dom), 300)
        x_a = g.get(hdd.CH_X)
    finally:
        g.close()
    srv = hdd.Server(setup_threads=4, groups=1, max_batch=1)
    try:
        st, x, rep = srv.wait(srv.submit(dom, cfg, _tol(dom), 300))
    finally:
        srv.close()
    print(f"[worker share] alone {it_a} served {rep['iterations']}, bit-identical {same_bits(x, x_a)}", flush=True)
    assert st == 0 and rep["iterations"] == it_a
    assert same_bits(x, x_a)
