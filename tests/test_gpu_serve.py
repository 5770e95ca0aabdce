"""Serving loop (continuous batching) against the oracle and against itself:
every result equals the same request evaluated alone, whatever else is in
flight. Tolerances: 1e-10 against the oracle (relative to max(1, |value|));
alone vs in the mix 1e-12 (the per-request reductions do not depend on the
batch they run in).
"""
import time

import numpy as np
import pytest

import oracle
from circuits import chain_circuit
import vqchip as vq
from vqchip import _lib
from vqchip.serve import Request, Server, plain_eval, poisson_arrivals, run_server_open_loop
from vqchip.workloads import config4_request, hea_circuit, tfim

pytestmark = pytest.mark.gpu


def oracle_eval(req):
    psi0 = np.zeros(1 << req.n, dtype=complex)
    psi0[0] = 1
    E, g, _ = oracle.energy_grad(psi0, req.n, 0, req.ops, req.nops, req.pool, req.terms, req.nterms, req.hpool,
                                 req.nparams)
    return E, g


def close(a, b, tol):
    return abs(a - b) <= tol * max(1.0, abs(b))


def make_requests(seed, ns, depth_max=6):
    rng = np.random.default_rng(seed)
    out = []
    for n in ns:
        d = int(rng.integers(1, depth_max + 1))
        c = hea_circuit(n, d, rng.uniform(0, 2 * np.pi, 3 * n * d))
        out.append(Request(c, tfim(n, float(rng.uniform(0.5, 1.5))), n))
    return out


@pytest.mark.parametrize("solo_qubits", [19, 13])
def test_served_results_match_oracle_and_alone(solo_qubits):
    # solo_qubits=13: every request above 12 qubits runs alone through the tiled phase kernels
    reqs = make_requests(11, [5, 13, 8, 16, 10, 12, 14, 6, 15, 9, 11, 13])
    rng = np.random.default_rng(12)
    reqs.append(Request(chain_circuit(12, rng), tfim(12, 0.7), 12))
    reqs.append(Request(chain_circuit(15, rng), tfim(15, 1.1), 15))
    with Server(hbm_budget=8 << 30, lanes=3, max_batch=8, solo_qubits=solo_qubits) as srv:
        tickets = [srv.submit(r) for r in reqs]
        mix = [srv.wait(t) for t in tickets]
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["pending"] == 0 and st["running"] == 0
        alone = []
        for r in reqs:  # one at a time: each request is its own batch
            alone.append(srv.wait(srv.submit(r)))
    for r, (E, g, tm), (E1, g1, _) in zip(reqs, mix, alone):
        assert tm[0] <= tm[1] <= tm[2]
        assert abs(E - E1) <= 1e-12 * max(1, abs(E1)) and np.max(np.abs(g - g1)) <= 1e-12 * max(1, np.max(np.abs(g1)))
        if r.n <= 14:
            Er, gr = oracle_eval(r)
            assert close(E, Er, 1e-10), r.n
            assert np.max(np.abs(g - gr)) <= 1e-10 * max(1, np.max(np.abs(gr))), r.n
        else:
            Ep, gp = plain_eval(r, fused=False)
            assert close(E, Ep, 1e-10) and np.max(np.abs(g - gp)) <= 1e-10 * max(1, np.max(np.abs(gp)))


def test_requests_arriving_while_others_run_overlap():
    # a long request first, short ones arriving during it on other lanes
    big = Request(hea_circuit(20, 8, np.linspace(0, 6, 480)), tfim(20), 20)
    smalls = make_requests(21, [8, 10, 12, 9, 11, 13, 10, 8])
    with Server(hbm_budget=8 << 30, lanes=3, max_batch=4, grid_frac=0.75) as srv:
        tb = srv.submit(big)
        time.sleep(0.002)
        ts = [srv.submit(r) for r in smalls]
        res_small = [srv.wait(t) for t in ts]
        Eb, gb, tmb = srv.wait(tb)
        st = srv.stats()
    assert st["batches"] >= 2
    Ep, gp = plain_eval(big, fused=True)
    assert close(Eb, Ep, 1e-10) and np.max(np.abs(gb - gp)) <= 1e-10 * max(1, np.max(np.abs(gp)))
    for r, (E, g, tm) in zip(smalls, res_small):
        Er, gr = oracle_eval(r)
        assert close(E, Er, 1e-10) and np.max(np.abs(g - gr)) <= 1e-10 * max(1, np.max(np.abs(gr)))


def test_admission_by_memory_budget():
    # 32 MiB of states per n=20 request; a 96 MiB budget admits at most two at once
    reqs = make_requests(31, [20, 20, 20, 20, 20], depth_max=2)
    with Server(hbm_budget=96 << 20, lanes=3, max_batch=8) as srv:
        ts = [srv.submit(r) for r in reqs]
        res = [srv.wait(t) for t in ts]
        st = srv.stats()
        assert st["peak_live_bytes"] <= st["budget_bytes"]
        assert st["batches"] >= 3
        huge = make_requests(32, [23], depth_max=1)[0]  # 256 MiB of states: never admissible
        with pytest.raises(vq.VQCError):
            srv.wait(srv.submit(huge))
    for r, (E, g, _) in zip(reqs, res):
        Ep, gp = plain_eval(r, fused=True)
        assert close(E, Ep, 1e-10) and np.max(np.abs(g - gp)) <= 1e-10 * max(1, np.max(np.abs(gp)))


def test_open_loop_poisson_stream():
    rng = np.random.default_rng(41)
    reqs = [Request(*config4_request(rng, nmin=6, nmax=16, dmin=2, dmax=8)) for _ in range(40)]
    arr = poisson_arrivals(400.0, len(reqs), rng)
    with Server(hbm_budget=8 << 30) as srv:
        lat, span, results = run_server_open_loop(srv, reqs, arr)
    assert np.all(lat > 0) and len(results) == len(reqs)
    for r, (E, g) in zip(reqs[:10], results[:10]):
        Er, gr = oracle_eval(r)
        assert close(E, Er, 1e-10) and np.max(np.abs(g - gr)) <= 1e-10 * max(1, np.max(np.abs(gr)))


def test_bad_request_is_rejected():
    bad = make_requests(51, [6])[0]
    bad.inst.nqubits = 0
    with Server(hbm_budget=1 << 30) as srv:
        with pytest.raises(vq.VQCError):
            srv.submit(bad)


def test_submissions_racing_the_idle_scheduler():
    # a request queued while the scheduler finishes an idle pass must still run
    # (regression: it used to be failed as "does not fit" without being tried)
    reqs = make_requests(61, [4, 5, 6], depth_max=1)
    rng = np.random.default_rng(62)
    with Server(hbm_budget=1 << 30, lanes=2) as srv:
        for i in range(300):
            r = reqs[i % 3]
            E, g, _ = srv.wait(srv.submit(r))
            if i < 3:
                Er, gr = oracle_eval(r)
                assert close(E, Er, 1e-10)
            time.sleep(float(rng.uniform(0, 2e-4)))
        assert srv.stats()["completed"] == 300
