"""Single-state entry points called from several host threads at once. ctypes
releases the GIL, so the C calls overlap. Each host thread runs its calls on its
own stream with its own device buffers; the library lock covers host work
(planning, the plan cache, launching) and the fused evaluations release it while
they wait for their results, so different threads' GPU work overlaps. Every
thread's results must equal a serial run exactly (same kernels, same reduction
order)."""
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from brute import rand_state
import vqchip as vq
from vqchip.workloads import hea_circuit, tfim

pytestmark = pytest.mark.gpu


def test_single_state_calls_from_several_threads():
    n, nthreads = 14, 4
    rng = np.random.default_rng(123)
    c = hea_circuit(n, 2, rng.uniform(0, 6.3, 6 * n))
    H = tfim(n, 0.7)
    inits = [rand_state(n, rng) for _ in range(nthreads)]

    def work(i):
        s = vq.StateVector(inits[i])
        out = []
        for _ in range(15):
            vq.apply_(c, s)
            out.append(vq.expectation(H, s))
            out.append(s.dot(s))
            E, g = vq.energy_and_gradient(H, c, s)
            out.append(E)
            out.extend(g.tolist())
        return out

    serial = [work(i) for i in range(nthreads)]
    with ThreadPoolExecutor(nthreads) as ex:
        par = list(ex.map(work, range(nthreads)))
    for a, b in zip(serial, par):
        assert a == b


def test_mixed_sizes_and_density_matrices_concurrently():
    """Threads on different programs, sizes and state kinds (their plans share the
    plan cache, their work buffers are their own), each against its serial run."""
    rng = np.random.default_rng(7)
    jobs = []
    for n in (9, 12, 16, 18):
        th = rng.uniform(0, 2 * np.pi, 3 * n * 3)
        jobs.append((hea_circuit(n, 3, th), tfim(n, 0.9), vq.StateVector(rand_state(n, rng))))
    for n in (5, 6):
        th = rng.uniform(0, 2 * np.pi, 3 * n * 2)
        c = hea_circuit(n, 2, th)
        c.append(vq.Depolarizing(2, 0.05))
        c.append(vq.AmplitudeDamping(1, 0.02))
        jobs.append((c, tfim(n, 1.1), vq.DensityMatrix(n)))

    def work(j):
        c, H, s = jobs[j]
        return [vq.energy_and_gradient(H, c, s) for _ in range(6)]

    serial = [work(j) for j in range(len(jobs))]
    with ThreadPoolExecutor(len(jobs)) as ex:
        par = list(ex.map(work, range(len(jobs))))
    for a, b in zip(serial, par):
        for (Ea, ga), (Eb, gb) in zip(a, b):
            assert Ea == Eb and np.array_equal(ga, gb)


def test_small_evaluations_overlap_across_threads():
    """12-qubit evaluations (8 workgroups per phase: a small part of the chip) from
    4 threads run side by side. The results equal the serial ones; the throughput
    of both runs is printed (no timing is asserted)."""
    from vqchip.circuit import Program
    from vqchip.operators import pack_terms
    from vqchip import _lib
    n, nthreads, reps = 12, 4, 150
    rng = np.random.default_rng(9)
    c = hea_circuit(n, 4, rng.uniform(0, 2 * np.pi, 3 * n * 4))
    H = tfim(n, 1.0)
    prog = Program(c, n=n)
    ops, pool = prog.op_array(), prog.pool()
    terms, nt, hpool = pack_terms(H)

    def work(i):
        s = vq.StateVector(n)
        e2, g = np.zeros(2), np.zeros(prog.nparams)
        out = []
        for _ in range(reps):
            _lib.call("vqc_sv_energy_grad", s.handle, len(prog.ops), ops, _lib.dptr(pool), nt, terms, _lib.dptr(hpool),
                      prog.nparams, _lib.dptr(e2), _lib.dptr(g), None, 0)
            out.append((e2[0], g.copy()))
        return out

    work(0)
    t0 = time.perf_counter()
    serial = [work(i) for i in range(nthreads)]
    ts = time.perf_counter() - t0
    with ThreadPoolExecutor(nthreads) as ex:
        t0 = time.perf_counter()
        par = list(ex.map(work, range(nthreads)))
        tp = time.perf_counter() - t0
    for a, b in zip(serial, par):
        for (Ea, ga), (Eb, gb) in zip(a, b):
            assert Ea == Eb and np.array_equal(ga, gb)
    print(f"\n12-qubit evaluations: serial {nthreads * reps / ts:.0f}/s, {nthreads} threads {nthreads * reps / tp:.0f}/s",
          flush=True)


def test_thread_release_then_next_call():
    """vqc_thread_release frees the calling thread's stream and buffers; the thread's
    next call makes new ones and gets the same results. Threads of a pool release
    before they end."""
    import threading

    from vqchip import _lib

    n = 12
    rng = np.random.default_rng(321)
    c = hea_circuit(n, 3, rng.uniform(0, 6.3, 9 * n))
    H = tfim(n, 0.8)
    psi = rand_state(n, rng)
    E0, g0 = vq.energy_and_gradient(H, c, vq.StateVector(psi))
    res, errs = [], []

    def work():
        try:
            s = vq.StateVector(psi)
            a = vq.energy_and_gradient(H, c, s)
            _lib.call("vqc_thread_release")
            _lib.call("vqc_thread_release")  # a second release is a no-op
            b = vq.energy_and_gradient(H, c, s)
            res.append((a, b))
            del s
            _lib.call("vqc_thread_release")
        except Exception as e:  # reported in the main thread
            errs.append(e)

    ts = [threading.Thread(target=work) for _ in range(3)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errs, errs
    assert len(res) == 3
    for a, b in res:
        for E, g in (a, b):
            assert E == E0 and np.array_equal(g, g0)
