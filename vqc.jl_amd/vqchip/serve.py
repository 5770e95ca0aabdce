"""Serving: continuous batching of energy + adjoint-gradient requests (no
counterpart in the reference, which evaluates one circuit per call,
src/circuitdiff.jl:2-8, src/expecdiff.jl:28).

    srv = Server(hbm_budget=288e9)
    t = srv.submit(Request(circuit, H, n))     # returns at once
    E, grad, times = srv.wait(t)               # times = (submit, launch, done), seconds
    srv.close()

The scheduling loop runs in the library (one native thread, several HIP
streams); this module packs requests and drives open-loop load for measurement.
`plain_eval` is the one-request-per-call path the server is compared with.
"""
import ctypes
import threading
import time

import numpy as np

from . import _lib
from .circuit import Program
from .operators import pack_terms


class Request:
    """One packed energy + gradient request: |0..0> through `circuit`, energy of H."""

    def __init__(self, circuit, H, n):
        self.n = n
        prog = Program(circuit, n=n)
        self.nparams = prog.nparams
        self.nops = len(prog.ops)
        self.ops = prog.op_array()
        self.pool = np.ascontiguousarray(prog.pool())
        terms, nt, hpool = pack_terms(H)
        self.terms, self.nterms = terms, nt
        self.hpool = np.ascontiguousarray(hpool) if len(hpool) else np.zeros(1, complex)
        inst = _lib.Instance()
        inst.nqubits = n
        inst.nops = self.nops
        inst.op_begin = 0
        inst.nterms = nt
        inst.nparams = self.nparams
        inst.term_begin = 0
        inst.grad_begin = 0
        self.inst = inst
        self.cost = float(2 ** n) * (3 * self.nops + 2 * nt + 1)


class Server:
    def __init__(self, hbm_budget=288e9, lanes=3, max_batch=64, grid_frac=1.0, cost_ratio=64.0, poll_us=50,
                 solo_qubits=19):
        cfg = _lib.ServerConfig()
        _lib.call("vqc_server_default_config", ctypes.byref(cfg))
        cfg.hbm_budget = int(hbm_budget)
        cfg.lanes = lanes
        cfg.max_batch = max_batch
        cfg.grid_frac = grid_frac
        cfg.cost_ratio = cost_ratio
        cfg.poll_us = poll_us
        cfg.solo_qubits = solo_qubits
        h = ctypes.c_void_p()
        _lib.call("vqc_server_create", ctypes.byref(cfg), ctypes.byref(h))
        self._h = h
        self._nparams = {}

    def submit(self, req):
        t = ctypes.c_int64()
        _lib.call("vqc_server_submit", self._h, ctypes.byref(req.inst), req.ops, req.terms, _lib.dptr(req.pool),
                  len(req.pool), _lib.dptr(req.hpool), len(req.hpool) if req.nterms else 0, ctypes.byref(t))
        self._nparams[t.value] = req.nparams
        return t.value

    def _buffers(self, ticket):
        e = np.zeros(2)
        g = np.zeros(max(self._nparams[ticket], 1))
        tm = np.zeros(3)
        return e, g, tm

    def wait(self, ticket):
        e, g, tm = self._buffers(ticket)
        np_ = self._nparams.pop(ticket)
        _lib.call("vqc_server_wait", self._h, ticket, _lib.dptr(e), _lib.dptr(g), _lib.dptr(tm))
        return complex(e[0], e[1]), g[:np_], tuple(tm)

    def poll(self, ticket):
        """None while running, else (E, grad, times)."""
        e, g, tm = self._buffers(ticket)
        rc = _lib.call_status("vqc_server_poll", self._h, ticket, _lib.dptr(e), _lib.dptr(g), _lib.dptr(tm))
        if rc == 1:
            return None
        np_ = self._nparams.pop(ticket)
        return complex(e[0], e[1]), g[:np_], tuple(tm)

    def stats(self):
        s = _lib.ServerStats()
        _lib.call("vqc_server_stats", self._h, ctypes.byref(s))
        d = {k: getattr(s, k) for k, _ in s._fields_ if k != "batch_hist"}
        # batches by size: the upper end of each bin (1, 2, 4, ..., 64, and 65 for larger)
        d["batch_size_hist"] = {(1 << i) if i < 7 else 65: int(s.batch_hist[i]) for i in range(8) if s.batch_hist[i]}
        return d

    def close(self):
        if self._h:
            _lib.call("vqc_server_destroy", self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def plain_eval(req, state=None, fused=False):
    """The plain path: one request per call (|0..0> uploaded, circuit, energy and
    adjoint gradient, results back). fused=False applies the circuit gate by gate."""
    from .state import StateVector
    if state is None or state.nqubits() != req.n:
        state = StateVector(req.n)
    else:
        _lib.call("vqc_state_reset", state.handle)
    e = np.zeros(2)
    g = np.zeros(max(req.nparams, 1))
    _lib.call("vqc_sv_energy_grad", state.handle, req.nops, req.ops, _lib.dptr(req.pool), req.nterms, req.terms,
              _lib.dptr(req.hpool), req.nparams, _lib.dptr(e), _lib.dptr(g), None, 0 if fused else 1)
    return complex(e[0], e[1]), g[:req.nparams]


def poisson_arrivals(rate, count, rng):
    """Open-loop arrival times (seconds from 0) of a Poisson stream."""
    return np.cumsum(rng.exponential(1.0 / rate, count))


def run_plain_fifo(requests, arrivals, fused=False):
    """One request at a time in arrival order. Returns (latencies, busy seconds, span)."""
    from .state import StateVector
    states = {}  # one device state per qubit count, reused (allocation is not charged to the plain path)
    for req in requests:
        if req.n not in states:
            states[req.n] = StateVector(req.n)
    lat = np.zeros(len(requests))
    busy = 0.0
    t0 = time.perf_counter()
    for i, req in enumerate(requests):
        wait = t0 + arrivals[i] - time.perf_counter()
        if wait > 0:
            time.sleep(wait)
        s = time.perf_counter()
        plain_eval(req, states[req.n], fused=fused)
        e = time.perf_counter()
        busy += e - s
        lat[i] = e - (t0 + arrivals[i])
    return lat, busy, time.perf_counter() - t0


def run_server_open_loop(server, requests, arrivals, pack=None):
    """Submit each request at its arrival time from a feeder thread; wait for all.
    With `pack` (the (circuit, H, n) of every request) each request is packed at
    its arrival, so its latency counts the host packing: arrival -> packed and
    submitted (host clock) plus submit -> done (the server's clock).
    Returns (latencies, span, results)."""
    tickets = [None] * len(requests)
    lead = np.zeros(len(requests))  # arrival -> submit
    ready = threading.Semaphore(0)
    err = []
    t0 = time.perf_counter()

    def feed():
        try:
            for i, req in enumerate(requests):
                wait = t0 + arrivals[i] - time.perf_counter()
                if wait > 0:
                    time.sleep(wait)
                if pack is not None:
                    req = Request(*pack[i])
                lead[i] = time.perf_counter() - (t0 + arrivals[i])
                tickets[i] = server.submit(req)
                ready.release()
        except Exception as ex:  # surfaced by the caller below
            err.append(ex)
            for _ in range(len(requests)):
                ready.release()

    th = threading.Thread(target=feed)
    th.start()
    lat = np.zeros(len(requests))
    results = []
    for i in range(len(requests)):
        ready.acquire()
        if err:
            break
        E, g, (ts, tl, td) = server.wait(tickets[i])
        lat[i] = max(lead[i], 0.0) + (td - ts)
        results.append((E, g))
    th.join()
    if err:
        raise err[0]
    return lat, time.perf_counter() - t0, results


def run_server_closed_loop(server, requests):
    """Every (packed) request queued at once; the span from the first submit to
    the last result, i.e. the server's saturation throughput. Returns (span, results)."""
    t0 = time.perf_counter()
    tickets = [server.submit(r) for r in requests]
    res = [server.wait(t) for t in tickets]
    return time.perf_counter() - t0, [(E, g) for E, g, _ in res]
