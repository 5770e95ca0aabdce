"""Benchmark: 28-qubit ComplexF64 gate sweeps on one MI355X (plus the energy,
gradient, batched and serving figures of the other workloads).

One step = one pass of the 480-gate, 28-qubit Haar circuit (20 layers of
14 one-, 7 two- and 3 three-qubit gates) over the state resident in HBM, one
kernel per gate (the unfused drop-in path). `value` is the algorithmic HBM
bandwidth of those sweeps: every gate reads and writes all 2^28 amplitudes once,
2 * 16 * 2^28 bytes per gate.

Every workload also reports its error against an FP64 reference: the oracle's
values (computed in the CPU leg, which times them as the baseline, or frozen in
tests/golden/ for the workloads that take the oracle minutes), closed forms
(the Pauli energy of a product state), or the plain path where the oracle is out
of reach (served requests above 16 qubits).

Prints one JSON line. --gpus N > 1 is refused: this build runs on one GPU.
"""
import os

# CPU baseline threads: one per core this process may run on, pinned in order
# (read by the OpenMP runtime of the CPU reference when it loads)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

import argparse  # noqa: E402
import ctypes  # noqa: E402
import json  # noqa: E402
import sys  # noqa: E402
import time  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "vqc.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
METRIC = "28q gate HBM GB/s; energy+grad evals/s and p99 latency, batched vs plain"
GOLDEN = os.path.join(ROOT, "tests", "golden")


class HipTimer:
    """hipEvent pairs recorded on the library's stream."""

    def __init__(self, stream):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.stream = stream
        self.e0, self.e1 = ctypes.c_void_p(), ctypes.c_void_p()
        for e in (self.e0, self.e1):
            assert self.hip.hipEventCreate(ctypes.byref(e)) == 0

    def start(self):
        assert self.hip.hipEventRecord(self.e0, self.stream) == 0

    def stop_ms(self):
        assert self.hip.hipEventRecord(self.e1, self.stream) == 0
        assert self.hip.hipEventSynchronize(self.e1) == 0
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), self.e0, self.e1) == 0
        return ms.value


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b)) / max(1.0, float(np.max(np.abs(b)))))


def median_s(fn, reps, warm=1):
    """Median wall time of `reps` calls after `warm` untimed ones."""
    for _ in range(warm):
        fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def unfused_sweep(vq, _lib, state, mats):
    """One kernel per gate (vqc_sv_apply_dense)."""
    for k, pos, U, _keep in mats:  # _keep holds the array U points into
        _lib.call("vqc_sv_apply_dense", state.handle, k, pos, U)


def _prog_args(vq, circuit, H, n, dm=False):
    from vqchip.operators import pack_terms
    prog = vq.Program(circuit, n=n, dm=dm)
    terms, nt, hpool = pack_terms(H)
    return prog, prog.op_array(), prog.pool(), terms, nt, hpool


def _eval(vq, _lib, state, args, fused):
    prog, ops, pool, terms, nt, hpool = args
    g = np.zeros(max(prog.nparams, 1))
    e = np.zeros(2)
    fn = "vqc_dm_energy_grad" if state.is_dm else "vqc_sv_energy_grad"
    _lib.call(fn, state.handle, len(prog.ops), ops, _lib.dptr(pool), nt, terms, _lib.dptr(hpool), prog.nparams,
              _lib.dptr(e), _lib.dptr(g), None, 0 if fused else 1)
    return complex(e[0], e[1]), g[:prog.nparams]


def _evaluator(vq, _lib, state, args, fused):
    """One evaluation as a closure over preallocated outputs and converted pointers,
    as a compiled caller (Julia's ccall) makes it: the timed loop then measures the
    library call, not the building of Python arguments (about 10 us per call)."""
    prog, ops, pool, terms, nt, hpool = args
    g = np.zeros(max(prog.nparams, 1))
    e = np.zeros(2)
    fn = "vqc_dm_energy_grad" if state.is_dm else "vqc_sv_energy_grad"
    cargs = (state.handle, len(prog.ops), ops, _lib.dptr(pool), nt, terms, _lib.dptr(hpool), prog.nparams, _lib.dptr(e),
             _lib.dptr(g), None, 0 if fused else 1)

    def run():
        _lib.call(fn, *cargs)
        return e, g

    return run


def _time_evals(fn, min_s=1.0, max_n=5000):
    fn()  # warm
    n, t0 = 0, time.perf_counter()
    while n < max_n and (n < 3 or time.perf_counter() - t0 < min_s):
        fn()
        n += 1
    return n / (time.perf_counter() - t0)


def eval_flops(circuit, H, n, dm=False):
    """Algorithmic FP64 flops of one energy + adjoint-gradient evaluation: a dense
    k-qubit matrix costs 8 * 2^k flops per amplitude it sweeps; the forward pass
    applies every gate once, the seed applies every key twice (the energy and
    (M + M^dagger) psi), the reverse pass applies every gate to psi and lambda and
    every generator once. A density-matrix gate sweeps the rows and the columns."""
    amps = 4 ** n if dm else 2 ** n
    side = 2 if dm else 1
    f = 0.0
    for g in circuit:
        k = len(g.ordered_positions())
        if hasattr(g, "kraus"):
            f += 3 * 8 * 4 ** k  # superoperator on 2k bits, forward and reverse (S^-1 and S^dagger)
            continue
        np_ = g.nparameters()
        f += side * 8 * 2 ** k * (3 + np_)
    for key in H.data():
        f += 2 * 8 * 2 ** len(key)
    return f * amps


def fused_hbm_bytes(info, nbits):
    """Physical HBM bytes of one fused evaluation from its plan: a forward pass reads
    and writes the state, the Hamiltonian passes read psi and write lambda (and read
    it back after the first), a reverse pass reads and writes psi and lambda."""
    sb = 16 * 2 ** nbits
    ham = 2 * min(info.phases_ham, 1) + 3 * max(info.phases_ham - 1, 0)
    return sb * (2 * info.phases_fwd + ham + 4 * info.phases_rev)


def roofline_of(seconds, flops, hbm_bytes):
    fl = flops / seconds / 1e12
    gb = hbm_bytes / seconds / 1e9
    return {"fp64_TFps": round(fl, 3), "fp64_frac": round(fl / FP64_PEAK_TFS, 4), "hbm_GBps": round(gb, 1),
            "hbm_frac": round(gb / HBM_PEAK_GBS, 4), "flops_per_eval": flops, "hbm_bytes_per_eval": hbm_bytes}


def _threaded_evals(vq, _lib, args, n, nthreads, seconds=1.0):
    """Evaluations per second of `nthreads` host threads each evaluating the same
    program on its own state (ctypes releases the GIL during the call)."""
    import threading
    counts = [0] * nthreads
    start = threading.Barrier(nthreads + 1)
    stop = [False]

    def work(i):
        s = vq.StateVector(n)
        run = _evaluator(vq, _lib, s, args, True)
        run()
        start.wait()
        while not stop[0]:
            run()
            counts[i] += 1
        del run, s
        _lib.call("vqc_thread_release")  # this thread's stream and buffers, before it ends

    ts = [threading.Thread(target=work, args=(i,)) for i in range(nthreads)]
    for t in ts:
        t.start()
    start.wait()
    t0 = time.perf_counter()
    time.sleep(seconds)
    stop[0] = True
    for t in ts:
        t.join()
    return sum(counts) / (time.perf_counter() - t0), nthreads


def bench_config0(vq, _lib):
    """12 qubits, 4 HEA layers (144 angles), TFIM: energy + full adjoint gradient."""
    from vqchip.workloads import config0
    c, H, _ = config0()
    n = 12
    args = _prog_args(vq, c, H, n)
    psi0 = vq.StateVector(n)
    Ef, gf = _eval(vq, _lib, psi0, args, True)
    info = _lib.PlanInfo()
    _lib.call("vqc_plan_info", ctypes.byref(info))
    Ep, gp = _eval(vq, _lib, psi0, args, False)
    rate_default = _time_evals(_evaluator(vq, _lib, psi0, args, True))
    # the autotuner picks this size's tile shape (forward / reverse tile bits, low
    # bits) into the library's per-size tile table; the table is restored after
    from vqchip import autotune
    table_before = autotune.get_tile_table()
    t_tune = time.perf_counter()
    tuned = autotune.tune_sizes([n], workload=lambda _n: (c, H, psi0), reps=5, lows=(2, 3, 4, 5), spread=(-1, 0, 1))
    tune_s = time.perf_counter() - t_tune
    Et, gt = _eval(vq, _lib, psi0, args, True)
    _lib.call("vqc_plan_info", ctypes.byref(info))
    rate = _time_evals(_evaluator(vq, _lib, psi0, args, True))
    # the same evaluation from 4 host threads at once (each thread has its own
    # stream and buffers): independent evaluations of an optimizer overlap on the GPU
    rate_threads, nthreads = _threaded_evals(vq, _lib, args, n, 4)
    autotune.set_tile_table({n: table_before.get(n, (0, 0, 0))})
    out = {
        "workload": "12 qubits, 4 layers Rx/Ry/Rz + CNOT ladder, 144 angles, TFIM (23 keys)",
        "fused_evals_per_s": round(rate, 1),
        "fused_evals_per_s_default_tiles": round(rate_default, 1),
        f"fused_evals_per_s_{nthreads}_threads": round(rate_threads, 1),
        "autotuned_tiles": {"fwd_bits": tuned[n]["best"][0], "rev_bits": tuned[n]["best"][1], "low_bits": tuned[n]["best"][2],
                            "tune_s": round(tune_s, 2)},
        "plain_evals_per_s": round(_time_evals(_evaluator(vq, _lib, psi0, args, False)), 1),
        "energy": Ef.real,
        "max_abs_diff_plain_vs_fused": max(abs(Ep - Ef), float(np.max(np.abs(gp - gf))), abs(Et - Ef),
                                           float(np.max(np.abs(gt - gf)))),
        "roofline_fused": roofline_of(1.0 / rate, eval_flops(c, H, n), fused_hbm_bytes(info, n)),
        "tile_passes": {"fwd": info.phases_fwd, "ham": info.phases_ham, "rev": info.phases_rev},
    }
    log(f"config0: {out}")
    return out, (Ef, gf, Ep, gp, args)


def bench_config2(vq, _lib):
    """256 instances, n ~ U{8..20}, depth ~ U{2..12}, Heisenberg: batched graph vs plain per-instance loop."""
    from vqchip.batch import Batch
    from vqchip.workloads import config2_instances
    inst = config2_instances()
    t0 = time.perf_counter()
    b = Batch(inst)
    plan_s = time.perf_counter() - t0
    E_b, g_b = b.run()
    binfo = b.info()
    t_batch = median_s(b.run, 7)
    # plain: one circuit per call, one kernel per gate and per reduction
    prepared = [(_prog_args(vq, c, H, n), n) for c, H, n in inst]
    states = {}
    t0 = time.perf_counter()
    E_p, g_p = [], []
    for a, n in prepared:
        if n not in states:
            states[n] = vq.StateVector(n)
        E, g = _eval(vq, _lib, states[n], a, False)
        E_p.append(E)
        g_p.append(g)
    t_plain = time.perf_counter() - t0

    def fused_loop():
        for a, n in prepared:
            _eval(vq, _lib, states[n], a, True)

    t_fused = median_s(fused_loop, 3)
    err_p = max(max(abs(x - y) for x, y in zip(E_b, E_p)),
                max(float(np.max(np.abs(x - y))) if x.size else 0.0 for x, y in zip(g_b, g_p)))
    # every instance against the oracle's values (frozen: minutes of CPU work)
    fx = np.load(os.path.join(GOLDEN, "config2_energy_grad.npz"), allow_pickle=False)
    off = fx["grad_offsets"]
    err_o = max(max(rel_err(E_b[i], fx["energy"][i]), rel_err(g_b[i], fx["grad"][off[i]:off[i + 1]]))
                for i in range(len(inst)))
    flops = sum(eval_flops(c, H, n) for c, H, n in inst)
    out = {
        "workload": "256 instances, n~U{8..20}, depth~U{2..12}, Heisenberg J~N(0,1)",
        "batched_evals_per_s": round(len(inst) / t_batch, 1),
        "batched_ms": round(t_batch * 1e3, 2),
        "plain_evals_per_s": round(len(inst) / t_plain, 1),
        "plain_ms": round(t_plain * 1e3, 1),
        "fused_per_instance_evals_per_s": round(len(inst) / t_fused, 1),
        "fused_per_instance_ms": round(t_fused * 1e3, 1),
        "speedup_batched_vs_plain": round(t_plain / t_batch, 2),
        "speedup_batched_vs_fused_per_instance": round(t_fused / t_batch, 2),
        "plan_ms": round(plan_s * 1e3, 1),
        "arena_MiB": round(b.arena_bytes() / 2**20, 1),
        "batch_plan": binfo,
        "roofline_batched": roofline_of(t_batch, flops, binfo["hbm_bytes"]),
        "max_abs_diff_batched_vs_plain": err_p,
        "max_rel_err_vs_oracle_all_256": err_o,
    }
    log(f"config2: {out}")
    return out


def bench_config2_toffoli(vq, _lib):
    """The uneven batch with one Toffoli per layer: batches with 3-qubit gates run
    the persistent kernel built for 3-bit micro-ops (k_batch<3>)."""
    from vqchip.batch import Batch
    from vqchip.workloads import config2_toffoli_instances
    inst = config2_toffoli_instances()
    b = Batch(inst)
    E_b, g_b = b.run()
    t_batch = median_s(b.run, 5)
    prepared = [(_prog_args(vq, c, H, n), n) for c, H, n in inst]
    states = {n: vq.StateVector(n) for _, n in prepared}
    res = [_eval(vq, _lib, states[n], a, True) for a, n in prepared]

    def fused_loop():
        for a, n in prepared:
            _eval(vq, _lib, states[n], a, True)

    t_fused = median_s(fused_loop, 3)
    err = max(max(abs(x - y[0]) for x, y in zip(E_b, res)),
              max(float(np.max(np.abs(x - y[1]))) if x.size else 0.0 for x, y in zip(g_b, res)))
    out = {
        "workload": "the 256-instance batch with one Toffoli (3-qubit gate) per layer",
        "batched_evals_per_s": round(len(inst) / t_batch, 1),
        "batched_ms": round(t_batch * 1e3, 2),
        "fused_per_instance_evals_per_s": round(len(inst) / t_fused, 1),
        "speedup_batched_vs_fused_per_instance": round(t_fused / t_batch, 2),
        "batch_plan": b.info(),
        "max_abs_diff_batched_vs_fused": err,
    }
    log(f"config2 toffoli: {out}")
    return out


def bench_config3(vq, _lib):
    """12-qubit density matrix, 6 noisy layers, sum Z_i: energy + 216-angle gradient."""
    from vqchip.workloads import config3
    c, H, _ = config3()
    n = 12
    args = _prog_args(vq, c, H, n, dm=True)
    rho0 = vq.DensityMatrix(n)
    from vqchip import autotune
    Ef, gf = _eval(vq, _lib, rho0, args, True)
    info = _lib.PlanInfo()
    _lib.call("vqc_plan_info", ctypes.byref(info))
    Ep, gp = _eval(vq, _lib, rho0, args, False)
    tf_default = median_s(lambda: _eval(vq, _lib, rho0, args, True), 9)
    tp = median_s(lambda: _eval(vq, _lib, rho0, args, False), 3)
    # the autotuner picks the tile shape for this workload (tile bits, contiguous low bits)
    before = autotune.get_tuning()
    t0 = time.perf_counter()
    tuned = autotune.tune_energy_grad(H, c, rho0, fwd=(11, 12), rev=(10, 11, 12), low=(2, 3, 4, 6), reps=2)
    tune_s = time.perf_counter() - t0
    Et, gt = _eval(vq, _lib, rho0, args, True)
    info_t = _lib.PlanInfo()
    _lib.call("vqc_plan_info", ctypes.byref(info_t))
    tf = median_s(lambda: _eval(vq, _lib, rho0, args, True), 9)
    autotune.set_tuning(**{k: before[k] for k in autotune.FIELDS})
    fx = np.load(os.path.join(GOLDEN, "config3_energy_grad.npz"), allow_pickle=False)
    out = {
        "workload": "12-qubit DM (2^24 entries), 6 layers, depolarizing p=0.01 after each CNOT, "
                    "amplitude damping 0.005 per qubit per layer, sum Z_i, 216 angles",
        "fused_evals_per_s": round(1 / tf, 3),
        "fused_ms": round(tf * 1e3, 2),
        "fused_ms_default_tiles": round(tf_default * 1e3, 2),
        "autotuned_tiles": tuned["best"],
        "autotune_s": round(tune_s, 2),
        "plain_evals_per_s": round(1 / tp, 3),
        "plain_ms": round(tp * 1e3, 2),
        "speedup_fused_vs_plain": round(tp / tf, 2),
        "energy": Ef.real,
        "roofline_fused": roofline_of(tf, eval_flops(c, H, n, dm=True), fused_hbm_bytes(info_t, 2 * n)),
        "max_rel_err_vs_oracle": max(rel_err(Ef, fx["energy"]), rel_err(gf, fx["grad"]), rel_err(Et, fx["energy"]),
                                     rel_err(gt, fx["grad"])),
        "max_rel_err_plain_vs_oracle": max(rel_err(Ep, fx["energy"]), rel_err(gp, fx["grad"])),
    }
    log(f"config3: {out}")
    return out


# HBM bytes one one-qubit sweep launch moves at 28 qubits (k_dense_1x; the same
# for the pair-per-thread form), from two rocprofv3 --pmc passes (FETCH_SIZE x 2.0
# + WRITE_SIZE x 1.0, the factors a copy of known size reads;
# profiles/hbm_traffic_dense_28q.md). Algorithmic: 2 * 16 * 2^28 = 8.590e9.
TRAFFIC_1Q_BYTES = 8.5901e9


def bench_config4(vq, _lib, count, seed=20261203):
    """Serving: open-loop Poisson stream of energy + gradient requests, n log-uniform
    in 10..26, depth U{4..40}, TFIM h ~ U(0.5, 1.5). The rate is set so that the
    plain path (one request per call, one kernel per gate) is ~70% busy; the
    continuous-batching server gets the same arrival trace. The server's closed-loop
    saturation rate is measured too, and open-loop streams at 50%, 80% and 100% of
    it, with every request packed at its arrival (packing counts in its latency)."""
    from vqchip.serve import (Request, Server, poisson_arrivals, run_plain_fifo, run_server_open_loop, plain_eval,
                              run_server_closed_loop)
    from vqchip.workloads import config4_request
    rng = np.random.default_rng(seed)
    specs = [config4_request(rng) for _ in range(count)]
    t0 = time.perf_counter()
    reqs = [Request(*s) for s in specs]
    pack_s = time.perf_counter() - t0
    # service time of every request on the plain path, back to back (also its closed-loop throughput)
    states = {}
    svc = np.zeros(count)
    plain_res = []
    for i, r in enumerate(reqs):
        if r.n not in states:
            states[r.n] = vq.StateVector(r.n)
        s0 = time.perf_counter()
        plain_res.append(plain_eval(r, states[r.n], fused=False))
        svc[i] = time.perf_counter() - s0
    del states
    rate = 0.7 / float(np.mean(svc))
    arr = poisson_arrivals(rate, count, np.random.default_rng(seed + 1))
    lat_p, busy_p, span_p = run_plain_fifo(reqs, arr, fused=False)
    log(f"config4: plain open loop done ({span_p:.1f} s)")
    err = 0.0

    def check(res):
        nonlocal err
        for (E, g), (Ep, gp) in zip(res, plain_res):
            err = max(err, abs(E - Ep), float(np.max(np.abs(g - gp))) if gp.size else 0.0)

    pct = lambda v, q: round(float(np.percentile(v, q)) * 1e3, 2)

    def hist(st):
        return {int(k): int(v) for k, v in sorted(st["batch_size_hist"].items())}

    with Server(hbm_budget=288e9) as srv:
        lat_s, span_s, res_s = run_server_open_loop(srv, reqs, arr, pack=specs)
        st = srv.stats()
    check(res_s)
    # closed loop: every request queued at once, the server's saturation rate
    with Server(hbm_budget=288e9) as srv:
        span_c, res_c = run_server_closed_loop(srv, reqs)
        st_c = srv.stats()
    check(res_c)
    sat = count / span_c
    loads = {}
    for frac in (0.5, 0.8, 1.0):
        arr_f = poisson_arrivals(frac * sat, count, np.random.default_rng(seed + int(frac * 10)))
        with Server(hbm_budget=288e9) as srv:
            lat_f, span_f, res_f = run_server_open_loop(srv, reqs, arr_f, pack=specs)
            st_f = srv.stats()
        check(res_f)
        loads[f"{int(frac * 100)}pct"] = {"rate_per_s": round(frac * sat, 2), "served_evals_per_s": round(count / span_f, 2),
                                         "p50_ms": pct(lat_f, 50), "p99_ms": pct(lat_f, 99),
                                         "mean_batch_size": round(st_f["batched_requests"] / max(st_f["batches"], 1), 2),
                                         "batch_size_hist": hist(st_f)}
    out = {
        "workload": f"{count} requests, n=round(exp(U(ln10, ln26))), depth~U{{4..40}}, HEA layer, TFIM h~U(0.5,1.5), "
                    "Poisson arrivals",
        "rate_per_s": round(rate, 3),
        "plain_mean_service_ms": round(float(np.mean(svc)) * 1e3, 2),
        "plain_utilisation": round(busy_p / span_p, 3),
        "plain_p50_ms": pct(lat_p, 50), "plain_p99_ms": pct(lat_p, 99), "plain_mean_ms": round(float(np.mean(lat_p)) * 1e3, 2),
        "plain_evals_per_s": round(count / span_p, 3),
        "plain_closed_loop_evals_per_s": round(count / float(np.sum(svc)), 3),
        "served_p50_ms": pct(lat_s, 50), "served_p99_ms": pct(lat_s, 99), "served_mean_ms": round(float(np.mean(lat_s)) * 1e3, 2),
        "served_evals_per_s": round(count / span_s, 3),
        "latency_includes_packing": True,
        "p99_speedup": round(float(np.percentile(lat_p, 99) / np.percentile(lat_s, 99)), 2),
        "p50_speedup": round(float(np.percentile(lat_p, 50) / np.percentile(lat_s, 50)), 2),
        "server_batches": st["batches"],
        "mean_batch_size": round(st["batched_requests"] / max(st["batches"], 1), 2),
        "batch_size_hist": hist(st),
        "saturation_closed_loop_evals_per_s": round(sat, 2),
        "saturation_batch_size_hist": hist(st_c),
        "load_sweep": loads,
        "peak_live_GiB": round(st["peak_live_bytes"] / 2**30, 3),
        "budget_GiB": round(st["budget_bytes"] / 2**30, 1),
        "max_abs_diff_served_vs_plain": err,
        "qubits_histogram": {int(k): int(v) for k, v in zip(*np.unique([r.n for r in reqs], return_counts=True))},
        "pack_s": round(pack_s, 2),
    }
    log(f"config4: {out}")
    return out, reqs, plain_res


def pauli_product_state_check(vq, _lib, n, Hp, arr, nt, seed=281):
    """200 Pauli strings on a product state: <P> = c prod_q <u_q|sigma_q|u_q>, the
    FP64 closed form of config 1's energy sweep."""
    rng = np.random.default_rng(seed)
    prod = vq.StateVector(n)
    us = []
    for q in range(n):
        Z = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
        Q, R = np.linalg.qr(Z)
        U = Q * (np.diag(R) / np.abs(np.diag(R)))
        us.append(U[:, 0])
        _lib.call("vqc_sv_apply_dense", prod.handle, 1, _lib.i32ptr([q]), _lib.dptr(np.ascontiguousarray(U.reshape(-1, order="F"))))
    per = np.zeros(2 * nt)
    s2 = np.zeros(2)
    _lib.call("vqc_sv_expec_pauli", prod.handle, None, nt, arr, _lib.dptr(per), _lib.dptr(s2))
    sig = {"X": np.array([[0, 1], [1, 0]], complex), "Y": np.array([[0, -1j], [1j, 0]]), "Z": np.diag([1.0 + 0j, -1]),
           "I": np.eye(2, dtype=complex)}
    closed = []
    for t in Hp.terms:
        v = t.coeff
        for p, l in zip(t.positions, t.labels):
            v *= us[p - 1].conj() @ sig[l] @ us[p - 1]
        closed.append(v)
    return rel_err(per[0::2] + 1j * per[1::2], np.array(closed))


def cpu_baseline_leg(oracle, args, circuit, c0, c4):
    """The CPU reference (the oracle: a C++/OpenMP restatement of the reference's
    threaded kernels) on this host's cores, pinned, one thread per core: 1 warm-up,
    then the median of at least 5 repetitions of a fixed sample of each workload.
    It also yields the oracle's values that the device results are checked against."""
    cores = len(os.sched_getaffinity(0))
    oracle.set_threads(cores)
    out = {"threads": cores, "nproc": os.cpu_count(), "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
           "omp_places": os.environ.get("OMP_PLACES")}
    n = args.qubits
    # config 1: the first layer (14 + 7 + 3 gates) of the 28-qubit circuit
    v = np.zeros(1 << n, dtype=np.complex128)
    v[0] = 1
    layer = list(circuit[:24])

    def sweep_layer():
        for g in layer:
            pos = [q - 1 for q in g.ordered_positions()]
            oracle.apply_dense(v, len(pos), pos, g.ordered_mat())

    t1 = median_s(sweep_layer, args.cpu_reps)
    gbs = 2 * 16 * (1 << n) * len(layer) / t1 / 1e9
    out["config1"] = {"GBps": round(gbs, 2), "sample": f"the first layer ({len(layer)} gates) of the {n}-qubit circuit",
                      "median_s": round(t1, 3)}
    log(f"cpu config1: {gbs:.1f} GB/s on {cores} threads")
    del v
    # config 0: the whole evaluation
    Ef, gf, Ep, gp, a0 = c0
    prog, ops, pool, terms, nt, hpool = a0
    v0 = np.zeros(1 << 12, dtype=complex)
    v0[0] = 1
    res = {}

    def e0():
        res["r"] = oracle.energy_grad(v0, 12, 0, ops, len(prog.ops), pool, terms, nt, hpool, prog.nparams)

    t0 = median_s(e0, max(args.cpu_reps, 15))
    Er, gr, _ = res["r"]
    out["config0"] = {"evals_per_s": round(1 / t0, 2), "sample": "the whole 12-qubit evaluation", "median_s": round(t0, 5)}
    errs0 = {"max_rel_err_fused_vs_oracle": max(rel_err(Ef, Er), rel_err(gf, gr)),
             "max_rel_err_plain_vs_oracle": max(rel_err(Ep, Er), rel_err(gp, gr))}
    # config 2: the instances of up to 12 qubits of the batch (the whole evaluation each)
    from vqchip.workloads import config2_instances
    import vqchip as vq
    sample = [(vq.Program(c, n=nn), H, nn) for c, H, nn in config2_instances() if nn <= 12]
    from vqchip.operators import pack_terms
    packed = []
    for p, H, nn in sample:
        t, k, hp = pack_terms(H)
        z = np.zeros(1 << nn, dtype=complex)
        z[0] = 1
        packed.append((p, p.op_array(), p.pool(), t, k, hp, nn, z))

    def e2():
        for p, o, pl, t, k, hp, nn, z in packed:
            oracle.energy_grad(z, nn, 0, o, len(p.ops), pl, t, k, hp, p.nparams)

    t2 = median_s(e2, args.cpu_reps)
    out["config2"] = {"evals_per_s": round(len(packed) / t2, 2), "median_s": round(t2, 3),
                      "sample": f"the {len(packed)} instances of up to 12 qubits of the 256-instance batch"}
    # config 3: one noisy layer of the 12-qubit density-matrix circuit (the full six take ~25 s per evaluation)
    from vqchip.workloads import config3
    c3, H3, _ = config3(depth=1)
    p3 = vq.Program(c3, n=12, dm=True)
    t3_, k3, hp3 = pack_terms(H3)
    rho0 = np.zeros(1 << 24, dtype=complex)
    rho0[0] = 1
    o3, pl3 = p3.op_array(), p3.pool()
    t3 = median_s(lambda: oracle.energy_grad(rho0, 12, 1, o3, len(p3.ops), pl3, t3_, k3, hp3, p3.nparams), args.cpu_reps)
    out["config3"] = {"evals_per_s_one_layer": round(1 / t3, 3), "median_s": round(t3, 3),
                      "sample": "one of the six noisy layers (72 of the 432 ops), energy + 36 gradients"}
    # config 4: the oracle's values of the requests of up to 16 qubits (checks the served results)
    errs4 = 0.0
    if c4 is not None:
        reqs, plain_res = c4
        for r, (Ep4, gp4) in zip(reqs, plain_res):
            if r.n > 16:
                continue
            z = np.zeros(1 << r.n, dtype=complex)
            z[0] = 1
            Er4, gr4, _ = oracle.energy_grad(z, r.n, 0, r.ops, r.nops, r.pool, r.terms, r.nterms, r.hpool, r.nparams)
            errs4 = max(errs4, rel_err(Ep4, Er4), rel_err(gp4, gr4))
    log(f"cpu leg: {out}")
    return out, errs0, errs4


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=28)
    ap.add_argument("--layers", type=int, default=20)
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--skip-extra", action="store_true")
    ap.add_argument("--only", default="", help="comma list of parts to run: c0,c2,c3,c4,cpu (default all)")
    ap.add_argument("--serve-requests", type=int, default=100)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding of a state is a design note, not implemented)", file=sys.stderr)
        sys.exit(2)

    import vqchip as vq
    from vqchip import _lib
    from vqchip.workloads import config1_circuit, random_pauli_operator

    n = args.qubits
    circuit = config1_circuit(n=n, layers=args.layers)
    G = len(circuit)
    mats = []
    for g in circuit:
        pos = g.ordered_positions()
        U = np.ascontiguousarray(g.ordered_mat().reshape(-1, order="F"))
        mats.append((len(pos), _lib.i32ptr([q - 1 for q in pos]), _lib.dptr(U), U))
    log(f"config1: n={n} gates={G}")
    state = vq.StateVector(n)
    stream = ctypes.c_void_p()
    _lib.call("vqc_get_stream", ctypes.byref(stream))
    timer = HipTimer(stream)
    bytes_per_gate = 2 * 16 * (1 << n)

    # warmup
    for _ in range(args.warmup):
        unfused_sweep(vq, _lib, state, mats)
    _lib.call("vqc_sync")
    # timed: K steps bracketed by synchronisation
    t0 = time.perf_counter()
    timer.start()
    for _ in range(args.steps):
        unfused_sweep(vq, _lib, state, mats)
    ms = timer.stop_ms()
    _lib.call("vqc_sync")
    wall = time.perf_counter() - t0
    gbs = bytes_per_gate * G * args.steps / (ms * 1e-3) / 1e9
    log(f"unfused: {ms / args.steps:.1f} ms/step, {gbs:.0f} GB/s (wall {wall:.2f}s)")
    norm = state.norm()

    # dominant kernel class: the one-qubit dense sweeps, timed alone (k_dense_1x<256>
    # for targets up to bit 12, k_dense_1x<512> above)
    one_q = [m for m, g in zip(mats, circuit) if len(g.positions) == 1]
    timer.start()
    for k, p, U, _keep in one_q:
        _lib.call("vqc_sv_apply_dense", state.handle, k, p, U)
    ms1 = timer.stop_ms()
    avg1_ms = ms1 / len(one_q)
    ach1 = bytes_per_gate / (avg1_ms * 1e-3) / 1e9
    per_class = {}
    for kq in (1, 2, 3):
        sub = [m for m, g in zip(mats, circuit) if len(g.positions) == kq]
        timer.start()
        for k, p, U, _keep in sub:
            _lib.call("vqc_sv_apply_dense", state.handle, k, p, U)
        per_class[f"{kq}q"] = round(bytes_per_gate / (timer.stop_ms() / len(sub) * 1e-3) / 1e9, 1)
    log(f"per-class GB/s {per_class}")

    extra = {"norm_error": abs(norm - 1.0), "per_gate_class_GBps": per_class}
    # ceiling reference: a plain device-to-device copy of the same 2^n amplitudes
    other = vq.StateVector(n)
    _lib.call("vqc_state_copy", other.handle, state.handle)
    timer.start()
    for _ in range(3):
        _lib.call("vqc_state_copy", other.handle, state.handle)
    ms_copy = timer.stop_ms() / 3
    extra["copy_GBps"] = round(2 * 16 * (1 << n) / (ms_copy * 1e-3) / 1e9, 1)
    # streaming reference with the sweep's own traffic (read and write each amplitude
    # once, in place): the library's x <- a x with a = 1, which leaves the state as it is
    timer.start()
    for _ in range(5):
        _lib.call("vqc_state_scale", state.handle, 1.0, 0.0)
    ms_scale = timer.stop_ms() / 5
    extra["scale_inplace_GBps"] = round(bytes_per_gate / (ms_scale * 1e-3) / 1e9, 1)
    log(f"copy {extra['copy_GBps']} GB/s, in-place scale {extra['scale_inplace_GBps']} GB/s")
    # fused (LDS tile phases), from |0..0>, against the per-gate result of the same circuit
    prog = vq.Program(circuit, n=n)
    ops, pool = prog.op_array(), prog.pool()
    _lib.call("vqc_state_reset", other.handle)
    unfused_sweep(vq, _lib, other, mats)
    _lib.call("vqc_state_reset", state.handle)
    _lib.call("vqc_apply_circuit", state.handle, len(prog.ops), ops, _lib.dptr(pool))
    _lib.call("vqc_state_axpy", other.handle, -1.0, 0.0, state.handle)
    extra["fused_vs_plain_l2_diff"] = float(np.sqrt(other.dot(other).real))
    del other
    _lib.call("vqc_sync")
    # the autotuner picks this size's forward tile shape (tile bits, contiguous low
    # bits) on a copy of the state; the choice goes into the library's tile table
    from vqchip import autotune
    t_tune = time.perf_counter()
    tuned_fwd = autotune.tune_forward(circuit, state, fwd=(12,), lows=(3, 4), reps=1)
    extra["fused_autotuned_tile"] = {"fwd_bits": tuned_fwd["best"][0], "low_bits": tuned_fwd["best"][1],
                                     "tune_s": round(time.perf_counter() - t_tune, 2),
                                     "candidates_ms": {f"{f}/{l}": round(t * 1e3, 2) for (f, l), t in tuned_fwd["table"]}}
    _lib.call("vqc_apply_circuit", state.handle, len(prog.ops), ops, _lib.dptr(pool))
    _lib.call("vqc_sync")
    timer.start()
    for _ in range(args.steps):
        _lib.call("vqc_apply_circuit", state.handle, len(prog.ops), ops, _lib.dptr(pool))
    msf = timer.stop_ms() / args.steps
    extra["fused_ms_per_step"] = round(msf, 3)
    extra["fused_effective_gate_GBps"] = round(bytes_per_gate * G / (msf * 1e-3) / 1e9, 1)
    info = _lib.PlanInfo()
    _lib.call("vqc_plan_info", ctypes.byref(info))
    # physical HBM bandwidth of the fused path: each forward tile pass reads and writes the state once
    extra["fused_tile_passes"] = info.phases_fwd
    extra["fused_physical_GBps"] = round(bytes_per_gate * info.phases_fwd / (msf * 1e-3) / 1e9, 1)
    fl = sum(8 * 2 ** len(g.ordered_positions()) for g in circuit) * 2.0 ** n
    extra["fused_fp64_TFps"] = round(fl / (msf * 1e-3) / 1e12, 2)
    extra["fused_bound_ms"] = round(max(bytes_per_gate * info.phases_fwd / HBM_PEAK_GBS / 1e6, fl / FP64_PEAK_TFS / 1e9), 2)
    log(f"fused: {msf:.1f} ms/step, effective {extra['fused_effective_gate_GBps']} GB/s, "
        f"{info.phases_fwd} tile passes = {extra['fused_physical_GBps']} GB/s physical, {extra['fused_fp64_TFps']} TF/s")

    if not args.skip_extra:
        Hp = random_pauli_operator(n, 200)
        arr, nt = vq.pack_pauli(Hp)
        s2 = np.zeros(2)
        timer.start()
        _lib.call("vqc_sv_expec_pauli", state.handle, None, nt, arr, None, _lib.dptr(s2))
        mse = timer.stop_ms()
        extra["pauli200_energy"] = s2[0]
        extra["pauli200_ms"] = round(mse, 2)
        extra["pauli200_read_GBps"] = round(16 * (1 << n) * nt / (mse * 1e-3) / 1e9, 1)
        del state
        extra["pauli200_product_state_rel_err_vs_closed_form"] = pauli_product_state_check(vq, _lib, n, Hp, arr, nt)
        log(f"pauli energy {s2[0]:.6f} in {mse:.1f} ms, closed-form error "
            f"{extra['pauli200_product_state_rel_err_vs_closed_form']:.2e}")
    else:
        del state

    parts = set(args.only.split(",")) if args.only else {"c0", "c2", "c3", "c4", "cpu"}
    c0 = c4 = None
    if "c0" in parts:
        extra["config0"], c0 = bench_config0(vq, _lib)
    if "c2" in parts:
        extra["config2"] = bench_config2(vq, _lib)
        extra["config2_toffoli"] = bench_config2_toffoli(vq, _lib)
    if "c3" in parts:
        extra["config3"] = bench_config3(vq, _lib)
    if "c4" in parts:
        extra["config4"], reqs4, plain4 = bench_config4(vq, _lib, args.serve_requests)
        c4 = (reqs4, plain4)

    cpu = {"value": None, "unit": "GB/s", "cores": len(os.sched_getaffinity(0)), "kind": "port", "sample": "not run"}
    if "cpu" in parts and c0 is not None:
        import oracle
        leg, errs0, errs4 = cpu_baseline_leg(oracle, args, circuit, c0, c4)
        extra["config0"].update(errs0)
        if c4 is not None:
            extra["config4"]["max_rel_err_served_vs_oracle_n_le_16"] = errs4
        extra["cpu_baseline_all"] = leg
        cpu = {"value": leg["config1"]["GBps"], "unit": "GB/s", "cores": leg["threads"], "kind": "port",
               "sample": leg["config1"]["sample"] + f", median of {args.cpu_reps} after 1 warm-up, oracle restatement, "
                         "threads pinned (OMP_PROC_BIND=close, OMP_PLACES=cores)"}

    out = {
        "metric": METRIC,
        "value": round(gbs, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128 (f64)",
        "data": "synthetic (seeded Haar gates)",
        "config": {"workload": f"{n}-qubit state vector, {args.layers} layers x (14 1q + 7 2q + 3 3q) Haar gates = {G} "
                               "gates, one kernel per gate",
                   "qubits": n, "gates": G, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(ach1, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach1 / HBM_PEAK_GBS, 4), "traffic": TRAFFIC_1Q_BYTES if n == 28 else None,
                     "kernel": "1-qubit dense sweep: k_dense_1x<256> (targets 0-12), k_dense_1x<512> (13-27)",
                     "avg_launch_ms": round(avg1_ms, 4), "bytes_per_launch": bytes_per_gate},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out, default=float), flush=True)


if __name__ == "__main__":
    main()
