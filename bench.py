"""Benchmark: 28-qubit ComplexF64 gate sweeps on one MI355X (plus the energy,
gradient, batched and serving figures of the other workloads).

One step = one pass of the 480-gate, 28-qubit Haar circuit (20 layers of
14 one-, 7 two- and 3 three-qubit gates) over the state resident in HBM, one
kernel per gate (the unfused drop-in path). `value` is the algorithmic HBM
bandwidth of those sweeps: every gate reads and writes all 2^28 amplitudes once,
2 * 16 * 2^28 bytes per gate.

Prints one JSON line. --gpus N > 1 is refused: this build runs on one GPU.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "vqc.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "28q gate HBM GB/s; energy+grad evals/s and p99 latency, batched vs plain"


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


def _time_evals(fn, min_s=1.0, max_n=5000):
    fn()  # warm
    n, t0 = 0, time.perf_counter()
    while n < max_n and (n < 3 or time.perf_counter() - t0 < min_s):
        fn()
        n += 1
    return n / (time.perf_counter() - t0)


def bench_config0(vq, _lib, oracle, timer):
    """12 qubits, 4 HEA layers (144 angles), TFIM: energy + full adjoint gradient."""
    from vqchip.workloads import config0
    c, H, _ = config0()
    n = 12
    args = _prog_args(vq, c, H, n)
    psi0 = vq.StateVector(n)
    Ef, gf = _eval(vq, _lib, psi0, args, True)
    Ep, gp = _eval(vq, _lib, psi0, args, False)
    v0 = np.zeros(1 << n, dtype=complex)
    v0[0] = 1
    prog, ops, pool, terms, nt, hpool = args
    Er, gr, _ = oracle.energy_grad(v0, n, 0, ops, len(prog.ops), pool, terms, nt, hpool, prog.nparams)
    out = {
        "workload": "12 qubits, 4 layers Rx/Ry/Rz + CNOT ladder, 144 angles, TFIM (23 keys)",
        "fused_evals_per_s": round(_time_evals(lambda: _eval(vq, _lib, psi0, args, True)), 1),
        "plain_evals_per_s": round(_time_evals(lambda: _eval(vq, _lib, psi0, args, False)), 1),
        "cpu_evals_per_s": round(_time_evals(lambda: oracle.energy_grad(v0, n, 0, ops, len(prog.ops), pool, terms, nt,
                                                                        hpool, prog.nparams)), 1),
        "energy": Ef.real,
        "max_abs_err_energy_vs_oracle": abs(Ef - Er),
        "max_abs_err_grad_vs_oracle": float(np.max(np.abs(gf - gr))),
        "max_abs_err_grad_plain_vs_oracle": float(np.max(np.abs(gp - gr))),
    }
    log(f"config0: {out}")
    return out


def bench_config2(vq, _lib, oracle, timer):
    """256 instances, n ~ U{8..20}, depth ~ U{2..12}, Heisenberg: batched graph vs plain per-instance loop."""
    from vqchip.batch import Batch
    from vqchip.workloads import config2_instances
    inst = config2_instances()
    t0 = time.perf_counter()
    b = Batch(inst)
    plan_s = time.perf_counter() - t0
    E_b, g_b = b.run()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        b.run()
    t_batch = (time.perf_counter() - t0) / reps
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
    t0 = time.perf_counter()
    for a, n in prepared:
        _eval(vq, _lib, states[n], a, True)
    t_fused = time.perf_counter() - t0
    err_e = max(abs(x - y) for x, y in zip(E_b, E_p))
    err_g = max(float(np.max(np.abs(x - y))) if x.size else 0.0 for x, y in zip(g_b, g_p))
    # oracle on the small instances (n <= 14) of the batch
    err_o, n_o, t_o, cnt = 0.0, 0, 0.0, 0
    for i, ((a, n), (c, H, _)) in enumerate(zip(prepared, inst)):
        if n > 14:
            continue
        prog, ops, pool, terms, nt, hpool = a
        v0 = np.zeros(1 << n, dtype=complex)
        v0[0] = 1
        t1 = time.perf_counter()
        Er, gr, _ = oracle.energy_grad(v0, n, 0, ops, len(prog.ops), pool, terms, nt, hpool, prog.nparams)
        t_o += time.perf_counter() - t1
        cnt += 1
        err_o = max(err_o, abs(E_b[i] - Er), float(np.max(np.abs(g_b[i] - gr))) if gr.size else 0.0)
    out = {
        "workload": "256 instances, n~U{8..20}, depth~U{2..12}, Heisenberg J~N(0,1)",
        "batched_evals_per_s": round(len(inst) / t_batch, 1),
        "batched_ms": round(t_batch * 1e3, 2),
        "plain_evals_per_s": round(len(inst) / t_plain, 1),
        "plain_ms": round(t_plain * 1e3, 1),
        "fused_per_instance_evals_per_s": round(len(inst) / t_fused, 1),
        "speedup_batched_vs_plain": round(t_plain / t_batch, 2),
        "plan_ms": round(plan_s * 1e3, 1),
        "arena_MiB": round(b.arena_bytes() / 2**20, 1),
        "max_abs_diff_batched_vs_plain": max(err_e, err_g),
        "max_abs_err_vs_oracle_n_le_14": err_o,
        "oracle_checked_instances": cnt,
        "cpu_evals_per_s_n_le_14_sample": round(cnt / t_o, 1) if t_o else None,
    }
    log(f"config2: {out}")
    return out


def bench_config3(vq, _lib, oracle, timer):
    """12-qubit density matrix, 6 noisy layers, sum Z_i: energy + 216-angle gradient."""
    from vqchip.workloads import config3
    c, H, _ = config3()
    n = 12
    args = _prog_args(vq, c, H, n, dm=True)
    rho0 = vq.DensityMatrix(n)
    from vqchip import autotune
    Ef, gf = _eval(vq, _lib, rho0, args, True)
    Ep, gp = _eval(vq, _lib, rho0, args, False)
    tf_default = 1.0 / _time_evals(lambda: _eval(vq, _lib, rho0, args, True), min_s=1.0, max_n=60)
    tp = 1.0 / _time_evals(lambda: _eval(vq, _lib, rho0, args, False), min_s=1.0, max_n=5)
    # the autotuner picks the tile shape for this workload (tile bits, contiguous low bits)
    before = autotune.get_tuning()
    t0 = time.perf_counter()
    tuned = autotune.tune_energy_grad(H, c, rho0, fwd=(11, 12), rev=(10, 11, 12), low=(2, 3, 4, 6), reps=2)
    tune_s = time.perf_counter() - t0
    Et, gt = _eval(vq, _lib, rho0, args, True)
    tf = 1.0 / _time_evals(lambda: _eval(vq, _lib, rho0, args, True), min_s=1.0, max_n=60)
    autotune.set_tuning(**{k: before[k] for k in autotune.FIELDS})
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
        "energy": Ef.real,
        "max_abs_diff_fused_vs_plain": max(abs(Ef - Ep), float(np.max(np.abs(gf - gp)))),
        "max_abs_diff_autotuned_vs_plain": max(abs(Et - Ep), float(np.max(np.abs(gt - gp)))),
    }
    log(f"config3: {out}")
    return out


# HBM bytes one k_dense_small<1, 1, true> launch moves at 28 qubits, from two
# rocprofv3 --pmc passes (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the factors a
# copy of known size reads; profiles/hbm_traffic_dense_28q.md). Algorithmic:
# 2 * 16 * 2^28 = 8.590e9.
TRAFFIC_1Q_BYTES = 8.5913e9


def bench_config4(vq, _lib, count, seed=20261203):
    """Serving: open-loop Poisson stream of energy + gradient requests, n log-uniform
    in 10..26, depth U{4..40}, TFIM h ~ U(0.5, 1.5). The rate is set so that the
    plain path (one request per call, one kernel per gate) is ~70% busy; the
    continuous-batching server gets the same arrival trace."""
    from vqchip.serve import Request, Server, poisson_arrivals, run_plain_fifo, run_server_open_loop, plain_eval
    from vqchip.workloads import config4_request
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    reqs = [Request(*config4_request(rng)) for _ in range(count)]
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
    with Server(hbm_budget=288e9) as srv:
        lat_s, span_s, res_s = run_server_open_loop(srv, reqs, arr)
        st = srv.stats()
    err = 0.0
    for (E, g), (Ep, gp) in zip(res_s, plain_res):
        err = max(err, abs(E - Ep), float(np.max(np.abs(g - gp))) if gp.size else 0.0)
    # the same requests at 10x the rate: far beyond the plain path's capacity, so
    # requests pile up and the server batches them (continuous batching under load)
    arr_hi = poisson_arrivals(10 * rate, count, np.random.default_rng(seed + 2))
    with Server(hbm_budget=288e9) as srv:
        lat_h, span_h, res_h = run_server_open_loop(srv, reqs, arr_hi)
        st_h = srv.stats()
    for (E, g), (Ep, gp) in zip(res_h, plain_res):
        err = max(err, abs(E - Ep), float(np.max(np.abs(g - gp))) if gp.size else 0.0)
    pct = lambda v, q: round(float(np.percentile(v, q)) * 1e3, 2)
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
        "p99_speedup": round(float(np.percentile(lat_p, 99) / np.percentile(lat_s, 99)), 2),
        "p50_speedup": round(float(np.percentile(lat_p, 50) / np.percentile(lat_s, 50)), 2),
        "server_batches": st["batches"],
        "mean_batch_size": round(st["batched_requests"] / max(st["batches"], 1), 2),
        "peak_live_GiB": round(st["peak_live_bytes"] / 2**30, 3),
        "budget_GiB": round(st["budget_bytes"] / 2**30, 1),
        "max_abs_diff_served_vs_plain": err,
        "at_10x_rate": {"rate_per_s": round(10 * rate, 3), "served_evals_per_s": round(count / span_h, 3),
                        "served_p50_ms": pct(lat_h, 50), "served_p99_ms": pct(lat_h, 99),
                        "server_batches": st_h["batches"],
                        "mean_batch_size": round(st_h["batched_requests"] / max(st_h["batches"], 1), 2)},
        "qubits_histogram": {int(k): int(v) for k, v in zip(*np.unique([r.n for r in reqs], return_counts=True))},
        "pack_s": round(pack_s, 2),
    }
    log(f"config4: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=28)
    ap.add_argument("--layers", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--skip-extra", action="store_true")
    ap.add_argument("--only", default="", help="comma list of parts to run: c0,c1,c2,c3,c4 (default all)")
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

    # dominant kernel: the one-qubit dense sweep k_dense_small<1, 1, true>, timed
    # alone; one-qubit gates on bits 0..2 run the temporal variant and are left out
    one_q = [m for m, g in zip(mats, circuit) if len(g.positions) == 1 and min(g.positions) - 1 >= 3]
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
    del other
    log(f"copy {extra['copy_GBps']} GB/s")
    # fused (LDS tile phases)
    prog = vq.Program(circuit, n=n)
    ops, pool = prog.op_array(), prog.pool()
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
    log(f"fused: {msf:.1f} ms/step, effective {extra['fused_effective_gate_GBps']} GB/s, "
        f"{info.phases_fwd} tile passes = {extra['fused_physical_GBps']} GB/s physical")

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
        log(f"pauli energy {s2[0]:.6f} in {mse:.1f} ms")

    parts = set(args.only.split(",")) if args.only else {"c0", "c2", "c3", "c4"}
    import oracle
    cores = int(os.environ.get("VQC_CPU_THREADS", min(16, os.cpu_count() or 1)))
    oracle.set_threads(cores)
    del state  # free the 28-qubit state before the other workloads
    if "c0" in parts:
        extra["config0"] = bench_config0(vq, _lib, oracle, timer)
    if "c2" in parts:
        extra["config2"] = bench_config2(vq, _lib, oracle, timer)
    if "c3" in parts:
        extra["config3"] = bench_config3(vq, _lib, oracle, timer)
    if "c4" in parts:
        extra["config4"] = bench_config4(vq, _lib, args.serve_requests)

    # CPU baseline: the oracle (C++/OpenMP restatement of the reference's
    # threaded kernels) on the host cores, on the first gates of the same circuit
    v = np.zeros(1 << n, dtype=np.complex128)
    v[0] = 1
    done, t0 = 0, time.perf_counter()
    while done < G and time.perf_counter() - t0 < args.cpu_seconds:
        g = circuit[done]
        pos = [q - 1 for q in g.ordered_positions()]
        oracle.apply_dense(v, len(pos), pos, g.ordered_mat())
        done += 1
    cpu_t = time.perf_counter() - t0
    cpu_gbs = bytes_per_gate * done / cpu_t / 1e9
    log(f"cpu oracle: {done} gates in {cpu_t:.1f}s = {cpu_gbs:.1f} GB/s on {cores} threads")

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
                     "kernel": "k_dense_small<1, 1, true> (1-qubit dense sweep)",
                     "avg_launch_ms": round(avg1_ms, 4), "bytes_per_launch": bytes_per_gate},
        "cpu_baseline": {"value": round(cpu_gbs, 2), "unit": "GB/s", "cores": cores, "kind": "port",
                         "sample": f"first {done} gates of the same {n}-qubit circuit, oracle restatement"},
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
