"""Autotuner for the fusion width and tile shape of the tiled energy + gradient
path, and for the launch shape of the dense gate sweeps (no counterpart in the
reference, whose kernels have fixed shapes).

    best = tune_energy_grad(H, circuit, state0)     # sweeps tile bits / low bits (process-wide)
    table = tune_sizes([10, 12, ..., 20])           # per-size tile table (vqc_set_tile_table)
    best = tune_batch(batch)                        # the wide-batch entry of that table
    best = tune_dense(28)                           # sweeps groups per thread / NT
    load_cached(key) / the JSON cache at ~/.cache/vqchip/tuning.json

The per-size table is what the library's own callers read: single-state energy
gradients, the tiled instances of narrow batches and the serving loop's
requests take the entry of their size, wide batches the batch entry.

Each candidate is applied with vqc_set_tuning, timed on the device (median of
`reps` runs after one warm-up), and the fastest is left in force. Results are
identical across tunings up to floating-point summation order: every
candidate's energy is checked against the first one's to 1e-10.
"""
import ctypes
import json
import os
import time

import numpy as np

from . import _lib

CACHE = os.path.join(os.path.expanduser("~"), ".cache", "vqchip", "tuning.json")
FIELDS = ("tile_fwd_bits", "tile_rev_bits", "tile_small_bits", "tile_low_bits")


def get_tuning():
    t = _lib.Tuning()
    _lib.call("vqc_get_tuning", ctypes.byref(t))
    d = {k: getattr(t, k) for k in FIELDS}
    d["dense_groups"] = list(t.dense_groups)
    d["dense_nontemporal"] = list(t.dense_nontemporal)
    return d


def set_tuning(**kw):
    """Set any subset of the fields of vqc_tuning_t; the rest keep their values."""
    t = _lib.Tuning()
    for k in FIELDS:
        setattr(t, k, int(kw.get(k, -1)))
    for k in ("dense_groups", "dense_nontemporal"):
        v = kw.get(k, [-1, -1, -1])
        arr = getattr(t, k)
        for i in range(3):
            arr[i] = int(v[i])
    _lib.call("vqc_set_tuning", ctypes.byref(t))


def _prepared_eval(H, circuit, state0):
    """A callable running the fused energy + gradient of a packed program and
    returning the energy: the candidates are timed on the library call alone,
    without packing the circuit each time (host work of the same order as a small
    evaluation, which would hide the differences being measured). Operators with
    keys of more than 5 sites take the general route of energy_and_gradient."""
    from .api import LARGEST_SUPPORTED_NTERMS, _as_operator, energy_and_gradient
    from .circuit import Program
    from .gates import QuantumMap
    from .operators import pack_terms
    H = _as_operator(H)
    if H.largest_nterm() > LARGEST_SUPPORTED_NTERMS or (not state0.is_dm and any(
            isinstance(g, QuantumMap) for g in circuit)):
        return lambda: energy_and_gradient(H, circuit, state0)[0]
    dm = bool(state0.is_dm)
    prog = Program(circuit, n=state0.nqubits(), dm=dm)
    ops, pool = prog.op_array(), prog.pool()
    terms, nt, hpool = pack_terms(H)
    grad = np.zeros(max(prog.nparams, 1))
    e2 = np.zeros(2)
    fn = "vqc_dm_energy_grad" if dm else "vqc_sv_energy_grad"

    def run():
        _lib.call(fn, state0.handle, len(prog.ops), ops, _lib.dptr(pool), nt, terms, _lib.dptr(hpool), prog.nparams,
                  _lib.dptr(e2), _lib.dptr(grad), None, 0)
        return complex(e2[0], e2[1])
    return run


def _time(fn, reps):
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def tune_energy_grad(H, circuit, state0, fwd=(10, 11, 12), rev=(10, 11, 12), low=(3, 4, 5, 6), reps=3, apply=True,
                     cache_key=None):
    """Sweep the tile shape of the fused energy + gradient path on this workload.
    Returns {"best": {...}, "seconds": t, "table": [(params, seconds), ...]}."""
    evaluate = _prepared_eval(H, circuit, state0)
    before = get_tuning()
    table = []
    ref = None
    for f in fwd:
        for r in rev:
            for m in low:
                if m > min(f, r):
                    continue
                set_tuning(tile_fwd_bits=f, tile_rev_bits=r, tile_low_bits=m)
                out = {}

                def run():
                    out["E"] = evaluate()

                t = _time(run, reps)
                if ref is None:
                    ref = out["E"]
                elif abs(out["E"] - ref) > 1e-10 * max(1.0, abs(ref)):
                    raise RuntimeError(f"tuning {(f, r, m)} changed the energy: {out['E']} vs {ref}")
                table.append(({"tile_fwd_bits": f, "tile_rev_bits": r, "tile_low_bits": m}, t))
    best, tbest = min(table, key=lambda x: x[1])
    set_tuning(**(best if apply else before))
    if cache_key is not None:
        store_cached(cache_key, best)
    return {"best": best, "seconds": tbest, "table": table}


def tune_dense(n, reps=3, apply=True, cache_key=None, seed=7):
    """Sweep groups per thread and non-temporal hints of the 1-, 2- and 3-qubit
    dense sweeps on an n-qubit state (each class on a fixed target set)."""
    from .state import StateVector
    rng = np.random.default_rng(seed)
    s = StateVector(n)
    before = get_tuning()
    groups, nts, table = list(before["dense_groups"]), list(before["dense_nontemporal"]), []
    stream = ctypes.c_void_p()
    _lib.call("vqc_get_stream", ctypes.byref(stream))
    for k in (1, 2, 3):
        pos = _lib.i32ptr(sorted(rng.choice(n, size=k, replace=False).tolist()))
        D = 1 << k
        U = np.linalg.qr(rng.standard_normal((D, D)) + 1j * rng.standard_normal((D, D)))[0]
        Uf = np.ascontiguousarray(U.reshape(-1, order="F"))
        best = None
        for r in (1, 2, 4, 8):
            for nt in (0, 1):
                g = list(groups)
                ntl = list(nts)
                g[k - 1], ntl[k - 1] = r, nt
                set_tuning(dense_groups=g, dense_nontemporal=ntl)

                def run():
                    for _ in range(4):
                        _lib.call("vqc_sv_apply_dense", s.handle, k, pos, _lib.dptr(Uf))
                    _lib.call("vqc_sync")

                t = _time(run, reps) / 4
                table.append(({"k": k, "groups": r, "nontemporal": nt}, t))
                if best is None or t < best[1]:
                    best = ((r, nt), t)
        groups[k - 1], nts[k - 1] = best[0]
    chosen = {"dense_groups": groups, "dense_nontemporal": nts}
    set_tuning(**(chosen if apply else {"dense_groups": before["dense_groups"],
                                         "dense_nontemporal": before["dense_nontemporal"]}))
    if cache_key is not None:
        store_cached(cache_key, chosen)
    return {"best": chosen, "table": table, "bytes_per_sweep": 32 << n}


def load_cached(key, path=CACHE):
    """Apply and return the cached tuning for `key`, or None."""
    try:
        with open(path) as f:
            d = json.load(f).get(key)
    except (OSError, ValueError):
        return None
    if d:
        set_tuning(**d)
    return d


def store_cached(key, params, path=CACHE):
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        d = {}
    d[key] = params
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)


def get_tile_table():
    """{nbits: (fwd, rev, low)} of the entries in force (nbits -1: wide batches)."""
    arr = (_lib.TileEntry * 35)()
    _lib.call("vqc_get_tile_table", arr)
    return {(-1 if i == 0 else i): (arr[i].fwd, arr[i].rev, arr[i].low) for i in range(35)
            if arr[i].fwd or arr[i].rev or arr[i].low}


def set_tile_table(table):
    """Set entries {nbits: (fwd, rev, low)} (0 keeps the rule for a field; nbits -1: wide batches)."""
    items = sorted(table.items())
    nb = (ctypes.c_int32 * max(len(items), 1))(*[int(k) for k, _ in items])
    ents = (_lib.TileEntry * max(len(items), 1))()
    for i, (_, (f, r, l)) in enumerate(items):
        ents[i].fwd, ents[i].rev, ents[i].low = int(f), int(r), int(l)
    _lib.call("vqc_set_tile_table", len(items), nb, ents)


def _rule(nbits):
    """The built-in tile rule for one state of nbits index bits (forward, reverse)."""
    f = 9 if nbits <= 14 else (10 if nbits <= 18 else (11 if nbits <= 20 else 12))
    return f, min(f, 11)


def tune_sizes(sizes, workload=None, reps=3, lows=(3, 4, 5), spread=(-1, 0, 1), apply=True, cache_key=None):
    """Tile shape per state size: for each n, energy + gradient of `workload(n)`
    ((circuit, H, state0); by default a hardware-efficient ansatz of 4 layers with
    a transverse-field Ising H) is timed for forward / reverse tile bits around the
    built-in rule and the contiguous low bits given, and the fastest shape becomes
    the size's entry of the library's tile table. Sizes of at most 9 bits run in
    one workgroup and are left out. Returns {n: {"best": (fwd, rev, low), "table": [...]}}."""
    from .state import StateVector
    from .workloads import hea_circuit, tfim
    if workload is None:
        def workload(n):
            th = np.random.default_rng(n).uniform(0, 2 * np.pi, 3 * n * 4)
            return hea_circuit(n, 4, th), tfim(n, 1.0), StateVector(n)
    before = get_tuning()
    set_tuning(tile_fwd_bits=0, tile_rev_bits=0)  # per-size entries apply only without a process-wide choice
    out = {}
    for n in sizes:
        if n <= 9:
            continue
        c, H, s0 = workload(n)
        evaluate = _prepared_eval(H, c, s0)
        f0, r0 = _rule(n)
        table, ref = [], None
        for df in spread:
            for dr in spread:
                f, r = f0 + df, r0 + dr
                if not (6 <= f <= 13 and 6 <= r <= 12) or f > n or r > n:
                    continue
                for low in lows:
                    if low > min(f, r):
                        continue
                    set_tile_table({n: (f, r, low)})
                    res = {}

                    def run():
                        res["E"] = evaluate()

                    t = _time(run, reps)
                    if ref is None:
                        ref = res["E"]
                    elif abs(res["E"] - ref) > 1e-10 * max(1.0, abs(ref)):
                        raise RuntimeError(f"tile shape {(f, r, low)} changed the energy at n = {n}")
                    table.append(((f, r, low), t))
        best = min(table, key=lambda x: x[1])[0]
        set_tile_table({n: best if apply else (0, 0, 0)})
        out[n] = {"best": best, "table": table}
    set_tuning(tile_fwd_bits=before["tile_fwd_bits"], tile_rev_bits=before["tile_rev_bits"])
    if cache_key is not None:
        store_cached(cache_key, {"tile_table": {str(n): list(v["best"]) for n, v in out.items()}})
    return out


def tune_batch(batch, fwd=(12, 13), rev=(11, 12), lows=(4, 5, 6), reps=3, apply=True):
    """The wide-batch tile shape: the batch is planned again for each candidate
    (its packed records are kept) and replayed; the fastest becomes the table's
    wide-batch entry, which vqc_batch_create and the serving loop's wide batches
    read. Results are checked equal across candidates. Returns {"best": ..., "table": ...}."""
    before = get_tile_table().get(-1, (0, 0, 0))
    table, ref = [], None
    for f in fwd:
        for r in rev:
            for low in lows:
                set_tile_table({-1: (f, r, low)})
                batch.rebuild()
                E, _ = batch.run()
                if ref is None:
                    ref = E
                elif np.max(np.abs(E - ref)) > 1e-10 * max(1.0, float(np.max(np.abs(ref)))):
                    raise RuntimeError(f"batch tile shape {(f, r, low)} changed the energies")
                table.append(((f, r, low), _time(batch.run, reps)))
    best = min(table, key=lambda x: x[1])[0]
    set_tile_table({-1: best if apply else before})
    batch.rebuild()
    return {"best": best, "table": table}


def tune_forward(circuit, state, fwd=(11, 12, 13), lows=(3, 4, 5), reps=3, apply=True):
    """Tile shape of the fused forward pass (apply!(circuit, state) on the tiled
    path) for the state's size: each candidate's forward tile bits and contiguous
    low bits are timed on a copy of the state and the fastest becomes the size's
    entry of the tile table (its reverse field is kept). Results are checked equal
    across candidates. Returns {"best": (fwd, low), "table": [...]}."""
    from .api import apply_
    nb = state.nqubits() * (2 if state.is_dm else 1)
    keep = get_tile_table().get(nb, (0, 0, 0))
    table, ref = [], None
    for f in fwd:
        for low in lows:
            if low > f or f > nb:
                continue
            set_tile_table({nb: (f, keep[1], low)})
            work = state.copy()
            apply_(circuit, work)
            v = work.dot(work)
            if ref is None:
                ref = work.copy()  # the timing below keeps applying the circuit to work
            else:
                d = work.copy()
                from . import _lib
                _lib.call("vqc_state_axpy", d.handle, -1.0, 0.0, ref.handle)
                if abs(d.dot(d)) > 1e-20 * max(1.0, abs(v)):
                    raise RuntimeError(f"forward tile shape {(f, low)} changed the state")

            def run():
                apply_(circuit, work)
                from . import _lib
                _lib.call("vqc_sync")

            table.append(((f, low), _time(run, reps)))
    best = min(table, key=lambda x: x[1])[0]
    set_tile_table({nb: (best[0], keep[1], best[1]) if apply else keep})
    return {"best": best, "table": table}
