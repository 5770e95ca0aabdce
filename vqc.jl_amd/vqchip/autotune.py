"""Autotuner for the fusion width and tile shape of the tiled energy + gradient
path, and for the launch shape of the dense gate sweeps (no counterpart in the
reference, whose kernels have fixed shapes).

    best = tune_energy_grad(H, circuit, state0)     # sweeps tile bits / low bits
    best = tune_dense(28)                           # sweeps groups per thread / NT
    load_cached(key) / the JSON cache at ~/.cache/vqchip/tuning.json

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
    from .api import energy_and_gradient
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
                    out["E"] = energy_and_gradient(H, circuit, state0)[0]

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
