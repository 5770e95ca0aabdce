"""The library's host state under concurrent callers (CPU only, no device call).

Several threads plan programs at once (vqc_plan_check: the planner, the matrix
pools and their structure claims, the per-size tile table) while others rewrite
the tuning and the tile table (vqc_set_tuning, vqc_set_tile_table). ctypes drops
the interpreter lock around each call, so the calls overlap inside the library.
Every plan must equal the one the same program gets alone. The module also runs
against the ThreadSanitizer build of the library in tests/test_sanitizers.py,
which fails on any data race the overlap exposes.
"""
import ctypes
import threading

import numpy as np

from test_planner import counts, plan
from vqchip import _lib
from vqchip.autotune import get_tile_table, get_tuning, set_tile_table, set_tuning
from vqchip.workloads import config0, config2_instances, config3

ROUNDS = 6


def _programs():
    progs = [(*config0()[:2], 12, False), (*config3()[:2], 12, True)]
    progs += [(c, H, n, False) for c, H, n in config2_instances()[:6]]
    return progs


def test_concurrent_planning_matches_serial():
    progs = _programs()
    serial = [counts(plan(c, H, n, dm=dm)) for c, H, n, dm in progs]
    errors, results = [], {}

    def worker(k):
        try:
            for r in range(ROUNDS):
                i = (k + r) % len(progs)
                c, H, n, dm = progs[i]
                results.setdefault(i, []).append(counts(plan(c, H, n, dm=dm)))
        except Exception as e:  # reported below, in the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(6)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for i, got in results.items():
        for g in got:
            assert g == serial[i], (i, g, serial[i])


def test_planning_while_tables_change():
    """Planners read the tuning and the tile table while other threads write them;
    each plan sees one consistent table (the shapes may differ between plans)."""
    c, H, _ = config0()
    old_tuning, old_table = get_tuning(), get_tile_table()
    stop = threading.Event()
    errors = []

    def writer(k):
        try:
            r = 0
            while not stop.is_set():
                set_tuning(tile_low_bits=3 + (r + k) % 3)
                set_tile_table({12: (9 + r % 2, 8 + r % 2, 3)})
                get_tuning()
                get_tile_table()
                r += 1
        except Exception as e:
            errors.append(e)

    def planner():
        try:
            for _ in range(ROUNDS):
                info = plan(c, H, 12)
                assert info.phases_fwd > 0 and info.phases_rev > 0
        except Exception as e:
            errors.append(e)

    writers = [threading.Thread(target=writer, args=(k,)) for k in range(2)]
    planners = [threading.Thread(target=planner) for _ in range(4)]
    try:
        for t in writers + planners:
            t.start()
        for t in planners:
            t.join()
    finally:
        stop.set()
        for t in writers:
            t.join()
        set_tile_table({12: old_table.get(12, (0, 0, 0))})
        set_tuning(**old_tuning)
    assert not errors, errors


def test_plan_info_is_per_thread():
    """vqc_plan_info reports the calling thread's last plan (a thread-local record)."""
    info = _lib.PlanInfo()
    _lib.call("vqc_plan_info", ctypes.byref(info))
    assert np.all(np.array(list(info.mops)) >= 0)
