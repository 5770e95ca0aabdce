"""Host code under AddressSanitizer, UndefinedBehaviorSanitizer and
ThreadSanitizer (CPU only).

* The oracle: oracle/liboracle_san.so (the same sources, `make liboracle_san.so`)
  runs the oracle's own test modules.
* The library's host logic: vqc.jl_amd/_build/san/libvqc_hip_san.so (the planner,
  the batch planner, the tuning tables and the serving scheduler compiled with the
  sanitizers on their host side, build_lib.build_all) runs the planner tests, which
  make no device call (vqc_plan_check).
* The same units built with ThreadSanitizer (vqc.jl_amd/_build/tsan/
  libvqc_hip_tsan.so, gcc's runtime preloaded) run the concurrent-caller tests of
  tests/test_host_threads.py: planning from several threads while others rewrite
  the tuning and the tile table.

Each run is a subprocess with the sanitizer runtime preloaded (the Python
interpreter itself is not instrumented); a report, or any failing test, fails
the run. The builds are made by __graft_entry__.build().
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SAN = os.path.join(ROOT, "oracle", "liboracle_san.so")
LIB_SAN = os.path.join(ROOT, "vqc.jl_amd", "_build", "san", "libvqc_hip_san.so")
LIB_TSAN = os.path.join(ROOT, "vqc.jl_amd", "_build", "tsan", "libvqc_hip_tsan.so")
ENV = {"ASAN_OPTIONS": "detect_leaks=0:verify_asan_link_order=0:abort_on_error=0",
       "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
       "TSAN_OPTIONS": "halt_on_error=0:report_signal_unsafe=0"}
REPORTS = ("AddressSanitizer", "runtime error:", "ThreadSanitizer")


def _gcc_asan():
    r = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    p = r.stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def _gcc_tsan():
    r = subprocess.run(["gcc", "-print-file-name=libtsan.so"], capture_output=True, text=True)
    p = r.stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def _clang_asan():
    import glob
    found = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return found[-1] if found else None


def _run(preload, extra_env, tests, probe):
    # the sanitizer runtime first; whatever the environment already preloads stays
    kept = os.environ.get("LD_PRELOAD", "")
    env = dict(os.environ, LD_PRELOAD=preload + (":" + kept if kept else ""), **ENV, **extra_env)
    # the sanitized library must be the one the process maps
    r = subprocess.run([sys.executable, "-c", probe], cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("True"), r.stdout + r.stderr
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu"] + tests,
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert not any(w in out for w in REPORTS), out[-4000:]
    assert r.returncode == 0, out[-4000:]
    assert " passed" in out, out[-2000:]


def test_oracle_under_address_and_undefined_sanitizers():
    rt = _gcc_asan()
    if rt is None or not os.path.exists(ORACLE_SAN):
        pytest.skip("no gcc AddressSanitizer runtime or no oracle/liboracle_san.so (run __graft_entry__.build())")
    probe = ("import sys; sys.path.insert(0, 'oracle'); import oracle; oracle.lib(); "
             "print(any('liboracle_san.so' in l for l in open('/proc/self/maps')))")
    _run(rt, {"VQC_ORACLE_LIB": ORACLE_SAN},
         ["tests/test_oracle.py", "tests/test_vjp.py", "tests/test_golden.py", "tests/test_algorithms.py"], probe)


def test_planner_under_address_and_undefined_sanitizers():
    rt = _clang_asan()
    if rt is None or not os.path.exists(LIB_SAN):
        pytest.skip("no clang AddressSanitizer runtime or no libvqc_hip_san.so (run __graft_entry__.build())")
    probe = ("import sys; sys.path.insert(0, 'vqc.jl_amd'); from vqchip import _lib; _lib.load(); "
             "print(any('libvqc_hip_san.so' in l for l in open('/proc/self/maps')))")
    _run(rt, {"VQC_LIB_PATH": LIB_SAN}, ["tests/test_planner.py"], probe)


def test_concurrent_callers_under_thread_sanitizer():
    rt = _gcc_tsan()
    if rt is None or not os.path.exists(LIB_TSAN):
        pytest.skip("no gcc ThreadSanitizer runtime or no libvqc_hip_tsan.so (run __graft_entry__.build())")
    probe = ("import sys; sys.path.insert(0, 'vqc.jl_amd'); from vqchip import _lib; _lib.load(); "
             "print(any('libvqc_hip_tsan.so' in l for l in open('/proc/self/maps')))")
    _run(rt, {"VQC_LIB_PATH": LIB_TSAN}, ["tests/test_host_threads.py", "tests/test_planner.py"], probe)
