"""The tile planner on the host (vqc_plan_check: no device call), over the
benchmarked programs and random ones: every plan the fused energy + gradient
would run keeps its invariants (contiguous phases, tile bits partitioning the
index bits, micro-op bits inside their tile, matrices inside the pool,
accumulators inside their rows). The same tests run against the host sanitizer
build of the library in tests/test_sanitizers.py."""
import ctypes

import numpy as np
import pytest

import vqchip as vq
from brute import rand_matrix
from vqchip import _lib
from vqchip.circuit import Program
from vqchip.operators import pack_terms
from vqchip.workloads import config0, config1_circuit, config2_instances, config3, hea_circuit, tfim

FLAG_CHECKPOINT = 2


def plan(circuit, H, n, dm=False, flags=0):
    prog = Program(circuit, n=n, dm=dm)
    if H is not None:
        terms, nt, hpool = pack_terms(H)
    else:
        terms, nt, hpool = None, 0, np.zeros(2)
    info = _lib.PlanInfo()
    _lib.call("vqc_plan_check", 1 if dm else 0, n, len(prog.ops), prog.op_array(), _lib.dptr(prog.pool()), nt, terms,
              _lib.dptr(hpool), prog.nparams, flags, ctypes.byref(info))
    return info


def counts(info):
    return info.phases_fwd, info.phases_ham, info.phases_rev, list(info.mops)


def test_benchmarked_programs():
    c, H, _ = config0()
    info = plan(c, H, 12)
    assert info.phases_fwd > 0 and info.phases_rev > 0 and info.phases_ham > 0
    c, H, _ = config3()
    info = plan(c, H, 12, dm=True)
    assert info.phases_fwd > 0 and info.phases_rev > 0
    # the 28-qubit circuit's forward fuses dense blocks on the matrix cores (micro-op kind 7)
    info = plan(config1_circuit(n=28, layers=20), None, 28)
    assert info.mops[7] > 0


def test_uneven_batch_instances():
    for c, H, n in config2_instances()[:48]:
        plan(c, H, n)


@pytest.mark.parametrize("flags", [0, FLAG_CHECKPOINT])
def test_checkpointed_plans(flags):
    c, H, _ = config0()
    a, b = counts(plan(c, H, 12)), counts(plan(c, H, 12, flags=flags))
    if flags:
        # segments every ceil(sqrt(#ops)) ops: more phases than the plain plan
        assert b[0] > a[0] and b[2] > a[2]
    else:
        assert a == b
    c, H, _ = config3()
    plan(c, H, 12, dm=True, flags=flags)


@pytest.mark.parametrize("n", [3, 6, 9, 11, 14, 17, 20, 24])
def test_random_dense_circuits(n):
    """Random 1..5-qubit dense gates on random targets (every micro-op width and the
    dense-block fusion), with an Ising Hamiltonian, at sizes that cross the small /
    tiled boundary and the tile-size bands."""
    rng = np.random.default_rng(100 + n)
    c = vq.QCircuit()
    for _ in range(40):
        k = int(rng.integers(1, min(5, n) + 1))
        pos = [int(x) + 1 for x in rng.choice(n, size=k, replace=False)]
        c.append(vq.QuantumGate(pos, rand_matrix(k, rng)))
    th = rng.uniform(0, 2 * np.pi, 3 * n * 2)
    for g in hea_circuit(n, 2, th):
        c.append(g)
    plan(c, tfim(n, 0.7), n)


@pytest.mark.parametrize("n", [2, 4, 6, 7])
def test_random_noisy_density_matrices(n):
    rng = np.random.default_rng(200 + n)
    th = rng.uniform(0, 2 * np.pi, 3 * n * 3)
    c = vq.QCircuit()
    for i, g in enumerate(hea_circuit(n, 3, th)):
        c.append(g)
        if i % 5 == 4:
            q = int(rng.integers(1, n + 1))
            c.append(vq.Depolarizing(q, 0.05) if i % 2 else vq.AmplitudeDamping(q, 0.03))
    for k in (3, 4):
        if n >= k:
            pos = [int(x) + 1 for x in rng.choice(n, size=k, replace=False)]
            c.append(vq.QuantumGate(pos, rand_matrix(k, rng)))
    plan(c, tfim(n, 1.1), n, dm=True)
    plan(c, tfim(n, 1.1), n, dm=True, flags=FLAG_CHECKPOINT)


def test_bad_programs_are_rejected():
    c, H, _ = config0()
    prog = Program(c, n=12)
    ops = prog.op_array()
    bad = (_lib.Op * len(prog.ops))()
    ctypes.memmove(bad, ops, ctypes.sizeof(bad))
    bad[3].pos[0] = 12  # a qubit outside the 12-qubit state
    terms, nt, hpool = pack_terms(H)
    info = _lib.PlanInfo()
    with pytest.raises(_lib.VQCError):
        _lib.call("vqc_plan_check", 0, 12, len(prog.ops), bad, _lib.dptr(prog.pool()), nt, terms, _lib.dptr(hpool),
                  prog.nparams, 0, ctypes.byref(info))
    with pytest.raises(_lib.VQCError):
        _lib.call("vqc_plan_check", 0, 0, len(prog.ops), ops, _lib.dptr(prog.pool()), nt, terms, _lib.dptr(hpool),
                  prog.nparams, 0, ctypes.byref(info))
