"""Batched persistent-kernel evaluation against the oracle, per instance.

Tolerance: energies and gradients within 1e-10 (relative to max(1, |value|)).
"""
import numpy as np
import pytest

import oracle
from brute import rand_unitary
from circuits import chain_circuit, structured_circuit
import vqchip as vq
from vqchip.batch import Batch
from vqchip.circuit import Program
from vqchip.operators import pack_terms
from vqchip.workloads import hea_circuit, heisenberg, tfim

pytestmark = pytest.mark.gpu


def oracle_eval(c, H, n):
    prog = Program(c, n=n)
    terms, nt, hpool = pack_terms(H)
    psi0 = np.zeros(1 << n, dtype=complex)
    psi0[0] = 1
    E, g, _ = oracle.energy_grad(psi0, n, 0, prog.op_array(), len(prog.ops), prog.pool(), terms, nt, hpool, prog.nparams)
    return E, g


def make_instances(seed, ns):
    rng = np.random.default_rng(seed)
    out = []
    for n in ns:
        depth = int(rng.integers(1, 4))
        c = hea_circuit(n, depth, rng.uniform(0, 2 * np.pi, 3 * n * depth))
        if n >= 5 and rng.random() < 0.5:
            c.append(vq.QuantumGate((1, 3, n - 1, n), rand_unitary(4, rng)))
        if n >= 6 and rng.random() < 0.3:
            c.append(vq.QuantumGate((1, 2, 4, n - 1, n), rand_unitary(5, rng)))
        if rng.random() < 0.3:
            c.insert(1, vq.FSIMGate(1, n, 0.4, 0.9))
        H = heisenberg(n, rng.standard_normal(n - 1)) if rng.random() < 0.5 else tfim(n, float(rng.uniform(0.5, 1.5)))
        out.append((c, H, n))
    return out


def check(instances, E, grads):
    for i, (c, H, n) in enumerate(instances):
        Er, gr = oracle_eval(c, H, n)
        assert abs(E[i] - Er) < 1e-10 * max(1, abs(Er)), (i, n)
        assert grads[i].shape == gr.shape
        if gr.size:
            assert np.max(np.abs(grads[i] - gr)) < 1e-10 * max(1, np.max(np.abs(gr))), (i, n)


def test_mixed_batch_matches_oracle():
    ns = [3, 8, 12, 13, 5, 16, 9, 14, 11, 17, 4, 15, 10, 12, 6]
    inst = make_instances(1, ns)
    b = Batch(inst)
    E, grads = b.run()
    check(inst, E, grads)
    assert b.arena_bytes() == sum(32 << n for n in ns if n > 12)


def test_graph_replay_with_new_angles():
    ns = [8, 14, 12, 15, 7]
    rng = np.random.default_rng(2)
    Hs = [heisenberg(n, rng.standard_normal(n - 1)) for n in ns]
    inst = [(hea_circuit(n, 2, rng.uniform(0, 6.3, 6 * n)), H, n) for n, H in zip(ns, Hs)]
    b = Batch(inst)
    E1, g1 = b.run()
    check(inst, E1, g1)
    inst2 = [(hea_circuit(n, 2, rng.uniform(0, 6.3, 6 * n)), H, n) for n, H in zip(ns, Hs)]
    b.update([c for c, _, _ in inst2])
    for _ in range(3):  # every replay agrees with the oracle
        E2, g2 = b.run()
        check(inst2, E2, g2)


def test_alone_equals_in_mix():
    ns = [9, 13, 11, 15]
    inst = make_instances(4, ns)
    E_mix, g_mix = Batch(inst).run()
    for i, one in enumerate(inst):
        E1, g1 = Batch([one]).run()
        assert abs(E1[0] - E_mix[i]) < 1e-12
        assert np.max(np.abs(g1[0] - g_mix[i])) < 1e-12


def test_instance_without_ops_or_terms():
    inst = [(vq.QCircuit(), tfim(5), 5), (vq.QCircuit(), tfim(14), 14), (hea_circuit(6, 1, np.ones(18)), vq.QubitsOperator(), 6)]
    E, grads = Batch(inst).run()
    check(inst, E, grads)


def test_same_qubit_runs_in_batch():
    ns = (4, 9, 12, 14, 16)
    vr = np.random.default_rng(10)
    inst = [(chain_circuit(n, np.random.default_rng(n), value_rng=vr), tfim(n, 0.8), n) for n in ns]
    b = Batch(inst)
    E, grads = b.run()
    check(inst, E, grads)
    inst2 = [(chain_circuit(n, np.random.default_rng(n), value_rng=vr), H, n) for _, H, n in inst]
    b.update([c for c, _, _ in inst2])
    E2, g2 = b.run()
    check(inst2, E2, g2)


def test_structure_hints_in_batch():
    # batches take gate structure from the hints (their matrices arrive at launch)
    ns = (5, 9, 12, 14, 15)
    vr = np.random.default_rng(20)
    inst = [(structured_circuit(n, np.random.default_rng(n), value_rng=vr), heisenberg(n, np.ones(n - 1)), n) for n in ns]
    b = Batch(inst)
    check(inst, *b.run())
    inst2 = [(structured_circuit(n, np.random.default_rng(n), value_rng=vr), H, n) for _, H, n in inst]
    b.update([c for c, _, _ in inst2])
    check(inst2, *b.run())


def test_batch_beyond_memory_limit_runs_in_waves():
    ns = [13, 15, 9, 14, 16, 12, 15, 13]
    inst = make_instances(7, ns)
    one = Batch(inst)
    E1, g1 = one.run()
    limit = 3 << 19  # 1.5 MiB: a 16-qubit instance (2 MiB) goes alone, the rest in small groups
    b = Batch(inst, max_bytes=limit)
    assert b.nwaves() > 1 and b.arena_bytes() <= max(limit, 32 << 16)
    E2, g2 = b.run()
    assert np.max(np.abs(E2 - E1)) < 1e-12
    for x, y in zip(g1, g2):
        assert x.shape == y.shape and (x.size == 0 or np.max(np.abs(x - y)) < 1e-12)
    check(inst, E2, g2)


def test_wide_batch_matches_oracle():
    # 16 or more instances take the wide-batch tile shape (13 forward / 12 reverse
    # tile bits, 5 contiguous low bits); instances above 12 qubits run in tile phases
    ns = [13, 5, 14, 9, 15, 12, 16, 7, 13, 11, 14, 8, 15, 10, 13, 6, 16, 12, 14, 9]
    inst = make_instances(11, ns)
    b = Batch(inst)
    E, grads = b.run()
    check(inst, E, grads)
    E1, g1 = Batch(inst[:15]).run()  # the same instances in a narrow batch
    assert np.max(np.abs(np.asarray(E[:15]) - np.asarray(E1))) < 1e-12
    for x, y in zip(grads[:15], g1):
        assert x.size == 0 or np.max(np.abs(x - y)) < 1e-12


def test_batch_rejects_more_than_three_parameters():
    """vqc_batch_create validates every op as the single-state paths do: an op
    with nparams = 4 (slot[] holds 3) is rejected with VQC_EINVAL."""
    import ctypes
    from vqchip import _lib
    c = hea_circuit(3, 1, np.linspace(0.1, 0.9, 9))
    prog = Program(c, n=3)
    ops = prog.op_array()
    ops[0].nparams = 4
    terms, nt, hpool = pack_terms(tfim(3))
    inst = (_lib.Instance * 1)()
    inst[0].nqubits, inst[0].nops, inst[0].nterms, inst[0].nparams = 3, len(prog.ops), nt, prog.nparams
    h = ctypes.c_void_p()
    with pytest.raises(_lib.VQCError) as e:
        _lib.call("vqc_batch_create", 1, inst, ops, len(prog.ops), terms, nt, len(prog.pool()), len(hpool), 0,
                  ctypes.byref(h))
    assert e.value.code == -1


def test_toffoli_batch_matches_oracle():
    """The Toffoli variant of the uneven batch (3-qubit gates in every instance,
    the 3-bit persistent kernel), at sizes the oracle finishes in seconds."""
    from vqchip.workloads import config2_toffoli_instances
    inst = config2_toffoli_instances(count=24, nmin=5, nmax=13, dmax=6)
    E, grads = Batch(inst).run()
    worst = 0.0
    for i, (c, H, n) in enumerate(inst):
        Er, gr = oracle_eval(c, H, n)
        worst = max(worst, abs(E[i] - Er) / max(1.0, abs(Er)), float(np.max(np.abs(grads[i] - gr))) / max(1.0, float(np.max(np.abs(gr)))))
    print(f"toffoli batch: 24 instances, max error vs the oracle {worst:.2e}", flush=True)
    assert worst < 1e-10
