"""Fused dense blocks of 3 to 5 qubits on the matrix cores (MK_DENSE micro-ops):
the planner groups the gates of a tile phase into blocks and runs a block as one
FP64 MFMA product of its matrix. The blocks' results are checked against the
one-kernel-per-gate path and against the oracle, for state vectors (forward
and energy gradient) and density matrices, and across plan-cache reuse with new
matrix values. Tolerances (FP64): amplitudes 1e-12 absolute on normalised
states, energies and gradients 1e-10 relative to max(1, |value|)."""
import ctypes

import numpy as np
import pytest

import oracle
from brute import rand_state, rand_unitary, vec
import vqchip as vq
from vqchip import _lib
from vqchip.circuit import Program
from vqchip.operators import pack_terms
from vqchip.workloads import config1_circuit, config3, tfim

pytestmark = pytest.mark.gpu


def _plan_mops():
    info = _lib.PlanInfo()
    _lib.call("vqc_plan_info", ctypes.byref(info))
    return list(info.mops)


def _per_gate(circuit, n, psi):
    s = vq.StateVector(psi)
    for g in circuit:
        pos = g.ordered_positions()
        U = np.ascontiguousarray(g.ordered_mat().reshape(-1, order="F"))
        _lib.call("vqc_sv_apply_dense", s.handle, len(pos), _lib.i32ptr([q - 1 for q in pos]), _lib.dptr(U))
    return s.storage()


def _haar(n, layers, seed, rng):
    return config1_circuit(seed=seed, n=n, layers=layers)


@pytest.mark.parametrize("n,layers", [(13, 4), (16, 6), (20, 3), (22, 2)])
def test_fused_dense_forward_matches_per_gate_and_oracle(n, layers):
    rng = np.random.default_rng(n)
    c = config1_circuit(seed=n, n=n, layers=layers)
    psi = rand_state(n, rng)
    s = vq.StateVector(psi)
    vq.apply_(c, s)
    mops = _plan_mops()
    got = s.storage()
    assert np.max(np.abs(got - _per_gate(c, n, psi))) < 1e-12
    if n <= 16:
        prog = Program(c, n=n)
        ref = oracle.apply_circuit(psi.copy(), n, 0, prog.op_array(), len(prog.ops), prog.pool())
        assert np.max(np.abs(got - ref)) < 1e-12
    if n >= 16:
        assert mops[7] > 0, f"no dense blocks planned: {mops}"


def test_dense_blocks_of_3_4_and_5_qubits():
    """Circuits built to fill 3-, 4- and 5-qubit blocks (gates confined to a few
    qubit groups), forward against the per-gate path."""
    n = 18
    rng = np.random.default_rng(5)
    for width in (3, 4, 5):
        c = vq.QCircuit()
        groups = [list(range(1 + width * b, 1 + width * (b + 1))) for b in range(n // width)]
        for _ in range(6):
            for grp in groups:
                k = int(rng.integers(1, 4))
                pos = sorted(rng.choice(grp, size=min(k, len(grp)), replace=False).tolist())
                c.append(vq.QuantumGate(tuple(pos), rand_unitary(len(pos), rng)))
        psi = rand_state(n, rng)
        s = vq.StateVector(psi)
        vq.apply_(c, s)
        mops = _plan_mops()
        assert np.max(np.abs(s.storage() - _per_gate(c, n, psi))) < 1e-12
        assert mops[7] > 0


@pytest.mark.parametrize("n", [12, 15])
def test_energy_gradient_with_dense_forward_blocks(n):
    rng = np.random.default_rng(100 + n)
    c = config1_circuit(seed=3, n=n, layers=3)
    c.append(vq.RxGate(1, 0.3))
    c.append(vq.RyGate(n, 1.1))
    c.insert(10, vq.RzGate(2, 0.7))
    H = tfim(n, 0.9)
    psi = rand_state(n, rng)
    E, g = vq.energy_and_gradient(H, c, vq.StateVector(psi))
    prog = Program(c, n=n)
    terms, nt, hpool = pack_terms(H)
    Er, gr, _ = oracle.energy_grad(psi, n, 0, prog.op_array(), len(prog.ops), prog.pool(), terms, nt, hpool, prog.nparams)
    assert abs(E - Er) < 1e-10 * max(1, abs(Er))
    assert np.max(np.abs(g - gr)) < 1e-10 * max(1, np.max(np.abs(gr)))


def test_density_matrix_forward_with_dense_blocks():
    n = 7
    c, H, _ = config3(seed=2, n=n, depth=2, p=0.05, gamma=0.03)
    rng = np.random.default_rng(7)
    for _ in range(4):
        pos = tuple(sorted(rng.choice(np.arange(1, n + 1), size=3, replace=False).tolist()))
        c.append(vq.QuantumGate(pos, rand_unitary(3, rng)))
    rho0 = np.zeros(1 << (2 * n), dtype=complex)
    rho0[0] = 1
    E, g, out = vq.energy_and_gradient(H, c, vq.DensityMatrix(rho0), return_state=True)
    prog = Program(c, n=n, dm=True)
    terms, nt, hpool = pack_terms(H)
    Er, gr, rr = oracle.energy_grad(rho0, n, 1, prog.op_array(), len(prog.ops), prog.pool(), terms, nt, hpool, prog.nparams)
    assert np.max(np.abs(vec(out.storage()) - rr)) < 1e-12
    assert abs(E - Er) < 1e-10 and np.max(np.abs(g - gr)) < 1e-10 * max(1, np.max(np.abs(gr)))


def test_plan_cache_recomputes_dense_blocks_with_new_matrices():
    n = 16
    rng = np.random.default_rng(9)
    c1 = config1_circuit(seed=11, n=n, layers=3)
    # the same structure, other matrices
    c2 = vq.QCircuit(vq.QuantumGate(g.positions, rand_unitary(len(g.positions), rng)) for g in c1)
    psi = rand_state(n, rng)
    for c in (c1, c2, c1):
        s = vq.StateVector(psi)
        vq.apply_(c, s)
        assert np.max(np.abs(s.storage() - _per_gate(c, n, psi))) < 1e-12
    info = _lib.PlanInfo()
    _lib.call("vqc_plan_info", ctypes.byref(info))
    assert info.reused == 1
