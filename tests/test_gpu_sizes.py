"""Edge sizes. Tiny states (1 and 2 qubits) against the oracle; the full
28-qubit workload through size-independent properties: the fused and the
one-kernel-per-gate paths agree (|a - b|^2 = 2 - 2 Re<a|b> from a device dot),
the norm is kept, and the circuit followed by its inverse returns |0...0>.
A 24-qubit energy and gradient: the fused and plain paths agree, and the
gradient matches central differences of the energy."""
import copy

import numpy as np
import pytest

import oracle
from brute import rand_state, rand_unitary
import vqchip as vq
from vqchip import _lib
from vqchip.circuit import Program
from vqchip.operators import pack_terms
from vqchip.workloads import config1_circuit, hea_circuit, tfim

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 2])
def test_tiny_states(n):
    rng = np.random.default_rng(200 + n)
    psi = rand_state(n, rng)
    for k in range(1, n + 1):
        U = rand_unitary(k, rng)
        s = vq.StateVector(psi)
        _lib.call("vqc_sv_apply_dense", s.handle, k, _lib.i32ptr(list(range(k))),
                  _lib.dptr(np.ascontiguousarray(U.reshape(-1, order="F"))))
        ref = oracle.apply_dense(psi.copy(), k, list(range(k)), U)
        assert np.max(np.abs(s.storage() - ref)) < 1e-14
    c = vq.QCircuit([vq.RyGate(1, 0.3), vq.RzGate(1, 1.1)] + ([vq.CNOTGate(1, 2), vq.RxGate(2, 0.7)] if n == 2 else []))
    H = tfim(n, 0.9) if n > 1 else vq.QubitsOperator([vq.QubitsTerm({1: "X"}, coeff=0.5)])
    for fused in (True, False):
        E, g = vq.energy_and_gradient(H, c, vq.StateVector(psi), fused=fused)
        prog = Program(c, n=n)
        terms, nt, hpool = pack_terms(H)
        Er, gr, _ = oracle.energy_grad(psi, n, 0, prog.op_array(), len(prog.ops), prog.pool(), terms, nt, hpool,
                                       prog.nparams)
        assert abs(E - Er) < 1e-12 and np.max(np.abs(g - gr)) < 1e-12
    d = vq.DensityMatrix(vq.StateVector(psi))
    vq.apply_(vq.RxGate(1, 0.4), d)
    assert abs(np.trace(d.storage()) - 1) < 1e-14


def test_full_size_28_qubits_properties():
    n = 28
    c = config1_circuit(n=n, layers=2)  # 48 Haar gates on 1-3 qubits
    prog = Program(c, n=n)
    ops, pool = prog.op_array(), prog.pool()
    a = vq.StateVector(n)
    _lib.call("vqc_apply_circuit", a.handle, len(prog.ops), ops, _lib.dptr(pool))
    b = vq.StateVector(n)
    for g in c:
        pos = [q - 1 for q in g.ordered_positions()]
        U = np.ascontiguousarray(g.ordered_mat().reshape(-1, order="F"))
        _lib.call("vqc_sv_apply_dense", b.handle, len(pos), _lib.i32ptr(pos), _lib.dptr(U))
    ab = a.dot(b)
    assert abs(a.norm() - 1) < 1e-12 and abs(b.norm() - 1) < 1e-12
    assert 2 - 2 * ab.real < 1e-20  # |a - b|^2
    # the inverse circuit (U^dagger in reverse order) brings the fused result back to |0...0>
    for g in reversed(list(c)):
        pos = [q - 1 for q in g.ordered_positions()]
        U = np.ascontiguousarray(g.ordered_mat().conj().T.reshape(-1, order="F"))
        _lib.call("vqc_sv_apply_dense", a.handle, len(pos), _lib.i32ptr(pos), _lib.dptr(U))
    e0 = vq.StateVector(n)
    assert abs(a.dot(e0) - 1) < 1e-12
    del a, b, e0


def test_energy_gradient_24_qubits_properties():
    # beyond the oracle's sizes: the fused tile engine against the plain path (one
    # kernel per gate and per reduction), and central differences of the energy
    n = 24
    rng = np.random.default_rng(2400)
    c = hea_circuit(n, 1, rng.uniform(0, 6.3, 3 * n))
    H = tfim(n, 0.8)
    Ef, gf = vq.energy_and_gradient(H, c, vq.StateVector(n), fused=True)
    Ep, gp = vq.energy_and_gradient(H, c, vq.StateVector(n), fused=False)
    assert len(gf) == 3 * n
    assert abs(Ef - Ep) < 1e-10 * max(1, abs(Ep))
    assert np.max(np.abs(gf - gp)) < 1e-10 * max(1, np.max(np.abs(gp)))
    x = np.array(c.active_parameters(), dtype=float)
    h = 1e-4
    for j in (0, 3 * n // 2 + 1, 3 * n - 1):
        ends = []
        for s in (h, -h):
            cj = copy.deepcopy(c)
            xj = x.copy()
            xj[j] += s
            cj.reset_parameters(xj)
            ends.append(vq.energy_and_gradient(H, cj, vq.StateVector(n), fused=True)[0].real)
        assert abs((ends[0] - ends[1]) / (2 * h) - gf[j]) < 1e-6 * max(1, np.max(np.abs(gf)))
