"""Plan reuse: the fused entry points called again with the same program structure
and new matrix values, checked against the oracle on every call.

A program whose op and term tables are unchanged reuses its cached plan with the
new values (vqc_plan_info().reused == 1); values that break a structure class the
plan relied on (a zero angle makes Rx diagonal, or a run of rotations on one qubit
multiplies out to a diagonal matrix) plan again (reused == 0). Tolerances as in
the other parity tests: states within 1e-12, energies and gradients within 1e-10.
"""
import ctypes

import numpy as np
import pytest

import oracle
from brute import rand_density, rand_state, vec
import vqchip as vq
from vqchip import _lib
from vqchip.circuit import Program
from vqchip.operators import pack_terms
from vqchip.workloads import config3, hea_circuit, heisenberg, tfim

pytestmark = pytest.mark.gpu


def reused():
    info = _lib.PlanInfo()
    _lib.call("vqc_plan_info", ctypes.byref(info))
    return info.reused


def check_sv(c, H, n, psi0):
    s = vq.StateVector(np.ascontiguousarray(psi0))
    E, g, out = vq.energy_and_gradient(H, c, s, fused=True, return_state=True)
    r = reused()
    prog = Program(c, n=n)
    terms, nt, hpool = pack_terms(H)
    Er, gr, psi_r = oracle.energy_grad(psi0, n, 0, prog.op_array(), len(prog.ops), prog.pool(), terms, nt, hpool,
                                       prog.nparams)
    assert np.max(np.abs(out.storage() - psi_r)) < 1e-12
    assert abs(E - Er) < 1e-10 * max(1, abs(Er))
    assert np.max(np.abs(g - gr)) < 1e-10 * max(1, np.max(np.abs(gr)))
    return r


def rot_circuit(n, angles):
    """Rz, Rx, Ry on every qubit (a run of three on one qubit) and a CNOT ladder."""
    c = vq.QCircuit()
    k = 0
    for q in range(1, n + 1):
        c.append(vq.RzGate(q, angles[k]))
        c.append(vq.RxGate(q, angles[k + 1]))
        c.append(vq.RyGate(q, angles[k + 2]))
        k += 3
    for q in range(1, n):
        c.append(vq.CNOTGate(q, q + 1))
    for q in range(1, n + 1):
        c.append(vq.RxGate(q, angles[k]))
        k += 1
    return c


@pytest.mark.parametrize("n", [10, 15])  # whole state in one workgroup / tile phases
def test_new_angles_reuse_the_plan(n):
    rng = np.random.default_rng(300 + n)
    H = heisenberg(n, rng.standard_normal(n - 1))
    psi0 = rand_state(n, rng)
    flags = []
    for _ in range(3):
        c = hea_circuit(n, 2, rng.uniform(0, 2 * np.pi, 6 * n))
        flags.append(check_sv(c, H, n, psi0))
    assert flags[1:] == [1, 1]


@pytest.mark.parametrize("n", [8, 14])
def test_structure_change_plans_again(n):
    rng = np.random.default_rng(400 + n)
    H = tfim(n, 0.7)
    psi0 = rand_state(n, rng)
    a = rng.uniform(0.2, 3.0, 4 * n)
    zero_x = a.copy()
    zero_x[1::3][:n] = 0.0  # every Rx of the first layer is the identity (diagonal)
    zero_run = a.copy()
    zero_run[1] = zero_run[2] = 0.0  # Rz Rx(0) Ry(0) on qubit 1 multiplies out to a diagonal matrix
    check_sv(rot_circuit(n, a), H, n, psi0)
    assert check_sv(rot_circuit(n, a * 0.5), H, n, psi0) == 1
    assert check_sv(rot_circuit(n, zero_x), H, n, psi0) == 0
    assert check_sv(rot_circuit(n, a), H, n, psi0) == 0
    assert check_sv(rot_circuit(n, zero_run), H, n, psi0) == 0
    assert check_sv(rot_circuit(n, a * 0.3), H, n, psi0) == 0
    assert check_sv(rot_circuit(n, a * 0.7), H, n, psi0) == 1


def test_programs_interleaved():
    # alternating programs: each call must run its own micro-ops, not the last upload
    rng = np.random.default_rng(7)
    cases = []
    for n in (9, 13, 9):
        H = heisenberg(n, rng.standard_normal(n - 1)) if n == 13 else tfim(n, 0.4)
        cases.append((n, H, rand_state(n, rng), int(rng.integers(1, 3))))
    for rep in range(3):
        for n, H, psi0, depth in cases:
            c = hea_circuit(n, depth, rng.uniform(0, 6.3, 3 * n * depth))
            r = check_sv(c, H, n, psi0)
            if rep:
                assert r == 1
        # a forward-only program in between
        n = 11
        c = hea_circuit(n, 1, rng.uniform(0, 6.3, 3 * n))
        v = rand_state(n, rng)
        s = vq.StateVector(np.ascontiguousarray(v))
        vq.apply_(c, s)
        prog = Program(c, n=n)
        ref = oracle.apply_circuit(v.copy(), n, 0, prog.op_array(), len(prog.ops), prog.pool())
        assert np.max(np.abs(s.storage() - ref)) < 1e-12
        if rep:
            assert reused() == 1


def test_density_matrix_and_vjp_reuse():
    n = 5
    rho0 = np.zeros(1 << (2 * n), dtype=complex)
    rho0[0] = 1
    for k, seed in enumerate((1, 2, 3)):
        c, H, _ = config3(seed=seed, n=n, depth=2, p=0.05, gamma=0.03)
        d = vq.DensityMatrix(rho0)
        E, g, out = vq.energy_and_gradient(H, c, d, fused=True, return_state=True)
        if k:
            assert reused() == 1
        prog = Program(c, n=n, dm=True)
        terms, nt, hpool = pack_terms(H)
        Er, gr, rho_r = oracle.energy_grad(rho0, n, 1, prog.op_array(), len(prog.ops), prog.pool(), terms, nt, hpool,
                                           prog.nparams)
        assert np.max(np.abs(vec(out.storage()) - rho_r)) < 1e-12
        assert abs(E - Er) < 1e-10
        assert np.max(np.abs(g - gr)) < 1e-10 * max(1, np.max(np.abs(gr)))
    # the pullback of circuit * x with new angles and a new cotangent
    rng = np.random.default_rng(11)
    n = 12
    x = rand_state(n, rng)
    for k in range(3):
        c = hea_circuit(n, 2, rng.uniform(0, 6.3, 6 * n))
        dy = rand_state(n, rng)
        g, dx = vq.circuit_vjp(c, vq.StateVector(np.ascontiguousarray(x)), vq.StateVector(np.ascontiguousarray(dy)))
        if k:
            assert reused() == 1
        prog = Program(c, n=n)
        gr, dxr, _ = oracle.circuit_vjp(x, n, 0, prog.op_array(), len(prog.ops), prog.pool(), prog.nparams, dy)
        assert np.max(np.abs(g - gr)) < 1e-10 * max(1, np.max(np.abs(gr)))
        assert np.max(np.abs(dx.storage() - dxr)) < 1e-12


def test_channel_form_change_plans_again():
    # the same op table with a depolarizing channel (parity-block superoperator) and
    # then a mixed-unitary one (dense): the parity claim fails and the plan is redone
    from test_gpu_parity import mixed_unitary_channel
    n = 4
    rng = np.random.default_rng(21)
    rho0 = vec(rand_density(n, rng))
    H = tfim(n, 0.6)

    def circ(ch):
        c = vq.QCircuit()
        for q in range(1, n + 1):
            c.append(vq.RyGate(q, 0.3 * q))
        c.append(vq.CNOTGate(1, 2))
        c.append(ch)
        c.append(vq.RzGate(2, 0.7))
        return c

    expected = [None, 0, 1, 0, 1]
    channels = [vq.Depolarizing(2, 0.1), mixed_unitary_channel((2,), rng), mixed_unitary_channel((2,), rng),
                vq.AmplitudeDamping(2, 0.2), vq.Depolarizing(2, 0.05)]
    for ch, want in zip(channels, expected):
        c = circ(ch)
        E, g, out = vq.energy_and_gradient(H, c, vq.DensityMatrix(rho0), fused=True, return_state=True)
        r = reused()
        prog = Program(c, n=n, dm=True)
        terms, nt, hpool = pack_terms(H)
        Er, gr, rho_r = oracle.energy_grad(rho0, n, 1, prog.op_array(), len(prog.ops), prog.pool(), terms, nt, hpool,
                                           prog.nparams)
        assert np.max(np.abs(vec(out.storage()) - rho_r)) < 1e-12
        assert abs(E - Er) < 1e-10 and np.max(np.abs(g - gr)) < 1e-10
        if want is not None:
            assert r == want, (ch, r)
