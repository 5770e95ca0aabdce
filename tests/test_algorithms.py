"""The QFT builder and the phase-estimation known answers (test/algs.jl) on the
CPU: the QFT circuit against the closed-form discrete Fourier transform, and
phase estimation run through the oracle, whose measured bits must return the
encoded phase exactly."""
import numpy as np
import pytest

import oracle
from algorithms import phase_estimation_circuit, to_digits
from brute import full_op
from vqchip import QFT, CONTROL, R, XGate, CNOTGate
from vqchip.circuit import Program


def circuit_matrix(c, n):
    M = np.eye(1 << n, dtype=complex)
    for g in c:
        M = full_op(n, [p - 1 for p in g.ordered_positions()], g.ordered_mat()) @ M
    return M


def _rev(x, L):
    return int(format(x, f"0{L}b")[::-1], 2)


@pytest.mark.parametrize("L", [1, 2, 3, 5])
def test_qft_is_the_fourier_transform(L):
    N = 1 << L
    rev = np.array([_rev(x, L) for x in range(N)])  # qubit 1 as the most significant bit
    F = np.exp(2j * np.pi * np.outer(rev, rev) / N) / np.sqrt(N)
    assert np.max(np.abs(circuit_matrix(QFT(L), L) - F)) < 1e-12
    assert len(QFT(L)) == L * (L + 1) // 2
    assert np.max(np.abs(circuit_matrix(QFT(L).adjoint(), L) - F.conj().T)) < 1e-12


def test_control_is_cnot_on_x_and_r_is_a_phase_step():
    assert np.array_equal(CONTROL(XGate(1).mat()), CNOTGate(1, 2).mat())
    assert np.allclose(R(1), np.diag([1, -1])) and np.allclose(R(2), np.diag([1, 1j]))


def _run_oracle(bits, dm, auto_reset, rng):
    L = len(bits)
    n = L + 1
    c = phase_estimation_circuit(bits)
    prog = Program(c, n=n, dm=dm)
    v = np.zeros(1 << (2 * n if dm else n), complex)
    v[0] = 1
    oracle.apply_circuit(v, n, int(dm), prog.op_array(), len(prog.ops), prog.pool())
    res = []
    for q in range(n):
        i, p = oracle.measure_inplace(v, n, q, float(rng.random()), auto_reset, is_dm=dm)
        assert abs(p - 1) < 1e-10
        res.append(i)
    return res


@pytest.mark.parametrize("dm", [False, True])
@pytest.mark.parametrize("auto_reset", [False, True])
def test_phase_estimation_on_the_oracle(dm, auto_reset):
    rng = np.random.default_rng(11 + 2 * dm + auto_reset)
    for L in range(2, 6 if dm else 9):
        bits = [int(b) for b in rng.integers(0, 2, L)]
        res = _run_oracle(bits, dm, auto_reset, rng)
        assert to_digits(res) == to_digits(bits) and res[:L] == bits, (L, bits, res)
