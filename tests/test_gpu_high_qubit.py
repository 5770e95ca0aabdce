"""The reference's high-qubit checks (test/check_high_qubit_gate.jl) on the device:
4- and 5-qubit gates and terms on the reference's position permutations.

* A gate from_external(positions, kron(m_1, ..., m_k)) equals the operator of
  QubitsTerm(positions[0] => m_1, ...): applied by the device, against that
  operator's full 2^L matrix (check_fourbody / check_fivebody), and against H * state
  on the device (check_qham_*).
* expectation(t, state) and expectation(bra, t, state) of random 4- and 5-site terms
  against the full matrix (check_expec_*).

The reference's tolerance is 1e-6 (it runs ComplexF32 too); here FP64 within 1e-12.
"""
import numpy as np
import pytest

from brute import full_op, rand_state, rand_unitary
import vqchip as vq

pytestmark = pytest.mark.gpu
TOL = 1e-12


def sv(psi):
    return vq.StateVector(np.ascontiguousarray(psi))


def term_full(L, t):
    return full_op(L, [q - 1 for q in t.positions], t.matrix())


def perms4(i, j, k, l):
    return [(i, j, k, l), (i, j, l, k), (j, i, k, l), (j, i, l, k), (l, k, j, i), (k, i, l, j)]


def perms5_gate(i, j, k, l, n):
    return [(i, j, k, l, n), (i, j, l, n, k), (j, i, n, k, l), (j, n, i, l, k), (l, k, j, n, i), (n, k, i, l, j)]


def perms5_expec(i, j, k, l, n):
    return [(j, i, n, k, l), (j, n, i, l, k), (j, n, i, k, l), (l, k, j, n, i), (n, k, i, l, j), (i, j, k, l, n)]


def gate_and_term(pos, rng):
    ms = [rand_unitary(1, rng) for _ in pos]
    K = ms[0]
    for m in ms[1:]:
        K = np.kron(K, m)
    return vq.from_external(pos, K), vq.QubitsTerm(dict(zip(pos, ms)))


def gate_vs_matrix(L, pos, rng):
    psi = rand_state(L, rng)
    g, h = gate_and_term(pos, rng)
    s = sv(psi)
    vq.apply_(g, s)
    return np.max(np.abs(s.storage() - term_full(L, h) @ psi))


def gate_vs_qham(L, pos, rng):
    psi = rand_state(L, rng)
    g, h = gate_and_term(pos, rng)
    s = sv(psi)
    vq.apply_(g, s)
    return np.max(np.abs(s.storage() - vq.apply_operator(h, sv(psi)).storage()))


@pytest.mark.parametrize("check", [gate_vs_matrix, gate_vs_qham])
def test_fourbody(check):
    rng = np.random.default_rng(4)
    cases = [(4, p) for p in perms4(1, 2, 3, 4)]
    cases += [(10, p) for p in perms4(10, 7, 1, 4)] if check is gate_vs_matrix else [(10, p) for p in perms4(6, 7, 1, 4)]
    assert max(check(L, p, rng) for L, p in cases) < TOL


@pytest.mark.parametrize("check", [gate_vs_matrix, gate_vs_qham])
def test_fivebody(check):
    rng = np.random.default_rng(5)
    cases = [(5, p) for p in perms5_gate(1, 2, 3, 5, 4)]
    cases += [(10, p) for p in perms5_gate(9, 7, 10, 1, 4)] if check is gate_vs_matrix else [(11, p) for p in perms5_gate(9, 11, 10, 1, 4)]
    assert max(check(L, p, rng) for L, p in cases) < TOL


def expec_errors(L, perms, coeff, rng, braket):
    errs = []
    for pos in perms:
        psi, phi = rand_state(L, rng), rand_state(L, rng)
        t = vq.QubitsTerm({p: rand_unitary(1, rng) for p in pos}, coeff=coeff)
        Mpsi = term_full(L, t) @ psi
        if braket:
            errs.append(abs(vq.expectation(sv(phi), t, sv(psi)) - np.vdot(phi, Mpsi)))
        else:
            errs.append(abs(vq.expectation(t, sv(psi)) - np.vdot(psi, Mpsi)))
    return errs


@pytest.mark.parametrize("braket", [False, True])
def test_expectation_four_and_five_sites(braket):
    rng = np.random.default_rng(45 + braket)
    c4, c5 = (0.7, 1.23) if braket else (0.9, 0.77)
    errs = expec_errors(4, perms4(1, 2, 3, 4), c4, rng, braket)
    errs += expec_errors(10 if braket else 9, perms4(7, 10, 1, 5) if braket else perms4(7, 8, 1, 5), c4, rng, braket)
    errs += expec_errors(5, perms5_expec(1, 2, 3, 5, 4), c5, rng, braket)
    errs += expec_errors(10, perms5_expec(7, 10, 2, 5, 8), c5, rng, braket)
    assert max(errs) < TOL
