"""The reference's Hamiltonian checks (test/check_qubits_ham.jl) and density-matrix
expectation checks (test/check_dm.jl) on the device.

* H * state for a one-term operator equals applying the kron of its factors as a
  gate (check_qham_*), over the reference's position loops and sizes.
* H * state and expectation(H, state) against the operator's matrix (check_term_*,
  check_ham_*, including the 5- to 8-site X strings that take the Pauli route).
* expectation(H, state) and expectation(bra, H, state) for random one-, two- and
  three-site terms on every position triple (check_expec_*).
* expectation(h, psi) == expectation(h, DensityMatrix(psi)) (check_dm.jl).

The dense reference is the term's matrix applied by the oracle's sweep (any L), or
the full 2^L operator for L <= 8. The reference's tolerance is 1e-6; here 1e-12.
"""
import numpy as np
import pytest

import oracle
from brute import full_op, rand_state, rand_unitary
import vqchip as vq

pytestmark = pytest.mark.gpu
TOL = 1e-12


def sv(psi):
    return vq.StateVector(np.ascontiguousarray(psi))


def term_apply(t, psi):
    """coeff * (prod of t's factors) psi, by the oracle's dense sweep."""
    pos0 = [q - 1 for q in t.positions]
    return oracle.apply_dense(psi.copy(), len(pos0), pos0, t.matrix())


def op_full(L, H):
    H = H if isinstance(H, vq.QubitsOperator) else vq.QubitsOperator([H])
    out = np.zeros((1 << L, 1 << L), dtype=np.complex128)
    for t in H.terms:
        out += full_op(L, [q - 1 for q in t.positions], t.matrix())
    return out


def qham_single(L, positions, rng):
    """check_qham_*_single: QuantumGate(positions, kron(m_k, ..., m_1)) vs the term
    positions[0] => m_1, positions[1] => m_2, ... applied as H * state."""
    psi = rand_state(L, rng)
    ms = [rand_unitary(1, rng) for _ in positions]
    K = ms[-1]
    for m in reversed(ms[:-1]):
        K = np.kron(K, m)
    g = vq.QuantumGate(positions, K)
    h = vq.QubitsOperator([vq.QubitsTerm(dict(zip(positions, ms)))])
    out = vq.apply_operator(h, sv(psi)).storage()
    s = sv(psi)
    vq.apply_(g, s)
    return np.max(np.abs(out - s.storage()))


def test_qham_onebody():
    rng = np.random.default_rng(1)
    for L in (16, 15):
        assert max(qham_single(L, (c,), rng) for c in range(1, L + 1)) < TOL


@pytest.mark.parametrize("L", [5, 15])
def test_qham_twobody(L):
    rng = np.random.default_rng(L)
    errs = [qham_single(L, (c, t), rng) for c in range(1, L + 1, 2) for t in range(2, L + 1, 2) if c != t]
    assert max(errs) < TOL


@pytest.mark.parametrize("L", [16, 15])
def test_qham_threebody(L):
    rng = np.random.default_rng(100 + L)
    errs = [qham_single(L, (a, b, c), rng) for a in range(1, L + 1, 2) for b in range(2, L + 1, 3)
            for c in range(3, L + 1, 5) if len({a, b, c}) == 3]
    assert max(errs) < TOL


def test_qham_small_states():
    rng = np.random.default_rng(2)
    for L, pos in [(2, (2,)), (4, (2,)), (2, (2, 1)), (3, (1, 3)), (3, (3, 1, 2)), (4, (3, 4, 1))]:
        assert qham_single(L, pos, rng) < TOL, (L, pos)


def ladder_term(L):
    return vq.QubitsTerm({1: "+", L - 1: "-"}, coeff=0.77)


@pytest.mark.parametrize("L", [16, 15])
def test_term_matrix_and_expectation(L):
    rng = np.random.default_rng(L)
    psi = rand_state(L, rng)
    t = ladder_term(L)
    ref = term_apply(t, psi)
    assert np.max(np.abs(vq.apply_operator(t, sv(psi)).storage() - ref)) < TOL
    assert abs(vq.expectation(t, sv(psi)) - np.vdot(psi, ref)) < TOL


@pytest.mark.parametrize("L", [6, 3])
def test_ham_matrix_and_expectation(L):
    rng = np.random.default_rng(30 + L)
    psi = rand_state(L, rng)
    m = (vq.QubitsTerm({1: "+", L - 1: "-"}, coeff=0.77) + vq.QubitsTerm({1: "-", L - 1: "+"}, coeff=0.77)
         + vq.QubitsTerm({1: "X", L: "Y"}, coeff=0.36))
    M = op_full(L, m)
    assert np.max(np.abs(vq.apply_operator(m, sv(psi)).storage() - M @ psi)) < TOL
    e = (vq.QubitsTerm({1: "+", 2: "-"}, coeff=0.77) + vq.QubitsTerm({1: "-", 2: "+"}, coeff=0.77)
         + vq.QubitsTerm({1: "X", L - 1: "Y"}, coeff=0.36))
    assert abs(vq.expectation(e, sv(psi)) - np.vdot(psi, op_full(L, e) @ psi)) < TOL


@pytest.mark.parametrize("L,n", [(6, 5), (6, 6), (8, 8)])
def test_ham_expectation_long_strings(L, n):
    # n >= 6 sites take the Pauli-string route (longer than the dense terms)
    rng = np.random.default_rng(L * 10 + n)
    psi = rand_state(L, rng)
    m = (vq.QubitsTerm({i: "X" for i in range(1, n + 1)}, coeff=0.77) + vq.QubitsTerm({1: "X"}, coeff=1.2)
         + vq.QubitsTerm({1: "X", 2: "Z"}, coeff=0.3))
    assert abs(vq.expectation(m, sv(psi)) - np.vdot(psi, op_full(L, m) @ psi)) < TOL


def expec_cases(L, k):
    if k == 1:
        return [(c,) for c in range(1, L + 1)]
    if k == 2:
        return [(c, t) for c in range(1, L + 1, 2) for t in range(2, L + 1, 2) if c != t]
    return [(a, b, c) for a in range(1, L + 1) for b in range(1, L + 1) for c in range(1, L + 1) if len({a, b, c}) == 3]


@pytest.mark.parametrize("L", [5, 15])
@pytest.mark.parametrize("k", [1, 2, 3])
def test_expectation_generic_terms(L, k):
    rng = np.random.default_rng(L * 7 + k)
    coeff = {1: 1.0, 2: 0.73, 3: 0.9}[k]
    cases = expec_cases(L, k)
    errs = []
    # a fresh pair of states every 64 cases (the reference draws one per case)
    for i, pos in enumerate(cases):
        if i % 64 == 0:
            psi, phi = rand_state(L, rng), rand_state(L, rng)
            s_psi, s_phi = sv(psi), sv(phi)
        t = vq.QubitsTerm({p: rand_unitary(1, rng) for p in pos}, coeff=coeff)
        ref = term_apply(t, psi)
        errs.append(abs(vq.expectation(t, s_psi) - np.vdot(psi, ref)))
        errs.append(abs(vq.expectation(s_phi, t, s_psi) - np.vdot(phi, ref)))
    assert max(errs) < TOL


@pytest.mark.parametrize("L,term", [(7, {1: "X"}), (7, {1: "X", 2: "+"}), (7, {1: "X", 2: "+", 4: "-"}),
                                    (8, {1: "X", 2: "+", 4: "-", 8: "Z"})])
def test_density_matrix_expectation_equals_state(L, term):
    psi = rand_state(L, np.random.default_rng(L + len(term)))
    h = vq.QubitsTerm(term)
    a1 = vq.expectation(h, sv(psi))
    a2 = vq.expectation(h, vq.DensityMatrix(sv(psi)))
    assert abs(a1 - a2) < TOL
    assert abs(a1 - np.vdot(psi, term_apply(h, psi))) < TOL
