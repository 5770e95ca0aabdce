"""Terms of more than 5 sites on the device: the per-term route of the
reference (_apply_qterm_util!, apply_qterms.jl:47-59; dot(psi, m * psi),
expecs.jl:17,40,50,72; tr(m * rho), expecs_dm.jl:25; the expectation pullback
of long terms, expecdiff.jl:7-16,37-56) against the oracle, which
test_oracle.py pins to the dense operator. Tolerances (FP64): amplitudes 1e-12
absolute per unit of coefficient, expectations and gradients 1e-10 relative."""
import numpy as np
import pytest

import oracle
from brute import rand_density, rand_state, vec
from test_oracle import _rand_qterms
from test_vjp import dm_circuit, sv_circuit
import vqchip as vq
from vqchip import _lib
from vqchip.operators import pack_qterms
from vqchip.workloads import tfim

pytestmark = pytest.mark.gpu

KINDS = [("+", "-"), ("+", "-", "X", "Y", "Z", "I", "P0", "P1"), ("dense", "diag", "anti", "Z"), ("dense",),
         ("diag", "anti")]


def _terms(qt):
    """oracle (coef, sites, mats) tuples -> QubitsTerms (1-based positions)."""
    return [vq.QubitsTerm([(s + 1, m) for s, m in zip(sites, mats)], coeff=c) for c, sites, mats in qt]


def _rel(a, b):
    return abs(a - b) / max(1.0, abs(b))


@pytest.mark.parametrize("kinds", KINDS)
@pytest.mark.parametrize("n", [6, 9, 13])
def test_apply_and_expec_qterms(kinds, n):
    rng = np.random.default_rng(40 * n + len(kinds))
    qt = _rand_qterms(n, rng, kinds, nterms=4)
    terms = _terms(qt)
    v, vc = rand_state(n, rng), rand_state(n, rng)
    ref = oracle.apply_qterms(v, qt)
    s, b = vq.StateVector(v), vq.StateVector(vc)
    arr, nt, pos, pool = pack_qterms(terms)
    out = vq.StateVector(n)
    _lib.call("vqc_sv_apply_qterms", s.handle, out.handle, nt, arr, pos, _lib.dptr(pool), 0)
    assert np.max(np.abs(out.storage() - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
    # accumulate onto what out holds
    _lib.call("vqc_sv_apply_qterms", s.handle, out.handle, nt, arr, pos, _lib.dptr(pool), 1)
    assert np.max(np.abs(out.storage() - 2 * ref)) < 2e-12 * max(1.0, np.max(np.abs(ref)))
    per_r, tot_r = oracle.expec_qterms(v, qt)
    per = np.zeros(2 * nt)
    s2 = np.zeros(2)
    _lib.call("vqc_sv_expec_qterms", s.handle, None, nt, arr, pos, _lib.dptr(pool), _lib.dptr(per), _lib.dptr(s2))
    assert np.max(np.abs(per[0::2] + 1j * per[1::2] - per_r)) < 1e-10
    assert _rel(complex(*s2), tot_r) < 1e-10
    _, tot_b = oracle.expec_qterms(v, qt, vc=vc)
    _lib.call("vqc_sv_expec_qterms", s.handle, b.handle, nt, arr, pos, _lib.dptr(pool), None, _lib.dptr(s2))
    assert _rel(complex(*s2), tot_b) < 1e-10


@pytest.mark.parametrize("kinds", KINDS)
def test_dm_expec_qterms(kinds):
    n = 6
    rng = np.random.default_rng(900 + len(kinds))
    qt = _rand_qterms(n, rng, kinds, nterms=3, wmin=4)
    rho = rand_density(n, rng)
    per_r, tot_r = oracle.dm_expec_qterms(rho, n, qt)
    terms = _terms(qt)
    arr, nt, pos, pool = pack_qterms(terms)
    r = vq.DensityMatrix(rho)
    per = np.zeros(2 * nt)
    s2 = np.zeros(2)
    _lib.call("vqc_dm_expec_qterms", r.handle, nt, arr, pos, _lib.dptr(pool), _lib.dptr(per), _lib.dptr(s2))
    assert np.max(np.abs(per[0::2] + 1j * per[1::2] - per_r)) < 1e-10
    assert _rel(complex(*s2), tot_r) < 1e-10


def _mixed_operator(n):
    """The reference's long-term checks (test/check_qubits_ham.jl:73-101): sigma+ /
    sigma- terms across the chain, a long X string, and short keys beside them."""
    return (vq.QubitsTerm({1: "+", n - 1: "-"}, coeff=0.77) + vq.QubitsTerm({1: "-", n - 1: "+"}, coeff=0.77)
            + vq.QubitsTerm({1: "X", n: "Y"}, coeff=0.36)
            + vq.QubitsTerm({i: "X" for i in range(1, n + 1)}, coeff=0.77)
            + vq.QubitsTerm({i: ("+" if i % 2 else "-") for i in range(1, n + 1)}, coeff=0.41 - 0.2j)
            + vq.QubitsTerm({1: "X"}, coeff=1.2) + vq.QubitsTerm({1: "X", 2: "Z"}, coeff=0.3))


def _oracle_qt(op):
    return [(t.coeff, [p - 1 for p in t.positions], t.oplist) for t in op.terms]


@pytest.mark.parametrize("n", [7, 10])
def test_operator_with_long_terms(n):
    rng = np.random.default_rng(70 + n)
    H = _mixed_operator(n)
    v = rand_state(n, rng)
    s = vq.StateVector(v)
    ref = oracle.apply_qterms(v, _oracle_qt(H))
    assert np.max(np.abs(vq.apply_operator(H, s).storage() - ref)) < 1e-12
    assert np.max(np.abs((H * s).storage() - ref)) < 1e-12
    _, tot = oracle.expec_qterms(v, _oracle_qt(H))
    assert _rel(vq.expectation(H, s), tot) < 1e-10
    vc = rand_state(n, rng)
    _, totb = oracle.expec_qterms(v, _oracle_qt(H), vc=vc)
    assert _rel(vq.expectation(vq.StateVector(vc), H, s), totb) < 1e-10
    # one long term alone (expectation(m::QubitsTerm, state), expecs.jl:12-19)
    t = H.terms[4]
    _, t1 = oracle.expec_qterms(v, _oracle_qt(vq.QubitsOperator([t])))
    assert _rel(vq.expectation(t, s), t1) < 1e-10
    # the same operator on DensityMatrix(psi): tr(H rho) = <psi|H|psi> (test/check_dm.jl method)
    if n <= 7:
        assert _rel(vq.expectation(H, vq.DensityMatrix(s)), tot) < 1e-10


def _fd_energy(H, c, state0, k, h=1e-4):
    """d Re E / d theta_k by the 5-point central difference (central_fdm(5, 1), test/util.jl:8)."""
    import copy
    theta = np.array(c.active_parameters(), dtype=float)

    def E(dt):
        th = theta.copy()
        th[k] += dt
        cc = copy.deepcopy(c)
        cc.reset_parameters(th)
        return vq.expectation(H, cc * state0).real

    return (-E(2 * h) + 8 * E(h) - 8 * E(-h) + E(-2 * h)) / (12 * h)


@pytest.mark.parametrize("n", [7, 11])
def test_energy_gradient_with_long_terms(n):
    """energy_and_gradient through the expectation pullback of long terms
    (expecdiff.jl:37-56) against the oracle's pullback chain and central
    differences of the device energy."""
    rng = np.random.default_rng(1100 + n)
    c = sv_circuit(n, rng)
    H = _mixed_operator(n)
    x = rand_state(n, rng)
    E, g = vq.energy_and_gradient(H, c, vq.StateVector(x))
    # oracle chain: y = circuit x, delta = (H + H^dagger) y term by term, circuit pullback
    from vqchip.circuit import Program
    prog = Program(c, n=n)
    y = oracle.apply_circuit(x.copy(), n, 0, prog.op_array(), len(prog.ops), prog.pool())
    qt = _oracle_qt(H)
    _, Er = oracle.expec_qterms(y, qt)
    delta = oracle.apply_qterms(y, qt) + oracle.apply_qterms(y, _oracle_qt(H.adjoint()))
    gr, _, _ = oracle.circuit_vjp(x, n, 0, prog.op_array(), len(prog.ops), prog.pool(), prog.nparams, delta)
    assert _rel(E, Er) < 1e-10
    assert np.max(np.abs(g - gr)) < 1e-10 * max(1.0, np.max(np.abs(gr)))
    for k in (0, len(g) // 2, len(g) - 1):
        assert abs(g[k] - _fd_energy(H, c, vq.StateVector(x), k)) < 1e-6


def test_dm_energy_gradient_with_long_terms():
    """A noisy density matrix with 6-site terms (the tr(m * rho) route and its
    pullback z vec(H^dagger), expecdiff.jl:24-25,64-65) against central differences."""
    n = 6
    c, _ = dm_circuit(n)
    H = (vq.QubitsTerm({1: "+", 2: "Z", 3: "-", 4: "X", 5: "I", 6: "Y"}, coeff=0.6) + vq.QubitsTerm({1: "Z", 2: "Z"})
         + vq.QubitsTerm({i: "Y" for i in range(1, n + 1)}, coeff=-0.3)
         + vq.QubitsTerm({i: "X" for i in range(1, n + 1)}))
    rho0 = vq.DensityMatrix(n)
    E, g = vq.energy_and_gradient(H, c, rho0)
    assert abs(E - vq.expectation(H, c * rho0)) < 1e-12
    for k in range(0, len(g), max(1, len(g) // 5)):
        assert abs(g[k] - _fd_energy(H, c, rho0, k)) < 1e-6


def test_expectation_pullback_chain_matches_energy_and_gradient():
    """The Zygote chain of the reference, gradient(c -> real(expectation(H, c * psi0)), c):
    the expectation pullback (conj(z) H + z H^dagger) psi with z = 1 (expecdiff.jl:30-46),
    then the circuit pullback (circuitdiff.jl:2-8), equals the fused energy gradient."""
    n = 12
    rng = np.random.default_rng(77)
    c = sv_circuit(n, rng)
    H = tfim(n, 0.7) + vq.QubitsTerm({1: "X", 3: "Y"}, coeff=0.2 + 0.1j)  # a non-Hermitian key too
    psi0 = vq.StateVector(rand_state(n, rng))
    E, g = vq.energy_and_gradient(H, c, psi0)
    y = c * psi0
    delta = vq.expectation_vjp(H, y, 1.0)
    gv, _ = vq.circuit_vjp(c, psi0, delta)
    assert abs(vq.expectation(H, y) - E) < 1e-12 * max(1.0, abs(E))
    assert np.max(np.abs(gv - g)) < 1e-12 * max(1.0, np.max(np.abs(g)))
