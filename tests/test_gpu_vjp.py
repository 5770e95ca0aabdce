"""vqc_circuit_vjp (the pullback of y = circuit * x for any cotangent,
circuitdiff.jl:2-8, 23-78) against the oracle's, which test_vjp.py pins by
brute force and finite differences. Tolerances (FP64): states within 1e-12
absolute per unit of |delta|, gradients within 1e-10 relative."""
import numpy as np
import pytest

import oracle
from brute import rand_density, rand_state, vec
from test_vjp import dm_circuit, sv_circuit
import vqchip as vq
from vqchip.circuit import Program
from vqchip.workloads import tfim

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [4, 7, 12, 16])
def test_state_vector_vjp(n):
    rng = np.random.default_rng(300 + n)
    c = sv_circuit(n, rng)
    x = rand_state(n, rng)
    d = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    g, dx, y = vq.circuit_vjp(c, vq.StateVector(x), vq.StateVector(d), return_state=True)
    prog = Program(c, n=n)
    gr, dxr, yr = oracle.circuit_vjp(x, n, 0, prog.op_array(), len(prog.ops), prog.pool(), prog.nparams, d)
    scale = np.max(np.abs(d))
    assert np.max(np.abs(y.storage() - yr)) < 1e-12
    assert np.max(np.abs(dx.storage() - dxr)) < 1e-12 * scale
    assert np.max(np.abs(g - gr)) < 1e-10 * max(1.0, np.max(np.abs(gr)))


@pytest.mark.parametrize("n", [3, 5])
def test_density_matrix_vjp_with_channels(n):
    rng = np.random.default_rng(400 + n)
    c, _ = dm_circuit(n)
    x = vec(rand_density(n, rng))
    d = rng.standard_normal(1 << (2 * n)) + 1j * rng.standard_normal(1 << (2 * n))
    g, dx, y = vq.circuit_vjp(c, vq.DensityMatrix(x), vq.DensityMatrix(d), return_state=True)
    prog = Program(c, n=n, dm=True)
    gr, dxr, yr = oracle.circuit_vjp(x, n, 1, prog.op_array(), len(prog.ops), prog.pool(), prog.nparams, d)
    assert np.max(np.abs(vec(y.storage()) - yr)) < 1e-12
    assert np.max(np.abs(vec(dx.storage()) - dxr)) < 1e-12 * np.max(np.abs(d))
    assert np.max(np.abs(g - gr)) < 1e-10 * max(1.0, np.max(np.abs(gr)))


def test_vjp_with_the_energy_cotangent_matches_energy_and_gradient():
    n = 14
    rng = np.random.default_rng(500)
    c = sv_circuit(n, rng)
    H = tfim(n, 0.8)
    s = vq.StateVector(rand_state(n, rng))
    E, g, y = vq.energy_and_gradient(H, c, s, return_state=True)
    Hy = vq.apply_operator(H, y)
    Hy.scale_(2.0)  # (H + H^dagger) y for a Hermitian H
    gv, dx = vq.circuit_vjp(c, s, Hy)
    assert np.max(np.abs(gv - g)) < 1e-10 * max(1.0, np.max(np.abs(g)))


def test_vjp_rejects_mismatched_shapes():
    c = sv_circuit(5, np.random.default_rng(1))
    with pytest.raises(ValueError):
        vq.circuit_vjp(c, vq.StateVector(5), vq.StateVector(4))


def _dense(op, n):
    from brute import full_op
    M = np.zeros((1 << n, 1 << n), dtype=complex)
    for t in op.terms:
        M += full_op(n, [p - 1 for p in t.positions], t.matrix())
    return M


def _ops(L):
    return [vq.QubitsOperator([vq.QubitsTerm({1: "+", L - 1: "Y"}, coeff=0.36)]),
            vq.QubitsOperator([vq.QubitsTerm({1: "+", L - 1: "-"}, coeff=0.77), vq.QubitsTerm({1: "-", L - 1: "+"}, coeff=0.77),
                               vq.QubitsTerm({1: "X", L: "Y"}, coeff=0.36)])]


@pytest.mark.parametrize("L", [3, 4, 6, 7])
def test_expectation_vjp_matches_finite_differences(L):
    # loss(x) = abs(expectation(m, x)) as in test/check_ham_expec_diff.jl
    rng = np.random.default_rng(600 + L)
    ops = _ops(L)
    if L == 7:  # a 7-site Pauli string takes the Pauli route (more than 5 sites)
        ops.append(vq.QubitsOperator([vq.QubitsTerm({i: "X" for i in range(1, 8)}, coeff=0.77)]))
    for op in ops:
        x = rand_state(L, rng)
        s = vq.StateVector(x)
        E = vq.expectation(op, s)
        dx = vq.expectation_vjp(op, s, E / abs(E)).storage()
        M = _dense(op, L)
        loss = lambda v: abs(np.vdot(v, M @ v))
        fd = np.zeros(1 << L, dtype=complex)
        for k in range(1 << L):
            e = np.zeros(1 << L)
            e[k] = 1e-6
            fd[k] = complex((loss(x + e) - loss(x - e)) / 2e-6, (loss(x + 1j * e) - loss(x - 1j * e)) / 2e-6)
        assert abs(E - np.vdot(x, M @ x)) < 1e-12
        assert np.max(np.abs(dx - fd)) < 1e-7


def test_expectation_vjp_density_matrix():
    n = 3
    op = _ops(n)[1]
    rho = vq.DensityMatrix(vec(rand_density(n, np.random.default_rng(7))))
    z = 0.3 - 0.8j
    d = vq.expectation_vjp(op, rho, z).storage()
    assert np.max(np.abs(d - z * _dense(op, n).conj().T)) < 1e-13


def test_loss_chain_through_expectation_and_circuit():
    # Zygote's chain for real(expectation(H, circuit * x)): the expectation pullback
    # with z = 1 feeds the circuit pullback; equals the fused energy gradient
    n = 9
    rng = np.random.default_rng(700)
    c = sv_circuit(n, rng)
    H = tfim(n, 0.6)
    s = vq.StateVector(rand_state(n, rng))
    E, g, y = vq.energy_and_gradient(H, c, s, return_state=True)
    dy = vq.expectation_vjp(H, y, 1.0)
    gv, dx = vq.circuit_vjp(c, s, dy)
    assert np.max(np.abs(gv - g)) < 1e-10 * max(1.0, np.max(np.abs(g)))


@pytest.mark.parametrize("loss", ["real", "imag", "abs", "abs2", "distance"])
def test_circuit_gradients_of_the_reference_losses(loss):
    # test/circuitgrad.jl and crxgategrad.jl: loss(circuit * |0..0>) for a
    # variational circuit with FSIM and CRx layers; the loss's cotangent is formed
    # from the device output, the gradient comes from the device pullback and is
    # checked against 5-point differences of the brute-force loss
    from test_vjp import FD_H, FD_W, circuit_matrix
    from vqchip.utility import variational_circuit_1d
    L, depth = 5, 2
    rng = np.random.default_rng(800)
    c = variational_circuit_1d(L, depth, rng=rng)
    c.insert(4, vq.FSIMGate(2, 4, 0.4, 0.9))
    c.append(vq.CRxGate(1, L, 1.1))
    t = rand_state(L, rng)
    x0 = np.zeros(1 << L, complex)
    x0[0] = 1

    def value(y):
        z = np.vdot(t, y)
        if loss == "distance":
            return np.sqrt(abs(np.vdot(t, t).real + np.vdot(y, y).real - 2 * z.real))
        return {"real": z.real, "imag": z.imag, "abs": abs(z), "abs2": abs(z) ** 2}[loss]

    x = vq.StateVector(x0)
    ydev = vq.apply_(c, vq.StateVector(x0))
    yv = ydev.storage()
    z = np.vdot(t, yv)
    delta = {"real": t, "imag": 1j * t, "abs": z / abs(z) * t, "abs2": 2 * z * t,
             "distance": (yv - t) / value(yv)}[loss]
    g, _ = vq.circuit_vjp(c, x, vq.StateVector(delta))
    theta = np.array(c.active_parameters())
    assert len(g) == len(theta)
    for k in range(len(theta)):
        fd = 0.0
        for w, s in zip(FD_W, (-2, -1, 0, 1, 2)):
            if w:
                th = theta.copy()
                th[k] += s * FD_H
                c.reset_parameters(th)
                fd += w * value(circuit_matrix(c, L) @ x0) / FD_H
        c.reset_parameters(theta)
        assert abs(fd - g[k]) < 1e-8, (loss, k, fd, g[k])


def test_qubit_encoding_gradient():
    # test/check_ham_expec_diff.jl check_mps_ham_expec_grad:
    # loss(v) = real(expectation(Z_1 Z_3, qubit_encoding(v)))
    from vqchip.state import qubit_vector
    L = 5
    v = np.random.default_rng(900).standard_normal(L)
    obs = vq.QubitsOperator([vq.QubitsTerm({1: "Z", 3: "Z"})])
    s = vq.qubit_encoding(v)
    g = vq.qubit_encoding_vjp(v, vq.expectation_vjp(obs, s, 1.0))
    M = _dense(obs, L)
    loss = lambda w: np.vdot(qubit_vector(w), M @ qubit_vector(w)).real
    for k in range(L):
        e = np.zeros(L)
        e[k] = 1e-6
        assert abs((loss(v + e) - loss(v - e)) / 2e-6 - g[k]) < 1e-8
