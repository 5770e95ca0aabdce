"""The circuit pullback (@adjoint *(circuit, x) with back_propagate,
circuitdiff.jl:2-8, 23-78) in the oracle, pinned without the reference:

* with the energy's cotangent as the seed it reproduces the energy gradient;
* delta_x = U^dagger delta_y and y = U x against the brute-force circuit matrix
  (a density matrix: the product of superoperators, channels included);
* the parameter gradients against 5-point central differences (the reference's
  central_fdm(5, 1), test/util.jl:8) of the losses of test/stategrad.jl
  (real and imaginary parts of <target|circuit * x>) and of |<target|y>|^2.
"""
import numpy as np
import pytest

import oracle
from brute import full_op, rand_density, rand_state, rand_unitary, vec
import vqchip as vq
from vqchip.circuit import Program
from vqchip.operators import pack_terms
from vqchip.workloads import config3, hea_circuit, tfim

FD_H = 1e-3
FD_W = (1 / 12, -2 / 3, 0, 2 / 3, -1 / 12)  # 5-point central first derivative


def sv_circuit(n, rng):
    c = hea_circuit(n, 2, rng.uniform(0, 6.3, 6 * n))
    c.insert(2, vq.FSIMGate(1, n, 0.3, 1.2))
    c.insert(5, vq.CRxGate(n, 2, 0.8))
    c.insert(7, vq.RyGate(2, 0.4, isparas=False))
    c.append(vq.QuantumGate((1, 3, n), rand_unitary(3, rng)))
    c.append(vq.PHASEGate(n // 2 + 1, 0.9))
    return c


def circuit_matrix(c, nbits, dm=False, n=None):
    M = np.eye(1 << nbits, dtype=complex)
    for g in c:
        pos = [p - 1 for p in g.ordered_positions()]
        if hasattr(g, "kraus"):
            S = g.ordered_supermat()
            F = full_op(nbits, sorted(pos + [p + n for p in pos]), S)
        elif dm:
            U = g.ordered_mat()
            F = full_op(nbits, pos, U) @ full_op(nbits, [p + n for p in pos], U.conj())
        else:
            F = full_op(nbits, pos, g.ordered_mat())
        M = F @ M
    return M


def vjp(c, x, n, dm, delta):
    prog = Program(c, n=n, dm=dm)
    return oracle.circuit_vjp(x, n, int(dm), prog.op_array(), len(prog.ops), prog.pool(), prog.nparams, delta)


def test_sv_vjp_seeded_with_the_energy_cotangent_is_the_energy_gradient():
    rng = np.random.default_rng(1)
    n = 6
    c, H = sv_circuit(n, rng), tfim(n, 0.7)
    x = rand_state(n, rng)
    prog = Program(c, n=n)
    terms, nt, hpool = pack_terms(H)
    E, g, y = oracle.energy_grad(x, n, 0, prog.op_array(), len(prog.ops), prog.pool(), terms, nt, hpool, prog.nparams)
    # (H + H^dagger) y from the oracle's own term sweeps
    Hy = np.zeros_like(y)
    for j in range(nt):
        M = hpool[terms[j].mat: terms[j].mat + (1 << (2 * terms[j].nq))].reshape(1 << terms[j].nq, -1, order="F")
        pos = [terms[j].pos[k] for k in range(terms[j].nq)]
        Hy += full_op(n, pos, M + M.conj().T) @ y
    gv, dx, yv = vjp(c, x, n, False, Hy)
    assert np.max(np.abs(yv - y)) < 1e-13
    assert np.max(np.abs(gv - g)) < 1e-12


def test_sv_vjp_closed_form():
    rng = np.random.default_rng(2)
    n = 5
    c = sv_circuit(n, rng)
    U = circuit_matrix(c, n)
    x, d = rand_state(n, rng), rand_state(n, rng) * 3.0
    g, dx, y = vjp(c, x, n, False, d)
    assert np.max(np.abs(y - U @ x)) < 1e-13
    assert np.max(np.abs(dx - U.conj().T @ d)) < 1e-13


@pytest.mark.parametrize("loss", ["real", "imag", "abs2"])
def test_sv_vjp_parameter_gradient_matches_finite_differences(loss):
    rng = np.random.default_rng(3)
    n = 4
    c = sv_circuit(n, rng)
    x, t = rand_state(n, rng), rand_state(n, rng)

    def L(circ):
        y = circuit_matrix(circ, n) @ x
        z = np.vdot(t, y)
        return {"real": z.real, "imag": z.imag, "abs2": abs(z) ** 2}[loss]

    y = circuit_matrix(c, n) @ x
    z = np.vdot(t, y)
    delta = {"real": t, "imag": 1j * t, "abs2": 2 * z * t}[loss]
    g, dx, _ = vjp(c, x, n, False, delta)
    theta = np.array(c.active_parameters())
    assert len(g) == len(theta) > 10
    for k in range(len(theta)):
        fd = 0.0
        for w, s in zip(FD_W, (-2, -1, 0, 1, 2)):
            if w:
                th = theta.copy()
                th[k] += s * FD_H
                c.reset_parameters(th)
                fd += w * L(c) / FD_H
        c.reset_parameters(theta)
        assert abs(fd - g[k]) < 1e-9, (k, fd, g[k])


def test_sv_vjp_state_gradient_matches_finite_differences():
    # d Re<t|U x> / d Re x_k + i d / d Im x_k = (U^dagger t)_k
    rng = np.random.default_rng(4)
    n = 4
    c = sv_circuit(n, rng)
    U = circuit_matrix(c, n)
    x, t = rand_state(n, rng), rand_state(n, rng)
    _, dx, _ = vjp(c, x, n, False, t)
    for k in range(1 << n):
        e = np.zeros(1 << n)
        e[k] = 1e-6
        dre = (np.vdot(t, U @ (x + e)).real - np.vdot(t, U @ (x - e)).real) / 2e-6
        dim = (np.vdot(t, U @ (x + 1j * e)).real - np.vdot(t, U @ (x - 1j * e)).real) / 2e-6
        assert abs(complex(dre, dim) - dx[k]) < 1e-8


def dm_circuit(n):
    c, H, _ = config3(seed=n, n=n, depth=2, p=0.05, gamma=0.03)
    return c, H


def test_dm_vjp_seeded_with_h_dagger_is_the_energy_gradient():
    n = 3
    c, H = dm_circuit(n)
    rho0 = vec(rand_density(n, np.random.default_rng(5)))
    prog = Program(c, n=n, dm=True)
    terms, nt, hpool = pack_terms(H)
    E, g, y = oracle.energy_grad(rho0, n, 1, prog.op_array(), len(prog.ops), prog.pool(), terms, nt, hpool, prog.nparams)
    seed = np.zeros(1 << (2 * n), complex)  # vec(H^dagger): H on the row bits, applied to the identity
    ident = vec(np.eye(1 << n, dtype=complex))
    for j in range(nt):
        M = hpool[terms[j].mat: terms[j].mat + (1 << (2 * terms[j].nq))].reshape(1 << terms[j].nq, -1, order="F")
        seed += full_op(2 * n, [terms[j].pos[k] for k in range(terms[j].nq)], M.conj().T) @ ident
    gv, dx, yv = vjp(c, rho0, n, True, seed)
    assert np.max(np.abs(yv - y)) < 1e-13
    assert np.max(np.abs(gv - g)) < 1e-12


def test_dm_vjp_closed_form_with_channels():
    rng = np.random.default_rng(6)
    n = 3
    c, _ = dm_circuit(n)
    S = circuit_matrix(c, 2 * n, dm=True, n=n)
    x = vec(rand_density(n, rng))
    d = rng.standard_normal(1 << (2 * n)) + 1j * rng.standard_normal(1 << (2 * n))
    g, dx, y = vjp(c, x, n, True, d)
    assert np.max(np.abs(y - S @ x)) < 1e-13
    assert np.max(np.abs(dx - S.conj().T @ d)) < 1e-12


def test_dm_vjp_parameter_gradient_matches_finite_differences():
    rng = np.random.default_rng(7)
    n = 2
    c, _ = dm_circuit(n)
    x = vec(rand_density(n, rng))
    T = rng.standard_normal(1 << (2 * n)) + 1j * rng.standard_normal(1 << (2 * n))

    def L(circ):
        return np.vdot(T, circuit_matrix(circ, 2 * n, dm=True, n=n) @ x).real

    g, _, _ = vjp(c, x, n, True, T)
    theta = np.array(c.active_parameters())
    for k in range(len(theta)):
        fd = 0.0
        for w, s in zip(FD_W, (-2, -1, 0, 1, 2)):
            if w:
                th = theta.copy()
                th[k] += s * FD_H
                c.reset_parameters(th)
                fd += w * L(c) / FD_H
        c.reset_parameters(theta)
        assert abs(fd - g[k]) < 1e-9, (k, fd, g[k])
