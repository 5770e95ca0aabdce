"""Krylov ground state and time evolution on the device (groundstate.jl) against
dense numpy / scipy on the full 2^n operator. Tolerances: eigenvalue 1e-9
relative, residual |Hv - Ev| < 1e-6, evolved amplitudes 1e-9."""
import numpy as np
import pytest
import scipy.linalg

from brute import full_op, rand_state
import vqchip as vq
from vqchip import krylov
from vqchip.operators import pack_terms
from vqchip.utility import heisenberg_chain, ising_chain

pytestmark = pytest.mark.gpu


def dense(H, n):
    terms, nt, hpool = pack_terms(H)
    M = np.zeros((1 << n, 1 << n), complex)
    for j in range(nt):
        k = terms[j].nq
        D = 1 << k
        M += full_op(n, [terms[j].pos[q] for q in range(k)], hpool[terms[j].mat:terms[j].mat + D * D].reshape(D, D, order="F"))
    return M


@pytest.mark.parametrize("H,n", [(ising_chain(8, J=1.0, hz=0.7), 8), (heisenberg_chain(10, J=1.0, Jzz=0.5, hz=0.1), 10)])
def test_ground_state(H, n):
    E, v = krylov.ground_state(H, n, rng=np.random.default_rng(3))
    Hd = dense(H, n)
    E_ref = np.linalg.eigvalsh(Hd)[0]
    assert abs(E - E_ref) < 1e-9 * max(1, abs(E_ref))
    x = v.storage()
    assert abs(np.linalg.norm(x) - 1) < 1e-12
    assert np.linalg.norm(Hd @ x - E * x) < 1e-6


@pytest.mark.parametrize("t", [-0.8j, 0.3, 2.5 - 1.5j])
def test_time_evolution(t):
    n = 7
    H = heisenberg_chain(n, J=0.9, hz=0.2)
    rng = np.random.default_rng(5)
    x0 = rand_state(n, rng)
    out = krylov.time_evolution(H, t, vq.StateVector(x0))
    ref = scipy.linalg.expm(t * dense(H, n)) @ x0
    assert np.max(np.abs(out.storage() - ref)) < 1e-9 * max(1, np.max(np.abs(ref)))


def test_axpy():
    rng = np.random.default_rng(6)
    x, y = rand_state(12, rng), rand_state(12, rng)
    sx, sy = vq.StateVector(x), vq.StateVector(y)
    krylov._axpy(sy, 0.3 - 1.1j, sx)
    assert np.max(np.abs(sy.storage() - (y + (0.3 - 1.1j) * x))) < 1e-15
