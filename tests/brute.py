"""Brute-force references on the full 2^n operator (numpy), for tiny n.

Follows the reference tests' method of building matrix(L, h) and multiplying
(test/check_high_qubit_gate.jl:5-16, test/check_qubits_ham.jl:73-158).
"""
import numpy as np


def full_op(nbits, pos0, U):
    """2^nbits square operator of U on ascending 0-based positions pos0."""
    L = 1 << nbits
    idx = np.arange(L)
    sub = np.zeros(L, dtype=np.int64)
    for j, p in enumerate(pos0):
        sub |= ((idx >> p) & 1) << j
    mask = 0
    for p in pos0:
        mask |= 1 << p
    rest = idx & ~mask
    U = np.asarray(U)
    return U[sub[:, None], sub[None, :]] * (rest[:, None] == rest[None, :])


def pauli_full(nbits, x, z, coef):
    """c * prod sigma as a dense matrix (bit q of x/z selects X/Z, both = Y)."""
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    Y = np.array([[0, -1j], [1j, 0]])
    Z = np.diag([1.0 + 0j, -1])
    m = np.array([[1.0 + 0j]])
    for q in range(nbits):
        bx, bz = x >> q & 1, z >> q & 1
        op = Y if (bx and bz) else (X if bx else (Z if bz else np.eye(2)))
        m = np.kron(op, m)
    return coef * m


def rand_state(nbits, rng):
    v = rng.standard_normal(1 << nbits) + 1j * rng.standard_normal(1 << nbits)
    return v / np.linalg.norm(v)


def rand_matrix(k, rng):
    D = 1 << k
    return rng.standard_normal((D, D)) + 1j * rng.standard_normal((D, D))


def rand_unitary(k, rng):
    Q, R = np.linalg.qr(rand_matrix(k, rng))
    return Q * (np.diag(R) / np.abs(np.diag(R)))


def rand_density(n, rng):
    A = rand_matrix(n, rng)
    rho = A @ A.conj().T
    return rho / np.trace(rho)


def vec(rho):
    return np.ascontiguousarray(rho.reshape(-1, order="F"))


def unvec(v):
    N = int(round(np.sqrt(v.size)))
    return v.reshape(N, N, order="F")


def apply_np(v, pos0, U):
    """U on ascending 0-based positions through a tensor contraction (no 2^n x 2^n matrix)."""
    nb = int(round(np.log2(v.size)))
    k = len(pos0)
    t = v.reshape((2,) * nb)  # numpy axis a <-> bit nb-1-a
    Ut = np.asarray(U).reshape((2,) * (2 * k))  # axis i <-> bit k-1-i of the row (i<k) / column index
    in_axes = [k + i for i in range(k)]
    t_axes = [nb - 1 - pos0[k - 1 - i] for i in range(k)]
    out = np.tensordot(Ut, t, axes=(in_axes, t_axes))
    out = np.moveaxis(out, list(range(k)), t_axes)
    return np.ascontiguousarray(out.reshape(-1))
