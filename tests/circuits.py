"""Circuit generators shared by the GPU parity tests."""
import numpy as np

import vqchip as vq
from brute import rand_unitary


def chain_circuit(n, rng, layers=2, value_rng=None):
    """Runs of 1..7 one-qubit gates on the same qubit (fixed, active and inactive
    rotations, dense unitaries), separated by CNOT layers, so that the fusion of
    same-qubit runs meets every run length and the four-step cap.

    rng fixes the structure; value_rng (default rng) draws angles and matrices, so
    two calls with equal rng seeds and different value_rng share one structure."""
    vr = rng if value_rng is None else value_rng
    kinds = [lambda q, a: vq.RxGate(q, a), lambda q, a: vq.RyGate(q, a), lambda q, a: vq.RzGate(q, a),
             lambda q, a: vq.HGate(q), lambda q, a: vq.TGate(q), lambda q, a: vq.RzGate(q, a, isparas=False),
             lambda q, a: vq.PHASEGate(q, a), lambda q, a: vq.QuantumGate((q,), rand_unitary(1, vr))]
    c = vq.QCircuit()
    for layer in range(layers):
        for q in range(1, n + 1):
            for _ in range(1 + (q + layer) % 7):
                c.append(kinds[int(rng.integers(len(kinds)))](q, float(vr.uniform(0, 6.3))))
        for q in range(1 + layer % 2, n, 2):
            c.append(vq.CNOTGate(q, q + 1))
    return c


def _xlike(vr):
    """Y Rx(a) = [[-sin(a/2), -i cos(a/2)], [i cos(a/2), sin(a/2)]]: unitary, real
    diagonal, imaginary off-diagonal."""
    a = vr.uniform(0, 6.3)
    return np.array([[-np.sin(a / 2), -1j * np.cos(a / 2)], [1j * np.cos(a / 2), np.sin(a / 2)]])


def structured_circuit(n, rng, layers=3, value_rng=None):
    """Diagonal gates (Rz, PHASE, S, T, Z), real ones (Ry, H, a random orthogonal
    matrix), X-like ones (Rx, Y, a random real-diagonal / imaginary-off-diagonal
    matrix) next to dense ones and CNOTs with the control on either side, so that
    every structure fast path (diagonal, real and X-like steps and generators,
    CNOT as a register swap, control low or high) is exercised."""
    vr = rng if value_rng is None else value_rng
    c = vq.QCircuit()
    diag = [lambda q, a: vq.RzGate(q, a), lambda q, a: vq.PHASEGate(q, a), lambda q, a: vq.SGate(q),
            lambda q, a: vq.TGate(q), lambda q, a: vq.ZGate(q)]
    other = [lambda q, a: vq.RxGate(q, a), lambda q, a: vq.YGate(q), lambda q, a: vq.HGate(q),
             lambda q, a: vq.QuantumGate((q,), np.linalg.qr(vr.standard_normal((2, 2)))[0]),
             lambda q, a: vq.QuantumGate((q,), _xlike(vr)), lambda q, a: vq.RxGate(q, a, isparas=False)]
    for layer in range(layers):
        for q in range(1, n + 1):
            c.append(vq.RyGate(q, float(vr.uniform(0, 6.3))))
            c.append(diag[int(rng.integers(len(diag)))](q, float(vr.uniform(0, 6.3))))
            c.append(vq.RzGate(q, float(vr.uniform(0, 6.3))))
            c.append(other[int(rng.integers(len(other)))](q, float(vr.uniform(0, 6.3))))
        for q in range(1, n):
            c.append(vq.CNOTGate(q + 1, q) if (q + layer) % 2 else vq.CNOTGate(q, q + 1))
    return c
