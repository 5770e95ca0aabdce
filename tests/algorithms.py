"""The phase-estimation circuit of the reference's algorithm tests
(test/algs.jl:1-24), built from the host builders, shared by the CPU and GPU
tests."""
import numpy as np

from vqchip import CONTROL, HGate, QCircuit, QFT, QuantumGate


def to_digits(bits):
    """sum_i s_i 2^-i (test/algs.jl:1-7)."""
    return sum(b * 2.0 ** -(i + 1) for i, b in enumerate(bits))


def phase_estimation_circuit(bits):
    """H on qubits 1..L, controlled U^(2^(L-i)) from qubit i onto qubit L+1 with
    U = diag(exp(2 pi i phi), 1), then the inverse QFT on qubits 1..L."""
    L = len(bits)
    U = np.diag([np.exp(2j * np.pi * to_digits(bits)), 1.0])
    c = QCircuit(HGate(i) for i in range(1, L + 1))
    tmp = U
    for i in range(L, 0, -1):
        c.append(QuantumGate((i, L + 1), CONTROL(tmp)))
        tmp = tmp @ tmp
    c.extend(QFT(L).adjoint())
    return c
