"""VQC's own workload builders (src/utility/spin_hamiltonians.jl,
src/utility/variationalcircuit.jl, src/utility/qft.jl), so the device path is
driven by the same Hamiltonians and circuits a VQC.jl user builds. Positions
are 1-based."""
import numpy as np

from .circuit import QCircuit
from .gates import CONTROL, SWAP, CNOTGate, HGate, QuantumGate, R, RyGate, RzGate
from .operators import QubitsOperator, QubitsTerm


def _hopping(i, j, J, Jzz):
    t = QubitsTerm({i: "+", j: "-"}, coeff=2 * J)
    return [t, t.adjoint(), QubitsTerm({i: "Z", j: "Z"}, coeff=Jzz)]


def heisenberg_chain(L, J=1.0, Jzz=None, hz=0.0):
    """XXZ chain: hz sum Z_i + sum 2J (s+_i s-_{i+1} + h.c.) + Jzz Z_i Z_{i+1}
    (spin_hamiltonians.jl:4-20)."""
    Jzz = J if Jzz is None else Jzz
    terms = [QubitsTerm({i: "Z"}, coeff=hz) for i in range(1, L + 1)]
    for i in range(1, L):
        terms += _hopping(i, i + 1, J, Jzz)
    return QubitsOperator(terms)


heisenberg_1d = heisenberg_chain


def _index(m):
    """LinearIndices((m, n)): column-major, (i, j) -> i + (j - 1) m, 1-based."""
    return lambda i, j: i + (j - 1) * m


def heisenberg_2d(m, n=None, J=1.0, Jzz=None, hz=0.0):
    """The XXZ model on an m x n open lattice (spin_hamiltonians.jl:23-47)."""
    n = m if n is None else n
    Jzz = J if Jzz is None else Jzz
    idx = _index(m)
    terms = [QubitsTerm({i: "Z"}, coeff=hz) for i in range(1, m * n + 1)]
    for i in range(1, m + 1):
        for j in range(1, n):
            terms += _hopping(idx(i, j), idx(i, j + 1), J, Jzz)
    for i in range(1, m):
        for j in range(1, n + 1):
            terms += _hopping(idx(i, j), idx(i + 1, j), J, Jzz)
    return QubitsOperator(terms)


def ising_chain(L, J=1.0, hz=1.0):
    """hz sum Z_i + J sum X_i X_{i+1} (spin_hamiltonians.jl:51-62)."""
    terms = [QubitsTerm({i: "Z"}, coeff=hz) for i in range(1, L + 1)]
    terms += [QubitsTerm({i: "X", i + 1: "X"}, coeff=J) for i in range(1, L)]
    return QubitsOperator(terms)


ising_1d = ising_chain


def ising_2d(m, n=None, J=1.0, hz=1.0):
    """The transverse-field Ising model on an m x n open lattice (spin_hamiltonians.jl:65-84)."""
    n = m if n is None else n
    idx = _index(m)
    terms = [QubitsTerm({i: "Z"}, coeff=hz) for i in range(1, m * n + 1)]
    for i in range(1, m + 1):
        for j in range(1, n):
            terms.append(QubitsTerm({idx(i, j): "X", idx(i, j + 1): "X"}, coeff=J))
    for i in range(1, m):
        for j in range(1, n + 1):
            terms.append(QubitsTerm({idx(i, j): "X", idx(i + 1, j): "X"}, coeff=J))
    return QubitsOperator(terms)


def nparas(L, depth, real=False):
    """_nparas (variationalcircuit.jl:137-140)."""
    return (depth + 1) * L * (1 if real else 3)


def _entangle(c, L, layer):
    # odd layers ladder upwards, even layers downwards (variationalcircuit.jl:21-29)
    order = range(1, L) if layer % 2 == 1 else range(L - 1, 0, -1)
    for j in order:
        c.append(CNOTGate(j, j + 1))


def variational_circuit_1d(L, depth, thetas=None, rng=None):
    """Rz Ry Rz on every qubit, then depth x (CNOT ladder + Rz Ry Rz)
    (variationalcircuit.jl:5-38). Default angles: uniform in [0, 2 pi)."""
    if thetas is None:
        thetas = (rng or np.random.default_rng()).uniform(0, 2 * np.pi, nparas(L, depth))
    thetas = np.asarray(thetas, dtype=float)
    if thetas.size != nparas(L, depth):
        raise ValueError("wrong number of parameters.")
    it = iter(thetas)
    c = QCircuit()
    for layer in range(depth + 1):
        if layer > 0:
            _entangle(c, L, layer)
        for j in range(1, L + 1):
            c.append(RzGate(j, next(it)))
            c.append(RyGate(j, next(it)))
            c.append(RzGate(j, next(it)))
    return c


def real_variational_circuit_1d(L, depth, thetas=None, rng=None):
    """Ry on every qubit, then depth x (CNOT ladder + Ry) (variationalcircuit.jl:40-67)."""
    if thetas is None:
        thetas = (rng or np.random.default_rng()).uniform(0, 2 * np.pi, nparas(L, depth, real=True))
    thetas = np.asarray(thetas, dtype=float)
    if thetas.size != nparas(L, depth, real=True):
        raise ValueError("wrong number of parameters.")
    it = iter(thetas)
    c = QCircuit()
    for layer in range(depth + 1):
        if layer > 0:
            _entangle(c, L, layer)
        for j in range(1, L + 1):
            c.append(RyGate(j, next(it)))
    return c


def _qft_block(L):
    """H on qubit 1, then CONTROL(R(m)) * SWAP on (m, m-1) for m = 2..L: qubit 1's
    content is rotated by every later qubit while it moves up to qubit L
    (_nnqft_one_block, qft.jl:3-14)."""
    if L <= 0:
        raise ValueError("input size can not be 0")
    r = QCircuit([HGate(1)])
    for m in range(2, L + 1):
        r.append(QuantumGate((m, m - 1), CONTROL(R(m)) @ SWAP))
    return r


def QFT(L):
    """The quantum Fourier transform on qubits 1..L from nearest-neighbour gates
    only: blocks L, L-1, ..., 1 (qft.jl:19-25). With the bit reversal its swaps
    perform, it maps basis state j to sum_k exp(2 pi i j k / 2^L) |k> / 2^(L/2),
    where j and k are read with qubit 1 as the MOST significant bit."""
    r = QCircuit()
    for i in range(L, 0, -1):
        r.extend(_qft_block(i))
    return r
