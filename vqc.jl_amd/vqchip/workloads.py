"""Seeded synthetic workloads for the five measured configurations.

Random numbers come from numpy's PCG64 (`numpy.random.default_rng(seed)`);
per-instance streams use seed + instance index through SeedSequence. Haar
unitaries follow Mezzadri (2007): QR of a complex Gaussian matrix with the
phases of R's diagonal moved into Q.
"""
import numpy as np

from .circuit import QCircuit
from .gates import AmplitudeDamping, CNOTGate, Depolarizing, QuantumGate, RxGate, RyGate, RzGate
from .operators import QubitsOperator, QubitsTerm

DEFAULT_SEED = 20261203


def haar_unitary(k, rng):
    D = 1 << k
    Z = (rng.standard_normal((D, D)) + 1j * rng.standard_normal((D, D))) / np.sqrt(2)
    Q, R = np.linalg.qr(Z)
    d = np.diag(R)
    return Q * (d / np.abs(d))


def hea_layer(circuit, n, thetas, it):
    """Rx, Ry, Rz on every qubit, then CNOT(i, i+1) for i = 1..n-1."""
    for i in range(1, n + 1):
        circuit.append(RxGate(i, thetas[next(it)]))
        circuit.append(RyGate(i, thetas[next(it)]))
        circuit.append(RzGate(i, thetas[next(it)]))
    for i in range(1, n):
        circuit.append(CNOTGate(i, i + 1))


def hea_circuit(n, depth, thetas):
    c = QCircuit()
    it = iter(range(len(thetas)))
    for _ in range(depth):
        hea_layer(c, n, thetas, it)
    return c


def tfim(n, h=1.0):
    """H = sum_i Z_i Z_{i+1} + h sum_i X_i, open chain."""
    terms = [QubitsTerm({i: "Z", i + 1: "Z"}) for i in range(1, n)]
    terms += [QubitsTerm({i: "X"}, coeff=h) for i in range(1, n + 1)]
    return QubitsOperator(terms)


def heisenberg(n, J):
    """H = sum_i J_i (XX + YY + ZZ)_{i,i+1}, open chain."""
    terms = []
    for i in range(1, n):
        for p in "XYZ":
            terms.append(QubitsTerm({i: p, i + 1: p}, coeff=J[i - 1]))
    return QubitsOperator(terms)


def sum_z(n):
    return QubitsOperator([QubitsTerm({i: "Z"}) for i in range(1, n + 1)])


def config0(seed=DEFAULT_SEED, n=12, depth=4):
    """12 qubits, 4 HEA layers, 144 angles ~ U[0, 2pi), TFIM with h = 1."""
    rng = np.random.default_rng(seed)
    thetas = rng.uniform(0, 2 * np.pi, 3 * n * depth)
    return hea_circuit(n, depth, thetas), tfim(n, 1.0), thetas


def config1_circuit(seed=DEFAULT_SEED, n=28, layers=20, n1=14, n2=7, n3=3):
    """Per layer: n1 one-qubit, n2 two-qubit, n3 three-qubit Haar gates on
    uniformly drawn distinct targets (distinct within each gate)."""
    rng = np.random.default_rng(seed)
    c = QCircuit()
    for _ in range(layers):
        for k, cnt in ((1, n1), (2, n2), (3, n3)):
            for _ in range(cnt):
                pos = tuple(int(q) + 1 for q in rng.choice(n, size=k, replace=False))
                c.append(QuantumGate(pos, haar_unitary(k, rng)))
    return c


def random_pauli_operator(n, nterms, seed=DEFAULT_SEED + 1):
    """nterms Pauli strings, each qubit's Pauli uniform in {I,X,Y,Z}, coeff ~ N(0,1)."""
    rng = np.random.default_rng(seed)
    terms = []
    for _ in range(nterms):
        labels = rng.integers(0, 4, size=n)
        ops = {q + 1: "IXYZ"[l] for q, l in enumerate(labels) if l}
        if not ops:
            ops = {1: "I"}
        terms.append(QubitsTerm(ops, coeff=float(rng.standard_normal())))
    return QubitsOperator(terms)


def config2_instances(seed=DEFAULT_SEED, count=256, nmin=8, nmax=20, dmin=2, dmax=12):
    """Uneven batch: (circuit, Heisenberg H, n) per instance."""
    out = []
    for i in range(count):
        rng = np.random.default_rng([seed, i])
        n = int(rng.integers(nmin, nmax + 1))
        depth = int(rng.integers(dmin, dmax + 1))
        thetas = rng.uniform(0, 2 * np.pi, 3 * n * depth)
        J = rng.standard_normal(n - 1)
        out.append((hea_circuit(n, depth, thetas), heisenberg(n, J), n))
    return out


def toffoli_matrix():
    """CCX as an 8 x 8 matrix: the target flips when both controls are 1 (its
    two highest basis states swap), whatever the order of the three positions."""
    m = np.eye(8, dtype=complex)
    m[[6, 7]] = m[[7, 6]]
    return m


def config2_toffoli_instances(seed=DEFAULT_SEED, count=256, nmin=8, nmax=20, dmin=2, dmax=12):
    """The uneven batch with one Toffoli per layer (on qubits k, k+1, k+2 of
    layer k, cyclically), so that every instance carries 3-qubit gates."""
    out = []
    for i in range(count):
        rng = np.random.default_rng([seed, i])
        n = int(rng.integers(nmin, nmax + 1))
        depth = int(rng.integers(dmin, dmax + 1))
        thetas = rng.uniform(0, 2 * np.pi, 3 * n * depth)
        J = rng.standard_normal(n - 1)
        c = QCircuit()
        it = iter(range(len(thetas)))
        for k in range(depth):
            hea_layer(c, n, thetas, it)
            q = k % (n - 2)
            c.append(QuantumGate((q + 1, q + 2, q + 3), toffoli_matrix()))
        out.append((c, heisenberg(n, J), n))
    return out


def config3(seed=DEFAULT_SEED, n=12, depth=6, p=0.01, gamma=0.005):
    """HEA layers with Depolarizing(p) on each CNOT's target after the CNOT and
    AmplitudeDamping(gamma) on every qubit at the end of each layer; observable
    sum_i Z_i."""
    rng = np.random.default_rng(seed)
    thetas = rng.uniform(0, 2 * np.pi, 3 * n * depth)
    c = QCircuit()
    it = iter(range(len(thetas)))
    for _ in range(depth):
        for i in range(1, n + 1):
            c.append(RxGate(i, thetas[next(it)]))
            c.append(RyGate(i, thetas[next(it)]))
            c.append(RzGate(i, thetas[next(it)]))
        for i in range(1, n):
            c.append(CNOTGate(i, i + 1))
            c.append(Depolarizing(i + 1, p))
        for i in range(1, n + 1):
            c.append(AmplitudeDamping(i, gamma))
    return c, sum_z(n), thetas


def config4_request(rng, nmin=10, nmax=26, dmin=4, dmax=40):
    """One serving request: n log-uniform in [nmin, nmax], depth ~ U{dmin..dmax}, TFIM h ~ U(0.5, 1.5)."""
    n = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
    depth = int(rng.integers(dmin, dmax + 1))
    thetas = rng.uniform(0, 2 * np.pi, 3 * n * depth)
    h = float(rng.uniform(0.5, 1.5))
    return hea_circuit(n, depth, thetas), tfim(n, h), n
