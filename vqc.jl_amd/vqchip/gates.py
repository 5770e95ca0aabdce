"""Gates and channels: the part of QuantumCircuits.jl's model the VQC.jl path uses.

Positions are 1-based, as in Julia. A gate's matrix is defined over its
positions IN THE GIVEN ORDER with the first position as the least significant
bit (QuantumGate((j, k), kron(m1, m2)) acts as m2 on j and m1 on k,
test/check_qubits_ham.jl:22-31); `ordered_mat` re-expresses it over ascending
positions, which is the form every kernel takes (generic.jl:43-50).

Matrix conventions pinned by the reference's kernels or tests: CNOT and
FREDKIN controls come first (specific/czgate.jl:60-64, fredkingate.jl:65-74);
PHASE, CZ, CPHASE, CCPHASE, SWAP, iSWAP follow specific/*.jl; Ry(t) =
exp(-i t Y / 2) (circuitdiff.jl:10-16 with statevector.jl:85). Rx, Rz, H, S,
T, FSIM, CRx/CRy/CRz, Depolarizing and AmplitudeDamping are not pinned by any
reference file; they use the textbook definitions below.

Each parameterised gate also provides its generators G_j = U^dagger dU/dtheta_j
(the adjoint sweep evaluates Re <Delta|G_j|psi>, circuitdiff.jl:23-32).
"""
import numpy as np

from . import _lib

I2 = np.eye(2, dtype=np.complex128)
X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
H = np.array([[1, 1], [1, -1]], dtype=np.complex128) / np.sqrt(2)
P0 = np.array([[1, 0], [0, 0]], dtype=np.complex128)
P1 = np.array([[0, 0], [0, 1]], dtype=np.complex128)
SIGMA_PLUS = np.array([[0, 1], [0, 0]], dtype=np.complex128)
SIGMA_MINUS = np.array([[0, 0], [1, 0]], dtype=np.complex128)
ZERO = np.array([1, 0], dtype=np.complex128)
ONE = np.array([0, 1], dtype=np.complex128)
SWAP = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=np.complex128)


def kron_le(*ops):
    """Kronecker product with ops[0] on the least significant bit."""
    m = np.array([[1.0 + 0j]])
    for op in ops:
        m = np.kron(op, m)
    return m


def reorder(mat, order):
    """Matrix over bits given in `order` -> matrix over sorted bits.

    mat index bit b corresponds to qubit order[b]; the result's bit j
    corresponds to sorted(order)[j].
    """
    k = len(order)
    srt = sorted(order)
    if list(order) == srt:
        return np.ascontiguousarray(mat, dtype=np.complex128)
    where = [order.index(q) for q in srt]  # sorted bit j <- given bit where[j]
    D = 1 << k
    perm = np.empty(D, dtype=np.int64)
    for j in range(D):  # j: index over sorted bits -> index over given bits
        g = 0
        for b in range(k):
            if j >> b & 1:
                g |= 1 << where[b]
        perm[j] = g
    return np.ascontiguousarray(mat[np.ix_(perm, perm)], dtype=np.complex128)


def R(k):
    """diag(1, exp(2 pi i / 2^k)): the phase step of the QFT (src/utility/qft.jl:11).
    Its sign is fixed by the phase-estimation known answers (test/algs.jl:9-24),
    whose inverse QFT must read back exp(+2 pi i phi)."""
    return np.diag([1, np.exp(2j * np.pi / (1 << k))]).astype(np.complex128)


def CONTROL(U):
    """Controlled-U as a matrix over (control, targets...): the control is the
    first position, as for CNOT (specific/czgate.jl:60-64), so CONTROL(X) is
    CNOT's matrix. The phase-estimation circuit of test/algs.jl:17-20 puts its
    controls first in the same way."""
    U = np.asarray(U, dtype=np.complex128)
    eye = np.eye(U.shape[0], dtype=np.complex128)
    return kron_le(P0, eye) + kron_le(P1, U)


class Gate:
    """A unitary on `positions` (1-based), matrix `mat` over the given order."""

    name = "Gate"
    named_kind = None  # libvqc named-gate kind when a specialised kernel exists

    def __init__(self, positions, mat, params=(), isparas=None, gens=None):
        self.positions = tuple(int(p) for p in positions)
        if len(set(self.positions)) != len(self.positions):
            raise ValueError(f"repeated qubit position in {self.positions}")
        if min(self.positions) < 1:
            raise ValueError("qubit positions are 1-based")
        self._mat = np.ascontiguousarray(mat, dtype=np.complex128)
        self.params = tuple(float(x) for x in params)
        if isparas is None:
            isparas = [True] * len(self.params)
        elif isinstance(isparas, bool):
            isparas = [isparas] * len(self.params)
        self.isparas = list(isparas)
        self._gens = gens  # generators over the given order (list of matrices)

    # --- QuantumCircuits.jl accessors -------------------------------------
    def ordered_positions(self):
        return tuple(sorted(self.positions))

    def mat(self):
        return self._mat

    def ordered_mat(self):
        return reorder(self._mat, list(self.positions))

    def nparameters(self):
        return sum(1 for a in self.isparas if a)

    def parameters(self):
        return list(self.params)

    def active_parameters(self):
        return [p for p, a in zip(self.params, self.isparas) if a]

    def differentiate(self):
        """G_j = U^dagger dU/dtheta_j over ascending positions, for active parameters."""
        if not self._gens:
            return []
        return [reorder(g, list(self.positions)) for g, a in zip(self._gens, self.isparas) if a]

    def reset_parameters(self, values):
        raise NotImplementedError(f"{self.name} has no parameters")

    def adjoint(self):
        return Gate(self.positions, self._mat.conj().T)

    def __repr__(self):
        return f"{self.name}{self.positions}" + (f"{self.params}" if self.params else "")


def _rot(P, theta):
    return np.cos(theta / 2) * np.eye(P.shape[0]) - 1j * np.sin(theta / 2) * P


class _Param(Gate):
    """exp(-i theta A) style gates: rebuilt from their parameters."""

    def reset_parameters(self, values):
        vals = list(self.params)
        it = iter(values)
        for j, a in enumerate(self.isparas):
            if a:
                vals[j] = float(next(it))
        new = type(self)(*self.positions, *vals, isparas=self.isparas)
        self.__dict__.update(new.__dict__)


class _SelfAdjoint(Gate):
    """Gates equal to their adjoint: gate' keeps the gate (and its named kernel)."""

    def adjoint(self):
        return type(self)(*self.positions)


class _PhaseFamily(_Param):
    """PHASE, CPHASE, CCPHASE: gate' is the same gate at -phi (its named kernel
    too), with the parameter inactive as for any adjoint gate."""

    def adjoint(self):
        return type(self)(*self.positions, -self.params[0], isparas=False)


class XGate(_SelfAdjoint):
    name, named_kind = "X", _lib.GATE_X

    def __init__(self, q):
        super().__init__((q,), X)


class YGate(_SelfAdjoint):
    name = "Y"

    def __init__(self, q):
        super().__init__((q,), Y)


class ZGate(_SelfAdjoint):
    name = "Z"

    def __init__(self, q):
        super().__init__((q,), Z)


class HGate(_SelfAdjoint):
    name = "H"

    def __init__(self, q):
        super().__init__((q,), H)


class SGate(Gate):
    name = "S"

    def __init__(self, q):
        super().__init__((q,), np.diag([1, 1j]))


class TGate(Gate):
    name = "T"

    def __init__(self, q):
        super().__init__((q,), np.diag([1, np.exp(1j * np.pi / 4)]))


class RxGate(_Param):
    name = "Rx"

    def __init__(self, q, theta, isparas=True):
        super().__init__((q,), _rot(X, theta), (theta,), isparas, [-0.5j * X])


class RyGate(_Param):
    name = "Ry"

    def __init__(self, q, theta, isparas=True):
        super().__init__((q,), _rot(Y, theta), (theta,), isparas, [-0.5j * Y])


class RzGate(_Param):
    name = "Rz"

    def __init__(self, q, theta, isparas=True):
        super().__init__((q,), _rot(Z, theta), (theta,), isparas, [-0.5j * Z])


class PHASEGate(_PhaseFamily):
    """|1> -> e^{i phi}|1> (specific/phasegate.jl:6-22)."""

    name, named_kind = "PHASE", _lib.GATE_PHASE

    def __init__(self, q, phi, isparas=True):
        super().__init__((q,), np.diag([1, np.exp(1j * phi)]), (phi,), isparas, [1j * P1])


class CNOTGate(_SelfAdjoint):
    """control = first position (specific/czgate.jl:60-64)."""

    name, named_kind = "CNOT", _lib.GATE_CNOT

    def __init__(self, c, t):
        super().__init__((c, t), kron_le(P0, I2) + kron_le(P1, X))


class CZGate(_SelfAdjoint):
    name, named_kind = "CZ", _lib.GATE_CZ

    def __init__(self, a, b):
        super().__init__((a, b), np.diag([1, 1, 1, -1]))


class CPHASEGate(_PhaseFamily):
    name, named_kind = "CPHASE", _lib.GATE_CPHASE

    def __init__(self, a, b, phi, isparas=True):
        P11 = np.diag([0, 0, 0, 1]).astype(np.complex128)
        super().__init__((a, b), np.diag([1, 1, 1, np.exp(1j * phi)]), (phi,), isparas, [1j * P11])


class SWAPGate(_SelfAdjoint):
    name, named_kind = "SWAP", _lib.GATE_SWAP

    def __init__(self, a, b):
        super().__init__((a, b), SWAP)


class iSWAPGate(Gate):
    """|01>,|10> swapped with factor i (specific/swapgates.jl:25-49)."""

    name, named_kind = "iSWAP", _lib.GATE_ISWAP

    def __init__(self, a, b):
        super().__init__((a, b), np.array([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]]))

    def adjoint(self):
        return iSWAPdagGate(*self.positions)


class iSWAPdagGate(Gate):
    """iSWAPGate': |01>,|10> swapped with factor -i (specific/swapgates.jl:47-49)."""

    name, named_kind = "iSWAP'", _lib.GATE_ISWAPDAG

    def __init__(self, a, b):
        super().__init__((a, b), np.array([[1, 0, 0, 0], [0, 0, -1j, 0], [0, -1j, 0, 0], [0, 0, 0, 1]]))

    def adjoint(self):
        return iSWAPGate(*self.positions)


def _controlled(P, theta):
    return kron_le(P0, I2) + kron_le(P1, _rot(P, theta)), -0.5j * kron_le(P1, P)


class CRxGate(_Param):
    name = "CRx"

    def __init__(self, c, t, theta, isparas=True):
        m, g = _controlled(X, theta)
        super().__init__((c, t), m, (theta,), isparas, [g])


class CRyGate(_Param):
    name = "CRy"

    def __init__(self, c, t, theta, isparas=True):
        m, g = _controlled(Y, theta)
        super().__init__((c, t), m, (theta,), isparas, [g])


class CRzGate(_Param):
    name = "CRz"

    def __init__(self, c, t, theta, isparas=True):
        m, g = _controlled(Z, theta)
        super().__init__((c, t), m, (theta,), isparas, [g])


class FSIMGate(_Param):
    """[[1,0,0,0],[0,cos t,-i sin t,0],[0,-i sin t,cos t,0],[0,0,0,e^{-i phi}]]."""

    name = "FSIM"

    def __init__(self, a, b, theta, phi, isparas=True):
        m = np.array(
            [[1, 0, 0, 0], [0, np.cos(theta), -1j * np.sin(theta), 0], [0, -1j * np.sin(theta), np.cos(theta), 0],
             [0, 0, 0, np.exp(-1j * phi)]])
        A = np.zeros((4, 4), dtype=np.complex128)
        A[1, 2] = A[2, 1] = 1
        P11 = np.diag([0, 0, 0, 1]).astype(np.complex128)
        super().__init__((a, b), m, (theta, phi), isparas, [-1j * A, -1j * P11])


class TOFFOLIGate(_SelfAdjoint):
    name = "TOFFOLI"

    def __init__(self, a, b, c):
        m = np.eye(8, dtype=np.complex128)
        m[[3, 7]] = m[[7, 3]]
        super().__init__((a, b, c), m)


class FREDKINGate(_SelfAdjoint):
    """control = first position, swaps the other two (specific/fredkingate.jl:65-74)."""

    name, named_kind = "FREDKIN", _lib.GATE_FREDKIN

    def __init__(self, c, a, b):
        m = np.eye(8, dtype=np.complex128)
        m[[3, 5]] = m[[5, 3]]  # |c=1,a=1,b=0> (3) <-> |c=1,a=0,b=1> (5)
        super().__init__((c, a, b), m)


class CCPHASEGate(_PhaseFamily):
    name, named_kind = "CCPHASE", _lib.GATE_CCPHASE

    def __init__(self, a, b, c, phi, isparas=True):
        d = np.ones(8, dtype=np.complex128)
        d[7] = np.exp(1j * phi)
        g = np.zeros((8, 8), dtype=np.complex128)
        g[7, 7] = 1j
        super().__init__((a, b, c), np.diag(d), (phi,), isparas, [g])


def from_external(positions, mat):
    """A dense gate from a matrix in the external order: kron(m_1, ..., m_k) puts m_1
    on positions[0] (the first factor of the kron is the first position), the order
    test/check_high_qubit_gate.jl:7-8 uses against QubitsTerm(j => m_1, ...). It is
    QuantumGate over the reversed positions."""
    positions = tuple(positions) if not isinstance(positions, int) else (positions,)
    return QuantumGate(tuple(reversed(positions)), mat)


def CONTROLGate(positions, U):
    """Controlled-U on positions (control, targets...): CONTROL(U) over the given
    order, the control first (test/check_gateop.jl:104)."""
    positions = tuple(positions)
    U = np.asarray(U, dtype=np.complex128)
    if U.shape != (1 << (len(positions) - 1),) * 2:
        raise ValueError("U must act on the positions after the control")
    g = Gate(positions, CONTROL(U))
    g.name = "CONTROL"
    return g


def CONTROLCONTROLGate(a, b, c, U):
    """U on c when both controls a and b are 1 (test/check_gateop.jl:144)."""
    U = np.asarray(U, dtype=np.complex128)
    if U.shape != (2, 2):
        raise ValueError("U must be a 2 x 2 matrix")
    m = np.eye(8, dtype=np.complex128)
    m[np.ix_([3, 7], [3, 7])] = U  # over (a, b, c): a = b = 1 is index 3 (c = 0) and 7 (c = 1)
    g = Gate((a, b, c), m)
    g.name = "CONTROLCONTROL"
    return g


def QuantumGate(positions, mat):
    """A dense gate on `positions` (given order, first = least significant)."""
    positions = tuple(positions) if not isinstance(positions, int) else (positions,)
    m = np.asarray(mat, dtype=np.complex128)
    if m.shape != (1 << len(positions),) * 2:
        raise ValueError("matrix size does not match the number of positions")
    return Gate(positions, m)


# ---------------------------------------------------------------------------
# channels (QuantumCircuits.QuantumMap): Kraus operators -> superoperator
# ---------------------------------------------------------------------------
class QuantumMap:
    """A channel on `positions` given by Kraus operators over ascending positions.

    ordered_supermat() = sum_i conj(K_i) (x) K_i: the column bits (pos .+ n) are the
    high half and the row bits (pos) the low half, matching vec(rho) with rho[i,j]
    at i + j 2^n (apply_qmaps.jl:52-62).
    """

    name = "QuantumMap"

    def __init__(self, positions, kraus):
        self.positions = tuple(sorted(int(p) for p in (positions if not isinstance(positions, int) else (positions,))))
        self.kraus = [np.asarray(k, dtype=np.complex128) for k in kraus]

    def ordered_positions(self):
        return self.positions

    def ordered_supermat(self):
        return sum(np.kron(k.conj(), k) for k in self.kraus)

    def nparameters(self):
        return 0

    def __repr__(self):
        return f"{self.name}{self.positions}"


def Depolarizing(q, p):
    """rho -> (1-p) rho + p/3 (X rho X + Y rho Y + Z rho Z). The reference uses
    QuantumCircuits.Depolarizing(i, p=...) (test/check_noisy_grad.jl:6), whose
    parameterisation is not in the reference; this one is recorded as a choice."""
    m = QuantumMap((q,), [np.sqrt(1 - p) * I2, np.sqrt(p / 3) * X, np.sqrt(p / 3) * Y, np.sqrt(p / 3) * Z])
    m.name = "Depolarizing"
    return m


def AmplitudeDamping(q, gamma):
    """Kraus {[[1,0],[0,sqrt(1-g)]], [[0,sqrt(g)],[0,0]]} (Nielsen & Chuang 8.3.5)."""
    k0 = np.array([[1, 0], [0, np.sqrt(1 - gamma)]], dtype=np.complex128)
    k1 = np.array([[0, np.sqrt(gamma)], [0, 0]], dtype=np.complex128)
    m = QuantumMap((q,), [k0, k1])
    m.name = "AmplitudeDamping"
    return m
