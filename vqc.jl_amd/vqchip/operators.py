"""QubitsTerm / QubitsOperator (QuantumCircuits.jl) and their packing.

expectation(H, psi) merges all terms that share a position tuple into one
dense matrix and does one sweep per key when every term spans at most 5
sites (hamiltonian/expecs/expecs.jl:22-28, hamiltonian/util.jl:6-15).
Longer terms take the reference's per-term route (_apply_qterm_util!,
apply_qterms.jl:47-59): Pauli strings through the bitmask kernels, any other
product of 2x2 factors through the product-term kernels (vqc_qterm_t).
"""
import ctypes

import numpy as np

from . import _lib
from .gates import I2, SIGMA_MINUS, SIGMA_PLUS, X, Y, Z, kron_le

_NAMED = {"X": X, "Y": Y, "Z": Z, "I": I2, "+": SIGMA_PLUS, "-": SIGMA_MINUS}
LARGEST_SUPPORTED_NTERMS = 5  # apply_qterms.jl:36


def _op(o):
    if isinstance(o, str):
        return _NAMED[o]
    return np.asarray(o, dtype=np.complex128)


class QubitsTerm:
    """coeff * prod_i op_i on 1-based positions, e.g. QubitsTerm({1: "X", 3: "Z"}, coeff=0.5)."""

    def __init__(self, ops, coeff=1.0):
        if isinstance(ops, dict):
            items = sorted(ops.items())
        else:
            items = sorted(ops)
        self.positions = tuple(int(p) for p, _ in items)
        if len(set(self.positions)) != len(self.positions):
            raise ValueError("repeated site in a term")
        self.oplist = [_op(o) for _, o in items]
        self.labels = [o if isinstance(o, str) else None for _, o in items]
        self.coeff = complex(coeff)

    def __add__(self, other):
        return QubitsOperator([self]) + other

    def __rmul__(self, a):
        return QubitsTerm(list(zip(self.positions, self.oplist)), self.coeff * a)

    def matrix(self):
        """_get_mat(x::QubitsTerm) = _kron_ops(reverse(oplist)) * coeff (util.jl:4):
        the first site is the least significant bit."""
        return kron_le(*self.oplist) * self.coeff

    def is_pauli(self):
        return all(l in ("X", "Y", "Z", "I") for l in self.labels)

    def adjoint(self):
        if self.is_pauli():  # Pauli strings are Hermitian: keep the labels (the Pauli-string route needs them)
            return QubitsTerm(dict(zip(self.positions, self.labels)), np.conj(self.coeff))
        return QubitsTerm(list(zip(self.positions, [o.conj().T for o in self.oplist])), np.conj(self.coeff))


class QubitsOperator:
    """Sum of QubitsTerms, stored per position key (QuantumCircuits.QubitsOperator)."""

    def __init__(self, terms=()):
        self.terms = list(terms)

    def __add__(self, other):
        if isinstance(other, QubitsTerm):
            return QubitsOperator(self.terms + [other])
        return QubitsOperator(self.terms + list(other.terms))

    __radd__ = __add__

    def __rmul__(self, a):
        return QubitsOperator([a * t for t in self.terms])

    def adjoint(self):
        return QubitsOperator([t.adjoint() for t in self.terms])

    def largest_nterm(self):
        return max((len(t.positions) for t in self.terms), default=0)

    def data(self):
        """key -> merged dense matrix, keys in first-appearance order (util.jl:6-15)."""
        out = {}
        for t in self.terms:
            m = t.matrix()
            if t.positions in out:
                out[t.positions] = out[t.positions] + m
            else:
                out[t.positions] = m
        return out

    def ishermitian(self):
        return all(np.allclose(m, m.conj().T) for m in self.data().values())


def pack_terms(op, pool=None):
    """QubitsOperator (all terms <= 5 sites) -> (Term array, number of terms,
    complex pool). With a shared `pool` (a MatrixPool) the matrices are appended
    to it, the offsets index it, and the third value is None: the pool's owner
    concatenates it once."""
    data = op.data()
    chunks = _lib.MatrixPool() if pool is None else pool
    terms = (_lib.Term * max(len(data), 1))()
    for i, (key, m) in enumerate(data.items()):
        if len(key) > LARGEST_SUPPORTED_NTERMS:
            raise ValueError("term longer than 5 sites: use the Pauli form")
        terms[i].nq = len(key)
        for j, p in enumerate(key):
            terms[i].pos[j] = p - 1
        terms[i].mat = chunks.add(np.ascontiguousarray(m, dtype=np.complex128).reshape(-1, order="F"))
    return terms, len(data), (chunks.array() if pool is None else None)


def pack_pauli(op):
    """QubitsOperator of Pauli strings -> vqc_pauli_t array."""
    arr = (_lib.Pauli * max(len(op.terms), 1))()
    for i, t in enumerate(op.terms):
        if not t.is_pauli():
            raise ValueError("term is not a Pauli string")
        x = z = 0
        for p, l in zip(t.positions, t.labels):
            b = 1 << (p - 1)
            if l == "X":
                x |= b
            elif l == "Z":
                z |= b
            elif l == "Y":
                x |= b
                z |= b
        arr[i].x, arr[i].z = x, z
        arr[i].re, arr[i].im = t.coeff.real, t.coeff.imag
    return arr, len(op.terms)


def split_long(op):
    """(the terms of at most 5 sites as a QubitsOperator, the longer Pauli strings
    as a QubitsOperator, the longer other terms as a list)."""
    short, pauli, other = [], [], []
    for t in op.terms:
        if len(t.positions) <= LARGEST_SUPPORTED_NTERMS:
            short.append(t)
        elif t.is_pauli():
            pauli.append(t)
        else:
            other.append(t)
    return QubitsOperator(short), QubitsOperator(pauli), other


def pack_qterms(terms):
    """QubitsTerms -> (vqc_qterm_t array, count, int32 positions, complex pool):
    each term's 0-based sites and its 2x2 factors in site order."""
    arr = (_lib.QTerm * max(len(terms), 1))()
    pos, chunks = [], _lib.MatrixPool()
    for i, t in enumerate(terms):
        arr[i].nsites = len(t.positions)
        arr[i].pos = len(pos)
        arr[i].re, arr[i].im = t.coeff.real, t.coeff.imag
        pos.extend(p - 1 for p in t.positions)
        arr[i].mat = chunks.size
        for m in t.oplist:
            chunks.add(np.ascontiguousarray(np.asarray(m, dtype=np.complex128).reshape(-1, order="F")))
    parr = (ctypes.c_int32 * max(len(pos), 1))(*pos)
    return arr, len(terms), parr, chunks.array()
