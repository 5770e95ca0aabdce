"""QCircuit and its packing into the C-ABI program format (vqc_op_t + matrix pool)."""
import numpy as np

from . import _lib
from .gates import CNOTGate, Gate, PHASEGate, QuantumMap, RxGate, RyGate, RzGate, SGate, TGate, ZGate


class QCircuit(list):
    """An ordered list of gates and channels (QuantumCircuits.QCircuit)."""

    def nparameters(self):
        return sum(g.nparameters() for g in self)

    def active_parameters(self):
        out = []
        for g in self:
            if isinstance(g, Gate):
                out.extend(g.active_parameters())
        return out

    def reset_parameters(self, values):
        """reset_parameters!(circuit, values) (test/parameter.jl:39-43)."""
        values = list(values)
        if len(values) != self.nparameters():
            raise ValueError("wrong number of parameters")
        i = 0
        for g in self:
            n = g.nparameters()
            if n:
                g.reset_parameters(values[i:i + n])
                i += n
        return self

    def adjoint(self):
        return QCircuit(g.adjoint() for g in reversed(self))


def _structure_hint(g):
    """vqc_op_t.reserved: bit 0 U diagonal, bit 1 generators diagonal, bit 2 / 3
    U is a CNOT with control on the lower / higher position, bit 4 / 5 U real /
    X-like, bit 6 / 7 generators real / X-like (real diagonal, imaginary
    off-diagonal). Batches plan before their angles arrive, so the planner takes
    a gate's structure from its type; every angle keeps it (Ry(t) is real and
    its generator -iY/2 too; Rx(t) and -iX/2 are X-like)."""
    if isinstance(g, (RzGate, PHASEGate, ZGate, SGate, TGate)):
        return 1 | 2
    if isinstance(g, RyGate):
        return 16 | 64
    if isinstance(g, RxGate):
        return 32 | 128
    if isinstance(g, CNOTGate):
        c, t = g.positions
        return 4 if c < t else 8
    return 0


class Program:
    """A circuit packed as vqc_op_t records plus one complex matrix pool.

    The gradient slot of each active parameter is its index in
    active_parameters(circuit): forward circuit order, as the reference
    flattens gradients (circuitdiff.jl:61-75).
    """

    def __init__(self, circuit, n=None, dm=False, slot_base=0, pool=None):
        self.ops = []
        self._chunks = _lib.MatrixPool() if pool is None else pool
        slot = slot_base
        for g in circuit:
            op = _lib.Op()
            if isinstance(g, QuantumMap):
                if not dm:
                    raise ValueError("a channel in a state-vector program: the reference continues on "
                                     "DensityMatrix(state) (apply_qmaps.jl:19); pack it for a density matrix")
                S = g.ordered_supermat()
                op.kind = _lib.OP_CHANNEL
                op.mat = self._add(S)
                # S^-1 for the reverse sweep; a singular S (depolarizing at p = 3/4) has
                # none, and its reverse restores a forward-state checkpoint instead
                if np.linalg.cond(S) < 1e14:
                    op.aux = self._add(np.linalg.inv(S))
                else:
                    op.aux = -1
                op.nparams = 0
                pos = g.ordered_positions()
            else:
                op.kind = _lib.OP_UNITARY
                op.mat = self._add(g.ordered_mat())
                op.aux = -1
                pos = g.ordered_positions()
                gens = g.differentiate()
                op.nparams = len(gens)
                if len(gens) > 3:
                    raise ValueError("at most 3 parameters per gate")
                for j, G in enumerate(gens):
                    op.gen[j] = self._add(G)
                    op.slot[j] = slot
                    slot += 1
                for j in range(len(gens), 3):
                    op.slot[j] = -1
                    op.gen[j] = -1
            if len(pos) > 5:
                raise ValueError("only implement 5-qubit gates and less currently.")  # generic.jl:37
            if n is not None and (pos[0] < 1 or pos[-1] > n):
                raise ValueError(f"qubit positions {pos} outside 1..{n}")
            op.nq = len(pos)
            for j, p in enumerate(pos):
                op.pos[j] = p - 1
            if op.kind == _lib.OP_CHANNEL:
                for j in range(3):
                    op.slot[j] = -1
                    op.gen[j] = -1
            else:
                op.reserved = _structure_hint(g)
            self.ops.append(op)
        self.nparams = slot - slot_base
        self.slot_end = slot

    def _add(self, m):
        return self._chunks.add(np.ascontiguousarray(m, dtype=np.complex128).reshape(-1, order="F"))

    def pool(self):
        return self._chunks.array()

    def op_array(self):
        arr = (_lib.Op * max(len(self.ops), 1))()
        for i, o in enumerate(self.ops):
            arr[i] = o
        return arr
