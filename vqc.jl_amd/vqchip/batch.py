"""Batched energy + adjoint gradient of many independent circuits (no
counterpart in the reference, which evaluates one circuit per call).

    b = Batch([(circuit_1, H_1, n_1), (circuit_2, H_2, n_2), ...])
    energies, grads = b.run()            # one persistent-kernel graph launch
    b.update([circuit_1', ...])          # same structure, new angles

    b = Batch(instances, max_bytes=64 << 30)   # more state than that: waves

The adjoint keeps no forward states (the reverse sweep uncomputes), so an
instance above one tile holds exactly two states (psi and lambda, 32 * 2^n
bytes) for the length of the launch. A batch whose states exceed max_bytes runs
as consecutive waves that each fit; a wave is planned, allocated, run and freed
in turn, so device memory stays within the largest wave (its planning is redone
on every run).
"""
import ctypes

import numpy as np

from . import _lib
from .circuit import Program
from .operators import pack_terms


def instance_bytes(n, t_small=None):
    """Device bytes of one instance's states in a batch: psi and lambda above one
    tile, none when the whole evaluation stays in LDS."""
    if t_small is None:
        from .autotune import get_tuning
        t_small = get_tuning()["tile_small_bits"] or 12  # 0 (by size): batches keep up to 12 bits in LDS
    return 0 if n <= t_small else 32 << n


def _op_key(o):
    """Every field of a vqc_op_t record the batch plan depends on."""
    return (o.kind, o.nq, tuple(o.pos), o.nparams, tuple(o.slot), o.reserved, o.mat, o.aux, tuple(o.gen))


def repack(circuits, nqubits, nops, old_ops, pool_size):
    """The matrix pool of new circuits for a batch planned from old_ops.

    A batch is planned from its op records (kinds, positions, parameter slots,
    matrix offsets and the structure hints the device planner used), so every
    record of the new circuits must equal the stored one; only the matrix values
    may change. Anything else raises ValueError."""
    circuits = list(circuits)
    if len(circuits) != len(nqubits):
        raise ValueError(f"{len(circuits)} circuits for a batch of {len(nqubits)}")
    chunks = _lib.MatrixPool()
    k = 0
    for i, c in enumerate(circuits):
        prog = Program(c, n=nqubits[i], pool=chunks)
        if len(prog.ops) != nops[i]:
            raise ValueError(f"instance {i}: {len(prog.ops)} ops instead of {nops[i]}; create a new Batch")
        for j, o in enumerate(prog.ops):
            if _op_key(o) != _op_key(old_ops[k + j]):
                raise ValueError(f"instance {i}, op {j}: kind, qubits, parameters, matrix offsets or structure "
                                 "changed; create a new Batch")
        k += len(prog.ops)
    pool = chunks.array()
    if pool.size != pool_size:
        raise ValueError("circuit structure changed; create a new Batch")
    return pool


def waves(instances, max_bytes):
    """Consecutive groups of instances whose states fit max_bytes (an instance
    larger than that forms a wave of its own)."""
    out, cur, used = [], [], 0
    for inst in instances:
        b = instance_bytes(inst[2])
        if cur and used + b > max_bytes:
            out.append(cur)
            cur, used = [], 0
        cur.append(inst)
        used += b
    if cur:
        out.append(cur)
    return out


class Batch:
    """One persistent-kernel batch, or consecutive waves of them (max_bytes)."""

    def __init__(self, instances, max_bytes=None):
        instances = list(instances)
        if max_bytes is not None and sum(instance_bytes(n) for _, _, n in instances) > max_bytes:
            self._waves = waves(instances, max_bytes)
            self._part = None
        else:
            self._waves = None
            self._part = _Part(instances)
        self.n_inst = len(instances)

    def update(self, circuits):
        if self._part is not None:
            return self._part.update(circuits)
        circuits = list(circuits)
        k = 0
        for w in self._waves:
            for i, (_, H, n) in enumerate(w):
                w[i] = (circuits[k], H, n)
                k += 1

    def run(self):
        if self._part is not None:
            return self._part.run()
        E, grads = [], []
        for w in self._waves:
            part = _Part(w)
            e, g = part.run()
            part.close()
            E.append(e)
            grads += g
        return np.concatenate(E), grads

    def rebuild(self):
        if self._part is None:
            raise RuntimeError("a batch in waves is planned per run")
        self._part.rebuild()

    def launch(self):
        if self._part is None:
            raise RuntimeError("a batch in waves runs with run()")
        self._part.launch()

    def wait(self):
        return self._part.wait()

    def arena_bytes(self):
        if self._part is not None:
            return self._part.arena_bytes()
        return max(sum(instance_bytes(n) for _, _, n in w) for w in self._waves)

    def info(self):
        """Plan shape of a one-part batch (vqc_batch_info)."""
        if self._part is None:
            raise RuntimeError("a batch in waves has one plan per wave")
        return self._part.info()

    def nwaves(self):
        return 1 if self._part is not None else len(self._waves)


class _Part:
    def __init__(self, instances):
        self.n_inst = len(instances)
        if not self.n_inst:
            raise ValueError("empty batch")
        inst = (_lib.Instance * self.n_inst)()
        self._ops = []
        self._terms = []
        pool_chunks, hpool_chunks = _lib.MatrixPool(), _lib.MatrixPool()
        self._nparams = []
        grad = 0
        for i, (circuit, H, n) in enumerate(instances):
            prog = Program(circuit, n=n, pool=pool_chunks)
            ops_begin = len(self._ops)
            self._ops.extend(prog.ops)
            terms, nt, _ = pack_terms(H, pool=hpool_chunks)
            t_begin = len(self._terms)
            self._terms.extend(terms[j] for j in range(nt))
            inst[i].nqubits = n
            inst[i].nops = len(prog.ops)
            inst[i].op_begin = ops_begin
            inst[i].nterms = nt
            inst[i].nparams = prog.nparams
            inst[i].term_begin = t_begin
            inst[i].grad_begin = grad
            self._nparams.append(prog.nparams)
            grad += prog.nparams
        self.grad_total = grad
        self._inst = inst
        self._pool = pool_chunks.array()
        self._hpool = hpool_chunks.array()
        self._op_arr = (_lib.Op * max(len(self._ops), 1))(*self._ops)
        self._term_arr = (_lib.Term * max(len(self._terms), 1))(*self._terms)
        self._h = None
        self.rebuild()
        self._energies = np.zeros(2 * self.n_inst)
        self._grads = np.zeros(max(grad, 1))

    def rebuild(self):
        """Plan the batch again from its packed records (with the tile shapes in
        force now) and capture a new graph; the autotuner times candidates so."""
        self.close()
        h = ctypes.c_void_p()
        _lib.call("vqc_batch_create", self.n_inst, self._inst, self._op_arr, len(self._ops), self._term_arr,
                  len(self._terms), len(self._pool), len(self._hpool), 0, ctypes.byref(h))
        self._h = h

    def update(self, circuits):
        """New parameters for the same circuit structures (matrix pool rebuilt;
        see repack)."""
        self._pool = repack(circuits, [self._inst[i].nqubits for i in range(self.n_inst)],
                            [self._inst[i].nops for i in range(self.n_inst)], self._ops, self._pool.size)

    def launch(self):
        _lib.call("vqc_batch_launch", self._h, _lib.dptr(self._pool), _lib.dptr(self._hpool))

    def wait(self):
        _lib.call("vqc_batch_wait", self._h, _lib.dptr(self._energies), _lib.dptr(self._grads))
        return self._results()

    def run(self):
        _lib.call("vqc_batch_run", self._h, _lib.dptr(self._pool), _lib.dptr(self._hpool), _lib.dptr(self._energies),
                  _lib.dptr(self._grads))
        return self._results()

    def _results(self):
        E = self._energies[0::2] + 1j * self._energies[1::2]
        grads, g = [], 0
        for p in self._nparams:
            grads.append(self._grads[g:g + p].copy())
            g += p
        return E, grads

    def arena_bytes(self):
        v = ctypes.c_uint64()
        _lib.call("vqc_batch_arena_bytes", self._h, ctypes.byref(v))
        return v.value

    def info(self):
        v = _lib.BatchInfo()
        _lib.call("vqc_batch_info", self._h, ctypes.byref(v))
        return {k: getattr(v, k) for k, _ in v._fields_ if k != "reserved"}

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.call("vqc_batch_destroy", h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
