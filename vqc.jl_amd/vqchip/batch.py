"""Batched energy + adjoint gradient of many independent circuits (no
counterpart in the reference, which evaluates one circuit per call).

    b = Batch([(circuit_1, H_1, n_1), (circuit_2, H_2, n_2), ...])
    energies, grads = b.run()            # one persistent-kernel graph launch
    b.update([circuit_1', ...])          # same structure, new angles
"""
import ctypes

import numpy as np

from . import _lib
from .circuit import Program
from .operators import pack_terms


class Batch:
    def __init__(self, instances):
        self.n_inst = len(instances)
        if not self.n_inst:
            raise ValueError("empty batch")
        inst = (_lib.Instance * self.n_inst)()
        self._ops = []
        self._terms = []
        pool_chunks, hpool_chunks = [], []
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
        self._pool = np.ascontiguousarray(np.concatenate(pool_chunks)) if pool_chunks else np.zeros(1, complex)
        self._hpool = np.ascontiguousarray(np.concatenate(hpool_chunks)) if hpool_chunks else np.zeros(1, complex)
        ops = (_lib.Op * max(len(self._ops), 1))(*self._ops)
        terms = (_lib.Term * max(len(self._terms), 1))(*self._terms)
        h = ctypes.c_void_p()
        _lib.call("vqc_batch_create", self.n_inst, inst, ops, len(self._ops), terms, len(self._terms),
                  len(self._pool), len(self._hpool), 0, ctypes.byref(h))
        self._h = h
        self._energies = np.zeros(2 * self.n_inst)
        self._grads = np.zeros(max(grad, 1))

    def update(self, circuits):
        """New parameters for the same circuit structures (matrix pool rebuilt)."""
        chunks = []
        for c, n in zip(circuits, [self._inst[i].nqubits for i in range(self.n_inst)]):
            Program(c, n=n, pool=chunks)
        pool = np.ascontiguousarray(np.concatenate(chunks))
        if pool.size != self._pool.size:
            raise ValueError("circuit structure changed; create a new Batch")
        self._pool = pool

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

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.call("vqc_batch_destroy", h)
            except Exception:
                pass
            self._h = None
