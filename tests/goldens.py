"""Loads the fixtures of tests/golden/ (numpy .npz written by
tests/golden/make_golden.py, read with allow_pickle=False) and turns their
integer tables back into the C ABI's structures."""
import os

import numpy as np

from vqchip import _lib

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def ops_from_table(t):
    """(nops, 17) int64 rows [kind, nq, pos0..4, nparams, slot0..2, reserved, mat, aux, gen0..2] -> vqc_op_t[]."""
    arr = (_lib.Op * max(len(t), 1))()
    for i, r in enumerate(t.tolist()):
        o = arr[i]
        o.kind, o.nq = r[0], r[1]
        for j in range(5):
            o.pos[j] = r[2 + j]
        o.nparams = r[7]
        for j in range(3):
            o.slot[j] = r[8 + j]
        o.reserved, o.mat, o.aux = r[11], r[12], r[13]
        for j in range(3):
            o.gen[j] = r[14 + j]
    return arr


def terms_from_table(t):
    """(nterms, 7) int64 rows [nq, pos0..4, mat] -> vqc_term_t[]."""
    arr = (_lib.Term * max(len(t), 1))()
    for i, r in enumerate(t.tolist()):
        arr[i].nq = r[0]
        for j in range(5):
            arr[i].pos[j] = r[1 + j]
        arr[i].mat = r[6]
    return arr


def paulis_from(xz, coef):
    arr = (_lib.Pauli * max(len(xz), 1))()
    for i in range(len(xz)):
        arr[i].x, arr[i].z = int(xz[i, 0]), int(xz[i, 1])
        arr[i].re, arr[i].im = float(coef[i].real), float(coef[i].imag)
    return arr


def dense_cases(d):
    """[(k, pos tuple, U, expected output)] of dense_gates_n7.npz."""
    out = []
    for key in sorted(k for k in d if k.endswith("_pos")):
        base = key[:-4]
        out.append((len(d[key]), tuple(int(p) for p in d[key]), d[base + "_U"], d[base + "_out"]))
    return out
