"""ctypes wrapper of liboracle.so, the CPU restatement of VQC.jl's kernels.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, to compare the device results against. The product
path never calls it.
Every function here forwards to vqc_oracle.cpp, whose routines cite the
reference file:line they restate. Pinning (see tests/test_oracle.py): dense
gates, expectations, Pauli strings and DM maps against numpy brute force on the
full 2^n operator; the adjoint gradient against exact parameter shift and
central finite differences; layout against the reference's known-answer tests.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# VQC_ORACLE_LIB: another build of the same sources (the sanitizer build)
LIB = os.environ.get("VQC_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        _lib.orc_get_threads.restype = ctypes.c_int
    return _lib


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def _c(a):
    assert a.dtype == np.complex128 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def _ipos(pos):
    return (ctypes.c_int * len(pos))(*[int(p) for p in pos])


def _mat(M):
    return np.ascontiguousarray(np.asarray(M, dtype=np.complex128).reshape(-1, order="F"))


def _rc(rc):
    if rc != 0:
        raise ValueError("oracle rejected the arguments")


def apply_dense(v, k, pos0, U):
    """v (complex128, 2^nbits) updated in place; pos0 ascending 0-based."""
    nb = int(np.log2(v.size))
    m = _mat(U)
    _rc(lib().orc_apply_dense(nb, _c(v), k, _ipos(pos0), _c(m)))
    return v


def apply_named(v, kind, pos0_given, phi=0.0):
    nb = int(np.log2(v.size))
    _rc(lib().orc_apply_named(nb, _c(v), int(kind), _ipos(pos0_given), ctypes.c_double(phi)))
    return v


def accum_dense(v, vout, k, pos0, M):
    nb = int(np.log2(v.size))
    m = _mat(M)
    _rc(lib().orc_accum_dense(nb, _c(v), _c(vout), k, _ipos(pos0), _c(m)))
    return vout


def expec_dense(v, k, pos0, M, vc=None):
    nb = int(np.log2(v.size))
    m = _mat(M)
    out = np.zeros(2)
    _rc(lib().orc_expec_dense(nb, _c(v), _c(vc) if vc is not None else None, k, _ipos(pos0), _c(m),
                              out.ctypes.data_as(_D)))
    return complex(out[0], out[1])


def dot(a, b):
    out = np.zeros(2)
    lib().orc_dot(int(np.log2(a.size)), _c(a), _c(b), out.ctypes.data_as(_D))
    return complex(out[0], out[1])


def expec_pauli(v, paulis, nterms, vc=None):
    """Per-term <vc|P_t|v> and their sum, the reference's per-term route."""
    nb = int(np.log2(v.size))
    out = np.zeros(2 * max(nterms, 1))
    s = np.zeros(2)
    lib().orc_expec_pauli(nb, _c(v), _c(vc) if vc is not None else None, nterms, paulis,
                          out.ctypes.data_as(_D), s.ctypes.data_as(_D))
    return out[0:2 * nterms:2] + 1j * out[1:2 * nterms:2], complex(s[0], s[1])


def apply_pauli_sum(v, paulis, nterms):
    out = np.zeros_like(v)
    lib().orc_apply_pauli_sum(int(np.log2(v.size)), _c(v), _c(out), nterms, paulis)
    return out


def _qterm_args(qterms):
    """(terms array, positions, pool) of a list of (coef, [0-based sites], [2x2 factors])."""
    import vqchip._lib as vl
    arr = (vl.QTerm * max(len(qterms), 1))()
    pos, pool = [], []
    for i, (c, sites, mats) in enumerate(qterms):
        arr[i].nsites = len(sites)
        arr[i].pos = len(pos)
        arr[i].mat = sum(len(m) for m in pool)
        arr[i].re, arr[i].im = complex(c).real, complex(c).imag
        pos.extend(int(s) for s in sites)
        for m in mats:
            pool.append(_mat(m))
    pool = np.ascontiguousarray(np.concatenate(pool) if pool else np.zeros(1, dtype=np.complex128))
    return arr, (ctypes.c_int * max(len(pos), 1))(*pos), pool


def apply_qterms(v, qterms):
    """sum_t T_t v (apply_qterms.jl:20-25 over _apply_qterm_util!, :47-59)."""
    v = np.ascontiguousarray(v, dtype=np.complex128)
    out = np.zeros_like(v)
    arr, pos, pool = _qterm_args(qterms)
    nbits = int(round(np.log2(v.size)))
    _rc(lib().orc_apply_qterms(nbits, _c(v), _c(out), len(qterms), arr, pos, _c(pool)))
    return out


def expec_qterms(v, qterms, vc=None):
    """(per-term <vc|T_t|v>, their sum) (expecs.jl:17,40,50,72)."""
    v = np.ascontiguousarray(v, dtype=np.complex128)
    arr, pos, pool = _qterm_args(qterms)
    out = np.zeros(2 * max(len(qterms), 1))
    s2 = np.zeros(2)
    nbits = int(round(np.log2(v.size)))
    vcp = _c(np.ascontiguousarray(vc, dtype=np.complex128)) if vc is not None else None
    _rc(lib().orc_expec_qterms(nbits, _c(v), vcp, len(qterms), arr, pos, _c(pool), out.ctypes.data_as(_D), s2.ctypes.data_as(_D)))
    return out[0:2 * len(qterms):2] + 1j * out[1:2 * len(qterms):2], complex(s2[0], s2[1])


def dm_expec_qterms(rho, n, qterms):
    """(per-term tr(T_t rho), their sum) (expecs_dm.jl:10-26)."""
    rho = np.ascontiguousarray(np.asarray(rho, dtype=np.complex128).reshape(-1, order="F"))
    arr, pos, pool = _qterm_args(qterms)
    out = np.zeros(2 * max(len(qterms), 1))
    s2 = np.zeros(2)
    _rc(lib().orc_dm_expec_qterms(n, _c(rho), len(qterms), arr, pos, _c(pool), out.ctypes.data_as(_D), s2.ctypes.data_as(_D)))
    return out[0:2 * len(qterms):2] + 1j * out[1:2 * len(qterms):2], complex(s2[0], s2[1])


def dm_apply_gate(rho, n, k, pos0, U):
    m = _mat(U)
    _rc(lib().orc_dm_apply_gate(n, _c(rho), k, _ipos(pos0), _c(m)))
    return rho


def dm_apply_superop(rho, n, N, pos0, S, mode=0):
    m = _mat(S)
    _rc(lib().orc_dm_apply_superop(n, _c(rho), N, _ipos(pos0), _c(m), mode))
    return rho


def dm_expec(rho, n, k, pos0, M):
    m = _mat(M)
    out = np.zeros(2)
    _rc(lib().orc_dm_expec(n, _c(rho), k, _ipos(pos0), _c(m), out.ctypes.data_as(_D)))
    return complex(out[0], out[1])


def prob_zero(v, n, key0, is_dm=False):
    """compute_probability_zero (measure.jl:46-58, 70-82); v is vec(rho) for a DM."""
    f = lib().orc_prob_zero
    f.restype = ctypes.c_double
    return f(n, _c(v), int(is_dm), key0)


def measure_inplace(v, n, key0, u, auto_reset=True, is_dm=False):
    """measure! (measure.jl:4-25) with rand() = u; returns (outcome, prob)."""
    i = ctypes.c_int()
    p = ctypes.c_double()
    _rc(lib().orc_measure(n, _c(v), int(is_dm), key0, ctypes.c_double(u), int(auto_reset), ctypes.byref(i),
                          ctypes.byref(p)))
    return i.value, p.value


def post_select_inplace(v, n, key0, value):
    """post_select! (postselect.jl:3-8); returns the norm it divided by."""
    s = ctypes.c_double()
    _rc(lib().orc_post_select(n, _c(v), key0, value, ctypes.byref(s)))
    return s.value


def select(v, n, key0, value):
    """The unnormalised row `value` of the measured qubit (postselect.jl:47-53)."""
    out = np.zeros(1 << (n - 1), dtype=np.complex128)
    _rc(lib().orc_select(n, _c(v), key0, value, _c(out)))
    return out


def apply_circuit(v, n, is_dm, ops, nops, pool):
    _rc(lib().orc_apply_circuit(n, int(is_dm), _c(v), nops, ops, _c(pool)))
    return v


def circuit_vjp(x, n, is_dm, ops, nops, pool, nparams, delta_y):
    """(grad, delta_x, y) of the pullback of y = circuit * x for the cotangent
    delta_y (@adjoint *(circuit, x) with back_propagate, circuitdiff.jl:2-8, 23-78)."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    d = np.ascontiguousarray(delta_y, dtype=np.complex128)
    g = np.zeros(max(nparams, 1))
    dx = np.zeros_like(x)
    y = np.zeros_like(x)
    _rc(lib().orc_circuit_vjp(n, int(is_dm), _c(x), nops, ops, _c(pool), nparams, _c(d), g.ctypes.data_as(_D), _c(dx),
                              _c(y)))
    return g[:nparams], dx, y


def energy_grad(state0, n, is_dm, ops, nops, pool, terms, nterms, hpool, nparams):
    """(E, grad, final state) of the restated adjoint (circuitdiff.jl / expecdiff.jl)."""
    e2 = np.zeros(2)
    g = np.zeros(max(nparams, 1))
    out = np.zeros_like(state0)
    fn = lib().orc_dm_energy_grad if is_dm else lib().orc_sv_energy_grad
    _rc(fn(n, _c(np.ascontiguousarray(state0)), nops, ops, _c(pool), nterms, terms, _c(hpool), nparams,
           e2.ctypes.data_as(_D), g.ctypes.data_as(_D), _c(out)))
    return complex(e2[0], e2[1]), g[:nparams], out


def _julia_hermitian_sqrt(A):
    """sqrt of a Hermitian matrix as Julia's LinearAlgebra takes it: the
    eigendecomposition, roots sqrt(max(l, 0)) when every eigenvalue is at least
    -max|l| * eps * N (round-off), else sqrt(complex(l))."""
    A = 0.5 * (A + A.conj().T)
    l, V = np.linalg.eigh(A)
    l0 = -np.max(np.abs(l)) * np.finfo(float).eps * A.shape[0]
    r = np.sqrt(np.maximum(l, 0)) if np.all(l >= l0) else np.sqrt(l.astype(complex))
    return (V * r) @ V.conj().T


def fidelity_dm(x, y):
    """fidelity(x::DensityMatrix, y::DensityMatrix) = real(tr(sqrt(x) * sqrt(y)))
    (densitymatrix.jl:65) for two 2^n x 2^n matrices."""
    return float(np.real(np.trace(_julia_hermitian_sqrt(np.asarray(x)) @ _julia_hermitian_sqrt(np.asarray(y)))))
