"""Krylov ground state and time evolution on the device (src/utility/groundstate.jl,
which hands x -> H*x to KrylovKit's eigsolve / exponentiate).

Every vector of the Krylov basis lives in HBM; each step is one H*psi sweep
(vqc_sv_apply_terms, which also returns <v|Hv>), dots and axpys. Only the small
tridiagonal matrix is handled on the host. Lanczos runs with full
reorthogonalisation and restarts from the current Ritz vector.
"""
import numpy as np

from . import _lib
from .operators import LARGEST_SUPPORTED_NTERMS, pack_pauli, pack_terms
from .state import StateVector, rand_state


class _Hop:
    """H packed once; apply(x, out) sets out = H x and returns <x|out>."""

    def __init__(self, H):
        self.dense = H.largest_nterm() <= LARGEST_SUPPORTED_NTERMS
        if self.dense:
            self.terms, self.nt, self.pool = pack_terms(H)
        else:
            self.arr, self.nt = pack_pauli(H)

    def apply(self, x, out):
        if self.dense:
            e = np.zeros(2)
            _lib.call("vqc_sv_apply_terms", x.handle, out.handle, self.nt, self.terms, _lib.dptr(self.pool), _lib.dptr(e))
            return complex(e[0], e[1])
        _lib.call("vqc_sv_apply_pauli_sum", x.handle, out.handle, self.nt, self.arr)
        return x.dot(out)


def _axpy(y, a, x):
    a = complex(a)
    _lib.call("vqc_state_axpy", y.handle, a.real, a.imag, x.handle)


def _scale(x, a):
    a = complex(a)
    _lib.call("vqc_state_scale", x.handle, a.real, a.imag)


def _n_of(H):
    return max((max(t.positions) for t in H.terms), default=0)


def _lanczos(hop, v, m, tol, basis):
    """m Lanczos steps from the unit vector v (kept as basis[0]). Returns the
    tridiagonal (alphas, betas) and the number of basis vectors built."""
    alphas, betas = [], []
    _lib.call("vqc_state_copy", basis[0].handle, v.handle)
    for j in range(m):
        w = basis[j + 1]
        a = hop.apply(basis[j], w).real
        alphas.append(a)
        _axpy(w, -a, basis[j])
        if j > 0:
            _axpy(w, -betas[-1], basis[j - 1])
        for i in range(j + 1):  # full reorthogonalisation (classical Gram-Schmidt, one pass)
            c = basis[i].dot(w)
            if abs(c) > 0:
                _axpy(w, -c, basis[i])
        b = w.norm()
        if b <= tol * max(1.0, abs(a)):
            return np.array(alphas), np.array(betas), j + 1, b
        betas.append(b)
        _scale(w, 1.0 / b)
    return np.array(alphas), np.array(betas[: m - 1]), m, betas[m - 1] if len(betas) >= m else 0.0


def _tridiag(alphas, betas):
    k = len(alphas)
    T = np.diag(alphas).astype(float)
    for i in range(k - 1):
        T[i, i + 1] = T[i + 1, i] = betas[i]
    return T


def _combine(basis, coeffs, out):
    _scale(out, 0.0)
    for c, b in zip(coeffs, basis):
        _axpy(out, c, b)


def ground_state(H, n=None, tol=1e-12, krylovdim=30, maxiter=100, v0=None, rng=None):
    """(lowest eigenvalue, its eigenvector) of a Hermitian H (groundstate.jl:4-13).
    Starts from a random state like the reference (rand_state) unless v0 is given."""
    if not H.ishermitian():
        raise ValueError("input operator is not hermitian.")
    n = n or _n_of(H)
    hop = _Hop(H)
    v = v0.copy() if v0 is not None else rand_state(n, rng)
    _scale(v, 1.0 / v.norm())
    m = max(2, min(krylovdim, 1 << n))
    basis = [StateVector(n) for _ in range(m + 1)]
    theta = None
    for _ in range(maxiter):
        al, be, k, beta_last = _lanczos(hop, v, m, 1e-14, basis)
        evals, evecs = np.linalg.eigh(_tridiag(al, be[: k - 1]))
        theta, s = evals[0], evecs[:, 0]
        _combine(basis[:k], s, v)
        _scale(v, 1.0 / v.norm())
        resid = abs(beta_last * s[-1]) if k == m else 0.0
        if resid < tol * max(1.0, abs(theta)):
            return float(theta), v
    raise RuntimeError("eigsolve fails to converge.")


def time_evolution(H, t, v, tol=1e-12, krylovdim=30, maxsteps=10000):
    """exp(t H) v for a Hermitian H and a complex (or real) t (groundstate.jl:16-25);
    real-time evolution over tau is t = -1j * tau."""
    if not H.ishermitian():
        raise ValueError("input operator is not hermitian.")
    n = v.nqubits()
    if _n_of(H) > n:
        raise ValueError("number of qubits mismatch.")
    hop = _Hop(H)
    out = v.copy()
    nrm = out.norm()
    if nrm == 0:
        return out
    _scale(out, 1.0 / nrm)
    m = max(2, min(krylovdim, 1 << n))
    basis = [StateVector(n) for _ in range(m + 1)]
    t = complex(t)
    done, tau = 0.0, 1.0  # fractions of t
    for _ in range(maxsteps):
        if done >= 1.0:
            break
        step = min(tau, 1.0 - done)
        al, be, k, beta_last = _lanczos(hop, out, m, 1e-14, basis)
        lam, Q = np.linalg.eigh(_tridiag(al, be[: k - 1]))
        y = Q @ (np.exp(t * step * lam) * Q[0, :])  # exp(step t T) e1
        err = abs(beta_last * y[-1]) if k == m else 0.0
        if err > tol and step > 1e-12:
            tau = step / 2
            continue
        _combine(basis[:k], y, out)
        g = out.norm()  # exp(tH) changes the norm when t is not imaginary
        _scale(out, 1.0 / g)
        nrm *= g
        done += step
        tau = min(1.0, step * 2)
    else:
        raise RuntimeError("exponentiate fails to converge.")
    _scale(out, nrm)
    return out
