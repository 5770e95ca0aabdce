"""Loss-side reductions around the path (src/auxiliary/distance.jl,
src/statevector.jl:59-65, src/densitymatrix.jl:60-73, src/ptrace.jl), computed
on the device. Positions are 1-based."""
import ctypes

import numpy as np

from . import _lib
from .state import DensityMatrix, StateVector


def dot(x, y):
    """<x|y> over the storage (a density matrix's vec(rho) for two DMs)."""
    out = np.zeros(2)
    _lib.call("vqc_state_dot", x.handle, y.handle, _lib.dptr(out))
    return complex(out[0], out[1])


def distance2(x, y):
    """|x|^2 + |y|^2 - 2 Re<x|y>, absolute value (distance.jl:2-8)."""
    return abs(dot(x, x).real + dot(y, y).real - 2 * dot(x, y).real)


def distance(x, y):
    return float(np.sqrt(distance2(x, y)))


def tr(rho):
    """tr(rho): the diagonal, as P(0) + P(1) of any qubit (densitymatrix.jl:61)."""
    p0, p1 = ctypes.c_double(), ctypes.c_double()
    _lib.call("vqc_prob", rho.handle, 0, 0, ctypes.byref(p0))
    _lib.call("vqc_prob", rho.handle, 0, 1, ctypes.byref(p1))
    return p0.value + p1.value


def fidelity(x, y):
    """|<x|y>|^2 for state vectors (statevector.jl:63); <y|rho|y> for a density
    matrix and a state vector (densitymatrix.jl:70-71); real(tr(sqrt(x) sqrt(y)))
    for two density matrices (densitymatrix.jl:65), from Hermitian
    eigendecompositions on the device (rocSOLVER)."""
    if isinstance(x, StateVector) and isinstance(y, StateVector):
        return abs(dot(x, y)) ** 2
    if isinstance(x, StateVector):
        x, y = y, x
    if isinstance(x, DensityMatrix) and isinstance(y, StateVector):
        out = np.zeros(2)
        _lib.call("vqc_dm_sandwich", x.handle, y.handle, _lib.dptr(out))
        return out[0]
    out = ctypes.c_double()
    _lib.call("vqc_dm_fidelity", x.handle, y.handle, ctypes.byref(out))
    return out.value


def partial_tr(state, sites):
    """partial_tr(state, sites) (ptrace.jl:4-30): a DensityMatrix of the kept
    qubits (ascending), or a scalar when every qubit is traced."""
    n = state.nqubits()
    sites = [int(q) for q in sites]
    if len(set(sites)) != len(sites):
        raise ValueError("duplicate sites not allowed.")
    if any(not 1 <= q <= n for q in sites):
        raise ValueError("site out of range.")
    if len(sites) == n:
        return tr(state) if isinstance(state, DensityMatrix) else dot(state, state)
    if not sites and isinstance(state, DensityMatrix):
        return state
    out = DensityMatrix(n - len(sites))
    _lib.call("vqc_partial_trace", state.handle, len(sites), _lib.i32ptr([q - 1 for q in sites]), out.handle)
    return out
