"""Measurement and post-selection (src/measure.jl, src/postselect.jl) on the
device. Positions are 1-based as in the reference. The reference draws its
outcome with rand() (auxiliary/sampling.jl:2-18); here a uniform u in [0, 1)
comes from `rng` (a numpy Generator) or is given directly, and the outcome is 0
exactly when u < P(0)."""
import ctypes

import numpy as np

from . import _lib
from .state import DensityMatrix, StateVector

_default_rng = np.random.default_rng()


def _u(rng, u):
    if u is not None:
        return float(u)
    return float((rng or _default_rng).random())


def _key(state, key):
    n = state.nqubits()
    if not 1 <= key <= n:
        raise ValueError(f"qubit {key} out of range for {n} qubits")
    return key - 1


def probability_zero(state, key):
    """compute_probability_zero (measure.jl:46-58, 70-82)."""
    p = ctypes.c_double()
    _lib.call("vqc_prob", state.handle, _key(state, key), 0, ctypes.byref(p))
    return p.value


def measure_(state, key, auto_reset=True, rng=None, u=None):
    """measure!(state, key; auto_reset) (measure.jl:4-25): collapses in place and
    returns (outcome, probability)."""
    i = ctypes.c_int()
    p = ctypes.c_double()
    _lib.call("vqc_measure", state.handle, _key(state, key), _u(rng, u), int(auto_reset), ctypes.byref(i),
              ctypes.byref(p))
    return i.value, p.value


def measure(state, key, rng=None, u=None):
    """measure(state, key) (measure.jl:28-44): (collapsed state of n-1 qubits,
    outcome, probability); the input is unchanged."""
    if isinstance(state, DensityMatrix):
        raise TypeError("measure returns a reduced state vector; use measure_ on a density matrix")
    k = _key(state, key)
    p0 = probability_zero(state, key)
    i = 0 if _u(rng, u) < p0 else 1
    out = StateVector(state.nqubits() - 1)
    nrm = ctypes.c_double()
    _lib.call("vqc_sv_select", state.handle, k, i, out.handle, ctypes.byref(nrm))
    return out, i, nrm.value ** 2


def post_select_(state, key, value=0):
    """post_select!(state, key, value) (postselect.jl:3-8): projects in place and
    returns the norm it divided by."""
    s = ctypes.c_double()
    _lib.call("vqc_post_select", state.handle, _key(state, key), int(value), ctypes.byref(s))
    return s.value


def post_select(state, key, value=0, keep=False):
    """post_select(state, key, value; keep) (postselect.jl:11-15): (normalised
    selected state, its norm before normalisation). keep=False drops the qubit
    (qubit 1 takes its place, as the reference's swap! does); keep=True returns
    the projected n-qubit state."""
    if keep:
        s = state.copy()
        return s, post_select_(s, key, value)
    out = StateVector(state.nqubits() - 1)
    nrm = ctypes.c_double()
    _lib.call("vqc_sv_select", state.handle, _key(state, key), int(value), out.handle, ctypes.byref(nrm))
    return out, nrm.value
