"""apply!, expectation and the adjoint gradient, over libvqc_hip.

    apply_(gate | circuit | channel, state)      apply!(x, state)        generic.jl:6-31, apply_qmaps.jl:8-16
    circuit * state                              Base.:*                  generic.jl:33
    expectation(H, state)                        expectation(m, state)    expecs.jl:12-42, expecs_dm.jl:7-27
    expectation(bra, H, state)                   expectation(c, m, s)     expecs.jl:45-74
    H * state                                    (m::QubitsOperator)(v)   apply_qterms.jl:13-28
    energy_and_gradient(H, circuit, state0)      Zygote.gradient(c -> real(expectation(H, c * state0)), c)
                                                 through circuitdiff.jl:2-8,23-78 and expecdiff.jl:30-68
"""
import ctypes

import numpy as np

from . import _lib
from .circuit import Program, QCircuit
from .gates import Gate, QuantumMap
from .operators import (LARGEST_SUPPORTED_NTERMS, QubitsOperator, QubitsTerm, pack_pauli, pack_qterms, pack_terms,
                        split_long)
from .state import DensityMatrix, StateVector


def _mat_ptr(m):
    m = np.ascontiguousarray(m, dtype=np.complex128).reshape(-1, order="F")
    return m, _lib.dptr(np.ascontiguousarray(m))


def _check_pos_range(pos, n):
    if len(pos) > 5:
        raise ValueError("only implement 5-qubit gates and less currently.")  # generic.jl:37
    if pos[0] < 1 or pos[-1] > n:
        raise ValueError(f"qubit positions {pos} outside 1..{n}")  # generic.jl:7 @assert


def apply_(x, state, named=True):
    """apply!(x, state): in place; returns the state."""
    if isinstance(x, (list, QCircuit)):
        if len(x) == 0:
            return state
        if not state.is_dm:
            first = next((i for i, g in enumerate(x) if isinstance(g, QuantumMap)), None)
            if first is not None:
                # apply!(circuit, state) applies element by element (generic.jl:26-31), and a
                # channel on a state vector continues on DensityMatrix(state) (apply_qmaps.jl:19):
                # the gates before the first channel run on the state vector, the rest on rho
                if first:
                    apply_(QCircuit(x[:first]), state)
                return apply_(QCircuit(x[first:]), DensityMatrix(state))
        prog = Program(x, n=state.nqubits(), dm=bool(state.is_dm))
        ops = prog.op_array()
        pool = prog.pool()
        _lib.call("vqc_apply_circuit", state.handle, len(prog.ops), ops, _lib.dptr(pool))
        return state
    if isinstance(x, QuantumMap):
        if not state.is_dm:
            state = DensityMatrix(state)  # apply_qmaps.jl:19
        pos = x.ordered_positions()
        _check_pos_range(pos, state.nqubits())
        S = np.ascontiguousarray(x.ordered_supermat().reshape(-1, order="F"))
        p = _lib.i32ptr([q - 1 for q in pos])
        _lib.call("vqc_dm_apply_superop", state.handle, len(pos), p, _lib.dptr(S), 0)
        return state
    if not isinstance(x, Gate):
        raise TypeError(f"cannot apply {type(x)}")
    pos = x.ordered_positions()
    _check_pos_range(pos, state.nqubits())
    if state.is_dm:
        U = np.ascontiguousarray(x.ordered_mat().reshape(-1, order="F"))
        _lib.call("vqc_dm_apply_gate", state.handle, len(pos), _lib.i32ptr([q - 1 for q in pos]), _lib.dptr(U))
        return state
    if named and x.named_kind is not None:
        given = _lib.i32ptr([q - 1 for q in x.positions])
        phi = x.params[0] if x.params else 0.0
        _lib.call("vqc_sv_apply_named", state.handle, x.named_kind, given, float(phi))
        return state
    U = np.ascontiguousarray(x.ordered_mat().reshape(-1, order="F"))
    _lib.call("vqc_sv_apply_dense", state.handle, len(pos), _lib.i32ptr([q - 1 for q in pos]), _lib.dptr(U))
    return state


def _mul(circuit, state):
    return apply_(circuit, state.copy())


QCircuit.__mul__ = _mul
Gate.__mul__ = _mul


def _as_operator(m):
    return QubitsOperator([m]) if isinstance(m, QubitsTerm) else m


def _expec_keys(m, state, bra):
    """sum over the merged keys (<= 5 sites) of <bra|M_key|state> or tr(M_key rho)."""
    total = 0j
    for key, M in m.data().items():
        Mf = np.ascontiguousarray(M.reshape(-1, order="F"))
        out = np.zeros(2)
        p = _lib.i32ptr([q - 1 for q in key])
        if state.is_dm:
            _lib.call("vqc_dm_expec_dense", state.handle, len(key), p, _lib.dptr(Mf), _lib.dptr(out))
        else:
            _lib.call("vqc_sv_expec_dense", state.handle, bra.handle if bra else None, len(key), p, _lib.dptr(Mf),
                      _lib.dptr(out))
        total += complex(out[0], out[1])
    return total


def _expec_products(terms, state, bra):
    """sum_t <bra|T_t|state> (or tr(T_t rho)) for product terms of any length."""
    if not terms:
        return 0j
    arr, nt, pos, pool = pack_qterms(terms)
    s2 = np.zeros(2)
    if state.is_dm:
        _lib.call("vqc_dm_expec_qterms", state.handle, nt, arr, pos, _lib.dptr(pool), None, _lib.dptr(s2))
    else:
        _lib.call("vqc_sv_expec_qterms", state.handle, bra.handle if bra else None, nt, arr, pos, _lib.dptr(pool), None,
                  _lib.dptr(s2))
    return complex(s2[0], s2[1])


def expectation(*args):
    """expectation(H, state) or expectation(bra, H, state); returns a complex.

    Keys of at most 5 sites take one dense sweep each (expecs.jl:22-28; on a
    density matrix tr(M rho) over the diagonal band, expecs_dm.jl:18-23). Longer
    terms take the reference's per-term route, dot(psi, m * psi) (expecs.jl:40,72)
    or tr(m * rho) (expecs_dm.jl:25): Pauli strings through the bitmask kernels,
    other products through the product-term kernels."""
    if len(args) == 2:
        m, state, bra = args[0], args[1], None
    else:
        bra, m, state = args
    if isinstance(m, Gate):
        pos = m.ordered_positions()
        M = np.ascontiguousarray(m.ordered_mat().reshape(-1, order="F"))
        out = np.zeros(2)
        _lib.call("vqc_sv_expec_dense", state.handle, bra.handle if bra else None, len(pos),
                  _lib.i32ptr([q - 1 for q in pos]), _lib.dptr(M), _lib.dptr(out))
        return complex(out[0], out[1])
    if isinstance(m, QubitsTerm) and not m.positions:  # isempty(m) (expecs.jl:13, expecs_dm.jl:8)
        if state.is_dm:
            return state.tr()
        return (bra or state).dot(state)
    m = _as_operator(m)
    if state.is_dm and bra is not None:
        raise NotImplementedError("bra-ket expectation on density matrices")
    if m.largest_nterm() <= LARGEST_SUPPORTED_NTERMS:
        return _expec_keys(m, state, bra)
    short, pauli, other = split_long(m)
    total = _expec_keys(short, state, bra) if short.terms else 0j
    if state.is_dm:
        return total + _expec_products(pauli.terms + other, state, None)
    if pauli.terms:
        arr, nt = pack_pauli(pauli)
        s2 = np.zeros(2)
        _lib.call("vqc_sv_expec_pauli", state.handle, bra.handle if bra else None, nt, arr, None, _lib.dptr(s2))
        total += complex(s2[0], s2[1])
    return total + _expec_products(other, state, bra)


def apply_operator(m, state):
    """H * state as a new state ((m::QubitsOperator)(v), apply_qterms.jl:13-28). On a
    density matrix the keys act on the row bits of vec(rho) (apply_qterms_dm.jl:11-27).
    Terms of more than 5 sites take the per-term route (_apply_qterm_util!,
    apply_qterms.jl:47-59)."""
    m = _as_operator(m)
    out = DensityMatrix(state.nqubits()) if state.is_dm else StateVector(state.nqubits())
    if m.largest_nterm() <= LARGEST_SUPPORTED_NTERMS:
        terms, nt, pool = pack_terms(m)
        _lib.call("vqc_sv_apply_terms", state.handle, out.handle, nt, terms, _lib.dptr(pool), None)
        return out
    short, pauli, other = split_long(m)
    if not short.terms and not other and not state.is_dm:
        arr, nt = pack_pauli(pauli)
        _lib.call("vqc_sv_apply_pauli_sum", state.handle, out.handle, nt, arr)
        return out
    accumulate = 0
    if short.terms:
        terms, nt, pool = pack_terms(short)
        _lib.call("vqc_sv_apply_terms", state.handle, out.handle, nt, terms, _lib.dptr(pool), None)
        accumulate = 1
    arr, nt, pos, pool = pack_qterms(pauli.terms + other)
    _lib.call("vqc_sv_apply_qterms", state.handle, out.handle, nt, arr, pos, _lib.dptr(pool), accumulate)
    return out


def _qop_mul(m, state):
    return apply_operator(m, state)


QubitsOperator.__mul__ = _qop_mul
QubitsTerm.__mul__ = _qop_mul


def expectation_vjp(H, state, z=1.0):
    """The pullback of expectation(H, state) for the cotangent z of its value
    (@adjoint expectation, expecdiff.jl:4-68): (conj(z) H + z H^dagger) state for a
    state vector; z H^dagger as vec, i.e. H^dagger applied to the identity on the
    row bits, for a density matrix. A new state of the same kind."""
    H = _as_operator(H)
    z = complex(z)
    n = state.nqubits()
    if state.is_dm:
        ident = DensityMatrix(n)
        _lib.call("vqc_dm_identity", ident.handle)
        out = apply_operator(H.adjoint(), ident)
        out.scale_(z)
        return out
    out = apply_operator(H, state)
    out.scale_(np.conj(z))
    hd = apply_operator(H.adjoint(), state)
    _lib.call("vqc_state_axpy", out.handle, z.real, z.imag, hd.handle)
    return out


def qubit_encoding_vjp(thetas, delta):
    """The pullback of qubit_encoding(thetas) for a cotangent delta on the encoded
    state (@adjoint qubit_encoding, circuitdiff.jl:10-17): the encoding is
    Ry(pi theta_i) on every qubit of |0..0>, so the gradient is pi times that
    circuit's parameter gradient."""
    from .gates import RyGate
    thetas = [float(t) for t in thetas]
    c = QCircuit(RyGate(i + 1, np.pi * t) for i, t in enumerate(thetas))
    g, _ = circuit_vjp(c, StateVector(len(thetas)), delta)
    return g * np.pi


def circuit_vjp(circuit, x, delta_y, return_state=False, checkpoint=False):
    """The pullback of y = circuit * x (@adjoint *(circuit, x), circuitdiff.jl:2-8,
    with back_propagate, :23-78) for the cotangent delta_y on y (a state of x's
    kind): (grad, delta_x[, y]). grad holds d loss / d theta for the circuit's
    active parameters in forward order; delta_x is circuit^dagger delta_y, the
    cotangent on x."""
    dm = bool(x.is_dm)
    if bool(delta_y.is_dm) != dm or delta_y.nqubits() != x.nqubits():
        raise ValueError("delta_y must have the shape of x")
    prog = Program(circuit, n=x.nqubits(), dm=dm)
    pool = prog.pool()
    grad = np.zeros(max(prog.nparams, 1))
    new = (lambda: DensityMatrix(x.nqubits())) if dm else (lambda: StateVector(x.nqubits()))
    dx = new()
    y = new() if return_state else None
    _lib.call("vqc_circuit_vjp", x.handle, len(prog.ops), prog.op_array(), _lib.dptr(pool), prog.nparams, delta_y.handle,
              _lib.dptr(grad), dx.handle, y.handle if y else None, _lib.VQC_FLAG_CHECKPOINT if checkpoint else 0)
    res = (grad[: prog.nparams], dx)
    return res + (y,) if return_state else res


def energy_and_gradient(H, circuit, state0, fused=True, return_state=False, checkpoint=False):
    """E = expectation(H, circuit * state0) and dRe(E)/dtheta for every active
    parameter of the circuit, in forward order (circuitdiff.jl:61-75).

    A circuit holding a channel, applied to a state vector, runs on
    DensityMatrix(state0) (apply_qmaps.jl:19). Hamiltonians of keys of at most 5
    sites run in one fused call (forward, the seed (H + H^dagger)psi or H^dagger
    with the energy, reverse). With longer terms the reference's own chain is
    taken: y = circuit * state0, E = expectation(H, y), the expectation pullback
    (expecdiff.jl:37-56: the per-term route of _apply_qterm_util!) and the circuit
    pullback of that cotangent (circuitdiff.jl:2-8).

    checkpoint=True stores forward states every ceil(sqrt(#ops)) ops for the
    reverse sweep to restore (VQC_FLAG_CHECKPOINT); density matrices also
    checkpoint around channels that are singular or whose inverses would
    amplify the uncompute error beyond 1e3, with or without it."""
    H = _as_operator(H)
    if not state0.is_dm and any(isinstance(g, QuantumMap) for g in circuit):
        state0 = DensityMatrix(state0)
    dm = bool(state0.is_dm)
    if H.largest_nterm() > LARGEST_SUPPORTED_NTERMS:
        y = apply_(circuit, state0.copy())
        E = expectation(H, y)
        delta = expectation_vjp(H, y, 1.0)
        grad, _ = circuit_vjp(circuit, state0, delta, checkpoint=checkpoint)
        res = (E, grad)
        return res + (y,) if return_state else res
    prog = Program(circuit, n=state0.nqubits(), dm=dm)
    pool = prog.pool()
    terms, nt, hpool = pack_terms(H)
    grad = np.zeros(max(prog.nparams, 1))
    e2 = np.zeros(2)
    out = (DensityMatrix(state0.nqubits()) if dm else StateVector(state0.nqubits())) if return_state else None
    fn = "vqc_dm_energy_grad" if dm else "vqc_sv_energy_grad"
    _lib.call(fn, state0.handle, len(prog.ops), prog.op_array(), _lib.dptr(pool), nt, terms, _lib.dptr(hpool),
              prog.nparams, _lib.dptr(e2), _lib.dptr(grad), out.handle if out else None,
              (0 if fused else _lib.VQC_FLAG_UNFUSED) | (_lib.VQC_FLAG_CHECKPOINT if checkpoint else 0))
    res = (complex(e2[0], e2[1]), grad[: prog.nparams])
    return res + (out,) if return_state else res
