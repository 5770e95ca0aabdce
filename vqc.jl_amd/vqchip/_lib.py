"""ctypes binding of libvqc_hip.so (include/vqc_hip.h).

The library is the only compute path: if it cannot be loaded, or no gfx950
device is present, every call raises. Nothing here falls back to the CPU.
"""
import ctypes
import os

import numpy as np

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# VQC_LIB_PATH: another build of the same sources (the host sanitizer build the
# CPU tests load); the default is the in-tree library
LIB_PATH = os.environ.get("VQC_LIB_PATH") or os.path.join(_PKG_DIR, "libvqc_hip.so")

VQC_OK = 0
VQC_FLAG_UNFUSED = 1
VQC_FLAG_CHECKPOINT = 2

GATE_X, GATE_PHASE, GATE_CZ, GATE_CNOT, GATE_CPHASE = 1, 2, 3, 4, 5
GATE_SWAP, GATE_ISWAP, GATE_ISWAPDAG, GATE_FREDKIN, GATE_CCPHASE = 6, 7, 8, 9, 10
OP_UNITARY, OP_CHANNEL = 0, 1


class VQCError(RuntimeError):
    """A libvqc_hip call failed (the reference raises a Julia exception here)."""

    def __init__(self, code, msg):
        super().__init__(f"libvqc_hip error {code}: {msg}")
        self.code = code


class Op(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("nq", ctypes.c_int32),
        ("pos", ctypes.c_int32 * 5),
        ("nparams", ctypes.c_int32),
        ("slot", ctypes.c_int32 * 3),
        ("reserved", ctypes.c_int32),
        ("mat", ctypes.c_int64),
        ("aux", ctypes.c_int64),
        ("gen", ctypes.c_int64 * 3),
    ]


class Term(ctypes.Structure):
    _fields_ = [("nq", ctypes.c_int32), ("pos", ctypes.c_int32 * 5), ("mat", ctypes.c_int64)]


class MatrixPool(list):
    """Flattened complex matrices shared by several programs (the pool that
    vqc_op_t.mat / vqc_term_t.mat offsets index), with its running length kept
    so that packing one more program does not re-count the chunks before it."""

    def __init__(self):
        super().__init__()
        self.size = 0

    def add(self, flat):
        off = self.size
        self.append(flat)
        self.size += len(flat)
        return off

    def array(self):
        if not self:
            return np.zeros(1, dtype=np.complex128)
        return np.ascontiguousarray(np.concatenate(self))


class Pauli(ctypes.Structure):
    _fields_ = [("x", ctypes.c_uint64), ("z", ctypes.c_uint64), ("re", ctypes.c_double), ("im", ctypes.c_double)]


class QTerm(ctypes.Structure):
    _fields_ = [("nsites", ctypes.c_int32), ("reserved", ctypes.c_int32), ("pos", ctypes.c_int64), ("mat", ctypes.c_int64),
                ("re", ctypes.c_double), ("im", ctypes.c_double)]


class Instance(ctypes.Structure):
    _fields_ = [
        ("nqubits", ctypes.c_int32),
        ("nops", ctypes.c_int32),
        ("op_begin", ctypes.c_int64),
        ("nterms", ctypes.c_int32),
        ("nparams", ctypes.c_int32),
        ("term_begin", ctypes.c_int64),
        ("grad_begin", ctypes.c_int64),
    ]


class BatchInfo(ctypes.Structure):
    _fields_ = [("tiled_instances", ctypes.c_int32), ("small_instances", ctypes.c_int32), ("phases", ctypes.c_int64),
                ("items", ctypes.c_int64), ("grid", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("hbm_bytes", ctypes.c_uint64), ("arena_bytes", ctypes.c_uint64)]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("hbm_budget", ctypes.c_uint64),
        ("lanes", ctypes.c_int32),
        ("max_batch", ctypes.c_int32),
        ("grid_frac", ctypes.c_double),
        ("cost_ratio", ctypes.c_double),
        ("poll_us", ctypes.c_int32),
        ("solo_qubits", ctypes.c_int32),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int64) for k in ("completed", "batches", "batched_requests", "pending", "running")] + [
        (k, ctypes.c_uint64) for k in ("budget_bytes", "reserved_bytes", "peak_live_bytes")] + [
        ("batch_hist", ctypes.c_int64 * 8)]


class Tuning(ctypes.Structure):
    _fields_ = [
        ("tile_fwd_bits", ctypes.c_int32),
        ("tile_rev_bits", ctypes.c_int32),
        ("tile_small_bits", ctypes.c_int32),
        ("tile_low_bits", ctypes.c_int32),
        ("dense_groups", ctypes.c_int32 * 3),
        ("dense_nontemporal", ctypes.c_int32 * 3),
    ]


class TileEntry(ctypes.Structure):
    _fields_ = [("fwd", ctypes.c_int32), ("rev", ctypes.c_int32), ("low", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("phases_fwd", ctypes.c_int32), ("phases_ham", ctypes.c_int32), ("phases_rev", ctypes.c_int32),
                ("mops", ctypes.c_int32 * 8), ("reused", ctypes.c_int32)]


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.POINTER(ctypes.c_int32)
_U64 = ctypes.c_uint64

# name -> argtypes (every function returns int)
SIGNATURES = {
    "vqc_version": [],
    "vqc_device_info": [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint64)],
    "vqc_sync": [],
    "vqc_get_stream": [ctypes.POINTER(_P)],
    "vqc_thread_release": [],
    "vqc_set_tuning": [ctypes.POINTER(Tuning)],
    "vqc_get_tuning": [ctypes.POINTER(Tuning)],
    "vqc_set_tile_table": [_I, _I32, ctypes.POINTER(TileEntry)],
    "vqc_get_tile_table": [ctypes.POINTER(TileEntry)],
    "vqc_sv_create": [_I, ctypes.POINTER(_P)],
    "vqc_dm_create": [_I, ctypes.POINTER(_P)],
    "vqc_state_wrap": [_I, _I, _P, ctypes.POINTER(_P)],
    "vqc_state_destroy": [_P],
    "vqc_state_nbits": [_P, ctypes.POINTER(ctypes.c_int)],
    "vqc_state_device_ptr": [_P, ctypes.POINTER(_P)],
    "vqc_state_upload": [_P, _D, _U64],
    "vqc_state_download": [_P, _D, _U64],
    "vqc_state_reset": [_P],
    "vqc_state_copy": [_P, _P],
    "vqc_state_scale": [_P, ctypes.c_double, ctypes.c_double],
    "vqc_state_axpy": [_P, ctypes.c_double, ctypes.c_double, _P],
    "vqc_state_dot": [_P, _P, _D],
    "vqc_dm_from_sv": [_P, _P],
    "vqc_dm_identity": [_P],
    "vqc_sv_apply_dense": [_P, _I, _I32, _D],
    "vqc_sv_apply_named": [_P, _I, _I32, ctypes.c_double],
    "vqc_sv_accum_dense": [_P, _P, _I, _I32, _D],
    "vqc_sv_expec_dense": [_P, _P, _I, _I32, _D, _D],
    "vqc_sv_expec_pauli": [_P, _P, _I, ctypes.POINTER(Pauli), _D, _D],
    "vqc_sv_expec_pauli_noalloc": [_P, _P, _I, ctypes.POINTER(Pauli), _D, _D],
    "vqc_sv_apply_pauli_sum": [_P, _P, _I, ctypes.POINTER(Pauli)],
    "vqc_sv_apply_terms": [_P, _P, _I, ctypes.POINTER(Term), _D, _D],
    "vqc_sv_apply_qterms": [_P, _P, _I, ctypes.POINTER(QTerm), _I32, _D, _I],
    "vqc_sv_expec_qterms": [_P, _P, _I, ctypes.POINTER(QTerm), _I32, _D, _D, _D],
    "vqc_dm_expec_qterms": [_P, _I, ctypes.POINTER(QTerm), _I32, _D, _D, _D],
    "vqc_dm_apply_gate": [_P, _I, _I32, _D],
    "vqc_dm_apply_superop": [_P, _I, _I32, _D, _I],
    "vqc_dm_expec_dense": [_P, _I, _I32, _D, _D],
    "vqc_prob": [_P, _I, _I, _D],
    "vqc_measure": [_P, _I, ctypes.c_double, _I, ctypes.POINTER(ctypes.c_int), _D],
    "vqc_post_select": [_P, _I, _I, _D],
    "vqc_sv_select": [_P, _I, _I, _P, _D],
    "vqc_partial_trace": [_P, _I, _I32, _P],
    "vqc_dm_sandwich": [_P, _P, _D],
    "vqc_dm_fidelity": [_P, _P, _D],
    "vqc_apply_circuit": [_P, _I, ctypes.POINTER(Op), _D],
    "vqc_plan_info": [ctypes.POINTER(PlanInfo)],
    "vqc_plan_check": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Op), _D, ctypes.c_int, ctypes.POINTER(Term), _D,
                       ctypes.c_int, ctypes.c_int, ctypes.POINTER(PlanInfo)],
    "vqc_sv_energy_grad": [_P, _I, ctypes.POINTER(Op), _D, _I, ctypes.POINTER(Term), _D, _I, _D, _D, _P, _I],
    "vqc_circuit_vjp": [_P, _I, ctypes.POINTER(Op), _D, _I, _P, _D, _P, _P, _I],
    "vqc_dm_energy_grad": [_P, _I, ctypes.POINTER(Op), _D, _I, ctypes.POINTER(Term), _D, _I, _D, _D, _P, _I],
    "vqc_batch_create": [_I, ctypes.POINTER(Instance), ctypes.POINTER(Op), ctypes.c_int64, ctypes.POINTER(Term),
                         ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _I, ctypes.POINTER(_P)],
    "vqc_batch_run": [_P, _D, _D, _D, _D],
    "vqc_batch_launch": [_P, _D, _D],
    "vqc_batch_wait": [_P, _D, _D],
    "vqc_batch_arena_bytes": [_P, ctypes.POINTER(ctypes.c_uint64)],
    "vqc_batch_info": [_P, ctypes.POINTER(BatchInfo)],
    "vqc_batch_destroy": [_P],
    "vqc_server_default_config": [ctypes.POINTER(ServerConfig)],
    "vqc_server_create": [ctypes.POINTER(ServerConfig), ctypes.POINTER(_P)],
    "vqc_server_submit": [_P, ctypes.POINTER(Instance), ctypes.POINTER(Op), ctypes.POINTER(Term), _D, ctypes.c_int64, _D,
                          ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)],
    "vqc_server_poll": [_P, ctypes.c_int64, _D, _D, _D],
    "vqc_server_wait": [_P, ctypes.c_int64, _D, _D, _D],
    "vqc_server_stats": [_P, ctypes.POINTER(ServerStats)],
    "vqc_server_destroy": [_P],
}

_lib = None


def load():
    """Load libvqc_hip.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise VQCError(-4, f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    lib.vqc_last_error.restype = ctypes.c_char_p
    lib.vqc_last_error.argtypes = []
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        fn.argtypes = args
    _lib = lib
    return lib


def call(name, *args):
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != VQC_OK:
        raise VQCError(rc, lib.vqc_last_error().decode(errors="replace"))
    return rc


def call_status(name, *args):
    """Like call, for entry points whose positive return values are statuses."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc < 0:
        raise VQCError(rc, lib.vqc_last_error().decode(errors="replace"))
    return rc


def dptr(a):
    """double* view of a contiguous float64 or complex128 numpy array."""
    assert a.flags.c_contiguous and a.dtype in (np.float64, np.complex128)
    return a.ctypes.data_as(_D)


def i32ptr(seq):
    arr = (ctypes.c_int32 * len(seq))(*[int(x) for x in seq])
    return arr
