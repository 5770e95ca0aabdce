"""Device-resident StateVector / DensityMatrix (src/statevector.jl, src/densitymatrix.jl).

The amplitudes live in HBM behind a libvqc_hip handle; `storage()` copies them
to the host. Layout: qubit q (1-based) is bit q-1 of the index; a density
matrix stores vec(rho) column-major with rho[i, j] at i + j*2^n.
"""
import ctypes

import numpy as np

from . import _lib


class _DeviceState:
    is_dm = 0

    def __init__(self, nbits, handle):
        self._h = handle
        self._nbits = nbits

    @property
    def handle(self):
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.call("vqc_state_destroy", h)
            except Exception:
                pass
            self._h = None

    def _raw(self):
        out = np.empty(1 << self._nbits, dtype=np.complex128)
        _lib.call("vqc_state_download", self._h, _lib.dptr(out), out.size)
        return out

    def _upload(self, data):
        data = np.ascontiguousarray(data, dtype=np.complex128).reshape(-1)
        if data.size != 1 << self._nbits:
            raise ValueError(f"expected {1 << self._nbits} amplitudes, got {data.size}")
        _lib.call("vqc_state_upload", self._h, _lib.dptr(data), data.size)

    def reset_(self):
        """reset!(x) (statevector.jl:111-115)."""
        _lib.call("vqc_state_reset", self._h)
        return self

    def scale_(self, a):
        a = complex(a)
        _lib.call("vqc_state_scale", self._h, a.real, a.imag)
        return self


class StateVector(_DeviceState):
    """StateVector{ComplexF64} (statevector.jl:1-21)."""

    def __init__(self, data_or_n):
        if isinstance(data_or_n, (int, np.integer)):
            n = int(data_or_n)
            h = ctypes.c_void_p()
            _lib.call("vqc_sv_create", n, ctypes.byref(h))
            super().__init__(n, h)
        else:
            data = np.ascontiguousarray(data_or_n, dtype=np.complex128).reshape(-1)
            n = int(round(np.log2(data.size)))
            if 1 << n != data.size:
                raise ValueError("state can not be interpretted as a qubit state.")  # statevector.jl:219
            h = ctypes.c_void_p()
            _lib.call("vqc_sv_create", n, ctypes.byref(h))
            super().__init__(n, h)
            self._upload(data)

    def nqubits(self):
        return self._nbits

    def storage(self):
        return self._raw()

    def copy(self):
        s = StateVector(self._nbits)
        _lib.call("vqc_state_copy", s._h, self._h)
        return s

    def dot(self, other):
        out = np.zeros(2)
        _lib.call("vqc_state_dot", self._h, other._h, _lib.dptr(out))
        return complex(out[0], out[1])

    def norm(self):
        return float(np.sqrt(self.dot(self).real))


class DensityMatrix(_DeviceState):
    """DensityMatrix{ComplexF64} (densitymatrix.jl:1-42)."""

    is_dm = 1

    def __init__(self, data_or_n):
        h = ctypes.c_void_p()
        if isinstance(data_or_n, (int, np.integer)):
            n = int(data_or_n)
            _lib.call("vqc_dm_create", n, ctypes.byref(h))
            super().__init__(2 * n, h)
            return
        if isinstance(data_or_n, StateVector):  # DensityMatrix(psi) = kron(conj psi, psi), on the device
            n = data_or_n.nqubits()
            _lib.call("vqc_dm_create", n, ctypes.byref(h))
            super().__init__(2 * n, h)
            _lib.call("vqc_dm_from_sv", self._h, data_or_n.handle)
            return
        else:
            data = np.asarray(data_or_n, dtype=np.complex128)
            if data.ndim == 2:
                data = data.reshape(-1, order="F")
        nb = int(round(np.log2(data.size)))
        if 1 << nb != data.size or nb % 2:
            raise ValueError("DimensionMismatch")  # densitymatrix.jl:6
        _lib.call("vqc_dm_create", nb // 2, ctypes.byref(h))
        super().__init__(nb, h)
        self._upload(data)

    def nqubits(self):
        return self._nbits // 2

    def storage(self):
        """The 2^n x 2^n matrix (densitymatrix.jl:30)."""
        N = 1 << self.nqubits()
        return self._raw().reshape(N, N, order="F")

    def copy(self):
        s = DensityMatrix(self.nqubits())
        _lib.call("vqc_state_copy", s._h, self._h)
        return s

    def tr(self):
        return complex(np.trace(self.storage()))


def onehot_vector(bits):
    """Amplitudes of onehot_encoding(bits): index sum_i b_i 2^(i-1) (_sub2ind, statevector.jl:167-176)."""
    idx = 0
    for i, b in enumerate(bits):
        if b not in (0, 1):
            raise ValueError("indices must be 0 or 1")
        idx |= int(b) << i
    v = np.zeros(1 << len(bits), dtype=np.complex128)
    v[idx] = 1
    return v


def onehot_encoding(bits):
    """onehot_encoding([b1, b2, ...]): qubit i in state b_i (statevector.jl:67-70)."""
    if isinstance(bits, (int, np.integer)):
        return StateVector(int(bits))
    return StateVector(onehot_vector(bits))


def qubit_vector(thetas):
    """Product of [cos(pi s/2), sin(pi s/2)] with qubit 1 least significant (statevector.jl:85,93-100)."""
    if len(thetas) == 0:
        raise ValueError("empty input.")
    v = np.array([1.0 + 0j])
    for s in thetas:
        v = np.kron(np.array([np.cos(s * np.pi / 2), np.sin(s * np.pi / 2)], dtype=np.complex128), v)
    return v


def qubit_encoding(thetas):
    """qubit_encoding(thetas) (statevector.jl:93-100)."""
    return StateVector(qubit_vector(thetas))


def rand_state(n, rng=None):
    """Normalised complex Gaussian state (statevector.jl:152-158), from a seeded generator."""
    rng = np.random.default_rng(rng)
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return StateVector(v / np.linalg.norm(v))
