"""libvqc_hip (through the C ABI) against the committed fixtures of
tests/golden/: dense k = 1..5 gates, the 12-qubit HEA energy and its 144-entry
adjoint gradient (fused and per-gate), a noisy 6-qubit density matrix, Pauli
strings and measurement, plus the reference's known answers for post-selection
(test/check_measure.jl:31-41). Tolerances (FP64): amplitudes 1e-12 absolute,
energies and gradients 1e-10, outcomes exact."""
import ctypes

import numpy as np
import pytest

from goldens import dense_cases, load, ops_from_table, paulis_from, terms_from_table
import vqchip as vq
from vqchip import _lib

pytestmark = pytest.mark.gpu


def test_dense_gates_match_fixture():
    d = load("dense_gates_n7.npz")
    for k, pos, U, out in dense_cases(d):
        s = vq.StateVector(np.ascontiguousarray(d["psi"]))
        _lib.call("vqc_sv_apply_dense", s.handle, k, _lib.i32ptr(pos),
                  _lib.dptr(np.ascontiguousarray(U.reshape(-1, order="F"))))
        assert np.max(np.abs(s.storage() - out)) < 1e-12, pos


def _run_fixture(d, n, dm, flags):
    s0 = np.zeros(1 << (2 * n if dm else n), complex)
    s0[0] = 1
    st = vq.DensityMatrix(s0) if dm else vq.StateVector(s0)
    out = vq.DensityMatrix(n) if dm else vq.StateVector(n)
    npar = int(d["nparams"])
    g, e2 = np.zeros(max(npar, 1)), np.zeros(2)
    _lib.call("vqc_dm_energy_grad" if dm else "vqc_sv_energy_grad", st.handle, len(d["ops"]), ops_from_table(d["ops"]),
              _lib.dptr(np.ascontiguousarray(d["pool"])), len(d["terms"]), terms_from_table(d["terms"]),
              _lib.dptr(np.ascontiguousarray(d["hpool"])), npar, _lib.dptr(e2), _lib.dptr(g), out.handle, flags)
    return complex(e2[0], e2[1]), g[:npar], out.storage()


@pytest.mark.parametrize("flags", [0, 1])
def test_hea12_energy_gradient_matches_fixture(flags):
    d = load("hea12_energy_grad.npz")
    E, g, psi = _run_fixture(d, 12, False, flags)
    assert abs(E - complex(d["energy"])) < 1e-10
    assert g.shape == (144,) and np.max(np.abs(g - d["grad"])) < 1e-10
    assert np.max(np.abs(psi - d["psi_out"])) < 1e-12


@pytest.mark.parametrize("flags", [0, 1])
def test_noisy_density_matrix_matches_fixture(flags):
    d = load("dm6_noisy_energy_grad.npz")
    E, g, rho = _run_fixture(d, 6, True, flags)
    assert abs(E - complex(d["energy"])) < 1e-10
    assert np.max(np.abs(g - d["grad"])) < 1e-10
    assert np.max(np.abs(rho.reshape(-1, order="F") - d["rho_out"])) < 1e-12


def test_pauli_strings_match_fixture():
    d = load("pauli_n10.npz")
    arr, nt = paulis_from(d["xz"], d["coef"]), len(d["xz"])
    s = vq.StateVector(np.ascontiguousarray(d["psi"]))
    per, tot = np.zeros(2 * nt), np.zeros(2)
    _lib.call("vqc_sv_expec_pauli", s.handle, None, nt, arr, _lib.dptr(per), _lib.dptr(tot))
    assert np.max(np.abs(per[0::2] + 1j * per[1::2] - d["per_term"])) < 1e-12
    assert abs(complex(*tot) - complex(d["total"])) < 1e-11
    out = vq.StateVector(10)
    _lib.call("vqc_sv_apply_pauli_sum", s.handle, out.handle, nt, arr)
    assert np.max(np.abs(out.storage() - d["h_psi"])) < 1e-11


def test_measurement_matches_fixture():
    d = load("measure_n6.npz")
    psi = np.ascontiguousarray(d["psi"])
    for key0 in (0, 3, 5):
        assert abs(vq.probability_zero(vq.StateVector(psi), key0 + 1) - float(d[f"p0_{key0}"])) < 1e-12
        for u in (0.2, 0.8):
            for reset in (0, 1):
                s = vq.StateVector(psi)
                i, p = vq.measure_(s, key0 + 1, auto_reset=bool(reset), u=u)
                ip = d[f"m_{key0}_{u}_{reset}_ip"]
                assert i == int(ip[0]) and abs(p - ip[1]) < 1e-12
                assert np.max(np.abs(s.storage() - d[f"m_{key0}_{u}_{reset}"])) < 1e-12
        for val in (0, 1):
            src, out, nrm = vq.StateVector(psi), vq.StateVector(5), ctypes.c_double()
            _lib.call("vqc_sv_select", src.handle, key0, val, out.handle, ctypes.byref(nrm))
            row = d[f"sel_{key0}_{val}"]
            assert abs(nrm.value - np.linalg.norm(row)) < 1e-12
            assert np.max(np.abs(out.storage() * nrm.value - row)) < 1e-12
    r = vq.DensityMatrix(np.ascontiguousarray(d["rho4"]))
    i, p = vq.measure_(r, 3, auto_reset=True, u=0.6)
    assert i == int(d["rho4_m_ip"][0]) and abs(p - d["rho4_m_ip"][1]) < 1e-12
    assert np.max(np.abs(r.storage().reshape(-1, order="F") - d["rho4_m"])) < 1e-12


def test_post_select_known_answers():
    d = load("known_answers.npz")
    for value in (0, 1):
        r, _ = vq.post_select(vq.StateVector(np.ascontiguousarray(d["bell"])), 1, value)
        assert np.max(np.abs(r.storage() - d[f"bell_sel{value}"])) < 1e-15
