"""The committed fixtures of tests/golden/ on the CPU: the reference's own known
answers for the amplitude layout and post-selection (test/check_state_init.jl,
test/check_measure.jl:31-41) against the host encoders and the oracle, and the
oracle against its frozen outputs (so a change to the oracle, or to the host
packing of circuits and Hamiltonians, shows up as a fixture mismatch)."""
import numpy as np
import pytest

import oracle
from goldens import dense_cases, load, ops_from_table, paulis_from, terms_from_table
from vqchip.circuit import Program
from vqchip.operators import pack_terms
from vqchip.state import onehot_vector, qubit_vector
from vqchip.workloads import config0, config3


def test_known_answers_of_the_state_layout():
    d = load("known_answers.npz")
    for i in range(int(d["ncases"])):
        kind, x, e = str(d[f"c{i}_kind"]), d[f"c{i}_in"], d[f"c{i}_out"]
        if kind == "onehot":
            assert np.array_equal(onehot_vector([int(b) for b in x]), e), x  # exact (==) in the reference
        else:
            assert np.allclose(qubit_vector(list(x)), e, atol=1e-15, rtol=0), x  # isapprox in the reference


def test_known_answers_of_post_select():
    d = load("known_answers.npz")
    for value in (0, 1):
        row = oracle.select(d["bell"].copy(), 2, 0, value)
        assert np.allclose(row / np.linalg.norm(row), d[f"bell_sel{value}"], atol=1e-15, rtol=0)
        w = d["bell"].copy()
        nrm = oracle.post_select_inplace(w, 2, 0, value)
        assert abs(nrm - np.sqrt(0.5)) < 1e-15


def test_oracle_reproduces_dense_gates():
    d = load("dense_gates_n7.npz")
    cases = dense_cases(d)
    assert sorted(k for k, *_ in cases) == [1, 1, 2, 3, 4, 5]
    for k, pos, U, out in cases:
        assert np.max(np.abs(oracle.apply_dense(d["psi"].copy(), k, pos, U) - out)) < 1e-14, pos


def _energy_grad_fixture(name, n, is_dm, state_key):
    d = load(name)
    ops, terms = ops_from_table(d["ops"]), terms_from_table(d["terms"])
    s0 = np.zeros(1 << (2 * n if is_dm else n), complex)
    s0[0] = 1
    E, g, out = oracle.energy_grad(s0, n, is_dm, ops, len(d["ops"]), d["pool"], terms, len(d["terms"]), d["hpool"],
                                   int(d["nparams"]))
    assert abs(E - complex(d["energy"])) < 1e-13
    assert np.max(np.abs(g - d["grad"])) < 1e-13
    assert np.max(np.abs(out - d[state_key])) < 1e-14
    return d


def test_oracle_reproduces_hea12_energy_gradient():
    d = _energy_grad_fixture("hea12_energy_grad.npz", 12, 0, "psi_out")
    assert d["grad"].shape == (144,)
    # the host packing of today's config-0 workload still yields the frozen tables
    c, H, _ = config0()
    prog = Program(c, n=12)
    assert np.array_equal(d["ops"][:, :14], np.array([[o.kind, o.nq, *o.pos, o.nparams, *o.slot, o.reserved, o.mat,
                                                       o.aux] for o in prog.ops]))
    assert np.max(np.abs(prog.pool() - d["pool"])) < 1e-15
    terms, nt, hpool = pack_terms(H)
    assert nt == len(d["terms"]) and np.max(np.abs(hpool - d["hpool"])) < 1e-15


def test_oracle_reproduces_noisy_density_matrix():
    d = _energy_grad_fixture("dm6_noisy_energy_grad.npz", 6, 1, "rho_out")
    rho = d["rho_out"].reshape(64, 64, order="F")
    assert abs(np.trace(rho) - 1) < 1e-12 and np.max(np.abs(rho - rho.conj().T)) < 1e-14


def test_oracle_reproduces_pauli_strings():
    d = load("pauli_n10.npz")
    arr = paulis_from(d["xz"], d["coef"])
    per, tot = oracle.expec_pauli(d["psi"], arr, len(d["xz"]))
    assert np.max(np.abs(per - d["per_term"])) < 1e-14 and abs(tot - complex(d["total"])) < 1e-13
    assert np.max(np.abs(oracle.apply_pauli_sum(d["psi"], arr, len(d["xz"])) - d["h_psi"])) < 1e-13


def test_oracle_reproduces_measurement():
    d = load("measure_n6.npz")
    psi = d["psi"]
    for key0 in (0, 3, 5):
        assert abs(oracle.prob_zero(psi, 6, key0) - float(d[f"p0_{key0}"])) < 1e-15
        for u in (0.2, 0.8):
            for reset in (0, 1):
                w = psi.copy()
                i, p = oracle.measure_inplace(w, 6, key0, u, bool(reset))
                ip = d[f"m_{key0}_{u}_{reset}_ip"]
                assert i == int(ip[0]) and abs(p - ip[1]) < 1e-15
                assert np.max(np.abs(w - d[f"m_{key0}_{u}_{reset}"])) < 1e-15
        for val in (0, 1):
            assert np.max(np.abs(oracle.select(psi.copy(), 6, key0, val) - d[f"sel_{key0}_{val}"])) < 1e-15
    r = d["rho4"].copy()
    i, p = oracle.measure_inplace(r, 4, 2, 0.6, True, is_dm=True)
    assert i == int(d["rho4_m_ip"][0]) and np.max(np.abs(r - d["rho4_m"])) < 1e-15
