"""Pin the CPU oracle (oracle/vqc_oracle.cpp) before comparing the device results against it.

Every oracle routine is checked against something it does not share code with:
numpy brute force on the full 2^n operator, closed forms, the reference's
known-answer tests, exact parameter-shift rules and central finite differences.
The matrices are random and non-symmetric so a transposed operand, a swapped
bit order or a dropped conjugate fails.
"""
import itertools

import numpy as np
import pytest

import oracle
from brute import apply_np, full_op, pauli_full, rand_density, rand_matrix, rand_state, rand_unitary, unvec, vec
from vqchip import (CNOTGate, CPHASEGate, CZGate, Depolarizing, FREDKINGate, PHASEGate, QCircuit, QubitsOperator,
                    QubitsTerm, RxGate, RyGate, RzGate, SWAPGate, XGate, iSWAPGate, AmplitudeDamping, CCPHASEGate,
                    CRxGate, FSIMGate, QuantumGate)
from vqchip import _lib
from vqchip.circuit import Program
from vqchip.operators import pack_pauli, pack_terms
from vqchip.workloads import hea_circuit, tfim, heisenberg, sum_z

TOL = 1e-12


@pytest.fixture(scope="module", params=[1, 4])
def threads(request):
    oracle.set_threads(request.param)
    yield request.param
    oracle.set_threads(1)


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_dense_gate_all_targets(k, threads):
    rng = np.random.default_rng(100 + k)
    nb = 7 if k <= 3 else 6
    combos = list(itertools.combinations(range(nb), k))
    for pos in combos:
        v = rand_state(nb, rng)
        U = rand_matrix(k, rng)  # not unitary: exercises every matrix entry
        ref = full_op(nb, pos, U) @ v
        out = oracle.apply_dense(v.copy(), k, pos, U)
        assert np.max(np.abs(out - ref)) < TOL * 10, pos


def test_apply_np_matches_full_operator():
    # the tensor-contraction helper used at larger n is itself pinned here
    rng = np.random.default_rng(70)
    for k, pos in [(1, (3,)), (2, (0, 4)), (3, (1, 2, 5))]:
        v, U = rand_state(6, rng), rand_matrix(k, rng)
        assert np.max(np.abs(apply_np(v, pos, U) - full_op(6, pos, U) @ v)) < 1e-12


def test_dense_gate_threaded_chunks():
    # the chunked ("H") variant and the threaded partition above MIN_SIZE groups
    rng = np.random.default_rng(7)
    nb = 14
    v = rand_state(nb, rng)
    for k, pos in [(1, (5,)), (2, (3, 9)), (3, (2, 7, 13)), (1, (0,)), (2, (0, 13)), (4, (1, 4, 8, 12))]:
        U = rand_unitary(k, rng)
        oracle.set_threads(1)
        a = oracle.apply_dense(v.copy(), k, pos, U)
        oracle.set_threads(4)
        b = oracle.apply_dense(v.copy(), k, pos, U)
        oracle.set_threads(1)
        assert np.max(np.abs(a - apply_np(v, pos, U))) < 1e-12
        assert np.max(np.abs(a - b)) < 1e-15


def test_named_gates_match_closed_forms():
    rng = np.random.default_rng(11)
    nb = 6
    cases = [
        (_lib.GATE_X, (2,), XGate(3)),
        (_lib.GATE_PHASE, (4,), PHASEGate(5, 0.7)),
        (_lib.GATE_CZ, (1, 4), CZGate(2, 5)),
        (_lib.GATE_CNOT, (4, 1), CNOTGate(5, 2)),
        (_lib.GATE_CNOT, (0, 5), CNOTGate(1, 6)),
        (_lib.GATE_CPHASE, (3, 0), CPHASEGate(4, 1, -1.3)),
        (_lib.GATE_SWAP, (0, 3), SWAPGate(1, 4)),
        (_lib.GATE_ISWAP, (5, 2), iSWAPGate(6, 3)),
        (_lib.GATE_FREDKIN, (3, 0, 5), FREDKINGate(4, 1, 6)),
        (_lib.GATE_FREDKIN, (0, 3, 5), FREDKINGate(1, 4, 6)),
        (_lib.GATE_CCPHASE, (1, 4, 2), CCPHASEGate(2, 5, 3, 0.4)),
    ]
    for kind, pos, gate in cases:
        v = rand_state(nb, rng)
        phi = gate.params[0] if gate.params else 0.0
        out = oracle.apply_named(v.copy(), kind, pos, phi)
        gp = [q - 1 for q in gate.ordered_positions()]
        ref = full_op(nb, gp, gate.ordered_mat()) @ v
        assert np.max(np.abs(out - ref)) < TOL, gate
    # CNOT control is the first position: |c=1,t=0> -> |c=1,t=1>
    v = np.zeros(1 << nb, dtype=complex)
    v[1 << 4] = 1
    out = oracle.apply_named(v.copy(), _lib.GATE_CNOT, (4, 1))
    assert out[(1 << 4) | (1 << 1)] == 1


@pytest.mark.parametrize("k", [1, 2, 3, 5])
def test_accumulate_and_expectations(k):
    rng = np.random.default_rng(200 + k)
    nb = 7
    for pos in list(itertools.combinations(range(nb), k))[:12]:
        v, w = rand_state(nb, rng), rand_state(nb, rng)
        M = rand_matrix(k, rng)
        F = full_op(nb, pos, M)
        acc = rand_state(nb, rng)
        out = oracle.accum_dense(v, acc.copy(), k, pos, M)
        assert np.max(np.abs(out - (acc + F @ v))) < 1e-12
        assert abs(oracle.expec_dense(v, k, pos, M) - np.vdot(v, F @ v)) < 1e-12
        assert abs(oracle.expec_dense(v, k, pos, M, vc=w) - np.vdot(w, F @ v)) < 1e-12


def _pauli_op(nb, rng, nterms):
    terms = []
    for _ in range(nterms):
        labels = rng.integers(0, 4, size=nb)
        ops = {q + 1: "IXYZ"[l] for q, l in enumerate(labels) if l} or {1: "I"}
        terms.append(QubitsTerm(ops, coeff=complex(rng.standard_normal(), rng.standard_normal())))
    return QubitsOperator(terms)


def test_pauli_strings():
    rng = np.random.default_rng(5)
    nb = 6
    H = _pauli_op(nb, rng, 9)
    arr, nt = pack_pauli(H)
    v, w = rand_state(nb, rng), rand_state(nb, rng)
    per, tot = oracle.expec_pauli(v, arr, nt)
    dense = sum(pauli_full(nb, arr[i].x, arr[i].z, complex(arr[i].re, arr[i].im)) for i in range(nt))
    for i in range(nt):
        Pi = pauli_full(nb, arr[i].x, arr[i].z, complex(arr[i].re, arr[i].im))
        assert abs(per[i] - np.vdot(v, Pi @ v)) < 1e-12
        # the same term through its QubitsTerm matrix (kron of the site operators)
        t = H.terms[i]
        pos = [p - 1 for p in t.positions]
        assert np.allclose(full_op(nb, pos, t.matrix()), Pi, atol=1e-14)
    assert abs(tot - np.vdot(v, dense @ v)) < 1e-11
    _, tot2 = oracle.expec_pauli(v, arr, nt, vc=w)
    assert abs(tot2 - np.vdot(w, dense @ v)) < 1e-11
    assert np.max(np.abs(oracle.apply_pauli_sum(v, arr, nt) - dense @ v)) < 1e-11


def test_density_matrix_maps():
    rng = np.random.default_rng(9)
    n = 4
    rho = rand_density(n, rng)
    for k, pos in [(1, (2,)), (2, (0, 3)), (3, (0, 1, 3))]:
        U = rand_unitary(k, rng)
        Uf = full_op(n, pos, U)
        out = oracle.dm_apply_gate(vec(rho).copy(), n, k, pos, U)
        assert np.max(np.abs(unvec(out) - Uf @ rho @ Uf.conj().T)) < 1e-13
    # channels: superoperator = sum conj(K) (x) K acts as sum K rho K^dagger
    for ch in (Depolarizing(2, 0.1), AmplitudeDamping(4, 0.3)):
        pos = [q - 1 for q in ch.ordered_positions()]
        S = ch.ordered_supermat()
        out = oracle.dm_apply_superop(vec(rho).copy(), n, 1, pos, S)
        ref = sum(full_op(n, pos, K) @ rho @ full_op(n, pos, K).conj().T for K in ch.kraus)
        assert np.max(np.abs(unvec(out) - ref)) < 1e-13
        outd = oracle.dm_apply_superop(vec(rho).copy(), n, 1, pos, S, mode=1)
        refd = sum(full_op(n, pos, K).conj().T @ rho @ full_op(n, pos, K) for K in ch.kraus)
        assert np.max(np.abs(unvec(outd) - refd)) < 1e-13
        # the inverse superoperator undoes the channel (circuitdiff.jl:53-59)
        back = oracle.dm_apply_superop(out.copy(), n, 1, pos, np.linalg.inv(S))
        assert np.max(np.abs(unvec(back) - rho)) < 1e-12
    # tr(M rho) on the diagonal blocks (expec_dm_serial.jl)
    for k, pos in [(1, (1,)), (2, (1, 3)), (3, (0, 2, 3)), (4, (0, 1, 2, 3))]:
        M = rand_matrix(k, rng)
        assert abs(oracle.dm_expec(vec(rho), n, k, pos, M) - np.trace(full_op(n, pos, M) @ rho)) < 1e-12


def test_dm_of_pure_state_matches_sv():
    # tr(H rho) with rho = |psi><psi| equals <psi|H|psi> (test/check_dm.jl:4-60)
    rng = np.random.default_rng(31)
    n = 5
    psi = rand_state(n, rng)
    rho = np.kron(psi.conj(), psi)  # DensityMatrix(psi), densitymatrix.jl:27
    assert np.allclose(unvec(rho), np.outer(psi, psi.conj()))
    for k, pos in [(1, (3,)), (2, (0, 4)), (3, (1, 2, 4))]:
        M = rand_matrix(k, rng)
        assert abs(oracle.dm_expec(rho, n, k, pos, M) - oracle.expec_dense(psi, k, pos, M)) < 1e-12


def test_state_layout_known_answers():
    # test/check_state_init.jl:2-29: onehot_encoding([1,0]) == kron(ZERO, ONE) etc.
    from vqchip.state import onehot_vector, qubit_vector
    ZERO, ONE = np.array([1, 0]), np.array([0, 1])
    assert np.array_equal(onehot_vector([1]), ONE)
    assert np.array_equal(onehot_vector([1, 0]), np.kron(ZERO, ONE))
    assert np.array_equal(onehot_vector([0, 1]), np.kron(ONE, ZERO))
    assert np.array_equal(onehot_vector([1, 0, 0]), np.kron(np.kron(ZERO, ZERO), ONE))
    v = np.array([np.sqrt(2) / 2, np.sqrt(2) / 2])
    assert np.allclose(qubit_vector([0.5]), v)
    assert np.allclose(qubit_vector([0.5, 1]), np.kron(ONE, v))
    assert np.allclose(qubit_vector([1.0, 0.0]), onehot_vector([1, 0]))


def _energy(circuit, H, n, psi0):
    prog = Program(circuit, n=n)
    terms, nt, hpool = pack_terms(H)
    return oracle.energy_grad(psi0, n, 0, prog.op_array(), len(prog.ops), prog.pool(), terms, nt, hpool, prog.nparams)


def _dense_H(H, n):
    return sum(full_op(n, [p - 1 for p in key], M) for key, M in H.data().items())


def test_adjoint_gradient_parameter_shift():
    # exact: dE/dtheta = [E(theta + pi/2) - E(theta - pi/2)] / 2 for exp(-i theta P / 2)
    rng = np.random.default_rng(3)
    n, depth = 4, 2
    thetas = rng.uniform(0, 2 * np.pi, 3 * n * depth)
    H = heisenberg(n, rng.standard_normal(n - 1))
    psi0 = np.zeros(1 << n, dtype=complex)
    psi0[0] = 1
    E, g, psi = _energy(hea_circuit(n, depth, thetas), H, n, psi0)
    Hd = _dense_H(H, n)
    assert abs(E - np.vdot(psi, Hd @ psi)) < 1e-12
    for j in range(len(thetas)):
        tp, tm = thetas.copy(), thetas.copy()
        tp[j] += np.pi / 2
        tm[j] -= np.pi / 2
        Ep = _energy(hea_circuit(n, depth, tp), H, n, psi0)[0].real
        Em = _energy(hea_circuit(n, depth, tm), H, n, psi0)[0].real
        assert abs(g[j] - (Ep - Em) / 2) < 1e-11, j


def _fd(f, x, h=1e-4):
    # 5-point central difference (central_fdm(5, 1), test/util.jl:8)
    return (-f(x + 2 * h) + 8 * f(x + h) - 8 * f(x - h) + f(x - 2 * h)) / (12 * h)


def test_adjoint_gradient_general_gates_fd():
    # non-rotation generators (PHASE, CPHASE, CRx, FSIM) and a non-Hermitian observable
    rng = np.random.default_rng(8)
    n = 4
    ps = list(rng.uniform(0, 2 * np.pi, 7))

    def circ(p):
        return QCircuit([RyGate(1, p[0]), CRxGate(1, 3, p[1]), PHASEGate(2, p[2]), FSIMGate(4, 2, p[3], p[4]),
                         CPHASEGate(3, 1, p[5]), QuantumGate((2, 4), rand_unitary(2, np.random.default_rng(1))),
                         RxGate(3, p[6])])

    H = QubitsOperator([QubitsTerm({1: "+", 3: "-"}, coeff=0.7), QubitsTerm({2: "X", 4: "Y"}, coeff=0.36),
                        QubitsTerm({3: "Z"}, coeff=-0.4)])
    psi0 = rand_state(n, np.random.default_rng(2))
    E, g, _ = _energy(circ(ps), H, n, psi0)
    for j in range(len(ps)):
        def f(x):
            q = list(ps)
            q[j] = x
            return _energy(circ(q), H, n, psi0)[0].real
        assert abs(g[j] - _fd(f, ps[j])) < 1e-8, j


def _dm_energy(circuit, H, n, rho0):
    prog = Program(circuit, n=n, dm=True)
    terms, nt, hpool = pack_terms(H)
    return oracle.energy_grad(rho0, n, 1, prog.op_array(), len(prog.ops), prog.pool(), terms, nt, hpool, prog.nparams)


def _noisy(n, thetas):
    c = QCircuit()
    it = iter(range(len(thetas)))
    for _ in range(len(thetas) // (3 * n)):
        for i in range(1, n + 1):
            c.append(RxGate(i, thetas[next(it)]))
            c.append(RyGate(i, thetas[next(it)]))
            c.append(RzGate(i, thetas[next(it)]))
        for i in range(1, n):
            c.append(CNOTGate(i, i + 1))
            c.append(Depolarizing(i + 1, 0.1))
        for i in range(1, n + 1):
            c.append(AmplitudeDamping(i, 0.05))
    return c


@pytest.mark.parametrize("observable", ["sumz", "nonhermitian"])
def test_dm_adjoint_gradient_fd(observable):
    # test/check_noisy_grad.jl method: noisy circuit, gradient vs central differences
    rng = np.random.default_rng(12)
    n = 3
    thetas = rng.uniform(0, 2 * np.pi, 3 * n * 2)
    if observable == "sumz":
        H = sum_z(n)
    else:
        H = QubitsOperator([QubitsTerm({1: "+", 2: "-"}, coeff=0.7), QubitsTerm({1: "-", 2: "+"}, coeff=0.77),
                            QubitsTerm({1: "X", 3: "Y"}, coeff=0.36)])
    rho0 = np.zeros(1 << (2 * n), dtype=complex)
    rho0[0] = 1
    E, g, rho = _dm_energy(_noisy(n, thetas), H, n, rho0)
    assert abs(E - np.trace(_dense_H(H, n) @ unvec(rho))) < 1e-12
    assert abs(np.trace(unvec(rho)) - 1) < 1e-12
    for j in range(0, len(thetas), 2):
        def f(x):
            t = thetas.copy()
            t[j] = x
            return _dm_energy(_noisy(n, t), H, n, rho0)[0].real
        assert abs(g[j] - _fd(f, thetas[j])) < 1e-8, j


def test_adjoint_threads_agree():
    circuit, H, thetas = hea_circuit(10, 2, np.random.default_rng(4).uniform(0, 6, 60)), tfim(10), None
    psi0 = np.zeros(1 << 10, dtype=complex)
    psi0[0] = 1
    oracle.set_threads(1)
    E1, g1, _ = _energy(circuit, H, 10, psi0)
    oracle.set_threads(4)
    E4, g4, _ = _energy(circuit, H, 10, psi0)
    oracle.set_threads(1)
    assert abs(E1 - E4) < 1e-12 and np.max(np.abs(g1 - g4)) < 1e-12


# ---- measurement and post-selection (measure.jl, postselect.jl) -------------
def _bits(n, key0):
    return (np.arange(1 << n) >> key0) & 1


def test_measurement_probability_brute_force():
    rng = np.random.default_rng(61)
    for n, key0 in [(1, 0), (3, 0), (3, 2), (7, 4)]:
        v = rand_state(n, rng)
        b = _bits(n, key0)
        assert abs(oracle.prob_zero(v, n, key0) - np.sum(np.abs(v[b == 0]) ** 2)) < 1e-14
        rho = rand_density(n, rng)
        P0 = np.diag((b == 0).astype(float))
        assert abs(oracle.prob_zero(vec(rho), n, key0, is_dm=True) - np.trace(P0 @ rho).real) < 1e-14


def test_measure_collapse_brute_force():
    rng = np.random.default_rng(62)
    n = 5
    for key0 in (0, 2, 4):
        for auto_reset in (True, False):
            for u in (0.05, 0.95):
                v = rand_state(n, rng)
                b = _bits(n, key0)
                p0 = np.sum(np.abs(v[b == 0]) ** 2)
                i_exp = 0 if u < p0 else 1
                w = v.copy()
                i, p = oracle.measure_inplace(w, n, key0, u, auto_reset)
                assert i == i_exp and abs(p - (p0 if i == 0 else 1 - p0)) < 1e-14
                exp = np.zeros_like(v)
                src = v[b == i] / np.sqrt(p)
                exp[b == (0 if auto_reset else i)] = src
                assert np.max(np.abs(w - exp)) < 1e-14
    # density matrix: M rho M^dagger with M = |0><i| / sqrt(p)
    rho = rand_density(4, rng)
    r = vec(rho)
    i, p = oracle.measure_inplace(r, 4, 1, 0.5, True, is_dm=True)
    M = np.zeros((2, 2), complex)
    M[0, i] = 1 / np.sqrt(p)
    F = full_op(4, [1], M)
    assert np.max(np.abs(unvec(r) - F @ rho @ F.conj().T)) < 1e-14
    assert abs(np.trace(unvec(r)) - 1) < 1e-13


def test_measure_and_select_known_answers():
    # check_measure.jl: after measure! the same qubit is certain (p = 1)
    rng = np.random.default_rng(63)
    for key0, auto_reset in [(0, True), (1, False)]:
        v = rand_state(3, rng)
        i, p = oracle.measure_inplace(v, 3, key0, float(rng.random()), auto_reset)
        p0 = oracle.prob_zero(v, 3, key0)
        assert abs((p0 if (auto_reset or i == 0) else 1 - p0) - 1) < 1e-10
    # check_select_1/2: post_select! then measure gives the selected value with p = 1
    for key0, val in [(1, 0), (2, 1)]:
        v = rand_state(3, rng)
        oracle.post_select_inplace(v, 3, key0, val)
        p0 = oracle.prob_zero(v, 3, key0)
        assert abs((p0 if val == 0 else 1 - p0) - 1) < 1e-10
    # check_select_3/4: (|00> + |11>)/sqrt(2) post-selected on qubit 1 -> ZERO / ONE
    bell = np.zeros(4, complex)
    bell[0] = bell[3] = np.sqrt(0.5)
    for val, exp in [(0, [1, 0]), (1, [0, 1])]:
        r = oracle.select(bell.copy(), 2, 0, val)
        r = r / np.sqrt(np.vdot(r, r))
        assert np.allclose(r, exp, atol=1e-15)


def select_brute(v, n, key0, val):
    """Row `val` of the measured qubit with the reference's qubit order: the
    measured qubit leaves, and qubit 1 takes its place (the swap!(v, 1, pos) of
    measure.jl:30 / postselect.jl:48): new qubit j holds original qubit
    order[j]."""
    if key0 == 0:
        order = list(range(1, n))
    else:
        order = list(range(1, key0)) + [0] + list(range(key0 + 1, n))
    out = np.zeros(1 << (n - 1), dtype=complex)
    for c in range(1 << (n - 1)):
        src = val << key0
        for j, q in enumerate(order):
            src |= ((c >> j) & 1) << q
        out[c] = v[src]
    return out


def test_select_brute_force():
    rng = np.random.default_rng(64)
    n = 6
    for key0 in range(n):
        v = rand_state(n, rng)
        for val in (0, 1):
            w = v.copy()
            assert np.array_equal(oracle.select(w, n, key0, val), select_brute(v, n, key0, val))
            assert np.array_equal(w, v)  # the swaps are undone
    # the permutation matters: a product state |q1=1, q2=0, q3=0> selected on q3=0
    # gives |1, 0> in the reference's order (q2, q1), i.e. index 2, not index 1
    v = np.zeros(8, complex)
    v[1] = 1.0
    assert np.array_equal(oracle.select(v, 3, 2, 0), np.eye(1, 4, 2, dtype=complex).reshape(-1))


def _rand_factor(kind, rng):
    from vqchip.gates import SIGMA_MINUS, SIGMA_PLUS
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    Y = np.array([[0, -1j], [1j, 0]])
    Z = np.diag([1.0 + 0j, -1])
    if kind == "dense":
        return rand_matrix(1, rng)
    if kind == "diag":
        return np.diag(rng.standard_normal(2) + 1j * rng.standard_normal(2))
    if kind == "anti":
        return np.array([[0, rng.standard_normal() + 1j], [rng.standard_normal() - 0.5j, 0]])
    return {"+": SIGMA_PLUS, "-": SIGMA_MINUS, "X": X, "Y": Y, "Z": Z, "I": np.eye(2, dtype=complex),
            "P0": np.diag([1.0 + 0j, 0]), "P1": np.diag([0j, 1.0])}[kind]


def _rand_qterms(n, rng, kinds, nterms=3, wmin=6):
    out = []
    for _ in range(nterms):
        w = int(rng.integers(wmin, n + 1))
        sites = sorted(rng.choice(n, size=w, replace=False).tolist())
        mats = [_rand_factor(kinds[int(rng.integers(len(kinds)))], rng) for _ in sites]
        out.append((complex(rng.standard_normal(), rng.standard_normal()), sites, mats))
    return out


def _qterm_full(n, c, sites, mats):
    from vqchip.gates import kron_le
    return c * full_op(n, sites, kron_le(*mats))


@pytest.mark.parametrize("kinds", [("+", "-"), ("+", "-", "X", "Y", "Z", "I", "P0", "P1"), ("dense", "diag", "anti", "Z"),
                                   ("dense",)])
def test_qterms_brute_force(kinds):
    """Product terms of 6..n sites (_apply_qterm_util!, apply_qterms.jl:47-59) against
    the dense operator built site by site (test/check_qubits_ham.jl:73-101 method:
    sigma+ / sigma- strings, m * state and expectation against matrix(L, m))."""
    rng = np.random.default_rng(sum(map(ord, "".join(kinds))))
    for n in (7, 9):
        qt = _rand_qterms(n, rng, kinds)
        v = rand_state(n, rng)
        vc = rand_state(n, rng)
        H = sum(_qterm_full(n, c, s, m) for c, s, m in qt)
        assert np.allclose(oracle.apply_qterms(v, qt), H @ v, atol=1e-12)
        per, tot = oracle.expec_qterms(v, qt)
        assert np.allclose(per, [v.conj() @ _qterm_full(n, c, s, m) @ v for c, s, m in qt], atol=1e-12)
        assert abs(tot - v.conj() @ H @ v) < 1e-12
        _, tot2 = oracle.expec_qterms(v, qt, vc=vc)
        assert abs(tot2 - vc.conj() @ H @ v) < 1e-12
    # density matrices: tr(T rho) with T on the row bits (expecs_dm.jl:25, apply_qterms_dm.jl:11-27)
    n = 6
    qt = _rand_qterms(n, rng, kinds, wmin=4)
    rho = rand_density(n, rng)
    per, tot = oracle.dm_expec_qterms(rho, n, qt)
    assert np.allclose(per, [np.trace(_qterm_full(n, c, s, m) @ rho) for c, s, m in qt], atol=1e-12)
    assert abs(tot - sum(per)) < 1e-12


def test_fidelity_dm_pins():
    """fidelity(x::DensityMatrix, y::DensityMatrix) (densitymatrix.jl:65) in the oracle:
    against scipy's Schur-based sqrtm, and the closed forms tr(sqrt(rho) sqrt(rho)) =
    tr(rho) = 1, pure states |<psi|phi>|^2 (sqrt(rho) = rho), diagonal matrices
    sum sqrt(p_i q_i)."""
    from scipy.linalg import sqrtm
    rng = np.random.default_rng(65)
    for n in (1, 3, 5):
        x, y = rand_density(n, rng), rand_density(n, rng)
        f = oracle.fidelity_dm(x, y)
        assert abs(f - np.real(np.trace(sqrtm(x) @ sqrtm(y)))) < 1e-10
        assert abs(oracle.fidelity_dm(x, x) - 1.0) < 1e-12
        psi, phi = rand_state(n, rng), rand_state(n, rng)
        fp = oracle.fidelity_dm(np.outer(psi, psi.conj()), np.outer(phi, phi.conj()))
        # rank-deficient: round-off eigenvalues of order eps enter as sqrt(eps) ~ 1e-8,
        # in the reference's formula as here
        assert abs(fp - abs(psi.conj() @ phi) ** 2) < 1e-7
        p, q = rng.dirichlet(np.ones(1 << n)), rng.dirichlet(np.ones(1 << n))
        assert abs(oracle.fidelity_dm(np.diag(p), np.diag(q)) - np.sum(np.sqrt(p * q))) < 1e-12
