"""libvqc_hip (through the C ABI) against the pinned CPU oracle on the same seeded inputs.

Tolerances (FP64 throughout): amplitudes and expectation values within 1e-12
absolute on normalised states of up to 2^16 amplitudes (the target is 1e-10
relative); gradients within 1e-10 of the oracle's adjoint. Integer/index work
(named permutations) must match exactly.
"""
import itertools

import numpy as np
import pytest

import oracle
from circuits import chain_circuit, structured_circuit
from brute import full_op, rand_density, rand_matrix, rand_state, rand_unitary, unvec, vec
import vqchip as vq
from vqchip import _lib
from vqchip.circuit import Program
from vqchip.operators import pack_pauli, pack_terms
from vqchip.workloads import config0, config3, hea_circuit, heisenberg, tfim, sum_z, config1_circuit

pytestmark = pytest.mark.gpu
AMP_TOL = 1e-12


def sv_from(v):
    return vq.StateVector(np.ascontiguousarray(v))


@pytest.mark.parametrize("k,nb", [(1, 8), (2, 8), (3, 8), (4, 7), (5, 7), (1, 9), (2, 9), (3, 9), (4, 10)])
def test_dense_gate_every_target_class(k, nb):
    """Every target set of a k-qubit gate on nb qubits. From k + 6 qubits on, a 2- or
    3-qubit gate takes the 2^(k+6)-amplitude one-amplitude-per-thread tile (below, the
    group-per-thread form), and the sets with every target at bit 6 or above take its
    direct-store form. 4- and 5-qubit gates run group-per-thread at every size."""
    rng = np.random.default_rng(1000 + k + 100 * nb)
    for pos in itertools.combinations(range(nb), k):
        v = rand_state(nb, rng)
        U = rand_unitary(k, rng)
        s = sv_from(v)
        _lib.call("vqc_sv_apply_dense", s.handle, k, _lib.i32ptr(pos), _lib.dptr(np.ascontiguousarray(U.reshape(-1, order="F"))))
        ref = oracle.apply_dense(v.copy(), k, pos, U)
        assert np.max(np.abs(s.storage() - ref)) < AMP_TOL, pos


@pytest.mark.parametrize("nb,k,pos", [(18, 1, (0,)), (18, 1, (17,)), (18, 2, (3, 16)), (18, 3, (0, 9, 17)),
                                      (18, 2, (12, 17)), (18, 3, (6, 13, 17)), (18, 3, (2, 4, 5)),
                                      (18, 4, (6, 9, 12, 17)), (18, 4, (0, 3, 8, 15)),
                                      (17, 4, (1, 5, 9, 16)), (17, 5, (0, 2, 7, 11, 16)), (20, 1, (10,))])
def test_dense_gate_multiblock(nb, k, pos):
    rng = np.random.default_rng(nb * 7 + k)
    v = rand_state(nb, rng)
    U = rand_unitary(k, rng)
    s = sv_from(v)
    _lib.call("vqc_sv_apply_dense", s.handle, k, _lib.i32ptr(pos), _lib.dptr(np.ascontiguousarray(U.reshape(-1, order="F"))))
    ref = oracle.apply_dense(v.copy(), k, pos, U)
    assert np.max(np.abs(s.storage() - ref)) < AMP_TOL


def test_named_gates_bit_exact():
    rng = np.random.default_rng(5)
    nb = 10
    gates = [vq.XGate(3), vq.XGate(1), vq.PHASEGate(10, 0.37), vq.CZGate(2, 9), vq.CNOTGate(9, 2), vq.CNOTGate(1, 2),
             vq.CPHASEGate(4, 1, -0.7), vq.SWAPGate(1, 10), vq.iSWAPGate(6, 3), vq.FREDKINGate(5, 1, 8),
             vq.FREDKINGate(1, 5, 8), vq.CCPHASEGate(2, 7, 4, 1.1)]
    for g in gates:
        v = rand_state(nb, rng)
        s = sv_from(v)
        vq.apply_(g, s)
        ref = oracle.apply_named(v.copy(), g.named_kind, [q - 1 for q in g.positions], g.params[0] if g.params else 0.0)
        got = s.storage()
        if g.name in ("X", "CNOT", "SWAP", "FREDKIN", "CZ"):
            assert np.array_equal(got, ref), g  # pure permutations / sign flips
        else:
            assert np.max(np.abs(got - ref)) < AMP_TOL, g
        # and the same as the dense kernel on the ordered matrix
        s2 = sv_from(v)
        vq.apply_(g, s2, named=False)
        assert np.max(np.abs(s2.storage() - ref)) < AMP_TOL, g
    # iSWAP adjoint
    v = rand_state(nb, rng)
    s = sv_from(v)
    _lib.call("vqc_sv_apply_named", s.handle, _lib.GATE_ISWAPDAG, _lib.i32ptr([4, 7]), 0.0)
    assert np.max(np.abs(s.storage() - oracle.apply_named(v.copy(), _lib.GATE_ISWAPDAG, (4, 7)))) < AMP_TOL


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_accumulate_and_expectations(k):
    rng = np.random.default_rng(300 + k)
    nb = 9
    for pos in list(itertools.combinations(range(nb), k))[::7][:10]:
        v, w, o = rand_state(nb, rng), rand_state(nb, rng), rand_state(nb, rng)
        M = rand_matrix(k, rng)
        Mf = np.ascontiguousarray(M.reshape(-1, order="F"))
        sv, sw, so = sv_from(v), sv_from(w), sv_from(o)
        _lib.call("vqc_sv_accum_dense", sv.handle, so.handle, k, _lib.i32ptr(pos), _lib.dptr(Mf))
        assert np.max(np.abs(so.storage() - oracle.accum_dense(v, o.copy(), k, pos, M))) < AMP_TOL
        out = np.zeros(2)
        _lib.call("vqc_sv_expec_dense", sv.handle, None, k, _lib.i32ptr(pos), _lib.dptr(Mf), _lib.dptr(out))
        assert abs(complex(*out) - oracle.expec_dense(v, k, pos, M)) < AMP_TOL
        _lib.call("vqc_sv_expec_dense", sv.handle, sw.handle, k, _lib.i32ptr(pos), _lib.dptr(Mf), _lib.dptr(out))
        assert abs(complex(*out) - oracle.expec_dense(v, k, pos, M, vc=w)) < AMP_TOL


def test_expectation_api_and_hamiltonian_apply():
    rng = np.random.default_rng(17)
    n = 11
    H = heisenberg(n, rng.standard_normal(n - 1)) + tfim(n, 0.8)
    v = rand_state(n, rng)
    s = sv_from(v)
    terms, nt, hpool = pack_terms(H)
    ref = sum(oracle.expec_dense(v, t.nq, list(t.pos)[: t.nq], hpool[t.mat: t.mat + (1 << (2 * t.nq))].reshape(
        1 << t.nq, 1 << t.nq, order="F")) for t in terms[:nt])
    assert abs(vq.expectation(H, s) - ref) < 1e-11
    out = vq.apply_operator(H, s).storage()
    acc = np.zeros_like(v)
    for t in terms[:nt]:
        D = 1 << t.nq
        oracle.accum_dense(v, acc, t.nq, list(t.pos)[: t.nq], hpool[t.mat: t.mat + D * D].reshape(D, D, order="F"))
    assert np.max(np.abs(out - acc)) < 1e-11


@pytest.mark.parametrize("nb", [7, 14])
def test_pauli_expectation_and_apply(nb):
    rng = np.random.default_rng(nb)
    terms = []
    for i in range(23):
        labels = rng.integers(0, 4, size=nb)
        if i % 7 == 0:
            labels = np.where(labels % 2 == 1, 0, labels)  # some z-only strings (x = 0 pass)
        ops = {q + 1: "IXYZ"[l] for q, l in enumerate(labels) if l} or {1: "Z"}
        terms.append(vq.QubitsTerm(ops, coeff=complex(rng.standard_normal(), 0.3 * rng.standard_normal())))
    terms.append(terms[3])  # a repeated x mask
    H = vq.QubitsOperator(terms)
    arr, nt = pack_pauli(H)
    v, w = rand_state(nb, rng), rand_state(nb, rng)
    s, sw = sv_from(v), sv_from(w)
    per_ref, tot_ref = oracle.expec_pauli(v, arr, nt)
    per, tot = np.zeros(2 * nt), np.zeros(2)
    _lib.call("vqc_sv_expec_pauli", s.handle, None, nt, arr, _lib.dptr(per), _lib.dptr(tot))
    assert np.max(np.abs(per[0::2] + 1j * per[1::2] - per_ref)) < 1e-12
    assert abs(complex(*tot) - tot_ref) < 1e-11
    _, tot2_ref = oracle.expec_pauli(v, arr, nt, vc=w)
    _lib.call("vqc_sv_expec_pauli", s.handle, sw.handle, nt, arr, None, _lib.dptr(tot))
    assert abs(complex(*tot) - tot2_ref) < 1e-11
    out = vq.StateVector(nb)
    _lib.call("vqc_sv_apply_pauli_sum", s.handle, out.handle, nt, arr)
    assert np.max(np.abs(out.storage() - oracle.apply_pauli_sum(v, arr, nt))) < 1e-11


def test_density_matrix_gate_channel_expectation():
    rng = np.random.default_rng(23)
    n = 6
    for k, pos in [(1, (0,)), (1, (5,)), (2, (1, 4)), (3, (0, 2, 5)), (4, (0, 1, 3, 5))]:
        rho = vec(rand_density(n, rng))
        U = rand_unitary(k, rng)
        d = vq.DensityMatrix(rho)
        _lib.call("vqc_dm_apply_gate", d.handle, k, _lib.i32ptr(pos), _lib.dptr(np.ascontiguousarray(U.reshape(-1, order="F"))))
        ref = oracle.dm_apply_gate(rho.copy(), n, k, pos, U)
        assert np.max(np.abs(vec(d.storage()) - ref)) < AMP_TOL
    for N, pos in [(1, (3,)), (2, (0, 5)), (2, (2, 3))]:
        rho = vec(rand_density(n, rng))
        S = rand_matrix(2 * N, rng)
        for mode in (0, 1):
            d = vq.DensityMatrix(rho)
            _lib.call("vqc_dm_apply_superop", d.handle, N, _lib.i32ptr(pos),
                      _lib.dptr(np.ascontiguousarray(S.reshape(-1, order="F"))), mode)
            ref = oracle.dm_apply_superop(rho.copy(), n, N, pos, S, mode)
            assert np.max(np.abs(vec(d.storage()) - ref)) < 1e-11
        # mode 2 is S^-1: it undoes mode 0 (apply_inverse!, apply_qmaps.jl:40-48)
        d = vq.DensityMatrix(rho)
        Sf = _lib.dptr(np.ascontiguousarray(S.reshape(-1, order="F")))
        _lib.call("vqc_dm_apply_superop", d.handle, N, _lib.i32ptr(pos), Sf, 0)
        _lib.call("vqc_dm_apply_superop", d.handle, N, _lib.i32ptr(pos), Sf, 2)
        assert np.max(np.abs(vec(d.storage()) - rho)) < 1e-10
        ref = oracle.dm_apply_superop(rho.copy(), n, N, pos, np.linalg.inv(S), 0)
        d = vq.DensityMatrix(rho)
        _lib.call("vqc_dm_apply_superop", d.handle, N, _lib.i32ptr(pos), Sf, 2)
        assert np.max(np.abs(vec(d.storage()) - ref)) < 1e-10
    rho = vec(rand_density(n, rng))
    d = vq.DensityMatrix(rho)
    for k, pos in [(1, (2,)), (2, (0, 5)), (3, (1, 3, 4))]:
        M = rand_matrix(k, rng)
        out = np.zeros(2)
        _lib.call("vqc_dm_expec_dense", d.handle, k, _lib.i32ptr(pos), _lib.dptr(np.ascontiguousarray(M.reshape(-1, order="F"))),
                  _lib.dptr(out))
        assert abs(complex(*out) - oracle.dm_expec(rho, n, k, pos, M)) < AMP_TOL


def test_hamiltonian_times_density_matrix():
    # (H * rho): each key acts on the row bits of vec(rho) (apply_qterms_dm.jl:11-27)
    rng = np.random.default_rng(29)
    n = 5
    H = tfim(n, 0.8) + vq.QubitsOperator([vq.QubitsTerm({1: "+", 4: "-"}, coeff=0.3),
                                          vq.QubitsTerm({2: "X", 3: "Y", 5: "Z"}, coeff=0.6)])
    terms, nt, hpool = pack_terms(H)
    Hf = np.zeros((1 << n, 1 << n), dtype=complex)
    for j in range(nt):
        k = terms[j].nq
        D = 1 << k
        M = hpool[terms[j].mat:terms[j].mat + D * D].reshape(D, D, order="F")
        Hf += full_op(n, [terms[j].pos[q] for q in range(k)], M)
    rho = rand_density(n, rng)
    out = vq.apply_operator(H, vq.DensityMatrix(vec(rho)))
    assert out.is_dm
    assert np.max(np.abs(out.storage() - Hf @ rho)) < AMP_TOL


@pytest.mark.parametrize("n,tiles", [(10, "small"), (16, "tiled"), (18, "tiled")])
def test_fused_circuit_forward(n, tiles):
    rng = np.random.default_rng(n)
    c = config1_circuit(seed=n, n=n, layers=3)
    c.append(vq.QuantumGate((1, 3, n - 2, n), rand_unitary(4, rng)))
    c.append(vq.QuantumGate((2, 4, 5, n - 1, n), rand_unitary(5, rng)))
    v = rand_state(n, rng)
    s = sv_from(v)
    vq.apply_(c, s)
    prog = Program(c, n=n)
    ref = oracle.apply_circuit(v.copy(), n, 0, prog.op_array(), len(prog.ops), prog.pool())
    assert np.max(np.abs(s.storage() - ref)) < 1e-11


def _sv_energy_grad(c, H, n, psi0, fused):
    s = sv_from(psi0)
    E, g, out = vq.energy_and_gradient(H, c, s, fused=fused, return_state=True)
    prog = Program(c, n=n)
    terms, nt, hpool = pack_terms(H)
    Er, gr, psi_r = oracle.energy_grad(psi0, n, 0, prog.op_array(), len(prog.ops), prog.pool(), terms, nt, hpool,
                                       prog.nparams)
    return E, g, out.storage(), Er, gr, psi_r


@pytest.mark.parametrize("fused", [True, False])
def test_config0_energy_gradient(fused):
    c, H, _ = config0()
    psi0 = np.zeros(1 << 12, dtype=complex)
    psi0[0] = 1
    E, g, psi, Er, gr, psi_r = _sv_energy_grad(c, H, 12, psi0, fused)
    assert np.max(np.abs(psi - psi_r)) < 1e-12
    assert abs(E - Er) < 1e-10 * max(1, abs(Er))
    assert len(g) == 144 and np.max(np.abs(g - gr)) < 1e-10


@pytest.mark.parametrize("n,fused", [(7, True), (14, True), (17, True), (14, False)])
def test_general_energy_gradient(n, fused):
    # 4- and 5-qubit gates, multi-parameter gates, inactive parameters, non-Hermitian H
    rng = np.random.default_rng(40 + n)
    c = hea_circuit(n, 2, rng.uniform(0, 6.3, 6 * n))
    c.insert(3, vq.FSIMGate(2, n, 0.3, 1.2))
    c.insert(9, vq.CRyGate(n - 1, 1, 0.8))
    c.insert(10, vq.RxGate(2, 0.4, isparas=False))
    c.append(vq.QuantumGate((1, 2, n - 1, n), rand_unitary(4, rng)))
    c.append(vq.QuantumGate((1, 3, 4, n - 2, n), rand_unitary(5, rng)))
    c.append(vq.PHASEGate(n // 2, 0.9))
    H = heisenberg(n, rng.standard_normal(n - 1)) + vq.QubitsOperator(
        [vq.QubitsTerm({1: "+", n: "-"}, coeff=0.7), vq.QubitsTerm({2: "X", 3: "Y", n: "Z"}, coeff=0.2)])
    psi0 = rand_state(n, rng)
    E, g, psi, Er, gr, psi_r = _sv_energy_grad(c, H, n, psi0, fused)
    assert np.max(np.abs(psi - psi_r)) < 1e-12
    assert abs(E - Er) < 1e-10 * max(1, abs(Er))
    assert np.max(np.abs(g - gr)) < 1e-10 * max(1, np.max(np.abs(gr)))


@pytest.mark.parametrize("n", [6, 12, 16])
def test_same_qubit_runs_energy_gradient(n):
    rng = np.random.default_rng(70 + n)
    c = chain_circuit(n, rng)
    H = tfim(n, 0.6)
    psi0 = rand_state(n, rng)
    E, g, psi, Er, gr, psi_r = _sv_energy_grad(c, H, n, psi0, True)
    assert np.max(np.abs(psi - psi_r)) < 1e-12
    assert abs(E - Er) < 1e-10 * max(1, abs(Er))
    assert np.max(np.abs(g - gr)) < 1e-10 * max(1, np.max(np.abs(gr)))
    s = sv_from(psi0)
    vq.apply_(c, s)
    assert np.max(np.abs(s.storage() - psi_r)) < 1e-12


@pytest.mark.parametrize("n", [6, 12, 16])
def test_structured_gates_energy_gradient(n):
    rng = np.random.default_rng(90 + n)
    c = structured_circuit(n, rng)
    H = tfim(n, 0.8)
    psi0 = rand_state(n, rng)
    E, g, psi, Er, gr, psi_r = _sv_energy_grad(c, H, n, psi0, True)
    assert np.max(np.abs(psi - psi_r)) < 1e-12
    assert abs(E - Er) < 1e-10 * max(1, abs(Er))
    assert np.max(np.abs(g - gr)) < 1e-10 * max(1, np.max(np.abs(gr)))


@pytest.mark.parametrize("n,fused", [(3, True), (5, True), (7, True), (5, False)])
def test_dm_noisy_energy_gradient(n, fused):
    c, H, _ = config3(seed=n, n=n, depth=2, p=0.05, gamma=0.03)
    if n == 5:
        H = H + vq.QubitsOperator([vq.QubitsTerm({1: "+", 4: "-"}, coeff=0.7), vq.QubitsTerm({2: "X", 5: "Y"}, coeff=0.36)])
    rho0 = np.zeros(1 << (2 * n), dtype=complex)
    rho0[0] = 1
    d = vq.DensityMatrix(rho0)
    E, g, out = vq.energy_and_gradient(H, c, d, fused=fused, return_state=True)
    prog = Program(c, n=n, dm=True)
    terms, nt, hpool = pack_terms(H)
    Er, gr, rho_r = oracle.energy_grad(rho0, n, 1, prog.op_array(), len(prog.ops), prog.pool(), terms, nt, hpool,
                                       prog.nparams)
    assert np.max(np.abs(vec(out.storage()) - rho_r)) < 1e-12
    assert abs(E - Er) < 1e-10
    assert np.max(np.abs(g - gr)) < 1e-10 * max(1, np.max(np.abs(gr)))


def test_errors_are_reported_not_crashes():
    s = vq.StateVector(5)
    U = np.ascontiguousarray(np.eye(2, dtype=complex).reshape(-1))
    with pytest.raises(vq.VQCError, match="out of range"):
        _lib.call("vqc_sv_apply_dense", s.handle, 1, _lib.i32ptr([5]), _lib.dptr(U))
    with pytest.raises(vq.VQCError, match="ascending"):
        _lib.call("vqc_sv_apply_dense", s.handle, 2, _lib.i32ptr([3, 1]), _lib.dptr(np.eye(4, dtype=complex).reshape(-1)))
    with pytest.raises(vq.VQCError):
        _lib.call("vqc_sv_apply_dense", s.handle, 6, _lib.i32ptr([0, 1, 2, 3, 4, 5]), _lib.dptr(U))
    with pytest.raises(ValueError):
        vq.apply_(vq.RxGate(6, 0.1), s)
    # the state is untouched by rejected calls
    assert np.array_equal(s.storage(), np.eye(1, 32, 0, dtype=complex).reshape(-1))
    d = vq.DensityMatrix(3)
    S0 = np.zeros(16, dtype=complex)
    with pytest.raises(vq.VQCError, match="singular"):
        _lib.call("vqc_dm_apply_superop", d.handle, 1, _lib.i32ptr([1]), _lib.dptr(S0), 2)
    with pytest.raises(vq.VQCError, match="mode"):
        _lib.call("vqc_dm_apply_superop", d.handle, 1, _lib.i32ptr([1]), _lib.dptr(S0), 3)
    with pytest.raises(vq.VQCError, match="density matrix"):
        _lib.call("vqc_dm_apply_superop", s.handle, 1, _lib.i32ptr([1]), _lib.dptr(S0), 0)


@pytest.mark.parametrize("n,reverse_ladder", [(12, False), (12, True), (16, False), (16, True)])
def test_ansatz_block_shapes_energy_gradient(n, reverse_ladder):
    # Rx Ry Rz on every qubit, then a CNOT ladder: the reverse sweep runs the
    # straight-line block for CNOT^dagger + Rz/Ry/Rx^dagger on the CNOT's target.
    # CNOT(i, i+1) ascending puts the target on the pair's high bit; CNOT(i+1, i)
    # descending puts it on the low bit (the mirrored shape).
    rng = np.random.default_rng(1200 + n + reverse_ladder)
    c = vq.QCircuit()
    for _ in range(3):
        for q in range(1, n + 1):
            c.append(vq.RxGate(q, float(rng.uniform(0, 6.3))))
            c.append(vq.RyGate(q, float(rng.uniform(0, 6.3))))
            c.append(vq.RzGate(q, float(rng.uniform(0, 6.3))))
        ladder = [(q + 1, q) for q in range(n - 1, 0, -1)] if reverse_ladder else [(q, q + 1) for q in range(1, n)]
        for a, b in ladder:
            c.append(vq.CNOTGate(a, b))
    H = heisenberg(n, rng.standard_normal(n - 1))
    psi0 = rand_state(n, rng)
    E, g, psi, Er, gr, psi_r = _sv_energy_grad(c, H, n, psi0, True)
    assert np.max(np.abs(psi - psi_r)) < 1e-12
    assert abs(E - Er) < 1e-10 * max(1, abs(Er))
    assert np.max(np.abs(g - gr)) < 1e-10 * max(1, np.max(np.abs(gr)))


@pytest.mark.parametrize("n,reverse_ladder", [(5, False), (5, True), (7, True)])
def test_dm_ansatz_block_shapes_energy_gradient(n, reverse_ladder):
    # the noisy layer of the density-matrix workload (Depolarizing after each CNOT,
    # AmplitudeDamping per qubit per layer): its reverse sweep runs the straight-line
    # block for a CNOT row (or column) uncompute, then Rz^dagger, Ry^dagger and
    # Rx^dagger with their gradient terms on the CNOT's target, in both orientations
    # of the ladder
    rng = np.random.default_rng(1300 + n + reverse_ladder)
    c = vq.QCircuit()
    for _ in range(2):
        for q in range(1, n + 1):
            c.append(vq.RxGate(q, float(rng.uniform(0, 6.3))))
            c.append(vq.RyGate(q, float(rng.uniform(0, 6.3))))
            c.append(vq.RzGate(q, float(rng.uniform(0, 6.3))))
        ladder = [(q + 1, q) for q in range(n - 1, 0, -1)] if reverse_ladder else [(q, q + 1) for q in range(1, n)]
        for a, b in ladder:
            c.append(vq.CNOTGate(a, b))
            c.append(vq.Depolarizing(b, 0.02))
        for q in range(1, n + 1):
            c.append(vq.AmplitudeDamping(q, 0.01))
    H = heisenberg(n, rng.standard_normal(n - 1))
    rho0 = vec(rand_density(n, rng))
    d = vq.DensityMatrix(rho0)
    E, g, out = vq.energy_and_gradient(H, c, d, fused=True, return_state=True)
    prog = Program(c, n=n, dm=True)
    terms, nt, hpool = pack_terms(H)
    Er, gr, rho_r = oracle.energy_grad(rho0, n, 1, prog.op_array(), len(prog.ops), prog.pool(), terms, nt, hpool,
                                       prog.nparams)
    assert np.max(np.abs(vec(out.storage()) - rho_r)) < 1e-12
    assert abs(E - Er) < 1e-10 * max(1, abs(Er))
    assert np.max(np.abs(g - gr)) < 1e-10 * max(1, np.max(np.abs(gr)))


def mixed_unitary_channel(positions, rng, k=3):
    """sum_i w_i U_i rho U_i^dagger with random unitaries: a dense superoperator, so
    it has no parity-block form (unlike depolarizing and amplitude damping)."""
    w = rng.uniform(0.2, 1.0, k)
    w /= w.sum()
    return vq.QuantumMap(positions, [np.sqrt(x) * rand_unitary(len(positions), rng) for x in w])


@pytest.mark.parametrize("n", [4, 6])
def test_dm_general_and_parity_channels(n):
    # the fused path takes the parity-block sweep for depolarizing / amplitude damping
    # and the general 4 x 4 and 16 x 16 superoperator sweeps for the other channels
    rng = np.random.default_rng(1400 + n)
    c = vq.QCircuit()
    for layer in range(2):
        for q in range(1, n + 1):
            c.append(vq.RyGate(q, float(rng.uniform(0, 6.3))))
            c.append(vq.RzGate(q, float(rng.uniform(0, 6.3))))
        for q in range(1, n):
            c.append(vq.CNOTGate(q, q + 1))
            c.append(vq.Depolarizing(q + 1, 0.03) if q % 2 else mixed_unitary_channel((q + 1,), rng))
        c.append(mixed_unitary_channel((1, n), rng))
        c.append(vq.AmplitudeDamping(2, 0.05))
    H = heisenberg(n, rng.standard_normal(n - 1))
    rho0 = vec(rand_density(n, rng))
    prog = Program(c, n=n, dm=True)
    terms, nt, hpool = pack_terms(H)
    Er, gr, rho_r = oracle.energy_grad(rho0, n, 1, prog.op_array(), len(prog.ops), prog.pool(), terms, nt, hpool,
                                       prog.nparams)
    d = vq.DensityMatrix(rho0)
    vq.apply_(c, d)
    assert np.max(np.abs(vec(d.storage()) - rho_r)) < 1e-12
    E, g, out = vq.energy_and_gradient(H, c, vq.DensityMatrix(rho0), fused=True, return_state=True)
    assert np.max(np.abs(vec(out.storage()) - rho_r)) < 1e-12
    assert abs(E - Er) < 1e-10 * max(1, abs(Er))
    assert np.max(np.abs(g - gr)) < 1e-10 * max(1, np.max(np.abs(gr)))


@pytest.mark.parametrize("n", [20, 21])
def test_energy_gradient_in_the_larger_tile_bands(n):
    # single states of 19-20 qubits take 11-bit tiles, above that 12 forward / 11
    # reverse: one shallow layered circuit in each band against the oracle
    rng = np.random.default_rng(1500 + n)
    c = hea_circuit(n, 1, rng.uniform(0, 6.3, 3 * n))
    c.append(vq.QuantumGate((2, n // 2, n), rand_unitary(3, rng)))
    H = tfim(n, 0.7)
    psi0 = rand_state(n, rng)
    E, g, psi, Er, gr, psi_r = _sv_energy_grad(c, H, n, psi0, True)
    assert np.max(np.abs(psi - psi_r)) < 1e-12
    assert abs(E - Er) < 1e-10 * max(1, abs(Er))
    assert np.max(np.abs(g - gr)) < 1e-10 * max(1, np.max(np.abs(gr)))


@pytest.mark.parametrize("n", [9, 10])
def test_dm_noisy_energy_gradient_in_the_larger_tile_bands(n):
    # vec(rho) of 18 bits takes 10-bit tiles, of 20 bits 11-bit tiles; a mixed
    # initial state so every entry of the reverse pass is non-trivial
    rng = np.random.default_rng(1600 + n)
    c, H, _ = config3(seed=n, n=n, depth=2, p=0.05, gamma=0.03)
    rho0 = vec(rand_density(n, rng))
    E, g, out = vq.energy_and_gradient(H, c, vq.DensityMatrix(rho0), fused=True, return_state=True)
    prog = Program(c, n=n, dm=True)
    terms, nt, hpool = pack_terms(H)
    Er, gr, rho_r = oracle.energy_grad(rho0, n, 1, prog.op_array(), len(prog.ops), prog.pool(), terms, nt, hpool,
                                       prog.nparams)
    assert np.max(np.abs(vec(out.storage()) - rho_r)) < 1e-12
    assert abs(E - Er) < 1e-10
    assert np.max(np.abs(g - gr)) < 1e-10 * max(1, np.max(np.abs(gr)))


@pytest.mark.parametrize("fused", [True, False])
def test_dm_energy_gradient_with_large_and_multi_parameter_gates(fused):
    # 3-, 4- and 5-qubit dense gates act on both sides of vec(rho) (up to 10 of its
    # 12 bits), between parametric gates with 1-3 generators and a channel
    n = 6
    rng = np.random.default_rng(1700 + fused)
    c = hea_circuit(n, 1, rng.uniform(0, 6.3, 3 * n))
    c.append(vq.QuantumGate((1, 3, 6), rand_unitary(3, rng)))
    c.append(vq.FSIMGate(2, 5, *rng.uniform(0, 6.3, 2)))
    c.append(vq.QuantumGate((1, 2, 4, 6), rand_unitary(4, rng)))
    c.append(vq.Depolarizing(3, 0.07))
    c.append(vq.CRxGate(6, 1, float(rng.uniform(0, 6.3))))
    c.append(vq.QuantumGate((1, 2, 3, 5, 6), rand_unitary(5, rng)))
    c.append(vq.RyGate(4, float(rng.uniform(0, 6.3))))
    H = heisenberg(n, rng.standard_normal(n - 1)) + vq.QubitsOperator([vq.QubitsTerm({1: "X", 4: "Y", 6: "Z"}, coeff=0.4)])
    rho0 = vec(rand_density(n, rng))
    E, g, out = vq.energy_and_gradient(H, c, vq.DensityMatrix(rho0), fused=fused, return_state=True)
    prog = Program(c, n=n, dm=True)
    terms, nt, hpool = pack_terms(H)
    Er, gr, rho_r = oracle.energy_grad(rho0, n, 1, prog.op_array(), len(prog.ops), prog.pool(), terms, nt, hpool,
                                       prog.nparams)
    assert prog.nparams == 3 * n + 2 + 1 + 1
    assert np.max(np.abs(vec(out.storage()) - rho_r)) < 1e-12
    assert abs(E - Er) < 1e-10
    assert np.max(np.abs(g - gr)) < 1e-10 * max(1, np.max(np.abs(gr)))


@pytest.mark.parametrize("n", [3, 5])
def test_channel_on_state_vector_continues_on_its_density_matrix(n):
    """apply!(circuit, psi) applies element by element (generic.jl:26-31) and a
    channel on a state vector continues on DensityMatrix(psi) (apply_qmaps.jl:19):
    the result, and the energy gradient, equal the whole circuit on |psi><psi|."""
    rng = np.random.default_rng(60 + n)
    c, H, _ = config3(seed=n, n=n, depth=2, p=0.05, gamma=0.03)
    first = next(i for i, g in enumerate(c) if isinstance(g, vq.QuantumMap))
    assert first > 0
    psi = rand_state(n, rng)
    rho = np.outer(psi, psi.conj())
    out = vq.apply_(c, vq.StateVector(psi))
    assert out.is_dm
    prog = Program(c, n=n, dm=True)
    ref = oracle.apply_circuit(vec(rho), n, 1, prog.op_array(), len(prog.ops), prog.pool())
    assert np.max(np.abs(vec(out.storage()) - ref)) < 1e-12
    # the device DensityMatrix(psi) itself: kron(conj(psi), psi) (densitymatrix.jl:27)
    assert np.max(np.abs(vec(vq.DensityMatrix(vq.StateVector(psi)).storage()) - np.kron(psi.conj(), psi))) < 1e-15
    E, g = vq.energy_and_gradient(H, c, vq.StateVector(psi))
    E2, g2 = vq.energy_and_gradient(H, c, vq.DensityMatrix(rho))
    assert abs(E - E2) < 1e-12 and np.max(np.abs(g - g2)) < 1e-12
