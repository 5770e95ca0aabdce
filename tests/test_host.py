"""Host-side logic: gate matrices and generators, packing, the library's exported ABI."""
import ctypes
import os
import re

import numpy as np
import pytest

from brute import full_op, rand_matrix
from vqchip import (CNOTGate, CRxGate, FSIMGate, PHASEGate, QCircuit, QuantumGate, QubitsOperator, QubitsTerm, RxGate,
                    RyGate, RzGate, CPHASEGate, CCPHASEGate, Depolarizing, AmplitudeDamping)
from vqchip import _lib
from vqchip.circuit import Program
from vqchip.gates import reorder
from vqchip.operators import pack_pauli, pack_terms
from vqchip.workloads import config0, config1_circuit, config2_instances, config3, hea_circuit

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_reorder_matches_bit_permutation():
    rng = np.random.default_rng(0)
    M = rand_matrix(3, rng)
    given = [5, 1, 3]  # matrix bit b <-> qubit given[b]
    # apply on a 6-qubit space through the given order, compare with the sorted form
    v = rng.standard_normal(64) + 1j * rng.standard_normal(64)
    idx = np.arange(64)
    sub = sum(((idx >> q) & 1) << b for b, q in enumerate(given))
    rest = idx & ~sum(1 << q for q in given)
    F = M[sub[:, None], sub[None, :]] * (rest[:, None] == rest[None, :])
    assert np.allclose(full_op(6, sorted(given), reorder(M, given)) @ v, F @ v)


def test_generators_are_udagger_du():
    h = 1e-6
    for make, nparam in [(lambda p: RxGate(1, *p), 1), (lambda p: RyGate(2, *p), 1), (lambda p: RzGate(1, *p), 1),
                         (lambda p: PHASEGate(1, *p), 1), (lambda p: CRxGate(3, 1, *p), 1),
                         (lambda p: CPHASEGate(1, 2, *p), 1), (lambda p: FSIMGate(2, 1, *p), 2),
                         (lambda p: CCPHASEGate(3, 1, 2, *p), 1)]:
        p0 = list(np.linspace(0.3, 1.1, nparam))
        g = make(p0)
        U = g.ordered_mat()
        assert np.allclose(U.conj().T @ U, np.eye(U.shape[0]), atol=1e-14)
        for j, G in enumerate(g.differentiate()):
            pp, pm = list(p0), list(p0)
            pp[j] += h
            pm[j] -= h
            dU = (make(pp).ordered_mat() - make(pm).ordered_mat()) / (2 * h)
            assert np.allclose(U.conj().T @ dU, G, atol=1e-8), (g, j)


def test_adjoints_and_controlled_gates():
    # gate' is the conjugate transpose; the named gates keep a named kernel (check_gateop.jl:68-158)
    import vqchip as vq
    rng = np.random.default_rng(8)
    U = np.linalg.qr(rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2)))[0]
    gates = [vq.XGate(2), vq.YGate(1), vq.ZGate(3), vq.TGate(2), vq.HGate(1), vq.PHASEGate(2, 0.3), vq.CNOTGate(3, 1),
             vq.CZGate(1, 2), vq.CPHASEGate(4, 1, -0.5), vq.SWAPGate(2, 5), vq.iSWAPGate(2, 5), vq.FSIMGate(1, 3, 0.2, 0.9),
             vq.CRxGate(2, 1, 0.4), vq.TOFFOLIGate(3, 1, 2), vq.FREDKINGate(1, 2, 3), vq.CCPHASEGate(1, 2, 3, 0.7),
             vq.CONTROLGate((2, 1), U), vq.CONTROLCONTROLGate(1, 3, 2, U)]
    for g in gates:
        a = g.adjoint()
        assert a.ordered_positions() == g.ordered_positions()
        assert np.array_equal(a.ordered_mat(), g.ordered_mat().conj().T), g
        assert (a.named_kind is None) == (g.named_kind is None), g
        assert np.array_equal(a.adjoint().ordered_mat(), g.ordered_mat()), g
    assert vq.iSWAPGate(1, 2).adjoint().named_kind == _lib.GATE_ISWAPDAG
    assert vq.PHASEGate(1, 0.3).adjoint().params == (-0.3,) and vq.PHASEGate(1, 0.3).adjoint().nparameters() == 0
    X = np.array([[0, 1], [1, 0]])
    assert np.array_equal(vq.CONTROLGate((3, 1), X).ordered_mat(), vq.CNOTGate(3, 1).ordered_mat())
    assert np.array_equal(vq.CONTROLCONTROLGate(2, 4, 1, X).ordered_mat(), vq.TOFFOLIGate(2, 4, 1).ordered_mat())
    # CONTROLCONTROL(U) acts as U on the target exactly when both controls are 1
    M = vq.CONTROLCONTROLGate(1, 2, 3, U).mat()
    for a in (0, 1):
        for b in (0, 1):
            blk = M[np.ix_([a + 2 * b, a + 2 * b + 4], [a + 2 * b, a + 2 * b + 4])]
            assert np.array_equal(blk, U if a and b else np.eye(2))


def test_from_external_is_the_term_operator():
    # check_high_qubit_gate.jl:7-9: from_external(pos, kron(m_1, ..., m_k)) == QubitsTerm(pos[i] => m_i)
    import vqchip as vq
    rng = np.random.default_rng(9)
    for pos in [(3, 1, 4, 2), (2, 5, 1, 4, 3), (2, 1)]:
        ms = [np.linalg.qr(rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2)))[0] for _ in pos]
        K = ms[0]
        for m in ms[1:]:
            K = np.kron(K, m)
        g = vq.from_external(pos, K)
        t = QubitsTerm(dict(zip(pos, ms)))
        n = max(pos)
        A = full_op(n, [q - 1 for q in g.ordered_positions()], g.ordered_mat())
        B = full_op(n, [q - 1 for q in t.positions], t.matrix())
        assert np.max(np.abs(A - B)) < 1e-14


def test_inactive_parameters_and_slots():
    # test/parameter.jl: isparas=false parameters are not counted and get no slot
    c = QCircuit([RxGate(1, 0.1), RxGate(1, 1.1, isparas=False), CNOTGate(1, 2), RyGate(2, 0.2)])
    assert c.nparameters() == 2 and c.active_parameters() == [0.1, 0.2]
    prog = Program(c, n=2)
    assert prog.nparams == 2
    assert [o.slot[0] for o in prog.ops] == [0, -1, -1, 1]
    c.reset_parameters([0.5, 0.6])
    assert c.active_parameters() == [0.5, 0.6] and c[1].params == (1.1,)


def test_operator_merge_and_packing():
    H = QubitsOperator([QubitsTerm({1: "Z", 2: "Z"}), QubitsTerm({1: "X"}, coeff=0.5),
                        QubitsTerm({1: "Y", 2: "Y"}, coeff=2.0)])
    data = H.data()
    assert list(data) == [(1, 2), (1,)]
    Z, Y = np.diag([1, -1]), np.array([[0, -1j], [1j, 0]])
    assert np.allclose(data[(1, 2)], np.kron(Z, Z) + 2 * np.kron(Y, Y))
    terms, nt, pool = pack_terms(H)
    assert nt == 2 and terms[0].nq == 2 and list(terms[0].pos)[:2] == [0, 1]
    # first site is the least significant bit: QubitsTerm(1=>"X", 2=>"Z") = kron(Z, X)
    t = QubitsTerm({1: "X", 2: "Z"})
    assert np.allclose(t.matrix(), np.kron(Z, np.array([[0, 1], [1, 0]])))
    arr, n = pack_pauli(QubitsOperator([QubitsTerm({1: "X", 3: "Y", 4: "Z"}, coeff=0.5)]))
    assert (arr[0].x, arr[0].z, arr[0].re) == (0b0101, 0b1100, 0.5)


def test_superoperator_layout():
    ch = Depolarizing(1, 0.2)
    S = ch.ordered_supermat()
    rho = np.array([[0.7, 0.1 - 0.2j], [0.1 + 0.2j, 0.3]])
    out = (S @ rho.reshape(-1, order="F")).reshape(2, 2, order="F")
    ref = sum(K @ rho @ K.conj().T for K in ch.kraus)
    assert np.allclose(out, ref)
    ad = AmplitudeDamping(1, 0.3)
    assert np.allclose(sum(K.conj().T @ K for K in ad.kraus), np.eye(2))


def test_workload_shapes():
    c0, H0, th = config0()
    assert len(c0) == 188 and len(th) == 144 and len(H0.data()) == 23
    c1 = config1_circuit(layers=2)
    assert len(c1) == 48 and sorted({len(g.positions) for g in c1}) == [1, 2, 3]
    inst = config2_instances(count=4)
    assert all(8 <= n <= 20 for _, _, n in inst)
    c3, H3, th3 = config3(n=4, depth=2)
    assert len(th3) == 24 and sum(1 for g in c3 if g.nparameters() == 0 and not hasattr(g, "kraus")) == 6


def _header_functions():
    text = open(os.path.join(ROOT, "include", "vqc_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|const char\*)\s+(vqc_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib_path = _lib.LIB_PATH
    if not os.path.exists(lib_path):
        pytest.skip("libvqc_hip.so not built")
    lib = ctypes.CDLL(lib_path)  # loads without a GPU; no compute call is made
    names = _header_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.SIGNATURES) | {"vqc_last_error"} >= set(names) - set()


def test_tuning_roundtrip_and_validation():
    # vqc_set_tuning / vqc_get_tuning are host-only: no device is touched
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libvqc_hip.so not built")
    from vqchip import autotune
    before = autotune.get_tuning()
    try:
        autotune.set_tuning(tile_fwd_bits=11, tile_low_bits=4, dense_groups=[2, -1, -1])
        t = autotune.get_tuning()
        assert t["tile_fwd_bits"] == 11 and t["tile_low_bits"] == 4 and t["dense_groups"][0] == 2
        assert t["tile_rev_bits"] == before["tile_rev_bits"]  # -1 keeps the value
        for bad in ({"tile_fwd_bits": 14}, {"tile_rev_bits": 13}, {"tile_low_bits": 12}, {"dense_groups": [3, -1, -1]},
                    {"dense_nontemporal": [2, -1, -1]}, {"tile_small_bits": 13}, {"tile_fwd_bits": 5}):
            with pytest.raises(_lib.VQCError):
                autotune.set_tuning(**bad)
        assert autotune.get_tuning()["tile_fwd_bits"] == 11  # a rejected call changes nothing
    finally:
        autotune.set_tuning(**{k: before[k] for k in autotune.FIELDS}, dense_groups=before["dense_groups"],
                            dense_nontemporal=before["dense_nontemporal"])
    assert autotune.get_tuning() == before


def _dense(H, n):
    terms, nt, hpool = pack_terms(H)
    M = np.zeros((1 << n, 1 << n), complex)
    for j in range(nt):
        k = terms[j].nq
        D = 1 << k
        M += full_op(n, [terms[j].pos[q] for q in range(k)], hpool[terms[j].mat:terms[j].mat + D * D].reshape(D, D, order="F"))
    return M


def test_spin_hamiltonian_builders():
    # spin_hamiltonians.jl: 2J(s+ s- + h.c.) = J(XX + YY) with s+- = (X +- iY)/2
    from vqchip.utility import heisenberg_2d, heisenberg_chain, ising_2d, ising_chain
    X = np.array([[0, 1], [1, 0]], complex)
    Y = np.array([[0, -1j], [1j, 0]])
    Z = np.diag([1.0 + 0j, -1.0])
    L, J, Jzz, hz = 4, 0.7, -0.4, 0.3
    ref = np.zeros((16, 16), complex)
    for i in range(L):
        ref += hz * full_op(L, [i], Z)
    for i in range(L - 1):
        for P, c in ((X, J), (Y, J), (Z, Jzz)):
            ref += c * full_op(L, [i, i + 1], np.kron(P, P))
    assert np.allclose(_dense(heisenberg_chain(L, J=J, Jzz=Jzz, hz=hz), L), ref, atol=1e-14)
    ref = sum(1.2 * full_op(3, [i], Z) for i in range(3)) + sum(0.5 * full_op(3, [i, i + 1], np.kron(X, X)) for i in range(2))
    assert np.allclose(_dense(ising_chain(3, J=0.5, hz=1.2), 3), ref, atol=1e-14)
    # 2 x 3 lattice, LinearIndices column-major: bonds (1,3),(3,5),(2,4),(4,6) and (1,2),(3,4),(5,6)
    H = ising_2d(2, 3, J=1.0, hz=0.0)
    keys = sorted(tuple(k) for k in H.data() if len(k) == 2)
    assert keys == [(1, 2), (1, 3), (2, 4), (3, 4), (3, 5), (4, 6), (5, 6)]
    assert len(heisenberg_2d(2, 3).data()) == 6 + 7


def test_variational_circuit_builders():
    from vqchip.utility import nparas, real_variational_circuit_1d, variational_circuit_1d
    th = np.arange(nparas(3, 2), dtype=float)
    c = variational_circuit_1d(3, 2, th)
    names = [type(g).__name__ for g in c]
    assert names[:9] == ["RzGate", "RyGate", "RzGate"] * 3
    cn = [g.positions for g in c if type(g).__name__ == "CNOTGate"]
    assert cn == [(1, 2), (2, 3), (2, 3), (1, 2)]  # odd layers upwards, even layers downwards
    assert Program(c, n=3).nparams == 27
    r = real_variational_circuit_1d(4, 3, np.zeros(nparas(4, 3, real=True)))
    assert sum(type(g).__name__ == "RyGate" for g in r) == 16
    with pytest.raises(ValueError):
        variational_circuit_1d(3, 2, np.zeros(5))


def test_krylov_rejects_non_hermitian_before_any_device_call():
    from vqchip import krylov
    H = QubitsOperator([QubitsTerm({1: "+", 2: "-"}, coeff=1.0)])  # s+ s- alone is not Hermitian
    with pytest.raises(ValueError, match="hermitian"):
        krylov.ground_state(H, 2)


def test_batch_wave_grouping():
    from vqchip.batch import instance_bytes, waves
    assert instance_bytes(12, t_small=12) == 0 and instance_bytes(13, t_small=12) == 32 << 13
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libvqc_hip.so not built")
    inst = [(None, None, n) for n in (13, 15, 9, 14, 16, 12, 15, 13)]
    w = waves(inst, 3 << 19)
    assert [x for g in w for x in g] == inst  # order kept, nothing lost
    for g in w:
        tot = sum(instance_bytes(n) for _, _, n in g)
        assert tot <= 3 << 19 or len(g) == 1
    assert [len(g) for g in w][0] >= 2


def test_shared_matrix_pool_offsets():
    # programs packed into one shared pool (a batch) point at the same matrices as
    # the same programs packed alone, shifted by the pool's length before them
    pool, hpool = _lib.MatrixPool(), _lib.MatrixPool()
    insts = config2_instances(count=6, seed=5)
    for c, H, n in insts:
        start, hstart = pool.size, hpool.size
        shared = Program(c, n=n, pool=pool)
        alone = Program(c, n=n)
        big = pool.array()
        own = alone.pool()
        assert alone.ops and pool.size == start + own.size
        for a, b in zip(shared.ops, alone.ops):
            assert a.mat == b.mat + start
            nm = 1 << (2 * a.nq)
            assert np.array_equal(big[a.mat:a.mat + nm], own[b.mat:b.mat + nm])
            for j in range(a.nparams):
                assert a.gen[j] == b.gen[j] + start
        terms, nt, arr = pack_terms(H, pool=hpool)
        terms1, nt1, arr1 = pack_terms(H)
        assert arr is None and nt == nt1
        hbig = hpool.array()
        for i in range(nt):
            nm = 1 << (2 * terms[i].nq)
            assert terms[i].mat == terms1[i].mat + hstart
            assert np.array_equal(hbig[terms[i].mat:terms[i].mat + nm], arr1[terms1[i].mat:terms1[i].mat + nm])


def test_batch_repack_rejects_structure_changes():
    """A batch is planned from its op records; new angles may change only the
    matrix values (vqchip.batch.repack), anything else raises."""
    from vqchip.batch import repack
    n = 4
    base = hea_circuit(n, 2, np.linspace(0.1, 2.0, 3 * n * 2))
    old = Program(base, n=n)
    size = old.pool().size
    fresh = hea_circuit(n, 2, np.linspace(0.5, 2.5, 3 * n * 2))
    pool = repack([fresh], [n], [len(old.ops)], old.ops, size)
    assert pool.size == size
    # Ry replaced by Rx: the same pool length, another structure hint
    swapped = QCircuit(RxGate(g.positions[0], 0.3) if isinstance(g, RyGate) else g for g in base)
    with pytest.raises(ValueError):
        repack([swapped], [n], [len(old.ops)], old.ops, size)
    # a rotation moved to another qubit
    moved = QCircuit(base)
    moved[0] = RxGate(2, 0.7)
    with pytest.raises(ValueError):
        repack([moved], [n], [len(old.ops)], old.ops, size)
    # one gate fewer
    with pytest.raises(ValueError):
        repack([QCircuit(base[:-1])], [n], [len(old.ops)], old.ops, size)
