"""Parity at the measured workload sizes (the shapes the benchmark times).

* 256-instance uneven batch (8..20 qubits, depth 2..12, Heisenberg): every
  energy and gradient of the persistent-kernel batch against the oracle's,
  frozen in tests/golden/config2_energy_grad.npz (minutes of CPU work, made by
  tests/golden/make_golden.py configs; the fixture holds a digest of the packed
  inputs, checked here).
* 12-qubit noisy density matrix, 6 layers, 216 angles: energy and gradient with
  the default and the autotuned tile shape against the frozen oracle result, and
  against 5-point central differences of the device energy on 4 angles (the
  method of test/check_noisy_grad.jl:3-60, central_fdm(5, 1) of test/util.jl:8).
* 28 qubits, 20 layers of Haar gates: the fused and one-kernel-per-gate paths
  agree over the whole 480-gate circuit, the norm is kept, the inverse circuit
  returns |0...0>, and the 200-term Pauli energy of a product state equals the
  closed form prod_q <u_q|sigma_q|u_q> term by term.
* Serving: requests of up to 26 qubits and depth 40 through the server equal the
  plain path; those of up to 16 qubits also equal the oracle.

Tolerances (FP64): energies and gradients 1e-10 relative to max(1, |value|);
amplitudes 1e-10 absolute after 960 gates; Pauli terms 1e-12 relative.
"""
import hashlib
import os

import numpy as np
import pytest

import oracle
from brute import rand_unitary
import vqchip as vq
from vqchip import _lib
from vqchip.batch import Batch
from vqchip.circuit import Program
from vqchip.operators import pack_pauli, pack_terms
from vqchip.serve import Request, Server, plain_eval
from vqchip.workloads import config1_circuit, config2_instances, config3, hea_circuit, random_pauli_operator, tfim

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b)) / max(1.0, np.max(np.abs(b))))


def _digest(packed):
    from golden.make_golden import op_table, term_table
    h = hashlib.sha256()
    for prog, terms, nt, hpool in packed:
        h.update(op_table(prog).tobytes())
        h.update(np.ascontiguousarray(prog.pool()).tobytes())
        h.update(term_table(terms, nt).tobytes())
        h.update(np.ascontiguousarray(hpool).tobytes())
    return h.hexdigest()


def test_config2_every_instance_against_the_oracle():
    fx = np.load(os.path.join(GOLDEN, "config2_energy_grad.npz"), allow_pickle=False)
    inst = config2_instances()
    packed = []
    for c, H, n in inst:
        terms, nt, hpool = pack_terms(H)
        packed.append((Program(c, n=n), terms, nt, hpool))
    assert _digest(packed) == str(fx["inputs_sha256"]), "the seeded generator no longer makes the fixture's inputs"
    assert list(fx["nqubits"]) == [n for _, _, n in inst]
    E, grads = Batch(inst).run()
    off = fx["grad_offsets"]
    worst = 0.0
    for i in range(len(inst)):
        gr = fx["grad"][off[i]:off[i + 1]]
        worst = max(worst, _rel_err(E[i], fx["energy"][i]), _rel_err(grads[i], gr))
    print(f"config2: 256 instances, max error vs the oracle {worst:.2e}", flush=True)
    assert worst < 1e-10


def _fd(H, c, rho0, k, h=1e-3):
    import copy
    theta = np.array(c.active_parameters(), dtype=float)

    def E(dt):
        th = theta.copy()
        th[k] += dt
        cc = copy.deepcopy(c)
        cc.reset_parameters(th)
        return vq.expectation(H, cc * rho0).real

    return (-E(2 * h) + 8 * E(h) - 8 * E(-h) + E(-2 * h)) / (12 * h)


@pytest.mark.parametrize("tiles", [None, (12, 11, 2)])
def test_config3_full_size(tiles):
    fx = np.load(os.path.join(GOLDEN, "config3_energy_grad.npz"), allow_pickle=False)
    c, H, _ = config3()
    prog = Program(c, n=12, dm=True)
    terms, nt, hpool = pack_terms(H)
    from golden.make_golden import inputs_digest
    assert inputs_digest([(prog, terms, nt, hpool)]) == str(fx["inputs_sha256"])
    from vqchip import autotune
    before = autotune.get_tuning()
    try:
        if tiles:
            autotune.set_tuning(tile_fwd_bits=tiles[0], tile_rev_bits=tiles[1], tile_low_bits=tiles[2])
        rho0 = vq.DensityMatrix(12)
        E, g = vq.energy_and_gradient(H, c, rho0)
        eE, eg = _rel_err(E, fx["energy"]), _rel_err(g, fx["grad"])
        print(f"config3 tiles={tiles}: |dE| {eE:.2e}, max |dgrad| {eg:.2e} vs the oracle", flush=True)
        assert eE < 1e-10 and eg < 1e-10
        if tiles is None:
            for k in (0, 71, 143, 215):
                fd = _fd(H, c, rho0, k)
                assert abs(g[k] - fd) < 1e-8, (k, g[k], fd)
    finally:
        autotune.set_tuning(**{k: before[k] for k in autotune.FIELDS})


def test_config1_full_circuit_fused_plain_inverse_and_pauli_closed_form():
    n = 28
    circuit = config1_circuit(n=n, layers=20)
    assert len(circuit) == 480
    fused = vq.StateVector(n)
    vq.apply_(circuit, fused)
    plain = vq.StateVector(n)
    for g in circuit:
        pos = g.ordered_positions()
        U = np.ascontiguousarray(g.ordered_mat().reshape(-1, order="F"))
        _lib.call("vqc_sv_apply_dense", plain.handle, len(pos), _lib.i32ptr([q - 1 for q in pos]), _lib.dptr(U))
    d = fused.copy()
    _lib.call("vqc_state_axpy", d.handle, -1.0, 0.0, plain.handle)
    diff = np.sqrt(d.dot(d).real)
    del d
    assert diff < 1e-10, diff
    assert abs(fused.norm() - 1.0) < 1e-12
    del plain
    # the inverse circuit brings the state back to |0...0>
    vq.apply_(circuit.adjoint(), fused)
    basis = vq.StateVector(n)
    a0 = basis.dot(fused)
    assert abs(a0 - 1.0) < 1e-10, a0
    del basis
    # 200 random Pauli strings on a product state: <P> = c prod_q <u_q|sigma_q|u_q>
    rng = np.random.default_rng(281)
    prod = vq.StateVector(n)
    us = []
    for q in range(n):
        U = rand_unitary(1, rng)
        us.append(U[:, 0])
        _lib.call("vqc_sv_apply_dense", prod.handle, 1, _lib.i32ptr([q]), _lib.dptr(np.ascontiguousarray(U.reshape(-1, order="F"))))
    Hp = random_pauli_operator(n, 200)
    arr, nt = pack_pauli(Hp)
    per = np.zeros(2 * nt)
    s2 = np.zeros(2)
    _lib.call("vqc_sv_expec_pauli", prod.handle, None, nt, arr, _lib.dptr(per), _lib.dptr(s2))
    sig = {"X": np.array([[0, 1], [1, 0]], complex), "Y": np.array([[0, -1j], [1j, 0]]), "Z": np.diag([1.0 + 0j, -1]),
           "I": np.eye(2, dtype=complex)}
    closed = []
    for t in Hp.terms:
        v = t.coeff
        for p, l in zip(t.positions, t.labels):
            u = us[p - 1]
            v *= u.conj() @ sig[l] @ u
        closed.append(v)
    closed = np.array(closed)
    got = per[0::2] + 1j * per[1::2]
    err = np.max(np.abs(got - closed) / np.maximum(1.0, np.abs(closed)))
    print(f"config1: fused-plain {diff:.2e}, inverse |<0|psi>-1| {abs(a0 - 1):.2e}, Pauli closed form {err:.2e}", flush=True)
    assert err < 1e-12
    assert abs(complex(*s2) - closed.sum()) < 1e-12 * max(1.0, np.sum(np.abs(closed)))


def test_config4_large_requests_served_equal_the_plain_path():
    rng = np.random.default_rng(44)
    shapes = [(26, 40), (24, 31), (22, 17), (20, 40), (16, 40), (14, 25), (12, 9), (10, 4)]
    reqs = []
    for n, d in shapes:
        c = hea_circuit(n, d, rng.uniform(0, 2 * np.pi, 3 * n * d))
        reqs.append(Request(c, tfim(n, float(rng.uniform(0.5, 1.5))), n))
    with Server(hbm_budget=64 << 30) as srv:
        tickets = [srv.submit(r) for r in reqs]
        res = [srv.wait(t) for t in tickets]
    worst_p, worst_o = 0.0, 0.0
    for r, (E, g, _) in zip(reqs, res):
        Ep, gp = plain_eval(r, fused=False)
        worst_p = max(worst_p, _rel_err(E, Ep), _rel_err(g, gp))
        if r.n <= 16:
            psi0 = np.zeros(1 << r.n, dtype=complex)
            psi0[0] = 1
            Er, gr, _ = oracle.energy_grad(psi0, r.n, 0, r.ops, r.nops, r.pool, r.terms, r.nterms, r.hpool, r.nparams)
            worst_o = max(worst_o, _rel_err(E, Er), _rel_err(g, gr))
    print(f"config4: served vs plain {worst_p:.2e}, vs the oracle (n <= 16) {worst_o:.2e}", flush=True)
    assert worst_p < 1e-10 and worst_o < 1e-10
