"""Writes the golden fixtures of tests/golden/ (numpy .npz, no pickles).

* known_answers.npz: the reference's own known answers for the amplitude layout,
  transcribed as data from test/check_state_init.jl:1-30 (each expected vector is
  the kron expression the reference test states, Gates.ZERO = [1, 0],
  Gates.ONE = [0, 1], v = [sqrt(2)/2, sqrt(2)/2]).
* The other files hold oracle outputs on small seeded inputs (n <= 12). The
  oracle is pinned first by tests/test_oracle.py (brute force, closed forms,
  parameter shift, finite differences); these files then freeze its outputs so
  that the device path and the oracle are both checked against fixed numbers.

Inputs are stored explicitly (states, matrices, the packed op and term tables),
so a fixture does not depend on the generators that made it.

    python tests/golden/make_golden.py           # from the repository root
    python tests/golden/make_golden.py configs   # the measured workloads (minutes)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "vqc.jl_amd"), os.path.join(ROOT, "oracle"), os.path.dirname(HERE)]

import oracle  # noqa: E402
from brute import rand_density, rand_matrix, rand_state, vec  # noqa: E402
from vqchip.circuit import Program  # noqa: E402
from vqchip.operators import pack_pauli, pack_terms  # noqa: E402
from vqchip.workloads import config0, config3, random_pauli_operator  # noqa: E402

OP_COLS = ["kind", "nq", "pos0", "pos1", "pos2", "pos3", "pos4", "nparams", "slot0", "slot1", "slot2", "reserved",
           "mat", "aux", "gen0", "gen1", "gen2"]


def op_table(prog):
    rows = []
    for o in prog.ops:
        rows.append([o.kind, o.nq, *o.pos, o.nparams, *o.slot, o.reserved, o.mat, o.aux, *o.gen])
    return np.array(rows, dtype=np.int64).reshape(-1, len(OP_COLS))


def term_table(terms, nt):
    return np.array([[terms[j].nq, *terms[j].pos, terms[j].mat] for j in range(nt)], dtype=np.int64).reshape(-1, 7)


def pauli_table(arr, nt):
    return (np.array([[arr[j].x, arr[j].z] for j in range(nt)], dtype=np.uint64).reshape(-1, 2),
            np.array([complex(arr[j].re, arr[j].im) for j in range(nt)], dtype=np.complex128))


def known_answers():
    Z, O = np.array([1, 0], complex), np.array([0, 1], complex)
    v = np.array([np.sqrt(2) / 2, np.sqrt(2) / 2], complex)
    k = lambda *xs: np.kron(xs[0], k(*xs[1:])) if len(xs) > 1 else xs[0]
    # (kind, input, expected): onehot bits / qubit angles, check_state_init.jl:2-27
    cases = [
        ("onehot", [1], O), ("onehot", [0], Z), ("onehot", [1, 0], k(Z, O)), ("onehot", [0, 1], k(O, Z)),
        ("onehot", [1, 1], k(O, O)), ("onehot", [1, 0, 0], k(Z, Z, O)),
        ("qubit", [0.5], v), ("qubit", [0.5, 1.0], k(O, v)), ("qubit", [0.0, 0.5], k(v, Z)),
        ("qubit", [0.0, 0.5, 0.5], k(v, v, Z)), ("qubit", [0.0, 0.5, 0.5, 1.0, 0.0], k(Z, O, v, v, Z)),
    ]
    out = {}
    for i, (kind, x, e) in enumerate(cases):
        out[f"c{i}_kind"] = np.array(kind)
        out[f"c{i}_in"] = np.array(x, dtype=float)
        out[f"c{i}_out"] = np.asarray(e, dtype=np.complex128)
    out["ncases"] = np.array(len(cases))
    # check_measure.jl:31-41: post_select((|00> + |11>)/sqrt(2), 1, value) keeps |value>
    bell = (np.sqrt(2) / 2) * (k(Z, Z) + k(O, O))
    out["bell"] = bell
    out["bell_sel0"], out["bell_sel1"] = Z, O
    return out


def main():
    rng = np.random.default_rng(20261203)
    np.savez(os.path.join(HERE, "known_answers.npz"), **known_answers())

    # dense k-qubit gates on a 7-qubit state, every k, mixed target classes
    n = 7
    psi = rand_state(n, rng)
    d = {"psi": psi}
    for k, pos in [(1, (0,)), (1, (6,)), (2, (1, 5)), (3, (0, 3, 6)), (4, (0, 2, 3, 6)), (5, (1, 2, 4, 5, 6))]:
        U = rand_matrix(k, rng)
        key = f"k{k}_" + "_".join(map(str, pos))
        d[key + "_U"] = U
        d[key + "_pos"] = np.array(pos)
        d[key + "_out"] = oracle.apply_dense(psi.copy(), k, list(pos), U)
    np.savez(os.path.join(HERE, "dense_gates_n7.npz"), **d)

    # the 12-qubit HEA energy and adjoint gradient (4 layers, 144 angles, TFIM)
    c, H, _ = config0()
    prog = Program(c, n=12)
    terms, nt, hpool = pack_terms(H)
    psi0 = np.zeros(1 << 12, complex)
    psi0[0] = 1
    E, g, psi_out = oracle.energy_grad(psi0, 12, 0, prog.op_array(), len(prog.ops), prog.pool(), terms, nt, hpool,
                                       prog.nparams)
    np.savez(os.path.join(HERE, "hea12_energy_grad.npz"), ops=op_table(prog), pool=prog.pool(), terms=term_table(terms, nt),
             hpool=hpool, nparams=np.array(prog.nparams), energy=np.array(E), grad=g, psi_out=psi_out)

    # a noisy 6-qubit density matrix: 2 layers, depolarizing and amplitude damping, sum Z_i
    c, H, _ = config3(n=6, depth=2)
    prog = Program(c, n=6, dm=True)
    terms, nt, hpool = pack_terms(H)
    rho0 = np.zeros(1 << 12, complex)
    rho0[0] = 1
    E, g, rho_out = oracle.energy_grad(rho0, 6, 1, prog.op_array(), len(prog.ops), prog.pool(), terms, nt, hpool,
                                       prog.nparams)
    np.savez(os.path.join(HERE, "dm6_noisy_energy_grad.npz"), ops=op_table(prog), pool=prog.pool(),
             terms=term_table(terms, nt), hpool=hpool, nparams=np.array(prog.nparams), energy=np.array(E), grad=g,
             rho_out=rho_out)

    # 24 random Pauli strings on 10 qubits
    n = 10
    psi = rand_state(n, rng)
    arr, nt = pack_pauli(random_pauli_operator(n, 24, seed=5))
    xz, coef = pauli_table(arr, nt)
    per, total = oracle.expec_pauli(psi, arr, nt)
    np.savez(os.path.join(HERE, "pauli_n10.npz"), psi=psi, xz=xz, coef=coef, per_term=per, total=np.array(total),
             h_psi=oracle.apply_pauli_sum(psi, arr, nt))

    # measurement: P(0), measure! with given u, the selected row of post_select
    n = 6
    psi = rand_state(n, rng)
    rho = vec(rand_density(4, rng))
    m = {"psi": psi, "rho4": rho}
    for key0 in (0, 3, 5):
        m[f"p0_{key0}"] = np.array(oracle.prob_zero(psi, n, key0))
        for u in (0.2, 0.8):
            for reset in (0, 1):
                w = psi.copy()
                i, p = oracle.measure_inplace(w, n, key0, u, bool(reset))
                m[f"m_{key0}_{u}_{reset}"] = w
                m[f"m_{key0}_{u}_{reset}_ip"] = np.array([i, p])
        for val in (0, 1):
            m[f"sel_{key0}_{val}"] = oracle.select(psi.copy(), n, key0, val)
    r = rho.copy()
    i, p = oracle.measure_inplace(r, 4, 2, 0.6, True, is_dm=True)
    m["rho4_m"] = r
    m["rho4_m_ip"] = np.array([i, p])
    np.savez(os.path.join(HERE, "measure_n6.npz"), **m)


def inputs_digest(progs_and_terms):
    """sha256 over the packed op tables, term tables and pools of a workload: a
    fixture of outputs alone states which inputs it belongs to."""
    import hashlib
    h = hashlib.sha256()
    for prog, terms, nt, hpool in progs_and_terms:
        h.update(op_table(prog).tobytes())
        h.update(np.ascontiguousarray(prog.pool()).tobytes())
        h.update(term_table(terms, nt).tobytes())
        h.update(np.ascontiguousarray(hpool).tobytes())
    return h.hexdigest()


def config_fixtures():
    """Oracle outputs at the measured workload sizes (minutes of CPU work; run with
    the argument `configs`): every instance of the 256-instance uneven batch, and the
    12-qubit noisy density matrix of 6 layers (216 angles). Inputs come from the
    seeded generators of vqchip.workloads; the fixture holds their digest."""
    from vqchip.workloads import config2_instances
    oracle.set_threads(os.cpu_count() or 1)
    inst = config2_instances()
    packed, E, grads, ns = [], [], [], []
    for c, H, n in inst:
        prog = Program(c, n=n)
        terms, nt, hpool = pack_terms(H)
        packed.append((prog, terms, nt, hpool))
        psi0 = np.zeros(1 << n, complex)
        psi0[0] = 1
        e, g, _ = oracle.energy_grad(psi0, n, 0, prog.op_array(), len(prog.ops), prog.pool(), terms, nt, hpool,
                                     prog.nparams)
        E.append(e)
        grads.append(g)
        ns.append(n)
    off = np.concatenate([[0], np.cumsum([len(g) for g in grads])])
    np.savez_compressed(os.path.join(HERE, "config2_energy_grad.npz"), nqubits=np.array(ns), energy=np.array(E),
                        grad=np.concatenate(grads), grad_offsets=off, inputs_sha256=np.array(inputs_digest(packed)))
    c, H, _ = config3()
    prog = Program(c, n=12, dm=True)
    terms, nt, hpool = pack_terms(H)
    rho0 = np.zeros(1 << 24, complex)
    rho0[0] = 1
    e, g, _ = oracle.energy_grad(rho0, 12, 1, prog.op_array(), len(prog.ops), prog.pool(), terms, nt, hpool, prog.nparams)
    np.savez_compressed(os.path.join(HERE, "config3_energy_grad.npz"), energy=np.array(e), grad=g,
                        inputs_sha256=np.array(inputs_digest([(prog, terms, nt, hpool)])))


if __name__ == "__main__":
    if sys.argv[1:] == ["configs"]:
        config_fixtures()
    else:
        main()
