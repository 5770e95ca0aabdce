"""Forward-state checkpoints in the adjoint sweep: the reverse pass restores
stored forward states instead of uncomputing through S^-1 (circuitdiff.jl:53-59
inverts each channel). Checked against central differences of the oracle's
forward energy (no inverse involved) for strong (p = 0.5) and singular
(p = 3/4, no inverse at all) depolarizing channels, and checkpointed against
plain results where both are exact. Tolerances: gradients 1e-8 absolute against
5-point central differences with h = 1e-3; 1e-10 relative between two exact routes."""
import copy

import numpy as np
import pytest

import oracle
from brute import rand_state, rand_unitary, vec
import vqchip as vq
from vqchip.circuit import Program
from vqchip.operators import pack_terms
from vqchip.workloads import hea_circuit, sum_z, tfim

pytestmark = pytest.mark.gpu


def noisy_circuit(n, depth, p, seed):
    rng = np.random.default_rng(seed)
    c = vq.QCircuit()
    for _ in range(depth):
        for i in range(1, n + 1):
            c.append(vq.RxGate(i, float(rng.uniform(0, 6.3))))
            c.append(vq.RyGate(i, float(rng.uniform(0, 6.3))))
        for i in range(1, n):
            c.append(vq.CNOTGate(i, i + 1))
            c.append(vq.Depolarizing(i + 1, p))
        c.append(vq.AmplitudeDamping(1, 0.2))
    return c


def oracle_energy(c, H, n):
    """tr(H rho) of the oracle's forward pass (no inverses anywhere)."""
    prog = Program(c, n=n, dm=True)
    rho = np.zeros(1 << (2 * n), dtype=complex)
    rho[0] = 1
    rho = oracle.apply_circuit(rho, n, 1, prog.op_array(), len(prog.ops), prog.pool())
    return sum(oracle.dm_expec(rho, n, len(k), [q - 1 for q in k], M) for k, M in H.data().items()).real


def fd(c, H, n, k, h=1e-3):
    theta = np.array(c.active_parameters(), dtype=float)

    def E(dt):
        th = theta.copy()
        th[k] += dt
        cc = copy.deepcopy(c)
        cc.reset_parameters(th)
        return oracle_energy(cc, H, n)

    return (-E(2 * h) + 8 * E(h) - 8 * E(-h) + E(-2 * h)) / (12 * h)


@pytest.mark.parametrize("p", [0.5, 0.75])
def test_strong_and_singular_depolarizing_gradient_against_fd(p):
    n, depth = 5, 4
    c = noisy_circuit(n, depth, p, seed=int(p * 100))
    if p == 0.75:  # singular superoperator: no inverse is packed
        assert any(o.kind == 1 and o.aux == -1 for o in Program(c, n=n, dm=True).ops)
    H = sum_z(n) + vq.QubitsTerm({1: "X", 2: "X"}, coeff=0.4)
    E, g = vq.energy_and_gradient(H, c, vq.DensityMatrix(n))
    assert abs(E.real - oracle_energy(c, H, n)) < 1e-12
    worst = 0.0
    for k in range(0, len(g), 3):
        worst = max(worst, abs(g[k] - fd(c, H, n, k)))
    print(f"p={p}: max |grad - fd| {worst:.2e} over {len(range(0, len(g), 3))} angles", flush=True)
    assert worst < 1e-8


def test_checkpoint_flag_agrees_with_the_uncomputing_sweep():
    """VQC_FLAG_CHECKPOINT (segments every ceil(sqrt(#ops)) ops) on state vectors
    and on a mildly noisy density matrix, where uncomputing is exact too."""
    rng = np.random.default_rng(3)
    n = 14
    c = hea_circuit(n, 6, rng.uniform(0, 6.3, 3 * n * 6))
    for _ in range(5):
        pos = tuple(sorted(rng.choice(np.arange(1, n + 1), size=3, replace=False).tolist()))
        c.append(vq.QuantumGate(pos, rand_unitary(3, rng)))
    H = tfim(n, 0.8)
    psi = vq.StateVector(rand_state(n, rng))
    E0, g0 = vq.energy_and_gradient(H, c, psi)
    E1, g1 = vq.energy_and_gradient(H, c, psi, checkpoint=True)
    assert abs(E1 - E0) < 1e-12 and np.max(np.abs(g1 - g0)) < 1e-10 * max(1, np.max(np.abs(g0)))
    prog = Program(c, n=n)
    terms, nt, hpool = pack_terms(H)
    Er, gr, _ = oracle.energy_grad(psi.storage(), n, 0, prog.op_array(), len(prog.ops), prog.pool(), terms, nt, hpool,
                                   prog.nparams)
    assert np.max(np.abs(g1 - gr)) < 1e-10 * max(1, np.max(np.abs(gr)))
    # vjp with checkpoints (state vector)
    d = vq.StateVector(rand_state(n, rng))
    ga, dxa = vq.circuit_vjp(c, psi, d)
    gb, dxb = vq.circuit_vjp(c, psi, d, checkpoint=True)
    assert np.max(np.abs(ga - gb)) < 1e-10 * max(1, np.max(np.abs(ga)))
    assert np.max(np.abs(dxa.storage() - dxb.storage())) < 1e-12
    # density matrix, weak noise
    m = 4
    cd = noisy_circuit(m, 3, 0.02, seed=8)
    Hd = sum_z(m)
    Ea, gda = vq.energy_and_gradient(Hd, cd, vq.DensityMatrix(m))
    Eb, gdb = vq.energy_and_gradient(Hd, cd, vq.DensityMatrix(m), checkpoint=True)
    assert abs(Ea - Eb) < 1e-12 and np.max(np.abs(gda - gdb)) < 1e-10 * max(1, np.max(np.abs(gda)))
