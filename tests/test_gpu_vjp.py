"""vqc_circuit_vjp (the pullback of y = circuit * x for any cotangent,
circuitdiff.jl:2-8, 23-78) against the oracle's, which test_vjp.py pins by
brute force and finite differences. Tolerances (FP64): states within 1e-12
absolute per unit of |delta|, gradients within 1e-10 relative."""
import numpy as np
import pytest

import oracle
from brute import rand_density, rand_state, vec
from test_vjp import dm_circuit, sv_circuit
import vqchip as vq
from vqchip.circuit import Program
from vqchip.workloads import tfim

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [4, 7, 12, 16])
def test_state_vector_vjp(n):
    rng = np.random.default_rng(300 + n)
    c = sv_circuit(n, rng)
    x = rand_state(n, rng)
    d = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    g, dx, y = vq.circuit_vjp(c, vq.StateVector(x), vq.StateVector(d), return_state=True)
    prog = Program(c, n=n)
    gr, dxr, yr = oracle.circuit_vjp(x, n, 0, prog.op_array(), len(prog.ops), prog.pool(), prog.nparams, d)
    scale = np.max(np.abs(d))
    assert np.max(np.abs(y.storage() - yr)) < 1e-12
    assert np.max(np.abs(dx.storage() - dxr)) < 1e-12 * scale
    assert np.max(np.abs(g - gr)) < 1e-10 * max(1.0, np.max(np.abs(gr)))


@pytest.mark.parametrize("n", [3, 5])
def test_density_matrix_vjp_with_channels(n):
    rng = np.random.default_rng(400 + n)
    c, _ = dm_circuit(n)
    x = vec(rand_density(n, rng))
    d = rng.standard_normal(1 << (2 * n)) + 1j * rng.standard_normal(1 << (2 * n))
    g, dx, y = vq.circuit_vjp(c, vq.DensityMatrix(x), vq.DensityMatrix(d), return_state=True)
    prog = Program(c, n=n, dm=True)
    gr, dxr, yr = oracle.circuit_vjp(x, n, 1, prog.op_array(), len(prog.ops), prog.pool(), prog.nparams, d)
    assert np.max(np.abs(vec(y.storage()) - yr)) < 1e-12
    assert np.max(np.abs(vec(dx.storage()) - dxr)) < 1e-12 * np.max(np.abs(d))
    assert np.max(np.abs(g - gr)) < 1e-10 * max(1.0, np.max(np.abs(gr)))


def test_vjp_with_the_energy_cotangent_matches_energy_and_gradient():
    n = 14
    rng = np.random.default_rng(500)
    c = sv_circuit(n, rng)
    H = tfim(n, 0.8)
    s = vq.StateVector(rand_state(n, rng))
    E, g, y = vq.energy_and_gradient(H, c, s, return_state=True)
    Hy = vq.apply_operator(H, y)
    Hy.scale_(2.0)  # (H + H^dagger) y for a Hermitian H
    gv, dx = vq.circuit_vjp(c, s, Hy)
    assert np.max(np.abs(gv - g)) < 1e-10 * max(1.0, np.max(np.abs(g)))


def test_vjp_rejects_mismatched_shapes():
    c = sv_circuit(5, np.random.default_rng(1))
    with pytest.raises(ValueError):
        vq.circuit_vjp(c, vq.StateVector(5), vq.StateVector(4))
