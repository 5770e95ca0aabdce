"""The autotuner: every candidate tile shape gives the same energy and gradient
(to 1e-10), the fastest is left in force, and the previous tuning can be put back."""
import numpy as np
import pytest

import vqchip as vq
from vqchip import autotune
from vqchip.workloads import config3, hea_circuit, tfim

pytestmark = pytest.mark.gpu


@pytest.fixture
def restore_tuning():
    before = autotune.get_tuning()
    yield before
    autotune.set_tuning(**{k: before[k] for k in autotune.FIELDS}, dense_groups=before["dense_groups"],
                        dense_nontemporal=before["dense_nontemporal"])
    assert autotune.get_tuning() == before


def test_tune_energy_grad_state_vector(restore_tuning):
    n = 15
    c = hea_circuit(n, 3, np.random.default_rng(5).uniform(0, 6.3, 9 * n))
    H = tfim(n, 0.9)
    s = vq.StateVector(n)
    E0, g0 = vq.energy_and_gradient(H, c, s)
    res = autotune.tune_energy_grad(H, c, s, fwd=(11, 12), rev=(10, 11), low=(4, 6), reps=1)
    assert len(res["table"]) == 8
    t = autotune.get_tuning()
    assert {k: t[k] for k in res["best"]} == res["best"]
    E1, g1 = vq.energy_and_gradient(H, c, s)
    assert abs(E1 - E0) < 1e-10 and np.max(np.abs(g1 - g0)) < 1e-10


def test_tune_energy_grad_density_matrix(restore_tuning):
    c, H, _ = config3(n=6, depth=2)
    rho = vq.DensityMatrix(6)
    E0, g0 = vq.energy_and_gradient(H, c, rho)
    res = autotune.tune_energy_grad(H, c, rho, fwd=(11, 12), rev=(11,), low=(2, 6), reps=1)
    E1, g1 = vq.energy_and_gradient(H, c, rho)
    assert abs(E1 - E0) < 1e-10 and np.max(np.abs(g1 - g0)) < 1e-10
    assert res["best"]["tile_low_bits"] in (2, 6)


def test_tune_dense(restore_tuning):
    res = autotune.tune_dense(16, reps=1)
    assert len(res["table"]) == 3 * 8
    t = autotune.get_tuning()
    assert t["dense_groups"] == res["best"]["dense_groups"]
    s = vq.StateVector(np.ones(1 << 10, complex) / 32)
    vq.apply_(vq.HGate(3), s)  # still a valid sweep with the tuned shape
    assert abs(np.linalg.norm(s.storage()) - 1) < 1e-12
