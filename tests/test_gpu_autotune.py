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


@pytest.fixture
def restore_table():
    before = autotune.get_tile_table()
    yield before
    autotune.set_tile_table({n: (0, 0, 0) for n in list(range(1, 35)) + [-1]})
    autotune.set_tile_table(before)


def test_tune_sizes_sets_the_per_size_table(restore_tuning, restore_table):
    """The per-size entries the library's own callers read (single states,
    narrow batches, the serving loop): every candidate gives the same energy and
    gradient, and the chosen entry is in force."""
    res = autotune.tune_sizes([12, 15], reps=1, lows=(3, 4), spread=(0, 1))
    assert set(res) == {12, 15}
    table = autotune.get_tile_table()
    for n, r in res.items():
        assert table[n] == r["best"]
    c = hea_circuit(15, 3, np.random.default_rng(8).uniform(0, 6.3, 9 * 15))
    H = tfim(15, 0.7)
    E, g = vq.energy_and_gradient(H, c, vq.StateVector(15))
    autotune.set_tile_table({15: (0, 0, 0)})
    E2, g2 = vq.energy_and_gradient(H, c, vq.StateVector(15))
    assert abs(E - E2) < 1e-10 and np.max(np.abs(g - g2)) < 1e-10


def test_tune_batch_sets_the_wide_batch_entry(restore_table):
    from vqchip.batch import Batch
    from vqchip.workloads import config2_instances
    inst = config2_instances(count=24, nmin=12, nmax=16, dmax=4)
    b = Batch(inst)
    E0, g0 = b.run()
    res = autotune.tune_batch(b, fwd=(12, 13), rev=(12,), lows=(4, 5), reps=1)
    assert autotune.get_tile_table()[-1] == res["best"]
    E1, g1 = b.run()
    assert np.max(np.abs(E1 - E0)) < 1e-10
    assert max(float(np.max(np.abs(x - y))) if x.size else 0.0 for x, y in zip(g0, g1)) < 1e-10


def test_tune_forward(restore_table):
    from vqchip.workloads import config1_circuit
    n = 18
    c = config1_circuit(n=n, layers=3)
    res = autotune.tune_forward(c, vq.StateVector(n), fwd=(11, 12), lows=(3, 4), reps=1)
    assert autotune.get_tile_table()[n][0] == res["best"][0]
