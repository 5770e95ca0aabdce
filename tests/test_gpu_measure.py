"""Measurement and post-selection on the device against the oracle (which is
pinned by brute force and the reference's known answers in test_oracle.py).
Tolerance: probabilities and amplitudes within 1e-12 absolute; outcomes exact."""
import numpy as np
import pytest

import oracle
from brute import rand_density, rand_state, vec
import vqchip as vq

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [3, 11, 20])
def test_probability_zero(n):
    rng = np.random.default_rng(80 + n)
    v = rand_state(n, rng)
    s = vq.StateVector(v)
    for key in sorted({1, 2, n // 2 + 1, n}):
        assert abs(vq.probability_zero(s, key) - oracle.prob_zero(v, n, key - 1)) < 1e-12
    if n <= 11:
        rho = vec(rand_density(min(n, 6), rng))
        m = min(n, 6)
        d = vq.DensityMatrix(rho)
        for key in (1, m):
            assert abs(vq.probability_zero(d, key) - oracle.prob_zero(rho, m, key - 1, is_dm=True)) < 1e-12


@pytest.mark.parametrize("n", [4, 18])
def test_measure_in_place(n):
    rng = np.random.default_rng(90 + n)
    for key in (1, 3, n):
        for auto_reset in (True, False):
            v = rand_state(n, rng)
            u = float(rng.random())
            s = vq.StateVector(v)
            i, p = vq.measure_(s, key, auto_reset=auto_reset, u=u)
            w = v.copy()
            i_r, p_r = oracle.measure_inplace(w, n, key - 1, u, auto_reset)
            assert i == i_r and abs(p - p_r) < 1e-12
            assert np.max(np.abs(s.storage() - w)) < 1e-12
    # density matrix (measure! accepts both, measure.jl:4)
    rho = vec(rand_density(5, rng))
    d = vq.DensityMatrix(rho)
    i, p = vq.measure_(d, 2, auto_reset=True, u=0.3)
    r = rho.copy()
    i_r, p_r = oracle.measure_inplace(r, 5, 1, 0.3, True, is_dm=True)
    assert i == i_r and abs(p - p_r) < 1e-12
    assert np.max(np.abs(vec(d.storage()) - r)) < 1e-12


@pytest.mark.parametrize("n", [2, 9, 17])
def test_select_and_post_select(n):
    rng = np.random.default_rng(100 + n)
    v = rand_state(n, rng)
    s = vq.StateVector(v)
    for key in sorted({1, 2, n}):
        for val in (0, 1):
            r, nrm = vq.post_select(s, key, val)
            ref = oracle.select(v.copy(), n, key - 1, val)
            nref = np.sqrt(np.vdot(ref, ref).real)
            assert abs(nrm - nref) < 1e-12
            assert np.max(np.abs(r.storage() - ref / nref)) < 1e-12
            s2 = vq.StateVector(v)
            nrm2 = vq.post_select_(s2, key, val)
            w = v.copy()
            assert abs(nrm2 - oracle.post_select_inplace(w, n, key - 1, val)) < 1e-12
            assert np.max(np.abs(s2.storage() - w)) < 1e-12
    # measure (non-mutating) returns the selected row of the sampled outcome
    out, i, p = vq.measure(s, n, u=0.25)
    p0 = oracle.prob_zero(v, n, n - 1)
    assert i == (0 if 0.25 < p0 else 1)
    ref = oracle.select(v.copy(), n, n - 1, i)
    assert abs(p - np.vdot(ref, ref).real) < 1e-12
    assert np.max(np.abs(out.storage() - ref / np.sqrt(p))) < 1e-12
    assert np.array_equal(s.storage(), v)


def test_measurement_known_answers_and_errors():
    # check_select_3/4 (test/check_measure.jl): Bell state post-selected on qubit 1
    bell = np.zeros(4, complex)
    bell[0] = bell[3] = np.sqrt(0.5)
    for val, exp in [(0, [1, 0]), (1, [0, 1])]:
        r, p = vq.post_select(vq.StateVector(bell), 1, val)
        assert np.allclose(r.storage(), exp, atol=1e-15) and abs(p - np.sqrt(0.5)) < 1e-15
    s = vq.StateVector(3)  # |000>: P(1) on any qubit is 0
    with pytest.raises(vq.VQCError, match="zero probability"):
        vq.post_select_(s, 2, 1)
    with pytest.raises(ValueError):
        vq.measure_(s, 4)
    i, p = vq.measure_(s, 2, u=0.999)
    assert i == 0 and p == 1.0
