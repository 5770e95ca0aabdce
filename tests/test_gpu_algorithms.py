"""The phase-estimation known answers of the reference (test/algs.jl:26-87) on
the device: the circuit runs through the fused tile path (dense 2-qubit gates on
unsorted positions included), then every qubit is measured in turn; the
outcomes must spell the encoded phase exactly. The reduce-as-you-measure variant
ends on a 0-qubit state whose one amplitude is 1."""
import numpy as np
import pytest

from algorithms import phase_estimation_circuit, to_digits
import vqchip as vq

pytestmark = pytest.mark.gpu


def _bits(rng, L):
    return [int(b) for b in rng.integers(0, 2, L)]


@pytest.mark.parametrize("auto_reset", [False, True])
def test_phase_estimation_measure_in_place(auto_reset):
    rng = np.random.default_rng(21 + auto_reset)
    for L in range(2, 16):
        bits = _bits(rng, L)
        s = vq.StateVector(L + 1)
        vq.apply_(phase_estimation_circuit(bits), s)
        res = [vq.measure_(s, q, auto_reset=auto_reset, u=float(rng.random()))[0] for q in range(1, L + 2)]
        assert to_digits(res) == to_digits(bits) and res[:L] == bits, (L, bits, res)


@pytest.mark.parametrize("auto_reset", [False, True])
def test_phase_estimation_density_matrix(auto_reset):
    rng = np.random.default_rng(31 + auto_reset)
    for L in range(2, 7):
        bits = _bits(rng, L)
        d = vq.DensityMatrix(L + 1)
        vq.apply_(phase_estimation_circuit(bits), d)
        res = [vq.measure_(d, q, auto_reset=auto_reset, u=float(rng.random()))[0] for q in range(1, L + 2)]
        assert to_digits(res) == to_digits(bits) and res[:L] == bits, (L, bits, res)
        assert abs(d.tr() - 1) < 1e-12


def test_phase_estimation_reducing_measure():
    rng = np.random.default_rng(41)
    for L in range(2, 16):
        bits = _bits(rng, L)
        s = vq.StateVector(L + 1)
        vq.apply_(phase_estimation_circuit(bits), s)
        res = []
        for _ in range(L + 1):
            s, i, p = vq.measure(s, 1, u=float(rng.random()))
            res.append(i)
            assert abs(p - 1) < 1e-10
        assert s.nqubits() == 0 and s.storage().size == 1 and abs(s.storage()[0] - 1) < 1e-10
        assert to_digits(res) == to_digits(bits) and res[:L] == bits, (L, bits, res)


def test_zero_qubit_state_supports_whole_state_calls_only():
    s = vq.StateVector(np.array([0.6 + 0.8j]))
    assert s.nqubits() == 0 and abs(s.norm() - 1) < 1e-15
    t = s.copy()
    t.scale_(2.0)
    assert abs(t.dot(s) - 2.0) < 1e-15
    with pytest.raises(vq.VQCError):
        vq._lib.call("vqc_prob", s.handle, 0, 0, vq._lib.dptr(np.zeros(1)))
    with pytest.raises(vq.VQCError):
        vq._lib.call("vqc_apply_circuit", s.handle, 0, None, None)
