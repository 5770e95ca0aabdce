"""The reference's gate-operation checks (test/check_gateop.jl) on the device.

* Generic gates (check_gateop.jl:1-64, 180-188): a random k-qubit unitary on every
  position pattern the reference loops over, applied by the device and by the
  oracle's serial sweep (the reference compares apply! with apply_serial!).
* Specialized gates (check_gateop.jl:66-178, 190-194): each gate and its adjoint,
  applied through its own path (a named kernel where the library has one), against
  the dense gate of its ordered matrix on the same state (the reference's
  check_single_gate), and against the oracle.

The reference's tolerance is 1e-6 (it runs ComplexF32 too); here everything is
FP64 and must agree within 1e-12.
"""
import numpy as np
import pytest

import oracle
from brute import rand_state, rand_unitary
import vqchip as vq

pytestmark = pytest.mark.gpu
TOL = 1e-12


def device_apply(gate, psi):
    s = vq.StateVector(np.ascontiguousarray(psi))
    vq.apply_(gate, s)
    return s.storage()


def check_generic(L, positions, rng):
    psi = rand_state(L, rng)
    U = rand_unitary(len(positions), rng)
    g = vq.QuantumGate(positions, U)
    out = device_apply(g, psi)
    ref = oracle.apply_dense(psi.copy(), len(positions), [q - 1 for q in g.ordered_positions()], g.ordered_mat())
    return np.max(np.abs(out - ref))


def test_generic_onebody():
    rng = np.random.default_rng(1)
    L = 15
    assert max(check_generic(L, (c,), rng) for c in range(1, L + 1)) < TOL


@pytest.mark.parametrize("L", [5, 15])
def test_generic_twobody(L):
    rng = np.random.default_rng(L)
    errs = [check_generic(L, (c, t), rng) for c in range(1, L + 1, 2) for t in range(2, L + 1, 2) if c != t]
    assert errs and max(errs) < TOL


def test_generic_threebody():
    rng = np.random.default_rng(3)
    L = 15
    errs = [check_generic(L, (a, b, c), rng) for a in range(1, L + 1, 2) for b in range(2, L + 1, 3)
            for c in range(3, L + 1, 5) if len({a, b, c}) == 3]
    assert len(errs) > 50 and max(errs) < TOL


def check_single_gate(L, gt, rng):
    """gt (its own path) against the dense gate of its ordered matrix, and both against the oracle."""
    psi = rand_state(L, rng)
    out = device_apply(gt, psi)
    generic = device_apply(vq.QuantumGate(gt.ordered_positions(), gt.ordered_mat()), psi)
    ref = oracle.apply_dense(psi.copy(), len(gt.positions), [q - 1 for q in gt.ordered_positions()], gt.ordered_mat())
    return max(np.max(np.abs(out - generic)), np.max(np.abs(out - ref)))


def both(gt):
    return [gt, gt.adjoint()]


def test_specialized_onebody_gates():
    rng = np.random.default_rng(11)
    L = 16
    errs = []
    for c in range(1, L + 1):
        for gt in (vq.XGate(c), vq.YGate(c), vq.ZGate(c), vq.TGate(c), vq.PHASEGate(c, float(rng.standard_normal()))):
            errs += [check_single_gate(L, g, rng) for g in both(gt)]
    assert max(errs) < TOL


def test_specialized_twobody_gates():
    rng = np.random.default_rng(12)
    L = 16
    errs = []
    for c in range(1, L + 1, 2):
        for t in range(2, L + 1, 2):
            gates = [vq.CONTROLGate((c, t), rand_unitary(1, rng)), vq.CRxGate(c, t, float(rng.standard_normal())),
                     vq.CRyGate(c, t, float(rng.standard_normal())), vq.CRzGate(c, t, float(rng.standard_normal())),
                     vq.CZGate(c, t), vq.CNOTGate(c, t), vq.CPHASEGate(c, t, float(rng.standard_normal())),
                     vq.SWAPGate(c, t), vq.iSWAPGate(c, t), vq.FSIMGate(c, t, *rng.standard_normal(2))]
            for gt in gates:
                errs += [check_single_gate(L, g, rng) for g in both(gt)]
    assert len(errs) == 8 * 8 * 10 * 2 and max(errs) < TOL


def test_specialized_threebody_gates():
    rng = np.random.default_rng(13)
    L = 15
    errs = []
    for a in range(1, L + 1, 2):
        for b in range(2, L + 1, 3):
            for c in range(3, L + 1, 5):
                if len({a, b, c}) < 3:
                    continue
                gates = [vq.CONTROLCONTROLGate(a, b, c, rand_unitary(1, rng)), vq.TOFFOLIGate(a, b, c),
                         vq.FREDKINGate(a, b, c), vq.CCPHASEGate(a, b, c, float(rng.standard_normal()))]
                for gt in gates:
                    errs += [check_single_gate(L, g, rng) for g in both(gt)]
    assert len(errs) > 400 and max(errs) < TOL
