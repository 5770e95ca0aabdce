"""vqc_sv_expec_pauli_noalloc, the entry point that allocates nothing, against
vqc_sv_expec_pauli (which grows the library's scratch and pinned buffers): two
calls in a row, the second on a smaller state with fewer terms, each equal to the
allocating entry point's result (same kernels and summation order: bit-identical;
the tolerance stated is 1e-14 relative). Sizes beyond the stated maximum are
rejected."""
import numpy as np
import pytest

from brute import rand_state
import vqchip as vq
from vqchip import _lib
from vqchip.operators import pack_pauli
from vqchip.workloads import random_pauli_operator

pytestmark = pytest.mark.gpu


def _both(s, arr, nt, bra=None):
    per_a, per_b = np.zeros(2 * nt), np.zeros(2 * nt)
    sa, sb = np.zeros(2), np.zeros(2)
    b = bra.handle if bra is not None else None
    _lib.call("vqc_sv_expec_pauli", s.handle, b, nt, arr, _lib.dptr(per_a), _lib.dptr(sa))
    _lib.call("vqc_sv_expec_pauli_noalloc", s.handle, b, nt, arr, _lib.dptr(per_b), _lib.dptr(sb))
    return per_a, per_b, sa, sb


def test_noalloc_pauli_expectation_matches_the_allocating_entry_point():
    rng = np.random.default_rng(2024)
    for n, nterms in ((20, 200), (9, 37)):  # the second call: smaller state, fewer terms
        s = vq.StateVector(rand_state(n, rng))
        arr, nt = pack_pauli(random_pauli_operator(n, nterms, seed=n))
        per_a, per_b, sa, sb = _both(s, arr, nt)
        scale = max(1.0, float(np.max(np.abs(per_a))))
        assert np.max(np.abs(per_a - per_b)) <= 1e-14 * scale
        assert np.max(np.abs(sa - sb)) <= 1e-14 * max(1.0, float(np.max(np.abs(sa))))
        bra = vq.StateVector(rand_state(n, rng))
        per_a, per_b, sa, sb = _both(s, arr, nt, bra)
        assert np.max(np.abs(per_a - per_b)) <= 1e-14 * max(1.0, float(np.max(np.abs(per_a))))


def test_noalloc_rejects_more_terms_than_its_maximum():
    s = vq.StateVector(6)
    arr, nt = pack_pauli(random_pauli_operator(6, 257, seed=3))
    out = np.zeros(2 * nt)
    with pytest.raises(_lib.VQCError) as e:
        _lib.call("vqc_sv_expec_pauli_noalloc", s.handle, None, nt, arr, _lib.dptr(out), None)
    assert e.value.code == -1
