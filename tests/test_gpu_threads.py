"""Single-state entry points called from several host threads at once. ctypes
releases the GIL, so the C calls overlap; they share the library's stream and
staging buffers, which one library lock serialises. Every thread's results must
equal a serial run exactly (same kernels, same reduction order)."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from brute import rand_state
import vqchip as vq
from vqchip.workloads import hea_circuit, tfim

pytestmark = pytest.mark.gpu


def test_single_state_calls_from_several_threads():
    n, nthreads = 14, 4
    rng = np.random.default_rng(123)
    c = hea_circuit(n, 2, rng.uniform(0, 6.3, 6 * n))
    H = tfim(n, 0.7)
    inits = [rand_state(n, rng) for _ in range(nthreads)]

    def work(i):
        s = vq.StateVector(inits[i])
        out = []
        for _ in range(15):
            vq.apply_(c, s)
            out.append(vq.expectation(H, s))
            out.append(s.dot(s))
            E, g = vq.energy_and_gradient(H, c, s)
            out.append(E)
            out.extend(g.tolist())
        return out

    serial = [work(i) for i in range(nthreads)]
    with ThreadPoolExecutor(nthreads) as ex:
        par = list(ex.map(work, range(nthreads)))
    for a, b in zip(serial, par):
        assert a == b
