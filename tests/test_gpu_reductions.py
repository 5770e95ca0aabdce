"""Partial trace, fidelity and distance on the device (ptrace.jl, distance.jl,
statevector.jl:59-65, densitymatrix.jl:60-73) against numpy on explicit index
maps and against closed forms (product states). Tolerance 1e-12."""
import numpy as np
import pytest

from brute import rand_density, rand_state, unvec, vec
import vqchip as vq
from vqchip import reductions as R

pytestmark = pytest.mark.gpu


def index_map(n, sites0):
    keep = [q for q in range(n) if q not in sites0]
    trace = sorted(sites0)
    M, K = 1 << len(keep), 1 << len(trace)
    a = np.arange(M)[:, None]
    k = np.arange(K)[None, :]
    idx = np.zeros((M, K), dtype=np.int64)
    for j, q in enumerate(keep):
        idx |= ((a >> j) & 1) << q
    for j, q in enumerate(trace):
        idx |= ((k >> j) & 1) << q
    return idx


@pytest.mark.parametrize("sites", [[1], [8], [2, 5, 7], [3, 1], [1, 2, 3, 4, 5, 6, 7]])
def test_partial_trace_state_vector(sites):
    n = 8
    psi = rand_state(n, np.random.default_rng(sum(sites)))
    out = R.partial_tr(vq.StateVector(psi), sites)
    A = psi[index_map(n, [q - 1 for q in sites])]
    assert out.nqubits() == n - len(sites)
    assert np.max(np.abs(out.storage() - A @ A.conj().T)) < 1e-12


@pytest.mark.parametrize("sites", [[1, 2], [2, 5], [6, 3]])
def test_partial_trace_state_equals_its_density_matrix(sites):
    # check_utility.jl:4-13: partial_tr(a, sites) == partial_tr(DensityMatrix(a), sites)
    psi = rand_state(6, np.random.default_rng(sites[0] * 10 + sites[1]))
    a = R.partial_tr(vq.StateVector(psi), sites)
    b = R.partial_tr(vq.DensityMatrix(vq.StateVector(psi)), sites)
    assert a.nqubits() == b.nqubits() == 4
    assert np.max(np.abs(a.storage() - b.storage())) < 1e-12


def test_partial_trace_density_matrix_and_scalars():
    n = 6
    rng = np.random.default_rng(12)
    rho = rand_density(n, rng)
    for sites in ([2], [1, 6], [4, 2, 5]):
        out = R.partial_tr(vq.DensityMatrix(vec(rho)), sites)
        idx = index_map(n, [q - 1 for q in sites])
        ref = rho[idx[:, None, :], idx[None, :, :]].sum(-1)
        assert np.max(np.abs(out.storage() - ref)) < 1e-12
        assert abs(R.tr(out) - 1) < 1e-12
    assert abs(R.partial_tr(vq.DensityMatrix(vec(rho)), list(range(1, n + 1))) - 1) < 1e-12
    psi = rand_state(n, rng)
    assert abs(R.partial_tr(vq.StateVector(psi), list(range(1, n + 1))) - 1) < 1e-12
    with pytest.raises(ValueError):
        R.partial_tr(vq.StateVector(psi), [2, 2])


def test_partial_trace_of_product_state_is_the_factor():
    # psi = a (qubits 1..3) x b (qubits 4..6): tracing 4..6 leaves |a><a| exactly
    rng = np.random.default_rng(13)
    a, b = rand_state(3, rng), rand_state(3, rng)
    psi = np.kron(b, a)  # qubit 1 is the least significant bit
    out = R.partial_tr(vq.StateVector(psi), [4, 5, 6])
    assert np.max(np.abs(out.storage() - np.outer(a, a.conj()))) < 1e-12
    out = R.partial_tr(vq.StateVector(psi), [1, 2, 3])
    assert np.max(np.abs(out.storage() - np.outer(b, b.conj()))) < 1e-12


def test_fidelity_and_distance():
    rng = np.random.default_rng(14)
    n = 7
    x, y = rand_state(n, rng), rand_state(n, rng)
    sx, sy = vq.StateVector(x), vq.StateVector(y)
    assert abs(R.fidelity(sx, sy) - abs(np.vdot(x, y)) ** 2) < 1e-14
    assert abs(R.distance2(sx, sy) - np.linalg.norm(x - y) ** 2) < 1e-12
    rho = rand_density(n, rng)
    d = vq.DensityMatrix(vec(rho))
    assert abs(R.fidelity(d, sy) - np.vdot(y, rho @ y).real) < 1e-13
    assert abs(R.fidelity(sy, d) - np.vdot(y, rho @ y).real) < 1e-13
    rho2 = rand_density(n, rng)
    assert abs(R.distance(d, vq.DensityMatrix(vec(rho2))) - np.linalg.norm(rho - rho2)) < 1e-12


@pytest.mark.parametrize("n", [1, 4, 7, 10])
def test_fidelity_of_two_density_matrices(n):
    """fidelity(x::DensityMatrix, y::DensityMatrix) = real(tr(sqrt(x) sqrt(y)))
    (densitymatrix.jl:65) on the device against the oracle (pinned to scipy's
    sqrtm and closed forms); full-rank inputs to 1e-10, pure states (square roots
    of round-off eigenvalues, ~1e-8 in any implementation of the formula) to 1e-7."""
    import oracle
    rng = np.random.default_rng(650 + n)
    x, y = rand_density(n, rng), rand_density(n, rng)
    f = R.fidelity(vq.DensityMatrix(x), vq.DensityMatrix(y))
    assert abs(f - oracle.fidelity_dm(x, y)) < 1e-10
    assert abs(R.fidelity(vq.DensityMatrix(x), vq.DensityMatrix(x)) - 1.0) < 1e-10
    psi, phi = rand_state(n, rng), rand_state(n, rng)
    fp = R.fidelity(vq.DensityMatrix(vq.StateVector(psi)), vq.DensityMatrix(vq.StateVector(phi)))
    assert abs(fp - abs(psi.conj() @ phi) ** 2) < 1e-7
