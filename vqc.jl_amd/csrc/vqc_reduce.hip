// vqc_reduce.hip -- loss-side reductions around the path: partial trace
// (src/ptrace.jl) and <psi|rho|psi> (the density-matrix / state-vector
// fidelity of src/densitymatrix.jl:70).
//
// Partial trace, in the reference's index order (ptrace.jl:24-53 with
// move_selected_index_backward, auxiliary/indexop.jl:46-70): the kept qubits,
// ascending, become the qubits of the result; out is vec of a 2^m x 2^m matrix.
//   * density matrix: out[a, b] = sum_k rho[(a,k), (b,k)], one thread per entry;
//   * state vector:   out[a, b] = sum_k psi(a,k) conj(psi(b,k)) = (A A^dagger)[a, b]
//     with A[a, k] = psi(a,k): an LDS-tiled Hermitian product, 16 x 16 entries
//     per workgroup, the traced index staged through LDS 16 at a time.
#include "vqc_internal.h"

namespace {

constexpr int kRB = 256;
constexpr int kTile = 16;

struct BitMap {
    int keep[32];   // positions of the kept qubits, ascending
    int trace[32];  // positions of the traced qubits
    int m, t;
};

__device__ __forceinline__ uint64_t scatter(uint64_t v, const int* pos, int cnt) {
    uint64_t r = 0;
    for (int j = 0; j < cnt; ++j) r |= ((v >> j) & 1) << pos[j];
    return r;
}

__global__ void __launch_bounds__(kRB) k_ptrace_dm(const double2* __restrict__ rho, int n, BitMap bm, double2* __restrict__ out) {
    const uint64_t M = uint64_t(1) << bm.m;
    const uint64_t e = uint64_t(blockIdx.x) * kRB + threadIdx.x;
    if (e >= M * M) return;
    const uint64_t a = e & (M - 1), b = e >> bm.m;
    const uint64_t ia = scatter(a, bm.keep, bm.m), ib = scatter(b, bm.keep, bm.m);
    double2 s = make_double2(0.0, 0.0);
    const uint64_t K = uint64_t(1) << bm.t;
    for (uint64_t k = 0; k < K; ++k) {
        const uint64_t sk = scatter(k, bm.trace, bm.t);
        const double2 v = rho[(ia | sk) + ((ib | sk) << n)];
        s.x += v.x;
        s.y += v.y;
    }
    out[e] = s;
}

// 16 x 16 block of out = A A^dagger per workgroup of 256 threads
__global__ void __launch_bounds__(kRB) k_ptrace_sv(const double2* __restrict__ psi, BitMap bm, double2* __restrict__ out) {
    __shared__ double2 Ta[kTile][kTile + 1];  // A[a0 + r, k0 + c]
    __shared__ double2 Tb[kTile][kTile + 1];  // A[b0 + r, k0 + c]
    const uint64_t M = uint64_t(1) << bm.m;
    const uint64_t K = uint64_t(1) << bm.t;
    const uint64_t a0 = uint64_t(blockIdx.x) * kTile, b0 = uint64_t(blockIdx.y) * kTile;
    const int tr = threadIdx.x % kTile, tc = threadIdx.x / kTile;  // load: row tr, column tc
    const int ra = threadIdx.x % kTile, rb = threadIdx.x / kTile;  // compute: out[a0 + ra, b0 + rb]
    const bool in_a = a0 + tr < M, in_b = b0 + tr < M;
    const uint64_t base_a = in_a ? scatter(a0 + tr, bm.keep, bm.m) : 0;
    const uint64_t base_b = in_b ? scatter(b0 + tr, bm.keep, bm.m) : 0;
    double sr = 0.0, si = 0.0;
    for (uint64_t k0 = 0; k0 < K; k0 += kTile) {
        const uint64_t k = k0 + tc;
        const uint64_t sk = k < K ? scatter(k, bm.trace, bm.t) : 0;
        Ta[tr][tc] = (in_a && k < K) ? psi[base_a | sk] : make_double2(0.0, 0.0);
        Tb[tr][tc] = (in_b && k < K) ? psi[base_b | sk] : make_double2(0.0, 0.0);
        __syncthreads();
#pragma unroll
        for (int c = 0; c < kTile; ++c) {
            const double2 x = Ta[ra][c], y = Tb[rb][c];  // x conj(y)
            sr = fma(x.x, y.x, fma(x.y, y.y, sr));
            si = fma(x.y, y.x, fma(-x.x, y.y, si));
        }
        __syncthreads();
    }
    if (a0 + ra < M && b0 + rb < M) out[(a0 + ra) + (b0 + rb) * M] = make_double2(sr, si);
}

// sum_{i,j} conj(psi_i) rho_ij psi_j, per-block partials
__global__ void __launch_bounds__(kRB) k_sandwich(const double2* __restrict__ rho, const double2* __restrict__ psi, int n,
                                                  double2* parts) {
    __shared__ double red[kRB / 64];
    const uint64_t N = uint64_t(1) << n;
    double sr = 0.0, si = 0.0;
    for (uint64_t e = uint64_t(blockIdx.x) * kRB + threadIdx.x; e < N * N; e += uint64_t(gridDim.x) * kRB) {
        const uint64_t i = e & (N - 1), j = e >> n;
        const double2 r = rho[e], a = psi[i], b = psi[j];
        // conj(a) * r * b
        const double2 t = make_double2(a.x * r.x + a.y * r.y, a.x * r.y - a.y * r.x);
        sr = fma(t.x, b.x, fma(-t.y, b.y, sr));
        si = fma(t.x, b.y, fma(t.y, b.x, si));
    }
    sr = block_sum(sr, red);
    si = block_sum(si, red);
    if (threadIdx.x == 0) parts[blockIdx.x] = make_double2(sr, si);
}

}  // namespace

extern "C" int vqc_partial_trace(const vqc_state* s, int nsites, const int32_t* sites, vqc_state* out) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_state(s, "vqc_partial_trace"));
    VQC_TRY(vqc_check_state(out, "vqc_partial_trace"));
    const int n = s->nqubits;
    if (nsites < 0 || nsites >= n || (nsites && !sites))
        return vqc_fail(VQC_EINVAL, "vqc_partial_trace: trace 0..n-1 of the %d qubits (all of them is a trace)", n);
    if (!out->is_dm || out->nqubits != n - nsites)
        return vqc_fail(VQC_EINVAL, "vqc_partial_trace: out must be a density matrix of %d qubits", n - nsites);
    BitMap bm;
    bool traced[32] = {false};
    for (int j = 0; j < nsites; ++j) {
        if (sites[j] < 0 || sites[j] >= n) return vqc_fail(VQC_EINVAL, "site out of range.");  // ptrace.jl:44-48
        if (traced[sites[j]]) return vqc_fail(VQC_EINVAL, "duplicate sites not allowed.");
        traced[sites[j]] = true;
    }
    bm.m = 0;
    bm.t = 0;
    for (int q = 0; q < n; ++q) {
        if (traced[q]) bm.trace[bm.t++] = q;
        else bm.keep[bm.m++] = q;
    }
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    const uint64_t M = uint64_t(1) << bm.m;
    if (s->is_dm) {
        hipLaunchKernelGGL(k_ptrace_dm, dim3(grid_for(M * M, kRB)), dim3(kRB), 0, c->stream, s->d, n, bm, out->d);
    } else {
        const unsigned g = unsigned((M + kTile - 1) / kTile);
        hipLaunchKernelGGL(k_ptrace_sv, dim3(g, g), dim3(kRB), 0, c->stream, s->d, bm, out->d);
    }
    VQC_HIP(hipGetLastError());
    return VQC_OK;
}

extern "C" int vqc_dm_sandwich(const vqc_state* rho, const vqc_state* psi, double* out2) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_state(rho, "vqc_dm_sandwich"));
    VQC_TRY(vqc_check_state(psi, "vqc_dm_sandwich"));
    if (!rho->is_dm || psi->is_dm || psi->nqubits != rho->nqubits || !out2)
        return vqc_fail(VQC_EINVAL, "vqc_dm_sandwich: a density matrix and a state vector of the same size");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    constexpr int kParts = 1024;
    void* scr;
    VQC_TRY(vqc_scratch(size_t(kParts + 1) * 16, &scr));
    double2* parts = (double2*)scr;
    const uint64_t NN = uint64_t(1) << (2 * rho->nqubits);
    const unsigned grid = unsigned(std::min<uint64_t>(grid_for(NN, kRB), kParts));
    hipLaunchKernelGGL(k_sandwich, dim3(grid), dim3(kRB), 0, c->stream, rho->d, psi->d, rho->nqubits, parts);
    vqc_reduce_partials(parts, 1, int(grid), parts + kParts, c->stream);
    VQC_HIP(hipGetLastError());
    return vqc_fetch(parts + kParts, 1, out2);
}
