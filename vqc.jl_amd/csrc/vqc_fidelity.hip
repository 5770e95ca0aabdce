// vqc_fidelity.hip -- fidelity(x::DensityMatrix, y::DensityMatrix) =
// real(tr(sqrt(x) * sqrt(y))) (src/densitymatrix.jl:65) on the device.
//
// With x = V diag(l) V^dagger and y = W diag(m) W^dagger (Hermitian
// eigendecompositions), sqrt(x) = V diag(s) V^dagger and
//   tr(sqrt(x) sqrt(y)) = sum_ij s_i t_j |C_ij|^2,   C = V^dagger W,
// so the trace needs two eigendecompositions (rocSOLVER zheevd), one complex
// product (rocBLAS zgemm, a plain library GEMM) and one weighted reduction
// (below). The square roots follow Julia's sqrt of a Hermitian matrix, which the
// reference's sqrt(storage(x)) reaches for a Hermitian density matrix: when every
// eigenvalue is at least -max|l| * eps * N (round-off) the roots are
// sqrt(max(l, 0)); otherwise they are sqrt(complex(l)), i sqrt(-l) for l < 0.
// The inputs are Hermitised first, (x + x^dagger) / 2, which is what zheevd's
// one-triangle read assumes.
#include <cfloat>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "vqc_internal.h"

namespace {

constexpr int kFB = 256;

// A <- (A + A^dagger) / 2 for an N x N column-major matrix, one (i, j), i <= j, pair per thread
__global__ void __launch_bounds__(kFB) k_hermitise(double2* __restrict__ A, int n) {
    const uint64_t N = uint64_t(1) << n;
    for (uint64_t e = uint64_t(blockIdx.x) * kFB + threadIdx.x; e < N * N; e += uint64_t(gridDim.x) * kFB) {
        const uint64_t i = e & (N - 1), j = e >> n;
        if (i > j) continue;
        const double2 a = A[i + j * N], b = A[j + i * N];
        const double2 h = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
        A[i + j * N] = h;
        A[j + i * N] = make_double2(h.x, -h.y);
    }
}

// Square roots of the eigenvalues l[0..N) as (re, im) pairs, with Julia's rule
// for round-off negatives (one workgroup).
__global__ void __launch_bounds__(kFB) k_eig_sqrt(const double* __restrict__ l, int N, double2* __restrict__ s) {
    __shared__ double red[kFB / 64];
    double mx = 0.0;
    for (int i = threadIdx.x; i < N; i += kFB) mx = fmax(mx, fabs(l[i]));
    // block max through the sum helper's layout: reduce by waves, then serially
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    mx = 0.0;
    for (int w = 0; w < kFB / 64; ++w) mx = fmax(mx, red[w]);
    const double l0 = -mx * DBL_EPSILON * N;
    int neg = 0;
    for (int i = threadIdx.x; i < N; i += kFB) neg |= l[i] < l0;
    neg = __syncthreads_or(neg);
    for (int i = threadIdx.x; i < N; i += kFB) {
        const double v = l[i];
        if (!neg) s[i] = make_double2(sqrt(fmax(v, 0.0)), 0.0);
        else s[i] = v >= 0 ? make_double2(sqrt(v), 0.0) : make_double2(0.0, sqrt(-v));
    }
}

// per-block partials of sum_ij Re(s_i t_j) |C_ij|^2
__global__ void __launch_bounds__(kFB) k_fid_sum(const double2* __restrict__ C, const double2* __restrict__ s,
                                                 const double2* __restrict__ t, int n, double2* __restrict__ parts) {
    __shared__ double red[kFB / 64];
    const uint64_t N = uint64_t(1) << n;
    double acc = 0.0;
    for (uint64_t e = uint64_t(blockIdx.x) * kFB + threadIdx.x; e < N * N; e += uint64_t(gridDim.x) * kFB) {
        const uint64_t i = e & (N - 1), j = e >> n;
        const double2 c = C[e];
        const double2 si = s[i], tj = t[j];
        acc = fma(fma(si.x, tj.x, -si.y * tj.y), fma(c.x, c.x, c.y * c.y), acc);
    }
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) parts[blockIdx.x] = make_double2(acc, 0.0);
}

rocblas_handle g_blas = nullptr;

struct FidWork {
    void* buf = nullptr;
    size_t cap = 0;
} g_fw;

}  // namespace

extern "C" int vqc_dm_fidelity(const vqc_state* x, const vqc_state* y, double* out) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_state(x, "vqc_dm_fidelity"));
    VQC_TRY(vqc_check_state(y, "vqc_dm_fidelity"));
    if (!out) return vqc_fail(VQC_EINVAL, "vqc_dm_fidelity: null output");
    if (!x->is_dm || !y->is_dm || x->nqubits != y->nqubits)
        return vqc_fail(VQC_EINVAL, "vqc_dm_fidelity: needs two density matrices of one size");
    const int n = x->nqubits;
    if (n > 14) return vqc_fail(VQC_EUNSUPPORTED, "vqc_dm_fidelity: at most 14 qubits (2^14 x 2^14 eigenproblems)");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    if (!g_blas) {
        if (rocblas_create_handle(&g_blas) != rocblas_status_success) return vqc_fail(VQC_EHIP, "rocblas_create_handle failed");
    }
    if (rocblas_set_stream(g_blas, c->stream) != rocblas_status_success) return vqc_fail(VQC_EHIP, "rocblas_set_stream failed");
    const size_t N = size_t(1) << n, NN = N * N;
    const unsigned nparts = 1024;
    // A, B (matrices, then eigenvectors), C = A^dagger B, eigenvalues l, m,
    // off-diagonals, roots s, t, info, partials
    const size_t bytes = 3 * NN * 16 + 4 * N * 8 + 2 * N * 16 + 64 + (nparts + 1) * 16;
    if (bytes > g_fw.cap) {
        if (g_fw.buf) {
            VQC_HIP(hipStreamSynchronize(c->stream));
            VQC_HIP(hipFree(g_fw.buf));
            g_fw.buf = nullptr;
            g_fw.cap = 0;
        }
        if (hipMalloc(&g_fw.buf, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return vqc_fail(VQC_ENOMEM, "vqc_dm_fidelity: cannot allocate %zu bytes of work space", bytes);
        }
        g_fw.cap = bytes;
    }
    char* p = (char*)g_fw.buf;
    double2* A = (double2*)p;
    double2* B = A + NN;
    double2* C = B + NN;
    double* l = (double*)(C + NN);
    double* m = l + N;
    double* e1 = m + N;
    double* e2 = e1 + N;
    double2* s = (double2*)(e2 + N);
    double2* t = s + N;
    rocblas_int* info = (rocblas_int*)(t + N);
    double2* parts = (double2*)(info + 16);
    VQC_HIP(hipMemcpyAsync(A, x->d, NN * 16, hipMemcpyDeviceToDevice, c->stream));
    VQC_HIP(hipMemcpyAsync(B, y->d, NN * 16, hipMemcpyDeviceToDevice, c->stream));
    const unsigned g = unsigned(std::min<size_t>((NN + kFB - 1) / kFB, 65536));
    hipLaunchKernelGGL(k_hermitise, dim3(g), dim3(kFB), 0, c->stream, A, n);
    hipLaunchKernelGGL(k_hermitise, dim3(g), dim3(kFB), 0, c->stream, B, n);
    VQC_HIP(hipGetLastError());
    if (rocsolver_zheevd(g_blas, rocblas_evect_original, rocblas_fill_lower, rocblas_int(N), (rocblas_double_complex*)A,
                         rocblas_int(N), l, e1, info) != rocblas_status_success ||
        rocsolver_zheevd(g_blas, rocblas_evect_original, rocblas_fill_lower, rocblas_int(N), (rocblas_double_complex*)B,
                         rocblas_int(N), m, e2, info + 1) != rocblas_status_success)
        return vqc_fail(VQC_EHIP, "vqc_dm_fidelity: rocsolver_zheevd failed");
    const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
    if (rocblas_zgemm(g_blas, rocblas_operation_conjugate_transpose, rocblas_operation_none, rocblas_int(N), rocblas_int(N),
                      rocblas_int(N), &one, (const rocblas_double_complex*)A, rocblas_int(N), (const rocblas_double_complex*)B,
                      rocblas_int(N), &zero, (rocblas_double_complex*)C, rocblas_int(N)) != rocblas_status_success)
        return vqc_fail(VQC_EHIP, "vqc_dm_fidelity: rocblas_zgemm failed");
    hipLaunchKernelGGL(k_eig_sqrt, dim3(1), dim3(kFB), 0, c->stream, l, int(N), s);
    hipLaunchKernelGGL(k_eig_sqrt, dim3(1), dim3(kFB), 0, c->stream, m, int(N), t);
    const unsigned np = unsigned(std::min<size_t>((NN + kFB - 1) / kFB, nparts));
    hipLaunchKernelGGL(k_fid_sum, dim3(np), dim3(kFB), 0, c->stream, C, s, t, n, parts);
    vqc_reduce_partials(parts, 1, int(np), parts + nparts, c->stream);
    VQC_HIP(hipGetLastError());
    double r[2];
    VQC_TRY(vqc_fetch(parts + nparts, 1, r));
    rocblas_int hinfo[2] = {0, 0};
    VQC_HIP(hipMemcpy(hinfo, info, sizeof hinfo, hipMemcpyDeviceToHost));
    if (hinfo[0] || hinfo[1]) return vqc_fail(VQC_EHIP, "vqc_dm_fidelity: the eigensolver did not converge (info %d, %d)", hinfo[0], hinfo[1]);
    *out = r[0];
    return VQC_OK;
}
