// vqc_measure.hip -- single-qubit measurement and post-selection
// (src/measure.jl, src/postselect.jl).
//
//   * P(bit q = v): a read-only sweep over the half of the state with bit q = v
//     (state vector) or over the diagonal of rho (density matrix); per-block
//     partials summed in a fixed order;
//   * the collapse is a one-qubit (non-unitary) gate through the tuned dense
//     sweep (state vector) or the one-pass U rho U^dagger sweep (density matrix);
//   * selecting an outcome out of the state (n -> n-1 qubits) is a compaction.
#include "vqc_internal.h"

int vqc_dense_sweep(double2* v, int nbits, int k, const int32_t* pos, const double* U, hipStream_t st);
int vqc_dm_gate_sweep(double2* v, int n, int k, const int32_t* pos, const double* U, hipStream_t st);

namespace {

constexpr int kMB = 256;
constexpr int kMParts = 1024;

__device__ __forceinline__ uint64_t with_bit(uint64_t g, int pos, int v) {
    const uint64_t lo = g & ((uint64_t(1) << pos) - 1);
    return ((g - lo) << 1) | lo | (uint64_t(v) << pos);
}

// sum over b with bit pos == v of |psi_b|^2 (compute_probability_zero, measure.jl:46-58)
__global__ void __launch_bounds__(kMB) k_prob_sv(const double2* __restrict__ psi, uint64_t half, int pos, int v,
                                                 double2* parts) {
    __shared__ double red[kMB / 64];
    double s = 0.0;
    for (uint64_t g = uint64_t(blockIdx.x) * kMB + threadIdx.x; g < half; g += uint64_t(gridDim.x) * kMB) {
        const double2 a = psi[with_bit(g, pos, v)];
        s = fma(a.x, a.x, fma(a.y, a.y, s));
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) parts[blockIdx.x] = make_double2(s, 0.0);
}

// sum over l with bit pos == v of real(rho[l, l]) (measure.jl:70-82)
__global__ void __launch_bounds__(kMB) k_prob_dm(const double2* __restrict__ rho, int n, uint64_t half, int pos, int v,
                                                 double2* parts) {
    __shared__ double red[kMB / 64];
    double s = 0.0;
    for (uint64_t g = uint64_t(blockIdx.x) * kMB + threadIdx.x; g < half; g += uint64_t(gridDim.x) * kMB) {
        const uint64_t l = with_bit(g, pos, v);
        s += rho[l + (l << n)].x;
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) parts[blockIdx.x] = make_double2(s, 0.0);
}

// out[c] = scale * in[src(c)] (n -> n-1 qubits) in the reference's qubit order:
// swap!(v, 1, pos), then row v of reshape(v, 2, L/2) (measure.jl:30-33,
// postselect.jl:47-50). So x = 2c + v, and src is x with bits 0 and pos
// exchanged: the measured qubit leaves and qubit 0 takes its place.
__global__ void __launch_bounds__(kMB) k_select(const double2* __restrict__ in, double2* __restrict__ out, uint64_t half, int pos,
                                                int v, double scale) {
    for (uint64_t c = uint64_t(blockIdx.x) * kMB + threadIdx.x; c < half; c += uint64_t(gridDim.x) * kMB) {
        uint64_t x = (c << 1) | uint64_t(v);
        if (((x ^ (x >> pos)) & 1) != 0) x ^= (uint64_t(1) << pos) | 1;
        const double2 a = in[x];
        out[c] = make_double2(a.x * scale, a.y * scale);
    }
}

int prob_of(const vqc_state* s, int pos, int v, double* p, hipStream_t st) {
    void* scr;
    VQC_TRY(vqc_scratch(size_t(kMParts + 1) * 16, &scr));
    double2* parts = (double2*)scr;
    const int n = s->nqubits;
    const uint64_t half = uint64_t(1) << (n - 1);
    const unsigned grid = unsigned(std::min<uint64_t>(grid_for(half, kMB), kMParts));
    if (s->is_dm) hipLaunchKernelGGL(k_prob_dm, dim3(grid), dim3(kMB), 0, st, s->d, n, half, pos, v, parts);
    else hipLaunchKernelGGL(k_prob_sv, dim3(grid), dim3(kMB), 0, st, s->d, half, pos, v, parts);
    vqc_reduce_partials(parts, 1, int(grid), parts + kMParts, st);
    VQC_HIP(hipGetLastError());
    double out2[2];
    VQC_TRY(vqc_fetch(parts + kMParts, 1, out2));
    *p = out2[0];
    return VQC_OK;
}

int check_qubit(const vqc_state* s, int pos, const char* who) {
    VQC_TRY(vqc_check_state(s, who));
    if (pos < 0 || pos >= s->nqubits)
        return vqc_fail(VQC_EINVAL, "%s: qubit %d out of range for %d qubits", who, pos, s->nqubits);
    return VQC_OK;
}

// a one-qubit matrix on a state: dense sweep (SV) or M rho M^dagger (DM)
int apply_1q(vqc_state* s, int pos, const double* M, hipStream_t st) {
    const int32_t p[1] = {pos};
    if (s->is_dm) return vqc_dm_gate_sweep(s->d, s->nqubits, 1, p, M, st);
    return vqc_dense_sweep(s->d, s->nbits, 1, p, M, st);
}

}  // namespace

extern "C" int vqc_prob(const vqc_state* s, int pos, int value, double* p) {
    VQC_API_LOCK;
    VQC_TRY(check_qubit(s, pos, "vqc_prob"));
    if ((value != 0 && value != 1) || !p) return vqc_fail(VQC_EINVAL, "vqc_prob: value must be 0 or 1 and p non-null");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    return prob_of(s, pos, value, p, c->stream);
}

extern "C" int vqc_measure(vqc_state* s, int pos, double u, int auto_reset, int* outcome, double* prob) {
    VQC_API_LOCK;
    VQC_TRY(check_qubit(s, pos, "vqc_measure"));
    if (!(u >= 0.0 && u < 1.0)) return vqc_fail(VQC_EINVAL, "vqc_measure: u must lie in [0, 1)");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    double p0;
    VQC_TRY(prob_of(s, pos, 0, &p0, c->stream));
    // discrete_sample([p0, 1 - p0]) with rand() = u (sampling.jl:2-18)
    const int i = u < p0 ? 0 : 1;
    const double p = i == 0 ? p0 : 1.0 - p0;
    if (!(p > 0.0)) return vqc_fail(VQC_EINVAL, "vqc_measure: sampled outcome has zero probability");
    // |0><i| / sqrt(p) with auto_reset, |i><i| / sqrt(p) without (measure.jl:7-12)
    double M[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int row = auto_reset ? 0 : i;
    M[2 * (row + 2 * i)] = 1.0 / std::sqrt(p);
    VQC_TRY(apply_1q(s, pos, M, c->stream));
    if (outcome) *outcome = i;
    if (prob) *prob = p;
    return VQC_OK;
}

extern "C" int vqc_post_select(vqc_state* s, int pos, int value, double* norm) {
    VQC_API_LOCK;
    VQC_TRY(check_qubit(s, pos, "vqc_post_select"));
    if (s->is_dm) return vqc_fail(VQC_EINVAL, "vqc_post_select: state vectors only (postselect.jl:3-8)");
    if (value != 0 && value != 1) return vqc_fail(VQC_EINVAL, "vqc_post_select: value must be 0 or 1");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    double p;
    VQC_TRY(prob_of(s, pos, value, &p, c->stream));
    if (!(p > 0.0)) return vqc_fail(VQC_EINVAL, "vqc_post_select: the selected outcome has zero probability");
    // the UP / DOWN projector, then the state divided by its norm (postselect.jl:3-8)
    double M[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    M[2 * (value + 2 * value)] = 1.0 / std::sqrt(p);
    VQC_TRY(apply_1q(s, pos, M, c->stream));
    if (norm) *norm = std::sqrt(p);
    return VQC_OK;
}

extern "C" int vqc_sv_select(const vqc_state* s, int pos, int value, vqc_state* out, double* norm) {
    VQC_API_LOCK;
    VQC_TRY(check_qubit(s, pos, "vqc_sv_select"));
    VQC_TRY(vqc_check_state(out, "vqc_sv_select"));
    if (s->is_dm || out->is_dm) return vqc_fail(VQC_EINVAL, "vqc_sv_select: state vectors only");
    if (value != 0 && value != 1) return vqc_fail(VQC_EINVAL, "vqc_sv_select: value must be 0 or 1");
    if (s->nqubits < 1 || out->nqubits != s->nqubits - 1)
        return vqc_fail(VQC_EINVAL, "vqc_sv_select: out must have one qubit less than the input");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    double p;
    VQC_TRY(prob_of(s, pos, value, &p, c->stream));
    if (!(p > 0.0)) return vqc_fail(VQC_EINVAL, "vqc_sv_select: the selected outcome has zero probability");
    const uint64_t half = uint64_t(1) << (s->nqubits - 1);
    const unsigned grid = unsigned(std::min<uint64_t>(grid_for(half, kMB), 65536));
    hipLaunchKernelGGL(k_select, dim3(grid), dim3(kMB), 0, c->stream, s->d, out->d, half, pos, value, 1.0 / std::sqrt(p));
    VQC_HIP(hipGetLastError());
    if (norm) *norm = std::sqrt(p);
    return VQC_OK;
}
