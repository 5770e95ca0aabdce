// vqc_qterm.hip -- product terms of any length: c * m_1 (x) ... (x) m_w on w
// sites, the QubitsTerm that the reference applies one site at a time
// (_apply_qterm_util!, src/hamiltonian/apply_qterms/apply_qterms.jl:47-59:
// vout = c v, then a one-qubit sweep per site) whenever an operator holds a
// term of more than 5 sites (apply_qterms.jl:20-25, expecs.jl:17,40,50,72,
// expecdiff.jl:7-16,37-56).
//
// Most factors of such terms (I, X, Y, Z, sigma+, sigma-, projectors, phases)
// have one non-zero per row: m[a, a ^ f] for a flip bit f. A term made only of
// those is applied in ONE pass over the state:
//   (T v)[a] = c' * (-1)^popcount(a & z) * [(a & pm) == pv] * prod_g r_g(a_g) * v[a ^ x]
// where the constant factors (I, X) are folded into c', the sign factors (Z,
// Y: r(1) = -r(0)) into c' and the z mask, the projectors (sigma+-, |0><0|,
// |1><1|: one of r(0), r(1) zero) into c' and the (pm, pv) test, and any other
// one-per-row factor is kept as a pair (r_g(0), r_g(1)) on its bit. Dense
// factors (H, Rx, ...) act on other sites, so they commute with that part: the
// term is the one-pass part followed by dense sweeps of up to 5 factors each
// (their kron product), the tuned gate kernels of vqc_gates.hip.
#include <cmath>

#include "vqc_internal.h"

int vqc_dense_sweep(double2* v, int nbits, int k, const int32_t* pos, const double* U, hipStream_t st);
void vqc_zero(double2* v, uint64_t len, hipStream_t st);

constexpr int kQB = 256;         // threads per block
constexpr int kQGen = 8;         // kept one-per-row factors per pass
constexpr int kQParts = 1024;    // partial sums per term

struct QtermDesc {
    uint64_t x, z, pm, pv;
    double2 c;
    int ng;
    int gbit[kQGen];
    double2 r0[kQGen], r1[kQGen];
};

// row factor of output index a (everything but the flip)
__device__ __forceinline__ double2 qterm_row(const QtermDesc& d, uint64_t a) {
    if ((a & d.pm) != d.pv) return make_double2(0.0, 0.0);
    double2 f = d.c;
    if (__popcll(a & d.z) & 1) f = make_double2(-f.x, -f.y);
    for (int g = 0; g < d.ng; ++g) f = cmul(f, (a >> d.gbit[g]) & 1 ? d.r1[g] : d.r0[g]);
    return f;
}

// out[a] (=|+=) f(a) in[a ^ x]
__global__ void __launch_bounds__(kQB) k_qterm_apply(const double2* __restrict__ in, double2* __restrict__ out, int nbits,
                                                     QtermDesc d, int accumulate) {
    const uint64_t L = uint64_t(1) << nbits;
    for (uint64_t a = uint64_t(xcd_block()) * kQB + threadIdx.x; a < L; a += uint64_t(gridDim.x) * kQB) {
        double2 r = cmul(qterm_row(d, a), in[a ^ d.x]);
        if (accumulate) r = cadd(r, out[a]);
        out[a] = r;
    }
}

// per-block partials of sum_a conj(bra[a]) f(a) ket[a ^ x]
__global__ void __launch_bounds__(kQB) k_qterm_expec(const double2* __restrict__ ket, const double2* __restrict__ bra, int nbits,
                                                     QtermDesc d, double2* __restrict__ parts) {
    __shared__ double red[kQB / 64];
    const uint64_t L = uint64_t(1) << nbits;
    double2 acc = make_double2(0.0, 0.0);
    for (uint64_t a = uint64_t(xcd_block()) * kQB + threadIdx.x; a < L; a += uint64_t(gridDim.x) * kQB) {
        const double2 f = qterm_row(d, a);
        if (f.x == 0.0 && f.y == 0.0) continue;
        acc = cfmaj(bra[a], cmul(f, ket[a ^ d.x]), acc);
    }
    const double sr = block_sum(acc.x, red);
    const double si = block_sum(acc.y, red);
    if (threadIdx.x == 0) parts[blockIdx.x] = make_double2(sr, si);
}

// per-block partials of tr(T rho) = sum_i f(i) rho[i ^ x, i] for T on the row
// bits of an n-qubit density matrix (rho[r, c] at r + c 2^n)
__global__ void __launch_bounds__(kQB) k_qterm_dm_expec(const double2* __restrict__ rho, int n, QtermDesc d,
                                                        double2* __restrict__ parts) {
    __shared__ double red[kQB / 64];
    const uint64_t N = uint64_t(1) << n;
    double2 acc = make_double2(0.0, 0.0);
    for (uint64_t i = uint64_t(xcd_block()) * kQB + threadIdx.x; i < N; i += uint64_t(gridDim.x) * kQB) {
        const double2 f = qterm_row(d, i);
        if (f.x == 0.0 && f.y == 0.0) continue;
        acc = cfma(f, rho[(i ^ d.x) + (i << n)], acc);
    }
    const double sr = block_sum(acc.x, red);
    const double si = block_sum(acc.y, red);
    if (threadIdx.x == 0) parts[blockIdx.x] = make_double2(sr, si);
}

// ---------------------------------------------------------------------------
// host: classify the factors of one term
// ---------------------------------------------------------------------------
namespace {

struct QtermPlan {
    QtermDesc d;
    std::vector<int32_t> dpos;     // dense factors: ascending positions
    std::vector<double> dmats;     // their 2x2 matrices, interleaved complex, column-major
};

inline bool czero(const double* m, int i) { return m[2 * i] == 0.0 && m[2 * i + 1] == 0.0; }
inline double2 cget(const double* m, int i) { return make_double2(m[2 * i], m[2 * i + 1]); }
inline double2 hmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

QtermPlan plan_term(const vqc_qterm_t& t, const int32_t* pos, const double* pool) {
    QtermPlan p;
    memset(&p.d, 0, sizeof p.d);
    p.d.c = make_double2(t.re, t.im);
    for (int j = 0; j < t.nsites; ++j) {
        const int b = pos[t.pos + j];
        const double* m = pool + 2 * (t.mat + 4 * int64_t(j));  // m00, m10, m01, m11
        const uint64_t bit = uint64_t(1) << b;
        int flip;
        if (czero(m, 1) && czero(m, 2)) flip = 0;        // diagonal
        else if (czero(m, 0) && czero(m, 3)) flip = 1;   // anti-diagonal
        else flip = -1;                                  // dense
        if (flip < 0) {
            p.dpos.push_back(b);
            for (int e = 0; e < 4; ++e) {
                p.dmats.push_back(m[2 * e]);
                p.dmats.push_back(m[2 * e + 1]);
            }
            continue;
        }
        // row values r(a) = m[a, a ^ flip]: column-major m[r + 2 c]
        const double2 r0 = cget(m, flip ? 2 : 0), r1 = cget(m, flip ? 1 : 3);
        if (flip) p.d.x |= bit;
        const bool z0 = r0.x == 0.0 && r0.y == 0.0, z1 = r1.x == 0.0 && r1.y == 0.0;
        if (z0 && z1) {  // the zero matrix: the term vanishes
            p.d.c = make_double2(0.0, 0.0);
        } else if (z1) {  // row 1 empty: output bit must be 0
            p.d.pm |= bit;
            p.d.c = hmul(p.d.c, r0);
        } else if (z0) {  // row 0 empty: output bit must be 1
            p.d.pm |= bit;
            p.d.pv |= bit;
            p.d.c = hmul(p.d.c, r1);
        } else if (r0.x == r1.x && r0.y == r1.y) {
            p.d.c = hmul(p.d.c, r0);
        } else if (r0.x == -r1.x && r0.y == -r1.y) {
            p.d.z |= bit;
            p.d.c = hmul(p.d.c, r0);
        } else if (p.d.ng < kQGen) {
            p.d.gbit[p.d.ng] = b;
            p.d.r0[p.d.ng] = r0;
            p.d.r1[p.d.ng] = r1;
            ++p.d.ng;
        } else {  // beyond the kept factors: a (diagonal or anti-diagonal) dense sweep
            p.d.x &= ~bit;
            p.dpos.push_back(b);
            for (int e = 0; e < 4; ++e) {
                p.dmats.push_back(m[2 * e]);
                p.dmats.push_back(m[2 * e + 1]);
            }
        }
    }
    return p;
}

// kron of the 2x2 factors f_0..f_{k-1} (factor j on the j-th smallest position:
// index bit j), column-major 2^k x 2^k, interleaved complex
std::vector<double> kron_factors(const double* f, int k) {
    const int D = 1 << k;
    std::vector<double> M(size_t(2) * D * D);
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) {
            double2 v = make_double2(1.0, 0.0);
            for (int j = 0; j < k; ++j) {
                const int e = ((r >> j) & 1) + 2 * ((c >> j) & 1);
                v = hmul(v, make_double2(f[8 * j + 2 * e], f[8 * j + 2 * e + 1]));
            }
            M[2 * (r + size_t(c) * D)] = v.x;
            M[2 * (r + size_t(c) * D) + 1] = v.y;
        }
    return M;
}

int check_qterms(int nsites_max, int nterms, const vqc_qterm_t* terms, const int32_t* pos, const double* pool) {
    if (nterms < 0 || (nterms && (!terms || !pos || !pool))) return vqc_fail(VQC_EINVAL, "bad product-term array");
    for (int t = 0; t < nterms; ++t) {
        const vqc_qterm_t& q = terms[t];
        if (q.nsites < 0 || q.nsites > 64 || q.pos < 0 || q.mat < 0)
            return vqc_fail(VQC_EINVAL, "product term %d: bad site count or offsets", t);
        for (int j = 0; j < q.nsites; ++j) {
            const int b = pos[q.pos + j];
            if (b < 0 || b >= nsites_max) return vqc_fail(VQC_EINVAL, "product term %d: site %d outside 0..%d", t, b, nsites_max - 1);
            if (j && b <= pos[q.pos + j - 1]) return vqc_fail(VQC_EINVAL, "product term %d: sites must be strictly ascending", t);
        }
    }
    return VQC_OK;
}

unsigned qgrid(uint64_t len) { return unsigned(std::max<uint64_t>(1, std::min<uint64_t>((len + kQB - 1) / kQB, 1u << 16))); }
unsigned qparts(uint64_t len) { return unsigned(std::max<uint64_t>(1, std::min<uint64_t>((len + kQB - 1) / kQB, kQParts))); }

void* g_qwork = nullptr;   // state-sized work buffer of terms with dense factors
size_t g_qwork_cap = 0;

int qwork(size_t bytes, double2** out) {
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    if (bytes > g_qwork_cap) {
        if (g_qwork) {
            VQC_HIP(hipStreamSynchronize(c->stream));
            VQC_HIP(hipFree(g_qwork));
            g_qwork = nullptr;
            g_qwork_cap = 0;
        }
        VQC_HIP(hipMalloc(&g_qwork, bytes));
        g_qwork_cap = bytes;
    }
    *out = (double2*)g_qwork;
    return VQC_OK;
}

// dst (=|+=) T src over nbits index bits, T's sites below nbits
int apply_one(const QtermPlan& p, const double2* src, double2* dst, int nbits, int accumulate, hipStream_t st) {
    const uint64_t L = uint64_t(1) << nbits;
    if (p.dpos.empty()) {
        hipLaunchKernelGGL(k_qterm_apply, dim3(qgrid(L)), dim3(kQB), 0, st, src, dst, nbits, p.d, accumulate);
        VQC_HIP(hipGetLastError());
        return VQC_OK;
    }
    double2* w = dst;
    if (accumulate) VQC_TRY(qwork(L * 16, &w));
    hipLaunchKernelGGL(k_qterm_apply, dim3(qgrid(L)), dim3(kQB), 0, st, src, w, nbits, p.d, 0);
    VQC_HIP(hipGetLastError());
    const int nd = int(p.dpos.size());
    for (int j0 = 0; j0 < nd; j0 += 5) {
        const int k = std::min(5, nd - j0);
        const std::vector<double> M = kron_factors(p.dmats.data() + 8 * j0, k);
        VQC_TRY(vqc_dense_sweep(w, nbits, k, p.dpos.data() + j0, M.data(), st));
    }
    if (accumulate) {
        // dst += w (the axpy of vqc_state_axpy on raw buffers)
        QtermDesc one;
        memset(&one, 0, sizeof one);
        one.c = make_double2(1.0, 0.0);
        hipLaunchKernelGGL(k_qterm_apply, dim3(qgrid(L)), dim3(kQB), 0, st, (const double2*)w, dst, nbits, one, 1);
        VQC_HIP(hipGetLastError());
    }
    return VQC_OK;
}

}  // namespace

// out (=|+=) sum_t T_t in. On a density matrix the sites are qubits and act on
// the row bits of vec(rho) (apply_qterms_dm.jl:11-27).
extern "C" int vqc_sv_apply_qterms(const vqc_state* in, vqc_state* out, int nterms, const vqc_qterm_t* terms, const int32_t* pos,
                                   const double* pool, int accumulate) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_qubits(in, "vqc_sv_apply_qterms"));
    VQC_TRY(vqc_check_state(out, "vqc_sv_apply_qterms"));
    if (in->nbits != out->nbits || in->is_dm != out->is_dm || in->d == out->d)
        return vqc_fail(VQC_EINVAL, "vqc_sv_apply_qterms: needs distinct states of one shape");
    VQC_TRY(check_qterms(in->nqubits, nterms, terms, pos, pool));
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    if (!nterms && !accumulate) vqc_zero(out->d, out->len(), c->stream);
    for (int t = 0; t < nterms; ++t) {
        const QtermPlan p = plan_term(terms[t], pos, pool);
        VQC_TRY(apply_one(p, in->d, out->d, in->nbits, accumulate || t > 0, c->stream));
    }
    VQC_HIP(hipGetLastError());
    return VQC_OK;
}

// <bra|T_t|ket> per term (bra = ket when null) and their sum; tr(T_t rho) per
// term on a density matrix (expecs_dm.jl:18: tr(m * rho) for long terms)
static int expec_qterms(const vqc_state* s, const vqc_state* bra, int nterms, const vqc_qterm_t* terms, const int32_t* pos,
                        const double* pool, double* out, double* sum2, const char* who) {
    VQC_TRY(vqc_check_qubits(s, who));
    if (bra && (bra->nbits != s->nbits || bra->is_dm != s->is_dm)) return vqc_fail(VQC_EINVAL, "%s: bra and ket differ", who);
    VQC_TRY(check_qterms(s->nqubits, nterms, terms, pos, pool));
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    const bool dm = s->is_dm;
    const uint64_t L = dm ? (uint64_t(1) << s->nqubits) : s->len();
    const unsigned np = qparts(L);
    std::vector<double> res(2 * size_t(std::max(nterms, 1)), 0.0);
    for (int t = 0; t < nterms; ++t) {
        const QtermPlan p = plan_term(terms[t], pos, pool);
        double2* w = nullptr;
        const double2* src = s->d;
        QtermDesc d = p.d;
        if (!p.dpos.empty()) {
            // dense factors: w = T s, then a plain overlap (bra or identity row factor)
            VQC_TRY(qwork(s->len() * 16, &w));
            VQC_TRY(apply_one(p, s->d, w, s->nbits, 0, c->stream));
            src = w;
            memset(&d, 0, sizeof d);
            d.c = make_double2(1.0, 0.0);
        }
        void* scr;
        VQC_TRY(vqc_scratch(size_t(np + 1) * 16, &scr));
        double2* parts = (double2*)scr;
        if (dm)
            hipLaunchKernelGGL(k_qterm_dm_expec, dim3(np), dim3(kQB), 0, c->stream, src, s->nqubits, d, parts);
        else
            hipLaunchKernelGGL(k_qterm_expec, dim3(np), dim3(kQB), 0, c->stream, src, bra ? bra->d : s->d, s->nbits, d, parts);
        vqc_reduce_partials(parts, 1, int(np), parts + np, c->stream);
        VQC_HIP(hipGetLastError());
        VQC_TRY(vqc_fetch(parts + np, 1, res.data() + 2 * t));
    }
    double sr = 0, si = 0;
    for (int t = 0; t < nterms; ++t) {
        if (out) {
            out[2 * t] = res[2 * t];
            out[2 * t + 1] = res[2 * t + 1];
        }
        sr += res[2 * t];
        si += res[2 * t + 1];
    }
    if (sum2) {
        sum2[0] = sr;
        sum2[1] = si;
    }
    return VQC_OK;
}

extern "C" int vqc_sv_expec_qterms(const vqc_state* ket, const vqc_state* bra, int nterms, const vqc_qterm_t* terms,
                                   const int32_t* pos, const double* pool, double* out, double* sum2) {
    VQC_API_LOCK;
    if (ket && ket->is_dm) return vqc_fail(VQC_EINVAL, "vqc_sv_expec_qterms: a state vector is needed");
    return expec_qterms(ket, bra, nterms, terms, pos, pool, out, sum2, "vqc_sv_expec_qterms");
}

extern "C" int vqc_dm_expec_qterms(const vqc_state* rho, int nterms, const vqc_qterm_t* terms, const int32_t* pos,
                                   const double* pool, double* out, double* sum2) {
    VQC_API_LOCK;
    if (rho && !rho->is_dm) return vqc_fail(VQC_EINVAL, "vqc_dm_expec_qterms: a density matrix is needed");
    return expec_qterms(rho, nullptr, nterms, terms, pos, pool, out, sum2, "vqc_dm_expec_qterms");
}
