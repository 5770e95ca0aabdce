// vqc_oracle.cpp -- CPU restatement of VQC.jl's gate / expectation / adjoint
// kernels in FP64, used ONLY as test infrastructure: the parity reference that
// libvqc_hip.so's results are compared against, and the timed CPU baseline
// ("reference algorithm, restated").
// Nothing in the product library links or calls this file.
//
// The reference is Julia (no Julia toolchain in this image), so each routine
// below restates one reference function; the file:line it follows is given
// with it. Threading follows src/auxiliary/parallel_for.jl: one contiguous
// block of the iteration space per thread, serial when the iteration count is
// below MIN_SIZE = 1024 (parallel_for.jl:3,12-23). Reductions add per-thread
// partial sums (parallel_for.jl:28-61); they are summed in thread order here
// instead of through atomics so that repeated runs agree bit for bit.
//
// Layout conventions are those of include/vqc_hip.h (the same program structs
// are reused so the GPU engine and this oracle consume identical inputs).

#include <omp.h>

#include <algorithm>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../include/vqc_hip.h"

typedef std::complex<double> cd;

static int g_threads = 1;
static const int64_t MIN_SIZE = 1024;  // parallel_for.jl:3

extern "C" void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }
extern "C" int orc_get_threads(void) { return g_threads; }

// parallel_run (parallel_for.jl:12-23): f(ist, ifn) over contiguous blocks.
template <class F>
static void parallel_run(int64_t total, F&& f) {
    const int nt = g_threads;
    if (total >= MIN_SIZE && nt > 1) {
#pragma omp parallel for num_threads(nt) schedule(static, 1)
        for (int t = 0; t < nt; ++t) {
            int64_t ist = total * t / nt;
            int64_t ifn = total * (t + 1) / nt - 1;
            f(ist, ifn);
        }
    } else if (total > 0) {
        f(0, total - 1);
    }
}

// parallel_sum for complex results (parallel_for.jl:44-61).
template <class F>
static cd parallel_sum(int64_t total, F&& f) {
    const int nt = g_threads;
    if (total >= MIN_SIZE && nt > 1) {
        std::vector<cd> part(nt);
#pragma omp parallel for num_threads(nt) schedule(static, 1)
        for (int t = 0; t < nt; ++t) {
            int64_t ist = total * t / nt;
            int64_t ifn = total * (t + 1) / nt - 1;
            part[t] = f(ist, ifn);
        }
        cd r = 0;
        for (int t = 0; t < nt; ++t) r += part[t];
        return r;
    }
    return total > 0 ? f(0, total - 1) : cd(0);
}

// Index of the i-th group: zero bits inserted at the ascending positions. This
// is the reference's mask form l = (4i & m3) | (2i & m2) | (i & m1)
// (e.g. specific/czgate.jl:36-43, threaded_long_range.jl:8-17).
static inline uint64_t insert_zeros(uint64_t i, const int* pos, int k) {
    for (int j = 0; j < k; ++j) {
        uint64_t lo = i & ((uint64_t(1) << pos[j]) - 1);
        i = ((i - lo) << 1) | lo;
    }
    return i;
}

static void group_offsets(const int* pos, int k, uint64_t* off) {
    for (int j = 0; j < (1 << k); ++j) {
        uint64_t o = 0;
        for (int b = 0; b < k; ++b)
            if (j >> b & 1) o |= uint64_t(1) << pos[b];
        off[j] = o;
    }
}

static bool sorted_valid(int nbits, int k, const int* pos) {
    if (k < 1 || k > 6) return false;
    for (int j = 0; j < k; ++j) {
        if (pos[j] < 0 || pos[j] >= nbits) return false;
        if (j && pos[j] <= pos[j - 1]) return false;
    }
    return true;
}

// ---------------------------------------------------------------------------
// Dense k-qubit gate, v <- U_pos v.
// Restates _apply_gate_threaded2!(key, U, v): threaded_short_range.jl:7-39
// (1q "H": 8 contiguous amplitudes per iteration when key > 3), :44-98
// (1q "L"), :113-362 (2q HH/LH/LL), :378-751 (3q), and
// threaded_long_range.jl:3-132 (4q/5q gather-multiply-scatter), with the
// serial loops of serial_short_range.jl:2-71 below 32 amplitudes. The tiling
// variants only reorder independent groups; the arithmetic per group is the
// same 2^k x 2^k product, so one chunked loop states all of them: C
// consecutive groups per iteration when every target lies above log2(C).
// ---------------------------------------------------------------------------
template <int C>
static void apply_dense_chunked(int nbits, cd* v, int k, const int* pos, const cd* U) {
    const int K = 1 << k;
    uint64_t off[64];
    group_offsets(pos, k, off);
    const int64_t total = (int64_t(1) << (nbits - k)) / C;
    parallel_run(total, [&](int64_t ist, int64_t ifn) {
        cd in[64], out[64];
        for (int64_t i = ist; i <= ifn; ++i) {
            uint64_t base = insert_zeros(uint64_t(i) * C, pos, k);
            for (int c = 0; c < C; ++c) {
                uint64_t b = base + c;
                for (int j = 0; j < K; ++j) in[j] = v[b + off[j]];
                for (int r = 0; r < K; ++r) {
                    cd s = 0;
                    for (int j = 0; j < K; ++j) s += U[r + j * K] * in[j];
                    out[r] = s;
                }
                for (int j = 0; j < K; ++j) v[b + off[j]] = out[j];
            }
        }
    });
}

extern "C" int orc_apply_dense(int nbits, cd* v, int k, const int* pos, const cd* U) {
    if (!sorted_valid(nbits, k, pos)) return -1;
    // generic.jl:50: serial below 32 amplitudes; threaded_short_range.jl picks
    // the "H" tiling (8 contiguous) when the lowest target is above 3 (1-based).
    if (k <= 2 && pos[0] >= 3)
        apply_dense_chunked<8>(nbits, v, k, pos, U);
    else if (k == 3 && pos[0] >= 2)
        apply_dense_chunked<4>(nbits, v, k, pos, U);
    else
        apply_dense_chunked<1>(nbits, v, k, pos, U);
    return 0;
}

// ---------------------------------------------------------------------------
// Named gates: src/applygates/specific/*.jl. pos is in the gate's own order.
// ---------------------------------------------------------------------------
extern "C" int orc_apply_named(int nbits, cd* v, int kind, const int* pos_in, double phi) {
    int k = 1;
    if (kind == VQC_GATE_CZ || kind == VQC_GATE_CNOT || kind == VQC_GATE_CPHASE || kind == VQC_GATE_SWAP ||
        kind == VQC_GATE_ISWAP || kind == VQC_GATE_ISWAPDAG)
        k = 2;
    else if (kind == VQC_GATE_FREDKIN || kind == VQC_GATE_CCPHASE)
        k = 3;
    else if (kind != VQC_GATE_X && kind != VQC_GATE_PHASE)
        return -1;
    int pos[3];
    for (int j = 0; j < k; ++j) pos[j] = pos_in[j];
    std::sort(pos, pos + k);
    if (!sorted_valid(nbits, k, pos)) return -1;
    const int64_t total = int64_t(1) << (nbits - k);
    const cd eph = std::exp(cd(0, phi));
    const uint64_t p0 = uint64_t(1) << pos_in[0];
    const uint64_t p1 = k > 1 ? uint64_t(1) << pos_in[1] : 0;
    const uint64_t p2 = k > 2 ? uint64_t(1) << pos_in[2] : 0;
    parallel_run(total, [&](int64_t ist, int64_t ifn) {
        for (int64_t i = ist; i <= ifn; ++i) {
            uint64_t l = insert_zeros(uint64_t(i), pos, k);
            switch (kind) {
                case VQC_GATE_X:  // phasegate.jl:27-43
                    std::swap(v[l], v[l + p0]);
                    break;
                case VQC_GATE_PHASE:  // phasegate.jl:6-22
                    v[l + p0] *= eph;
                    break;
                case VQC_GATE_CZ:  // czgate.jl:6-25
                    v[l + p0 + p1] = -v[l + p0 + p1];
                    break;
                case VQC_GATE_CNOT:  // czgate.jl:30-68 (control = positions(g)[1])
                    std::swap(v[l + p0], v[l + p0 + p1]);
                    break;
                case VQC_GATE_CPHASE:  // czgate.jl:73-95
                    v[l + p0 + p1] *= eph;
                    break;
                case VQC_GATE_SWAP:  // swapgates.jl:3-23
                    std::swap(v[l + p0], v[l + p1]);
                    break;
                case VQC_GATE_ISWAP:
                case VQC_GATE_ISWAPDAG: {  // swapgates.jl:25-49
                    cd c = kind == VQC_GATE_ISWAP ? cd(0, 1) : cd(0, -1);
                    cd a = v[l + p0], b = v[l + p1];
                    v[l + p0] = c * b;
                    v[l + p1] = c * a;
                    break;
                }
                case VQC_GATE_FREDKIN:  // fredkingate.jl:4-77 (control = positions(g)[1])
                    std::swap(v[l + p0 + p1], v[l + p0 + p2]);
                    break;
                case VQC_GATE_CCPHASE:  // fredkingate.jl:80-104
                    v[l + p0 + p1 + p2] *= eph;
                    break;
            }
        }
    });
    return 0;
}

// ---------------------------------------------------------------------------
// vout += M_pos v. Restates the accumulate kernels behind
// _apply_gate_threaded2!(key, U, v, vout): short_range_threaded.jl:7-1515,
// long_range_threaded.jl:3-177, short_range_serial.jl:3-96.
// ---------------------------------------------------------------------------
extern "C" int orc_accum_dense(int nbits, const cd* v, cd* vout, int k, const int* pos, const cd* M) {
    if (!sorted_valid(nbits, k, pos)) return -1;
    const int K = 1 << k;
    uint64_t off[64];
    group_offsets(pos, k, off);
    parallel_run(int64_t(1) << (nbits - k), [&](int64_t ist, int64_t ifn) {
        cd in[64];
        for (int64_t i = ist; i <= ifn; ++i) {
            uint64_t b = insert_zeros(uint64_t(i), pos, k);
            for (int j = 0; j < K; ++j) in[j] = v[b + off[j]];
            for (int r = 0; r < K; ++r) {
                cd s = 0;
                for (int j = 0; j < K; ++j) s += M[r + j * K] * in[j];
                vout[b + off[r]] += s;
            }
        }
    });
    return 0;
}

// ---------------------------------------------------------------------------
// <vc|M_pos|v> (vc = v when null). Restates expectation_value(key, M, v[, v_c]):
// expec_threaded.jl:5-648, expec_threaded_2.jl:7-788,
// expec_high_qubit_terms(_2).jl, expec_serial.jl:2-128 (dot(vc_i, M, v_i) per group).
// ---------------------------------------------------------------------------
extern "C" int orc_expec_dense(int nbits, const cd* v, const cd* vc, int k, const int* pos, const cd* M,
                               double* out2) {
    if (!sorted_valid(nbits, k, pos)) return -1;
    if (!vc) vc = v;
    const int K = 1 << k;
    uint64_t off[64];
    group_offsets(pos, k, off);
    cd r = parallel_sum(int64_t(1) << (nbits - k), [&](int64_t ist, int64_t ifn) {
        cd in[64], s = 0;
        for (int64_t i = ist; i <= ifn; ++i) {
            uint64_t b = insert_zeros(uint64_t(i), pos, k);
            for (int j = 0; j < K; ++j) in[j] = v[b + off[j]];
            for (int rr = 0; rr < K; ++rr) {
                cd t = 0;
                for (int j = 0; j < K; ++j) t += M[rr + j * K] * in[j];
                s += std::conj(vc[b + off[rr]]) * t;
            }
        }
        return s;
    });
    out2[0] = r.real();
    out2[1] = r.imag();
    return 0;
}

// <a|b> (LinearAlgebra.dot over storage, statevector.jl:58).
extern "C" void orc_dot(int nbits, const cd* a, const cd* b, double* out2) {
    cd r = parallel_sum(int64_t(1) << nbits, [&](int64_t ist, int64_t ifn) {
        cd s = 0;
        for (int64_t i = ist; i <= ifn; ++i) s += std::conj(a[i]) * b[i];
        return s;
    });
    out2[0] = r.real();
    out2[1] = r.imag();
}

// ---------------------------------------------------------------------------
// One QubitsTerm with more sites than the dense kernels take: vout = c * v,
// then each site's 2x2 operator as a one-qubit gate
// (_apply_qterm_util!, apply_qterms.jl:47-59).
// ---------------------------------------------------------------------------
static void apply_qterm(int nbits, const cd* v, cd* vout, int nsites, const int* sites, const cd* ops, cd coef) {
    const int64_t L = int64_t(1) << nbits;
    parallel_run(L, [&](int64_t ist, int64_t ifn) {
        for (int64_t i = ist; i <= ifn; ++i) vout[i] = coef * v[i];
    });
    for (int s = 0; s < nsites; ++s) orc_apply_dense(nbits, vout, 1, &sites[s], ops + 4 * s);
}

// Product terms of any length (vqc_qterm_t: c, sites, 2x2 factors in a complex
// pool). out = sum_t T_t v with the workspace of (m::QubitsOperator)(v) for
// long terms: vout = 0, then _apply_qterm_util!(term, v, workspace) and
// vout .+= workspace per term (apply_qterms.jl:20-25). On a density matrix the
// vector is vec(rho) and the sites are row bits (apply_qterms_dm.jl:11-27).
extern "C" int orc_apply_qterms(int nbits, const cd* v, cd* vout, int nterms, const vqc_qterm_t* terms, const int* pos,
                                const cd* pool) {
    const int64_t L = int64_t(1) << nbits;
    std::vector<cd> work(L);
    std::fill(vout, vout + L, cd(0));
    for (int t = 0; t < nterms; ++t) {
        apply_qterm(nbits, v, work.data(), terms[t].nsites, pos + terms[t].pos, pool + terms[t].mat, cd(terms[t].re, terms[t].im));
        parallel_run(L, [&](int64_t ist, int64_t ifn) {
            for (int64_t i = ist; i <= ifn; ++i) vout[i] += work[i];
        });
    }
    return 0;
}

// <vc|T_t|v> per term: dot(state_c, m * state) (expecs.jl:17,50 for one long
// term; :40,72 for the operator, summed in term order).
extern "C" int orc_expec_qterms(int nbits, const cd* v, const cd* vc, int nterms, const vqc_qterm_t* terms, const int* pos,
                                const cd* pool, double* out, double* sum2) {
    if (!vc) vc = v;
    const int64_t L = int64_t(1) << nbits;
    std::vector<cd> work(L);
    cd total = 0;
    for (int t = 0; t < nterms; ++t) {
        apply_qterm(nbits, v, work.data(), terms[t].nsites, pos + terms[t].pos, pool + terms[t].mat, cd(terms[t].re, terms[t].im));
        double r2[2];
        orc_dot(nbits, vc, work.data(), r2);
        if (out) {
            out[2 * t] = r2[0];
            out[2 * t + 1] = r2[1];
        }
        total += cd(r2[0], r2[1]);
    }
    sum2[0] = total.real();
    sum2[1] = total.imag();
    return 0;
}

// tr(T_t rho) per term: tr(m * rho) (expecs_dm.jl:10-26), m * rho through the
// row-bit route above, then the trace of the result.
extern "C" int orc_dm_expec_qterms(int n, const cd* rho, int nterms, const vqc_qterm_t* terms, const int* pos, const cd* pool,
                                   double* out, double* sum2) {
    const int64_t N = int64_t(1) << n, L = N * N;
    std::vector<cd> work(L);
    cd total = 0;
    for (int t = 0; t < nterms; ++t) {
        apply_qterm(2 * n, rho, work.data(), terms[t].nsites, pos + terms[t].pos, pool + terms[t].mat, cd(terms[t].re, terms[t].im));
        cd tr = 0;
        for (int64_t i = 0; i < N; ++i) tr += work[i + i * N];
        if (out) {
            out[2 * t] = tr.real();
            out[2 * t + 1] = tr.imag();
        }
        total += tr;
    }
    sum2[0] = total.real();
    sum2[1] = total.imag();
    return 0;
}

// Pauli strings through the reference's >5-site route: state_2 = sum_t P_t v
// term by term (apply_qterms.jl:13-28), then dot(bra, state_2) (expecs.jl:40,68).
// out gets one complex per term; sum2 the total.
extern "C" int orc_expec_pauli(int nbits, const cd* v, const cd* vc, int nterms, const vqc_pauli_t* terms,
                               double* out, double* sum2) {
    if (!vc) vc = v;
    const int64_t L = int64_t(1) << nbits;
    std::vector<cd> work(L);
    const cd I(0, 1);
    const cd X[4] = {0, 1, 1, 0}, Y[4] = {0, I, -I, 0}, Z[4] = {1, 0, 0, -1};
    cd total = 0;
    for (int t = 0; t < nterms; ++t) {
        std::vector<int> sites;
        std::vector<cd> ops;
        for (int q = 0; q < nbits; ++q) {
            bool x = terms[t].x >> q & 1, z = terms[t].z >> q & 1;
            if (!x && !z) continue;
            sites.push_back(q);
            const cd* m = x && z ? Y : (x ? X : Z);
            ops.insert(ops.end(), m, m + 4);
        }
        apply_qterm(nbits, v, work.data(), int(sites.size()), sites.data(), ops.data(), cd(terms[t].re, terms[t].im));
        double r2[2];
        orc_dot(nbits, vc, work.data(), r2);
        out[2 * t] = r2[0];
        out[2 * t + 1] = r2[1];
        total += cd(r2[0], r2[1]);
    }
    sum2[0] = total.real();
    sum2[1] = total.imag();
    return 0;
}

// out = sum_t c_t P_t v, per term as above ((m::QubitsOperator)(v) fallback).
extern "C" int orc_apply_pauli_sum(int nbits, const cd* v, cd* vout, int nterms, const vqc_pauli_t* terms) {
    const int64_t L = int64_t(1) << nbits;
    std::vector<cd> work(L);
    const cd I(0, 1);
    const cd X[4] = {0, 1, 1, 0}, Y[4] = {0, I, -I, 0}, Z[4] = {1, 0, 0, -1};
    std::fill(vout, vout + L, cd(0));
    for (int t = 0; t < nterms; ++t) {
        std::vector<int> sites;
        std::vector<cd> ops;
        for (int q = 0; q < nbits; ++q) {
            bool x = terms[t].x >> q & 1, z = terms[t].z >> q & 1;
            if (!x && !z) continue;
            sites.push_back(q);
            const cd* m = x && z ? Y : (x ? X : Z);
            ops.insert(ops.end(), m, m + 4);
        }
        apply_qterm(nbits, v, work.data(), int(sites.size()), sites.data(), ops.data(), cd(terms[t].re, terms[t].im));
        parallel_run(L, [&](int64_t ist, int64_t ifn) {
            for (int64_t i = ist; i <= ifn; ++i) vout[i] += work[i];
        });
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Density matrices.
// ---------------------------------------------------------------------------
// _dm_apply_threaded! (generic.jl:54-64): U on pos, then conj(U) on pos .+ n.
extern "C" int orc_dm_apply_gate(int n, cd* rho, int k, const int* pos, const cd* U) {
    if (!sorted_valid(n, k, pos)) return -1;
    const int K = 1 << k;
    std::vector<cd> Uc(K * K);
    for (int i = 0; i < K * K; ++i) Uc[i] = std::conj(U[i]);
    int pos2[5];
    for (int j = 0; j < k; ++j) pos2[j] = pos[j] + n;
    orc_apply_dense(2 * n, rho, k, pos, U);
    orc_apply_dense(2 * n, rho, k, pos2, Uc.data());
    return 0;
}

// _qmap_apply_threaded! / _dagger / _inverse (apply_qmaps.jl:52-86): one
// 2N-site dense sweep with the supermatrix on (pos..., pos.+n...). mode 1 uses S'.
extern "C" int orc_dm_apply_superop(int n, cd* rho, int N, const int* pos, const cd* S, int mode) {
    if (!sorted_valid(n, N, pos) || N > 2) return -1;
    int all[4];
    for (int j = 0; j < N; ++j) {
        all[j] = pos[j];
        all[j + N] = pos[j] + n;
    }
    const int K = 1 << (2 * N);
    std::vector<cd> m(S, S + K * K);
    if (mode == 1)
        for (int r = 0; r < K; ++r)
            for (int c = 0; c < K; ++c) m[r + c * K] = std::conj(S[c + r * K]);
    return orc_apply_dense(2 * n, rho, 2 * N, all, m.data());
}

// tr(M rho) for a k-site key: restates expectation_value_serial on the
// 2^k x 2^k diagonal blocks, r += tr(m * block) (expec_dm_serial.jl:23-129).
// The reference takes this route for k <= 3 and tr(m * rho) above (same sum).
extern "C" int orc_dm_expec(int n, const cd* rho, int k, const int* pos, const cd* M, double* out2) {
    if (!sorted_valid(n, k, pos)) return -1;
    const int K = 1 << k;
    const uint64_t N = uint64_t(1) << n;
    uint64_t off[32];
    group_offsets(pos, k, off);
    cd r = 0;  // the reference's DM expectation is serial (expecs_dm.jl:17-27)
    for (uint64_t i = 0; i < (N >> k); ++i) {
        uint64_t b = insert_zeros(i, pos, k);
        for (int a = 0; a < K; ++a)
            for (int c = 0; c < K; ++c) r += M[a + c * K] * rho[(b + off[c]) + (b + off[a]) * N];
    }
    out2[0] = r.real();
    out2[1] = r.imag();
    return 0;
}

// ---------------------------------------------------------------------------
// Circuits and their adjoints.
// ---------------------------------------------------------------------------
static void dagger(const cd* U, int K, cd* out) {
    for (int r = 0; r < K; ++r)
        for (int c = 0; c < K; ++c) out[r + c * K] = std::conj(U[c + r * K]);
}

static int check_ops(int nbits_state, int is_dm, int nops, const vqc_op_t* ops) {
    for (int i = 0; i < nops; ++i) {
        const vqc_op_t& o = ops[i];
        if (o.kind == VQC_OP_CHANNEL && !is_dm) return -1;
        if (!sorted_valid(nbits_state, o.nq, o.pos)) return -1;
        if (o.nparams < 0 || o.nparams > 3) return -1;
    }
    return 0;
}

// apply!(circuit, state) (generic.jl:26-31; apply_qmaps.jl:8-16 for channels).
extern "C" int orc_apply_circuit(int n, int is_dm, cd* v, int nops, const vqc_op_t* ops, const cd* pool) {
    if (check_ops(n, is_dm, nops, ops)) return -1;
    for (int i = 0; i < nops; ++i) {
        const vqc_op_t& o = ops[i];
        if (o.kind == VQC_OP_CHANNEL)
            orc_dm_apply_superop(n, v, o.nq, o.pos, pool + o.mat, 0);
        else if (is_dm)
            orc_dm_apply_gate(n, v, o.nq, o.pos, pool + o.mat);
        else
            orc_apply_dense(n, v, o.nq, o.pos, pool + o.mat);
    }
    return 0;
}

// back_propagate(Delta, circuit, y) for a state vector (circuitdiff.jl:23-32,
// 61-78): per op in reverse, Delta <- m' Delta, y <- m' y, then
// grad = real(expectation(y, item, Delta)) for item in differentiate(m). The op
// carries G = m' dm/dtheta; item = G' so that real(<y|G'|Delta>) =
// real(<Delta|G|y>), which is what is evaluated here. y and delta are updated
// in place (they end as the circuit's input and the cotangent on it).
static void sv_back_propagate(int n, int nops, const vqc_op_t* ops, const cd* pool, int nparams, cd* y, cd* delta,
                              double* grad) {
    for (int p = 0; p < nparams; ++p) grad[p] = 0;
    std::vector<cd> ud;
    for (int i = nops - 1; i >= 0; --i) {
        const vqc_op_t& o = ops[i];
        const int K = 1 << o.nq;
        ud.assign(K * K, 0);
        dagger(pool + o.mat, K, ud.data());
        orc_apply_dense(n, delta, o.nq, o.pos, ud.data());
        orc_apply_dense(n, y, o.nq, o.pos, ud.data());
        for (int j = 0; j < o.nparams; ++j) {
            if (o.slot[j] < 0) continue;
            double g2[2];
            orc_expec_dense(n, y, delta, o.nq, o.pos, pool + o.gen[j], g2);
            grad[o.slot[j]] += g2[0];
        }
    }
}

// Zygote.gradient(c -> real(expectation(H, c * psi0)), c) for a state vector:
//  forward   y = circuit * psi0                       (circuitdiff.jl:2-4)
//  energy    E = sum_keys expectation_value(key, M, y) (expecs.jl:22-28)
//  seed      Delta = (conj(z) H + z H') y, z = 1      (expecdiff.jl:30-46)
//  reverse   per op: Delta <- m' Delta, y <- m' y, then
//            grad = real(expectation(y, item, Delta)) for item in differentiate(m)
//            (circuitdiff.jl:23-32), collected in forward order (:61-75).
// The op carries G = m' dm/dtheta; item = G' so that
// real(<y|G'|Delta>) = real(<Delta|G|y>), which is what is evaluated here.
extern "C" int orc_sv_energy_grad(int n, const cd* psi0, int nops, const vqc_op_t* ops, const cd* pool,
                                  int nterms, const vqc_term_t* terms, const cd* hpool, int nparams,
                                  double* energy2, double* grad, cd* psi_out) {
    if (check_ops(n, 0, nops, ops)) return -1;
    const int64_t L = int64_t(1) << n;
    std::vector<cd> y(psi0, psi0 + L), delta(L, cd(0));
    orc_apply_circuit(n, 0, y.data(), nops, ops, pool);
    if (psi_out) std::copy(y.begin(), y.end(), psi_out);
    cd E = 0;
    std::vector<cd> md;
    for (int t = 0; t < nterms; ++t) {
        const vqc_term_t& h = terms[t];
        double e2[2];
        orc_expec_dense(n, y.data(), nullptr, h.nq, h.pos, hpool + h.mat, e2);
        E += cd(e2[0], e2[1]);
        const int K = 1 << h.nq;
        md.assign(K * K, 0);
        dagger(hpool + h.mat, K, md.data());
        for (int i = 0; i < K * K; ++i) md[i] += hpool[h.mat + i];
        orc_accum_dense(n, y.data(), delta.data(), h.nq, h.pos, md.data());
    }
    energy2[0] = E.real();
    energy2[1] = E.imag();
    sv_back_propagate(n, nops, ops, pool, nparams, y.data(), delta.data(), grad);
    return 0;
}

// tr(x * y) with both stored as vec (column-major N x N), no conjugation: the
// reference's dot(storage(x)', storage(yc)) (circuitdiff.jl:36-40).
static cd trace_prod(int n, const cd* x, const cd* y) {
    const uint64_t N = uint64_t(1) << n;
    return parallel_sum(int64_t(N), [&](int64_t ist, int64_t ifn) {
        cd s = 0;
        for (int64_t i = ist; i <= ifn; ++i)
            for (uint64_t j = 0; j < N; ++j) s += x[i + j * N] * y[j + i * N];
        return s;
    });
}

static void dm_back_propagate(int n, int nops, const vqc_op_t* ops, const cd* pool, int nparams, cd* yp, cd* deltap,
                              double* grad);

// Density-matrix version (circuitdiff.jl:36-59, expecdiff.jl:64-68):
//  energy  E = sum_keys tr(M rho)           (expecs_dm.jl:17-27)
//  seed    Delta = z H' applied to the identity, z = 1
//  gate    Delta <- m' Delta m, y <- m' y m,
//          grad = real(tr(y item Delta) + tr(Delta item' y)), item = G'
//  channel Delta <- S' Delta, y <- S^-1 y.
extern "C" int orc_dm_energy_grad(int n, const cd* rho0, int nops, const vqc_op_t* ops, const cd* pool,
                                  int nterms, const vqc_term_t* terms, const cd* hpool, int nparams,
                                  double* energy2, double* grad, cd* rho_out) {
    if (check_ops(n, 1, nops, ops)) return -1;
    const uint64_t N = uint64_t(1) << n;
    const int64_t L = int64_t(N * N);
    std::vector<cd> y(rho0, rho0 + L), delta(L, cd(0)), ident(L, cd(0));
    orc_apply_circuit(n, 1, y.data(), nops, ops, pool);
    if (rho_out) std::copy(y.begin(), y.end(), rho_out);
    for (uint64_t i = 0; i < N; ++i) ident[i + i * N] = 1;
    cd E = 0;
    std::vector<cd> md;
    for (int t = 0; t < nterms; ++t) {
        const vqc_term_t& h = terms[t];
        double e2[2];
        orc_dm_expec(n, y.data(), h.nq, h.pos, hpool + h.mat, e2);
        E += cd(e2[0], e2[1]);
        const int K = 1 << h.nq;
        md.assign(K * K, 0);
        dagger(hpool + h.mat, K, md.data());
        orc_accum_dense(2 * n, ident.data(), delta.data(), h.nq, h.pos, md.data());
    }
    energy2[0] = E.real();
    energy2[1] = E.imag();
    dm_back_propagate(n, nops, ops, pool, nparams, y.data(), delta.data(), grad);
    return 0;
}

// back_propagate for a density matrix (circuitdiff.jl:42-59): per gate
// Delta <- m' Delta m, y <- m' y m, grad = real(tr(y item Delta) + tr(Delta item' y)),
// item = G'; per channel Delta <- S' Delta, y <- S^-1 y. In place, like the
// state-vector version.
static void dm_back_propagate(int n, int nops, const vqc_op_t* ops, const cd* pool, int nparams, cd* yp, cd* deltap,
                              double* grad) {
    const uint64_t N = uint64_t(1) << n;
    const int64_t L = int64_t(N * N);
    std::vector<cd> tmp(L);
    cd* y = yp;
    cd* delta = deltap;
    for (int p = 0; p < nparams; ++p) grad[p] = 0;
    std::vector<cd> ud, gd;
    for (int i = nops - 1; i >= 0; --i) {
        const vqc_op_t& o = ops[i];
        if (o.kind == VQC_OP_CHANNEL) {
            orc_dm_apply_superop(n, delta, o.nq, o.pos, pool + o.mat, 1);
            orc_dm_apply_superop(n, y, o.nq, o.pos, pool + o.aux, 0);
            continue;
        }
        const int K = 1 << o.nq;
        ud.assign(K * K, 0);
        dagger(pool + o.mat, K, ud.data());
        orc_dm_apply_gate(n, delta, o.nq, o.pos, ud.data());
        orc_dm_apply_gate(n, y, o.nq, o.pos, ud.data());
        for (int j = 0; j < o.nparams; ++j) {
            if (o.slot[j] < 0) continue;
            const cd* G = pool + o.gen[j];
            gd.assign(K * K, 0);
            dagger(G, K, gd.data());  // item = G'
            // tr(y * (item on rows of Delta))
            std::copy(delta, delta + L, tmp.begin());
            orc_apply_dense(2 * n, tmp.data(), o.nq, o.pos, gd.data());
            cd a = trace_prod(n, y, tmp.data());
            // tr(Delta * (item' on rows of y)), item' = G
            std::copy(y, y + L, tmp.begin());
            orc_apply_dense(2 * n, tmp.data(), o.nq, o.pos, G);
            cd b = trace_prod(n, delta, tmp.data());
            grad[o.slot[j]] += (a + b).real();
        }
    }
}

// The pullback of y = circuit * x (@adjoint *(circuit, x), circuitdiff.jl:2-8)
// for a cotangent delta_y on y: grad (the circuit's active parameters, forward
// order) and delta_x, the cotangent on x. y_out (may be null) receives y.
extern "C" int orc_circuit_vjp(int n, int is_dm, const cd* x, int nops, const vqc_op_t* ops, const cd* pool, int nparams,
                               const cd* delta_y, double* grad, cd* delta_x, cd* y_out) {
    if (check_ops(n, is_dm, nops, ops)) return -1;
    const int64_t L = int64_t(1) << (is_dm ? 2 * n : n);
    std::vector<cd> y(x, x + L), delta(delta_y, delta_y + L);
    orc_apply_circuit(n, is_dm, y.data(), nops, ops, pool);
    if (y_out) std::copy(y.begin(), y.end(), y_out);
    if (is_dm) dm_back_propagate(n, nops, ops, pool, nparams, y.data(), delta.data(), grad);
    else sv_back_propagate(n, nops, ops, pool, nparams, y.data(), delta.data(), grad);
    if (delta_x) std::copy(delta.begin(), delta.end(), delta_x);
    return 0;
}

// ---------------------------------------------------------------------------
// Measurement and post-selection (src/measure.jl, src/postselect.jl). Serial,
// like the reference's loops.
// ---------------------------------------------------------------------------
// compute_probability_zero (measure.jl:46-58 for a vector, :70-82 for a
// density matrix): sum over the blocks i = 0, stri, 2 stri, ... of the first
// `pos` entries (|v|^2, or the diagonal of rho).
extern "C" double orc_prob_zero(int n, const cd* v, int is_dm, int key0) {
    const uint64_t L = uint64_t(1) << n;
    const uint64_t pos = uint64_t(1) << key0, stri = pos * 2;
    double r = 0.0;
    for (uint64_t i = 0; i < L; i += stri)
        for (uint64_t j = 0; j < pos; ++j) {
            const uint64_t l = i + j;
            r += is_dm ? v[l + l * L].real() : std::norm(v[l]);
        }
    return r;
}

// discrete_sample(l) with rand() = u (auxiliary/sampling.jl:2-18): the first i
// with cumulative(l)[i] <= u < cumulative(l)[i+1], over l normalised by its sum.
static int discrete_sample(const double* l, int L, double u) {
    double s = 0.0;
    for (int i = 0; i < L; ++i) s += l[i];
    double c = 0.0;
    for (int i = 0; i < L; ++i) {
        const double c1 = c + l[i] / s;
        if (u >= c && u < c1) return i;
        c = c1;
    }
    return -1;  // the reference errors here
}

// measure! (measure.jl:4-25): sample the outcome, then apply the 1-qubit
// matrix reshape(kron(e_i, ZERO) or kron(e_i, e_i), 2, 2) / sqrt(p) as a gate
// (U rho U^dagger on a density matrix).
extern "C" int orc_measure(int n, cd* v, int is_dm, int key0, double u, int auto_reset, int* outcome, double* prob) {
    if (key0 < 0 || key0 >= n) return -1;
    const double p0 = orc_prob_zero(n, v, is_dm, key0);
    const double l[2] = {p0, 1.0 - p0};
    const int i = discrete_sample(l, 2, u);
    if (i < 0) return -1;
    const double p = l[i];
    cd e[2] = {0.0, 0.0}, z[2] = {1.0, 0.0};
    e[i] = 1.0;
    const cd* second = auto_reset ? z : e;
    cd k4[4];  // kron(e, second): k4[a*2 + b] = e[a] * second[b]
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) k4[a * 2 + b] = e[a] * second[b] / std::sqrt(p);
    // reshape(k4, (2, 2)) is column-major: m[r, c] = k4[r + 2c]
    cd m[4] = {k4[0], k4[1], k4[2], k4[3]};
    const int pos[1] = {key0};
    if (is_dm) orc_dm_apply_gate(n, v, 1, pos, m);
    else orc_apply_dense(n, v, 1, pos, m);
    *outcome = i;
    *prob = p;
    return 0;
}

// post_select! (postselect.jl:3-8): the UP or DOWN projector as a gate, then
// division by the norm, which is returned.
extern "C" int orc_post_select(int n, cd* v, int key0, int value, double* norm) {
    if (key0 < 0 || key0 >= n || (value != 0 && value != 1)) return -1;
    const cd P[4] = {value == 0 ? 1.0 : 0.0, 0.0, 0.0, value == 1 ? 1.0 : 0.0};
    const int pos[1] = {key0};
    orc_apply_dense(n, v, 1, pos, P);
    const uint64_t L = uint64_t(1) << n;
    double s = 0.0;
    for (uint64_t b = 0; b < L; ++b) s += std::norm(v[b]);
    s = std::sqrt(s);
    for (uint64_t b = 0; b < L; ++b) v[b] /= s;
    *norm = s;
    return 0;
}

// _apply_throw_impl (postselect.jl:47-53) and the collapse of measure
// (measure.jl:28-39): swap qubit 1 with the measured qubit, take row `value`
// of reshape(v, 2, L/2), swap back. out gets 2^(n-1) unnormalised entries.
extern "C" int orc_select(int n, cd* v, int key0, int value, cd* out) {
    if (key0 < 0 || key0 >= n || n < 2 || (value != 0 && value != 1)) return -1;
    const cd S[16] = {1, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 1};  // SWAP, column-major
    const int sw[2] = {0, key0};
    if (key0 != 0) orc_apply_dense(n, v, 2, sw, S);
    const uint64_t half = uint64_t(1) << (n - 1);
    for (uint64_t j = 0; j < half; ++j) out[j] = v[2 * j + value];
    if (key0 != 0) orc_apply_dense(n, v, 2, sw, S);
    return 0;
}
