// vqc_expec.hip -- Hamiltonian sweeps over a state in HBM: k-local
// accumulate (H*psi), k-local expectations <bra|M|ket>, Pauli-string
// expectations and H*psi, and tr(M rho) on density matrices.
//
// Reductions are deterministic: each block writes one partial per term, and a
// second kernel sums the partials in a fixed order (the reference sums
// per-thread partials through atomics, parallel_for.jl:28-61, whose order
// varies run to run).
#include "vqc_internal.h"

template <int K>
struct EPos {
    int p[K];
};

constexpr int kEB = 256;        // threads per block
constexpr int kMaxParts = 1024; // partial sums per term

static unsigned parts_grid(uint64_t work) {
    uint64_t b = (work + kEB - 1) / kEB;
    return unsigned(std::max<uint64_t>(1, std::min<uint64_t>(b, kMaxParts)));
}

// ---------------------------------------------------------------------------
// out += M_pos in  (accumulate kernels of src/hamiltonian/apply_qterms/*.jl)
// ---------------------------------------------------------------------------
template <int K>
__global__ void __launch_bounds__(kEB) k_accum(const double2* __restrict__ in, double2* __restrict__ out, uint64_t ngroups,
                                               EPos<K> gp, const double2* __restrict__ Ug) {
    constexpr int D = 1 << K;
    __shared__ double2 Us[D * D];
    for (int i = threadIdx.x; i < D * D; i += kEB) Us[i] = Ug[i];
    __syncthreads();
    for (uint64_t g = uint64_t(xcd_block()) * kEB + threadIdx.x; g < ngroups; g += uint64_t(gridDim.x) * kEB) {
        const uint64_t base = insert_zeros<K>(g, gp.p);
        double2 a[D];
#pragma unroll
        for (int j = 0; j < D; ++j) a[j] = in[base + sub_offset<K>(j, gp.p)];
#pragma unroll
        for (int r = 0; r < D; ++r) {
            const uint64_t o = base + sub_offset<K>(r, gp.p);
            double2 s = out[o];
#pragma unroll
            for (int c = 0; c < D; ++c) s = cfma(Us[r + c * D], a[c], s);
            out[o] = s;
        }
    }
}

// per-block partials of <bra|M_pos|ket>
template <int K, bool BRA>
__global__ void __launch_bounds__(kEB) k_expec(const double2* __restrict__ ket, const double2* __restrict__ bra, uint64_t ngroups,
                                               EPos<K> gp, const double2* __restrict__ Ug, double2* __restrict__ parts) {
    constexpr int D = 1 << K;
    __shared__ double2 Us[D * D];
    __shared__ double red[kEB / 64];
    for (int i = threadIdx.x; i < D * D; i += kEB) Us[i] = Ug[i];
    __syncthreads();
    double2 acc = make_double2(0.0, 0.0);
    for (uint64_t g = uint64_t(xcd_block()) * kEB + threadIdx.x; g < ngroups; g += uint64_t(gridDim.x) * kEB) {
        const uint64_t base = insert_zeros<K>(g, gp.p);
        double2 a[D];
#pragma unroll
        for (int j = 0; j < D; ++j) a[j] = ket[base + sub_offset<K>(j, gp.p)];
#pragma unroll
        for (int r = 0; r < D; ++r) {
            double2 s = make_double2(0.0, 0.0);
#pragma unroll
            for (int c = 0; c < D; ++c) s = cfma(Us[r + c * D], a[c], s);
            const double2 b = BRA ? bra[base + sub_offset<K>(r, gp.p)] : a[r];
            acc = cfmaj(b, s, acc);
        }
    }
    const double sr = block_sum(acc.x, red);
    const double si = block_sum(acc.y, red);
    if (threadIdx.x == 0) parts[blockIdx.x] = make_double2(sr, si);
}

static thread_local DeviceBuffer g_emat;

template <int K>
static void launch_accum(const double2* in, double2* out, int nbits, const int32_t* pos, const double2* Ud, hipStream_t st) {
    EPos<K> gp;
    for (int j = 0; j < K; ++j) gp.p[j] = pos[j];
    const uint64_t ng = uint64_t(1) << (nbits - K);
    const unsigned grid = unsigned(std::min<uint64_t>((ng + kEB - 1) / kEB, 1u << 20));
    hipLaunchKernelGGL((k_accum<K>), dim3(grid), dim3(kEB), 0, st, in, out, ng, gp, Ud);
}

int vqc_accum_sweep(const double2* in, double2* out, int nbits, int k, const int32_t* pos, const double2* Ud, hipStream_t st) {
    switch (k) {
        case 1: launch_accum<1>(in, out, nbits, pos, Ud, st); break;
        case 2: launch_accum<2>(in, out, nbits, pos, Ud, st); break;
        case 3: launch_accum<3>(in, out, nbits, pos, Ud, st); break;
        case 4: launch_accum<4>(in, out, nbits, pos, Ud, st); break;
        case 5: launch_accum<5>(in, out, nbits, pos, Ud, st); break;
        default: return vqc_fail(VQC_EINVAL, "accumulate on %d qubits", k);
    }
    VQC_HIP(hipGetLastError());
    return VQC_OK;
}

template <int K>
static void launch_expec(const double2* ket, const double2* bra, int nbits, const int32_t* pos, const double2* Ud,
                         double2* parts, unsigned grid, hipStream_t st) {
    EPos<K> gp;
    for (int j = 0; j < K; ++j) gp.p[j] = pos[j];
    const uint64_t ng = uint64_t(1) << (nbits - K);
    if (bra)
        hipLaunchKernelGGL((k_expec<K, true>), dim3(grid), dim3(kEB), 0, st, ket, bra, ng, gp, Ud, parts);
    else
        hipLaunchKernelGGL((k_expec<K, false>), dim3(grid), dim3(kEB), 0, st, ket, ket, ng, gp, Ud, parts);
}

// Enqueue <bra|M|ket> partials for one term into parts[0..grid) and return grid.
int vqc_expec_sweep(const double2* ket, const double2* bra, int nbits, int k, const int32_t* pos, const double2* Ud,
                    double2* parts, unsigned* grid_out, hipStream_t st) {
    const unsigned grid = parts_grid(uint64_t(1) << (nbits - k));
    switch (k) {
        case 1: launch_expec<1>(ket, bra, nbits, pos, Ud, parts, grid, st); break;
        case 2: launch_expec<2>(ket, bra, nbits, pos, Ud, parts, grid, st); break;
        case 3: launch_expec<3>(ket, bra, nbits, pos, Ud, parts, grid, st); break;
        case 4: launch_expec<4>(ket, bra, nbits, pos, Ud, parts, grid, st); break;
        case 5: launch_expec<5>(ket, bra, nbits, pos, Ud, parts, grid, st); break;
        default: return vqc_fail(VQC_EINVAL, "expectation on %d qubits", k);
    }
    VQC_HIP(hipGetLastError());
    *grid_out = grid;
    return VQC_OK;
}

extern "C" int vqc_sv_accum_dense(const vqc_state* in, vqc_state* out, int k, const int32_t* pos, const double* M) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_qubits(in, "vqc_sv_accum_dense"));
    VQC_TRY(vqc_check_state(out, "vqc_sv_accum_dense"));
    if (in->nbits != out->nbits) return vqc_fail(VQC_EINVAL, "accumulate between states of different size");
    if (in->d == out->d) return vqc_fail(VQC_EINVAL, "accumulate needs distinct input and output states");
    VQC_TRY(vqc_check_positions(in->nbits, k, pos, 5));
    if (!M) return vqc_fail(VQC_EINVAL, "null matrix");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    const int D = 1 << k;
    VQC_TRY(vqc_upload_buffer(&g_emat.p, &g_emat.cap, M, size_t(D) * D * 16));
    return vqc_accum_sweep(in->d, out->d, in->nbits, k, pos, (const double2*)g_emat.p, c->stream);
}

extern "C" int vqc_sv_expec_dense(const vqc_state* ket, const vqc_state* bra, int k, const int32_t* pos, const double* M,
                                  double* out2) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_qubits(ket, "vqc_sv_expec_dense"));
    if (bra && bra->nbits != ket->nbits) return vqc_fail(VQC_EINVAL, "bra and ket differ in size");
    VQC_TRY(vqc_check_positions(ket->nbits, k, pos, 5));
    if (!M || !out2) return vqc_fail(VQC_EINVAL, "null matrix or output");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    const int D = 1 << k;
    VQC_TRY(vqc_upload_buffer(&g_emat.p, &g_emat.cap, M, size_t(D) * D * 16));
    void* scr;
    VQC_TRY(vqc_scratch(size_t(kMaxParts + 1) * 16, &scr));
    double2* parts = (double2*)scr;
    unsigned grid;
    VQC_TRY(vqc_expec_sweep(ket->d, bra ? bra->d : nullptr, ket->nbits, k, pos, (const double2*)g_emat.p, parts, &grid, c->stream));
    vqc_reduce_partials(parts, 1, grid, parts + kMaxParts, c->stream);
    VQC_HIP(hipGetLastError());
    return vqc_fetch(parts + kMaxParts, 1, out2);
}

// ---------------------------------------------------------------------------
// Pauli strings. P|b> = c i^{nY} (-1)^{popcount(b & z)} |b ^ x>. Terms that
// share an x mask share one sweep. With x != 0 the sweep visits each pair
// (b, b^x) once, b with the top bit of x clear, so every amplitude is read
// exactly once per distinct x mask.
// ---------------------------------------------------------------------------
constexpr int kPauliChunk = 8;  // terms per sweep

struct PauliChunk {
    uint64_t x;
    int hbit;              // highest set bit of x (-1 when x == 0)
    int nt;
    uint64_t z[kPauliChunk];
    double2 c[kPauliChunk];  // c * i^{nY}
};

__device__ __forceinline__ double sgn(uint64_t b, uint64_t z) { return (__popcll(b & z) & 1) ? -1.0 : 1.0; }

// parts layout: [term slot t][block]; slot = chunk_first + t
template <bool BRA>
__global__ void __launch_bounds__(kEB) k_pauli_expec(const double2* __restrict__ ket, const double2* __restrict__ bra, int nbits,
                                                     PauliChunk pc, double2* __restrict__ parts, int nparts, int first) {
    __shared__ double red[kEB / 64];
    double2 acc[kPauliChunk];
#pragma unroll
    for (int t = 0; t < kPauliChunk; ++t) acc[t] = make_double2(0.0, 0.0);
    const uint64_t L = uint64_t(1) << nbits;
    if (pc.x == 0) {
        for (uint64_t b = uint64_t(xcd_block()) * kEB + threadIdx.x; b < L; b += uint64_t(gridDim.x) * kEB) {
            const double2 k = ket[b];
            const double2 q = BRA ? bra[b] : k;
            double2 w = make_double2(0.0, 0.0);
            w = cfmaj(q, k, w);
#pragma unroll
            for (int t = 0; t < kPauliChunk; ++t)
                if (t < pc.nt) {
                    const double s = sgn(b, pc.z[t]);
                    acc[t].x += s * w.x;
                    acc[t].y += s * w.y;
                }
        }
    } else {
        const uint64_t half = L >> 1;
        for (uint64_t g = uint64_t(xcd_block()) * kEB + threadIdx.x; g < half; g += uint64_t(gridDim.x) * kEB) {
            const uint64_t lo = g & ((uint64_t(1) << pc.hbit) - 1);
            const uint64_t b = ((g - lo) << 1) | lo;
            const uint64_t a = b ^ pc.x;
            const double2 kb = ket[b], ka = ket[a];
            const double2 qb = BRA ? bra[b] : kb;
            const double2 qa = BRA ? bra[a] : ka;
            double2 w1 = make_double2(0.0, 0.0), w2 = make_double2(0.0, 0.0);
            w1 = cfmaj(qa, kb, w1);  // conj(bra_a) ket_b, weight s(b)
            w2 = cfmaj(qb, ka, w2);  // conj(bra_b) ket_a, weight s(a)
#pragma unroll
            for (int t = 0; t < kPauliChunk; ++t)
                if (t < pc.nt) {
                    const double s1 = sgn(b, pc.z[t]), s2 = sgn(a, pc.z[t]);
                    acc[t].x += s1 * w1.x + s2 * w2.x;
                    acc[t].y += s1 * w1.y + s2 * w2.y;
                }
        }
    }
#pragma unroll
    for (int t = 0; t < kPauliChunk; ++t) {
        if (t >= pc.nt) break;
        const double sr = block_sum(acc[t].x, red);
        const double si = block_sum(acc[t].y, red);
        if (threadIdx.x == 0) parts[size_t(first + t) * nparts + blockIdx.x] = cmul(pc.c[t], make_double2(sr, si));
    }
}

// out (=|+=) sum_t c_t P_t in for the terms of one chunk (pair sweep when x != 0)
__global__ void __launch_bounds__(kEB) k_pauli_apply(const double2* __restrict__ in, double2* __restrict__ out, int nbits,
                                                     PauliChunk pc, int accumulate) {
    const uint64_t L = uint64_t(1) << nbits;
    if (pc.x == 0) {
        for (uint64_t b = uint64_t(xcd_block()) * kEB + threadIdx.x; b < L; b += uint64_t(gridDim.x) * kEB) {
            double2 f = make_double2(0.0, 0.0);
            for (int t = 0; t < pc.nt; ++t) {
                const double s = sgn(b, pc.z[t]);
                f.x += s * pc.c[t].x;
                f.y += s * pc.c[t].y;
            }
            double2 r = cmul(f, in[b]);
            if (accumulate) r = cadd(r, out[b]);
            out[b] = r;
        }
        return;
    }
    const uint64_t half = L >> 1;
    for (uint64_t g = uint64_t(xcd_block()) * kEB + threadIdx.x; g < half; g += uint64_t(gridDim.x) * kEB) {
        const uint64_t lo = g & ((uint64_t(1) << pc.hbit) - 1);
        const uint64_t b = ((g - lo) << 1) | lo;
        const uint64_t a = b ^ pc.x;
        double2 fb = make_double2(0.0, 0.0), fa = make_double2(0.0, 0.0);
        for (int t = 0; t < pc.nt; ++t) {
            const double sb = sgn(b, pc.z[t]), sa = sgn(a, pc.z[t]);
            fb.x += sb * pc.c[t].x;
            fb.y += sb * pc.c[t].y;
            fa.x += sa * pc.c[t].x;
            fa.y += sa * pc.c[t].y;
        }
        // (P in)_a = f(b) in_b ; (P in)_b = f(a) in_a
        double2 ra = cmul(fb, in[b]);
        double2 rb = cmul(fa, in[a]);
        if (accumulate) {
            ra = cadd(ra, out[a]);
            rb = cadd(rb, out[b]);
        }
        out[a] = ra;
        out[b] = rb;
    }
}

// Group terms by x mask into chunks of at most kPauliChunk terms; returns the
// chunks and, for each term, its slot (chunk order).
static void pauli_chunks(int nterms, const vqc_pauli_t* terms, std::vector<PauliChunk>& chunks, std::vector<int>& slot_of) {
    std::vector<int> order(nterms);
    for (int i = 0; i < nterms; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return terms[a].x < terms[b].x; });
    slot_of.assign(nterms, 0);
    int slot = 0;
    for (int i = 0; i < nterms;) {
        PauliChunk pc;
        pc.x = terms[order[i]].x;
        pc.hbit = pc.x ? 63 - __builtin_clzll(pc.x) : -1;
        pc.nt = 0;
        while (i < nterms && terms[order[i]].x == pc.x && pc.nt < kPauliChunk) {
            const vqc_pauli_t& t = terms[order[i]];
            const int ny = __builtin_popcountll(t.x & t.z) & 3;
            double re = t.re, im = t.im;
            for (int r = 0; r < ny; ++r) {  // multiply by i
                double tmp = re;
                re = -im;
                im = tmp;
            }
            pc.z[pc.nt] = t.z;
            pc.c[pc.nt] = make_double2(re, im);
            slot_of[order[i]] = slot++;
            ++pc.nt;
            ++i;
        }
        for (int t = pc.nt; t < kPauliChunk; ++t) {
            pc.z[t] = 0;
            pc.c[t] = make_double2(0, 0);
        }
        chunks.push_back(pc);
    }
}

static int check_paulis(int nbits, int nterms, const vqc_pauli_t* terms) {
    if (nterms < 0 || (nterms && !terms)) return vqc_fail(VQC_EINVAL, "bad Pauli term array");
    const uint64_t valid = nbits >= 64 ? ~uint64_t(0) : ((uint64_t(1) << nbits) - 1);
    for (int i = 0; i < nterms; ++i)
        if ((terms[i].x | terms[i].z) & ~valid) return vqc_fail(VQC_EINVAL, "Pauli term %d acts outside the %d qubits", i, nbits);
    return VQC_OK;
}

extern "C" int vqc_sv_expec_pauli(const vqc_state* ket, const vqc_state* bra, int nterms, const vqc_pauli_t* terms,
                                  double* out, double* sum2) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_qubits(ket, "vqc_sv_expec_pauli"));
    if (bra && bra->nbits != ket->nbits) return vqc_fail(VQC_EINVAL, "bra and ket differ in size");
    VQC_TRY(check_paulis(ket->nbits, nterms, terms));
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    std::vector<PauliChunk> chunks;
    std::vector<int> slot_of;
    pauli_chunks(nterms, terms, chunks, slot_of);
    const unsigned nparts = parts_grid(ket->len() >> 1);
    void* scr;
    VQC_TRY(vqc_scratch(size_t(nterms) * (nparts + 1) * 16 + 16, &scr));
    double2* parts = (double2*)scr;
    double2* res = parts + size_t(nterms) * nparts;
    int first = 0;
    for (const PauliChunk& pc : chunks) {
        if (bra)
            hipLaunchKernelGGL(k_pauli_expec<true>, dim3(nparts), dim3(kEB), 0, c->stream, ket->d, bra->d, ket->nbits, pc, parts,
                               int(nparts), first);
        else
            hipLaunchKernelGGL(k_pauli_expec<false>, dim3(nparts), dim3(kEB), 0, c->stream, ket->d, ket->d, ket->nbits, pc, parts,
                               int(nparts), first);
        first += pc.nt;
    }
    if (nterms) vqc_reduce_partials(parts, nterms, nparts, res, c->stream);
    VQC_HIP(hipGetLastError());
    std::vector<double> r(2 * size_t(std::max(nterms, 1)));
    if (nterms) VQC_TRY(vqc_fetch(res, nterms, r.data()));
    double sr = 0, si = 0;
    for (int i = 0; i < nterms; ++i) {
        const int s = slot_of[i];
        if (out) {
            out[2 * i] = r[2 * s];
            out[2 * i + 1] = r[2 * s + 1];
        }
    }
    for (int i = 0; i < nterms; ++i) {  // sum in the caller's term order
        sr += r[2 * slot_of[i]];
        si += r[2 * slot_of[i] + 1];
    }
    if (sum2) {
        sum2[0] = sr;
        sum2[1] = si;
    }
    return VQC_OK;
}

// ---------------------------------------------------------------------------
// vqc_sv_expec_pauli_noalloc: the same sweeps and reductions with the partial
// sums in a file-scope device array instead of the library's growable scratch,
// and the results copied straight to the caller's arrays (no pinned staging),
// so the call allocates nothing.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocDoubles2 = size_t(1) << 18;  // 4 MiB of partials and results per device
static __device__ __align__(256) double2 scratch[kNoAllocDoubles2];
constexpr int kMaxDevices = 64;
static double2* g_noalloc_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up once per
// device, with that device current; the array's host-side name is not it).
static int noalloc_scratch(double2** out) {
    int dev = 0;
    VQC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return vqc_fail(VQC_EUNSUPPORTED, "device %d beyond the no-allocation table", dev);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        VQC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_noalloc_ptr[dev] = (double2*)p;
    }
    *out = g_noalloc_ptr[dev];
    return VQC_OK;
}

constexpr int kNoAllocMaxTerms = 256;

// pauli_chunks without heap use: a stable insertion sort of the terms by x mask
// into caller arrays of kNoAllocMaxTerms entries
static int pauli_chunks_fixed(int nterms, const vqc_pauli_t* terms, PauliChunk* chunks, int* order, int* slot_of) {
    for (int i = 0; i < nterms; ++i) {
        int j = i;
        while (j > 0 && terms[order[j - 1]].x > terms[i].x) {
            order[j] = order[j - 1];
            --j;
        }
        order[j] = i;
    }
    int nch = 0, slot = 0;
    for (int i = 0; i < nterms;) {
        PauliChunk& pc = chunks[nch++];
        pc.x = terms[order[i]].x;
        pc.hbit = pc.x ? 63 - __builtin_clzll(pc.x) : -1;
        pc.nt = 0;
        while (i < nterms && terms[order[i]].x == pc.x && pc.nt < kPauliChunk) {
            const vqc_pauli_t& t = terms[order[i]];
            const int ny = __builtin_popcountll(t.x & t.z) & 3;
            double re = t.re, im = t.im;
            for (int r = 0; r < ny; ++r) {  // multiply by i
                const double tmp = re;
                re = -im;
                im = tmp;
            }
            pc.z[pc.nt] = t.z;
            pc.c[pc.nt] = make_double2(re, im);
            slot_of[order[i]] = slot++;
            ++pc.nt;
            ++i;
        }
        for (int t = pc.nt; t < kPauliChunk; ++t) {
            pc.z[t] = 0;
            pc.c[t] = make_double2(0, 0);
        }
    }
    return nch;
}

// Host-side work arrays of the call, used under the library lock (no heap use per call).
static PauliChunk g_na_chunks[kNoAllocMaxTerms];
static int g_na_order[kNoAllocMaxTerms], g_na_slot[kNoAllocMaxTerms];
static double g_na_res[2 * kNoAllocMaxTerms];

extern "C" int vqc_sv_expec_pauli_noalloc(const vqc_state* ket, const vqc_state* bra, int nterms, const vqc_pauli_t* terms,
                                          double* out, double* sum2) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_qubits(ket, "vqc_sv_expec_pauli_noalloc"));
    if (bra && bra->nbits != ket->nbits) return vqc_fail(VQC_EINVAL, "bra and ket differ in size");
    VQC_TRY(check_paulis(ket->nbits, nterms, terms));
    if (nterms > kNoAllocMaxTerms)
        return vqc_fail(VQC_EINVAL, "vqc_sv_expec_pauli_noalloc: %d terms, the maximum is %d", nterms, kNoAllocMaxTerms);
    const unsigned nparts = parts_grid(ket->len() >> 1);
    const size_t need = size_t(nterms) * (nparts + 1);
    if (need > kNoAllocDoubles2)  // cannot happen for kNoAllocMaxTerms * 1025 <= 2^18; kept as the stated limit
        return vqc_fail(VQC_EINVAL, "vqc_sv_expec_pauli_noalloc: %zu partial sums, the maximum is %zu", need, kNoAllocDoubles2);
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    double2* parts;
    VQC_TRY(noalloc_scratch(&parts));
    const int nch = pauli_chunks_fixed(nterms, terms, g_na_chunks, g_na_order, g_na_slot);
    double2* res = parts + size_t(nterms) * nparts;
    int first = 0;
    for (int k = 0; k < nch; ++k) {
        const PauliChunk& pc = g_na_chunks[k];
        if (bra)
            hipLaunchKernelGGL(k_pauli_expec<true>, dim3(nparts), dim3(kEB), 0, c->stream, ket->d, bra->d, ket->nbits, pc, parts,
                               int(nparts), first);
        else
            hipLaunchKernelGGL(k_pauli_expec<false>, dim3(nparts), dim3(kEB), 0, c->stream, ket->d, ket->d, ket->nbits, pc, parts,
                               int(nparts), first);
        first += pc.nt;
    }
    if (nterms) {
        vqc_reduce_partials(parts, nterms, nparts, res, c->stream);
        VQC_HIP(hipGetLastError());
        VQC_HIP(hipMemcpyAsync(g_na_res, res, size_t(nterms) * 16, hipMemcpyDeviceToHost, c->stream));
        VQC_HIP(hipStreamSynchronize(c->stream));
    }
    double sr = 0, si = 0;
    for (int i = 0; i < nterms; ++i) {  // per term in the caller's order, summed in that order (as vqc_sv_expec_pauli)
        const int sl = g_na_slot[i];
        if (out) {
            out[2 * i] = g_na_res[2 * sl];
            out[2 * i + 1] = g_na_res[2 * sl + 1];
        }
        sr += g_na_res[2 * sl];
        si += g_na_res[2 * sl + 1];
    }
    if (sum2) {
        sum2[0] = sr;
        sum2[1] = si;
    }
    return VQC_OK;
}

__global__ void __launch_bounds__(256) k_zero(double2* v, uint64_t len) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < len; i += uint64_t(gridDim.x) * blockDim.x)
        v[i] = make_double2(0.0, 0.0);
}

void vqc_zero(double2* v, uint64_t len, hipStream_t st) {
    hipLaunchKernelGGL(k_zero, dim3(unsigned(std::min<uint64_t>((len + 255) / 256, 65536))), dim3(256), 0, st, v, len);
}

extern "C" int vqc_sv_apply_pauli_sum(const vqc_state* in, vqc_state* out, int nterms, const vqc_pauli_t* terms) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_qubits(in, "vqc_sv_apply_pauli_sum"));
    VQC_TRY(vqc_check_state(out, "vqc_sv_apply_pauli_sum"));
    if (in->nbits != out->nbits) return vqc_fail(VQC_EINVAL, "states differ in size");
    if (in->d == out->d) return vqc_fail(VQC_EINVAL, "H*psi needs distinct input and output states");
    VQC_TRY(check_paulis(in->nbits, nterms, terms));
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    std::vector<PauliChunk> chunks;
    std::vector<int> slot_of;
    pauli_chunks(nterms, terms, chunks, slot_of);
    if (chunks.empty()) {
        vqc_zero(out->d, out->len(), c->stream);
        VQC_HIP(hipGetLastError());
        return VQC_OK;
    }
    const unsigned grid = unsigned(std::min<uint64_t>((in->len() / 2 + kEB - 1) / kEB, 1u << 16));
    int accumulate = 0;
    for (const PauliChunk& pc : chunks) {
        hipLaunchKernelGGL(k_pauli_apply, dim3(grid), dim3(kEB), 0, c->stream, in->d, out->d, in->nbits, pc, accumulate);
        accumulate = 1;
    }
    VQC_HIP(hipGetLastError());
    return VQC_OK;
}

// out = sum_keys M_key in ; energy = <in|out>  ((m::QubitsOperator)(v) and
// expectation(m, v), apply_qterms.jl:13-18 and expecs.jl:22-28)
static thread_local DeviceBuffer g_hpool;
void vqc_release_expec_thread() {
    g_emat.release();
    g_hpool.release();
}

extern "C" int vqc_sv_apply_terms(const vqc_state* in, vqc_state* out, int nterms, const vqc_term_t* terms, const double* pool,
                                  double* energy2) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_qubits(in, "vqc_sv_apply_terms"));
    VQC_TRY(vqc_check_state(out, "vqc_sv_apply_terms"));
    if (in->nbits != out->nbits || in->d == out->d) return vqc_fail(VQC_EINVAL, "need distinct states of equal size");
    if (nterms < 0 || (nterms && (!terms || !pool))) return vqc_fail(VQC_EINVAL, "bad term array");
    int64_t plen = 0;
    for (int t = 0; t < nterms; ++t) {
        VQC_TRY(vqc_check_positions(in->nbits, terms[t].nq, terms[t].pos, 5));
        if (terms[t].mat < 0) return vqc_fail(VQC_EINVAL, "negative matrix offset");
        plen = std::max<int64_t>(plen, terms[t].mat + (int64_t(1) << (2 * terms[t].nq)));
    }
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    VQC_TRY(vqc_upload_buffer(&g_hpool.p, &g_hpool.cap, pool, size_t(plen) * 16));
    vqc_zero(out->d, out->len(), c->stream);
    for (int t = 0; t < nterms; ++t)
        VQC_TRY(vqc_accum_sweep(in->d, out->d, in->nbits, terms[t].nq, terms[t].pos, (const double2*)g_hpool.p + terms[t].mat, c->stream));
    if (energy2) return vqc_state_dot(in, out, energy2);
    return VQC_OK;
}

// ---------------------------------------------------------------------------
// tr(M rho) on the 2^k x 2^k diagonal blocks (expec_dm_serial.jl:23-129):
// sum over row groups b of sum_{a,c} M[a,c] rho[b+off_c, b+off_a]
// ---------------------------------------------------------------------------
template <int K>
__global__ void __launch_bounds__(kEB) k_dm_expec(const double2* __restrict__ rho, int n, uint64_t ngroups, EPos<K> gp,
                                                  const double2* __restrict__ Ug, double2* __restrict__ parts) {
    constexpr int D = 1 << K;
    __shared__ double2 Us[D * D];
    __shared__ double red[kEB / 64];
    for (int i = threadIdx.x; i < D * D; i += kEB) Us[i] = Ug[i];
    __syncthreads();
    double2 acc = make_double2(0.0, 0.0);
    for (uint64_t g = uint64_t(xcd_block()) * kEB + threadIdx.x; g < ngroups; g += uint64_t(gridDim.x) * kEB) {
        const uint64_t base = insert_zeros<K>(g, gp.p);
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const uint64_t col = (base + sub_offset<K>(a, gp.p)) << n;
#pragma unroll
            for (int cc = 0; cc < D; ++cc) acc = cfma(Us[a + cc * D], rho[col + base + sub_offset<K>(cc, gp.p)], acc);
        }
    }
    const double sr = block_sum(acc.x, red);
    const double si = block_sum(acc.y, red);
    if (threadIdx.x == 0) parts[blockIdx.x] = make_double2(sr, si);
}

int vqc_dm_expec_sweep(const double2* rho, int n, int k, const int32_t* pos, const double2* Ud, double2* parts, unsigned* grid_out,
                       hipStream_t st) {
    const uint64_t ng = uint64_t(1) << (n - k);
    const unsigned grid = parts_grid(ng);
#define DMX(KK)                                                                                                              \
    {                                                                                                                        \
        EPos<KK> gp;                                                                                                         \
        for (int j = 0; j < KK; ++j) gp.p[j] = pos[j];                                                                       \
        hipLaunchKernelGGL((k_dm_expec<KK>), dim3(grid), dim3(kEB), 0, st, rho, n, ng, gp, Ud, parts);                       \
    }
    switch (k) {
        case 1: DMX(1); break;
        case 2: DMX(2); break;
        case 3: DMX(3); break;
        case 4: DMX(4); break;
        case 5: DMX(5); break;
        default: return vqc_fail(VQC_EINVAL, "DM expectation on %d qubits", k);
    }
#undef DMX
    VQC_HIP(hipGetLastError());
    *grid_out = grid;
    return VQC_OK;
}

extern "C" int vqc_dm_expec_dense(const vqc_state* rho, int k, const int32_t* pos, const double* M, double* out2) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_state(rho, "vqc_dm_expec_dense"));
    if (!rho->is_dm) return vqc_fail(VQC_EINVAL, "vqc_dm_expec_dense: state is not a density matrix");
    VQC_TRY(vqc_check_positions(rho->nqubits, k, pos, 5));
    if (!M || !out2) return vqc_fail(VQC_EINVAL, "null matrix or output");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    const int D = 1 << k;
    VQC_TRY(vqc_upload_buffer(&g_emat.p, &g_emat.cap, M, size_t(D) * D * 16));
    void* scr;
    VQC_TRY(vqc_scratch(size_t(kMaxParts + 1) * 16, &scr));
    double2* parts = (double2*)scr;
    unsigned grid;
    VQC_TRY(vqc_dm_expec_sweep(rho->d, rho->nqubits, k, pos, (const double2*)g_emat.p, parts, &grid, c->stream));
    vqc_reduce_partials(parts, 1, grid, parts + kMaxParts, c->stream);
    VQC_HIP(hipGetLastError());
    return vqc_fetch(parts + kMaxParts, 1, out2);
}
