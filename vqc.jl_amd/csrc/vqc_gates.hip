// vqc_gates.hip -- one-pass gate sweeps over a state in HBM.
//
// Dense k-qubit gates (k = 1..5) replace _apply_gate_threaded2!(key, U, v)
// (threaded_short_range.jl:100,364,753; threaded_long_range.jl:51,131). The
// reference tiles each target class differently for CPU SIMD; here one kernel
// covers every class: thread t owns group g = t (the amplitudes whose target
// bits run over all 2^k values), so lanes of a wavefront read consecutive base
// indices and every one of the 2^k loads is a contiguous 1 KiB wavefront
// access whenever the lowest target is >= 6, and stays cache-line complete
// below that. Each amplitude is read and written exactly once (32 B per
// amplitude of HBM traffic), which is the roofline of an unfused gate.
#include <cmath>
#include <cstdlib>

#include "vqc_internal.h"

template <int K>
struct GPos {
    int p[K];
};
template <int K>
struct GMat {
    double2 m[(1 << K) * (1 << K)];
};

constexpr int kBlock = 256;

typedef double d2v __attribute__((ext_vector_type(2)));

template <bool NT>
__device__ __forceinline__ double2 ld(const double2* p) {
    if constexpr (NT) {
        const d2v v = __builtin_nontemporal_load(reinterpret_cast<const d2v*>(p));
        return make_double2(v.x, v.y);
    } else {
        return *p;
    }
}
template <bool NT>
__device__ __forceinline__ void st(double2* p, double2 x) {
    if constexpr (NT) {
        d2v v;
        v.x = x.x;
        v.y = x.y;
        __builtin_nontemporal_store(v, reinterpret_cast<d2v*>(p));
    } else {
        *p = x;
    }
}

// K <= 3: the matrix travels in the kernel arguments (<= 1 KiB, read through
// the scalar cache). Each thread applies the gate to R groups, issuing all
// R * 2^K loads before the first multiply; NT marks the streaming accesses
// non-temporal (each amplitude is touched once per sweep). A barrier separates
// the workgroup's loads from its stores, so that HBM sees each workgroup's reads
// together and then its writes: one-qubit sweeps of a 28-qubit state went from
// 5.62-6.21 to 5.85-6.45 TB/s over targets 3 to 27 (a standalone probe).
template <int K, int R, bool NT>
__global__ void __launch_bounds__(kBlock) k_dense_small(double2* __restrict__ v, uint64_t ngroups, GPos<K> gp, GMat<K> U) {
    constexpr int D = 1 << K;
    uint64_t off[D];
#pragma unroll
    for (int j = 0; j < D; ++j) off[j] = sub_offset<K>(j, gp.p);
    const uint64_t g0 = uint64_t(xcd_block()) * (R * kBlock) + threadIdx.x;
    double2 a[R][D];
    uint64_t base[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint64_t g = g0 + uint64_t(r) * kBlock;
        base[r] = insert_zeros<K>(g, gp.p);
        if (g < ngroups) {
#pragma unroll
            for (int j = 0; j < D; ++j) a[r][j] = ld<NT>(v + base[r] + off[j]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (g0 + uint64_t(r) * kBlock >= ngroups) continue;
#pragma unroll
        for (int row = 0; row < D; ++row) {
            double2 s = make_double2(0.0, 0.0);
#pragma unroll
            for (int c = 0; c < D; ++c) s = cfma(U.m[row + c * D], a[r][c], s);
            st<NT>(v + base[r] + off[row], s);
        }
    }
}

// K = 1 with one amplitude per thread: a workgroup of BS threads holds BS / 2 whole
// pairs (the BS contiguous amplitudes around a target below log2(BS), else BS / 2
// consecutive amplitudes from each side of the pair). Each thread loads its own
// amplitude, takes its partner's from LDS and stores its own row: one load and one
// store per thread over contiguous addresses, the access shape of a plain in-place
// stream, whatever the target. Standalone probe on a 28-qubit state: 6.65-6.86
// TB/s for targets 0-12 with 256 threads, against 6.35-6.72 for the pair-per-thread
// form (4.2-5.1 for targets 0-2, which had to stay temporal there); for targets
// 13-27, 512 threads (4 KiB from each side) average 6373 GB/s against 6273 for the
// pair form and 6289 with 256 threads.
constexpr int kDense1xWideFrom = 13;  // targets from this bit on take 512 threads
constexpr bool kDenseKx = true;  // K = 2, 3 in the one-amplitude form below
template <int BS>
__global__ void __launch_bounds__(BS) k_dense_1x(double2* __restrict__ v, int q, GMat<1> U) {
    __shared__ double2 xs[BS];
    constexpr int LB = BS == 512 ? 9 : 8;  // log2(BS)
    const uint32_t t = threadIdx.x;
    const uint64_t b = xcd_block();
    uint64_t a;
    uint32_t pt;  // the partner's thread
    if (q < LB) {
        a = b * BS + t;
        pt = t ^ (1u << q);
    } else {
        const uint64_t g = b * (BS / 2) + (t & (BS / 2 - 1));
        const uint64_t lo = g & ((uint64_t(1) << q) - 1);
        a = (((g - lo) << 1) | lo) + (uint64_t(t / (BS / 2)) << q);
        pt = t ^ (BS / 2);
    }
    const double2 x = ld<true>(v + a);
    xs[t] = x;
    __syncthreads();
    const double2 y = xs[pt];
    const bool hi = (a >> q) & 1;  // row of this amplitude: x0 = low, x1 = high of the pair
    const double2 x0 = hi ? y : x, x1 = hi ? x : y;
    const double2 r = hi ? cfma(U.m[3], x1, cmul(U.m[1], x0)) : cfma(U.m[2], x1, cmul(U.m[0], x0));
    st<true>(v + a, r);
}

// K = 2 and 3 with one amplitude per thread. A workgroup holds a tile of 2^LB
// amplitudes (LB = 8 for K = 2, 9 for K = 3): the m lowest state bits in place and
// the targets at or above m on the next local bits (m = LB - that count), so a wave
// still reads contiguous amplitudes. Each thread loads one amplitude into LDS. Wave w
// then computes row w of the gate for the 64 groups of the tile, lane l for group l:
// the row is wave-uniform, so its 2^K entries come from the kernel arguments through
// scalar loads, and only the group's members are read from LDS. When the targets sit
// on local bits 6 and up (LP_TOP), the output of (group l, row w) is lane l's own
// amplitude and is stored at once; otherwise it goes back through LDS to the thread
// that loaded that position. The sum runs over the same terms in the same order as
// the group-per-thread form, so the results are bit-identical to it. Standalone probe
// on a 28-qubit state, 16 random target sets: 6146 -> 6405 GB/s for K = 2 and
// 5920 -> 6387 for K = 3 against reading the row from LDS as well (the copy-only
// access pattern of the same tiles: 6398 and 6436).
template <int K, bool LP_TOP>
__global__ void __launch_bounds__(64 << K) k_dense_kx(double2* __restrict__ v, GPos<K + 6> tile, GPos<K> lp, GMat<K> U) {
    constexpr int LB = K + 6;
    constexpr uint32_t BS = 1u << LB;
    __shared__ double2 xs[BS];
    __shared__ double2 ys[LP_TOP ? 1 : BS];
    const uint32_t t = threadIdx.x;
    uint64_t a = insert_zeros<LB>(xcd_block(), tile.p);
#pragma unroll
    for (int k = 0; k < LB; ++k) a |= uint64_t((t >> k) & 1) << tile.p[k];
    xs[t] = ld<true>(v + a);
    __syncthreads();
    const uint32_t w = __builtin_amdgcn_readfirstlane(t >> 6);  // the row: uniform per wave
    uint32_t o[K];
#pragma unroll
    for (int j = 0; j < K; ++j) o[j] = 1u << lp.p[j];
    const uint32_t g = LP_TOP ? (t & 63u) : uint32_t(insert_zeros<K>(t & 63u, lp.p));
    double2 s = make_double2(0.0, 0.0);
#pragma unroll
    for (int c = 0; c < (1 << K); ++c) {
        uint32_t e = g;
#pragma unroll
        for (int j = 0; j < K; ++j) e |= ((c >> j) & 1) ? o[j] : 0u;
        s = cfma(U.m[w + (1u << K) * c], xs[e], s);
    }
    if constexpr (LP_TOP) {
        st<true>(v + a, s);
    } else {
        uint32_t e = g;
#pragma unroll
        for (int j = 0; j < K; ++j) e |= ((w >> j) & 1) ? o[j] : 0u;
        ys[e] = s;
        __syncthreads();
        st<true>(v + a, ys[t]);
    }
}

// K = 4, 5: the matrix (4 or 16 KiB) is staged once per block in LDS and read
// as a broadcast; one group per thread.
template <int K>
__global__ void __launch_bounds__(kBlock) k_dense_big(double2* __restrict__ v, uint64_t ngroups, GPos<K> gp,
                                                       const double2* __restrict__ Ug) {
    constexpr int D = 1 << K;
    __shared__ double2 Us[D * D];
    for (int i = threadIdx.x; i < D * D; i += kBlock) Us[i] = Ug[i];
    __syncthreads();
    const uint64_t g = uint64_t(xcd_block()) * kBlock + threadIdx.x;
    if (g >= ngroups) return;
    const uint64_t base = insert_zeros<K>(g, gp.p);
    double2 a[D];
#pragma unroll
    for (int j = 0; j < D; ++j) a[j] = v[base + sub_offset<K>(j, gp.p)];
#pragma unroll
    for (int row = 0; row < D; ++row) {
        double2 s = make_double2(0.0, 0.0);
#pragma unroll
        for (int c = 0; c < D; ++c) s = cfma(Us[row + c * D], a[c], s);
        v[base + sub_offset<K>(row, gp.p)] = s;
    }
}

// Device copy of a host matrix for the K >= 4 kernels (stream ordered, so a
// later call may reuse the buffer).
static thread_local DeviceBuffer g_mat;  // this thread's matrix upload buffer
void vqc_release_gates_thread() { g_mat.release(); }

template <int K>
static int launch_dense(double2* v, int nbits, const int32_t* pos, const double* Uh, hipStream_t st) {
    constexpr int D = 1 << K;
    GPos<K> gp;
    for (int j = 0; j < K; ++j) gp.p[j] = pos[j];
    const uint64_t ngroups = uint64_t(1) << (nbits - K);
    if constexpr (K <= 3) {
        GMat<K> U;
        for (int i = 0; i < D * D; ++i) U.m[i] = make_double2(Uh[2 * i], Uh[2 * i + 1]);
        // Groups per thread and non-temporal hints (vqc_set_tuning). Measured on
        // MI355X at 28 qubits: one group per thread is best for every K.
        // Non-temporal accesses pay when every target is at bit 3 or above (a
        // 3-qubit gate on bits 20-22: 5.1 -> 5.9 TB/s) and cost when one is
        // below (bits 0, 13, 27: 5.3 -> 3.5 TB/s), where each lane's accesses
        // fall in a few 128-byte lines; those sweeps stay temporal.
        const vqc_tuning_t tu = vqc_tuning_now();
        if constexpr (K == 1) {
            if (nbits >= 9 && tu.dense_groups[0] == 1 && tu.dense_nontemporal[0] != 0) {
                if (gp.p[0] < kDense1xWideFrom)
                    hipLaunchKernelGGL((k_dense_1x<256>), dim3(unsigned((uint64_t(1) << nbits) / 256)), dim3(256), 0, st, v, gp.p[0], U);
                else
                    hipLaunchKernelGGL((k_dense_1x<512>), dim3(unsigned((uint64_t(1) << nbits) / 512)), dim3(512), 0, st, v, gp.p[0], U);
                if (vqc_trace_on()) vqc_trace("k_dense_1x nbits %d target %d", nbits, gp.p[0]);
                VQC_HIP(hipGetLastError());
                return VQC_OK;
            }
        }
        if constexpr (K == 2 || K == 3) {
            constexpr int LB = K + 6;
            if (nbits >= LB && kDenseKx && tu.dense_groups[K - 1] == 1 && tu.dense_nontemporal[K - 1] != 0) {
                // the m lowest bits plus the targets at or above m: LB bits in all
                int m = LB;
                for (;;) {
                    int above = 0;
                    for (int j = 0; j < K; ++j) above += gp.p[j] >= m;
                    if (above + m <= LB) break;
                    --m;
                }
                GPos<LB> tile;
                GPos<K> lp;
                int k = 0;
                for (int b = 0; b < m; ++b) tile.p[k++] = b;
                for (int j = 0; j < K; ++j) {
                    if (gp.p[j] < m) lp.p[j] = gp.p[j];
                    else {
                        lp.p[j] = k;
                        tile.p[k++] = gp.p[j];
                    }
                }
                if (k != LB) return vqc_fail(VQC_EINTERNAL, "dense sweep tile of %d bits instead of %d", k, LB);
                bool top = true;
                for (int j = 0; j < K; ++j) top = top && lp.p[j] == 6 + j;
                const dim3 grid(unsigned((uint64_t(1) << nbits) >> LB)), blk(1u << LB);
                if (top) hipLaunchKernelGGL((k_dense_kx<K, true>), grid, blk, 0, st, v, tile, lp, U);
                else hipLaunchKernelGGL((k_dense_kx<K, false>), grid, blk, 0, st, v, tile, lp, U);
                if (vqc_trace_on())
                    vqc_trace("k_dense_kx<%d,%d> nbits %d targets %d %d %d", K, int(top), nbits, gp.p[0], gp.p[1],
                              K > 2 ? gp.p[K - 1] : -1);
                VQC_HIP(hipGetLastError());
                return VQC_OK;
            }
        }
        const int R = tu.dense_groups[K - 1];
        const bool NT = tu.dense_nontemporal[K - 1] != 0 && gp.p[0] >= 3;
        const dim3 blk(kBlock);
#define VQC_DENSE_LAUNCH(RR)                                                                                             \
    if (R == RR) {                                                                                                      \
        if (NT) hipLaunchKernelGGL((k_dense_small<K, RR, true>), dim3(grid_for(ngroups, RR * kBlock)), blk, 0, st, v,     \
                                   ngroups, gp, U);                                                                       \
        else hipLaunchKernelGGL((k_dense_small<K, RR, false>), dim3(grid_for(ngroups, RR * kBlock)), blk, 0, st, v,      \
                                ngroups, gp, U);                                                                          \
    }
        VQC_DENSE_LAUNCH(1) else VQC_DENSE_LAUNCH(2) else VQC_DENSE_LAUNCH(4) else VQC_DENSE_LAUNCH(8) else return vqc_fail(
            VQC_EINVAL, "dense groups per thread must be 1, 2, 4 or 8");
#undef VQC_DENSE_LAUNCH
        if (vqc_trace_on())
            vqc_trace("k_dense_small<%d,%d,%d> nbits %d lowest target %d grid %u", K, R, int(NT), nbits, gp.p[0],
                      grid_for(ngroups, R * kBlock));
    } else {
        VQC_TRY(vqc_upload_buffer(&g_mat.p, &g_mat.cap, Uh, size_t(D) * D * 16));
        hipLaunchKernelGGL((k_dense_big<K>), dim3(grid_for(ngroups, kBlock)), dim3(kBlock), 0, st, v, ngroups, gp,
                           (const double2*)g_mat.p);
        if (vqc_trace_on()) vqc_trace("k_dense_big<%d> nbits %d grid %u", K, nbits, grid_for(ngroups, kBlock));
    }
    VQC_HIP(hipGetLastError());
    return VQC_OK;
}

int vqc_dense_sweep(double2* v, int nbits, int k, const int32_t* pos, const double* U, hipStream_t st) {
    switch (k) {
        case 1: return launch_dense<1>(v, nbits, pos, U, st);
        case 2: return launch_dense<2>(v, nbits, pos, U, st);
        case 3: return launch_dense<3>(v, nbits, pos, U, st);
        case 4: return launch_dense<4>(v, nbits, pos, U, st);
        case 5: return launch_dense<5>(v, nbits, pos, U, st);
    }
    return vqc_fail(VQC_EINVAL, "dense gate on %d qubits", k);
}

extern "C" int vqc_sv_apply_dense(vqc_state* s, int k, const int32_t* pos, const double* U) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_qubits(s, "vqc_sv_apply_dense"));
    VQC_TRY(vqc_check_positions(s->nbits, k, pos, 5));
    if (!U) return vqc_fail(VQC_EINVAL, "null gate matrix");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    return vqc_dense_sweep(s->d, s->nbits, k, pos, U, c->stream);
}

// ---------------------------------------------------------------------------
// Named gates (src/applygates/specific/*.jl): permutations and phases that
// touch only the amplitudes they change.
// ---------------------------------------------------------------------------
template <int KIND, int K>
__global__ void __launch_bounds__(kBlock) k_named(double2* __restrict__ v, uint64_t ngroups, GPos<K> sorted,
                                                  GPos<K> given, double2 ph) {
    const uint64_t g = uint64_t(xcd_block()) * kBlock + threadIdx.x;
    if (g >= ngroups) return;
    const uint64_t l = insert_zeros<K>(g, sorted.p);
    const uint64_t p0 = uint64_t(1) << given.p[0];
    const uint64_t p1 = K > 1 ? uint64_t(1) << given.p[K > 1 ? 1 : 0] : 0;
    const uint64_t p2 = K > 2 ? uint64_t(1) << given.p[K > 2 ? 2 : 0] : 0;
    if constexpr (KIND == VQC_GATE_X) {
        double2 a = v[l], b = v[l + p0];
        v[l] = b;
        v[l + p0] = a;
    } else if constexpr (KIND == VQC_GATE_PHASE) {
        v[l + p0] = cmul(ph, v[l + p0]);
    } else if constexpr (KIND == VQC_GATE_CZ) {
        double2 a = v[l + p0 + p1];
        v[l + p0 + p1] = make_double2(-a.x, -a.y);
    } else if constexpr (KIND == VQC_GATE_CNOT) {
        double2 a = v[l + p0], b = v[l + p0 + p1];
        v[l + p0] = b;
        v[l + p0 + p1] = a;
    } else if constexpr (KIND == VQC_GATE_CPHASE) {
        v[l + p0 + p1] = cmul(ph, v[l + p0 + p1]);
    } else if constexpr (KIND == VQC_GATE_SWAP) {
        double2 a = v[l + p0], b = v[l + p1];
        v[l + p0] = b;
        v[l + p1] = a;
    } else if constexpr (KIND == VQC_GATE_ISWAP) {  // ph = +i or -i
        double2 a = v[l + p0], b = v[l + p1];
        v[l + p0] = cmul(ph, b);
        v[l + p1] = cmul(ph, a);
    } else if constexpr (KIND == VQC_GATE_FREDKIN) {
        double2 a = v[l + p0 + p1], b = v[l + p0 + p2];
        v[l + p0 + p1] = b;
        v[l + p0 + p2] = a;
    } else if constexpr (KIND == VQC_GATE_CCPHASE) {
        v[l + p0 + p1 + p2] = cmul(ph, v[l + p0 + p1 + p2]);
    }
}

template <int KIND, int K>
static int launch_named(double2* v, int nbits, const int32_t* pos, double2 ph, hipStream_t st) {
    GPos<K> sorted, given;
    for (int j = 0; j < K; ++j) sorted.p[j] = given.p[j] = pos[j];
    std::sort(sorted.p, sorted.p + K);
    for (int j = 0; j < K; ++j) {
        if (sorted.p[j] < 0 || sorted.p[j] >= nbits) return vqc_fail(VQC_EINVAL, "qubit position %d out of range", sorted.p[j]);
        if (j && sorted.p[j] == sorted.p[j - 1]) return vqc_fail(VQC_EINVAL, "repeated qubit position %d", sorted.p[j]);
    }
    const uint64_t ngroups = uint64_t(1) << (nbits - K);
    hipLaunchKernelGGL((k_named<KIND, K>), dim3(grid_for(ngroups, kBlock)), dim3(kBlock), 0, st, v, ngroups, sorted, given, ph);
    VQC_HIP(hipGetLastError());
    return VQC_OK;
}

extern "C" int vqc_sv_apply_named(vqc_state* s, int kind, const int32_t* pos, double phi) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_qubits(s, "vqc_sv_apply_named"));
    if (!pos) return vqc_fail(VQC_EINVAL, "null position array");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    const double2 e = make_double2(cos(phi), sin(phi));
    hipStream_t st = c->stream;
    switch (kind) {
        case VQC_GATE_X: return launch_named<VQC_GATE_X, 1>(s->d, s->nbits, pos, e, st);
        case VQC_GATE_PHASE: return launch_named<VQC_GATE_PHASE, 1>(s->d, s->nbits, pos, e, st);
        case VQC_GATE_CZ: return launch_named<VQC_GATE_CZ, 2>(s->d, s->nbits, pos, e, st);
        case VQC_GATE_CNOT: return launch_named<VQC_GATE_CNOT, 2>(s->d, s->nbits, pos, e, st);
        case VQC_GATE_CPHASE: return launch_named<VQC_GATE_CPHASE, 2>(s->d, s->nbits, pos, e, st);
        case VQC_GATE_SWAP: return launch_named<VQC_GATE_SWAP, 2>(s->d, s->nbits, pos, e, st);
        case VQC_GATE_ISWAP: return launch_named<VQC_GATE_ISWAP, 2>(s->d, s->nbits, pos, make_double2(0, 1), st);
        case VQC_GATE_ISWAPDAG: return launch_named<VQC_GATE_ISWAP, 2>(s->d, s->nbits, pos, make_double2(0, -1), st);
        case VQC_GATE_FREDKIN: return launch_named<VQC_GATE_FREDKIN, 3>(s->d, s->nbits, pos, e, st);
        case VQC_GATE_CCPHASE: return launch_named<VQC_GATE_CCPHASE, 3>(s->d, s->nbits, pos, e, st);
    }
    return vqc_fail(VQC_EINVAL, "unknown named gate kind %d", kind);
}

// ---------------------------------------------------------------------------
// Density matrices. rho <- U rho U^dagger in ONE sweep: the reference runs two
// k-qubit sweeps, U on pos then conj(U) on pos+n (generic.jl:54-64); both act
// on the same 4^k block, so each block is loaded once, U applied to its row
// index and conj(U) to its column index, and stored once (32 B per entry).
// ---------------------------------------------------------------------------
template <int K>
__global__ void __launch_bounds__(kBlock) k_dm_gate(double2* __restrict__ v, uint64_t ngroups, GPos<2 * K> gp, GMat<K> U) {
    constexpr int D = 1 << K;
    const uint64_t g = uint64_t(xcd_block()) * kBlock + threadIdx.x;
    if (g >= ngroups) return;
    const uint64_t base = insert_zeros<2 * K>(g, gp.p);
    double2 x[D][D];  // [row][col]
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) x[r][c] = v[base + sub_offset<2 * K>(r | (c << K), gp.p)];
    double2 t[D][D];
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) {
            double2 s = make_double2(0.0, 0.0);
#pragma unroll
            for (int q = 0; q < D; ++q) s = cfma(U.m[r + q * D], x[q][c], s);
            t[r][c] = s;
        }
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) {
            double2 s = make_double2(0.0, 0.0);
#pragma unroll
            for (int q = 0; q < D; ++q) s = cfma(cconj(U.m[c + q * D]), t[r][q], s);
            v[base + sub_offset<2 * K>(r | (c << K), gp.p)] = s;
        }
}

template <int K>
static int launch_dm_gate(vqc_state* s, const int32_t* pos, const double* Uh, hipStream_t st) {
    constexpr int D = 1 << K;
    GPos<2 * K> gp;
    for (int j = 0; j < K; ++j) {
        gp.p[j] = pos[j];
        gp.p[j + K] = pos[j] + s->nqubits;
    }
    GMat<K> U;
    for (int i = 0; i < D * D; ++i) U.m[i] = make_double2(Uh[2 * i], Uh[2 * i + 1]);
    const uint64_t ngroups = uint64_t(1) << (s->nbits - 2 * K);
    hipLaunchKernelGGL((k_dm_gate<K>), dim3(grid_for(ngroups, kBlock)), dim3(kBlock), 0, st, s->d, ngroups, gp, U);
    VQC_HIP(hipGetLastError());
    return VQC_OK;
}

// rho <- U rho U^dagger on a raw vec(rho) of 2n index bits.
int vqc_dm_gate_sweep(double2* v, int n, int k, const int32_t* pos, const double* U, hipStream_t st) {
    vqc_state tmp;
    tmp.nbits = 2 * n;
    tmp.nqubits = n;
    tmp.is_dm = 1;
    tmp.owns = 0;
    tmp.d = v;
    switch (k) {
        case 1: return launch_dm_gate<1>(&tmp, pos, U, st);
        case 2: return launch_dm_gate<2>(&tmp, pos, U, st);
        case 3: return launch_dm_gate<3>(&tmp, pos, U, st);
    }
    const int D = 1 << k;
    std::vector<double> Uc(size_t(2) * D * D);
    for (int i = 0; i < D * D; ++i) {
        Uc[2 * i] = U[2 * i];
        Uc[2 * i + 1] = -U[2 * i + 1];
    }
    int32_t pos2[5];
    for (int j = 0; j < k; ++j) pos2[j] = pos[j] + n;
    VQC_TRY(vqc_dense_sweep(v, 2 * n, k, pos, U, st));
    return vqc_dense_sweep(v, 2 * n, k, pos2, Uc.data(), st);
}

extern "C" int vqc_dm_apply_gate(vqc_state* rho, int k, const int32_t* pos, const double* U) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_state(rho, "vqc_dm_apply_gate"));
    if (!rho->is_dm) return vqc_fail(VQC_EINVAL, "vqc_dm_apply_gate: state is not a density matrix");
    VQC_TRY(vqc_check_positions(rho->nqubits, k, pos, 5));
    if (!U) return vqc_fail(VQC_EINVAL, "null gate matrix");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    return vqc_dm_gate_sweep(rho->d, rho->nqubits, k, pos, U, c->stream);
}

extern "C" int vqc_dm_apply_superop(vqc_state* rho, int N, const int32_t* pos, const double* S, int mode) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_state(rho, "vqc_dm_apply_superop"));
    if (!rho->is_dm) return vqc_fail(VQC_EINVAL, "vqc_dm_apply_superop: state is not a density matrix");
    VQC_TRY(vqc_check_positions(rho->nqubits, N, pos, 2));
    if (!S) return vqc_fail(VQC_EINVAL, "null superoperator");
    if (mode < 0 || mode > 2) return vqc_fail(VQC_EINVAL, "superoperator mode must be 0 (S), 1 (S^dagger) or 2 (S^-1)");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    const int D = 1 << (2 * N);
    int32_t all[4];
    for (int j = 0; j < N; ++j) {
        all[j] = pos[j];
        all[j + N] = pos[j] + rho->nqubits;
    }
    if (mode == 0) return vqc_dense_sweep(rho->d, rho->nbits, 2 * N, all, S, c->stream);
    if (mode == 2) {
        // apply_inverse! (apply_qmaps.jl:40-48): inv(S) by Gauss-Jordan elimination
        // with partial pivoting on the host (S is at most 16 x 16)
        std::vector<double> a(S, S + size_t(2) * D * D), inv(size_t(2) * D * D, 0.0);
        auto A = [&](int r, int cc) -> double* { return &a[2 * size_t(r + cc * D)]; };
        auto I = [&](int r, int cc) -> double* { return &inv[2 * size_t(r + cc * D)]; };
        for (int i = 0; i < D; ++i) I(i, i)[0] = 1.0;
        double scale = 0;
        for (int i = 0; i < D * D; ++i) scale = std::max(scale, std::hypot(a[2 * i], a[2 * i + 1]));
        for (int col = 0; col < D; ++col) {
            int piv = col;
            for (int r = col + 1; r < D; ++r)
                if (std::hypot(A(r, col)[0], A(r, col)[1]) > std::hypot(A(piv, col)[0], A(piv, col)[1])) piv = r;
            const double pr = A(piv, col)[0], pi = A(piv, col)[1];
            const double m2 = pr * pr + pi * pi;
            if (!(std::sqrt(m2) > 1e-13 * std::max(scale, 1e-300)))
                return vqc_fail(VQC_EINVAL, "vqc_dm_apply_superop: the superoperator is singular; it has no inverse");
            for (int cc = 0; cc < D; ++cc) {
                std::swap(A(piv, cc)[0], A(col, cc)[0]);
                std::swap(A(piv, cc)[1], A(col, cc)[1]);
                std::swap(I(piv, cc)[0], I(col, cc)[0]);
                std::swap(I(piv, cc)[1], I(col, cc)[1]);
            }
            const double ir = pr / m2, ii = -pi / m2;  // 1 / pivot
            for (int cc = 0; cc < D; ++cc) {
                double* x = A(col, cc);
                double* y = I(col, cc);
                const double xr = x[0] * ir - x[1] * ii, xi = x[0] * ii + x[1] * ir;
                const double yr = y[0] * ir - y[1] * ii, yi = y[0] * ii + y[1] * ir;
                x[0] = xr, x[1] = xi, y[0] = yr, y[1] = yi;
            }
            for (int r = 0; r < D; ++r) {
                if (r == col) continue;
                const double fr = A(r, col)[0], fi = A(r, col)[1];
                if (fr == 0.0 && fi == 0.0) continue;
                for (int cc = 0; cc < D; ++cc) {
                    const double* x = A(col, cc);
                    const double* y = I(col, cc);
                    A(r, cc)[0] -= fr * x[0] - fi * x[1];
                    A(r, cc)[1] -= fr * x[1] + fi * x[0];
                    I(r, cc)[0] -= fr * y[0] - fi * y[1];
                    I(r, cc)[1] -= fr * y[1] + fi * y[0];
                }
            }
        }
        return vqc_dense_sweep(rho->d, rho->nbits, 2 * N, all, inv.data(), c->stream);
    }
    std::vector<double> Sd(size_t(2) * D * D);
    for (int r = 0; r < D; ++r)
        for (int cc = 0; cc < D; ++cc) {
            Sd[2 * (r + cc * D)] = S[2 * (cc + r * D)];
            Sd[2 * (r + cc * D) + 1] = -S[2 * (cc + r * D) + 1];
        }
    return vqc_dense_sweep(rho->d, rho->nbits, 2 * N, all, Sd.data(), c->stream);
}
