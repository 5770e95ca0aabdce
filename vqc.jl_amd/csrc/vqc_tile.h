// vqc_tile.h -- the LDS tile engine shared by single-circuit and batched
// evaluation.
//
// A state of `nbits` index bits is cut into tiles: a tile fixes every index
// bit outside a chosen set S of t "tile bits" and holds the 2^t amplitudes
// that vary over S. S always contains the contiguous low bits 0..m-1 (so a
// wavefront's 64 loads cover consecutive addresses) plus the targets of a run
// of consecutive circuit operations; every operation whose targets lie in S
// can then be applied to the tile in LDS, so a whole run of gates costs one
// HBM round trip instead of one per gate. A "phase" is one such run with its
// tile bits; a phase is executed by streaming all 2^(nbits-t) tiles through LDS.
//
// Micro-operations (MicroOp) are expressed in local tile coordinates:
//   MK_APPLY   A <- M A                       (forward gate / channel sweep)
//   MK_APPLY2  A <- M A, B <- M B, then for each generator G_q:
//              acc[q] += Re sum conj(B) (G_q A)   (adjoint reverse of a gate)
//   MK_APPLY2X A <- M A, B <- M2 B                (channel reverse: S^-1, S^dagger)
//   MK_GRAD    acc += Re sum conj(B) (G A)        (density-matrix gradient terms)
//   MK_HAM     B += (M + M^dagger) A, acc += sum conj(A) (M A)   (adjoint seed, energy)
//   MK_HAMS    the MK_HAM of nchain terms of at most 2 bits at once, one amplitude
//              per thread; the terms are the nchain MK_HAM records that follow it,
//              and acc receives the sum of their energies
//   MK_CHAIN   up to 4 one-bit MK_APPLY2 steps on the same bit, done in registers
//              (one LDS round trip for e.g. the Rz, Ry, Rx of one qubit)
//   MK_DENSE   A <- M A for a dense 2^K x 2^K matrix on K = 3..5 bits, the product
//              of a run of gates fused by the planner, as FP64 MFMA products
//              (v_mfma_f64_16x16x4f64) over the tile's groups
//   MK_BLOCK   up to 4 steps on one pair of bits, done in registers: each step a
//              4x4 on both bits or a 2x2 on either bit (or no matrix), on A alone
//              (forward phases) or on A and B with an optional generator per step
//              (reverse phases), e.g. a CNOT and the rotations of one of its qubits,
//              or the row and column uncompute of a density-matrix gate with its
//              two gradient terms
// Matrices are read from a device pool in which the planner has already stored
// every variant a micro-op needs (U, U^dagger, conj(U), U^T, M + M^dagger), so
// the inner loops are plain complex matrix-vector products.
#pragma once
#include <unordered_map>

#include "vqc_internal.h"

enum { MK_APPLY = 0, MK_APPLY2 = 1, MK_APPLY2X = 2, MK_GRAD = 3, MK_HAM = 4, MK_CHAIN = 5, MK_BLOCK = 6, MK_DENSE = 7, MK_HAMS = 8 };
enum { BS_BOTH = 0, BS_LO = 1, BS_HI = 2 };  // MK_BLOCK step targets
// matrix structure; codes 2 and 3 name CNOT patterns for 4x4 matrices and
// real / "X-like" (real diagonal, imaginary off-diagonal: Rx and its generator)
// forms for 2x2 ones
enum { MS_DENSE = 0, MS_DIAG = 1, MS_CX_LO = 2, MS_CX_HI = 3, MS_REAL = 2, MS_XLIKE = 3 };
enum { MF_CONJ = 1, MF_TRANS = 2 };
enum { PH_FWD = 0, PH_REV = 1, PH_HAM = 2 };
// kind sets a kernel is built for (bit k = micro-op kind k)
enum {
    KS_FWD = (1 << MK_APPLY) | (1 << MK_BLOCK),
    KS_HAM = (1 << MK_HAM) | (1 << MK_HAMS),
    // batches, which plan without the values and keep a micro-op per term
    KS_HAM1 = 1 << MK_HAM,
    KS_REV = (1 << MK_APPLY2) | (1 << MK_CHAIN) | (1 << MK_BLOCK),  // state-vector reverse
    KS_REVX = (1 << MK_APPLY) | (1 << MK_APPLY2) | (1 << MK_APPLY2X) | (1 << MK_GRAD) | (1 << MK_CHAIN) | (1 << MK_BLOCK),
    // forward, with the straight-line block for the ansatz shape (chosen per
    // phase by the host, so other forward phases keep the plain forward kernel)
    KS_FWDH = KS_FWD | (1 << 9),
    // forward, with fused dense blocks of 3 to 5 bits on the matrix cores
    KS_FWDD = KS_FWD | (1 << MK_DENSE),
};

constexpr int kTileThreads = 512;
constexpr int kTileWaves = kTileThreads / 64;
constexpr int kLoadBatch = 8;      // global loads in flight per thread when a tile is read
constexpr int kMaxTileBits = 13;   // one tile buffer in LDS (forward phases): 128 KiB
constexpr int kMaxTileBits2 = 12;  // two tile buffers (Hamiltonian / reverse phases, whole small evaluations)
constexpr int kMaxAcc = 48;  // accumulator entries per phase
constexpr int kWideBatch = 16;  // batches of at least this many instances take the batch tile shape
// MicroOp::mstr flag of a 2-bit MK_APPLY / MK_APPLY2X whose 4 x 4 matrices only
// couple indices of equal parity (b0 xor b1): entries {00, 11} x {00, 11} and
// {01, 10} x {01, 10}. Depolarizing and amplitude-damping superoperators have this
// form, and so do their inverses and adjoints.
constexpr int kMsParity = 1 << 16;

struct MicroOp {
    int32_t kind;
    int32_t kb;         // local bits touched (1..5)
    int32_t buf;        // MK_APPLY: 0 = A, 1 = B
    int32_t nchain;     // MK_CHAIN / MK_BLOCK: steps (2..4); MK_HAMS: term records that follow
    int32_t lpos[5];    // ascending local bit positions
    int32_t ng;         // partial sums produced (MK_APPLY2: 0..3, MK_GRAD / MK_HAM: 1, MK_CHAIN: 0..4)
    int32_t acc[4];     // accumulator index of each partial
    int32_t smask;      // MK_BLOCK: target of step j in bits 2j..2j+1 (BS_BOTH / BS_LO / BS_HI)
    int32_t mstr;       // structure (MS_*) of step j's matrix in bits 4j..4j+1, of its generator in 4j+2..4j+3
                        // (step 0 = the single matrix of MK_APPLY / MK_APPLY2)
    int64_t mat;
    int64_t mat2;
    int64_t gen[4];     // generators; MK_CHAIN / MK_BLOCK: one per step, -1 when the step has none
    int64_t cmat[4];    // MK_CHAIN / MK_BLOCK: the steps' matrices, in order
};

struct PhaseDesc {
    int32_t kind;        // PH_FWD / PH_REV / PH_HAM
    int32_t nbits;
    int32_t t;           // tile bits
    int32_t m;           // contiguous low bits in the tile
    int32_t nhb;         // tile bits above the low run
    int32_t nob;         // index bits outside the tile
    int32_t hb[kMaxTileBits];
    int32_t ob[40];
    int32_t op_begin;
    int32_t op_end;
    int32_t nacc;        // accumulator entries used by this phase
    int32_t row_begin;   // first partial row of those entries
    int32_t ham_accum;   // PH_HAM: 1 = add onto B already in HBM, 0 = overwrite
    int32_t flags;       // PF_INIT_ZERO: the phase reads |0...0> instead of srcA
    // pool entries [pf_lo, pf_lo + pf_n) cover every matrix the phase's micro-ops
    // read (pf_n = 0: too spread out to be worth reading ahead, see warm_phase)
    int32_t pf_n;
    int64_t pf_lo;
};
enum { PF_INIT_ZERO = 1 };
// largest pool range (complex entries) a phase reads ahead: 64 KiB
constexpr int kPrefetchMax = 4096;

template <int K>
__device__ __forceinline__ uint32_t ins32(uint32_t i, const int32_t* p) {
#pragma unroll
    for (int j = 0; j < K; ++j) {
        uint32_t lo = i & ((1u << p[j]) - 1);
        i = ((i - lo) << 1) | lo;
    }
    return i;
}
template <int K>
__device__ __forceinline__ uint32_t off32(int j, const int32_t* p) {
    uint32_t o = 0;
#pragma unroll
    for (int b = 0; b < K; ++b)
        if (j >> b & 1) o |= 1u << p[b];
    return o;
}

__device__ __forceinline__ uint64_t tile_base(uint64_t tile, const PhaseDesc& p) {
    uint64_t b = 0;
    for (int i = 0; i < p.nob; ++i) b |= ((tile >> i) & 1) << p.ob[i];
    return b;
}
// State-index bits of the tile-local bits [b0, b1) of j: tile-local bit b sits at
// state bit b below p.m and at p.hb[b - p.m] above. The deposit is linear over
// disjoint bits, so the index j = u * kTileThreads + threadIdx.x of a thread's
// u-th amplitude is the OR of a per-thread part (the bits below
// kTileThreadBits, once per tile) and a part uniform over the workgroup (the bits
// of u, scalar work): a few instructions per amplitude instead of a loop over the
// tile's high bits.
constexpr int kTileThreadBits = 9;
static_assert((1 << kTileThreadBits) == kTileThreads, "kTileThreadBits must match kTileThreads");
__device__ __forceinline__ uint64_t tile_deposit(uint32_t j, int b0, int b1, const PhaseDesc& p) {
    const int lo = std::min(b1, p.m);  // bits below p.m stay where they are
    uint64_t r = b0 < lo ? uint64_t(j & ((1u << lo) - 1) & ~((1u << b0) - 1)) : 0;
    for (int b = std::max(b0, p.m); b < b1; ++b)
        if (j >> b & 1) r |= uint64_t(1) << p.hb[b - p.m];
    return r;
}

// Reductions of the gradient / energy terms. A micro-op leaves one partial sum
// per wave in red[buf][q][wave] (a plain store by lane 0 of each wave); at the
// start of the NEXT micro-op, wave 0 adds those partials, in wave order, onto
// gacc[row]. Two red buffers alternate, so no extra barrier is needed and the
// sums are the same on every run.
constexpr int kRedSlots = 4;                                // partials per micro-op
constexpr int kRedDoubles2 = 2 * kRedSlots * kTileWaves;  // double2 entries of red[]

struct AccState {
    int pend_n;          // partials waiting in red[pend_buf] (0..4)
    int pend_buf;
    int e0, e1, e2, e3;  // accumulator rows they belong to
    int cur_buf;         // red buffer the next micro-op writes
};

__device__ __forceinline__ void acc_init(AccState& st) {
    st.pend_n = 0;
    st.pend_buf = 0;
    st.cur_buf = 0;
    st.e0 = st.e1 = st.e2 = st.e3 = 0;
}

// (every lane of the wave must be active: called after the group loops)
__device__ __forceinline__ void wave_partial(double2* red, int buf, int q, double re, double im) {
    re = wave_sum_full(re);
    im = wave_sum_full(im);
    if ((threadIdx.x & 63) == 0) red[(buf * kRedSlots + q) * kTileWaves + (threadIdx.x >> 6)] = make_double2(re, im);
}
// a real partial (gradient terms)
__device__ __forceinline__ void wave_partial_re(double2* red, int buf, int q, double re) {
    re = wave_sum_full(re);
    if ((threadIdx.x & 63) == 0) red[(buf * kRedSlots + q) * kTileWaves + (threadIdx.x >> 6)] = make_double2(re, 0.0);
}

// fold pending partials into gacc (callers order it with barriers). One thread
// folds them in turn: two partials of one micro-op may belong to the same row
// (the row and column terms of a density-matrix gradient). The folding thread is
// lane 0 of the last wave: micro-ops hand out groups from thread 0 up, so on
// small tiles (a 9-bit tile holds 128 groups of a 2-bit micro-op: waves 0 and 1)
// the last wave has no group of its own and the fold runs beside the micro-op
// instead of before it.
__device__ __forceinline__ void acc_fold(AccState& st, const double2* red, double2* gacc) {
    if (st.pend_n && threadIdx.x == kTileThreads - 64) {
        for (int q = 0; q < st.pend_n; ++q) {
            const int e = q == 0 ? st.e0 : (q == 1 ? st.e1 : (q == 2 ? st.e2 : st.e3));
            double2 s = gacc[e];
            for (int w = 0; w < kTileWaves; ++w) s = cadd(s, red[(st.pend_buf * kRedSlots + q) * kTileWaves + w]);
            gacc[e] = s;
        }
    }
    st.pend_n = 0;
}

// after a micro-op that wrote n partials into red[st.cur_buf]
__device__ __forceinline__ void acc_push(AccState& st, int n, const int32_t* e) {
    if (!n) return;
    st.pend_n = n;
    st.pend_buf = st.cur_buf;
    st.e0 = e[0];
    st.e1 = n > 1 ? e[1] : 0;
    st.e2 = n > 2 ? e[2] : 0;
    st.e3 = n > 3 ? e[3] : 0;
    st.cur_buf ^= 1;
}

// Row loops are unrolled for up to 3 local bits; for 4 and 5 bits only the
// inner product is unrolled (it indexes the register-held group).
// Matrix structure of a micro-op's matrices, decided by the planner (from the
// values when it has them, else from the caller's per-op hints) and stored in
// MicroOp::mstr: diagonal 2x2 matrices (Rz, phases and their generators) take
// two products instead of a 2x2 product, and the exact CNOT patterns (control
// on the low or the high bit of the pair) become a register swap.
__device__ __forceinline__ int ms_mat(const MicroOp& op, int j) { return (op.mstr >> (4 * j)) & 3; }
__device__ __forceinline__ int ms_gen(const MicroOp& op, int j) { return (op.mstr >> (4 * j + 2)) & 3; }
__device__ __forceinline__ void dswap(double2& x, double2& y) {
    const double2 t = x;
    x = y;
    y = t;
}

// 2x2 matrices by structure. The four entries are read once per step and the
// structure is dispatched once per step into straight-line code over every pair
// the step updates. A real matrix or an X-like one takes half the multiplies of
// a complex one; entries that the structure says are zero are exactly zero, so
// results equal the dense ones. Pairs are passed and returned by value: every
// amplitude stays in a register whatever branch the target bit takes.
struct Mat2 {
    double2 m00, m10, m01, m11;
};
struct Pair2 {
    double2 x0, x1;
};
__device__ __forceinline__ Mat2 load_mat2(const double2* __restrict__ M) { return Mat2{M[0], M[1], M[2], M[3]}; }
// M (x0, x1)
template <int MS>
__device__ __forceinline__ Pair2 m2(const Mat2& m, double2 x0, double2 x1) {
    if constexpr (MS == MS_DIAG) {
        return Pair2{cmul(m.m00, x0), cmul(m.m11, x1)};
    } else if constexpr (MS == MS_REAL) {
        return Pair2{make_double2(fma(m.m01.x, x1.x, m.m00.x * x0.x), fma(m.m01.x, x1.y, m.m00.x * x0.y)),
                     make_double2(fma(m.m11.x, x1.x, m.m10.x * x0.x), fma(m.m11.x, x1.y, m.m10.x * x0.y))};
    } else if constexpr (MS == MS_XLIKE) {
        // M = [[d0, i u], [i l, d1]] with d0, d1, u, l real
        const double d0 = m.m00.x, l = m.m10.y, u = m.m01.y, d1 = m.m11.x;
        return Pair2{make_double2(fma(-u, x1.y, d0 * x0.x), fma(u, x1.x, d0 * x0.y)),
                     make_double2(fma(-l, x0.y, d1 * x1.x), fma(l, x0.x, d1 * x1.y))};
    } else {
        return Pair2{cfma(m.m01, x1, cmul(m.m00, x0)), cfma(m.m11, x1, cmul(m.m10, x0))};
    }
}
// Re (conj(b0) s0 + conj(b1) s1) with s = M a
template <int MS>
__device__ __forceinline__ double gdot2(const Mat2& m, double2 a0, double2 a1, double2 b0, double2 b1) {
    const Pair2 s = m2<MS>(m, a0, a1);
    return fma(b0.x, s.x0.x, fma(b0.y, s.x0.y, fma(b1.x, s.x1.x, b1.y * s.x1.y)));
}
// in-place update of one pair
template <int MS>
__device__ __forceinline__ void m2_ip(const Mat2& m, double2& x0, double2& x1) {
    const Pair2 r = m2<MS>(m, x0, x1);
    x0 = r.x0;
    x1 = r.x1;
}

// MK_APPLY on one bit with a 2x2 of structure MS
template <int MS>
__device__ __forceinline__ void apply1(double2* __restrict__ s, uint32_t ng, const MicroOp& op, const Mat2 m) {
    const uint32_t o1 = 1u << op.lpos[0];
    for (uint32_t g = threadIdx.x; g < ng; g += kTileThreads) {
        const uint32_t base = ins32<1>(g, op.lpos);
        const Pair2 r = m2<MS>(m, s[base], s[base + o1]);
        s[base] = r.x0;
        s[base + o1] = r.x1;
    }
}

// y = M x for a 4 x 4 M with the parity-block form (kMsParity), x over (lpos[0],
// lpos[1]): two 2 x 2 products, on (x0, x3) and on (x1, x2)
struct Par4 {
    double2 m00, m30, m03, m33, m11, m21, m12, m22;
};
__device__ __forceinline__ Par4 load_par4(const double2* __restrict__ M) {
    return Par4{M[0], M[3], M[12], M[15], M[5], M[6], M[9], M[10]};
}
__device__ __forceinline__ void par4_ip(const Par4& m, double2& x0, double2& x1, double2& x2, double2& x3) {
    const double2 y0 = cfma(m.m03, x3, cmul(m.m00, x0)), y3 = cfma(m.m33, x3, cmul(m.m30, x0));
    const double2 y1 = cfma(m.m12, x2, cmul(m.m11, x1)), y2 = cfma(m.m22, x2, cmul(m.m21, x1));
    x0 = y0;
    x1 = y1;
    x2 = y2;
    x3 = y3;
}
// MK_APPLY (one buffer) or MK_APPLY2X (A <- M A, B <- M2 B) of a parity-block channel
template <bool TWO>
__device__ __forceinline__ void mop_parity(double2* __restrict__ sa, double2* __restrict__ sb, int t, const MicroOp& op,
                                           const double2* __restrict__ pool) {
    const uint32_t ng = 1u << (t - 2);
    const uint32_t o0 = 1u << op.lpos[0], o1 = 1u << op.lpos[1];
    const Par4 m = load_par4(pool + op.mat);
    const Par4 m2 = TWO ? load_par4(pool + op.mat2) : m;
    for (uint32_t g = threadIdx.x; g < ng; g += kTileThreads) {
        const uint32_t base = ins32<2>(g, op.lpos);
        const uint32_t i1 = base + o0, i2 = base + o1, i3 = base + o0 + o1;
        double2 x0 = sa[base], x1 = sa[i1], x2 = sa[i2], x3 = sa[i3];
        par4_ip(m, x0, x1, x2, x3);
        sa[base] = x0;
        sa[i1] = x1;
        sa[i2] = x2;
        sa[i3] = x3;
        if (TWO) {
            double2 y0 = sb[base], y1 = sb[i1], y2 = sb[i2], y3 = sb[i3];
            par4_ip(m2, y0, y1, y2, y3);
            sb[base] = y0;
            sb[i1] = y1;
            sb[i2] = y2;
            sb[i3] = y3;
        }
    }
}

// ROLL keeps the row loop rolled at every width (reverse kernels, where a lone
// MK_APPLY is rare and unrolled 8 x 8 rows beside the reverse micro-ops spill).
template <int KB, bool ROLL = false>
__device__ __forceinline__ void mop_apply(double2* __restrict__ s, int t, const MicroOp& op, const double2* __restrict__ pool) {
    constexpr int D = 1 << KB;
    constexpr int RU = KB <= 3 && !ROLL ? D : 1;
    const double2* __restrict__ M = pool + op.mat;
    const uint32_t ng = 1u << (t - KB);
    if constexpr (KB == 1) {
        // forward one-bit ops are mostly fused products (dense); diagonal ones
        // (lone Rz / phases) take two products per pair
        if (ms_mat(op, 0) == MS_DIAG) {
            apply1<MS_DIAG>(s, ng, op, load_mat2(M));
            return;
        }
    }
    if constexpr (KB == 2) {
        const int ms = ms_mat(op, 0);
        if (ms == MS_CX_LO || ms == MS_CX_HI) {
            for (uint32_t g = threadIdx.x; g < ng; g += kTileThreads) {
                const uint32_t base = ins32<KB>(g, op.lpos);
                const uint32_t o1 = off32<KB>(1, op.lpos), o2 = off32<KB>(2, op.lpos), o3 = off32<KB>(3, op.lpos);
                const uint32_t x = base + (ms == MS_CX_LO ? o1 : o2), y = base + o3;
                const double2 vx = s[x], vy = s[y];
                s[x] = vy;
                s[y] = vx;
            }
            return;
        }
    }
    for (uint32_t g = threadIdx.x; g < ng; g += kTileThreads) {
        const uint32_t base = ins32<KB>(g, op.lpos);
        double2 a[D];
#pragma unroll
        for (int j = 0; j < D; ++j) a[j] = s[base + off32<KB>(j, op.lpos)];
#pragma unroll RU
        for (int r = 0; r < D; ++r) {
            double2 v = make_double2(0.0, 0.0);
#pragma unroll
            for (int c = 0; c < D; ++c) v = cfma(M[r + c * D], a[c], v);
            s[base + off32<KB>(r, op.lpos)] = v;
        }
    }
}

// MK_APPLY2 / MK_APPLY2X (KB <= 4) and MK_GRAD (KB <= 5): A <- M A, B <- M2 B,
// then acc[q] += Re sum conj(B) (G_q A). New values go to LDS as soon as a row
// is done (the group's inputs are in registers and groups are disjoint between
// threads) and are re-read for the generator terms. MK_APPLY2X uses a second
// matrix (mat2) for B (channels), MK_GRAD applies nothing. Row
// loops of 2 and more bits stay rolled: the matrix rows are read through the
// scalar cache one at a time instead of all being hoisted into registers.
// The same for 4 and 5 bits with one buffer's group in registers at a time: A's
// D amplitudes, then B's, then A's new ones for the generator terms, each B entry
// of those read from LDS as its row comes. Holding both groups (2 D complex values,
// 256 VGPRs at 5 bits) spills.
template <int KB>
__device__ __forceinline__ void mop_pair_wide(double2* __restrict__ sa, double2* __restrict__ sb, int t, const MicroOp& op,
                                              const double2* __restrict__ pool, double2* red, int buf) {
    constexpr int D = 1 << KB;
    const bool APPLY = op.kind != MK_GRAD;
    const double2* __restrict__ M = pool + op.mat;
    const double2* __restrict__ M2 = op.kind == MK_APPLY2X ? pool + op.mat2 : M;
    const int ngen = op.ng;
    double gs0 = 0.0, gs1 = 0.0, gs2 = 0.0;
    const uint32_t ng = 1u << (t - KB);
    for (uint32_t g = threadIdx.x; g < ng; g += kTileThreads) {
        const uint32_t base = ins32<KB>(g, op.lpos);
        double2 x[D];
        if (APPLY) {
#pragma unroll
            for (int j = 0; j < D; ++j) x[j] = sa[base + off32<KB>(j, op.lpos)];
#pragma unroll 1
            for (int r = 0; r < D; ++r) {
                double2 s = make_double2(0.0, 0.0);
#pragma unroll
                for (int c = 0; c < D; ++c) s = cfma(M[r + c * D], x[c], s);
                sa[base + off32<KB>(r, op.lpos)] = s;
            }
#pragma unroll
            for (int j = 0; j < D; ++j) x[j] = sb[base + off32<KB>(j, op.lpos)];
#pragma unroll 1
            for (int r = 0; r < D; ++r) {
                double2 u = make_double2(0.0, 0.0);
#pragma unroll
                for (int c = 0; c < D; ++c) u = cfma(M2[r + c * D], x[c], u);
                sb[base + off32<KB>(r, op.lpos)] = u;
            }
        }
        if (ngen) {
#pragma unroll
            for (int j = 0; j < D; ++j) x[j] = sa[base + off32<KB>(j, op.lpos)];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                if (q >= ngen) break;
                const double2* __restrict__ G = pool + op.gen[q];
                double acc = 0.0;
#pragma unroll 1
                for (int r = 0; r < D; ++r) {
                    double2 s = make_double2(0.0, 0.0);
#pragma unroll
                    for (int c = 0; c < D; ++c) s = cfma(G[r + c * D], x[c], s);
                    const double2 br = sb[base + off32<KB>(r, op.lpos)];
                    acc = fma(br.x, s.x, fma(br.y, s.y, acc));  // Re(conj(b_r) s)
                }
                if (q == 0) gs0 += acc;
                else if (q == 1) gs1 += acc;
                else gs2 += acc;
            }
        }
    }
    if (ngen > 0) wave_partial_re(red, buf, 0, gs0);
    if (ngen > 1) wave_partial_re(red, buf, 1, gs1);
    if (ngen > 2) wave_partial_re(red, buf, 2, gs2);
}

template <int KB>
__device__ __forceinline__ void mop_pair(double2* __restrict__ sa, double2* __restrict__ sb, int t, const MicroOp& op,
                                         const double2* __restrict__ pool, double2* red, int buf) {
    if constexpr (KB >= 4) {
        mop_pair_wide<KB>(sa, sb, t, op, pool, red, buf);
        return;
    }
    constexpr int D = 1 << KB;
    constexpr int RU = KB == 1 ? D : 1;  // unrolled rows would hoist every matrix out of the group loop
    const bool APPLY = op.kind != MK_GRAD;
    const bool X = op.kind == MK_APPLY2X;
    const double2* __restrict__ M = pool + op.mat;
    const double2* __restrict__ M2 = X ? pool + op.mat2 : M;
    const int ngen = op.ng;
    double gs0 = 0.0, gs1 = 0.0, gs2 = 0.0;
    const uint32_t ng = 1u << (t - KB);
    for (uint32_t g = threadIdx.x; g < ng; g += kTileThreads) {
        const uint32_t base = ins32<KB>(g, op.lpos);
        double2 a[D], b[D];
#pragma unroll
        for (int j = 0; j < D; ++j) {
            a[j] = sa[base + off32<KB>(j, op.lpos)];
            b[j] = sb[base + off32<KB>(j, op.lpos)];
        }
        if (APPLY) {
#pragma unroll RU
            for (int r = 0; r < D; ++r) {
                double2 s = make_double2(0.0, 0.0), u = make_double2(0.0, 0.0);
#pragma unroll
                for (int c = 0; c < D; ++c) {
                    s = cfma(M[r + c * D], a[c], s);
                    u = cfma(M2[r + c * D], b[c], u);
                }
                sa[base + off32<KB>(r, op.lpos)] = s;
                sb[base + off32<KB>(r, op.lpos)] = u;
            }
            if (ngen) {
#pragma unroll
                for (int j = 0; j < D; ++j) {
                    a[j] = sa[base + off32<KB>(j, op.lpos)];
                    b[j] = sb[base + off32<KB>(j, op.lpos)];
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (q >= ngen) break;
            const double2* __restrict__ G = pool + op.gen[q];
            double acc = 0.0;
#pragma unroll RU
            for (int r = 0; r < D; ++r) {
                double2 s = make_double2(0.0, 0.0);
#pragma unroll
                for (int c = 0; c < D; ++c) s = cfma(G[r + c * D], a[c], s);
                double2 br = b[0];
#pragma unroll
                for (int j = 1; j < D; ++j)
                    if (j == r) br = b[j];
                acc = fma(br.x, s.x, fma(br.y, s.y, acc));  // Re(conj(b_r) s)
            }
            if (q == 0) gs0 += acc;
            else if (q == 1) gs1 += acc;
            else gs2 += acc;
        }
    }
    if (ngen > 0) wave_partial_re(red, buf, 0, gs0);
    if (ngen > 1) wave_partial_re(red, buf, 1, gs1);
    if (ngen > 2) wave_partial_re(red, buf, 2, gs2);
}

// MK_CHAIN: for each step j, (A, B) <- U_j^dagger (A, B) on one bit and, when the
// step has a generator, partial += Re conj(B) (G_j A); the pair stays in
// registers across the steps.
__device__ __forceinline__ void mop_chain(double2* __restrict__ sa, double2* __restrict__ sb, int t, const MicroOp& op,
                                          const double2* __restrict__ pool, double2* red, int buf) {
    const int nch = op.nchain;
    double g0 = 0.0, g1 = 0.0, g2 = 0.0, g3 = 0.0;
    const uint32_t ng = 1u << (t - 1);
    const uint32_t o1 = 1u << op.lpos[0];
    for (uint32_t g = threadIdx.x; g < ng; g += kTileThreads) {
        const uint32_t base = ins32<1>(g, op.lpos);
        double2 a0 = sa[base], a1 = sa[base + o1], b0 = sb[base], b1 = sb[base + o1];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j >= nch) break;
            const Mat2 mj = load_mat2(pool + op.cmat[j]);
            const int ms = ms_mat(op, j);
            if (ms == MS_DIAG) {
                m2_ip<MS_DIAG>(mj, a0, a1);
                m2_ip<MS_DIAG>(mj, b0, b1);
            } else if (ms == MS_REAL) {
                m2_ip<MS_REAL>(mj, a0, a1);
                m2_ip<MS_REAL>(mj, b0, b1);
            } else if (ms == MS_XLIKE) {
                m2_ip<MS_XLIKE>(mj, a0, a1);
                m2_ip<MS_XLIKE>(mj, b0, b1);
            } else {
                m2_ip<MS_DENSE>(mj, a0, a1);
                m2_ip<MS_DENSE>(mj, b0, b1);
            }
            if (op.gen[j] >= 0) {
                const Mat2 gj = load_mat2(pool + op.gen[j]);
                const int gs = ms_gen(op, j);
                const double v = gs == MS_DIAG    ? gdot2<MS_DIAG>(gj, a0, a1, b0, b1)
                                 : gs == MS_REAL  ? gdot2<MS_REAL>(gj, a0, a1, b0, b1)
                                 : gs == MS_XLIKE ? gdot2<MS_XLIKE>(gj, a0, a1, b0, b1)
                                                  : gdot2<MS_DENSE>(gj, a0, a1, b0, b1);
                if (j == 0) g0 += v;
                else if (j == 1) g1 += v;
                else if (j == 2) g2 += v;
                else g3 += v;
            }
        }
        sa[base] = a0;
        sa[base + o1] = a1;
        sb[base] = b0;
        sb[base + o1] = b1;
    }
    int q = 0;
    if (nch > 0 && op.gen[0] >= 0) wave_partial_re(red, buf, q++, g0);
    if (nch > 1 && op.gen[1] >= 0) wave_partial_re(red, buf, q++, g1);
    if (nch > 2 && op.gen[2] >= 0) wave_partial_re(red, buf, q++, g2);
    if (nch > 3 && op.gen[3] >= 0) wave_partial_re(red, buf, q++, g3);
}

// MK_BLOCK: the amplitudes of one pair of bits (lo, hi) in registers, x0..x3
// with index bit_lo + 2 bit_hi; each step updates them in place, and a reverse
// step with a generator adds Re sum conj(b) (G a) over the pairs it touched.
__device__ __forceinline__ void blk_4x4(const double2* __restrict__ M, double2& x0, double2& x1, double2& x2, double2& x3) {
    double2 y[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) y[r] = cfma(M[r + 12], x3, cfma(M[r + 8], x2, cfma(M[r + 4], x1, cmul(M[r], x0))));
    x0 = y[0];
    x1 = y[1];
    x2 = y[2];
    x3 = y[3];
}
__device__ __forceinline__ double blk_grad4(const double2* __restrict__ G, double2 a0, double2 a1, double2 a2, double2 a3,
                                            double2 b0, double2 b1, double2 b2, double2 b3) {
    double2 s[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) s[r] = cfma(G[r + 12], a3, cfma(G[r + 8], a2, cfma(G[r + 4], a1, cmul(G[r], a0))));
    double v = fma(b0.x, s[0].x, b0.y * s[0].y);
    v = fma(b1.x, s[1].x, fma(b1.y, s[1].y, v));
    v = fma(b2.x, s[2].x, fma(b2.y, s[2].y, v));
    return fma(b3.x, s[3].x, fma(b3.y, s[3].y, v));
}
// A 2x2 step of structure MS on the first bit of the register quad: pairs
// (x0, x1) and (x2, x3) of A and, when `two`, of B.
template <int MS>
__device__ __forceinline__ void blk_step2(const Mat2& m, bool two, double2& a0, double2& a1, double2& a2, double2& a3, double2& b0,
                                          double2& b1, double2& b2, double2& b3) {
    m2_ip<MS>(m, a0, a1);
    m2_ip<MS>(m, a2, a3);
    if (two) {
        m2_ip<MS>(m, b0, b1);
        m2_ip<MS>(m, b2, b3);
    }
}
template <int MS>
__device__ __forceinline__ double blk_grad2(const Mat2& m, double2 a0, double2 a1, double2 a2, double2 a3, double2 b0, double2 b1,
                                            double2 b2, double2 b3) {
    return gdot2<MS>(m, a0, a1, b0, b1) + gdot2<MS>(m, a2, a3, b2, b3);
}
// The quad is held as x0..x3 = (00, 01, 10, 11) over (first bit, second bit),
// where the first bit is lo while `sw` is false and hi while it is true: a
// one-bit step on the other bit first exchanges x1 and x2 (a uniform branch
// with fixed registers), so every one-bit step runs the same code on the
// pairs (x0, x1) and (x2, x3). Two-bit steps and the final stores use sw = false.
__device__ __forceinline__ void mop_block(double2* __restrict__ sa, double2* __restrict__ sb, int t, const MicroOp& op,
                                          const double2* __restrict__ pool, double2* red, int buf, bool two) {
    const int nst = op.nchain;
    double g0 = 0.0, g1 = 0.0, g2s = 0.0, g3 = 0.0;
    const uint32_t ng = 1u << (t - 2);
    const uint32_t olo = 1u << op.lpos[0], ohi = 1u << op.lpos[1];
    const double2 zero = make_double2(0.0, 0.0);
    for (uint32_t g = threadIdx.x; g < ng; g += kTileThreads) {
        const uint32_t base = ins32<2>(g, op.lpos);
        double2 a0 = sa[base], a1 = sa[base + olo], a2 = sa[base + ohi], a3 = sa[base + olo + ohi];
        double2 b0 = zero, b1 = zero, b2 = zero, b3 = zero;
        if (two) {
            b0 = sb[base];
            b1 = sb[base + olo];
            b2 = sb[base + ohi];
            b3 = sb[base + olo + ohi];
        }
        bool sw = false;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j >= nst) break;
            const int tgt = (op.smask >> (2 * j)) & 3;
            const int skj = ms_mat(op, j), gkj = ms_gen(op, j);
            if ((tgt == BS_BOTH && sw) || (tgt == BS_LO && sw) || (tgt == BS_HI && !sw)) {
                dswap(a1, a2);
                if (two) dswap(b1, b2);
                sw = !sw;
            }
            if (op.cmat[j] < 0) {
                // a gradient-only step: the generator on the current state
            } else if (tgt == BS_BOTH) {
                if (skj == MS_CX_LO) {
                    dswap(a1, a3);
                    if (two) dswap(b1, b3);
                } else if (skj == MS_CX_HI) {
                    dswap(a2, a3);
                    if (two) dswap(b2, b3);
                } else {
                    const double2* __restrict__ M = pool + op.cmat[j];
                    blk_4x4(M, a0, a1, a2, a3);
                    if (two) blk_4x4(M, b0, b1, b2, b3);
                }
            } else {
                const Mat2 m = load_mat2(pool + op.cmat[j]);
                if (skj == MS_DIAG) blk_step2<MS_DIAG>(m, two, a0, a1, a2, a3, b0, b1, b2, b3);
                else if (skj == MS_REAL) blk_step2<MS_REAL>(m, two, a0, a1, a2, a3, b0, b1, b2, b3);
                else if (skj == MS_XLIKE) blk_step2<MS_XLIKE>(m, two, a0, a1, a2, a3, b0, b1, b2, b3);
                else blk_step2<MS_DENSE>(m, two, a0, a1, a2, a3, b0, b1, b2, b3);
            }
            if (two && op.gen[j] >= 0) {
                const double2* __restrict__ G = pool + op.gen[j];
                double v;
                if (tgt == BS_BOTH) {
                    v = blk_grad4(G, a0, a1, a2, a3, b0, b1, b2, b3);
                } else {
                    const Mat2 gm = load_mat2(G);
                    v = gkj == MS_DIAG    ? blk_grad2<MS_DIAG>(gm, a0, a1, a2, a3, b0, b1, b2, b3)
                        : gkj == MS_REAL  ? blk_grad2<MS_REAL>(gm, a0, a1, a2, a3, b0, b1, b2, b3)
                        : gkj == MS_XLIKE ? blk_grad2<MS_XLIKE>(gm, a0, a1, a2, a3, b0, b1, b2, b3)
                                          : blk_grad2<MS_DENSE>(gm, a0, a1, a2, a3, b0, b1, b2, b3);
                }
                if (j == 0) g0 += v;
                else if (j == 1) g1 += v;
                else if (j == 2) g2s += v;
                else g3 += v;
            }
        }
        if (sw) {
            dswap(a1, a2);
            if (two) dswap(b1, b2);
        }
        sa[base] = a0;
        sa[base + olo] = a1;
        sa[base + ohi] = a2;
        sa[base + olo + ohi] = a3;
        if (two) {
            sb[base] = b0;
            sb[base + olo] = b1;
            sb[base + ohi] = b2;
            sb[base + olo + ohi] = b3;
        }
    }
    if (two) {
        int q = 0;
        if (nst > 0 && op.gen[0] >= 0) wave_partial_re(red, buf, q++, g0);
        if (nst > 1 && op.gen[1] >= 0) wave_partial_re(red, buf, q++, g1);
        if (nst > 2 && op.gen[2] >= 0) wave_partial_re(red, buf, q++, g2s);
        if (nst > 3 && op.gen[3] >= 0) wave_partial_re(red, buf, q++, g3);
    }
}

// The reverse block shape that hardware-efficient ansatz layers (Rx Ry Rz on
// each qubit, then a CNOT ladder) produce almost exclusively (440 of 500 reverse
// micro-ops at 24 qubits), as straight-line code: CNOT^dagger, then Rz^dagger,
// Ry^dagger, Rx^dagger on the CNOT's target, each with its gradient (diagonal,
// real, X-like classes), and its forward counterpart, a CNOT followed by a fused
// 2x2 on its control (KS_FWDH kernels only). CXLO: the control is the pair's
// low bit. Density matrices run the same reverse shape twice per CNOT, on its row
// bits and on its column bits (each side's gradient term rides with that side's
// uncompute).
constexpr int kHeaRevSmask = (BS_HI << 2) | (BS_HI << 4) | (BS_HI << 6);  // BOTH, then the target three times
constexpr int kHeaRevSmaskHi = (BS_LO << 2) | (BS_LO << 4) | (BS_LO << 6);
constexpr int kHeaRevMstr = MS_CX_LO | ((MS_DIAG | MS_DIAG << 2) << 4) | ((MS_REAL | MS_REAL << 2) << 8) |
                            ((MS_XLIKE | MS_XLIKE << 2) << 12);
constexpr int kHeaRevMstrHi = MS_CX_HI | (kHeaRevMstr & ~3);
// forward: CNOT, then one dense 2x2 (a fused rotation product) on its control
constexpr int kHeaFwdSmask = BS_LO << 2, kHeaFwdSmaskHi = BS_HI << 2;
constexpr int kHeaFwdMstr = MS_CX_LO, kHeaFwdMstrHi = MS_CX_HI;

__device__ __forceinline__ int hea_rev_shape(const MicroOp& op) {
    if (op.nchain != 4 || op.cmat[0] < 0 || op.gen[0] >= 0 || op.gen[1] < 0 || op.gen[2] < 0 || op.gen[3] < 0) return 0;
    if (op.smask == kHeaRevSmask && op.mstr == kHeaRevMstr) return 1;
    if (op.smask == kHeaRevSmaskHi && op.mstr == kHeaRevMstrHi) return 2;
    return 0;
}

__device__ __forceinline__ int hea_fwd_shape(const MicroOp& op) {
    if (op.nchain != 2 || op.cmat[0] < 0 || op.cmat[1] < 0) return 0;
    if (op.smask == kHeaFwdSmask && op.mstr == kHeaFwdMstr) return 1;
    if (op.smask == kHeaFwdSmaskHi && op.mstr == kHeaFwdMstrHi) return 2;
    return 0;
}

// one rotation step of the reverse shape on the pairs (p0, p1), (q0, q1) of A
// and B: the matrix, then Re sum conj(b) (G a)
template <int MS>
__device__ __forceinline__ double hea_step(const double2* __restrict__ pool, int64_t mat, int64_t gen, double2& p0, double2& p1,
                                           double2& q0, double2& q1, double2& r0, double2& r1, double2& s0, double2& s1) {
    const Mat2 m = load_mat2(pool + mat);
    m2_ip<MS>(m, p0, p1);
    m2_ip<MS>(m, q0, q1);
    m2_ip<MS>(m, r0, r1);
    m2_ip<MS>(m, s0, s1);
    const Mat2 g = load_mat2(pool + gen);
    return gdot2<MS>(g, p0, p1, r0, r1) + gdot2<MS>(g, q0, q1, s0, s1);
}

template <bool CXLO>
__device__ __forceinline__ void mop_hea_rev(double2* __restrict__ sa, double2* __restrict__ sb, int t, const MicroOp& op,
                                            const double2* __restrict__ pool, double2* red, int buf) {
    double g1 = 0.0, g2 = 0.0, g3 = 0.0;
    const uint32_t ng = 1u << (t - 2);
    const uint32_t olo = 1u << op.lpos[0], ohi = 1u << op.lpos[1];
    const int64_t m1 = op.cmat[1], m2 = op.cmat[2], m3 = op.cmat[3], e1 = op.gen[1], e2 = op.gen[2], e3 = op.gen[3];
    for (uint32_t g = threadIdx.x; g < ng; g += kTileThreads) {
        const uint32_t base = ins32<2>(g, op.lpos);
        // x00, x01, x10, x11 over (lo, hi) after the CNOT: the control's set half
        // has its target flipped, i.e. two registers trade names
        const uint32_t i00 = base, i10 = base + olo, i01 = base + ohi, i11 = base + olo + ohi;
        double2 a00 = sa[i00], a10, a01, a11, b00 = sb[i00], b10, b01, b11;
        if (CXLO) {  // control lo, target hi: (lo=1, hi=0) <-> (lo=1, hi=1)
            a10 = sa[i11];
            a11 = sa[i10];
            a01 = sa[i01];
            b10 = sb[i11];
            b11 = sb[i10];
            b01 = sb[i01];
        } else {  // control hi, target lo: (lo=0, hi=1) <-> (lo=1, hi=1)
            a01 = sa[i11];
            a11 = sa[i01];
            a10 = sa[i10];
            b01 = sb[i11];
            b11 = sb[i01];
            b10 = sb[i10];
        }
        if (CXLO) {  // rotations on the hi bit: pairs (00, 01), (10, 11)
            g1 += hea_step<MS_DIAG>(pool, m1, e1, a00, a01, a10, a11, b00, b01, b10, b11);
            g2 += hea_step<MS_REAL>(pool, m2, e2, a00, a01, a10, a11, b00, b01, b10, b11);
            g3 += hea_step<MS_XLIKE>(pool, m3, e3, a00, a01, a10, a11, b00, b01, b10, b11);
        } else {  // rotations on the lo bit: pairs (00, 10), (01, 11)
            g1 += hea_step<MS_DIAG>(pool, m1, e1, a00, a10, a01, a11, b00, b10, b01, b11);
            g2 += hea_step<MS_REAL>(pool, m2, e2, a00, a10, a01, a11, b00, b10, b01, b11);
            g3 += hea_step<MS_XLIKE>(pool, m3, e3, a00, a10, a01, a11, b00, b10, b01, b11);
        }
        sa[i00] = a00;
        sa[i10] = a10;
        sa[i01] = a01;
        sa[i11] = a11;
        sb[i00] = b00;
        sb[i10] = b10;
        sb[i01] = b01;
        sb[i11] = b11;
    }
    wave_partial_re(red, buf, 0, g1);
    wave_partial_re(red, buf, 1, g2);
    wave_partial_re(red, buf, 2, g3);
}

template <bool CXLO>
__device__ __forceinline__ void mop_hea_fwd(double2* __restrict__ sa, int t, const MicroOp& op, const double2* __restrict__ pool) {
    const uint32_t ng = 1u << (t - 2);
    const uint32_t olo = 1u << op.lpos[0], ohi = 1u << op.lpos[1];
    const Mat2 m = load_mat2(pool + op.cmat[1]);
    for (uint32_t g = threadIdx.x; g < ng; g += kTileThreads) {
        const uint32_t base = ins32<2>(g, op.lpos);
        const uint32_t i00 = base, i10 = base + olo, i01 = base + ohi, i11 = base + olo + ohi;
        double2 a00 = sa[i00], a10, a01, a11;
        if (CXLO) {  // CNOT with control lo, then the 2x2 on lo: pairs (00, 10), (01, 11)
            a10 = sa[i11];
            a11 = sa[i10];
            a01 = sa[i01];
            m2_ip<MS_DENSE>(m, a00, a10);
            m2_ip<MS_DENSE>(m, a01, a11);
        } else {  // control hi, then the 2x2 on hi: pairs (00, 01), (10, 11)
            a01 = sa[i11];
            a11 = sa[i01];
            a10 = sa[i10];
            m2_ip<MS_DENSE>(m, a00, a01);
            m2_ip<MS_DENSE>(m, a10, a11);
        }
        sa[i00] = a00;
        sa[i10] = a10;
        sa[i01] = a01;
        sa[i11] = a11;
    }
}

// MK_HAM: B += Dm A with Dm = M + M^dagger (op.mat2); acc += sum conj(A) (M A)
template <int KB>
__device__ __forceinline__ void mop_ham(const double2* __restrict__ sa, double2* __restrict__ sb, int t, const MicroOp& op,
                                        const double2* __restrict__ pool, double2* red, int buf) {
    constexpr int D = 1 << KB;
    constexpr int RU = KB == 1 ? D : 1;
    const double2* __restrict__ M = pool + op.mat;
    const double2* __restrict__ Dm = pool + op.mat2;
    double er = 0.0, ei = 0.0;
    const uint32_t ng = 1u << (t - KB);
    for (uint32_t g = threadIdx.x; g < ng; g += kTileThreads) {
        const uint32_t base = ins32<KB>(g, op.lpos);
        double2 a[D];
#pragma unroll
        for (int j = 0; j < D; ++j) a[j] = sa[base + off32<KB>(j, op.lpos)];
#pragma unroll RU
        for (int r = 0; r < D; ++r) {
            double2 s = make_double2(0.0, 0.0), h = make_double2(0.0, 0.0);
#pragma unroll
            for (int c = 0; c < D; ++c) {
                s = cfma(M[r + c * D], a[c], s);
                h = cfma(Dm[r + c * D], a[c], h);
            }
            const uint32_t o = base + off32<KB>(r, op.lpos);
            const double2 ar = sa[o];
            er = fma(ar.x, s.x, fma(ar.y, s.y, er));
            ei = fma(ar.x, s.y, fma(-ar.y, s.x, ei));
            sb[o] = cadd(sb[o], h);
        }
    }
    wave_partial(red, buf, 0, er, ei);
}

// MK_HAMS: the nrec MK_HAM records at rec (terms of 1 or 2 bits) in one pass.
// Each thread takes whole amplitudes j of the tile and sums, over the terms, row
// j of (M + M^dagger) A into B[j] and conj(A[j]) (M A)[j] into one energy
// partial. A is only read and each B[j] is written by the thread that owns j, so
// the terms need no barrier between them: one micro-op instead of one per term.
// A diagonal term (mstr MS_DIAG, read from the values by the planner) reads no
// other amplitude; a dense one reads the 2 or 4 amplitudes of j's group.
__device__ __forceinline__ void ham_term1(const double2* __restrict__ sa, uint32_t j, double2 aj, const MicroOp& op,
                                          const double2* __restrict__ pool, double2& s, double2& h) {
    const double2* __restrict__ M = pool + op.mat;
    const double2* __restrict__ Dm = pool + op.mat2;
    const uint32_t o0 = 1u << op.lpos[0];
    const bool r = (j & o0) != 0;
    if (ms_mat(op, 0) == MS_DIAG) {
        s = cfma(r ? M[3] : M[0], aj, s);
        h = cfma(r ? Dm[3] : Dm[0], aj, h);
        return;
    }
    // row r of the 2 x 2 products over the pair (a0, a1) holding j
    const double2 ao = sa[j ^ o0];
    const double2 a0 = r ? ao : aj, a1 = r ? aj : ao;
    s = cfma(r ? M[3] : M[2], a1, cfma(r ? M[1] : M[0], a0, s));
    h = cfma(r ? Dm[3] : Dm[2], a1, cfma(r ? Dm[1] : Dm[0], a0, h));
}
__device__ __forceinline__ void ham_term2(const double2* __restrict__ sa, uint32_t j, double2 aj, const MicroOp& op,
                                          const double2* __restrict__ pool, double2& s, double2& h) {
    const uint32_t o0 = 1u << op.lpos[0], o1 = 1u << op.lpos[1];
    const int r = ((j & o0) ? 1 : 0) | ((j & o1) ? 2 : 0);
    // row r of column-major 4 x 4 matrices: entries r, r + 4, r + 8, r + 12
    const double2* __restrict__ M = pool + op.mat + r;
    const double2* __restrict__ Dm = pool + op.mat2 + r;
    if (ms_mat(op, 0) == MS_DIAG) {
        s = cfma(M[4 * r], aj, s);  // entry (r, r)
        h = cfma(Dm[4 * r], aj, h);
        return;
    }
    const uint32_t base = j & ~(o0 | o1);
    const double2 x0 = sa[base], x1 = sa[base | o0], x2 = sa[base | o1], x3 = sa[base | o0 | o1];
    s = cfma(M[12], x3, cfma(M[8], x2, cfma(M[4], x1, cfma(M[0], x0, s))));
    h = cfma(Dm[12], x3, cfma(Dm[8], x2, cfma(Dm[4], x1, cfma(Dm[0], x0, h))));
}
__device__ __forceinline__ void mop_hams(const double2* __restrict__ sa, double2* __restrict__ sb, int t, const MicroOp* rec,
                                         int nrec, const double2* __restrict__ pool, double2* red, int buf) {
    const uint32_t T = 1u << t;
    double er = 0.0, ei = 0.0;
    for (uint32_t j = threadIdx.x; j < T; j += kTileThreads) {
        const double2 aj = sa[j];
        double2 s = make_double2(0.0, 0.0), h = make_double2(0.0, 0.0);
        for (int q = 0; q < nrec; ++q) {
            if (rec[q].kb == 1) ham_term1(sa, j, aj, rec[q], pool, s, h);
            else ham_term2(sa, j, aj, rec[q], pool, s, h);
        }
        sb[j] = cadd(sb[j], h);
        er = fma(aj.x, s.x, fma(aj.y, s.y, er));
        ei = fma(aj.x, s.y, fma(-aj.y, s.x, ei));
    }
    wave_partial(red, buf, 0, er, ei);
}

// MK_DENSE: A <- M A with the 2^K x 2^K complex M (column-major over lpos[0..K),
// ascending) as a real GEMM on the matrix cores. A group (2^K amplitudes, the
// other tile bits fixed) is one column of X; in real form M is 2D x 2D with
// entries (row 2i + p, column 2j + q) = Re M_ij, -Im M_ij, Im M_ij, Re M_ij for
// (p, q) = (0,0), (0,1), (1,0), (1,1), which acts on the interleaved (re, im)
// amplitudes as stored. v_mfma_f64_16x16x4f64 takes a 16 x 4 slice of that
// matrix (lane l: row l % 16, k = l / 16), a 4 x 16 slice of X (lane l: k = l / 16,
// column l % 16) and accumulates a 16 x 16 block (lane l, register i: row
// l / 16 + 4 i, column l % 16; checked on gfx950 with a lane-layout probe). The
// k order is chosen so that one 16-byte LDS read serves two steps: steps 2s and
// 2s + 1 take, in lane group g = l / 16, the real and imaginary parts of
// amplitude j = 4s + g. The row order is chosen so that each lane's output is
// whole amplitudes: block row rho = g + 4 i of row tile rt is amplitude
// 8 rt + 2 g + i / 2, part i % 2, so registers (0, 1) and (2, 3) are amplitudes
// 8 rt + 2 g and 8 rt + 2 g + 1. A wave owns whole column tiles (16 groups),
// computes every row tile of them, and only then writes them back: the reads and
// writes of one column stay in one wave, so the in-place update needs no extra
// barrier.
typedef double vqc_v4d __attribute__((ext_vector_type(4)));
template <int K>
__device__ __forceinline__ void mop_dense(double2* __restrict__ s, int t, const MicroOp& op, const double2* __restrict__ M) {
    constexpr int D = 1 << K;          // amplitudes per group
    constexpr int RT = (2 * D) / 16;   // 16-row tiles of the real matrix
    constexpr int NS = D / 4;          // pairs of k steps
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, g = lane >> 4;
    const uint32_t ncols = 1u << (t - K);  // groups
    const uint32_t nct = ncols >> 4;       // column tiles (the planner keeps t - K >= 7)
    // A-operand entries of this lane: block row rho = lane % 16 = g' + 4 i' is complex
    // row 8 rt + 2 g' + i' / 2, part i' % 2; step 2q + h reads complex column 4q + g, part h
    const int prow = (col >> 2) & 1, rrow = 2 * (col & 3) + (col >> 3);
    const double2* __restrict__ Mrow = M + rrow + size_t(g) * D;  // + 8 rt + 4 q D
    const uint32_t offg = off32<K>(g, op.lpos);
    if constexpr (K <= 4) {
        // A operands of every step in registers for the whole micro-op (K = 4: 32
        // VGPRs); column tiles are taken CT at a time, so that the matrix cores see
        // RT * CT independent accumulation chains
        constexpr int CT = K == 3 ? 2 : 1;
        double a0v[NS][RT], a1v[NS][RT];
#pragma unroll
        for (int q = 0; q < NS; ++q)
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) {
                const double2 m = Mrow[8 * rt + size_t(4 * q) * D];
                // (p, h) = (0, 0): Re m, (0, 1): -Im m, (1, 0): Im m, (1, 1): Re m
                a0v[q][rt] = prow ? m.y : m.x;
                a1v[q][rt] = prow ? m.x : -m.y;
            }
        uint32_t xo[NS];
#pragma unroll
        for (int q = 0; q < NS; ++q) xo[q] = offg + off32<K>(4 * q, op.lpos);
        for (uint32_t ct0 = wave; ct0 < nct; ct0 += CT * kTileWaves) {
            uint32_t ob[CT];
            bool live[CT];
            double2 x[CT][NS];
#pragma unroll
            for (int c = 0; c < CT; ++c) {
                const uint32_t ct = ct0 + c * kTileWaves;
                live[c] = ct < nct;  // uniform over the wave
                ob[c] = ins32<K>((live[c] ? ct : ct0) * 16 + col, op.lpos);
#pragma unroll
                for (int q = 0; q < NS; ++q) x[c][q] = s[ob[c] + xo[q]];
            }
            vqc_v4d acc[CT][RT];
#pragma unroll
            for (int c = 0; c < CT; ++c)
#pragma unroll
                for (int rt = 0; rt < RT; ++rt) acc[c][rt] = vqc_v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < NS; ++q)
#pragma unroll
                for (int rt = 0; rt < RT; ++rt)
#pragma unroll
                    for (int c = 0; c < CT; ++c) {
                        acc[c][rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0v[q][rt], x[c][q].x, acc[c][rt], 0, 0, 0);
                        acc[c][rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1v[q][rt], x[c][q].y, acc[c][rt], 0, 0, 0);
                    }
#pragma unroll
            for (int c = 0; c < CT; ++c) {
                if (!live[c]) continue;
#pragma unroll
                for (int rt = 0; rt < RT; ++rt) {
                    s[ob[c] + off32<K>(8 * rt + 2 * g, op.lpos)] = make_double2(acc[c][rt][0], acc[c][rt][1]);
                    s[ob[c] + off32<K>(8 * rt + 2 * g + 1, op.lpos)] = make_double2(acc[c][rt][2], acc[c][rt][3]);
                }
            }
        }
        return;
    }
    for (uint32_t ct = wave; ct < nct; ct += kTileWaves) {
        const uint32_t ob = ins32<K>(ct * 16 + col, op.lpos);
        const uint32_t base = ob + offg;
        vqc_v4d acc[RT];
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) acc[rt] = vqc_v4d{0.0, 0.0, 0.0, 0.0};
        // the k loop stays rolled for the largest blocks, so that the matrix entries
        // are read per step instead of all held in registers
#pragma unroll 2
        for (int q = 0; q < NS; ++q) {
            const double2 x = s[base + off32<K>(4 * q, op.lpos)];
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) {
                const double2 m = Mrow[8 * rt + size_t(4 * q) * D];
                const double a0 = prow ? m.y : m.x;
                const double a1 = prow ? m.x : -m.y;
                acc[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, x.x, acc[rt], 0, 0, 0);
                acc[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, x.y, acc[rt], 0, 0, 0);
            }
        }
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
            s[ob + off32<K>(8 * rt + 2 * g, op.lpos)] = make_double2(acc[rt][0], acc[rt][1]);
            s[ob + off32<K>(8 * rt + 2 * g + 1, op.lpos)] = make_double2(acc[rt][2], acc[rt][3]);
        }
    }
}

// One micro-op followed by a barrier. `red` holds kRedDoubles2 entries and
// `gacc` the accumulator rows addressed by op.acc[]. MAXKB is the widest
// micro-op the calling kernel is built for: kernels for phases of at most 3
// local bits carry no 4/5-bit code, which keeps them free of register spills.
// `op` is an element of the phase's micro-op array; the return value is the
// number of records after it that it consumed (the terms of an MK_HAMS), which
// the caller skips.
template <int MAXKB, int KS>
__device__ __forceinline__ int run_mop(double2* sa, double2* sb, int t, const MicroOp& op, const double2* pool, double2* red,
                                       double2* gacc, AccState& st) {
    acc_fold(st, red, gacc);
    const int kb = op.kb;
    const int buf = st.cur_buf;
    int nred = 0, nrec = 0;
    switch (op.kind) {
        case MK_APPLY:
            if constexpr ((KS & (1 << MK_APPLY)) != 0) {
            constexpr bool roll = (KS & (1 << MK_APPLY2)) != 0;  // a reverse kernel
            double2* s = op.buf ? sb : sa;
            if (kb == 1) mop_apply<1, roll>(s, t, op, pool);
            else if (kb == 2 && (op.mstr & kMsParity)) mop_parity<false>(s, sb, t, op, pool);
            else if (kb == 2) mop_apply<2, roll>(s, t, op, pool);
            else if constexpr (MAXKB >= 3) {
                if (kb == 3) mop_apply<3, roll>(s, t, op, pool);
                else if constexpr (MAXKB >= 5) {
                    if (kb == 4) mop_apply<4, roll>(s, t, op, pool);
                    else mop_apply<5, roll>(s, t, op, pool);
                }
            }
            }
            break;
        case MK_APPLY2:
        case MK_APPLY2X:
        case MK_GRAD:
            if constexpr ((KS & ((1 << MK_APPLY2) | (1 << MK_APPLY2X) | (1 << MK_GRAD))) != 0) {
                bool done = false;
                if constexpr ((KS & (1 << MK_APPLY2X)) != 0) {
                    if (kb == 2 && op.kind == MK_APPLY2X && op.ng == 0 && (op.mstr & kMsParity)) {
                        mop_parity<true>(sa, sb, t, op, pool);
                        done = true;
                    }
                }
                if (done) {
                } else if (kb == 1) mop_pair<1>(sa, sb, t, op, pool, red, buf);
                else if (kb == 2) mop_pair<2>(sa, sb, t, op, pool, red, buf);
                else if constexpr (MAXKB >= 3) {
                    if (kb == 3) mop_pair<3>(sa, sb, t, op, pool, red, buf);
                    else if constexpr (MAXKB >= 5) {
                        if (kb == 4) mop_pair<4>(sa, sb, t, op, pool, red, buf);
                        // 5 bits: gradient-only terms (MK_GRAD), general reverse kernels only
                        else if constexpr ((KS & (1 << MK_GRAD)) != 0) mop_pair<5>(sa, sb, t, op, pool, red, buf);
                    }
                }
                nred = op.ng;
            }
            break;
        case MK_CHAIN:
            if constexpr ((KS & (1 << MK_CHAIN)) != 0) {
                mop_chain(sa, sb, t, op, pool, red, buf);
                nred = op.ng;
            }
            break;
        case MK_BLOCK:
            if constexpr ((KS & (1 << MK_BLOCK)) != 0) {
                // forward micro-ops act on A alone; reverse ones on A and B
                constexpr bool two = KS != KS_FWD && KS != KS_FWDH && KS != KS_FWDD;
                if constexpr (KS == KS_REV) {  // state-vector reverse: the ansatz shape first
                    const int shape = hea_rev_shape(op);
                    if (shape == 1) mop_hea_rev<true>(sa, sb, t, op, pool, red, buf);
                    else if (shape == 2) mop_hea_rev<false>(sa, sb, t, op, pool, red, buf);
                    else mop_block(sa, sb, t, op, pool, red, buf, two);
                } else if constexpr (KS == KS_REVX) {  // general reverse (density matrices): the same shape
                    const int shape = hea_rev_shape(op);
                    if (shape == 1) mop_hea_rev<true>(sa, sb, t, op, pool, red, buf);
                    else if (shape == 2) mop_hea_rev<false>(sa, sb, t, op, pool, red, buf);
                    else mop_block(sa, sb, t, op, pool, red, buf, two);
                } else if constexpr (KS == KS_FWDH) {
                    const int shape = hea_fwd_shape(op);
                    if (shape == 1) mop_hea_fwd<true>(sa, t, op, pool);
                    else if (shape == 2) mop_hea_fwd<false>(sa, t, op, pool);
                    else mop_block(sa, sb, t, op, pool, red, buf, two);
                } else {
                    mop_block(sa, sb, t, op, pool, red, buf, two);
                }
                nred = two ? op.ng : 0;
            }
            break;
        case MK_DENSE:
            if constexpr ((KS & (1 << MK_DENSE)) != 0) {
                const double2* M = pool + op.mat;
                if (kb == 3) mop_dense<3>(sa, t, op, M);
                else if (kb == 4) mop_dense<4>(sa, t, op, M);
                else mop_dense<5>(sa, t, op, M);
            }
            break;
        case MK_HAM:
            if constexpr ((KS & (1 << MK_HAM)) != 0) {
            if (kb == 1) mop_ham<1>(sa, sb, t, op, pool, red, buf);
            else if (kb == 2) mop_ham<2>(sa, sb, t, op, pool, red, buf);
            else if constexpr (MAXKB >= 3) {
                if (kb == 3) mop_ham<3>(sa, sb, t, op, pool, red, buf);
                else if constexpr (MAXKB >= 5) {
                    if (kb == 4) mop_ham<4>(sa, sb, t, op, pool, red, buf);
                    else mop_ham<5>(sa, sb, t, op, pool, red, buf);
                }
            }
            nred = 1;
            }
            break;
        case MK_HAMS:
            if constexpr ((KS & (1 << MK_HAMS)) != 0) {
                nrec = op.nchain;
                mop_hams(sa, sb, t, &op + 1, nrec, pool, red, buf);
                nred = 1;
            }
            break;
    }
    __syncthreads();
    acc_push(st, nred, op.acc);
    return nrec;
}

// Fold the last pending partials; ends with a barrier so gacc is complete.
__device__ __forceinline__ void acc_finish(AccState& st, const double2* red, double2* gacc) {
    acc_fold(st, red, gacc);
    __syncthreads();
}

// Constant-memory view of a buffer that no thread writes while the kernel runs
// (micro-ops, phases, the matrix pool once the variant kernels are done). A
// kernel with fences and atomics (the batch work queue) otherwise makes the
// compiler treat every read as possibly clobbered, so uniform reads become
// vector loads held in VGPRs; reads through these pointers are scalar loads.
template <class T>
using const_ptr = const __attribute__((address_space(4))) T*;
template <class T>
__device__ __forceinline__ const_ptr<T> as_const(const T* p) {
    return (const_ptr<T>)(uintptr_t)p;
}

// Read ahead what micro-ops [op_begin, op_end) will read: their descriptors and
// the pool range [pf_lo, pf_lo + pf_n). Each micro-op reads its descriptor and
// then its matrices with scalar loads, two dependent trips to memory per
// micro-op; issued here together by all threads as vector loads, they bring
// every such line into the L2 in one trip. The values are written to `sink`
// (64 entries of LDS that the caller overwrites before reading them), so that
// the loads are kept.
__device__ __forceinline__ void warm_ops(const MicroOp* __restrict__ ops, int op_begin, int op_end, const double2* __restrict__ pool,
                                         int64_t pf_lo, int32_t pf_n, double2* sink) {
    double s = 0.0;
    const double2* __restrict__ q = pool + pf_lo;
    for (int i = threadIdx.x; i < pf_n; i += kTileThreads) s += q[i].x;
    // descriptors as raw 8-byte words (sizeof(MicroOp) is a multiple of 8)
    const double* __restrict__ od = reinterpret_cast<const double*>(ops + op_begin);
    const int on = (op_end - op_begin) * int(sizeof(MicroOp) / 8);
    for (int i = threadIdx.x; i < on; i += kTileThreads) s += od[i];
    sink[threadIdx.x & 63] = make_double2(s, 0.0);
}

// this thread's part of every tile index of phase p (see tile_deposit)
__device__ __forceinline__ uint64_t thread_deposit(const PhaseDesc& p) {
    return tile_deposit(threadIdx.x, 0, std::min(p.t, kTileThreadBits), p);
}

// Tile data is read and written once per pass, so its loads and stores carry the
// non-temporal hint (streaming, no reuse in the caches): 28-qubit fused circuit
// 89.7 -> 88.6 ms, 256-instance batch 52.0 -> 49.5 ms (same-box A/B).
typedef double tile_d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ld_stream(const double2* p) {
    const tile_d2v v = __builtin_nontemporal_load(reinterpret_cast<const tile_d2v*>(p));
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st_stream(double2* p, double2 x) {
    tile_d2v v;
    v.x = x.x;
    v.y = x.y;
    __builtin_nontemporal_store(v, reinterpret_cast<tile_d2v*>(p));
}

// Stream one tile of phase p through LDS.
//  PH_FWD: sa <- srcA tile; ops; dstA tile <- sa
//  PH_REV: sa <- A, sb <- B; ops; A <- sa, B <- sb
//  PH_HAM: sa <- A, sb <- (ham_accum ? B : 0); ops; B <- sb
// Global loads are issued kLoadBatch at a time per thread before any LDS store
// so that each wave keeps several 1 KiB requests in flight.
template <int MAXKB, int KS>
__device__ __forceinline__ void phase_tile(const PhaseDesc& p, uint64_t tile, uint64_t tdep, const double2* srcA, double2* A,
                                           double2* B, const MicroOp* __restrict__ ops, const double2* __restrict__ pool,
                                           double2* sa, double2* sb, double2* red, double2* gacc, AccState& st) {
    const uint32_t T = 1u << p.t;
    const bool loadB = p.kind == PH_REV || (p.kind == PH_HAM && p.ham_accum);
    const bool zeroB = p.kind == PH_HAM && !p.ham_accum;
    const bool init = p.flags & PF_INIT_ZERO;
    // this thread's state-index bits: the tile's base and the deposit of threadIdx.x
    // (tdep = thread_deposit(p), computed once per phase by the caller)
    const uint64_t tbase = tile_base(tile, p) | tdep;
    // per thread: T / kTileThreads amplitudes (uniform), read kLoadBatch at a time
    const uint32_t per = (T + kTileThreads - 1) / kTileThreads;
    for (uint32_t u0 = 0; u0 < per; u0 += kLoadBatch) {
        double2 va[kLoadBatch], vb[kLoadBatch];
#pragma unroll
        for (int u = 0; u < kLoadBatch; ++u) {
            va[u] = make_double2(0.0, 0.0);
            vb[u] = make_double2(0.0, 0.0);
            const uint32_t j = (u0 + u) * kTileThreads + threadIdx.x;
            if (u0 + u < per && j < T) {
                const uint64_t gi = tbase | tile_deposit((u0 + u) << kTileThreadBits, kTileThreadBits, p.t, p);
                if (init) va[u].x = gi == 0 ? 1.0 : 0.0;
                else va[u] = ld_stream(srcA + gi);
                if (loadB) vb[u] = ld_stream(B + gi);
            }
        }
#pragma unroll
        for (int u = 0; u < kLoadBatch; ++u) {
            const uint32_t j = (u0 + u) * kTileThreads + threadIdx.x;
            if (u0 + u < per && j < T) {
                sa[j] = va[u];
                if (loadB || zeroB) sb[j] = vb[u];
            }
        }
    }
    __syncthreads();
    for (int o = p.op_begin; o < p.op_end; ++o) o += run_mop<MAXKB, KS>(sa, sb, p.t, ops[o], pool, red, gacc, st);
    for (uint32_t u = 0; u < per; ++u) {
        const uint32_t j = u * kTileThreads + threadIdx.x;
        if (j < T) {
            const uint64_t gi = tbase | tile_deposit(u << kTileThreadBits, kTileThreadBits, p.t, p);
            if (p.kind != PH_HAM) st_stream(A + gi, sa[j]);
            if (p.kind != PH_FWD) st_stream(B + gi, sb[j]);
        }
    }
    __syncthreads();
}

// Whole evaluation of a state held in one tile: sa <- src (|0...0> when src is
// null), then the forward micro-ops [mop_begin, ham_begin), the Hamiltonian
// ones [ham_begin, rev_begin) into sb (zeroed first) and the reverse ones
// [rev_begin, mop_end), each stage in its own loop; A receives the forward state. The accumulator index
// of every micro-op is a local row; row e is written to rows[e * row_stride].
// LDS: sa, sb (2^nbits each), red (kRedDoubles2), gacc (nrows).
template <int MAXKB, int RKS, int HKS = KS_HAM>
__device__ __forceinline__ void small_eval_body(int nbits, const double2* src, double2* A, const MicroOp* __restrict__ ops,
                                                int mop_begin, int ham_begin, int rev_begin, int mop_end,
                                                const double2* __restrict__ pool, double2* sa, double2* sb, double2* red,
                                                double2* gacc, int nrows, double2* rows, size_t row_stride) {
    const uint32_t T = 1u << nbits;
    for (int i = threadIdx.x; i < nrows; i += kTileThreads) gacc[i] = make_double2(0.0, 0.0);
    for (uint32_t j = threadIdx.x; j < T; j += kTileThreads) {
        sa[j] = src ? src[j] : make_double2(j == 0 ? 1.0 : 0.0, 0.0);
        sb[j] = make_double2(0.0, 0.0);
    }
    __syncthreads();
    AccState st;
    acc_init(st);
    for (int o = mop_begin; o < ham_begin; ++o) o += run_mop<MAXKB, KS_FWD>(sa, sb, nbits, ops[o], pool, red, gacc, st);
    if (A) {
        for (uint32_t j = threadIdx.x; j < T; j += kTileThreads) A[j] = sa[j];
    }
    for (int o = ham_begin; o < rev_begin; ++o) o += run_mop<MAXKB, HKS>(sa, sb, nbits, ops[o], pool, red, gacc, st);
    for (int o = rev_begin; o < mop_end; ++o) o += run_mop<MAXKB, RKS>(sa, sb, nbits, ops[o], pool, red, gacc, st);
    acc_finish(st, red, gacc);
    for (int e = threadIdx.x; e < nrows; e += kTileThreads) rows[size_t(e) * row_stride] = gacc[e];
    __syncthreads();
}

// ---------------------------------------------------------------------------
// host-side planning
// ---------------------------------------------------------------------------
struct TileTuning {
    int t_fwd = 12;     // tile bits of forward phases
    int t_rev = 11;     // tile bits of reverse / Hamiltonian phases (two buffers)
    int t_small = 12;   // whole state in one tile up to this many bits (both buffers)
    int mlow = 4;       // contiguous low bits kept in every tile
};
TileTuning vqc_tile_tuning(int nbits = -1);  // nbits: one state of that size; -1: a batch

struct OpBits {
    std::vector<int> bits;  // global index bits the op touches
};

// Host copy of the device matrix pool: the caller's matrices first, then the
// variants the micro-ops use (cached per (offset, variant)).
struct XPoolEntry {
    int64_t src, dst;
    int32_t D, v;        // v: MF_CONJ|MF_TRANS bits, 4 = M + M^dagger, 8 = product of srcs
    int64_t srcs[4];     // v = 8: factors in application order (first applied first)
    int32_t nsrc;
    int32_t pad;
};
// The product of a run of gates as one dense matrix on the union of their bits
// (MK_DENSE): factor f acts on nb[f] of the K block bits, its matrix bit j on
// block bit map[5 f + j]; the product is column-major over the block bits in
// ascending order.
struct XPoolBlock {
    int64_t dst;
    int32_t K;
    std::vector<int64_t> mats;  // factor matrices, in application order
    std::vector<int32_t> nb;
    std::vector<int32_t> map;
};
struct XPoolClaim {
    int64_t off;
    int32_t D, ms;  // what was found for the D x D matrix at off: its class (MS_*), or for
    int32_t test;   // test 1, whether it is a parity-block channel (1) or not (0), for
                    // test 2, whether it is diagonal (1) or not (0)
    int32_t pad;
};
struct XPool {
    std::vector<double> h;  // interleaved complex
    std::unordered_map<uint64_t, int64_t> cache;  // (source offset, variant) -> offset
    // deferred: only offsets are assigned, to be computed on the device once the
    // caller's matrices arrive
    bool deferred = false;
    int64_t len = 0;
    // every derived matrix in the order it was made (deferred or not)
    std::vector<XPoolEntry> table;
    // every structure class the planner read from values (not deferred)
    std::vector<XPoolClaim> claims;
    // dense block products (not deferred), remade after `table` on reload
    std::vector<XPoolBlock> blocks;
    // the product of a dense block, recorded (host pools only)
    int64_t dense_block(XPoolBlock b);
    int64_t size() const { return deferred ? len : int64_t(h.size() / 2); }
    // variant: MF_CONJ / MF_TRANS combination, or 4 for M + M^dagger
    int64_t variant(int64_t off, int D, int v);
    // U_n ... U_2 U_1 for D x D factors given in application order (n <= 4)
    int64_t product(const int64_t* srcs, int n, int D);
    // the structure class (MS_*) of the D x D matrix at off, from its values, recorded
    int classify(int64_t off, int D);
    // whether the 4 x 4 matrix at off only couples indices of equal bit parity, recorded
    bool parity_blocks(int64_t off);
    // whether the D x D matrix at off is diagonal, recorded
    bool diagonal(int64_t off, int D);
    // New caller matrices (pool, then hpool, the lengths the pool was made with):
    // every derived matrix is recomputed from them in the order it was made.
    // False when a structure class the planner read no longer holds.
    bool reload(const double* pool, int64_t plen, const double* hpool, int64_t hlen);
};

struct PlanBuilder {
    XPool* xp = nullptr;
    // forward phases may hold MK_DENSE micro-ops (the caller runs them with the
    // tile-phase kernels, not the whole-state small kernels)
    bool dense_fwd = false;
    std::vector<MicroOp> mops;
    std::vector<PhaseDesc> phases;
    int nrows = 0;
    // row -> (gradient slot, or -1 for an energy row)
    std::vector<int> row_slot;
};

// Choose tile bits S for a set of required bits; fills the contiguous low bits.
void vqc_make_phase(PhaseDesc& p, int kind, int nbits, int t, const std::vector<int>& req);

// Plan the forward phases of ops (SV when !is_dm) onto nbits with tile bits t.
// pool offsets are used as given.
int vqc_plan_forward(PlanBuilder& pb, int nbits, int n, int is_dm, int nops, const vqc_op_t* ops, int t, int mlow);
// Plan H phases: MK_HAM micro-ops for each term; one accumulator (energy row) per term.
int vqc_plan_ham(PlanBuilder& pb, int nbits, int nterms, const vqc_term_t* terms, int64_t hoff, int t, int mlow);
// Plan the reverse (adjoint) phases.
int vqc_plan_reverse(PlanBuilder& pb, int nbits, int n, int is_dm, int nops, const vqc_op_t* ops, int t, int mlow);
// Merge runs of one-bit micro-ops on the same bit: forward MK_APPLY runs become one
// MK_APPLY with the product matrix, reverse MK_APPLY2 runs become MK_CHAIN.
void vqc_fuse_chains(PlanBuilder& pb, size_t phase_begin, size_t phase_end);
// first micro-op of phase i (the micro-op count when i is past the last phase)
inline int vqc_phase_op_start(const PlanBuilder& pb, size_t i) {
    return i < pb.phases.size() ? pb.phases[i].op_begin : int(pb.mops.size());
}
// The pool range [*pf_lo, *pf_lo + *pf_n) holding every matrix that micro-ops
// [op_begin, op_end) read (*pf_n = 0 when it spans more than kPrefetchMax entries).
void vqc_prefetch_range(const std::vector<MicroOp>& mops, int op_begin, int op_end, int64_t plen, int64_t* pf_lo, int32_t* pf_n);
// micro-ops a KS_REV kernel runs
inline bool vqc_rev_simple(const MicroOp& m) { return m.kind == MK_APPLY2 || m.kind == MK_CHAIN || m.kind == MK_BLOCK; }
