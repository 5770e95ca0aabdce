// vqc_circuit.hip -- circuits, energies and adjoint gradients.
//
// Fused path: the circuit is cut into LDS tile phases (vqc_tile.h); the
// forward pass, the Hamiltonian pass (adjoint seed and energy) and the reverse
// adjoint sweep each cost one HBM round trip per phase instead of one per gate.
// States of at most t_small bits run the whole evaluation in one workgroup
// with both the state and the adjoint held in LDS.
//
// Plain path (VQC_FLAG_UNFUSED): one kernel per gate and per reduction, the
// same sequence of sweeps as the reference's back_propagate loop
// (circuitdiff.jl:23-32, 61-78); it is the per-circuit baseline the fused and
// batched forms are measured against.
#include <cstdlib>
#include <memory>

#include "vqc_tile.h"

int vqc_dense_sweep(double2* v, int nbits, int k, const int32_t* pos, const double* U, hipStream_t st);
int vqc_accum_sweep(const double2* in, double2* out, int nbits, int k, const int32_t* pos, const double2* Ud, hipStream_t st);
int vqc_expec_sweep(const double2* ket, const double2* bra, int nbits, int k, const int32_t* pos, const double2* Ud,
                    double2* parts, unsigned* grid_out, hipStream_t st);
int vqc_dm_expec_sweep(const double2* rho, int n, int k, const int32_t* pos, const double2* Ud, double2* parts,
                       unsigned* grid_out, hipStream_t st);
void vqc_zero(double2* v, uint64_t len, hipStream_t st);

// Tile shape. Tile-bit fields left at 0 resolve by what is evaluated. One
// state of nbits index bits (nbits >= 0) runs on one workgroup up to 9 bits,
// else in tiles of 9 bits up to 14, 10 up to 18, 11 up to 22 and 12 (forward) /
// 11 (reverse) above, with 3 contiguous low bits from 21 bits on (22 qubits:
// 22.2 -> 19.5 ms at 11 / 11 against 12 / 11): 2^(nbits - t) tiles must keep enough CUs busy, while
// each phase, one launch, covers fewer gates the smaller its tiles (the sizes
// measured fastest for layered circuits of 8 to 20 qubits on MI355X). A wide
// batch (nbits < 0), whose instances fill the device together, takes the
// largest tiles, 13 / 12: fewer phases and HBM passes per instance (256 mixed
// instances: 90 ms at 12 / 11, 66 ms at 13 / 12), a whole-state workgroup up to
// 12 bits, and 5 contiguous low bits unless tile_low_bits was given (64.0 ms
// against 65.8 at 4).
TileTuning vqc_tile_tuning(int nbits) {
    const vqc_tuning_t g = vqc_tuning_now();
    int fwd = kMaxTileBits, rev = kMaxTileBits2, small = 12;
    if (nbits >= 0) {
        small = 9;
        fwd = nbits <= 14 ? 9 : (nbits <= 18 ? 10 : (nbits <= 22 ? 11 : 12));
        rev = std::min(fwd, 11);
    }
    // the autotuner's entry for this size (vqc_set_tile_table) replaces the rule
    // above; explicit process-wide fields (vqc_set_tuning) replace both
    const vqc_tile_entry_t te = vqc_tile_entry(nbits);
    if (te.fwd) fwd = te.fwd;
    if (te.rev) rev = te.rev;
    TileTuning t;
    t.t_fwd = std::min(kMaxTileBits, std::max(6, g.tile_fwd_bits ? int(g.tile_fwd_bits) : fwd));
    t.t_rev = std::min(kMaxTileBits2, std::max(6, g.tile_rev_bits ? int(g.tile_rev_bits) : rev));
    t.t_small = std::min(kMaxTileBits2, std::max(1, g.tile_small_bits ? int(g.tile_small_bits) : small));
    t.mlow = std::min(8, std::max(0, int(g.tile_low_bits)));
    if (nbits < 0 && !vqc_tuning_low_bits_given()) t.mlow = 5;
    // states of 21 qubits and more: 3 contiguous low bits (HEA energy + gradient, one
    // box: 24 qubits 75.8 -> 73.4 ms, 26 qubits 375.7 -> 356.9 ms)
    if (nbits >= 21 && !vqc_tuning_low_bits_given()) t.mlow = 3;
    if (te.low && !vqc_tuning_low_bits_given()) t.mlow = te.low;
    return t;
}

// ---------------------------------------------------------------------------
// planning
// ---------------------------------------------------------------------------
void vqc_make_phase(PhaseDesc& p, int kind, int nbits, int t, const std::vector<int>& req) {
    memset(&p, 0, sizeof p);
    t = std::min(t, nbits);
    std::vector<char> in(nbits, 0);
    int cnt = 0;
    for (int b : req)
        if (!in[b]) {
            in[b] = 1;
            ++cnt;
        }
    for (int b = 0; b < nbits && cnt < t; ++b)
        if (!in[b]) {
            in[b] = 1;
            ++cnt;
        }
    p.kind = kind;
    p.nbits = nbits;
    p.t = cnt;
    int m = 0;
    while (m < nbits && in[m]) ++m;
    p.m = m;
    p.nhb = 0;
    for (int b = m; b < nbits; ++b)
        if (in[b]) p.hb[p.nhb++] = b;
    p.nob = 0;
    for (int b = 0; b < nbits; ++b)
        if (!in[b]) p.ob[p.nob++] = b;
}

static int local_pos(const PhaseDesc& p, int b) {
    if (b < p.m) return b;
    for (int i = 0; i < p.nhb; ++i)
        if (p.hb[i] == b) return p.m + i;
    return -1;
}

static std::vector<int> op_bits(const vqc_op_t& o, int n, int is_dm) {
    std::vector<int> b(o.pos, o.pos + o.nq);
    if (is_dm)
        for (int j = 0; j < o.nq; ++j) b.push_back(o.pos[j] + n);
    return b;
}

// can `bits` join `cur` in a tile of t bits that keeps mlow contiguous low bits?
static bool fits(const std::vector<int>& cur, const std::vector<int>& bits, int nbits, int t, int mlow) {
    if (nbits <= t) return true;
    std::vector<int> u(cur);
    for (int b : bits)
        if (std::find(u.begin(), u.end(), b) == u.end()) u.push_back(b);
    int high = 0;
    for (int b : u)
        if (b >= mlow) ++high;
    return high <= t - std::min(mlow, t);
}

static void set_lpos(MicroOp& m, const PhaseDesc& p, const int* gbits, int kb) {
    m.kb = kb;
    for (int j = 0; j < 5; ++j) m.lpos[j] = 0;
    for (int j = 0; j < kb; ++j) m.lpos[j] = local_pos(p, gbits[j]);
}

static MicroOp mop_init(int kind) {
    MicroOp m;
    memset(&m, 0, sizeof m);
    m.kind = kind;
    m.mat2 = -1;
    for (int q = 0; q < 4; ++q) {
        m.gen[q] = -1;
        m.acc[q] = 0;
        m.cmat[q] = -1;
    }
    return m;
}

// Structure (MS_*) of a D x D pool matrix: from its values when the pool is on
// the host, else from the op's hint bits (vqc_op_t::reserved): bit 0 U
// diagonal, bit 1 generators diagonal, bit 2 / 3 U is a CNOT whose control is
// the lower / higher position, bit 4 / 5 U real / X-like, bit 6 / 7 generators
// real / X-like (X-like: real diagonal, imaginary off-diagonal). Values decide
// whenever they are known, so a hint can only matter for batches, whose
// matrices arrive at launch. Every class is closed under the conjugate,
// transpose and adjoint variants the planner derives.
static int host_mstruct(const double* M, int D) {
    auto z = [&](int r, int c) { return M[2 * (r + c * D)] == 0.0 && M[2 * (r + c * D) + 1] == 0.0; };
    auto one = [&](int r, int c) { return M[2 * (r + c * D)] == 1.0 && M[2 * (r + c * D) + 1] == 0.0; };
    auto re0 = [&](int r, int c) { return M[2 * (r + c * D)] == 0.0; };
    auto im0 = [&](int r, int c) { return M[2 * (r + c * D) + 1] == 0.0; };
    if (D == 2) {
        if (z(1, 0) && z(0, 1)) return MS_DIAG;
        if (im0(0, 0) && im0(1, 0) && im0(0, 1) && im0(1, 1)) return MS_REAL;
        if (im0(0, 0) && im0(1, 1) && re0(1, 0) && re0(0, 1)) return MS_XLIKE;
        return MS_DENSE;
    }
    if (D != 4) return MS_DENSE;
    bool lo = true, hi = true;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            const int plo = c == 1 ? 3 : (c == 3 ? 1 : c), phi = c == 2 ? 3 : (c == 3 ? 2 : c);
            lo = lo && (r == plo ? one(r, c) : z(r, c));
            hi = hi && (r == phi ? one(r, c) : z(r, c));
        }
    return lo ? MS_CX_LO : (hi ? MS_CX_HI : MS_DENSE);
}

static int mstruct_of(XPool& xp, int64_t off, int D, int hint_ms) {
    if (off < 0) return MS_DENSE;
    if (!xp.deferred) return xp.classify(off, D);
    return hint_ms;
}
static int hint_mat(const vqc_op_t& o) {
    if (o.nq == 1 && (o.reserved & 1)) return MS_DIAG;
    if (o.nq == 1 && (o.reserved & 16)) return MS_REAL;
    if (o.nq == 1 && (o.reserved & 32)) return MS_XLIKE;
    if (o.nq == 2 && (o.reserved & 4)) return MS_CX_LO;
    if (o.nq == 2 && (o.reserved & 8)) return MS_CX_HI;
    return MS_DENSE;
}
static int hint_gen(const vqc_op_t& o) {
    if (o.nq != 1) return MS_DENSE;
    if (o.reserved & 2) return MS_DIAG;
    if (o.reserved & 64) return MS_REAL;
    if (o.reserved & 128) return MS_XLIKE;
    return MS_DENSE;
}

// Group items into phases; emit(phase, item) appends micro-ops.
//
// Lookahead: scanning the unplaced items in order, an item joins the current
// phase when its bits fit the tile and none of its bits belongs to an item
// skipped earlier in this scan. Such an item acts on qubits disjoint from every
// item it overtakes, so it commutes with them exactly, and so do the generators
// of a gradient (U_j G_k U_j^dagger = G_k for disjoint supports): states,
// energies and gradients are unchanged. all_commute (Hamiltonian terms, which
// are summed) drops the dependence test.
constexpr int kLookahead = 128;

// One phase's scan with the high tile bits fixed to the set S (bits below mlow
// are always in the tile): an item is taken when its high bits are in S, no bit
// of it belongs to an item skipped before, and the accumulators allow it. Items
// and sets are bit masks (at most 64 index bits), so a scan is a few operations
// per item: the coverage choice below runs thousands of them per phase.
static size_t scan_fixed(const std::vector<int>& rest, const std::vector<uint64_t>& mask,
                         const std::vector<std::vector<int>>& bits, int nbits, int mlow, int a, uint64_t S,
                         std::vector<int>* take, std::vector<int>* keep, size_t* idx_out, std::vector<int>* cur) {
    const uint64_t all = nbits >= 64 ? ~uint64_t(0) : (uint64_t(1) << nbits) - 1;
    const uint64_t hi = all & ~((uint64_t(1) << mlow) - 1);
    uint64_t blocked = 0;
    int acc = 0;
    size_t cnt = 0, idx = 0;
    for (; idx < rest.size() && int(idx) < kLookahead && blocked != all; ++idx) {
        const int j = rest[idx];
        const uint64_t m = mask[j];
        if (acc + a <= kMaxAcc && !(m & blocked) && !(m & hi & ~S)) {
            ++cnt;
            if (take) take->push_back(j);
            if (cur)
                for (int x : bits[j])
                    if (std::find(cur->begin(), cur->end(), x) == cur->end()) cur->push_back(x);
            acc += a;
            if (a > 0 && acc + a > kMaxAcc) {
                ++idx;
                break;
            }
        } else {
            if (keep) keep->push_back(j);
            blocked |= m;
        }
    }
    if (idx_out) *idx_out = idx;
    return cnt;
}

// Coverage choice of a phase's high tile bits: starting from none, repeatedly add
// the move (the missing high bits of one of the next 64 items) that makes the
// most items of the window executable per added bit, while the tile has room.
static size_t cover_scan(const std::vector<int>& rest, const std::vector<uint64_t>& mask,
                         const std::vector<std::vector<int>>& bits, int nbits, int t, int mlow, int a,
                         std::vector<int>& take, std::vector<int>& keep, size_t& idx, std::vector<int>& cur) {
    const int budget = t - std::min(mlow, t);
    const uint64_t all = nbits >= 64 ? ~uint64_t(0) : (uint64_t(1) << nbits) - 1;
    const uint64_t hi = all & ~((uint64_t(1) << mlow) - 1);
    uint64_t S = 0;
    int used = 0;
    size_t base = scan_fixed(rest, mask, bits, nbits, mlow, a, S, nullptr, nullptr, nullptr, nullptr);
    const size_t W = std::min<size_t>(rest.size(), 64);
    std::vector<uint64_t> seen;
    for (;;) {
        uint64_t best = 0;
        double best_score = 0;
        size_t best_cnt = base;
        seen.clear();
        for (size_t w = 0; w < W; ++w) {
            const uint64_t mv = mask[rest[w]] & hi & ~S;
            const int nmv = __builtin_popcountll(mv);
            if (!mv || used + nmv > budget) continue;
            if (std::find(seen.begin(), seen.end(), mv) != seen.end()) continue;
            seen.push_back(mv);
            const size_t c = scan_fixed(rest, mask, bits, nbits, mlow, a, S | mv, nullptr, nullptr, nullptr, nullptr);
            const double score = double(c - base) / double(nmv);
            if (c > base && score > best_score) {
                best = mv;
                best_score = score;
                best_cnt = c;
            }
        }
        if (!best) break;
        S |= best;
        used += __builtin_popcountll(best);
        base = best_cnt;
    }
    return scan_fixed(rest, mask, bits, nbits, mlow, a, S, &take, &keep, &idx, &cur);
}

template <class BitsFn, class EmitFn>
static void plan_groups(PlanBuilder& pb, int kind, int nbits, int nitems, int t, int mlow, int max_acc_per_item,
                        BitsFn bits_of, EmitFn emit, bool all_commute = false) {
    std::vector<std::vector<int>> bits(nitems);
    std::vector<uint64_t> mask(nitems, 0);
    for (int j = 0; j < nitems; ++j) {
        bits[j] = bits_of(j);
        for (int x : bits[j]) mask[j] |= uint64_t(1) << x;
    }
    const int a = std::max(max_acc_per_item, 0);
    std::vector<int> rest(nitems);
    for (int j = 0; j < nitems; ++j) rest[j] = j;
    std::vector<char> blocked(size_t(std::max(nbits, 1)), 0);
    while (!rest.empty()) {
        std::vector<int> cur, take, keep;
        std::fill(blocked.begin(), blocked.end(), 0);
        int nblocked = 0, acc = 0;
        size_t idx = 0;
        for (; idx < rest.size() && int(idx) < kLookahead && nblocked < nbits; ++idx) {
            const int j = rest[idx];
            const std::vector<int>& b = bits[j];
            bool ok = true;
            if (!all_commute)
                for (int x : b)
                    if (blocked[x]) ok = false;
            if (ok && !take.empty()) ok = fits(cur, b, nbits, t, mlow) && acc + a <= kMaxAcc;
            if (ok) {
                take.push_back(j);
                for (int x : b)
                    if (std::find(cur.begin(), cur.end(), x) == cur.end()) cur.push_back(x);
                acc += a;
                if (a > 0 && acc + a > kMaxAcc) {
                    ++idx;
                    break;
                }
            } else {
                keep.push_back(j);
                if (!all_commute)
                    for (int x : b)
                        if (!blocked[x]) {
                            blocked[x] = 1;
                            ++nblocked;
                        }
            }
        }
        if (!all_commute && nbits > t && !getenv("VQC_NO_COVER")) {
            // Coverage choice of the tile bits: grow the set of high bits greedily by
            // the item moves (an item's missing high bits) that make the most items of
            // the window executable per added bit, under the same dependence rule.
            // It replaces the in-order first fit when it runs more items.
            std::vector<int> take2, keep2, cur2;
            size_t idx2 = 0;
            if (cover_scan(rest, mask, bits, nbits, t, mlow, a, take2, keep2, idx2, cur2) > take.size()) {
                take.swap(take2);
                keep.swap(keep2);
                cur.swap(cur2);
                idx = idx2;
            }
        }
        keep.insert(keep.end(), rest.begin() + idx, rest.end());
        rest.swap(keep);
        PhaseDesc p;
        vqc_make_phase(p, kind, nbits, t, cur);
        p.op_begin = int(pb.mops.size());
        p.row_begin = pb.nrows;
        p.nacc = 0;
        for (int q : take) emit(p, q);
        p.op_end = int(pb.mops.size());
        pb.phases.push_back(p);
    }
}

// R <- the variant v of the D x D matrix M (R and M do not overlap)
static void make_variant(const double* M, int D, int v, double* R) {
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) {
            const size_t dst = 2 * size_t(r + c * D);
            const size_t src = (v & MF_TRANS) ? 2 * size_t(c + r * D) : dst;
            double re = M[src], im = M[src + 1];
            if (v == 4) {  // M + M^dagger
                const size_t tr = 2 * size_t(c + r * D);
                re = M[dst] + M[tr];
                im = M[dst + 1] - M[tr + 1];
            } else if (v & MF_CONJ) {
                im = -im;
            }
            R[dst] = re;
            R[dst + 1] = im;
        }
}
// out <- U_n ... U_1 for the factors at h + 2 srcs[f], in application order
static void make_product(const double* h, const int64_t* srcs, int n, int D, double* out) {
    std::vector<double> P(size_t(2) * D * D, 0.0), T(size_t(2) * D * D);
    for (int i = 0; i < D; ++i) P[2 * (i + i * D)] = 1.0;
    for (int f = 0; f < n; ++f) {  // P <- U_f P
        const double* U = h + 2 * srcs[f];
        for (int r = 0; r < D; ++r)
            for (int c = 0; c < D; ++c) {
                double re = 0, im = 0;
                for (int k = 0; k < D; ++k) {
                    const double ur = U[2 * (r + k * D)], ui = U[2 * (r + k * D) + 1];
                    const double pr = P[2 * (k + c * D)], pi = P[2 * (k + c * D) + 1];
                    re += ur * pr - ui * pi;
                    im += ur * pi + ui * pr;
                }
                T[2 * (r + c * D)] = re;
                T[2 * (r + c * D) + 1] = im;
            }
        P.swap(T);
    }
    std::copy(P.begin(), P.end(), out);
}

int64_t XPool::variant(int64_t off, int D, int v) {
    if (v == 0) return off;
    const uint64_t key = (uint64_t(off) << 4) | uint64_t(v);
    const auto hit = cache.find(key);
    if (hit != cache.end()) return hit->second;
    const int64_t out = size();
    cache.emplace(key, out);
    table.push_back({off, out, D, v});
    if (deferred) {
        len += int64_t(D) * D;
        return out;
    }
    h.resize(h.size() + size_t(2) * D * D);
    make_variant(h.data() + 2 * off, D, v, h.data() + 2 * out);
    return out;
}

int64_t XPool::product(const int64_t* srcs, int n, int D) {
    const int64_t out = size();
    XPoolEntry e{};
    e.src = srcs[0];
    e.dst = out;
    e.D = D;
    e.v = 8;
    e.nsrc = n;
    for (int i = 0; i < 4; ++i) e.srcs[i] = i < n ? srcs[i] : -1;
    table.push_back(e);
    if (deferred) {
        len += int64_t(D) * D;
        return out;
    }
    h.resize(h.size() + size_t(2) * D * D);
    make_product(h.data(), srcs, n, D, h.data() + 2 * out);
    return out;
}

// out <- the product of b's factors (application order), each embedded in the
// 2^K block space
static void make_block_product(const double* h, const XPoolBlock& b, double* out) {
    const int D = 1 << b.K;
    std::vector<double> P(size_t(2) * D * D, 0.0), col(size_t(2) * D);
    for (int i = 0; i < D; ++i) P[2 * (i + i * D)] = 1.0;
    for (size_t f = 0; f < b.mats.size(); ++f) {
        const int k = b.nb[f], d = 1 << k;
        const double* U = h + 2 * b.mats[f];
        const int* mp = b.map.data() + 5 * f;
        int fmask = 0;
        for (int j = 0; j < k; ++j) fmask |= 1 << mp[j];
        for (int c = 0; c < D; ++c) {  // column c of P <- U_f (column c of P)
            double* pc = P.data() + 2 * size_t(c) * D;
            std::copy(pc, pc + 2 * D, col.begin());
            for (int r = 0; r < D; ++r) {
                if (r & fmask) continue;  // r: the group base (factor bits zero)
                double in[64], o[64];
                for (int a = 0; a < d; ++a) {
                    int idx = r;
                    for (int j = 0; j < k; ++j)
                        if (a >> j & 1) idx |= 1 << mp[j];
                    in[2 * a] = col[2 * idx];
                    in[2 * a + 1] = col[2 * idx + 1];
                }
                for (int a = 0; a < d; ++a) {
                    double re = 0, im = 0;
                    for (int e = 0; e < d; ++e) {
                        const double ur = U[2 * (a + e * d)], ui = U[2 * (a + e * d) + 1];
                        re += ur * in[2 * e] - ui * in[2 * e + 1];
                        im += ur * in[2 * e + 1] + ui * in[2 * e];
                    }
                    o[2 * a] = re;
                    o[2 * a + 1] = im;
                }
                for (int a = 0; a < d; ++a) {
                    int idx = r;
                    for (int j = 0; j < k; ++j)
                        if (a >> j & 1) idx |= 1 << mp[j];
                    pc[2 * idx] = o[2 * a];
                    pc[2 * idx + 1] = o[2 * a + 1];
                }
            }
        }
    }
    std::copy(P.begin(), P.end(), out);
}

int64_t XPool::dense_block(XPoolBlock b) {
    const int64_t D = int64_t(1) << b.K;
    b.dst = size();
    h.resize(h.size() + size_t(2) * D * D);
    make_block_product(h.data(), b, h.data() + 2 * b.dst);
    blocks.push_back(std::move(b));
    return blocks.back().dst;
}

int XPool::classify(int64_t off, int D) {
    const int ms = host_mstruct(h.data() + 2 * off, D);
    claims.push_back({off, D, ms, 0, 0});
    return ms;
}

// entries (r, c) of a 4 x 4 with parity(r) != parity(c) are all exactly zero
static bool host_parity4(const double* M) {
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            const bool same = ((r ^ (r >> 1)) & 1) == ((c ^ (c >> 1)) & 1);
            if (!same && (M[2 * (r + 4 * c)] != 0.0 || M[2 * (r + 4 * c) + 1] != 0.0)) return false;
        }
    return true;
}

bool XPool::parity_blocks(int64_t off) {
    const bool p = host_parity4(h.data() + 2 * off);
    claims.push_back({off, 4, p ? 1 : 0, 1, 0});
    return p;
}

// every off-diagonal entry of the D x D matrix is exactly zero
static bool host_diag(const double* M, int D) {
    for (int c = 0; c < D; ++c)
        for (int r = 0; r < D; ++r)
            if (r != c && (M[2 * (r + c * D)] != 0.0 || M[2 * (r + c * D) + 1] != 0.0)) return false;
    return true;
}

bool XPool::diagonal(int64_t off, int D) {
    const bool d = host_diag(h.data() + 2 * off, D);
    claims.push_back({off, D, d ? 1 : 0, 2, 0});
    return d;
}

bool XPool::reload(const double* pool, int64_t plen, const double* hpool, int64_t hlen) {
    if (plen) std::copy(pool, pool + 2 * plen, h.begin());
    if (hlen) std::copy(hpool, hpool + 2 * hlen, h.begin() + 2 * plen);
    for (const XPoolEntry& e : table) {
        if (e.v == 8) make_product(h.data(), e.srcs, e.nsrc, e.D, h.data() + 2 * e.dst);
        else make_variant(h.data() + 2 * e.src, e.D, e.v, h.data() + 2 * e.dst);
    }
    for (const XPoolBlock& b : blocks) make_block_product(h.data(), b, h.data() + 2 * b.dst);
    for (const XPoolClaim& c : claims) {
        const int now = c.test == 1   ? int(host_parity4(h.data() + 2 * c.off))
                        : c.test == 2 ? int(host_diag(h.data() + 2 * c.off, c.D))
                                      : host_mstruct(h.data() + 2 * c.off, c.D);
        if (now != c.ms) return false;
    }
    return true;
}

// Register blocks on a pair of bits (MK_BLOCK). Within one phase's micro-ops,
// a one-bit op C on bit a moves to a block position q when no op between them
// touches a, so C commutes with everything it is moved across:
//   * q holds an op or block whose bits include a (reverse: the last op before C
//     touching a; forward: the next op after C touching a, else the last before);
//   * or, moving backwards, q is the nearest one-bit op or one-bit block on
//     another bit with no op touching a in between; the block gets both bits
//     (e.g. the row and column uncompute of a density-matrix gate).
// Reverse one-bit ops are uncomputes with at most one gradient (MK_APPLY2,
// MK_CHAIN) and gradient terms (MK_GRAD, a step without a matrix). A block
// holds at most 4 steps.
namespace {
struct BStep {
    int tgt;  // filled when the block's bits are final
    int bit;
    int64_t mat, gen;
    int32_t acc;
    int ms;  // MS_* of mat in bits 0..1, of gen in bits 2..3
};
bool touches(const MicroOp& m, int bit) {
    for (int j = 0; j < m.kb; ++j)
        if (m.lpos[j] == bit) return true;
    return false;
}
// the steps of a one-bit op (empty when it cannot be a step)
std::vector<BStep> one_bit_steps(const MicroOp& C, bool rev) {
    std::vector<BStep> cs;
    if (C.kb != 1) return cs;
    const int a = C.lpos[0];
    if (rev) {
        if (C.kind == MK_CHAIN) {
            int k = 0;
            for (int s2 = 0; s2 < C.nchain; ++s2) {
                const bool g = C.gen[s2] >= 0;
                cs.push_back({0, a, C.cmat[s2], g ? C.gen[s2] : -1, g ? C.acc[k] : 0, (C.mstr >> (4 * s2)) & 15});
                k += g;
            }
        } else if (C.kind == MK_APPLY2 && C.ng <= 1) {
            cs.push_back({0, a, C.mat, C.ng ? C.gen[0] : -1, C.ng ? C.acc[0] : 0, C.mstr & 15});
        } else if (C.kind == MK_GRAD && C.ng == 1) {
            cs.push_back({0, a, -1, C.gen[0], C.acc[0], C.mstr & 12});
        }
    } else if (C.kind == MK_APPLY) {
        cs.push_back({0, a, C.mat, -1, 0, C.mstr & 3});
    }
    return cs;
}
void form_blocks(std::vector<MicroOp>& out, size_t b, int phase_kind) {
    const size_t L = out.size() - b;
    if (L < 2 || (phase_kind != PH_FWD && phase_kind != PH_REV)) return;
    const bool rev = phase_kind == PH_REV;
    std::vector<MicroOp> seg(out.begin() + b, out.end());
    std::vector<std::vector<BStep>> steps(L);
    std::vector<std::vector<int>> bits(L);
    std::vector<char> dead(L, 0), isblk(L, 0);
    auto touches_k = [&](size_t k, int bit) {
        if (isblk[k]) return std::find(bits[k].begin(), bits[k].end(), bit) != bits[k].end();
        return touches(seg[k], bit);
    };
    // can seg[q] become (or grow into) a block holding bit a and `extra` more steps?
    auto can_host = [&](size_t q, int a, size_t extra) -> bool {
        if (dead[q]) return false;
        if (isblk[q])  // blocks always hold two bits
            return steps[q].size() + extra <= 4 && std::find(bits[q].begin(), bits[q].end(), a) != bits[q].end();
        const MicroOp& P = seg[q];
        if (P.kb == 2) {
            if (!(P.lpos[0] == a || P.lpos[1] == a)) return false;
            if (rev ? !(P.kind == MK_APPLY2 && P.ng <= 1) : P.kind != MK_APPLY) return false;
            return 1 + extra <= 4;
        }
        // a one-bit op hosts only an op on another bit (same-bit runs are chains already)
        if (P.lpos[0] == a) return false;
        const std::vector<BStep> ps = one_bit_steps(P, rev);
        return !ps.empty() && ps.size() + extra <= 4;
    };
    auto make_block = [&](size_t q) {
        if (isblk[q]) return;
        const MicroOp& P = seg[q];
        if (P.kb == 2) {
            steps[q].push_back({BS_BOTH, -1, P.mat, rev && P.ng ? P.gen[0] : -1, rev && P.ng ? P.acc[0] : 0, P.mstr & 15});
            bits[q] = {P.lpos[0], P.lpos[1]};
        } else {
            steps[q] = one_bit_steps(P, rev);
            bits[q] = {P.lpos[0]};
        }
        isblk[q] = 1;
    };
    for (size_t j = 0; j < L; ++j) {
        const MicroOp& C = seg[j];
        std::vector<BStep> cs = one_bit_steps(C, rev);
        if (cs.empty() || (!rev && seg[j].kind != MK_APPLY)) continue;
        const int a = C.lpos[0];
        size_t q = L;
        bool before = false;  // C's steps run before the host's (forward, next op)
        if (!rev) {
            for (size_t k = j + 1; k < L; ++k)
                if (!dead[k] && touches_k(k, a)) {
                    if (seg[k].buf == C.buf && (isblk[k] || seg[k].kb == 2) && can_host(k, a, cs.size())) {
                        q = k;
                        before = true;
                    }
                    break;
                }
        }
        if (q == L) {
            // backwards: the last op touching a, else the nearest one-bit host
            for (size_t k = j; k-- > 0;) {
                if (dead[k]) continue;
                if (touches_k(k, a)) {
                    if (seg[k].buf == C.buf && can_host(k, a, cs.size())) q = k;
                    break;
                }
                if (!isblk[k] && seg[k].kb == 1 && seg[k].buf == C.buf && can_host(k, a, cs.size())) {
                    q = k;
                    break;
                }
            }
        }
        if (q == L) continue;
        make_block(q);
        if (std::find(bits[q].begin(), bits[q].end(), a) == bits[q].end()) bits[q].push_back(a);
        if (before) steps[q].insert(steps[q].begin(), cs.begin(), cs.end());
        else steps[q].insert(steps[q].end(), cs.begin(), cs.end());
        dead[j] = 1;
    }
    out.resize(b);
    for (size_t k = 0; k < L; ++k) {
        if (dead[k]) continue;
        if (!isblk[k]) {
            out.push_back(seg[k]);
            continue;
        }
        const int lo = std::min(bits[k][0], bits[k][1]), hi = std::max(bits[k][0], bits[k][1]);
        MicroOp m = mop_init(MK_BLOCK);
        m.kb = 2;
        m.buf = seg[k].buf;
        m.lpos[0] = lo;
        m.lpos[1] = hi;
        m.nchain = int(steps[k].size());
        m.ng = 0;
        m.smask = 0;
        m.mstr = 0;
        for (int j = 0; j < m.nchain; ++j) {
            const BStep& st = steps[k][j];
            const int tgt = st.bit < 0 ? BS_BOTH : (st.bit == lo ? BS_LO : BS_HI);
            m.smask |= tgt << (2 * j);
            m.mstr |= st.ms << (4 * j);
            m.cmat[j] = st.mat;
            m.gen[j] = st.gen;
            if (st.gen >= 0) m.acc[m.ng++] = st.acc;
        }
        out.push_back(m);
    }
}
}  // namespace

// ---------------------------------------------------------------------------
// Dense-block fusion of forward phases (MK_DENSE). The phase's MK_APPLY
// micro-ops are grouped greedily, in order, into blocks of at most kmax bits:
// an op joins the blocks that last touched its bits when their union fits and
// every older block among them is still the last one on all of its bits (so it
// commutes with everything it is moved past); blocks keep the position of the
// newest. A block of 3 to 5 bits becomes one MK_DENSE micro-op (its product
// matrix, run on the matrix cores) when that is cheaper by a cost per amplitude
// of 8 * 2^k flops scaled for each engine plus a fixed cost per micro-op;
// otherwise its ops stay as they are. kmax = 5, 4 and 3 are tried and the
// cheapest grouping is kept. The tile must hold at least 128 groups of the
// block (t - K >= 7: 8 column tiles of 16 groups, one per wave).
// ---------------------------------------------------------------------------
namespace {
struct DenseCost {
    double av = 2.0, cv = 64.0, am = 1.25, cm = 64.0;  // VALU / MFMA scale and per-micro-op cost
};
DenseCost dense_cost() {
    DenseCost c;
    if (const char* e = getenv("VQC_DENSE_COST")) sscanf(e, "%lf,%lf,%lf,%lf", &c.av, &c.cv, &c.am, &c.cm);
    return c;
}
struct DBlock {
    std::vector<int> bits;  // local bits
    std::vector<int> ops;   // indices into the phase's op list, in order
    bool fusable = true;
};
double op_cost_valu(const MicroOp& m, const DenseCost& c) { return c.av * 8.0 * double(1 << m.kb) + c.cv; }

std::vector<DBlock> group_dense(const std::vector<MicroOp>& ph, int kmax) {
    std::vector<DBlock> bl;
    std::vector<int> last(64, -1);
    for (size_t i = 0; i < ph.size(); ++i) {
        const MicroOp& m = ph[i];
        const bool fusable = m.kind == MK_APPLY && m.buf == 0;
        std::vector<int> cands;
        for (int j = 0; j < m.kb; ++j)
            if (last[m.lpos[j]] >= 0 && std::find(cands.begin(), cands.end(), last[m.lpos[j]]) == cands.end())
                cands.push_back(last[m.lpos[j]]);
        std::sort(cands.begin(), cands.end());
        bool ok = fusable && !cands.empty();
        std::vector<int> u(m.lpos, m.lpos + m.kb);
        if (ok) {
            for (int cidx : cands) {
                if (!bl[cidx].fusable) ok = false;
                for (int b : bl[cidx].bits)
                    if (std::find(u.begin(), u.end(), b) == u.end()) u.push_back(b);
            }
            ok = ok && int(u.size()) <= kmax;
            for (size_t q = 0; ok && q + 1 < cands.size(); ++q)
                for (int b : bl[cands[q]].bits)
                    if (last[b] != cands[q]) ok = false;
        }
        if (ok) {
            const int keep = cands.back();
            std::vector<int> ops;
            for (int cidx : cands) {
                ops.insert(ops.end(), bl[cidx].ops.begin(), bl[cidx].ops.end());
                if (cidx != keep) {
                    // an older block was the last one on all its bits: they move with it
                    for (int b : bl[cidx].bits) last[b] = keep;
                    bl[cidx].ops.clear();
                    bl[cidx].bits.clear();
                }
            }
            // merged ops: the older blocks' ops commute with the newer ones'
            // (disjoint bits), so concatenation in block order is exact
            ops.push_back(int(i));
            DBlock& nb = bl[keep];
            nb.ops = ops;
            std::sort(u.begin(), u.end());
            nb.bits = u;
            // only the new op's bits: a bit of the kept block that a later block
            // touched keeps that block as its last
            for (int j = 0; j < m.kb; ++j) last[m.lpos[j]] = keep;
        } else {
            DBlock b;
            b.bits.assign(m.lpos, m.lpos + m.kb);
            std::sort(b.bits.begin(), b.bits.end());
            b.ops = {int(i)};
            b.fusable = fusable;
            bl.push_back(b);
            for (int x : b.bits) last[x] = int(bl.size()) - 1;
        }
    }
    return bl;
}

bool block_dense(const DBlock& b, const std::vector<MicroOp>& ph, int t, const DenseCost& c, double* cost) {
    const int K = int(b.bits.size());
    double split = 0;
    for (int o : b.ops) split += op_cost_valu(ph[o], c);
    const double dense = c.am * 8.0 * double(1 << K) + c.cm;
    const bool use = b.fusable && K >= 3 && K <= 5 && t - K >= 7 && b.ops.size() >= 1 && dense < split;
    *cost = use ? dense : split;
    return use;
}
}  // namespace

static std::vector<MicroOp> fuse_dense(const std::vector<MicroOp>& ph, XPool& xp, int t) {
    if (ph.size() < 1 || getenv("VQC_NO_DENSE")) return ph;
    const DenseCost c = dense_cost();
    std::vector<DBlock> best;
    double best_cost = 0;
    for (int kmax : {5, 4, 3}) {
        std::vector<DBlock> bl = group_dense(ph, kmax);
        double tot = 0;
        for (const DBlock& b : bl) {
            if (b.ops.empty()) continue;
            double cst;
            block_dense(b, ph, t, c, &cst);
            tot += cst;
        }
        if (best.empty() || tot < best_cost) {
            best = bl;
            best_cost = tot;
        }
    }
    bool any = false;
    for (const DBlock& b : best) {
        double cst;
        if (!b.ops.empty() && block_dense(b, ph, t, c, &cst)) any = true;
    }
    if (!any) return ph;  // no block pays: the phase keeps its order
    std::vector<MicroOp> out;
    for (const DBlock& b : best) {
        if (b.ops.empty()) continue;
        double cst;
        if (!block_dense(b, ph, t, c, &cst)) {
            for (int o : b.ops) out.push_back(ph[o]);
            continue;
        }
        XPoolBlock xb;
        xb.K = int(b.bits.size());
        for (int o : b.ops) {
            const MicroOp& m = ph[o];
            xb.mats.push_back(m.mat);
            xb.nb.push_back(m.kb);
            for (int j = 0; j < 5; ++j)
                xb.map.push_back(j < m.kb ? int(std::find(b.bits.begin(), b.bits.end(), m.lpos[j]) - b.bits.begin()) : 0);
        }
        MicroOp d = mop_init(MK_DENSE);
        d.kb = xb.K;
        for (int j = 0; j < 5; ++j) d.lpos[j] = j < xb.K ? b.bits[j] : 0;
        d.mat = xp.dense_block(std::move(xb));
        out.push_back(d);
    }
    return out;
}

// The pool range a phase's micro-ops read matrices from (PhaseDesc::pf_lo,
// pf_n), for the kernels to read ahead once per phase: every micro-op otherwise
// waits for its descriptor and then for its matrices, two dependent round trips
// to memory, and on small tiles those waits are most of the micro-op's time.
// Sizes are upper bounds (4^kb entries per matrix, 16 for register-block steps),
// clamped to the pool; ranges over kPrefetchMax entries are not read ahead.
void vqc_prefetch_range(const std::vector<MicroOp>& mops, int op_begin, int op_end, int64_t plen, int64_t* pf_lo, int32_t* pf_n) {
    int64_t lo = INT64_MAX, hi = 0;
    auto use = [&](int64_t off, int64_t n) {
        if (off < 0) return;
        lo = std::min(lo, off);
        hi = std::max(hi, off + n);
    };
    for (int o = op_begin; o < op_end; ++o) {
        const MicroOp& m = mops[o];
        const int64_t D2 = int64_t(1) << (2 * m.kb);
        switch (m.kind) {
            case MK_CHAIN:
            case MK_BLOCK:
                for (int q = 0; q < 4; ++q) {
                    use(m.cmat[q], 16);
                    use(m.gen[q], 16);
                }
                break;
            case MK_GRAD:
                for (int q = 0; q < m.ng; ++q) use(m.gen[q], D2);
                break;
            default:  // MK_APPLY, MK_APPLY2, MK_APPLY2X, MK_HAM, MK_DENSE
                use(m.mat, D2);
                use(m.mat2, D2);
                if (m.kind == MK_APPLY2)
                    for (int q = 0; q < m.ng; ++q) use(m.gen[q], D2);
                break;
        }
    }
    hi = std::min(hi, plen);
    *pf_lo = 0;
    *pf_n = 0;
    if (lo < hi && hi - lo <= kPrefetchMax) {
        *pf_lo = lo;
        *pf_n = int32_t(hi - lo);
    }
}
static void set_prefetch(PhaseDesc& p, const std::vector<MicroOp>& mops, int64_t plen) {
    vqc_prefetch_range(mops, p.op_begin, p.op_end, plen, &p.pf_lo, &p.pf_n);
}

void vqc_fuse_chains(PlanBuilder& pb, size_t phase_begin, size_t phase_end) {
    // Phases own contiguous, ascending micro-op ranges; the range's phases are
    // rewritten in place and later phases are shifted.
    if (phase_begin >= phase_end || phase_end > pb.phases.size()) return;
    XPool& xp = *pb.xp;
    const int first = pb.phases[phase_begin].op_begin;
    const int last = pb.phases[phase_end - 1].op_end;
    std::vector<MicroOp> out(pb.mops.begin(), pb.mops.begin() + first);
    out.reserve(pb.mops.size());
    for (size_t i = phase_begin; i < phase_end; ++i) {
        PhaseDesc& p = pb.phases[i];
        const int b = int(out.size());
        std::vector<MicroOp> ph(pb.mops.begin() + p.op_begin, pb.mops.begin() + p.op_end);
        if (p.kind == PH_FWD && pb.dense_fwd && !xp.deferred) ph = fuse_dense(ph, xp, p.t);
        const int np = int(ph.size());
        if (p.kind == PH_FWD) {
            // One-bit steps on different bits commute exactly. Within each run of
            // consecutive one-bit MK_APPLY on one buffer, order them by bit (the order
            // on any one bit kept), so that each bit's steps form one run for the
            // product below: the row and column steps of a density-matrix gate
            // alternate between two bits, and so do gates on different qubits.
            for (int o = 0; o < np;) {
                int e = o;
                while (e < np && ph[e].kind == MK_APPLY && ph[e].kb == 1 && ph[e].buf == ph[o].buf) ++e;
                if (e - o > 1)
                    std::stable_sort(ph.begin() + o, ph.begin() + e,
                                     [](const MicroOp& x, const MicroOp& y) { return x.lpos[0] < y.lpos[0]; });
                o = std::max(e, o + 1);
            }
        }
        for (int o = 0; o < np;) {
            const MicroOp& m = ph[o];
            int run = 1;
            if (p.kind == PH_FWD && m.kind == MK_APPLY && m.kb == 1) {
                // U_k ... U_1 on one bit of one buffer -> one 2x2 product
                while (o + run < np && run < 4) {
                    const MicroOp& x = ph[o + run];
                    if (x.kind != MK_APPLY || x.kb != 1 || x.lpos[0] != m.lpos[0] || x.buf != m.buf) break;
                    ++run;
                }
                if (run > 1) {
                    int64_t srcs[4];
                    for (int j = 0; j < run; ++j) srcs[j] = ph[o + j].mat;
                    MicroOp f = m;
                    f.mat = xp.product(srcs, run, 2);
                    f.mstr = xp.deferred ? MS_DENSE : xp.classify(f.mat, 2);
                    out.push_back(f);
                    o += run;
                    continue;
                }
            } else if (p.kind == PH_REV && m.kind == MK_APPLY2 && m.kb == 1 && m.ng <= 1) {
                // reverse steps on one bit stay separate (each gradient needs its
                // intermediate pair) but share one pass over the tile
                while (o + run < np && run < 4) {
                    const MicroOp& x = ph[o + run];
                    if (x.kind != MK_APPLY2 || x.kb != 1 || x.lpos[0] != m.lpos[0] || x.ng > 1) break;
                    ++run;
                }
                if (run > 1) {
                    MicroOp f = mop_init(MK_CHAIN);
                    f.kb = 1;
                    f.buf = m.buf;
                    f.lpos[0] = m.lpos[0];
                    f.nchain = run;
                    f.ng = 0;
                    f.mstr = 0;
                    for (int j = 0; j < run; ++j) {
                        const MicroOp& x = ph[o + j];
                        f.cmat[j] = x.mat;
                        f.gen[j] = x.ng ? x.gen[0] : -1;
                        if (x.ng) f.acc[f.ng++] = x.acc[0];
                        f.mstr |= (x.mstr & 15) << (4 * j);
                    }
                    out.push_back(f);
                    o += run;
                    continue;
                }
            }
            out.push_back(m);
            ++o;
        }
        if (!getenv("VQC_NO_BLOCKS")) form_blocks(out, size_t(b), p.kind);
        p.op_begin = b;
        p.op_end = int(out.size());
    }
    const int shift = int(out.size()) - last;
    out.insert(out.end(), pb.mops.begin() + last, pb.mops.end());
    for (size_t i = phase_end; i < pb.phases.size(); ++i) {
        pb.phases[i].op_begin += shift;
        pb.phases[i].op_end += shift;
    }
    pb.mops.swap(out);
    for (size_t i = phase_begin; i < phase_end; ++i) set_prefetch(pb.phases[i], pb.mops, xp.size());
}

int vqc_plan_forward(PlanBuilder& pb, int nbits, int n, int is_dm, int nops, const vqc_op_t* ops, int t, int mlow) {
    XPool& xp = *pb.xp;
    plan_groups(
        pb, PH_FWD, nbits, nops, t, mlow, 0, [&](int j) { return op_bits(ops[j], n, is_dm); },
        [&](PhaseDesc& p, int q) {
            const vqc_op_t& o = ops[q];
            if (!is_dm || o.kind == VQC_OP_CHANNEL) {
                std::vector<int> b = op_bits(o, n, o.kind == VQC_OP_CHANNEL);
                MicroOp m = mop_init(MK_APPLY);
                set_lpos(m, p, b.data(), int(b.size()));
                m.mat = o.mat;
                if (o.kind != VQC_OP_CHANNEL) m.mstr = mstruct_of(xp, m.mat, 1 << o.nq, hint_mat(o));
                else if (o.nq == 1 && !xp.deferred && xp.parity_blocks(m.mat)) m.mstr = kMsParity;
                pb.mops.push_back(m);
            } else {
                int cols[5];
                for (int j = 0; j < o.nq; ++j) cols[j] = o.pos[j] + n;
                MicroOp m = mop_init(MK_APPLY);
                set_lpos(m, p, o.pos, o.nq);
                m.mat = o.mat;
                m.mstr = mstruct_of(xp, m.mat, 1 << o.nq, hint_mat(o));
                pb.mops.push_back(m);
                m = mop_init(MK_APPLY);
                set_lpos(m, p, cols, o.nq);
                m.mat = xp.variant(o.mat, 1 << o.nq, MF_CONJ);
                m.mstr = mstruct_of(xp, m.mat, 1 << o.nq, hint_mat(o));
                pb.mops.push_back(m);
            }
        });
    return VQC_OK;
}

// Rewrite the Hamiltonian phases [first_phase, end) that plan_groups made (one
// MK_HAM and one energy row per term): in each phase, the terms of at most two
// bits become one MK_HAMS followed by their MK_HAM records, with one energy row
// for their sum; wider terms keep their own micro-op and row. Rows and micro-ops
// of the phases are renumbered in order. A term's matrix is marked diagonal
// (MS_DIAG) when its values are known and say so.
static void merge_ham_terms(PlanBuilder& pb, size_t first_phase) {
    if (first_phase >= pb.phases.size()) return;
    XPool& xp = *pb.xp;
    const int op0 = pb.phases[first_phase].op_begin, row0 = pb.phases[first_phase].row_begin;
    const std::vector<MicroOp> old(pb.mops.begin() + op0, pb.mops.end());
    const std::vector<int> old_slot(pb.row_slot.begin() + row0, pb.row_slot.end());
    pb.mops.resize(size_t(op0));
    pb.row_slot.resize(size_t(row0));
    pb.nrows = row0;
    for (size_t i = first_phase; i < pb.phases.size(); ++i) {
        PhaseDesc& p = pb.phases[i];
        std::vector<MicroOp> narrow, wide;
        std::vector<int> narrow_slot, wide_slot;
        for (int o = p.op_begin; o < p.op_end; ++o) {
            MicroOp m = old[size_t(o - op0)];
            const int slot = old_slot[size_t(p.row_begin + m.acc[0] - row0)];
            if (m.kb <= 2) {
                narrow.push_back(m);
                narrow_slot.push_back(slot);
            } else {
                wide.push_back(m);
                wide_slot.push_back(slot);
            }
        }
        if (narrow.size() < 2) {  // nothing to merge: keep the term as it is
            wide.insert(wide.begin(), narrow.begin(), narrow.end());
            wide_slot.insert(wide_slot.begin(), narrow_slot.begin(), narrow_slot.end());
            narrow.clear();
        }
        p.op_begin = int(pb.mops.size());
        p.row_begin = pb.nrows;
        p.nacc = 0;
        if (!narrow.empty()) {
            MicroOp h = mop_init(MK_HAMS);
            h.nchain = int(narrow.size());
            h.kb = 1;
            for (const MicroOp& m : narrow) h.kb = std::max(h.kb, m.kb);
            for (int j = 0; j < h.kb; ++j) h.lpos[j] = j;
            h.mat = -1;
            h.ng = 1;
            h.acc[0] = p.nacc++;
            pb.row_slot.push_back(narrow_slot[0]);
            pb.nrows++;
            pb.mops.push_back(h);
            for (MicroOp m : narrow) {
                m.ng = 0;
                m.acc[0] = 0;
                m.mstr = !xp.deferred && xp.diagonal(m.mat, 1 << m.kb) ? MS_DIAG : MS_DENSE;
                pb.mops.push_back(m);
            }
        }
        for (size_t k = 0; k < wide.size(); ++k) {
            MicroOp m = wide[k];
            m.acc[0] = p.nacc++;
            pb.row_slot.push_back(wide_slot[k]);
            pb.nrows++;
            pb.mops.push_back(m);
        }
        p.op_end = int(pb.mops.size());
    }
}

int vqc_plan_ham(PlanBuilder& pb, int nbits, int nterms, const vqc_term_t* terms, int64_t hoff, int t, int mlow) {
    XPool& xp = *pb.xp;
    const size_t first_phase = pb.phases.size();
    plan_groups(
        pb, PH_HAM, nbits, nterms, t, mlow, 1, [&](int j) { return std::vector<int>(terms[j].pos, terms[j].pos + terms[j].nq); },
        [&](PhaseDesc& p, int q) {
            MicroOp m = mop_init(MK_HAM);
            set_lpos(m, p, terms[q].pos, terms[q].nq);
            m.mat = hoff + terms[q].mat;
            m.mat2 = xp.variant(m.mat, 1 << terms[q].nq, 4);
            m.ng = 1;
            m.acc[0] = p.nacc++;
            pb.row_slot.push_back(-1 - q);
            pb.nrows++;
            pb.mops.push_back(m);
        },
        true);
    for (size_t i = first_phase; i < pb.phases.size(); ++i) pb.phases[i].ham_accum = i > first_phase;
    // batches plan before the values exist, so every term would be dense: there the
    // merged pass measured 1% slower than a micro-op per term (config 2, same box)
    if (!getenv("VQC_NO_HAMS") && !xp.deferred) merge_ham_terms(pb, first_phase);
    return VQC_OK;
}

int vqc_plan_reverse(PlanBuilder& pb, int nbits, int n, int is_dm, int nops, const vqc_op_t* ops, int t, int mlow) {
    XPool& xp = *pb.xp;
    auto nact = [&](const vqc_op_t& o) {
        int a = 0;
        for (int q = 0; q < o.nparams; ++q) a += o.slot[q] >= 0;
        return a;
    };
    int maxa = 0;
    for (int i = 0; i < nops; ++i) maxa = std::max(maxa, nact(ops[i]));
    // items are the ops in reverse order
    plan_groups(
        pb, PH_REV, nbits, nops, t, mlow, std::max(maxa, 1), [&](int j) { return op_bits(ops[nops - 1 - j], n, is_dm); },
        [&](PhaseDesc& p, int q) {
            const vqc_op_t& o = ops[nops - 1 - q];
            const int D = 1 << o.nq;
            if (is_dm && o.kind == VQC_OP_CHANNEL) {
                std::vector<int> b = op_bits(o, n, 1);
                MicroOp m = mop_init(MK_APPLY2X);
                set_lpos(m, p, b.data(), int(b.size()));
                m.mat = o.aux;                                           // state: S^-1
                m.mat2 = xp.variant(o.mat, D * D, MF_CONJ | MF_TRANS);   // adjoint: S^dagger
                if (o.nq == 1 && !xp.deferred && xp.parity_blocks(m.mat) && xp.parity_blocks(m.mat2)) m.mstr = kMsParity;
                pb.mops.push_back(m);
                return;
            }
            int accs[3], na = 0;
            int64_t gens[3];
            for (int j = 0; j < o.nparams; ++j) {
                if (o.slot[j] < 0) continue;
                accs[na] = p.nacc++;
                gens[na] = o.gen[j];
                pb.row_slot.push_back(o.slot[j]);
                pb.nrows++;
                ++na;
            }
            auto uncompute = [&](const int* bits, int64_t mat, bool with_grads) {
                if (o.nq <= 4) {
                    MicroOp m = mop_init(MK_APPLY2);
                    set_lpos(m, p, bits, o.nq);
                    m.mat = m.mat2 = mat;
                    m.mstr = mstruct_of(xp, mat, D, hint_mat(o));
                    if (with_grads) {
                        m.ng = na;
                        for (int j = 0; j < na; ++j) {
                            m.gen[j] = gens[j];
                            m.acc[j] = accs[j];
                        }
                        if (na == 1) m.mstr |= mstruct_of(xp, gens[0], D, hint_gen(o)) << 2;
                    }
                    pb.mops.push_back(m);
                    return;
                }
                for (int buf = 0; buf < 2; ++buf) {
                    MicroOp m = mop_init(MK_APPLY);
                    set_lpos(m, p, bits, o.nq);
                    m.mat = mat;
                    m.buf = buf;
                    pb.mops.push_back(m);
                }
                if (with_grads)
                    for (int j = 0; j < na; ++j) {
                        MicroOp g = mop_init(MK_GRAD);
                        set_lpos(g, p, bits, o.nq);
                        g.ng = 1;
                        g.gen[0] = gens[j];
                        g.acc[0] = accs[j];
                        pb.mops.push_back(g);
                    }
            };
            const int64_t udag = xp.variant(o.mat, D, MF_CONJ | MF_TRANS);
            if (!is_dm) {
                uncompute(o.pos, udag, true);
                return;
            }
            int cols[5];
            for (int j = 0; j < o.nq; ++j) cols[j] = o.pos[j] + n;
            if (o.nq == 1 && na == 1) {
                // A one-qubit, one-parameter gate: each side's gradient term right after
                // that side's uncompute, in the same micro-op. The other side's uncompute
                // acts on another bit and is unitary, so it leaves the term unchanged:
                // <C D| G_row C y> = <D| G_row y> for C on the column bit.
                auto side = [&](const int* bits, int64_t mat, int64_t gen) {
                    MicroOp m = mop_init(MK_APPLY2);
                    set_lpos(m, p, bits, 1);
                    m.mat = m.mat2 = mat;
                    m.mstr = mstruct_of(xp, mat, D, hint_mat(o)) | (mstruct_of(xp, gen, D, hint_gen(o)) << 2);
                    m.ng = 1;
                    m.gen[0] = gen;
                    m.acc[0] = accs[0];
                    pb.mops.push_back(m);
                };
                side(o.pos, udag, gens[0]);                                                  // rows
                side(cols, xp.variant(o.mat, D, MF_TRANS), xp.variant(gens[0], D, MF_CONJ));  // columns
                return;
            }
            uncompute(o.pos, udag, false);                             // U^dagger on rows
            uncompute(cols, xp.variant(o.mat, D, MF_TRANS), false);    // conj(U^dagger) = U^T on columns
            for (int j = 0; j < na; ++j)                               // <D|G_row y> + <D|conj(G)_col y>
                for (int side = 0; side < 2; ++side) {
                    MicroOp g = mop_init(MK_GRAD);
                    set_lpos(g, p, side ? cols : o.pos, o.nq);
                    g.ng = 1;
                    g.gen[0] = side ? xp.variant(gens[j], D, MF_CONJ) : gens[j];
                    g.acc[0] = accs[j];
                    // the generator's structure (its conjugate keeps it), for one-bit
                    // terms that become gradient-only steps of a register block
                    if (o.nq == 1) g.mstr = mstruct_of(xp, g.gen[0], D, hint_gen(o)) << 2;
                    pb.mops.push_back(g);
                }
        });
    return VQC_OK;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
struct PhaseBufs {
    const double2* srcA;
    double2* A;
    double2* B;
    const MicroOp* ops;
    const double2* pool;
    double2* parts;   // [row][stride]
    int stride;
};

template <int MAXKB, int KS>
__global__ void __launch_bounds__(kTileThreads, MAXKB <= 3 ? 4 : 1) k_phase(const PhaseDesc* __restrict__ pd, PhaseBufs b) {
    extern __shared__ double2 smem[];
    const PhaseDesc& p = *pd;
    const uint32_t T = 1u << p.t;
    double2* sa = smem;
    double2* sb = p.kind == PH_FWD ? smem : smem + T;
    double2* red = p.kind == PH_FWD ? smem + T : smem + 2 * T;
    double2* gacc = red + kRedDoubles2;
    warm_ops(b.ops, p.op_begin, p.op_end, b.pool, p.pf_lo, p.pf_n, red);
    for (int i = threadIdx.x; i < p.nacc; i += kTileThreads) gacc[i] = make_double2(0.0, 0.0);
    __syncthreads();
    AccState st;
    acc_init(st);
    const uint64_t ntiles = uint64_t(1) << (p.nbits - p.t);
    const uint64_t tdep = thread_deposit(p);
    // tiles in XCD-contiguous order: consecutive tiles hold neighbouring runs of
    // amplitudes, so each XCD streams one region
    for (uint64_t tile = xcd_block(); tile < ntiles; tile += gridDim.x)
        phase_tile<MAXKB, KS>(p, tile, tdep, b.srcA, b.A, b.B, b.ops, b.pool, sa, sb, red, gacc, st);
    acc_finish(st, red, gacc);
    for (int e = threadIdx.x; e < p.nacc; e += kTileThreads) b.parts[size_t(p.row_begin + e) * b.stride + blockIdx.x] = gacc[e];
}

// Whole evaluation of a state that fits one tile: forward, Hamiltonian, reverse.
// Accumulator index of every micro-op is its global partial row.
struct SmallArgs {
    int nbits;
    int nops;           // micro-ops
    int ham_begin;      // first MK_HAM micro-op
    int rev_begin;      // first reverse micro-op
    int nrows;
    const double2* src; // psi0
    double2* A;         // final state (may be null)
    const MicroOp* ops;
    const double2* pool;
    double2* rows;      // [nrows] row sums (one column)
    int64_t pf_lo;      // pool range every micro-op's matrices lie in (pf_n = 0: not read ahead)
    int32_t pf_n;
};

template <int MAXKB, int RKS>
__global__ void __launch_bounds__(kTileThreads) k_small_eval(SmallArgs a) {
    extern __shared__ double2 smem[];
    const uint32_t T = 1u << a.nbits;
    warm_ops(a.ops, 0, a.nops, a.pool, a.pf_lo, a.pf_n, smem + 2 * T);
    small_eval_body<MAXKB, RKS>(a.nbits, a.src, a.A, a.ops, 0, a.ham_begin, a.rev_begin, a.nops, a.pool, smem, smem + T, smem + 2 * T,
                    smem + 2 * T + kRedDoubles2, a.nrows, a.rows, 1);
}

// ---------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------
// Work buffers of the fused paths: one set per host thread (each thread runs its
// calls on its own stream), freed when the thread ends.
struct Work {
    void* buf = nullptr;
    size_t cap = 0;
    Work() = default;
    Work(const Work&) = delete;
    Work& operator=(const Work&) = delete;
    void release() {
        if (buf) (void)hipFree(buf);
        buf = nullptr;
        cap = 0;
    }
    ~Work() { release(); }
};
static thread_local Work g_wA, g_wB, g_wparts, g_wpool, g_wmops, g_wrows, g_wphases;

// Pinned host staging for the per-call matrix upload: a copy from pageable
// memory makes the host wait for the transfer before it can launch the first
// phase; from pinned memory it is queued like a kernel. The event marks the end
// of the last transfer out of the staging, which the next upload waits for
// before overwriting it (a call that returns without waiting for its stream may
// leave it in flight).
struct UploadStage {
    double* buf = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
    bool pending = false;
    UploadStage() = default;
    UploadStage(const UploadStage&) = delete;
    UploadStage& operator=(const UploadStage&) = delete;
    void release() {
        if (pending) (void)hipEventSynchronize(done);
        pending = false;
        if (done) (void)hipEventDestroy(done);
        done = nullptr;
        if (buf) (void)hipHostFree(buf);
        buf = nullptr;
        cap = 0;
    }
    ~UploadStage() { release(); }
};
static thread_local UploadStage g_upload;

static int upload_staged(void* dev, const double* host, size_t bytes, hipStream_t st) {
    UploadStage& u = g_upload;
    if (u.pending) {
        VQC_HIP(hipEventSynchronize(u.done));
        u.pending = false;
    }
    if (bytes > u.cap) {
        if (u.buf) VQC_HIP(hipHostFree(u.buf));
        u.buf = nullptr;
        u.cap = 0;
        const size_t nb = std::max<size_t>(bytes, size_t(1) << 16);
        VQC_HIP(hipHostMalloc((void**)&u.buf, nb, hipHostMallocDefault));
        u.cap = nb;
    }
    if (!u.done) VQC_HIP(hipEventCreateWithFlags(&u.done, hipEventDisableTiming));
    memcpy(u.buf, host, bytes);
    VQC_HIP(hipMemcpyAsync(dev, u.buf, bytes, hipMemcpyHostToDevice, st));
    VQC_HIP(hipEventRecord(u.done, st));
    u.pending = true;
    return VQC_OK;
}

static int ensure(Work& w, size_t bytes) {
    if (bytes <= w.cap) return VQC_OK;
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    if (w.buf) {
        VQC_HIP(hipStreamSynchronize(c->stream));
        VQC_HIP(hipFree(w.buf));
        w.buf = nullptr;
        w.cap = 0;
    }
    VQC_HIP(hipMalloc(&w.buf, bytes));
    w.cap = bytes;
    return VQC_OK;
}

static int check_program(int nbits_q, int is_dm, int nops, const vqc_op_t* ops, int64_t* pool_len, int nparams) {
    if (nops < 0 || (nops && !ops)) return vqc_fail(VQC_EINVAL, "bad op array");
    int64_t len = 0;
    for (int i = 0; i < nops; ++i) {
        const vqc_op_t& o = ops[i];
        if (o.kind != VQC_OP_UNITARY && o.kind != VQC_OP_CHANNEL) return vqc_fail(VQC_EINVAL, "op %d: unknown kind %d", i, o.kind);
        if (o.kind == VQC_OP_CHANNEL && !is_dm) return vqc_fail(VQC_EINVAL, "op %d: channels act on density matrices only", i);
        VQC_TRY(vqc_check_positions(nbits_q, o.nq, o.pos, o.kind == VQC_OP_CHANNEL ? 2 : 5));
        if (o.nparams < 0 || o.nparams > 3) return vqc_fail(VQC_EINVAL, "op %d: %d parameters", i, o.nparams);
        const int64_t D2 = o.kind == VQC_OP_CHANNEL ? (int64_t(1) << (4 * o.nq)) : (int64_t(1) << (2 * o.nq));
        if (o.mat < 0) return vqc_fail(VQC_EINVAL, "op %d: negative matrix offset", i);
        len = std::max(len, o.mat + D2);
        if (o.kind == VQC_OP_CHANNEL && o.aux >= 0) len = std::max(len, o.aux + D2);  // aux < 0: no inverse (checkpointed)
        for (int j = 0; j < o.nparams; ++j) {
            if (o.slot[j] >= nparams) return vqc_fail(VQC_EINVAL, "op %d: gradient slot %d >= nparams %d", i, o.slot[j], nparams);
            if (o.slot[j] >= 0) {
                if (o.gen[j] < 0) return vqc_fail(VQC_EINVAL, "op %d: active parameter without generator", i);
                len = std::max(len, o.gen[j] + D2);
            }
        }
    }
    *pool_len = len;
    return VQC_OK;
}

static int check_terms(int nq, int nterms, const vqc_term_t* terms, int64_t* len) {
    if (nterms < 0 || (nterms && !terms)) return vqc_fail(VQC_EINVAL, "bad term array");
    int64_t l = 0;
    for (int t = 0; t < nterms; ++t) {
        VQC_TRY(vqc_check_positions(nq, terms[t].nq, terms[t].pos, 5));
        if (terms[t].mat < 0) return vqc_fail(VQC_EINVAL, "term %d: negative matrix offset", t);
        l = std::max(l, terms[t].mat + (int64_t(1) << (2 * terms[t].nq)));
    }
    *len = l;
    return VQC_OK;
}

static size_t phase_lds(const PhaseDesc& p) {
    const size_t T = size_t(1) << p.t;
    return ((p.kind == PH_FWD ? T : 2 * T) + kRedDoubles2 + size_t(std::max(p.nacc, 1))) * 16;
}

// widest micro-op a phase kernel must carry code for (MK_DENSE blocks have their own)
int vqc_max_kb(const std::vector<MicroOp>& mops, int begin, int end) {
    int m = 1;
    for (int o = begin; o < end; ++o)
        if (mops[o].kind != MK_DENSE) m = std::max(m, int(mops[o].kb));
    return m;
}

#define VQC_FOR_KERNELS(X)                                                                                     \
    X(2, KS_FWD) X(2, KS_FWDH) X(2, KS_FWDD) X(2, KS_HAM) X(2, KS_REV) X(2, KS_REVX) X(3, KS_FWD) X(3, KS_FWDH)     \
    X(3, KS_FWDD) X(3, KS_HAM) X(3, KS_REV) X(3, KS_REVX) X(5, KS_FWD) X(5, KS_FWDD) X(5, KS_HAM) X(5, KS_REV)     \
    X(5, KS_REVX)

static int set_lds_attrs() {
    static bool done = false;
    if (done) return VQC_OK;
#define VQC_ATTR(M, K) VQC_HIP(hipFuncSetAttribute((const void*)k_phase<M, K>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    VQC_FOR_KERNELS(VQC_ATTR)
#undef VQC_ATTR
#define VQC_ATTR2(M, K) VQC_HIP(hipFuncSetAttribute((const void*)k_small_eval<M, K>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    VQC_ATTR2(2, KS_REV) VQC_ATTR2(2, KS_REVX) VQC_ATTR2(3, KS_REV) VQC_ATTR2(3, KS_REVX) VQC_ATTR2(5, KS_REV)
    VQC_ATTR2(5, KS_REVX)
#undef VQC_ATTR2
    done = true;
    return VQC_OK;
}

// kind set of the micro-ops of phase i
int vqc_phase_kindset(const PlanBuilder& pb, size_t i) {
    const PhaseDesc& p = pb.phases[i];
    if (p.kind == PH_FWD) {
        for (int o = p.op_begin; o < p.op_end; ++o)
            if (pb.mops[o].kind == MK_DENSE) return KS_FWDD;
        // the ansatz-shape kernel when a block of that shape is present (and the
        // phase needs no 4- or 5-bit micro-op kernel)
        if (vqc_max_kb(pb.mops, p.op_begin, p.op_end) > 3) return KS_FWD;
        for (int o = p.op_begin; o < p.op_end; ++o) {
            const MicroOp& m = pb.mops[o];
            if (m.kind == MK_BLOCK && m.nchain == 2 && m.cmat[0] >= 0 && m.cmat[1] >= 0 &&
                ((m.smask == kHeaFwdSmask && m.mstr == kHeaFwdMstr) || (m.smask == kHeaFwdSmaskHi && m.mstr == kHeaFwdMstrHi)))
                return KS_FWDH;
        }
        return KS_FWD;
    }
    if (p.kind == PH_HAM) return KS_HAM;
    for (int o = p.op_begin; o < p.op_end; ++o)
        if (!vqc_rev_simple(pb.mops[o])) return KS_REVX;
    return KS_REV;
}

// dpd: device copy of pb.phases
static int launch_phase(const PlanBuilder& pb, size_t i, const PhaseDesc* dpd, const PhaseBufs& b, unsigned grid, hipStream_t st) {
    const PhaseDesc& p = pb.phases[i];
    const size_t lds = phase_lds(p);
    VQC_TRY(set_lds_attrs());
    const int mkb = vqc_max_kb(pb.mops, p.op_begin, p.op_end);
    const int mk = mkb <= 2 ? 2 : (mkb <= 3 ? 3 : 5);
    const int ks = vqc_phase_kindset(pb, i);
    bool launched = false;
#define VQC_LAUNCH(M, K)                                                                           \
    if (mk == M && ks == K) {                                                                      \
        hipLaunchKernelGGL((k_phase<M, K>), dim3(grid), dim3(kTileThreads), lds, st, dpd + i, b); \
        launched = true;                                                                           \
    }
    VQC_FOR_KERNELS(VQC_LAUNCH)
#undef VQC_LAUNCH
    if (!launched) return vqc_fail(VQC_EINVAL, "no tile kernel for micro-op width %d and kind set %d", mk, ks);
    if (vqc_trace_on()) {
        // micro-ops as kind:bits, the MK_DENSE ones as D<bits>
        std::string mo;
        for (int o = p.op_begin; o < p.op_end; ++o) {
            const MicroOp& op = pb.mops[o];
            char w[24];
            if (op.kind == MK_DENSE) std::snprintf(w, sizeof w, " D%d", op.kb);
            else std::snprintf(w, sizeof w, " %d:%d", op.kind, op.kb);
            mo += w;
        }
        vqc_trace("k_phase<%d,%d> phase %zu kind %d nbits %d t %d low %d grid %u lds %zu nacc %d ops%s", mk, ks, i, p.kind,
                  p.nbits, p.t, p.m, grid, lds, p.nacc, mo.c_str());
    }
    VQC_HIP(hipGetLastError());
    return VQC_OK;
}

static unsigned phase_grid(const PhaseDesc& p, unsigned max_acc_grid) {
    const uint64_t ntiles = uint64_t(1) << (p.nbits - p.t);
    if (p.nacc == 0) return unsigned(std::min<uint64_t>(ntiles, 1u << 30));
    return unsigned(std::min<uint64_t>(ntiles, max_acc_grid));
}

// The columns of the partial rows that the phases write: workgroup b of a phase
// launched on phase_grid(p, stride) workgroups writes column b of each of the
// phase's rows, so rows come in segments of equal column counts (no zero fill
// needed). False when there are more segments than a kernel argument holds.
static bool row_segments(const PlanBuilder& pb, int stride, VqcRowSegs* s) {
    s->n = 0;
    int cur = 0;
    auto push = [&](int end, int cols) {
        if (end <= cur) return true;
        if (s->n > 0 && s->cols[s->n - 1] == cols) {
            s->end[s->n - 1] = end;
        } else {
            if (s->n == kMaxRowSegs) return false;
            s->end[s->n] = end;
            s->cols[s->n] = cols;
            ++s->n;
        }
        cur = end;
        return true;
    };
    for (const PhaseDesc& p : pb.phases) {
        if (p.nacc <= 0) continue;
        if (p.row_begin < cur) return false;  // rows out of order: not the layout assumed here
        if (!push(p.row_begin, 0) || !push(p.row_begin + p.nacc, int(phase_grid(p, unsigned(stride))))) return false;
    }
    return push(pb.nrows, 0);
}

// Run planned FWD phases from src into A.
static int run_forward(const PlanBuilder& pb, size_t ph_begin, size_t ph_end, const double2* src, double2* A,
                       const MicroOp* dmops, const double2* dpool, const PhaseDesc* dpd, hipStream_t st) {
    for (size_t i = ph_begin; i < ph_end; ++i) {
        PhaseBufs b{i == ph_begin ? src : A, A, nullptr, dmops, dpool, nullptr, 0};
        VQC_TRY(launch_phase(pb, i, dpd, b, phase_grid(pb.phases[i], 1), st));
    }
    return VQC_OK;
}

// The forward phases [0, nfwd) from src into A, storing checkpoint k (L
// amplitudes at ck + k L) after phase e.ck_store[k] - 1.
struct PlanEntry;
static int run_forward_ckpt(const PlanEntry& e, size_t nfwd, const double2* src, double2* A, double2* ck, uint64_t L,
                            const MicroOp* dmops, const double2* dpool, hipStream_t st);
static int ckpt_restore(const PlanEntry& e, size_t phase, double2* A, const double2* ck, uint64_t L, hipStream_t st);

// shape of the most recent plan uploaded by this thread (vqc_plan_info)
static thread_local vqc_plan_info_t g_last_plan{};
// whether the plan being run came from the plan cache
static bool g_plan_reused = false;

static void record_plan(const PlanBuilder& pb) {
    vqc_plan_info_t& r = g_last_plan;
    memset(&r, 0, sizeof r);
    for (const PhaseDesc& p : pb.phases) {
        if (p.kind == PH_FWD) ++r.phases_fwd;
        else if (p.kind == PH_HAM) ++r.phases_ham;
        else ++r.phases_rev;
    }
    for (const MicroOp& m : pb.mops)
        if (m.kind >= 0 && m.kind < 8) ++r.mops[m.kind];
    r.reused = g_plan_reused;
}

extern "C" int vqc_plan_info(vqc_plan_info_t* out) {
    VQC_API_LOCK;  // g_last_plan is per thread; the lock orders this read after the calling thread's last plan
    if (!out) return vqc_fail(VQC_EINVAL, "null plan info");
    *out = g_last_plan;
    return VQC_OK;
}

// cached plan whose micro-ops and phases the device buffers hold (0: unknown)
static thread_local uint64_t g_uploaded_plan = 0;

// The matrix pool always goes up; the micro-ops and phases only when the device
// does not hold them already (plan_id: the cached plan they belong to, 0 for an
// uncached one).
static int upload_program(const PlanBuilder& pb, hipStream_t st, uint64_t plan_id = 0) {
    record_plan(pb);
    const std::vector<double>& h = pb.xp->h;
    VQC_TRY(ensure(g_wpool, h.size() * 8 + 16));
    if (!h.empty()) VQC_TRY(upload_staged(g_wpool.buf, h.data(), h.size() * 8, st));
    if (plan_id != 0 && plan_id == g_uploaded_plan) return VQC_OK;
    g_uploaded_plan = 0;
    VQC_TRY(ensure(g_wmops, std::max<size_t>(pb.mops.size(), 1) * sizeof(MicroOp)));
    if (!pb.mops.empty())
        VQC_HIP(hipMemcpyAsync(g_wmops.buf, pb.mops.data(), pb.mops.size() * sizeof(MicroOp), hipMemcpyHostToDevice, st));
    VQC_TRY(ensure(g_wphases, std::max<size_t>(pb.phases.size(), 1) * sizeof(PhaseDesc)));
    if (!pb.phases.empty())
        VQC_HIP(hipMemcpyAsync(g_wphases.buf, pb.phases.data(), pb.phases.size() * sizeof(PhaseDesc), hipMemcpyHostToDevice, st));
    g_uploaded_plan = plan_id;
    return VQC_OK;
}

// ------------------------------- plan cache ---------------------------------
// Plans of the most recent programs. An optimizer evaluates one program again
// and again with new angles. A plan depends on the op and term tables (kinds,
// positions, pool offsets, hints), the pool lengths, the tile tuning and the
// entry point, and on the matrix values only through the derived matrices it
// made (XPool::table) and the structure classes it read (XPool::claims). A hit
// loads the new values, recomputes the derived matrices and checks every class
// again; when one no longer holds (an angle that makes a rotation diagonal, say)
// the program is planned again. The cache lives under the library lock.
struct PlanEntry {
    std::string key;
    XPool xp;
    PlanBuilder pb;
    int info[8] = {};  // entry-point data (launch shape, phase boundaries)
    uint64_t id = 0;
    // forward-state checkpoints: checkpoint k is stored once the forward phases
    // [0, ck_store[k]) have run and restored into A before phase ck_restore[k]
    std::vector<int> ck_store, ck_restore;
};
constexpr size_t kPlanCacheSize = 8;
static std::vector<std::unique_ptr<PlanEntry>> g_plans;  // most recent first
static uint64_t g_plan_ids = 0;
static bool g_plan_cache_on = std::getenv("VQC_PLAN_CACHE") == nullptr || std::atoi(std::getenv("VQC_PLAN_CACHE")) != 0;

template <class T>
static void key_put(std::string& k, const T& v) {
    k.append(reinterpret_cast<const char*>(&v), sizeof v);
}
static std::string plan_key(char entry, int nbits, int n, int extra, int nparams, int nops, const vqc_op_t* ops, int64_t plen,
                            int nterms, const vqc_term_t* terms, int64_t hlen) {
    const TileTuning tt = vqc_tile_tuning(nbits);
    std::string k;
    k.reserve(64 + size_t(nops) * sizeof(vqc_op_t) + size_t(nterms) * sizeof(vqc_term_t));
    k.push_back(entry);
    for (int v : {nbits, n, extra, nparams, nops, nterms, tt.t_fwd, tt.t_rev, tt.t_small, tt.mlow}) key_put(k, v);
    key_put(k, plen);
    key_put(k, hlen);
    if (nops) k.append(reinterpret_cast<const char*>(ops), size_t(nops) * sizeof(vqc_op_t));
    if (nterms) k.append(reinterpret_cast<const char*>(terms), size_t(nterms) * sizeof(vqc_term_t));
    return k;
}
// The cached plan for key with the caller's matrices loaded, or null.
static PlanEntry* plan_find(const std::string& key, const double* pool, int64_t plen, const double* hpool, int64_t hlen) {
    if (!g_plan_cache_on) return nullptr;
    for (size_t i = 0; i < g_plans.size(); ++i) {
        if (g_plans[i]->key != key) continue;
        std::unique_ptr<PlanEntry> e = std::move(g_plans[i]);
        g_plans.erase(g_plans.begin() + i);
        if (!e->xp.reload(pool, plen, hpool, hlen)) return nullptr;
        g_plans.insert(g_plans.begin(), std::move(e));
        g_plan_reused = true;
        return g_plans.front().get();
    }
    return nullptr;
}
// A new, empty plan whose pool holds the caller's matrices (pool, then hpool).
static std::unique_ptr<PlanEntry> plan_make(std::string key, const double* pool, int64_t plen, const double* hpool, int64_t hlen) {
    std::unique_ptr<PlanEntry> e(new PlanEntry);
    g_plan_reused = false;
    e->key = std::move(key);
    if (plen) e->xp.h.assign(pool, pool + 2 * plen);
    if (hlen) e->xp.h.insert(e->xp.h.end(), hpool, hpool + 2 * hlen);
    e->pb.xp = &e->xp;
    e->id = ++g_plan_ids;
    return e;
}
// Keep a finished plan (the least recent one beyond the cache size goes).
static PlanEntry* plan_store(std::unique_ptr<PlanEntry> e) {
    g_plans.insert(g_plans.begin(), std::move(e));
    if (g_plans.size() > kPlanCacheSize) g_plans.pop_back();
    return g_plans.front().get();
}

extern "C" int vqc_apply_circuit(vqc_state* s, int nops, const vqc_op_t* ops, const double* pool) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_qubits(s, "vqc_apply_circuit"));
    int64_t plen = 0;
    VQC_TRY(check_program(s->nqubits, s->is_dm, nops, ops, &plen, 1 << 30));
    if (nops == 0) return VQC_OK;
    if (!pool) return vqc_fail(VQC_EINVAL, "null matrix pool");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    std::string key = plan_key('F', s->nbits, s->nqubits, s->is_dm, 0, nops, ops, plen, 0, nullptr, 0);
    PlanEntry* e = plan_find(key, pool, plen, nullptr, 0);
    if (!e) {
        const TileTuning tt = vqc_tile_tuning(s->nbits);
        std::unique_ptr<PlanEntry> ne = plan_make(std::move(key), pool, plen, nullptr, 0);
        ne->pb.dense_fwd = true;
        VQC_TRY(vqc_plan_forward(ne->pb, s->nbits, s->nqubits, s->is_dm, nops, ops, tt.t_fwd, tt.mlow));
        vqc_fuse_chains(ne->pb, 0, ne->pb.phases.size());
        e = plan_store(std::move(ne));
    }
    const PlanBuilder& pb = e->pb;
    VQC_TRY(upload_program(pb, c->stream, e->id));
    return run_forward(pb, 0, pb.phases.size(), s->d, s->d, (const MicroOp*)g_wmops.buf, (const double2*)g_wpool.buf,
                       (const PhaseDesc*)g_wphases.buf, c->stream);
}

// Sum rows into energy and gradient (host, fixed order).
static void finish_rows(const std::vector<double>& rows, const std::vector<int>& row_slot, int nparams, double* energy2,
                        double* grad) {
    double er = 0, ei = 0;
    for (int p = 0; p < nparams; ++p) grad[p] = 0;
    for (size_t r = 0; r < row_slot.size(); ++r) {
        if (row_slot[r] < 0) {
            er += rows[2 * r];
            ei += rows[2 * r + 1];
        } else {
            grad[row_slot[r]] += rows[2 * r];
        }
    }
    if (energy2) {
        energy2[0] = er;
        energy2[1] = ei;
    }
}

// --------------------------- plain state-vector path ------------------------
static void dagger_host(const double* U, int D, double* out) {
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) {
            out[2 * (r + c * D)] = U[2 * (c + r * D)];
            out[2 * (r + c * D) + 1] = -U[2 * (c + r * D) + 1];
        }
}

static int sv_energy_grad_plain(const vqc_state* psi0, int nops, const vqc_op_t* ops, const double* pool, int64_t plen, int nterms,
                                const vqc_term_t* terms, const double* hpool, int nparams, double* energy2, double* grad,
                                double2* A, double2* B, double2* fwd_out, hipStream_t st) {
    const int nb = psi0->nbits;
    const uint64_t L = psi0->len();
    // device pool: [ops pool | for each term M, M + M^dagger]
    std::vector<double> dp(pool, pool + 2 * plen);
    std::vector<int64_t> toff(nterms), tdoff(nterms);
    for (int t = 0; t < nterms; ++t) {
        const int D = 1 << terms[t].nq;
        const double* M = hpool + 2 * terms[t].mat;
        toff[t] = int64_t(dp.size() / 2);
        dp.insert(dp.end(), M, M + 2 * D * D);
        tdoff[t] = int64_t(dp.size() / 2);
        std::vector<double> md(2 * D * D);
        dagger_host(M, D, md.data());
        for (int i = 0; i < 2 * D * D; ++i) md[i] += M[i];
        dp.insert(dp.end(), md.begin(), md.end());
    }
    VQC_TRY(ensure(g_wpool, dp.size() * 8));
    VQC_HIP(hipMemcpyAsync(g_wpool.buf, dp.data(), dp.size() * 8, hipMemcpyHostToDevice, st));
    const double2* dpool = (const double2*)g_wpool.buf;
    // rows: energy per term, then one per active parameter
    std::vector<int> row_slot;
    for (int t = 0; t < nterms; ++t) row_slot.push_back(-1 - t);
    for (int i = nops - 1; i >= 0; --i)
        for (int j = 0; j < ops[i].nparams; ++j)
            if (ops[i].slot[j] >= 0) row_slot.push_back(ops[i].slot[j]);
    const int nrows = int(row_slot.size());
    const int stride = 1024;
    VQC_TRY(ensure(g_wparts, size_t(std::max(nrows, 1)) * (stride + 1) * 16));
    double2* parts = (double2*)g_wparts.buf;
    double2* rowsum = parts + size_t(std::max(nrows, 1)) * stride;
    vqc_zero(parts, size_t(std::max(nrows, 1)) * stride, st);
    // forward: y = circuit * psi0 (generic.jl:26-33)
    VQC_HIP(hipMemcpyAsync(A, psi0->d, L * 16, hipMemcpyDeviceToDevice, st));
    for (int i = 0; i < nops; ++i) VQC_TRY(vqc_dense_sweep(A, nb, ops[i].nq, ops[i].pos, pool + 2 * ops[i].mat, st));
    if (fwd_out) VQC_HIP(hipMemcpyAsync(fwd_out, A, L * 16, hipMemcpyDeviceToDevice, st));
    // energy per key and seed Delta = (H + H^dagger) y
    int row = 0;
    unsigned grid;
    for (int t = 0; t < nterms; ++t, ++row)
        VQC_TRY(vqc_expec_sweep(A, nullptr, nb, terms[t].nq, terms[t].pos, dpool + toff[t], parts + size_t(row) * stride, &grid, st));
    vqc_zero(B, L, st);
    for (int t = 0; t < nterms; ++t) VQC_TRY(vqc_accum_sweep(A, B, nb, terms[t].nq, terms[t].pos, dpool + tdoff[t], st));
    // reverse sweep (circuitdiff.jl:23-32)
    std::vector<double> ud;
    for (int i = nops - 1; i >= 0; --i) {
        const vqc_op_t& o = ops[i];
        const int D = 1 << o.nq;
        ud.resize(2 * D * D);
        dagger_host(pool + 2 * o.mat, D, ud.data());
        VQC_TRY(vqc_dense_sweep(B, nb, o.nq, o.pos, ud.data(), st));
        VQC_TRY(vqc_dense_sweep(A, nb, o.nq, o.pos, ud.data(), st));
        for (int j = 0; j < o.nparams; ++j) {
            if (o.slot[j] < 0) continue;
            VQC_TRY(vqc_expec_sweep(A, B, nb, o.nq, o.pos, dpool + o.gen[j], parts + size_t(row) * stride, &grid, st));
            ++row;
        }
    }
    if (nrows) vqc_reduce_partials(parts, nrows, stride, rowsum, st);
    VQC_HIP(hipGetLastError());
    std::vector<double> rows(2 * size_t(std::max(nrows, 1)));
    if (nrows) VQC_TRY(vqc_fetch(rowsum, nrows, rows.data()));
    else VQC_HIP(hipStreamSynchronize(st));
    finish_rows(rows, row_slot, nparams, energy2, grad);
    return VQC_OK;
}

// --------------------------- fused state-vector path ------------------------
static int max_small_rows(int nbits) {
    const size_t T = size_t(1) << nbits;
    const size_t used = (2 * T + kRedDoubles2) * 16;
    return used >= 160 * 1024 ? 0 : int((160 * 1024 - used) / 16);
}

// ---------------------------------------------------------------------------
// Forward-state checkpoints. The reverse sweep recovers each earlier state by
// uncomputing (U^dagger for gates, S^-1 for channels, circuitdiff.jl:23-32,
// 53-59). A unitary uncompute is exact up to rounding, but S^-1 amplifies the
// error of what it is applied to by up to ||S^-1|| per channel, and a channel
// without an inverse (depolarizing at p = 3/4) cannot be uncomputed at all.
// The op list is then cut into segments; the forward pass stores the state
// after each segment but the last in spare device memory, and the reverse pass
// restores it at the start of that segment's reverse instead of trusting the
// uncomputed one. A segment starts right before every channel whose inverse is
// missing or would push the accumulated amplification of the channels
// uncomputed since the last restore above VQC_CKPT_AMP (1e3, three digits), and,
// with VQC_FLAG_CHECKPOINT, every ceil(sqrt(nops)) ops. Such a channel is the
// last op of its segment's reverse, planned alone and with the identity in
// place of S^-1: its input state is the next restore.
// ---------------------------------------------------------------------------
struct CkptCut {
    std::vector<int> begin;  // segment starts, begin[0] = 0
    std::vector<char> chan;  // the segment starts with a channel undone through the checkpoint
};
static double norm1_cols(const double* M, int D) {
    double best = 0;
    for (int c = 0; c < D; ++c) {
        double s = 0;
        for (int r = 0; r < D; ++r) s += std::hypot(M[2 * (r + c * D)], M[2 * (r + c * D) + 1]);
        best = std::max(best, s);
    }
    return best;
}
static CkptCut ckpt_cut(int nops, const vqc_op_t* ops, const double* pool, int is_dm, int flags) {
    std::vector<char> at(size_t(nops) + 1, 0), chan(size_t(nops) + 1, 0);
    if (flags & VQC_FLAG_CHECKPOINT) {
        const int seg = std::max(1, int(std::ceil(std::sqrt(double(nops)))));
        for (int i = seg; i < nops; i += seg) at[i] = 1;
    }
    if (is_dm) {
        double lim = 1e3;
        if (const char* e = getenv("VQC_CKPT_AMP")) lim = std::max(1.0, atof(e));
        double amp = 1.0;
        for (int i = nops - 1; i >= 0; --i) {
            const vqc_op_t& o = ops[i];
            if (o.kind != VQC_OP_CHANNEL) continue;
            const double a = o.aux < 0 ? INFINITY : norm1_cols(pool + 2 * o.aux, 1 << (2 * o.nq));
            if (o.aux < 0 || amp * a > lim) {
                at[i] = 1;
                chan[i] = 1;
                amp = 1.0;
            } else {
                amp *= std::max(a, 1.0);
            }
        }
    }
    CkptCut c;
    c.begin.push_back(0);
    c.chan.push_back(chan[0]);
    for (int i = 1; i < nops; ++i)
        if (at[i]) {
            c.begin.push_back(i);
            c.chan.push_back(chan[i]);
        }
    return c;
}
static std::string cut_key(const CkptCut& c) {
    std::string k;
    for (size_t i = 0; i < c.begin.size(); ++i) {
        key_put(k, c.begin[i]);
        k.push_back(c.chan[i]);
    }
    return k;
}
// The identity superoperator (4^N x 4^N) appended to a host pool, for the
// checkpointed channels' reverse; it does not depend on the caller's matrices.
static int64_t pool_identity(XPool& xp, int D) {
    const int64_t off = xp.size();
    xp.h.resize(xp.h.size() + size_t(2) * D * D, 0.0);
    for (int i = 0; i < D; ++i) xp.h[2 * (off + i + int64_t(i) * D)] = 1.0;
    return off;
}
static int run_forward_ckpt(const PlanEntry& e, size_t nfwd, const double2* src, double2* A, double2* ck, uint64_t L,
                            const MicroOp* dmops, const double2* dpool, hipStream_t st) {
    size_t k = 0;
    for (size_t i = 0; i < nfwd; ++i) {
        PhaseBufs b{i == 0 ? src : A, A, nullptr, dmops, dpool, nullptr, 0};
        VQC_TRY(launch_phase(e.pb, i, (const PhaseDesc*)g_wphases.buf, b, phase_grid(e.pb.phases[i], 1), st));
        while (k < e.ck_store.size() && size_t(e.ck_store[k]) == i + 1) {
            VQC_HIP(hipMemcpyAsync(ck + k * L, A, L * 16, hipMemcpyDeviceToDevice, st));
            ++k;
        }
    }
    return VQC_OK;
}
static int ckpt_restore(const PlanEntry& e, size_t phase, double2* A, const double2* ck, uint64_t L, hipStream_t st) {
    for (size_t k = 0; k < e.ck_restore.size(); ++k)
        if (size_t(e.ck_restore[k]) == phase) VQC_HIP(hipMemcpyAsync(A, ck + k * L, L * 16, hipMemcpyDeviceToDevice, st));
    return VQC_OK;
}

// Forward phases of every segment; checkpoint k after segment k's phases.
static int plan_forward_cut(PlanEntry& e, const CkptCut& cut, int nbits, int n, int is_dm, int nops, const vqc_op_t* ops, int t,
                            int mlow) {
    const size_t S = cut.begin.size();
    for (size_t s = 0; s < S; ++s) {
        const int b = cut.begin[s], end = s + 1 < S ? cut.begin[s + 1] : nops;
        VQC_TRY(vqc_plan_forward(e.pb, nbits, n, is_dm, end - b, ops + b, t, mlow));
        if (s + 1 < S) e.ck_store.push_back(int(e.pb.phases.size()));
    }
    return VQC_OK;
}
// Reverse phases, last segment first; checkpoint k is restored before segment k's.
static int plan_reverse_cut(PlanEntry& e, const CkptCut& cut, int nbits, int n, int is_dm, int nops, const vqc_op_t* ops, int t,
                            int mlow) {
    const size_t S = cut.begin.size();
    e.ck_restore.assign(S > 0 ? S - 1 : 0, 0);
    for (size_t s = S; s-- > 0;) {
        const int b = cut.begin[s], end = s + 1 < S ? cut.begin[s + 1] : nops;
        if (s + 1 < S) e.ck_restore[s] = int(e.pb.phases.size());
        if (cut.chan[s]) {
            // the channel opening the segment: its reverse runs last and alone, with
            // the identity on the state (the restore that follows replaces it)
            if (end - b > 1) VQC_TRY(vqc_plan_reverse(e.pb, nbits, n, is_dm, end - b - 1, ops + b + 1, t, mlow));
            vqc_op_t ch = ops[b];
            ch.aux = pool_identity(e.xp, 1 << (2 * ch.nq));
            VQC_TRY(vqc_plan_reverse(e.pb, nbits, n, is_dm, 1, &ch, t, mlow));
        } else {
            VQC_TRY(vqc_plan_reverse(e.pb, nbits, n, is_dm, end - b, ops + b, t, mlow));
        }
    }
    return VQC_OK;
}
// Checkpoint buffers of one call (grown on demand, kept for the next call)
static thread_local Work g_wckpt;

void vqc_release_circuit_thread() {
    for (Work* w : {&g_wA, &g_wB, &g_wparts, &g_wpool, &g_wmops, &g_wrows, &g_wphases, &g_wckpt}) w->release();
    g_upload.release();
    g_uploaded_plan = 0;  // the micro-op and phase buffers are gone
}
static int ckpt_buffers(const PlanEntry& e, uint64_t L, double2** out) {
    const size_t k = e.ck_store.size();
    if (!k) {
        *out = nullptr;
        return VQC_OK;
    }
    if (ensure(g_wckpt, k * L * 16) != VQC_OK)
        return vqc_fail(VQC_ENOMEM, "forward-state checkpoints: %zu states of %llu bytes do not fit", k,
                        (unsigned long long)(L * 16));
    *out = (double2*)g_wckpt.buf;
    return VQC_OK;
}

// Plan of the state-vector adjoint into e. The whole evaluation runs in one
// workgroup when allow_small and the state and the accumulators fit in LDS
// (info: 1, Hamiltonian start, reverse start, micro-op width, reverse kind set);
// otherwise in tile phases (info: 0, forward phases, forward + Hamiltonian phases).
static int plan_sv(PlanEntry& e, int nb, int nops, const vqc_op_t* ops, int nterms, const vqc_term_t* terms, int64_t plen,
                   int64_t hlen, bool allow_small, const CkptCut& cut) {
    const TileTuning tt = vqc_tile_tuning(nb);
    PlanBuilder& pb = e.pb;
    const bool cuts = cut.begin.size() > 1;
    if (allow_small && !cuts && nb <= tt.t_small) {
        VQC_TRY(vqc_plan_forward(pb, nb, nb, 0, nops, ops, nb, 0));
        const size_t nf = pb.phases.size();
        VQC_TRY(vqc_plan_ham(pb, nb, nterms, terms, plen, nb, 0));
        const size_t nh = pb.phases.size();
        VQC_TRY(vqc_plan_reverse(pb, nb, nb, 0, nops, ops, nb, 0));
        vqc_fuse_chains(pb, 0, pb.phases.size());
        if (pb.nrows <= max_small_rows(nb)) {
            // accumulator index := global row
            for (const PhaseDesc& p : pb.phases)
                for (int o = p.op_begin; o < p.op_end; ++o)
                    for (int q = 0; q < 4; ++q) pb.mops[o].acc[q] += p.row_begin;
            const int rev_begin = vqc_phase_op_start(pb, nh);
            const int mkb = vqc_max_kb(pb.mops, 0, int(pb.mops.size()));
            bool rev_simple = true;
            for (size_t o = size_t(rev_begin); o < pb.mops.size(); ++o) rev_simple &= vqc_rev_simple(pb.mops[o]);
            e.info[0] = 1;
            e.info[1] = vqc_phase_op_start(pb, nf);
            e.info[2] = rev_begin;
            e.info[3] = mkb <= 2 ? 2 : (mkb <= 3 ? 3 : 5);
            e.info[4] = rev_simple ? KS_REV : KS_REVX;
            return VQC_OK;
        }
        // too many accumulators for LDS: plan again in tiles, from the caller's matrices only
        pb = PlanBuilder();
        pb.xp = &e.xp;
        e.xp.h.resize(size_t(2) * (plen + hlen));
        e.xp.cache.clear();
        e.xp.table.clear();
        e.xp.claims.clear();
    }
    pb.dense_fwd = true;
    if (cuts) VQC_TRY(plan_forward_cut(e, cut, nb, nb, 0, nops, ops, tt.t_fwd, tt.mlow));
    else VQC_TRY(vqc_plan_forward(pb, nb, nb, 0, nops, ops, tt.t_fwd, tt.mlow));
    const size_t nfwd = pb.phases.size();
    VQC_TRY(vqc_plan_ham(pb, nb, nterms, terms, plen, tt.t_rev, tt.mlow));
    const size_t nham_end = pb.phases.size();
    if (cuts) VQC_TRY(plan_reverse_cut(e, cut, nb, nb, 0, nops, ops, tt.t_rev, tt.mlow));
    else VQC_TRY(vqc_plan_reverse(pb, nb, nb, 0, nops, ops, tt.t_rev, tt.mlow));
    vqc_fuse_chains(pb, 0, pb.phases.size());
    e.info[0] = 0;
    e.info[1] = int(nfwd);
    e.info[2] = int(nham_end);
    return VQC_OK;
}

// The state-vector adjoint sweep. The reverse buffer B starts as
// (H + H^dagger) psi (the energy's cotangent) or, when `seed` is given (no
// terms then), as a copy of it; `seed_out` (may be null) receives B after the
// reverse sweep: the cotangent on psi0.
static int sv_adjoint(const vqc_state* psi0, int nops, const vqc_op_t* ops, const double* pool, int nterms,
                      const vqc_term_t* terms, const double* hpool, int nparams, double* energy2, double* grad,
                      vqc_state* psi_out, int flags, const vqc_state* seed, vqc_state* seed_out) {
    VQC_TRY(vqc_check_qubits(psi0, "vqc_sv_energy_grad"));
    if (psi0->is_dm) return vqc_fail(VQC_EINVAL, "vqc_sv_energy_grad: state is a density matrix");
    if (nparams < 0 || (nparams && !grad)) return vqc_fail(VQC_EINVAL, "bad gradient output");
    if (psi_out && (psi_out->nbits != psi0->nbits || psi_out->is_dm)) return vqc_fail(VQC_EINVAL, "psi_out has the wrong shape");
    int64_t plen = 0, hlen = 0;
    VQC_TRY(check_program(psi0->nbits, 0, nops, ops, &plen, nparams));
    VQC_TRY(check_terms(psi0->nbits, nterms, terms, &hlen));
    if ((plen && !pool) || (hlen && !hpool)) return vqc_fail(VQC_EINVAL, "null matrix pool");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    hipStream_t st = c->stream;
    const int nb = psi0->nbits;
    const uint64_t L = psi0->len();
    VQC_TRY(ensure(g_wB, L * 16));
    VQC_TRY(ensure(g_wA, L * 16));
    double2* A = (double2*)g_wA.buf;  // the reverse sweep uncomputes A back to psi0
    double2* B = (double2*)g_wB.buf;
    double2* fwd_out = psi_out ? psi_out->d : nullptr;
    if (flags & VQC_FLAG_UNFUSED) {
        if (seed) return vqc_fail(VQC_EUNSUPPORTED, "vqc_circuit_vjp runs on the fused path only (flags must be 0)");
        VQC_TRY(sv_energy_grad_plain(psi0, nops, ops, pool, plen, nterms, terms, hpool, nparams, energy2, grad, A, B, fwd_out, st));
    } else {
        const CkptCut cut = ckpt_cut(nops, ops, pool, 0, flags);
        std::string key = plan_key('V', nb, nb, seed ? 1 : 0, nparams, nops, ops, plen, nterms, terms, hlen) + cut_key(cut);
        PlanEntry* e = plan_find(key, pool, plen, hpool, hlen);
        if (!e) {
            std::unique_ptr<PlanEntry> ne = plan_make(std::move(key), pool, plen, hpool, hlen);
            VQC_TRY(plan_sv(*ne, nb, nops, ops, nterms, terms, plen, hlen, !seed, cut));
            e = plan_store(std::move(ne));
        }
        const PlanBuilder& pb = e->pb;
        VQC_TRY(upload_program(pb, st, e->id));
        if (e->info[0]) {
            const int ham_begin = e->info[1], rev_begin = e->info[2], mk = e->info[3], ks = e->info[4];
            VQC_TRY(ensure(g_wrows, size_t(std::max(pb.nrows, 1)) * 16));
            SmallArgs a{nb, int(pb.mops.size()), ham_begin, rev_begin, pb.nrows, psi0->d, fwd_out, (const MicroOp*)g_wmops.buf,
                        (const double2*)g_wpool.buf, (double2*)g_wrows.buf, 0, 0};
            vqc_prefetch_range(pb.mops, 0, int(pb.mops.size()), pb.xp->size(), &a.pf_lo, &a.pf_n);
            VQC_TRY(set_lds_attrs());
            const size_t lds = (2 * L + kRedDoubles2 + size_t(std::max(pb.nrows, 1))) * 16;
#define VQC_SMALL(M, K) \
    if (mk == M && ks == K) hipLaunchKernelGGL((k_small_eval<M, K>), dim3(1), dim3(kTileThreads), lds, st, a);
            VQC_SMALL(2, KS_REV) VQC_SMALL(2, KS_REVX) VQC_SMALL(3, KS_REV) VQC_SMALL(3, KS_REVX) VQC_SMALL(5, KS_REV)
            VQC_SMALL(5, KS_REVX)
#undef VQC_SMALL
            VQC_HIP(hipGetLastError());
            std::vector<double> rows(2 * size_t(std::max(pb.nrows, 1)));
            // the wait releases the API lock: keep what the plan's rows mean first
            const std::vector<int> row_slot = pb.row_slot;
            if (pb.nrows) VQC_TRY(vqc_fetch_unlocked((double2*)g_wrows.buf, pb.nrows, rows.data()));
            else VQC_HIP(hipStreamSynchronize(st));
            finish_rows(rows, row_slot, nparams, energy2, grad);
            return VQC_OK;
        }
        const size_t nfwd = size_t(e->info[1]), nham_end = size_t(e->info[2]);
        const MicroOp* dm = (const MicroOp*)g_wmops.buf;
        const double2* dp = (const double2*)g_wpool.buf;
        const int stride = 1024;
        const int nrows = std::max(pb.nrows, 1);
        VQC_TRY(ensure(g_wparts, size_t(nrows) * (stride + 1) * 16));
        double2* parts = (double2*)g_wparts.buf;
        double2* rowsum = parts + size_t(nrows) * stride;
        VqcRowSegs segs;
        const bool by_segs = row_segments(pb, stride, &segs);
        if (!by_segs) vqc_zero(parts, size_t(nrows) * stride, st);
        double2* ck;
        VQC_TRY(ckpt_buffers(*e, L, &ck));
        if (nfwd == 0) VQC_HIP(hipMemcpyAsync(A, psi0->d, L * 16, hipMemcpyDeviceToDevice, st));
        else VQC_TRY(run_forward_ckpt(*e, nfwd, psi0->d, A, ck, L, dm, dp, st));
        if (fwd_out) VQC_HIP(hipMemcpyAsync(fwd_out, A, L * 16, hipMemcpyDeviceToDevice, st));
        if (seed) VQC_HIP(hipMemcpyAsync(B, seed->d, L * 16, hipMemcpyDeviceToDevice, st));
        else if (nham_end == nfwd) vqc_zero(B, L, st);
        for (size_t i = nfwd; i < pb.phases.size(); ++i) {
            VQC_TRY(ckpt_restore(*e, i, A, ck, L, st));
            PhaseBufs b{A, A, B, dm, dp, parts, stride};
            VQC_TRY(launch_phase(pb, i, (const PhaseDesc*)g_wphases.buf, b, phase_grid(pb.phases[i], stride), st));
        }
        if (seed_out) VQC_HIP(hipMemcpyAsync(seed_out->d, B, L * 16, hipMemcpyDeviceToDevice, st));
        std::vector<double> rows(2 * size_t(nrows));
        // the wait releases the API lock: keep what the plan's rows mean first
        const std::vector<int> row_slot = pb.row_slot;
        if (pb.nrows && by_segs) {
            // row sums straight into pinned host memory: no copy after the reduction
            double* pin;
            VQC_TRY(vqc_pinned_rows(size_t(pb.nrows) * 16, &pin));
            vqc_reduce_rows(parts, stride, segs, pb.nrows, (double2*)pin, st);
            VQC_HIP(hipGetLastError());
            VQC_TRY(vqc_wait_unlocked());
            std::copy(pin, pin + 2 * size_t(pb.nrows), rows.begin());
        } else if (pb.nrows) {
            vqc_reduce_partials(parts, pb.nrows, stride, rowsum, st);
            VQC_HIP(hipGetLastError());
            VQC_TRY(vqc_fetch_unlocked(rowsum, pb.nrows, rows.data()));
        } else {
            VQC_HIP(hipGetLastError());
            VQC_HIP(hipStreamSynchronize(st));
        }
        finish_rows(rows, row_slot, nparams, energy2, grad);
    }
    return VQC_OK;
}

extern "C" int vqc_sv_energy_grad(const vqc_state* psi0, int nops, const vqc_op_t* ops, const double* pool, int nterms,
                                  const vqc_term_t* terms, const double* hpool, int nparams, double* energy2, double* grad,
                                  vqc_state* psi_out, int flags) {
    VQC_API_LOCK;
    return sv_adjoint(psi0, nops, ops, pool, nterms, terms, hpool, nparams, energy2, grad, psi_out, flags, nullptr, nullptr);
}

// ------------------------------- density matrices ---------------------------
struct GPosDyn {
    int p[5];
    int k;
};

// Delta = H^dagger as a vec(matrix): zero, then each key adds conj(M[c,a]) at
// (b + off_a, b + off_c) for every b over the other row bits (expecdiff.jl:64-65).
template <int K>
__global__ void __launch_bounds__(256) k_dm_seed(double2* __restrict__ Dl, int n, uint64_t ngroups, GPosDyn p,
                                                 const double2* __restrict__ M) {
    constexpr int D = 1 << K;
    const uint64_t g = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (g >= ngroups) return;
    const uint64_t b = insert_zeros<K>(g, p.p);
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int c = 0; c < D; ++c) {
            const uint64_t idx = (b + sub_offset<K>(a, p.p)) + ((b + sub_offset<K>(c, p.p)) << n);
            Dl[idx] = cadd(Dl[idx], cconj(M[c + a * D]));
        }
}

static int dm_seed(double2* Dl, int n, const vqc_term_t& t, const double2* M, hipStream_t st) {
    GPosDyn p;
    for (int j = 0; j < 5; ++j) p.p[j] = j < t.nq ? t.pos[j] : 0;
    p.k = t.nq;
    const uint64_t ng = uint64_t(1) << (n - t.nq);
    const unsigned grid = grid_for(ng, 256);
    switch (t.nq) {
        case 1: hipLaunchKernelGGL(k_dm_seed<1>, dim3(grid), dim3(256), 0, st, Dl, n, ng, p, M); break;
        case 2: hipLaunchKernelGGL(k_dm_seed<2>, dim3(grid), dim3(256), 0, st, Dl, n, ng, p, M); break;
        case 3: hipLaunchKernelGGL(k_dm_seed<3>, dim3(grid), dim3(256), 0, st, Dl, n, ng, p, M); break;
        case 4: hipLaunchKernelGGL(k_dm_seed<4>, dim3(grid), dim3(256), 0, st, Dl, n, ng, p, M); break;
        case 5: hipLaunchKernelGGL(k_dm_seed<5>, dim3(grid), dim3(256), 0, st, Dl, n, ng, p, M); break;
    }
    VQC_HIP(hipGetLastError());
    return VQC_OK;
}

int vqc_dm_gate_sweep(double2* v, int n, int k, const int32_t* pos, const double* U, hipStream_t st);

// The fused density-matrix adjoint's plan (vec(rho) on 2n index bits): forward
// phases, then the reverse ones; rows 0..nterms-1 are the energies, which the
// trace kernels compute.
static int plan_dm(PlanEntry& e, int n, int nops, const vqc_op_t* ops, int nterms, const CkptCut& cut) {
    const int nb = 2 * n;
    const bool cuts = cut.begin.size() > 1 || cut.chan[0];
    const TileTuning tt = vqc_tile_tuning(nb);
    PlanBuilder& np = e.pb;
    np.dense_fwd = true;
    if (cuts) VQC_TRY(plan_forward_cut(e, cut, nb, n, 1, nops, ops, tt.t_fwd, tt.mlow));
    else VQC_TRY(vqc_plan_forward(np, nb, n, 1, nops, ops, tt.t_fwd, tt.mlow));
    vqc_fuse_chains(np, 0, np.phases.size());
    const size_t nf = np.phases.size();
    np.nrows = nterms;
    for (int t = 0; t < nterms; ++t) np.row_slot.push_back(-1 - t);
    if (cuts) VQC_TRY(plan_reverse_cut(e, cut, nb, n, 1, nops, ops, tt.t_rev, tt.mlow));
    else VQC_TRY(vqc_plan_reverse(np, nb, n, 1, nops, ops, tt.t_rev, tt.mlow));
    vqc_fuse_chains(np, nf, np.phases.size());
    e.info[0] = int(nf);
    return VQC_OK;
}

// The density-matrix adjoint sweep; `seed` / `seed_out` as for sv_adjoint
// (the seed replaces Delta = H^dagger).
static int dm_adjoint(const vqc_state* rho0, int nops, const vqc_op_t* ops, const double* pool, int nterms,
                      const vqc_term_t* terms, const double* hpool, int nparams, double* energy2, double* grad,
                      vqc_state* rho_out, int flags, const vqc_state* seed, vqc_state* seed_out) {
    VQC_TRY(vqc_check_state(rho0, "vqc_dm_energy_grad"));
    if (!rho0->is_dm) return vqc_fail(VQC_EINVAL, "vqc_dm_energy_grad: state is not a density matrix");
    if (nparams < 0 || (nparams && !grad)) return vqc_fail(VQC_EINVAL, "bad gradient output");
    if (rho_out && (rho_out->nbits != rho0->nbits || !rho_out->is_dm)) return vqc_fail(VQC_EINVAL, "rho_out has the wrong shape");
    const int n = rho0->nqubits, nb = rho0->nbits;
    int64_t plen = 0, hlen = 0;
    VQC_TRY(check_program(n, 1, nops, ops, &plen, nparams));
    VQC_TRY(check_terms(n, nterms, terms, &hlen));
    if ((plen && !pool) || (hlen && !hpool)) return vqc_fail(VQC_EINVAL, "null matrix pool");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    hipStream_t st = c->stream;
    const uint64_t L = rho0->len();
    VQC_TRY(ensure(g_wB, L * 16));
    VQC_TRY(ensure(g_wA, L * 16));
    double2* A = (double2*)g_wA.buf;  // the reverse sweep takes A back to rho0
    double2* B = (double2*)g_wB.buf;
    const bool plain = flags & VQC_FLAG_UNFUSED;
    // device pool [ops pool | hpool | variants]
    std::unique_ptr<PlanEntry> local;
    PlanEntry* e = nullptr;
    std::vector<int64_t> conj_gen(size_t(nops) * 3, -1);
    if (plain) {
        for (int i = 0; i < nops; ++i)
            if (ops[i].kind == VQC_OP_CHANNEL && ops[i].aux < 0)
                return vqc_fail(VQC_EUNSUPPORTED, "op %d: a channel without an inverse needs the fused path (its reverse "
                                "restores a forward-state checkpoint)", i);
        local = plan_make(std::string(), pool, plen, hpool, hlen);
        local->id = 0;
        e = local.get();
        PlanBuilder& lp = e->pb;
        for (int i = 0; i < nops; ++i)
            for (int j = 0; j < ops[i].nparams; ++j)
                if (ops[i].slot[j] >= 0) conj_gen[size_t(i) * 3 + j] = e->xp.variant(ops[i].gen[j], 1 << ops[i].nq, MF_CONJ);
        lp.nrows = nterms;  // rows 0..nterms-1: energies (DM energies come from the trace kernels)
        for (int t = 0; t < nterms; ++t) lp.row_slot.push_back(-1 - t);
        for (int i = nops - 1; i >= 0; --i)
            for (int j = 0; j < ops[i].nparams; ++j)
                if (ops[i].slot[j] >= 0) {
                    lp.row_slot.push_back(ops[i].slot[j]);
                    lp.nrows++;
                }
    } else {
        const CkptCut cut = ckpt_cut(nops, ops, pool, 1, flags);
        std::string key = plan_key('D', nb, n, seed ? 1 : 0, nparams, nops, ops, plen, nterms, terms, hlen) + cut_key(cut);
        e = plan_find(key, pool, plen, hpool, hlen);
        if (!e) {
            std::unique_ptr<PlanEntry> ne = plan_make(std::move(key), pool, plen, hpool, hlen);
            VQC_TRY(plan_dm(*ne, n, nops, ops, nterms, cut));
            e = plan_store(std::move(ne));
        }
    }
    const PlanBuilder& pb = e->pb;
    const size_t nfwd = plain ? 0 : size_t(e->info[0]);
    VQC_TRY(upload_program(pb, st, e->id));
    const MicroOp* dm = (const MicroOp*)g_wmops.buf;
    const double2* dp = (const double2*)g_wpool.buf;
    double2* ck = nullptr;    // forward-state checkpoints (fused path)
    const int stride = 2048;  // row and column gradient halves in the plain path
    const int nrows = std::max(pb.nrows, 1);
    VQC_TRY(ensure(g_wparts, size_t(nrows) * (stride + 1) * 16));
    double2* parts = (double2*)g_wparts.buf;
    double2* rowsum = parts + size_t(nrows) * stride;
    vqc_zero(parts, size_t(nrows) * stride, st);
    // forward
    if (plain) {
        VQC_HIP(hipMemcpyAsync(A, rho0->d, L * 16, hipMemcpyDeviceToDevice, st));
        for (int i = 0; i < nops; ++i) {
            const vqc_op_t& o = ops[i];
            if (o.kind == VQC_OP_CHANNEL) {
                int32_t all[4];
                for (int j = 0; j < o.nq; ++j) {
                    all[j] = o.pos[j];
                    all[j + o.nq] = o.pos[j] + n;
                }
                VQC_TRY(vqc_dense_sweep(A, nb, 2 * o.nq, all, pool + 2 * o.mat, st));
            } else {
                VQC_TRY(vqc_dm_gate_sweep(A, n, o.nq, o.pos, pool + 2 * o.mat, st));
            }
        }
    } else if (nfwd == 0) {
        VQC_HIP(hipMemcpyAsync(A, rho0->d, L * 16, hipMemcpyDeviceToDevice, st));
    } else {
        VQC_TRY(ckpt_buffers(*e, L, &ck));
        VQC_TRY(run_forward_ckpt(*e, nfwd, rho0->d, A, ck, L, dm, dp, st));
    }
    if (rho_out) VQC_HIP(hipMemcpyAsync(rho_out->d, A, L * 16, hipMemcpyDeviceToDevice, st));
    // energies tr(M rho) and the seed Delta = H^dagger
    unsigned grid;
    for (int t = 0; t < nterms; ++t)
        VQC_TRY(vqc_dm_expec_sweep(A, n, terms[t].nq, terms[t].pos, dp + plen + terms[t].mat, parts + size_t(t) * stride, &grid, st));
    if (seed) {
        VQC_HIP(hipMemcpyAsync(B, seed->d, L * 16, hipMemcpyDeviceToDevice, st));
    } else {
        vqc_zero(B, L, st);
        for (int t = 0; t < nterms; ++t) VQC_TRY(dm_seed(B, n, terms[t], dp + plen + terms[t].mat, st));
    }
    // reverse
    if (plain) {
        int row = nterms;
        std::vector<double> ud;
        for (int i = nops - 1; i >= 0; --i) {
            const vqc_op_t& o = ops[i];
            if (o.kind == VQC_OP_CHANNEL) {
                int32_t all[4];
                for (int j = 0; j < o.nq; ++j) {
                    all[j] = o.pos[j];
                    all[j + o.nq] = o.pos[j] + n;
                }
                const int D = 1 << (2 * o.nq);
                ud.resize(2 * D * D);
                dagger_host(pool + 2 * o.mat, D, ud.data());
                VQC_TRY(vqc_dense_sweep(B, nb, 2 * o.nq, all, ud.data(), st));
                VQC_TRY(vqc_dense_sweep(A, nb, 2 * o.nq, all, pool + 2 * o.aux, st));
                continue;
            }
            const int D = 1 << o.nq;
            ud.resize(2 * D * D);
            dagger_host(pool + 2 * o.mat, D, ud.data());
            VQC_TRY(vqc_dm_gate_sweep(B, n, o.nq, o.pos, ud.data(), st));
            VQC_TRY(vqc_dm_gate_sweep(A, n, o.nq, o.pos, ud.data(), st));
            int32_t cols[5];
            for (int j = 0; j < o.nq; ++j) cols[j] = o.pos[j] + n;
            for (int j = 0; j < o.nparams; ++j) {
                if (o.slot[j] < 0) continue;
                // rows and columns write their partials to two halves of the row
                VQC_TRY(vqc_expec_sweep(A, B, nb, o.nq, o.pos, dp + o.gen[j], parts + size_t(row) * stride, &grid, st));
                if (grid > unsigned(stride / 2)) return vqc_fail(VQC_EUNSUPPORTED, "partial buffer too small");
                unsigned g2;
                VQC_TRY(vqc_expec_sweep(A, B, nb, o.nq, cols, dp + conj_gen[size_t(i) * 3 + j], parts + size_t(row) * stride + grid, &g2, st));
                ++row;
            }
        }
    } else {
        for (size_t i = nfwd; i < pb.phases.size(); ++i) {
            VQC_TRY(ckpt_restore(*e, i, A, ck, L, st));
            PhaseBufs b{A, A, B, dm, dp, parts, stride};
            VQC_TRY(launch_phase(pb, i, (const PhaseDesc*)g_wphases.buf, b, phase_grid(pb.phases[i], stride / 2), st));
        }
    }
    if (seed_out) VQC_HIP(hipMemcpyAsync(seed_out->d, B, L * 16, hipMemcpyDeviceToDevice, st));
    if (pb.nrows) vqc_reduce_partials(parts, pb.nrows, stride, rowsum, st);
    VQC_HIP(hipGetLastError());
    std::vector<double> rows(2 * size_t(nrows));
    // the wait releases the API lock: keep what the plan's rows mean first
    const std::vector<int> row_slot = pb.row_slot;
    if (pb.nrows) VQC_TRY(vqc_fetch_unlocked(rowsum, pb.nrows, rows.data()));
    else VQC_HIP(hipStreamSynchronize(st));
    finish_rows(rows, row_slot, nparams, energy2, grad);
    return VQC_OK;
}

extern "C" int vqc_dm_energy_grad(const vqc_state* rho0, int nops, const vqc_op_t* ops, const double* pool, int nterms,
                                  const vqc_term_t* terms, const double* hpool, int nparams, double* energy2, double* grad,
                                  vqc_state* rho_out, int flags) {
    VQC_API_LOCK;
    return dm_adjoint(rho0, nops, ops, pool, nterms, terms, hpool, nparams, energy2, grad, rho_out, flags, nullptr, nullptr);
}

// ------------------------------- plan checks --------------------------------
// Invariants of a finished plan on nb index bits over a pool of plen complex
// entries; the first violation is described in *why.
static bool plan_invariants(const PlanEntry& e, int nb, int nterms, int nparams, std::string* why) {
    const PlanBuilder& pb = e.pb;
    const int64_t plen = e.xp.size();
    char buf[256];
    auto fail = [&](const char* fmt, size_t i, int o, long long a, long long b) {
        std::snprintf(buf, sizeof buf, fmt, i, o, a, b);
        *why = buf;
        return false;
    };
    const bool small = e.info[0] == 1 && pb.phases.size() && pb.phases[0].t == nb && nb <= kMaxTileBits2;
    if (int(pb.row_slot.size()) != pb.nrows) return fail("phase %zu op %d: %lld row slots for %lld rows", 0, 0, (long long)pb.row_slot.size(), pb.nrows);
    for (size_t r = 0; r < pb.row_slot.size(); ++r) {
        const int sl = pb.row_slot[r];
        if (sl >= nparams || sl < -nterms - 1) return fail("phase %zu op %d: row %lld has slot %lld", 0, 0, (long long)r, sl);
    }
    int prev_end = 0;
    for (size_t i = 0; i < pb.phases.size(); ++i) {
        const PhaseDesc& p = pb.phases[i];
        if (p.op_begin != prev_end || p.op_end < p.op_begin || p.op_end > int(pb.mops.size()))
            return fail("phase %zu op %d: micro-op range [%lld, %lld) is not the next one", i, prev_end, p.op_begin, p.op_end);
        prev_end = p.op_end;
        const int tmax = p.kind == PH_FWD ? kMaxTileBits : kMaxTileBits2;
        if (p.nbits != nb || p.t < 1 || p.t > std::min(nb, tmax)) return fail("phase %zu op %d: %lld tile bits of %lld", i, 0, p.t, p.nbits);
        if (p.m + p.nhb != p.t || p.nob != nb - p.t || p.m < 0) return fail("phase %zu op %d: tile split %lld + %lld", i, 0, p.m, p.nhb);
        std::vector<char> seen(size_t(nb), 0);
        for (int b = 0; b < p.m; ++b) seen[size_t(b)] = 1;
        for (int k = 0; k < p.nhb; ++k) {
            const int b = p.hb[k];
            if (b < p.m || b >= nb || seen[size_t(b)] || (k && b <= p.hb[k - 1])) return fail("phase %zu op %d: high tile bit %lld (%lld)", i, 0, b, k);
            seen[size_t(b)] = 1;
        }
        for (int k = 0; k < p.nob; ++k) {
            const int b = p.ob[k];
            if (b < 0 || b >= nb || seen[size_t(b)] || (k && b <= p.ob[k - 1])) return fail("phase %zu op %d: outside bit %lld (%lld)", i, 0, b, k);
            seen[size_t(b)] = 1;
        }
        if (p.nacc < 0 || p.nacc > kMaxAcc || (!small && p.row_begin + p.nacc > pb.nrows))
            return fail("phase %zu op %d: %lld accumulators from row %lld", i, 0, p.nacc, p.row_begin);
        if (p.pf_n < 0 || p.pf_lo < 0 || p.pf_lo + p.pf_n > plen) return fail("phase %zu op %d: read-ahead range %lld + %lld", i, 0, p.pf_lo, p.pf_n);
        for (int o = p.op_begin; o < p.op_end; ++o) {
            const MicroOp& m = pb.mops[size_t(o)];
            if (m.kind < MK_APPLY || m.kind > MK_HAMS || m.kb < 1 || m.kb > 5) return fail("phase %zu op %d: kind %lld on %lld bits", i, o, m.kind, m.kb);
            if (m.kind == MK_HAMS) {
                // the header of nchain term records (MK_HAM, 1 or 2 bits, no partials of their own)
                if (p.kind != PH_HAM || m.nchain < 2 || o + m.nchain >= p.op_end || m.ng != 1)
                    return fail("phase %zu op %d: merged terms %lld (%lld partials)", i, o, m.nchain, m.ng);
                for (int q = 1; q <= m.nchain; ++q) {
                    const MicroOp& r = pb.mops[size_t(o + q)];
                    if (r.kind != MK_HAM || r.kb < 1 || r.kb > std::min(2, m.kb) || r.ng != 0)
                        return fail("phase %zu op %d: term record of kind %lld on %lld bits", i, o + q, r.kind, r.kb);
                }
            }
            for (int j = 0; j < m.kb; ++j)
                if (m.lpos[j] < 0 || m.lpos[j] >= p.t || (j && m.lpos[j] <= m.lpos[j - 1]))
                    return fail("phase %zu op %d: local bit %lld (%lld)", i, o, m.lpos[j], j);
            if (m.kind == MK_DENSE && p.t - m.kb < 7) return fail("phase %zu op %d: dense block of %lld bits in a %lld-bit tile", i, o, m.kb, p.t);
            const int64_t D2 = int64_t(1) << (2 * m.kb);
            auto in_pool = [&](int64_t off, int64_t n) { return off >= 0 && off + n <= plen; };
            bool ok = true;
            switch (m.kind) {
                case MK_CHAIN:
                case MK_BLOCK:
                    for (int q = 0; q < m.nchain && q < 4; ++q) {
                        const int64_t sz = m.kind == MK_CHAIN ? 4 : 16;
                        if (m.cmat[q] >= 0 && !in_pool(m.cmat[q], m.kind == MK_BLOCK && ((m.smask >> (2 * q)) & 3) != BS_BOTH ? 4 : sz)) ok = false;
                        if (m.gen[q] >= 0 && !in_pool(m.gen[q], 4)) ok = false;
                    }
                    break;
                case MK_GRAD:
                    for (int q = 0; q < m.ng; ++q) ok = ok && in_pool(m.gen[q], D2);
                    break;
                case MK_HAMS:  // no matrix of its own
                    break;
                default:
                    ok = in_pool(m.mat, D2) && (m.mat2 < 0 || in_pool(m.mat2, D2));
                    if (m.kind == MK_APPLY2)
                        for (int q = 0; q < m.ng; ++q) ok = ok && in_pool(m.gen[q], D2);
                    break;
            }
            if (!ok) return fail("phase %zu op %d: a matrix of kind %lld lies outside the pool of %lld entries", i, o, m.kind, plen);
            if (m.ng < 0 || m.ng > 4) return fail("phase %zu op %d: %lld partials (%lld)", i, o, m.ng, 0);
            for (int q = 0; q < m.ng; ++q) {
                const int a = m.acc[q];
                if (a < 0 || (small ? a >= pb.nrows : a >= p.nacc)) return fail("phase %zu op %d: accumulator %lld of %lld", i, o, a, small ? pb.nrows : p.nacc);
            }
        }
    }
    if (prev_end != int(pb.mops.size())) return fail("phase %zu op %d: %lld micro-ops past the last phase", pb.phases.size(), prev_end, (long long)pb.mops.size(), 0);
    return true;
}

// Host only: plans the fused energy + gradient of a program as
// vqc_sv_energy_grad (is_dm = 0) or vqc_dm_energy_grad (is_dm = 1) would for an
// n-qubit state, checks the plan's invariants (contiguous phases, tile bits that
// partition the index bits, micro-op bits inside their tile, every matrix inside
// the pool, accumulators inside their rows) and returns its shape. It makes no
// device call and leaves the plan cache alone, so the planner can be exercised
// on a machine without a GPU, and under the sanitizers.
extern "C" int vqc_plan_check(int is_dm, int nqubits, int nops, const vqc_op_t* ops, const double* pool, int nterms,
                              const vqc_term_t* terms, const double* hpool, int nparams, int flags, vqc_plan_info_t* out) {
    VQC_API_LOCK;
    if (!out) return vqc_fail(VQC_EINVAL, "null plan info");
    if (nqubits < 1 || nqubits > (is_dm ? 17 : 34)) return vqc_fail(VQC_EINVAL, "unsupported state size: %d qubits", nqubits);
    if (nparams < 0) return vqc_fail(VQC_EINVAL, "bad parameter count");
    const int nb = is_dm ? 2 * nqubits : nqubits;
    int64_t plen = 0, hlen = 0;
    VQC_TRY(check_program(nqubits, is_dm, nops, ops, &plen, nparams));
    VQC_TRY(check_terms(nqubits, nterms, terms, &hlen));
    if ((plen && !pool) || (hlen && !hpool)) return vqc_fail(VQC_EINVAL, "null matrix pool");
    const CkptCut cut = ckpt_cut(nops, ops, pool, is_dm, flags);
    std::unique_ptr<PlanEntry> e = plan_make(std::string(), pool, plen, hpool, hlen);
    if (is_dm) VQC_TRY(plan_dm(*e, nqubits, nops, ops, nterms, cut));
    else VQC_TRY(plan_sv(*e, nb, nops, ops, nterms, terms, plen, hlen, true, cut));
    std::string why;
    if (!plan_invariants(*e, nb, nterms, nparams, &why)) return vqc_fail(VQC_EINTERNAL, "plan check: %s", why.c_str());
    memset(out, 0, sizeof *out);
    for (const PhaseDesc& p : e->pb.phases) {
        if (p.kind == PH_FWD) ++out->phases_fwd;
        else if (p.kind == PH_HAM) ++out->phases_ham;
        else ++out->phases_rev;
    }
    for (const MicroOp& m : e->pb.mops)
        if (m.kind >= 0 && m.kind < 8) ++out->mops[m.kind];
    return VQC_OK;
}

extern "C" int vqc_circuit_vjp(const vqc_state* x, int nops, const vqc_op_t* ops, const double* pool, int nparams,
                               const vqc_state* delta_y, double* grad, vqc_state* delta_x, vqc_state* y_out, int flags) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_state(x, "vqc_circuit_vjp"));
    VQC_TRY(vqc_check_state(delta_y, "vqc_circuit_vjp"));
    if (delta_y->nbits != x->nbits || delta_y->is_dm != x->is_dm || (delta_x && (delta_x->nbits != x->nbits || delta_x->is_dm != x->is_dm)))
        return vqc_fail(VQC_EINVAL, "vqc_circuit_vjp: delta_y and delta_x must have the shape of x");
    double e2[2];
    if (x->is_dm) return dm_adjoint(x, nops, ops, pool, 0, nullptr, nullptr, nparams, e2, grad, y_out, flags, delta_y, delta_x);
    return sv_adjoint(x, nops, ops, pool, 0, nullptr, nullptr, nparams, e2, grad, y_out, flags, delta_y, delta_x);
}

// ------------------------- one evaluation on a caller's stream ---------------
// The tiled energy + gradient path of vqc_sv_energy_grad for |0..0>, with every
// device buffer in one caller-owned block and every launch on a caller's
// stream, so that several evaluations can be in flight at once (serving loop).
struct vqc_solo {
    XPool xp;
    PlanBuilder pb;
    size_t nfwd = 0;
    int nbits = 0, nparams = 0, nrows = 0;
    static constexpr int kStride = 1024;
    size_t o_A = 0, o_B = 0, o_pool = 0, o_mops = 0, o_phases = 0, o_parts = 0, o_rowsum = 0, bytes = 0;
};

int vqc_solo_build(int n, int nops, const vqc_op_t* ops, const double* pool, int64_t pool_len, int nterms,
                   const vqc_term_t* terms, const double* hpool, int64_t hpool_len, int nparams, vqc_solo** out) {
    if (!out || n < 1 || n > 34 || nops < 0 || nterms < 0 || nparams < 0) return vqc_fail(VQC_EINVAL, "vqc_solo: bad arguments");
    int64_t plen = 0, hlen = 0;
    VQC_TRY(check_program(n, 0, nops, ops, &plen, nparams));
    VQC_TRY(check_terms(n, nterms, terms, &hlen));
    if (plen > pool_len || hlen > hpool_len) return vqc_fail(VQC_EINVAL, "vqc_solo: matrix offsets outside the pools");
    if ((plen && !pool) || (hlen && !hpool)) return vqc_fail(VQC_EINVAL, "vqc_solo: null matrix pool");
    const TileTuning tt = vqc_tile_tuning(n);
    vqc_solo* s = new vqc_solo;
    s->nbits = n;
    s->nparams = nparams;
    s->xp.h.assign(pool, pool + 2 * plen);
    s->xp.h.insert(s->xp.h.end(), hpool, hpool + 2 * hlen);
    PlanBuilder& pb = s->pb;
    pb.xp = &s->xp;
    int rc = vqc_plan_forward(pb, n, n, 0, nops, ops, tt.t_fwd, tt.mlow);
    if (rc == VQC_OK && pb.phases.empty()) {  // no ops: one empty phase writes |0..0> into A
        PhaseDesc p;
        vqc_make_phase(p, PH_FWD, n, tt.t_fwd, {});
        p.op_begin = p.op_end = 0;
        p.row_begin = pb.nrows;
        pb.phases.push_back(p);
    }
    if (rc == VQC_OK) {
        pb.phases[0].flags |= PF_INIT_ZERO;
        s->nfwd = pb.phases.size();
        rc = vqc_plan_ham(pb, n, nterms, terms, plen, tt.t_rev, tt.mlow);
    }
    if (rc == VQC_OK) rc = vqc_plan_reverse(pb, n, n, 0, nops, ops, tt.t_rev, tt.mlow);
    if (rc != VQC_OK) {
        delete s;
        return rc;
    }
    vqc_fuse_chains(pb, 0, pb.phases.size());
    s->nrows = pb.nrows;
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        const size_t o = off;
        off += (std::max<size_t>(bytes, 16) + 255) & ~size_t(255);
        return o;
    };
    const size_t L = size_t(1) << n;
    s->o_A = carve(L * 16);
    s->o_B = carve(L * 16);
    s->o_pool = carve(s->xp.h.size() * 8);
    s->o_mops = carve(pb.mops.size() * sizeof(MicroOp));
    s->o_phases = carve(pb.phases.size() * sizeof(PhaseDesc));
    s->o_parts = carve(size_t(std::max(s->nrows, 1)) * vqc_solo::kStride * 16);
    s->o_rowsum = carve(size_t(s->nrows + 1) * 16);
    s->bytes = off;
    *out = s;
    return VQC_OK;
}

size_t vqc_solo_device_bytes(const vqc_solo* s) { return s->bytes; }
size_t vqc_solo_rows_bytes(const vqc_solo* s) { return size_t(s->nrows + 1) * 16; }

int vqc_solo_enqueue(vqc_solo* s, void* block, hipStream_t st, double* h_rows) {
    char* base = (char*)block;
    double2* A = (double2*)(base + s->o_A);
    double2* B = (double2*)(base + s->o_B);
    const double2* dp = (const double2*)(base + s->o_pool);
    const MicroOp* dm = (const MicroOp*)(base + s->o_mops);
    const PhaseDesc* dpd = (const PhaseDesc*)(base + s->o_phases);
    double2* parts = (double2*)(base + s->o_parts);
    double2* rowsum = (double2*)(base + s->o_rowsum);
    const PlanBuilder& pb = s->pb;
    if (!s->xp.h.empty())
        VQC_HIP(hipMemcpyAsync(base + s->o_pool, s->xp.h.data(), s->xp.h.size() * 8, hipMemcpyHostToDevice, st));
    if (!pb.mops.empty())
        VQC_HIP(hipMemcpyAsync(base + s->o_mops, pb.mops.data(), pb.mops.size() * sizeof(MicroOp), hipMemcpyHostToDevice, st));
    VQC_HIP(hipMemcpyAsync(base + s->o_phases, pb.phases.data(), pb.phases.size() * sizeof(PhaseDesc), hipMemcpyHostToDevice, st));
    const int nrows = std::max(s->nrows, 1);
    vqc_zero(parts, size_t(nrows) * vqc_solo::kStride, st);
    VQC_TRY(run_forward(pb, 0, s->nfwd, nullptr, A, dm, dp, dpd, st));
    if (pb.phases.size() == s->nfwd || pb.phases[s->nfwd].kind != PH_HAM) vqc_zero(B, size_t(1) << s->nbits, st);
    for (size_t i = s->nfwd; i < pb.phases.size(); ++i) {
        PhaseBufs b{A, A, B, dm, dp, parts, vqc_solo::kStride};
        VQC_TRY(launch_phase(pb, i, dpd, b, phase_grid(pb.phases[i], vqc_solo::kStride), st));
    }
    if (s->nrows) vqc_reduce_partials(parts, s->nrows, vqc_solo::kStride, rowsum, st);
    VQC_HIP(hipGetLastError());
    VQC_HIP(hipMemcpyAsync(h_rows, rowsum, size_t(s->nrows + 1) * 16, hipMemcpyDeviceToHost, st));
    return VQC_OK;
}

int vqc_solo_collect(const vqc_solo* s, const double* h_rows, double* energy2, double* grad) {
    std::vector<double> rows(h_rows, h_rows + 2 * size_t(std::max(s->nrows, 1)));
    std::vector<double> g(size_t(std::max(s->nparams, 1)));
    finish_rows(rows, s->pb.row_slot, s->nparams, energy2, g.data());
    if (grad && s->nparams) memcpy(grad, g.data(), size_t(s->nparams) * sizeof(double));
    return VQC_OK;
}

void vqc_solo_free(vqc_solo* s) { delete s; }
