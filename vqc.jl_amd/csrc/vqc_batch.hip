// vqc_batch.hip -- many independent energy + adjoint-gradient evaluations in
// one persistent kernel, replayed as a HIP graph.
//
// Planning (once per batch shape):
//   * an instance whose state fits one tile (n <= t_small) becomes ONE work
//     item: the whole forward / Hamiltonian / reverse evaluation in LDS;
//   * a larger instance becomes a chain of phases (vqc_tile.h); each phase is
//     split into tile chunks, one work item per chunk; the items of phase p of
//     an instance may start only when every item of its phase p-1 is done.
// Items are put in one queue in the order of a list-scheduling simulation
// (longest remaining chain first, ready items only), so every item depends
// only on items dequeued before it: with workgroups taking items in queue
// order, the earliest unfinished item never waits, and the kernel drains.
// Matrix variants (U^dagger, conj U, U^T, M + M^dagger) are computed on the
// device from the caller's pools by a small kernel at the start of each run.
//
// Run = graph of: pool upload (pinned -> HBM), variant kernel, zero fill of
// partials and counters, persistent kernel, row reduction, row download.
#include <cstdlib>
#include <queue>

#include "vqc_batch_internal.h"
#include "vqc_tile.h"

void vqc_zero(double2* v, uint64_t len, hipStream_t st);

constexpr int kMaxChunks = 1024;  // tile chunks (and partial columns) per phase
constexpr int kItemSmall = 0, kItemPhase = 1;

struct BItem {
    int32_t kind;
    int32_t inst;
    int32_t phase;       // kItemPhase: global phase index
    int32_t chunk;       // column of the partial rows
    int32_t need;        // items of this instance that must be done first
    int32_t pad;
    int64_t tile_begin;
    int64_t tile_end;
};

struct BInst {
    int64_t a_off;       // arena offsets (complex elements) of A and B; -1 for small
    int64_t b_off;
    int32_t nbits;
    int32_t mop_begin;   // small: micro-op range and first Hamiltonian micro-op
    int32_t mop_end;
    int32_t ham_begin;
    int32_t rev_begin;
    int32_t row_begin;   // first global partial row
    int32_t nrows;
    int32_t pf_n;        // small: pool range [pf_lo, pf_lo + pf_n) of every micro-op's matrices (0: not read ahead)
    int64_t pf_lo;
};

struct BatchArgs {
    const BItem* items;
    int nitems;
    const BInst* inst;
    const PhaseDesc* phases;
    const MicroOp* mops;
    const double2* pool;
    double2* arena;
    double2* parts;      // row r's partials start at row_off[r] (one per chunk of its phase)
    const int64_t* row_off;
    unsigned* counters;  // [0] queue head, [1] error flag, [2 + i] done items of instance i
};

struct VarEntry {
    int64_t src, dst;
    int32_t D, v;
};

// one 2x2 chain product U_n ... U_1 per entry (its sources are pool matrices or
// variants written by k_variants before)
struct ProdEntry {
    int64_t dst;
    int64_t srcs[4];
    int32_t n, pad;
};

__global__ void __launch_bounds__(256) k_products(double2* pool, const ProdEntry* tab, int n) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const ProdEntry t = tab[e];
    double2 p00 = pool[t.srcs[0]], p10 = pool[t.srcs[0] + 1], p01 = pool[t.srcs[0] + 2], p11 = pool[t.srcs[0] + 3];
    for (int f = 1; f < t.n; ++f) {
        const double2* U = pool + t.srcs[f];
        const double2 u00 = U[0], u10 = U[1], u01 = U[2], u11 = U[3];
        const double2 q00 = cfma(u01, p10, cmul(u00, p00)), q10 = cfma(u11, p10, cmul(u10, p00));
        const double2 q01 = cfma(u01, p11, cmul(u00, p01)), q11 = cfma(u11, p11, cmul(u10, p01));
        p00 = q00;
        p10 = q10;
        p01 = q01;
        p11 = q11;
    }
    pool[t.dst] = p00;
    pool[t.dst + 1] = p10;
    pool[t.dst + 2] = p01;
    pool[t.dst + 3] = p11;
}

__global__ void __launch_bounds__(256) k_variants(double2* pool, const VarEntry* tab, int n) {
    for (int e = blockIdx.x; e < n; e += gridDim.x) {
        const VarEntry t = tab[e];
        const int D = t.D;
        for (int i = threadIdx.x; i < D * D; i += blockDim.x) {
            const int r = i % D, c = i / D;
            const double2 m = pool[t.src + r + c * D];
            const double2 mt = pool[t.src + c + r * D];
            double2 o;
            if (t.v == 4) o = make_double2(m.x + mt.x, m.y - mt.y);  // M + M^dagger
            else {
                o = (t.v & MF_TRANS) ? mt : m;
                if (t.v & MF_CONJ) o.y = -o.y;
            }
            pool[t.dst + i] = o;
        }
    }
}

__global__ void __launch_bounds__(256) k_zero_u32(unsigned* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// Bounded wait: the queue order makes every wait finite; the bound (5 s of the
// 100 MHz s_memrealtime clock) only turns a planning bug into a reported error.
constexpr uint64_t kWaitTicks = 500000000ull;

template <int MAXKB, int RKS>
__device__ __forceinline__ void batch_loop(const BatchArgs& a, int& s_item, int& s_ok) {
    extern __shared__ double2 smem[];
    // Every branch of this loop is wave-uniform. The queue head and the done
    // counters are touched by all 64 lanes of wave 0 (lane 0 adds 1, the others
    // add 0) instead of by thread 0 alone: a lane-divergent region next to the
    // loop's back-edge lets the compiler rotate the loop so that lane 0 (the
    // only lane that would dequeue and signal) leaves it while the other lanes
    // keep re-running the body with a stale item.
    const bool wave0 = (__builtin_amdgcn_readfirstlane(threadIdx.x) >> 6) == 0;
    // items, instances, phases, micro-ops, the pool and the row offsets are
    // read-only for the whole launch (the pool's variants and products are made
    // by the kernels before this one): constant memory, read with scalar loads
    const MicroOp* const mops = (const MicroOp*)as_const(a.mops);
    const double2* const pool = (const double2*)as_const(a.pool);
    const int64_t* const row_off = (const int64_t*)as_const(a.row_off);
    const BItem* const items = (const BItem*)as_const(a.items);
    const BInst* const insts = (const BInst*)as_const(a.inst);
    const PhaseDesc* const phases = (const PhaseDesc*)as_const(a.phases);
    const unsigned one = (threadIdx.x & 63) == 0 ? 1u : 0u;
    for (;;) {
        if (wave0) s_item = int(__builtin_amdgcn_readfirstlane(atomicAdd(&a.counters[0], one)));
        __syncthreads();
        const int idx = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (idx >= a.nitems) break;
        const BItem it = items[idx];
        if (wave0) {
            int ok = 1;
            const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
            // The polls are relaxed: an acquire load invalidates this XCD's L2 on
            // every poll, and hundreds of waiting workgroups polling that way kept
            // the L2 of every XCD cold for the workgroups doing the work. One
            // acquire fence after the wait orders this item after its producers.
            for (;;) {
                const unsigned done = __builtin_amdgcn_readfirstlane(
                    __hip_atomic_load(&a.counters[2 + it.inst], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
                if (done >= unsigned(it.need)) break;
                if (__builtin_amdgcn_s_memrealtime() - t0 > kWaitTicks) {
                    ok = 0;
                    atomicOr(&a.counters[1], one);
                    break;
                }
                // back off by the number of items still ahead: many workgroups
                // can wait on one instance, and their polls of its counter must
                // not crowd out the increments of the workgroups that finish
                const unsigned rem = unsigned(it.need) - done;
                if (rem > 32) {
                    __builtin_amdgcn_s_sleep(127);
                    __builtin_amdgcn_s_sleep(127);
                    __builtin_amdgcn_s_sleep(127);
                    __builtin_amdgcn_s_sleep(127);
                } else if (rem > 4) {
                    __builtin_amdgcn_s_sleep(127);
                } else {
                    __builtin_amdgcn_s_sleep(16);
                }
            }
            // one wave's acquire invalidates the CU's L1 and this XCD's stale L2
            // lines for the whole workgroup; the barrier below orders the rest
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            s_ok = ok;
        }
        __syncthreads();
        const int ok = __builtin_amdgcn_readfirstlane(s_ok);
        if (ok) {
            const BInst in = insts[it.inst];
            if (it.kind == kItemSmall) {
                const uint32_t T = 1u << in.nbits;
                warm_ops(mops, in.mop_begin, in.mop_end, pool, in.pf_lo, in.pf_n, smem + 2 * T);
                small_eval_body<MAXKB, RKS, KS_HAM1>(in.nbits, nullptr, nullptr, mops, in.mop_begin, in.ham_begin, in.rev_begin, in.mop_end, pool, smem,
                                smem + T, smem + 2 * T, smem + 2 * T + kRedDoubles2, in.nrows,
                                a.parts + row_off[in.row_begin], 1);
            } else {
                const PhaseDesc& p = phases[it.phase];
                const uint32_t T = 1u << p.t;
                double2* sa = smem;
                double2* sb = p.kind == PH_FWD ? smem : smem + T;
                double2* red = p.kind == PH_FWD ? smem + T : smem + 2 * T;
                double2* gacc = red + kRedDoubles2;
                warm_ops(mops, p.op_begin, p.op_end, pool, p.pf_lo, p.pf_n, red);
                for (int i = threadIdx.x; i < p.nacc; i += kTileThreads) gacc[i] = make_double2(0.0, 0.0);
                __syncthreads();
                AccState st;
                acc_init(st);
                const uint64_t tdep = thread_deposit(p);
                double2* A = a.arena + in.a_off;
                double2* B = a.arena + in.b_off;
                if (p.kind == PH_FWD) {
                    for (int64_t tile = it.tile_begin; tile < it.tile_end; ++tile)
                        phase_tile<MAXKB, KS_FWDH>(p, uint64_t(tile), tdep, A, A, B, mops, pool, sa, sb, red, gacc, st);
                } else if (p.kind == PH_HAM) {
                    for (int64_t tile = it.tile_begin; tile < it.tile_end; ++tile)
                        phase_tile<MAXKB, KS_HAM1>(p, uint64_t(tile), tdep, A, A, B, mops, pool, sa, sb, red, gacc, st);
                } else {
                    for (int64_t tile = it.tile_begin; tile < it.tile_end; ++tile)
                        phase_tile<MAXKB, RKS>(p, uint64_t(tile), tdep, A, A, B, mops, pool, sa, sb, red, gacc, st);
                }
                acc_finish(st, red, gacc);
                for (int e = threadIdx.x; e < p.nacc; e += kTileThreads)
                    a.parts[row_off[p.row_begin + e] + it.chunk] = gacc[e];
            }
        }
        // the barrier makes every wave's stores complete in L2; one agent-scope
        // release fence (L2 write-back) by wave 0 then publishes them to other
        // XCDs, and the relaxed increment after it carries that release (a
        // release increment would write the L2 back a second time)
        __syncthreads();
        if (wave0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_fetch_add(&a.counters[2 + it.inst], one, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
    }
}

// One workgroup per CU, any register count (5-qubit micro-ops need them).
template <int MAXKB, int RKS>
__global__ void __launch_bounds__(kTileThreads) k_batch(BatchArgs a) {
    __shared__ int s_item;
    __shared__ int s_ok;
    batch_loop<MAXKB, RKS>(a, s_item, s_ok);
}

// out[r] = sum of row r's partials, in a fixed order
__global__ void __launch_bounds__(256) k_reduce_rows(const double2* __restrict__ parts, const int64_t* __restrict__ off,
                                                     const int32_t* __restrict__ cnt, double2* out) {
    __shared__ double red[4];
    const double2* p = parts + off[blockIdx.x];
    const int n = cnt[blockIdx.x];
    double sr = 0, si = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        sr += p[i].x;
        si += p[i].y;
    }
    sr = block_sum(sr, red);
    si = block_sum(si, red);
    if (threadIdx.x == 0) out[blockIdx.x] = make_double2(sr, si);
}

// rows[nrows].x = the error flag, so one download brings results and status
__global__ void k_finish(const unsigned* counters, double2* flag_slot) {
    if (threadIdx.x == 0) *flag_slot = make_double2(double(counters[1]), 0.0);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct vqc_batch {
    int ninst = 0;
    std::vector<vqc_instance_t> inst;
    std::vector<int> row_slot;        // global row -> local gradient slot (>= 0) or energy (< 0)
    std::vector<int> row_inst;        // global row -> instance
    int nrows = 0;
    int nitems = 0;
    int64_t pool_len = 0, hpool_len = 0, xlen = 0;
    int64_t grad_total = 0;
    size_t lds = 0;
    unsigned grid = 0;
    int maxkb = 1;
    bool rev_simple = true;  // every reverse micro-op is MK_APPLY2 or MK_CHAIN
    BatchOpts opts;
    hipStream_t stream = nullptr;
    void* d_block = nullptr;  // every device buffer below is carved from this block
    size_t block_bytes = 0;
    // device
    void* d_items = nullptr;
    void* d_inst = nullptr;
    void* d_phases = nullptr;
    void* d_mops = nullptr;
    void* d_var = nullptr;
    int nvar = 0;
    void* d_prod = nullptr;
    int nprod = 0;
    double2* d_pool = nullptr;
    double2* d_arena = nullptr;
    double2* d_parts = nullptr;
    double2* d_rows = nullptr;
    void* d_row_off = nullptr;
    void* d_row_cnt = nullptr;
    int64_t nparts = 0;
    unsigned* d_counters = nullptr;
    uint64_t arena_bytes = 0;
    // plan shape: tiled and LDS-resident instances, phases, and the HBM bytes the
    // phases move per launch (each phase reads and writes its states once)
    int ntiled = 0, nsmall = 0;
    int64_t nphases = 0;
    uint64_t hbm_bytes = 0;
    // host pinned staging
    double* h_pool = nullptr;
    double* h_rows = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool launched = false;
};

typedef void (*BatchKernel)(BatchArgs);
static BatchKernel batch_kernel(const vqc_batch* b) {
    // One workgroup per CU and no register cap: capped at 128 VGPRs (two workgroups
    // per CU where their LDS fits) the kernel spills, and it measured no faster on
    // the serving loop or on 256 instances of 8 to 11 qubits (within 5%).
    if (b->maxkb <= 2) return b->rev_simple ? k_batch<2, KS_REV> : k_batch<2, KS_REVX>;
    if (b->maxkb <= 3) return b->rev_simple ? k_batch<3, KS_REV> : k_batch<3, KS_REVX>;
    return b->rev_simple ? k_batch<5, KS_REV> : k_batch<5, KS_REVX>;
}

static int batch_enqueue(vqc_batch* b, hipStream_t st, const double* h_pool, double* h_rows) {
    BatchArgs a{(const BItem*)b->d_items, b->nitems, (const BInst*)b->d_inst, (const PhaseDesc*)b->d_phases,
                (const MicroOp*)b->d_mops, b->d_pool, b->d_arena, b->d_parts, (const int64_t*)b->d_row_off, b->d_counters};
    if (b->pool_len + b->hpool_len)
        VQC_HIP(hipMemcpyAsync(b->d_pool, h_pool, size_t(b->pool_len + b->hpool_len) * 16, hipMemcpyHostToDevice, st));
    if (b->nvar) hipLaunchKernelGGL(k_variants, dim3(std::min(b->nvar, 4096)), dim3(256), 0, st, b->d_pool, (const VarEntry*)b->d_var, b->nvar);
    if (b->nprod)
        hipLaunchKernelGGL(k_products, dim3((b->nprod + 255) / 256), dim3(256), 0, st, b->d_pool, (const ProdEntry*)b->d_prod, b->nprod);
    vqc_zero(b->d_parts, uint64_t(std::max<int64_t>(b->nparts, 1)), st);
    hipLaunchKernelGGL(k_zero_u32, dim3(std::max(1, (b->ninst + 2 + 255) / 256)), dim3(256), 0, st, b->d_counters, b->ninst + 2);
    hipLaunchKernelGGL(batch_kernel(b), dim3(b->grid), dim3(kTileThreads), b->lds, st, a);
    if (vqc_trace_on())
        vqc_trace("k_batch maxkb %d instances %d items %d grid %u lds %zu", int(b->maxkb), int(b->ninst), int(b->nitems), unsigned(b->grid), size_t(b->lds));
    if (b->nrows)
        hipLaunchKernelGGL(k_reduce_rows, dim3(b->nrows), dim3(256), 0, st, b->d_parts, (const int64_t*)b->d_row_off,
                           (const int32_t*)b->d_row_cnt, b->d_rows);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(64), 0, st, b->d_counters, b->d_rows + b->nrows);
    VQC_HIP(hipMemcpyAsync(h_rows, b->d_rows, size_t(b->nrows + 1) * 16, hipMemcpyDeviceToHost, st));
    VQC_HIP(hipGetLastError());
    return VQC_OK;
}

static int batch_free(vqc_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->d_block) {
        if (b->opts.dev_free) b->opts.dev_free(b->opts.alloc_ctx, b->d_block);
        else (void)hipFree(b->d_block);
    }
    if (b->h_pool) (void)hipHostFree(b->h_pool);
    if (b->h_rows) (void)hipHostFree(b->h_rows);
    delete b;
    return VQC_OK;
}

int vqc_batch_release(vqc_batch* b) { return b ? batch_free(b) : VQC_OK; }

void vqc_batch_sizes(const vqc_batch* b, size_t* device_bytes, size_t* pool_staging_bytes, size_t* rows_staging_bytes) {
    if (device_bytes) *device_bytes = b->block_bytes;
    if (pool_staging_bytes) *pool_staging_bytes = size_t(std::max<int64_t>(b->pool_len + b->hpool_len, 1)) * 16;
    if (rows_staging_bytes) *rows_staging_bytes = size_t(b->nrows + 1) * 16;
}

extern "C" int vqc_batch_create(int ninst, const vqc_instance_t* inst, const vqc_op_t* ops, int64_t nops_total,
                                const vqc_term_t* terms, int64_t nterms_total, int64_t pool_len, int64_t hpool_len,
                                int flags, vqc_batch** out) {
    (void)flags;
    BatchOpts o;
    o.graph = getenv("VQC_NO_GRAPH") == nullptr;
    return vqc_batch_build(ninst, inst, ops, nops_total, terms, nterms_total, pool_len, hpool_len, o, out);
}

int vqc_batch_build(int ninst, const vqc_instance_t* inst, const vqc_op_t* ops, int64_t nops_total,
                    const vqc_term_t* terms, int64_t nterms_total, int64_t pool_len, int64_t hpool_len,
                    const BatchOpts& opts, vqc_batch** out) {
    if (!out || ninst < 1 || !inst || nops_total < 0 || nterms_total < 0 || pool_len < 0 || hpool_len < 0)
        return vqc_fail(VQC_EINVAL, "vqc_batch_create: bad arguments");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    // A wide batch fills the device with its instances and takes the batch tile
    // shape; a narrow one (the serving loop's usual single request) gives each
    // tiled instance the shape of a lone state of its size. Instances of up to
    // tt.t_small qubits run in one workgroup either way.
    const TileTuning tt = vqc_tile_tuning();
    const bool wide = ninst >= kWideBatch;
    vqc_batch* b = new vqc_batch;
    b->opts = opts;
    b->stream = opts.stream ? opts.stream : c->stream;
    b->ninst = ninst;
    b->inst.assign(inst, inst + ninst);
    b->pool_len = pool_len;
    b->hpool_len = hpool_len;
    XPool xp;
    xp.deferred = true;
    xp.len = pool_len + hpool_len;
    PlanBuilder pb;  // shared micro-op and phase arrays
    pb.xp = &xp;
    std::vector<BInst> binst(ninst);
    struct Chain {
        int first_phase, nphases;
        bool small;
        double cost;
    };
    std::vector<Chain> chains(ninst);
    int64_t arena = 0;
    auto fail = [&](int code, const char* msg, int i) {
        batch_free(b);
        return vqc_fail(code, "vqc_batch_create: instance %d: %s", i, msg);
    };
    for (int i = 0; i < ninst; ++i) {
        const vqc_instance_t& I = inst[i];
        const int n = I.nqubits;
        if (n < 1 || n > 30 || I.nops < 0 || I.nterms < 0 || I.nparams < 0 || I.op_begin < 0 || I.term_begin < 0 ||
            I.op_begin + I.nops > nops_total || I.term_begin + I.nterms > nterms_total || I.grad_begin < 0)
            return fail(VQC_EINVAL, "bad ranges", i);
        const vqc_op_t* iops = ops + I.op_begin;
        const vqc_term_t* iterms = terms + I.term_begin;
        for (int o = 0; o < I.nops; ++o) {
            const vqc_op_t& op = iops[o];
            if (op.kind != VQC_OP_UNITARY || vqc_check_positions(n, op.nq, op.pos, 5) != VQC_OK)
                return fail(VQC_EINVAL, "bad op (batched instances are state vectors)", i);
            const int64_t D2 = int64_t(1) << (2 * op.nq);
            if (op.mat < 0 || op.mat + D2 > pool_len) return fail(VQC_EINVAL, "op matrix outside the pool", i);
            if (op.nparams < 0 || op.nparams > 3) return fail(VQC_EINVAL, "an op has more than 3 parameters", i);
            for (int j = 0; j < op.nparams; ++j)
                if (op.slot[j] >= I.nparams || (op.slot[j] >= 0 && (op.gen[j] < 0 || op.gen[j] + D2 > pool_len)))
                    return fail(VQC_EINVAL, "bad parameter slot or generator", i);
        }
        for (int t = 0; t < I.nterms; ++t) {
            if (vqc_check_positions(n, iterms[t].nq, iterms[t].pos, 5) != VQC_OK) return fail(VQC_EINVAL, "bad term", i);
            if (iterms[t].mat < 0 || iterms[t].mat + (int64_t(1) << (2 * iterms[t].nq)) > hpool_len)
                return fail(VQC_EINVAL, "term matrix outside the pool", i);
        }
        b->grad_total = std::max<int64_t>(b->grad_total, I.grad_begin + I.nparams);
        BInst& bi = binst[i];
        bi.nbits = n;
        bi.row_begin = pb.nrows;
        const size_t ph0 = pb.phases.size();
        const int rows0 = pb.nrows;
        const size_t slot0 = pb.row_slot.size();
        bool small = false;
        if (n <= tt.t_small) {
            PlanBuilder sp;
            sp.xp = &xp;
            sp.nrows = 0;
            vqc_plan_forward(sp, n, n, 0, I.nops, iops, n, 0);
            const size_t nf = sp.phases.size();
            vqc_plan_ham(sp, n, I.nterms, iterms, pool_len, n, 0);
            const size_t nh = sp.phases.size();
            vqc_plan_reverse(sp, n, n, 0, I.nops, iops, n, 0);
            vqc_fuse_chains(sp, 0, sp.phases.size());
            const int hb = vqc_phase_op_start(sp, nf);
            const int rb = vqc_phase_op_start(sp, nh);
            const size_t T = size_t(1) << n;
            const size_t need = (2 * T + kRedDoubles2 + size_t(std::max(sp.nrows, 1))) * 16;
            if (need <= 160 * 1024) {
                small = true;
                for (const PhaseDesc& p : sp.phases)
                    for (int o = p.op_begin; o < p.op_end; ++o)
                        for (int q = 0; q < 4; ++q) sp.mops[o].acc[q] += p.row_begin;
                bi.mop_begin = int(pb.mops.size());
                bi.ham_begin = bi.mop_begin + hb;
                bi.rev_begin = bi.mop_begin + rb;
                pb.mops.insert(pb.mops.end(), sp.mops.begin(), sp.mops.end());
                bi.mop_end = int(pb.mops.size());
                vqc_prefetch_range(pb.mops, bi.mop_begin, bi.mop_end, xp.size(), &bi.pf_lo, &bi.pf_n);
                bi.nrows = sp.nrows;
                for (int r : sp.row_slot) pb.row_slot.push_back(r);
                pb.nrows += sp.nrows;
                bi.a_off = bi.b_off = -1;
                b->lds = std::max(b->lds, need);
                chains[i] = {-1, 0, true, double(T) * double(bi.mop_end - bi.mop_begin)};
                b->nsmall += 1;
            }
        }
        if (!small) {
            const TileTuning ti = wide ? tt : vqc_tile_tuning(n);
            // phases: forward (first reads |0..0>), Hamiltonian, reverse; rows are global
            vqc_plan_forward(pb, n, n, 0, I.nops, iops, ti.t_fwd, ti.mlow);
            if (pb.phases.size() == ph0) {  // no ops: one empty forward phase writes |0..0> into A
                PhaseDesc p;
                vqc_make_phase(p, PH_FWD, n, ti.t_fwd, {});
                p.op_begin = p.op_end = int(pb.mops.size());
                p.row_begin = pb.nrows;
                pb.phases.push_back(p);
            }
            pb.phases[ph0].flags |= PF_INIT_ZERO;
            vqc_plan_ham(pb, n, I.nterms, iterms, pool_len, ti.t_rev, ti.mlow);
            vqc_plan_reverse(pb, n, n, 0, I.nops, iops, ti.t_rev, ti.mlow);
            vqc_fuse_chains(pb, ph0, pb.phases.size());
            bi.nrows = pb.nrows - rows0;
            bi.a_off = arena;
            bi.b_off = arena + (int64_t(1) << n);
            arena += int64_t(2) << n;
            bi.mop_begin = bi.mop_end = bi.ham_begin = bi.rev_begin = 0;
            double cost = 0;
            b->ntiled += 1;
            b->nphases += int64_t(pb.phases.size() - ph0);
            for (size_t p = ph0; p < pb.phases.size(); ++p) {
                const PhaseDesc& ph = pb.phases[p];
                const uint64_t sb = uint64_t(16) << n;  // one state, read or written once
                if (ph.kind == PH_FWD) b->hbm_bytes += (ph.flags & PF_INIT_ZERO) ? sb : 2 * sb;
                else if (ph.kind == PH_HAM) b->hbm_bytes += (ph.ham_accum ? 3 : 2) * sb;
                else b->hbm_bytes += 4 * sb;
                const double nb = ph.kind == PH_FWD ? 2 : 4;
                cost += double(int64_t(1) << n) * (nb + 0.25 * (ph.op_end - ph.op_begin));
                b->lds = std::max(b->lds, ((ph.kind == PH_FWD ? 1 : 2) * (size_t(1) << ph.t) + kRedDoubles2 + size_t(std::max(ph.nacc, 1))) * 16);
            }
            chains[i] = {int(ph0), int(pb.phases.size() - ph0), false, cost};
        }
        for (size_t r = slot0; r < pb.row_slot.size(); ++r) b->row_inst.push_back(i);
        (void)rows0;
    }
    b->row_slot = pb.row_slot;
    b->nrows = pb.nrows;
    for (const MicroOp& m : pb.mops) b->maxkb = std::max(b->maxkb, int(m.kb));
    // reverse micro-ops: small instances [rev_begin, mop_end), large instances their PH_REV phases
    for (const BInst& bi : binst)
        for (int o = bi.rev_begin; o < bi.mop_end; ++o) b->rev_simple = b->rev_simple && vqc_rev_simple(pb.mops[o]);
    for (const PhaseDesc& ph : pb.phases)
        if (ph.kind == PH_REV)
            for (int o = ph.op_begin; o < ph.op_end; ++o) b->rev_simple = b->rev_simple && vqc_rev_simple(pb.mops[o]);

    // ---- persistent grid: as many workgroups as can be resident (times grid_frac)
    b->lds = std::max<size_t>(b->lds, 16 * 1024);
    const void* kfn = (const void*)batch_kernel(b);
    VQC_HIP(hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, int(b->lds)));
    int per_cu = 0;
    VQC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, kTileThreads, b->lds));
    if (per_cu < 1) {
        batch_free(b);
        return vqc_fail(VQC_EUNSUPPORTED, "vqc_batch_create: %zu bytes of LDS per workgroup do not fit", b->lds);
    }
    const int64_t resident = int64_t(per_cu) * c->cu_count;
    int64_t grid_cap = std::min<int64_t>(
        kMaxChunks, std::max<int64_t>(1, int64_t(double(resident) * std::min(1.0, std::max(opts.grid_frac, 0.0)))));
    if (const char* e = getenv("VQC_BATCH_GRID")) grid_cap = std::max<int64_t>(1, std::min<int64_t>(grid_cap, atoll(e)));

    // ---- items in list-scheduling order ---------------------------------
    // a phase is cut into one chunk per workgroup of the grid (or per tile when
    // it has fewer), so a phase that fills the chip ends in one round
    std::vector<BItem> items;
    struct PhaseItems {
        std::vector<BItem> v;
        double cost_each;
    };
    // per instance: list of phases, each a list of chunk items (need filled in order)
    std::vector<std::vector<PhaseItems>> plist(ninst);
    std::vector<double> remaining(ninst, 0.0);
    for (int i = 0; i < ninst; ++i) {
        if (chains[i].small) {
            BItem it{kItemSmall, i, -1, 0, 0, 0, 0, 0};
            plist[i].push_back({{it}, chains[i].cost});
            remaining[i] = chains[i].cost;
            continue;
        }
        int need = 0;
        for (int q = 0; q < chains[i].nphases; ++q) {
            const int pg = chains[i].first_phase + q;
            const PhaseDesc& ph = pb.phases[pg];
            const int64_t ntiles = int64_t(1) << (ph.nbits - ph.t);
            const int nch = int(std::min<int64_t>(ntiles, grid_cap));
            PhaseItems pi;
            const double nbuf = ph.kind == PH_FWD ? 2 : 4;
            const double pcost = double(int64_t(1) << ph.nbits) * (nbuf + 0.25 * (ph.op_end - ph.op_begin));
            pi.cost_each = pcost / nch;
            for (int ch = 0; ch < nch; ++ch) {
                BItem it{kItemPhase, i, pg, ch, need, 0, ntiles * ch / nch, ntiles * (ch + 1) / nch};
                pi.v.push_back(it);
            }
            need += nch;
            remaining[i] += pcost;
            plist[i].push_back(pi);
        }
    }
    // no more workgroups than the instances can keep busy at once (the widest
    // phase of each); extra ones would only wait
    int64_t width = 0;
    for (int i = 0; i < ninst; ++i) {
        int64_t w = 1;
        for (const PhaseItems& pi : plist[i]) w = std::max<int64_t>(w, int64_t(pi.v.size()));
        width += w;
    }
    const int workers = int(std::max<int64_t>(1, std::min(grid_cap, width)));
    // simulation: workers free at times; an instance's next phase is ready when all chunks of the previous are done
    std::priority_queue<double, std::vector<double>, std::greater<double>> free_at;
    for (int w = 0; w < workers; ++w) free_at.push(0.0);
    std::vector<size_t> ph_idx(ninst, 0), ch_idx(ninst, 0);
    std::vector<double> ready(ninst, 0.0), ph_end(ninst, 0.0);
    size_t total = 0;
    for (int i = 0; i < ninst; ++i)
        for (auto& p : plist[i]) total += p.v.size();
    items.reserve(total);
    while (items.size() < total) {
        double t = free_at.top();
        free_at.pop();
        int best = -1;
        double next_ready = 1e300;
        for (int i = 0; i < ninst; ++i) {
            if (ph_idx[i] >= plist[i].size()) continue;
            if (ready[i] <= t) {
                if (best < 0 || remaining[i] > remaining[best]) best = i;
            } else {
                next_ready = std::min(next_ready, ready[i]);
            }
        }
        if (best < 0) {  // idle until the next phase becomes ready
            free_at.push(next_ready);
            continue;
        }
        PhaseItems& pi = plist[best][ph_idx[best]];
        items.push_back(pi.v[ch_idx[best]]);
        const double end = t + pi.cost_each;
        free_at.push(end);
        ph_end[best] = std::max(ph_end[best], end);
        remaining[best] -= pi.cost_each;
        if (++ch_idx[best] == pi.v.size()) {
            ch_idx[best] = 0;
            ++ph_idx[best];
            ready[best] = ph_end[best];
        }
    }
    // every item must depend only on items of its instance queued before it
    {
        std::vector<int> seen(ninst, 0);
        for (const BItem& it : items) {
            if (it.need > seen[it.inst]) {
                batch_free(b);
                return vqc_fail(VQC_EINVAL, "vqc_batch_create: internal queue order error");
            }
            ++seen[it.inst];
        }
    }
    b->nitems = int(items.size());

    // ---- partial layout: one slot per (row, chunk of the row's phase) -----
    std::vector<int32_t> row_cnt(size_t(std::max(b->nrows, 1)), 1);
    for (int i = 0; i < ninst; ++i) {
        if (chains[i].small) continue;
        for (int q = 0; q < chains[i].nphases; ++q) {
            const PhaseDesc& ph = pb.phases[chains[i].first_phase + q];
            const int nch = int(std::min<int64_t>(int64_t(1) << (ph.nbits - ph.t), grid_cap));
            for (int e = 0; e < ph.nacc; ++e) row_cnt[ph.row_begin + e] = nch;
        }
    }
    std::vector<int64_t> row_off(row_cnt.size());
    int64_t np = 0;
    for (size_t r = 0; r < row_cnt.size(); ++r) {
        row_off[r] = np;
        np += row_cnt[r];
    }
    b->nparts = np;

    // ---- device buffers -------------------------------------------------
    b->xlen = xp.size();
    b->arena_bytes = uint64_t(arena) * 16;
    // variants first (one launch), then the chain products that may read them
    std::vector<VarEntry> vt;
    std::vector<ProdEntry> pt;
    for (const XPoolEntry& e : xp.table) {
        if (e.v == 8) {
            ProdEntry q{e.dst, {-1, -1, -1, -1}, e.nsrc, 0};
            for (int j = 0; j < e.nsrc; ++j) q.srcs[j] = e.srcs[j];
            pt.push_back(q);
        } else {
            vt.push_back({e.src, e.dst, e.D, e.v});
        }
    }
    b->nvar = int(vt.size());
    b->nprod = int(pt.size());
    // one device block, carved into 256-byte aligned buffers
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        const size_t o = off;
        off += (std::max<size_t>(bytes, 16) + 255) & ~size_t(255);
        return o;
    };
    const size_t o_items = carve(items.size() * sizeof(BItem));
    const size_t o_inst = carve(binst.size() * sizeof(BInst));
    const size_t o_phases = carve(pb.phases.size() * sizeof(PhaseDesc));
    const size_t o_mops = carve(pb.mops.size() * sizeof(MicroOp));
    const size_t o_var = carve(vt.size() * sizeof(VarEntry));
    const size_t o_prod = carve(pt.size() * sizeof(ProdEntry));
    const size_t o_pool = carve(size_t(b->xlen) * 16);
    const size_t o_arena = carve(size_t(arena) * 16);
    const size_t o_parts = carve(size_t(std::max<int64_t>(b->nparts, 1)) * 16);
    const size_t o_row_off = carve(row_off.size() * sizeof(int64_t));
    const size_t o_row_cnt = carve(row_cnt.size() * sizeof(int32_t));
    const size_t o_rows = carve(size_t(b->nrows + 1) * 16);
    const size_t o_counters = carve(size_t(ninst + 2) * 4);
    b->block_bytes = off;
    hipError_t e = hipSuccess;
    if (opts.dev_alloc) {
        b->d_block = opts.dev_alloc(opts.alloc_ctx, off);
        if (!b->d_block) e = hipErrorOutOfMemory;
    } else {
        e = hipMalloc(&b->d_block, off);
    }
    if (e == hipSuccess) {
        char* base = (char*)b->d_block;
        b->d_items = base + o_items;
        b->d_inst = base + o_inst;
        b->d_phases = base + o_phases;
        b->d_mops = base + o_mops;
        b->d_var = base + o_var;
        b->d_prod = base + o_prod;
        b->d_pool = (double2*)(base + o_pool);
        b->d_arena = (double2*)(base + o_arena);
        b->d_parts = (double2*)(base + o_parts);
        b->d_row_off = base + o_row_off;
        b->d_row_cnt = base + o_row_cnt;
        b->d_rows = (double2*)(base + o_rows);
        b->d_counters = (unsigned*)(base + o_counters);
    } else {
        b->d_block = nullptr;
    }
    if (e == hipSuccess && opts.pinned) {
        e = hipHostMalloc((void**)&b->h_pool, size_t(std::max<int64_t>(pool_len + hpool_len, 1)) * 16, hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&b->h_rows, size_t(b->nrows + 1) * 16, hipHostMallocDefault);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        batch_free(b);
        return vqc_fail(VQC_ENOMEM, "vqc_batch_create: device allocation of %zu bytes failed (%s)", off, hipGetErrorString(e));
    }
    // the batch kernels carry a micro-op per Hamiltonian term only (KS_HAM1)
    for (const MicroOp& m : pb.mops)
        if (m.kind == MK_HAMS) {
            batch_free(b);
            return vqc_fail(VQC_EINTERNAL, "vqc_batch_create: merged Hamiltonian terms in a batch plan");
        }
    hipStream_t st = b->stream;
    VQC_HIP(hipMemcpyAsync(b->d_items, items.data(), items.size() * sizeof(BItem), hipMemcpyHostToDevice, st));
    VQC_HIP(hipMemcpyAsync(b->d_inst, binst.data(), binst.size() * sizeof(BInst), hipMemcpyHostToDevice, st));
    if (!pb.phases.empty())
        VQC_HIP(hipMemcpyAsync(b->d_phases, pb.phases.data(), pb.phases.size() * sizeof(PhaseDesc), hipMemcpyHostToDevice, st));
    if (!pb.mops.empty())
        VQC_HIP(hipMemcpyAsync(b->d_mops, pb.mops.data(), pb.mops.size() * sizeof(MicroOp), hipMemcpyHostToDevice, st));
    if (!vt.empty()) VQC_HIP(hipMemcpyAsync(b->d_var, vt.data(), vt.size() * sizeof(VarEntry), hipMemcpyHostToDevice, st));
    if (!pt.empty()) VQC_HIP(hipMemcpyAsync(b->d_prod, pt.data(), pt.size() * sizeof(ProdEntry), hipMemcpyHostToDevice, st));
    VQC_HIP(hipMemcpyAsync(b->d_row_off, row_off.data(), row_off.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
    VQC_HIP(hipMemcpyAsync(b->d_row_cnt, row_cnt.data(), row_cnt.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    VQC_HIP(hipStreamSynchronize(st));

    b->grid = unsigned(std::min<int64_t>({grid_cap, std::max(b->nitems, 1), width}));

    // capture the run as a graph (replayed by vqc_batch_run / vqc_batch_launch)
    if (opts.graph && opts.pinned) {
        hipError_t ce = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        int rc = ce == hipSuccess ? batch_enqueue(b, st, b->h_pool, b->h_rows) : VQC_EHIP;
        hipGraph_t g = nullptr;
        hipError_t ee = hipStreamEndCapture(st, &g);
        if (ce == hipSuccess && rc == VQC_OK && ee == hipSuccess && g) {
            if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) == hipSuccess) b->graph = g;
            else {
                (void)hipGraphDestroy(g);
                b->exec = nullptr;
            }
        } else if (g) {
            (void)hipGraphDestroy(g);
        }
        (void)hipGetLastError();
    }
    *out = b;
    return VQC_OK;
}

int vqc_batch_enqueue_run(vqc_batch* b, const double* pool, const double* hpool, double* h_pool, double* h_rows) {
    if ((b->pool_len && !pool) || (b->hpool_len && !hpool)) return vqc_fail(VQC_EINVAL, "null matrix pool");
    const bool own = h_pool == nullptr;
    double* hp = own ? b->h_pool : h_pool;
    double* hr = h_rows ? h_rows : b->h_rows;
    if (!hp || !hr) return vqc_fail(VQC_EINVAL, "vqc_batch: no staging buffers");
    if (b->pool_len && hp != pool) memcpy(hp, pool, size_t(b->pool_len) * 16);
    if (b->hpool_len && hp + 2 * b->pool_len != hpool) memcpy(hp + 2 * b->pool_len, hpool, size_t(b->hpool_len) * 16);
    if (b->exec && own && hr == b->h_rows) VQC_HIP(hipGraphLaunch(b->exec, b->stream));
    else VQC_TRY(batch_enqueue(b, b->stream, hp, hr));
    return VQC_OK;
}

int vqc_batch_collect(const vqc_batch* b, const double* h_rows, double* energies2, double* grads) {
    const double* rows = h_rows ? h_rows : b->h_rows;
    if (rows[2 * b->nrows] != 0.0) return vqc_fail(VQC_EHIP, "vqc_batch: a work item exceeded its dependency wait (planning error)");
    if (energies2)
        for (int i = 0; i < 2 * b->ninst; ++i) energies2[i] = 0;
    if (grads)
        for (int64_t i = 0; i < b->grad_total; ++i) grads[i] = 0;
    for (int r = 0; r < b->nrows; ++r) {
        const int i = b->row_inst[r];
        const int s = b->row_slot[r];
        if (s < 0) {
            if (energies2) {
                energies2[2 * i] += rows[2 * r];
                energies2[2 * i + 1] += rows[2 * r + 1];
            }
        } else if (grads) {
            grads[b->inst[i].grad_begin + s] += rows[2 * r];
        }
    }
    return VQC_OK;
}

extern "C" int vqc_batch_launch(vqc_batch* b, const double* pool, const double* hpool) {
    if (!b) return vqc_fail(VQC_EINVAL, "null batch");
    if (b->launched) return vqc_fail(VQC_EINVAL, "vqc_batch_launch: previous launch not waited for");
    VQC_TRY(vqc_batch_enqueue_run(b, pool, hpool, nullptr, nullptr));
    b->launched = true;
    return VQC_OK;
}

extern "C" int vqc_batch_wait(vqc_batch* b, double* energies2, double* grads) {
    if (!b) return vqc_fail(VQC_EINVAL, "null batch");
    if (!b->launched) return vqc_fail(VQC_EINVAL, "vqc_batch_wait: nothing launched");
    b->launched = false;
    VQC_HIP(hipStreamSynchronize(b->stream));
    return vqc_batch_collect(b, nullptr, energies2, grads);
}

extern "C" int vqc_batch_run(vqc_batch* b, const double* pool, const double* hpool, double* energies2, double* grads) {
    VQC_TRY(vqc_batch_launch(b, pool, hpool));
    return vqc_batch_wait(b, energies2, grads);
}

extern "C" int vqc_batch_arena_bytes(const vqc_batch* b, uint64_t* bytes) {
    if (!b || !bytes) return vqc_fail(VQC_EINVAL, "null argument");
    *bytes = b->arena_bytes;
    return VQC_OK;
}

extern "C" int vqc_batch_info(const vqc_batch* b, vqc_batch_info_t* out) {
    if (!b || !out) return vqc_fail(VQC_EINVAL, "null argument");
    out->tiled_instances = b->ntiled;
    out->small_instances = b->nsmall;
    out->phases = b->nphases;
    out->items = b->nitems;
    out->grid = int32_t(b->grid);
    out->hbm_bytes = b->hbm_bytes;
    out->arena_bytes = b->arena_bytes;
    return VQC_OK;
}

extern "C" int vqc_batch_destroy(vqc_batch* b) {
    if (!b) return VQC_OK;
    VQC_HIP(hipStreamSynchronize(b->stream));
    return batch_free(b);
}
