/*
 * vqc_hip.h -- C ABI of libvqc_hip.so, the MI355X (gfx950) engine for the
 * gate-application -> Hamiltonian-expectation -> adjoint-gradient path of
 * VQC.jl.
 *
 * Conventions (all taken from the reference, cited per entry point):
 *   - Amplitudes are ComplexF64, passed as interleaved (re, im) doubles; this is
 *     byte-identical to Julia's Vector{ComplexF64} and C's double _Complex[].
 *   - Qubit q (0-based here, 1-based in Julia) is bit q of the amplitude index
 *     (little-endian, src/statevector.jl:167-176).
 *   - A k-qubit matrix is 2^k x 2^k, column-major, over ASCENDING positions:
 *     row/column index bit j corresponds to the j-th smallest position
 *     (ordered_positions / ordered_mat, src/applygates/generic/generic.jl:43-50).
 *   - A density matrix of n qubits is vec(rho), column-major, 4^n entries;
 *     rho[i,j] lives at i + j*2^n, row bits are 0..n-1 and column bits n..2n-1
 *     (src/densitymatrix.jl:27-30).
 *   - Every call returns 0 on success and a negative VQC_E* code on failure;
 *     vqc_last_error() gives a message (thread-local). This replaces the
 *     reference's Julia exceptions (@assert / throw / error).
 *   - Calls are ordered on one HIP stream per device. Calls that return data
 *     to the host (downloads, expectations, gradients) synchronise; the rest
 *     return once the work is enqueued. The single-state calls may come from
 *     any host thread: the library serialises them with one lock, since they
 *     share that stream and its staging buffers. A state used by two threads
 *     sees their calls in an unspecified order (the reference's apply! is not
 *     re-entrant either). Batches and servers own their streams and buffers
 *     and run beside the single-state calls.
 */
#ifndef VQC_HIP_H
#define VQC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define VQC_OK            0
#define VQC_EINVAL       -1   /* bad argument: qubit index, k, size, null pointer */
#define VQC_ENOMEM       -2   /* device allocation failed or budget exceeded */
#define VQC_EHIP         -3   /* a HIP runtime call failed */
#define VQC_EUNSUPPORTED -4   /* valid request this build does not handle */
#define VQC_EINTERNAL    -5   /* an internal consistency check failed (vqc_plan_check) */

/* ---- named gates (src/applygates/specific/) ------------------------------- */
#define VQC_GATE_X        1   /* phasegate.jl:27-43   pos[0] */
#define VQC_GATE_PHASE    2   /* phasegate.jl:6-22    |1> -> e^{i phi}|1> */
#define VQC_GATE_CZ       3   /* czgate.jl:6-25       |11> -> -|11> */
#define VQC_GATE_CNOT     4   /* czgate.jl:30-68      pos[0] = control, pos[1] = target */
#define VQC_GATE_CPHASE   5   /* czgate.jl:73-95      |11> -> e^{i phi}|11> */
#define VQC_GATE_SWAP     6   /* swapgates.jl:3-23 */
#define VQC_GATE_ISWAP    7   /* swapgates.jl:25-49   |01>,|10> swapped times  i */
#define VQC_GATE_ISWAPDAG 8   /* swapgates.jl:47-49   |01>,|10> swapped times -i */
#define VQC_GATE_FREDKIN  9   /* fredkingate.jl:4-77  pos[0] = control */
#define VQC_GATE_CCPHASE 10   /* fredkingate.jl:80-104 |111> -> e^{i phi}|111> */

/* ---- circuit program ---------------------------------------------------- */
#define VQC_OP_UNITARY 0      /* QuantumCircuits Gate: U (SV) or U.rho.U^dagger (DM) */
#define VQC_OP_CHANNEL 1      /* QuantumMap: superoperator on (pos, pos+n) of a DM */

/* One circuit element (a QCircuit entry, src/applygates/generic/generic.jl:26-31).
 * Matrices live in a caller-owned pool of interleaved complex doubles; the
 * offsets below count complex elements. */
typedef struct vqc_op {
    int32_t kind;        /* VQC_OP_UNITARY or VQC_OP_CHANNEL */
    int32_t nq;          /* qubits acted on: 1..5 for unitaries, 1..2 for channels */
    int32_t pos[5];      /* ascending 0-based qubit indices */
    int32_t nparams;     /* differentiable parameters of this op: 0..3 */
    int32_t slot[3];     /* gradient index of each parameter (-1: inactive) */
    int32_t reserved;    /* structure hints, optional (batches use them; other paths read the values):
                            bit 0 U diagonal, bit 1 generators diagonal, bit 2 / 3 U is a CNOT
                            with control on pos[0] / pos[1], bit 4 / 5 U real / X-like, bit 6 / 7
                            generators real / X-like (X-like: real diagonal, imaginary
                            off-diagonal, as Rx); 0 when unknown. A wrong hint gives wrong
                            results. */
    int64_t mat;         /* U (2^nq square) or superoperator S (4^nq square) */
    int64_t aux;         /* channels: S^-1 for the reverse sweep, or -1 when S is singular (the
                            reverse then restores a checkpoint); unitaries: -1 */
    int64_t gen[3];      /* G_j = U^dagger dU/dtheta_j (2^nq square), one per parameter */
} vqc_op_t;

/* One merged Hamiltonian key: sum of the terms sharing a position tuple as a
 * dense 2^nq matrix (src/hamiltonian/util.jl:6-15, expecs.jl:22-28). */
typedef struct vqc_term {
    int32_t nq;          /* 1..5 */
    int32_t pos[5];      /* ascending 0-based qubit indices */
    int64_t mat;         /* pool offset of M */
} vqc_term_t;

/* One Pauli string c * prod_q sigma_q: sigma_q = X (x bit only), Y (x and z),
 * Z (z bit only). Acts as P|b> = c * i^{popcount(x&z)} * (-1)^{popcount(b&z)} |b^x>. */
typedef struct vqc_pauli {
    uint64_t x;
    uint64_t z;
    double   re;
    double   im;
} vqc_pauli_t;

/* One product term c * m_1 (x) ... (x) m_w on w ascending sites: a QubitsTerm of any
 * length, which the reference applies one site at a time (_apply_qterm_util!,
 * src/hamiltonian/apply_qterms/apply_qterms.jl:47-59) whenever an operator holds a
 * term of more than 5 sites. Positions live in a caller array of int32 (0-based
 * qubits), factors in a pool of interleaved complex doubles. */
typedef struct vqc_qterm {
    int32_t nsites;      /* 0..64 (0: c times the identity) */
    int32_t reserved;
    int64_t pos;         /* offset of the nsites ascending positions in the position array */
    int64_t mat;         /* pool offset (complex elements) of nsites consecutive 2x2 column-major factors */
    double  re, im;      /* the coefficient c */
} vqc_qterm_t;

/* ---- library ----------------------------------------------------------- */
const char* vqc_last_error(void);
int vqc_version(void);
/* Device facts used by planners and benchmarks. */
int vqc_device_info(int* cu_count, int* lds_bytes_per_cu, uint64_t* hbm_bytes);
/* Wait for all enqueued work. */
int vqc_sync(void);
/* The HIP stream every call is ordered on (a hipStream_t), for callers that
 * time or order their own work against the library's (no reference counterpart). */
int vqc_get_stream(void** stream);
/* Free the calling thread's stream and work buffers now, after waiting for its
 * work (no reference counterpart; a thread's next call makes new ones). Threads
 * of a pool may call it before they end: the same release otherwise runs in the
 * thread's exit handlers, where a profiler that intercepts HIP calls may already
 * have torn down its own per-thread state. */
int vqc_thread_release(void);

/* ---- tuning (no counterpart in the reference) ----------------------------
 * Launch and fusion parameters, process-wide. A field set to -1 keeps its
 * current value; vqc_get_tuning reports the values in force. The VQC_TILE_* /
 * VQC_DENSE_* environment variables give the initial values. */
typedef struct vqc_tuning {
    int32_t tile_fwd_bits;          /* tile bits of forward phases, 6..13, or 0 (default): by size, below */
    int32_t tile_rev_bits;          /* tile bits of Hamiltonian / reverse phases, 6..12, or 0 (default) */
    int32_t tile_small_bits;        /* whole evaluation in one workgroup up to this many bits, 1..12, or 0 (default) */
    int32_t tile_low_bits;          /* contiguous low index bits in every tile, 0..8 (4; wide
                                       batches take 5 unless this field was set) */
    int32_t dense_groups[3];        /* amplitude groups per thread, 1-, 2-, 3-qubit sweeps: 1, 2, 4, 8 (1) */
    int32_t dense_nontemporal[3];   /* non-temporal loads/stores, 1-, 2-, 3-qubit sweeps: 0/1 (1, 1, 1);
                                       used only when every target is at bit 3 or above */
} vqc_tuning_t;
/* Tile bits left at 0 follow what is evaluated. A single state of b index bits
 * (2n for an n-qubit density matrix) runs as one workgroup up to b = 9, else in
 * tiles of 9 bits up to b = 14, 10 up to 18, 11 up to 20 and 12 forward / 11
 * reverse above: smaller tiles put more CUs on a small state, larger ones put
 * more gates in each pass. A batch (vqc_batch_*) of at least 16 instances takes
 * 13 / 12, the largest tiles, since its instances fill the device together; a
 * smaller batch (the serving loop's usual one or two requests) gives each tiled
 * instance the single-state shape of its size. Batch instances of up to 12 qubits
 * run in one workgroup each. */
int vqc_set_tuning(const vqc_tuning_t* t);
int vqc_get_tuning(vqc_tuning_t* t);
/* Tile shapes by state size, the autotuner's results (vqchip.autotune): entry nbits[i]
 * (index bits 1..34 of a single state, or of a tiled instance of a narrow batch or a
 * served request; -1 for wide batches of at least 16 instances) takes entries[i]. A
 * zero field keeps the built-in rule above for that size; process-wide fields set
 * by vqc_set_tuning take precedence over the table. */
typedef struct vqc_tile_entry {
    int32_t fwd, rev, low, reserved;
} vqc_tile_entry_t;
int vqc_set_tile_table(int nentries, const int32_t* nbits, const vqc_tile_entry_t* entries);
/* the 35 entries in force: index 0 is the wide-batch entry, index b the b-bit one */
int vqc_get_tile_table(vqc_tile_entry_t* out35);

/* ---- state handles (StateVector / DensityMatrix storage) ----------------
 * src/statevector.jl:1-21 (StateVector{T}(n) = |0..0>), src/densitymatrix.jl:38-42. */
typedef struct vqc_state vqc_state;

/* nqubits 0..34. A 0-qubit state vector holds one amplitude: what measure /
 * post_select leave after the last qubit (test/algs.jl:59-74). It supports the
 * whole-state calls (upload, download, copy, scale, axpy, dot, reset); calls that
 * act on qubits reject it. */
int vqc_sv_create(int nqubits, vqc_state** out);             /* |0...0> */
int vqc_dm_create(int nqubits, vqc_state** out);             /* |0...0><0...0| */
/* Borrow caller-owned device memory of 2^nbits complex doubles (not freed). */
int vqc_state_wrap(int nbits, int is_dm, void* device_ptr, vqc_state** out);
int vqc_state_destroy(vqc_state* s);
int vqc_state_nbits(const vqc_state* s, int* nbits);
int vqc_state_device_ptr(vqc_state* s, void** out);
int vqc_state_upload(vqc_state* s, const double* host, uint64_t ncomplex);
int vqc_state_download(const vqc_state* s, double* host, uint64_t ncomplex);
int vqc_state_reset(vqc_state* s);                            /* reset!, statevector.jl:111-115 */
int vqc_state_copy(vqc_state* dst, const vqc_state* src);     /* Base.copy, statevector.jl:35 */
int vqc_state_scale(vqc_state* s, double re, double im);      /* x * y, statevector.jl:50 */
/* y += a x (x + y and x - y, statevector.jl:44-45; the vector updates of the Krylov solvers
 * that groundstate.jl delegates to KrylovKit). */
int vqc_state_axpy(vqc_state* y, double re, double im, const vqc_state* x);
/* rho = |psi><psi|, i.e. vec(rho) = kron(conj(psi), psi): DensityMatrix(x::StateVector)
 * (densitymatrix.jl:27). rho must have psi's qubit count. */
int vqc_dm_from_sv(vqc_state* rho, const vqc_state* psi);
/* rho = the identity matrix (not normalised): DensityMatrix(one(storage(state)), n), on
 * which the expectation pullback of a density matrix applies H^dagger (expecdiff.jl:24-25). */
int vqc_dm_identity(vqc_state* rho);
/* <a|b> = sum conj(a_i) b_i (LinearAlgebra.dot, statevector.jl:58). */
int vqc_state_dot(const vqc_state* a, const vqc_state* b, double* out2);

/* ---- state-vector gates --------------------------------------------------
 * Replaces _apply_gate_threaded2!(key, U, v) for k = 1..5:
 * threaded_short_range.jl:100,364,753 and threaded_long_range.jl:51,131,
 * and the serial fallback _apply_gate_2! (serial_short_range.jl:68-71). */
int vqc_sv_apply_dense(vqc_state* s, int k, const int32_t* pos, const double* U);
/* Replaces apply_threaded!(::NamedGate, v) (the kernels of src/applygates/specific/).
 * pos is in the gate's own order (control first for CNOT and FREDKIN). */
int vqc_sv_apply_named(vqc_state* s, int kind, const int32_t* pos, double phi);

/* ---- Hamiltonians on state vectors -------------------------------------- */
/* out += M_pos * in. Replaces _apply_gate_threaded2!(key, U, v, vout):
 * short_range_threaded.jl:155,756,1504 and long_range_threaded.jl:68,176. */
int vqc_sv_accum_dense(const vqc_state* in, vqc_state* out, int k, const int32_t* pos, const double* M);
/* <bra|M|ket> (bra = ket when bra is NULL). Replaces expectation_value(key, M, v[, v_c]):
 * src/hamiltonian/expecs/expecs.jl:85-131 and the kernels it dispatches to. */
int vqc_sv_expec_dense(const vqc_state* ket, const vqc_state* bra, int k, const int32_t* pos,
                       const double* M, double* out2);
/* sum_t <bra|P_t|ket>, one complex per term written to out (2*nterms doubles) and
 * their sum to sum2. Replaces the >5-site fallback dot(psi, H*psi) of
 * expecs.jl:40 / apply_qterms.jl:20-25 for Pauli strings. */
int vqc_sv_expec_pauli(const vqc_state* ket, const vqc_state* bra, int nterms,
                       const vqc_pauli_t* terms, double* out, double* sum2);
/* vqc_sv_expec_pauli without any allocation (no counterpart in the reference): for
 * callers that cannot allocate device or pinned memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The same kernels and
 * summation order, bit-identical results. The partial sums live in a 4 MiB device
 * array of the library that each GPU holds until the process ends; host work arrays
 * are static. At most 256 terms (VQC_EINVAL beyond). Every call on one GPU shares
 * that array, so calls must not overlap on a GPU: the library's calls are ordered
 * on its one stream and serialised by its lock, which keeps them apart. The first
 * call of any entry point creates the library's stream. */
int vqc_sv_expec_pauli_noalloc(const vqc_state* ket, const vqc_state* bra, int nterms,
                               const vqc_pauli_t* terms, double* out, double* sum2);
/* out = sum_t c_t P_t in  ((m::QubitsOperator)(v), apply_qterms.jl:13-28). */
int vqc_sv_apply_pauli_sum(const vqc_state* in, vqc_state* out, int nterms, const vqc_pauli_t* terms);
/* out = sum_keys M_key in; energy2 = <in|out> (fused H*psi and expectation). */
int vqc_sv_apply_terms(const vqc_state* in, vqc_state* out, int nterms, const vqc_term_t* terms,
                       const double* pool, double* energy2);

/* out = sum_t T_t in, or out += sum_t T_t in when accumulate is 1: the per-term
 * route of (m::QubitsOperator)(v) for operators with terms of more than 5 sites
 * (apply_qterms.jl:20-25 over _apply_qterm_util!, :47-59). On a density matrix
 * the terms act on the row bits of vec(rho) (apply_qterms_dm.jl:11-27). */
int vqc_sv_apply_qterms(const vqc_state* in, vqc_state* out, int nterms, const vqc_qterm_t* terms,
                        const int32_t* pos, const double* pool, int accumulate);
/* <bra|T_t|ket> for every term (2*nterms doubles in out, may be NULL) and their sum
 * (bra = ket when bra is NULL): the dot(psi, m * psi) fallback of
 * expecs.jl:17,40,50,72 for long terms. */
int vqc_sv_expec_qterms(const vqc_state* ket, const vqc_state* bra, int nterms, const vqc_qterm_t* terms,
                        const int32_t* pos, const double* pool, double* out, double* sum2);
/* tr(T_t rho) for every term and their sum: tr(m * rho) of expecs_dm.jl:10-26
 * for terms beyond the dense-key kernels. */
int vqc_dm_expec_qterms(const vqc_state* rho, int nterms, const vqc_qterm_t* terms, const int32_t* pos,
                        const double* pool, double* out, double* sum2);

/* ---- density matrices ----------------------------------------------------
 * rho <- U rho U^dagger: U on pos, conj(U) on pos+n (generic.jl:54-64). */
int vqc_dm_apply_gate(vqc_state* rho, int k, const int32_t* pos, const double* U);
/* Superoperator S (4^N square, N = 1, 2) on (pos..., pos.+n...) (apply_qmaps.jl:52-86);
 * mode 0 applies S (apply!, :8-16), mode 1 S^dagger (apply_dagger!, :25-33), mode 2
 * S^-1 (apply_inverse!, :40-48; VQC_EINVAL when S is singular). */
int vqc_dm_apply_superop(vqc_state* rho, int N, const int32_t* pos, const double* S, int mode);
/* tr(M rho) for a k-site M (expecs_dm.jl:7-27, expec_dm_serial.jl:2-129). */
int vqc_dm_expec_dense(const vqc_state* rho, int k, const int32_t* pos, const double* M, double* out2);

/* ---- measurement and post-selection (src/measure.jl, src/postselect.jl) ----
 * Qubit positions are 0-based. */
/* P(qubit pos == value): sum of |psi_b|^2 (SV) or of rho[l,l] (DM) over the
 * indices with that bit (compute_probability_zero, measure.jl:46-58, 70-82). */
int vqc_prob(const vqc_state* s, int pos, int value, double* p);
/* measure! (measure.jl:4-25): outcome 0 when u < P(0), else 1 (discrete_sample,
 * auxiliary/sampling.jl:2-18, with the caller's uniform u in [0,1) standing for
 * rand()); the state collapses in place by |0><i|/sqrt(p) (auto_reset) or
 * |i><i|/sqrt(p) (on a density matrix: M rho M^dagger). */
int vqc_measure(vqc_state* s, int pos, double u, int auto_reset, int* outcome, double* prob);
/* post_select! (postselect.jl:3-8): project qubit pos on |value>, divide by the
 * norm; *norm receives that norm (the value post_select! returns). SV only. */
int vqc_post_select(vqc_state* s, int pos, int value, double* norm);
/* The amplitudes with bit pos == value, divided by their norm, as a state of one
 * qubit less (measure, measure.jl:28-44; post_select with keep=false,
 * postselect.jl:11-15, 47-53). The remaining qubits come in the reference's
 * order: qubit 0 takes the place of the measured one (its swap!(v, 1, pos)).
 * *norm receives that norm. SV only. */
int vqc_sv_select(const vqc_state* s, int pos, int value, vqc_state* out, double* norm);

/* ---- reductions around the path -----------------------------------------
 * partial_tr (ptrace.jl:4-53): trace out `sites` (0-based, distinct, fewer than
 * all n) of a state vector (out = A A^dagger, ptrace.jl:13-30) or a density
 * matrix (ptrace.jl:42-53). out is a density matrix of n - nsites qubits, which
 * are the kept qubits in ascending order (move_selected_index_backward,
 * auxiliary/indexop.jl:46-70). A trace over every qubit is tr(rho) or <psi|psi>. */
int vqc_partial_trace(const vqc_state* s, int nsites, const int32_t* sites, vqc_state* out);
/* <psi|rho|psi>: fidelity(::DensityMatrix, ::StateVector) (densitymatrix.jl:70). */
int vqc_dm_sandwich(const vqc_state* rho, const vqc_state* psi, double* out2);

/* real(tr(sqrt(x) sqrt(y))): fidelity(x::DensityMatrix, y::DensityMatrix)
 * (densitymatrix.jl:65), from Hermitian eigendecompositions of (x + x^dagger)/2 and
 * (y + y^dagger)/2 (square roots as Julia's sqrt of a Hermitian matrix takes them).
 * At most 14 qubits. */
int vqc_dm_fidelity(const vqc_state* x, const vqc_state* y, double* out);

/* ---- circuits -----------------------------------------------------------
 * apply!(circuit, state) for SV or DM (generic.jl:26-31, apply_qmaps.jl:8-16),
 * fused into LDS-tiled passes. */
int vqc_apply_circuit(vqc_state* s, int nops, const vqc_op_t* ops, const double* pool);

#define VQC_FLAG_UNFUSED 1    /* one kernel per gate and per reduction (the plain path) */
/* Forward-state checkpoints every ceil(sqrt(nops)) ops (fused energy gradients and
 * vjps): the reverse sweep restores each stored state instead of trusting the
 * uncomputed one. Density matrices also checkpoint, whatever the flags, before
 * every channel given without an inverse (vqc_op_t::aux = -1) and wherever the
 * accumulated ||S^-1||_1 of the channels uncomputed since the last restore would
 * pass VQC_CKPT_AMP (1e3); such a channel's reverse applies S^dagger to the
 * cotangent only (circuitdiff.jl:53-59 takes S^-1 of the state). The checkpoints
 * take one state of device memory each. */
#define VQC_FLAG_CHECKPOINT 2

/* Shape of the most recent fused plan run by the calling thread (no reference
 * counterpart): LDS tile passes by kind (each forward pass reads and writes the
 * state once, each Hamiltonian / reverse pass two states) and micro-ops by kind
 * (apply, apply2, apply2x, grad, ham, chain, block). reused is 1 when the plan
 * came from the library's plan cache.
 *
 * Plan cache: vqc_apply_circuit, vqc_sv_energy_grad, vqc_dm_energy_grad and
 * vqc_circuit_vjp keep the plans of the 8 most recent programs. A call whose op
 * and term tables, pool lengths, state shape and tile tuning equal a cached
 * plan's reuses it with the new matrix values: the matrices the plan derives
 * (adjoints, conjugates, products of runs on one qubit) are recomputed from them
 * and every structure class the plan relies on (diagonal, real, X-like, CNOT) is
 * checked again; a class that no longer holds (a zero angle that makes a
 * rotation diagonal, say) plans the program again. VQC_PLAN_CACHE=0 in the
 * environment turns the cache off. */
typedef struct vqc_plan_info {
    int32_t phases_fwd, phases_ham, phases_rev;
    int32_t mops[8];
    int32_t reused;
} vqc_plan_info_t;
int vqc_plan_info(vqc_plan_info_t* out);
/* Host only, no reference counterpart (the reference has no planner): plans the
 * fused energy + gradient of a program as vqc_sv_energy_grad (is_dm = 0) or
 * vqc_dm_energy_grad (is_dm = 1) would for an nqubits-qubit state, checks the
 * plan's invariants (contiguous phases, tile bits partitioning the index bits,
 * micro-op bits inside their tile, matrices inside the pool, accumulators inside
 * their rows) and writes its shape to *out (reused = 0). Makes no device call
 * and leaves the plan cache alone. VQC_EINTERNAL names a violated invariant. */
int vqc_plan_check(int is_dm, int nqubits, int nops, const vqc_op_t* ops, const double* pool, int nterms,
                   const vqc_term_t* terms, const double* hpool, int nparams, int flags, vqc_plan_info_t* out);

/* Energy E = sum_keys <psi|M|psi> with psi = circuit * psi0 and the adjoint
 * gradient dRe(E)/dtheta for every active parameter. Replaces Zygote.gradient(
 * c -> real(expectation(H, c * psi0)), circuit) through @adjoint *(circuit, x)
 * (circuitdiff.jl:2-8, 23-32, 61-78) and @adjoint expectation(H, psi)
 * (expecdiff.jl:30-68). grad has nparams entries, zeroed first. psi_out (may be
 * NULL) receives circuit * psi0. */
int vqc_sv_energy_grad(const vqc_state* psi0, int nops, const vqc_op_t* ops, const double* pool,
                       int nterms, const vqc_term_t* terms, const double* hpool,
                       int nparams, double* energy2, double* grad, vqc_state* psi_out, int flags);
/* Same for a density matrix: E = sum_keys tr(M rho), the adjoint seeded with
 * H^dagger (expecdiff.jl:64-68), gates and channels reversed as in
 * circuitdiff.jl:36-59 (channels undone with S^-1). */
int vqc_dm_energy_grad(const vqc_state* rho0, int nops, const vqc_op_t* ops, const double* pool,
                       int nterms, const vqc_term_t* terms, const double* hpool,
                       int nparams, double* energy2, double* grad, vqc_state* rho_out, int flags);
/* The pullback of y = circuit * x for a cotangent delta_y on y: replaces
 * @adjoint *(circuit, x) with back_propagate (circuitdiff.jl:2-8, 23-78) for
 * any loss. grad[slot] (nparams entries, zeroed first) receives
 * sum Re <Delta|G|y> over the active parameters, and delta_x (may be NULL) the
 * cotangent on x, circuit^dagger delta_y (channels undone with S^dagger). x may
 * be a state vector or a density matrix; delta_y and delta_x have its shape.
 * y_out (may be NULL) receives circuit * x and must not alias delta_y;
 * delta_x may alias delta_y or x. Runs on the fused path: flags must be 0 for
 * state vectors (density matrices also take VQC_FLAG_UNFUSED). */
int vqc_circuit_vjp(const vqc_state* x, int nops, const vqc_op_t* ops, const double* pool, int nparams,
                    const vqc_state* delta_y, double* grad, vqc_state* delta_x, vqc_state* y_out, int flags);

/* ---- batched evaluation (no counterpart in the reference) ----------------
 * Many independent state-vector instances of uneven size, each |0..0> pushed
 * through its own circuit, its energy and its full adjoint gradient, run as
 * one persistent kernel over a packed state arena and replayed as a HIP graph. */
typedef struct vqc_instance {
    int32_t nqubits;     /* 1..30 */
    int32_t nops;
    int64_t op_begin;    /* first op in the shared op array */
    int32_t nterms;
    int32_t nparams;
    int64_t term_begin;  /* first term in the shared term array */
    int64_t grad_begin;  /* first gradient slot in the shared gradient output */
} vqc_instance_t;

typedef struct vqc_batch vqc_batch;
/* Plan (fusion, tiles, work queue), allocate the arena and capture the graph.
 * ops[].mat/gen and terms[].mat index the pools given to vqc_batch_run. */
int vqc_batch_create(int ninst, const vqc_instance_t* inst, const vqc_op_t* ops, int64_t nops_total,
                     const vqc_term_t* terms, int64_t nterms_total, int64_t pool_len,
                     int64_t hpool_len, int flags, vqc_batch** out);
/* Upload the matrices (new angles) and run; energies2 gets 2 doubles per
 * instance, grads the concatenated gradients. */
int vqc_batch_run(vqc_batch* b, const double* pool, const double* hpool, double* energies2, double* grads);
/* Enqueue only (no wait); vqc_batch_wait copies results back. */
int vqc_batch_launch(vqc_batch* b, const double* pool, const double* hpool);
int vqc_batch_wait(vqc_batch* b, double* energies2, double* grads);
int vqc_batch_arena_bytes(const vqc_batch* b, uint64_t* bytes);
/* Plan shape of a batch: instances run in tile phases and in one workgroup's LDS,
 * phases and work items of the persistent queue, its grid, and the HBM bytes the
 * phases move per launch (every phase reads and writes its states once). */
typedef struct vqc_batch_info {
    int32_t tiled_instances, small_instances;
    int64_t phases, items;
    int32_t grid, reserved;
    uint64_t hbm_bytes, arena_bytes;
} vqc_batch_info_t;
int vqc_batch_info(const vqc_batch* b, vqc_batch_info_t* out);
int vqc_batch_destroy(vqc_batch* b);

/* ---- serving loop (no counterpart in the reference) ------------------------
 * Continuous batching of energy + adjoint-gradient requests of mixed size:
 * requests queue on submit; a scheduler thread packs pending requests of a
 * similar cost that fit the device-memory budget into persistent-kernel batches
 * on one of `lanes` HIP streams, several batches in flight at once, and hands
 * each result to its waiter when its batch completes. A request large enough
 * to fill the chip by itself runs alone on its lane through the tiled phase
 * kernels. Each request is one
 * vqc_instance_t whose op_begin / term_begin index the arrays given with it and
 * whose matrices index the pools given with it (grad_begin is ignored).
 * Results are identical to the same request evaluated alone. */
typedef struct vqc_server_config {
    uint64_t hbm_budget;  /* bytes of device memory the loop may hold (capped by free memory) */
    int32_t lanes;        /* concurrent batches (HIP streams), 1..8 */
    int32_t max_batch;    /* requests per batch */
    double grid_frac;     /* share of the resident-workgroup capacity one batch may take (0, 1] */
    double cost_ratio;    /* largest/smallest work estimate allowed in one batch (<= 1: no limit) */
    int32_t poll_us;      /* completion polling interval while batches run */
    int32_t solo_qubits;  /* requests of at least this many qubits run alone (0: never) */
} vqc_server_config_t;

typedef struct vqc_server_stats {
    int64_t completed, batches, batched_requests, pending, running;
    uint64_t budget_bytes, reserved_bytes, peak_live_bytes;
    int64_t batch_hist[8];  /* batches of 1, 2, 3-4, 5-8, 9-16, 17-32, 33-64, 65+ requests */
} vqc_server_stats_t;

typedef struct vqc_server vqc_server;
int vqc_server_default_config(vqc_server_config_t* cfg);
int vqc_server_create(const vqc_server_config_t* cfg, vqc_server** out);
/* Copy one request in and queue it; *ticket names it in poll / wait. */
int vqc_server_submit(vqc_server* s, const vqc_instance_t* inst, const vqc_op_t* ops, const vqc_term_t* terms,
                      const double* pool, int64_t pool_len, const double* hpool, int64_t hpool_len,
                      int64_t* ticket);
/* 1 while the request runs; 0 (and results) once done; < 0 on its failure.
 * times3 = (submit, launch, done) in seconds since the server started. A
 * finished ticket is released by the call that returns its result. */
int vqc_server_poll(vqc_server* s, int64_t ticket, double* energy2, double* grad, double* times3);
int vqc_server_wait(vqc_server* s, int64_t ticket, double* energy2, double* grad, double* times3);
int vqc_server_stats(vqc_server* s, vqc_server_stats_t* out);
/* Finishes every submitted request, then frees the server. */
int vqc_server_destroy(vqc_server* s);

#ifdef __cplusplus
}
#endif
#endif /* VQC_HIP_H */
