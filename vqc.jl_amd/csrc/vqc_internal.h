// vqc_internal.h -- shared definitions of libvqc_hip.so (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vqc_hip.h"

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
int vqc_fail(int code, const char* fmt, ...);
// Per-launch trace on stderr when the environment sets VQC_TRACE=1: one line per
// kernel launch with its shape (kernel, grid, tile bits, micro-ops).
bool vqc_trace_on();
void vqc_trace(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
#define VQC_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) return vqc_fail(VQC_EHIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)
#define VQC_TRY(call)               \
    do {                            \
        int r_ = (call);            \
        if (r_ != VQC_OK) return r_; \
    } while (0)

// One lock for the host work of the single-state entry points: the plan cache,
// the tuning and tile tables, planning and launching. Each host thread has its
// own stream and device buffers (VqcCtx, DeviceBuffer below), so the GPU work of
// calls from different threads overlaps: the fused evaluations release the lock
// while they wait for their results (vqc_fetch_unlocked). Recursive: an entry
// point may call another. Batches and the server own their buffers and streams
// and do not take it.
std::recursive_mutex& vqc_api_mutex();
int& vqc_api_depth();  // times this thread holds the lock
struct VqcApiGuard {
    VqcApiGuard() {
        vqc_api_mutex().lock();
        ++vqc_api_depth();
    }
    ~VqcApiGuard() {
        --vqc_api_depth();
        vqc_api_mutex().unlock();
    }
    VqcApiGuard(const VqcApiGuard&) = delete;
    VqcApiGuard& operator=(const VqcApiGuard&) = delete;
};
#define VQC_API_LOCK VqcApiGuard vqc_api_guard_

// ---------------------------------------------------------------------------
// library context of the calling host thread: the device's properties, this
// thread's stream (default priority), a growable device scratch area and pinned
// host staging for results. Freed when the thread ends.
// ---------------------------------------------------------------------------
struct VqcCtx {
    int device = -1;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    int lds_per_cu = 0;
    uint64_t hbm_bytes = 0;
    void* scratch = nullptr;      // device, grown on demand
    size_t scratch_bytes = 0;
    double* pinned = nullptr;     // host pinned staging
    size_t pinned_bytes = 0;
    double* rows = nullptr;       // host pinned, device-coherent: results kernels write directly
    size_t rows_bytes = 0;
};
int vqc_ctx(VqcCtx** out);
// A device buffer of one host thread (declared thread_local), grown on demand and
// freed when the thread ends.
struct DeviceBuffer {
    void* p = nullptr;
    size_t cap = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    ~DeviceBuffer() { release(); }
};
// Free the calling thread's buffers of each unit (vqc_thread_release; the
// thread's stream has been waited for).
void vqc_release_circuit_thread();
void vqc_release_expec_thread();
void vqc_release_gates_thread();
// the tuning in force (vqc_set_tuning; initial values from the environment)
vqc_tuning_t vqc_tuning_now();
// whether tile_low_bits was given (vqc_set_tuning or VQC_TILE_MLOW) rather than defaulted
bool vqc_tuning_low_bits_given();
// the autotuned tile shape for states of nbits index bits (nbits < 0: wide batches); zero fields: none
vqc_tile_entry_t vqc_tile_entry(int nbits);
int vqc_scratch(size_t bytes, void** out);      // device scratch, valid until the next call
int vqc_pinned(size_t bytes, double** out);     // pinned host staging
// pinned, device-coherent host memory of this thread that kernels write results
// into directly (no copy kernel): valid once the stream has been waited for
int vqc_pinned_rows(size_t bytes, double** out);
// wait for this thread's stream with the API lock released
int vqc_wait_unlocked();
// Upload host data into a dedicated device buffer of at least `bytes` bytes (grown on demand).
int vqc_upload_buffer(void** dev, size_t* cap, const void* host, size_t bytes);

struct vqc_state {
    int nbits = 0;       // log2 of the number of stored amplitudes
    int nqubits = 0;     // n (== nbits for SV, nbits/2 for DM)
    int is_dm = 0;
    int owns = 1;
    double2* d = nullptr;
    uint64_t len() const { return uint64_t(1) << nbits; }
};

// ---------------------------------------------------------------------------
// complex helpers (double2 = re, im)
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ double2 cmk(double a, double b) { return make_double2(a, b); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
// acc + a*b
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 acc) {
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
    return acc;
}
// acc + conj(a)*b
__device__ __forceinline__ double2 cfmaj(double2 a, double2 b, double2 acc) {
    acc.x = fma(a.x, b.x, fma(a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(-a.y, b.x, acc.y));
    return acc;
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// insert zero bits at ascending positions p[0] < p[1] < ...
template <int K>
__device__ __forceinline__ uint64_t insert_zeros(uint64_t i, const int* p) {
#pragma unroll
    for (int j = 0; j < K; ++j) {
        uint64_t lo = i & ((uint64_t(1) << p[j]) - 1);
        i = ((i - lo) << 1) | lo;
    }
    return i;
}
__device__ __forceinline__ uint64_t insert_zeros_dyn(uint64_t i, const int* p, int k) {
    for (int j = 0; j < k; ++j) {
        uint64_t lo = i & ((uint64_t(1) << p[j]) - 1);
        i = ((i - lo) << 1) | lo;
    }
    return i;
}
template <int K>
__device__ __forceinline__ uint64_t sub_offset(int j, const int* p) {
    uint64_t o = 0;
#pragma unroll
    for (int b = 0; b < K; ++b)
        if (j >> b & 1) o |= uint64_t(1) << p[b];
    return o;
}

// ---------------------------------------------------------------------------
// wavefront / block reductions (64-wide wavefronts)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// Lane permutes of a double through DPP (in the VALU, no LDS traffic): CTRL is
// a DPP control word (quad_perm 0x00-0xff, row_mirror 0x140, row_half_mirror
// 0x141). Every source lane must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double readlane_f64(double v, int lane) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                            __builtin_amdgcn_readlane(__double2loint(v), lane));
}
// Sum over the 64 lanes of a wave with every lane active (call it outside
// divergent code): pairs, quads, half rows and rows of 16 by DPP, then the four
// row sums read as scalars. Same result in every lane, in a fixed order.
__device__ __forceinline__ double wave_sum_full(double v) {
    v += dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_f64<0x141>(v);  // row_half_mirror
    v += dpp_f64<0x140>(v);  // row_mirror
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}
// Workgroups are dispatched to the 8 XCDs of an MI355X round-robin (workgroup i
// runs on XCD i % 8; each XCD has its own L2). Renumbering them so that XCD x
// takes a contiguous 1/8 of the workgroup range turns each XCD's accesses of a
// streaming sweep into one sequential stream instead of every 8th chunk: a
// 1-qubit sweep of a 28-qubit state went from 5.5-5.9 to 5.8-6.2 TB/s for most
// target bits. A bijection on [0, gridDim.x) for any grid size.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t nb = gridDim.x, i = blockIdx.x;
    const uint32_t q = nb / kXcds, r = nb % kXcds, x = i % kXcds, k = i / kXcds;
    return x * q + (x < r ? x : r) + k;
}

// Sum over a block of blockDim.x threads (multiple of 64); every thread gets
// the result. `red` holds at least blockDim.x/64 doubles of LDS.
__device__ __forceinline__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0;
    for (int i = 0; i < nw; ++i) s += red[i];
    return s;
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
inline unsigned grid_for(uint64_t work, unsigned per_block) {
    uint64_t b = (work + per_block - 1) / per_block;
    return unsigned(b ? b : 1);
}
int vqc_check_positions(int nbits, int k, const int32_t* pos, int kmax);
int vqc_check_state(const vqc_state* s, const char* what);
// vqc_check_state, and at least one qubit (a 0-qubit state has one amplitude)
int vqc_check_qubits(const vqc_state* s, const char* what);

// Reduce `nparts` per-block complex partials for each of `nterms` terms
// (layout [term][part]) into `out` (device, nterms complex), in a fixed order.
void vqc_reduce_partials(const double2* parts, int nterms, int nparts, double2* out, hipStream_t st);
// Partial rows whose written columns vary by row: segment k holds the rows below
// end[k] (from end[k-1] on) with cols[k] columns written, the rest never written.
constexpr int kMaxRowSegs = 64;
struct VqcRowSegs {
    int32_t n;
    int32_t end[kMaxRowSegs];
    int32_t cols[kMaxRowSegs];
};
// out[r] = sum of the written columns of row r (row stride `stride`), in a fixed
// order; out may be host memory from vqc_pinned_rows
void vqc_reduce_rows(const double2* parts, int stride, const VqcRowSegs& segs, int nrows, double2* out, hipStream_t st);
// Copy n complex device results to the host and wait.
int vqc_fetch(const double2* dev, int n, double* host);
// vqc_fetch with the API lock released while it waits for the stream (the
// caller holds no reference into shared host state across it)
int vqc_fetch_unlocked(const double2* dev, int n, double* host);
