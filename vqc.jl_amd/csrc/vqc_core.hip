// vqc_core.hip -- library context, errors, state handles and whole-state
// elementwise kernels (StateVector / DensityMatrix storage on the device).
#include <cstdarg>
#include <cstdlib>
#include <mutex>

#include "vqc_internal.h"

static thread_local std::string g_err;

int vqc_fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

bool vqc_trace_on() {
    static const bool on = [] {
        const char* e = std::getenv("VQC_TRACE");
        return e && std::atoi(e) != 0;
    }();
    return on;
}

void vqc_trace(const char* fmt, ...) {
    if (!vqc_trace_on()) return;
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    std::fprintf(stderr, "vqc-trace: %s\n", buf);
}

extern "C" const char* vqc_last_error(void) { return g_err.c_str(); }
extern "C" int vqc_version(void) { return 10000; }

static VqcCtx g_ctx;  // the device's properties (stream and buffers unused)
static std::mutex g_ctx_mu;

std::recursive_mutex& vqc_api_mutex() {
    static std::recursive_mutex mu;
    return mu;
}
int& vqc_api_depth() {
    static thread_local int depth = 0;
    return depth;
}

// This thread's context: its own stream and staging, released when the thread
// ends (before any static of the HIP runtime goes away, also for the main thread).
namespace {
struct ThreadCtx {
    VqcCtx c;
    void release() {
        if (c.stream) (void)hipStreamSynchronize(c.stream);
        if (c.scratch) (void)hipFree(c.scratch);
        if (c.pinned) (void)hipHostFree(c.pinned);
        if (c.rows) (void)hipHostFree(c.rows);
        if (c.stream) (void)hipStreamDestroy(c.stream);
        c = VqcCtx{};
    }
    ~ThreadCtx() { release(); }
};
thread_local ThreadCtx t_ctx;
}  // namespace

extern "C" int vqc_thread_release(void) {
    // only this thread's own state: no API lock
    VqcCtx& t = t_ctx.c;
    if (t.stream) VQC_HIP(hipStreamSynchronize(t.stream));
    vqc_release_circuit_thread();
    vqc_release_expec_thread();
    vqc_release_gates_thread();
    t_ctx.release();
    return VQC_OK;
}

int vqc_ctx(VqcCtx** out) {
    VqcCtx& t = t_ctx.c;
    if (t.device < 0) {
        {
            std::lock_guard<std::mutex> lk(g_ctx_mu);
            if (g_ctx.device < 0) {
                int ndev = 0;
                hipError_t e = hipGetDeviceCount(&ndev);
                if (e != hipSuccess || ndev < 1)
                    return vqc_fail(VQC_EHIP, "no HIP device available (%s)", hipGetErrorString(e));
                int dev = 0;
                VQC_HIP(hipGetDevice(&dev));
                hipDeviceProp_t prop;
                VQC_HIP(hipGetDeviceProperties(&prop, dev));
                if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos)
                    return vqc_fail(VQC_EUNSUPPORTED, "libvqc_hip is built for gfx950, device is %s", prop.gcnArchName);
                g_ctx.cu_count = prop.multiProcessorCount;
                g_ctx.lds_per_cu = int(prop.maxSharedMemoryPerMultiProcessor);
                g_ctx.hbm_bytes = prop.totalGlobalMem;
                g_ctx.device = dev;
            }
        }
        VQC_HIP(hipSetDevice(g_ctx.device));
        VQC_HIP(hipStreamCreateWithFlags(&t.stream, hipStreamNonBlocking));
        t.cu_count = g_ctx.cu_count;
        t.lds_per_cu = g_ctx.lds_per_cu;
        t.hbm_bytes = g_ctx.hbm_bytes;
        t.device = g_ctx.device;
    }
    *out = &t;
    return VQC_OK;
}

int vqc_scratch(size_t bytes, void** out) {
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    if (bytes > c->scratch_bytes) {
        if (c->scratch) {
            VQC_HIP(hipStreamSynchronize(c->stream));
            VQC_HIP(hipFree(c->scratch));
            c->scratch = nullptr;
            c->scratch_bytes = 0;
        }
        size_t nb = std::max<size_t>(bytes, size_t(1) << 20);
        VQC_HIP(hipMalloc(&c->scratch, nb));
        c->scratch_bytes = nb;
    }
    *out = c->scratch;
    return VQC_OK;
}

int vqc_pinned(size_t bytes, double** out) {
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    if (bytes > c->pinned_bytes) {
        if (c->pinned) {
            VQC_HIP(hipStreamSynchronize(c->stream));
            VQC_HIP(hipHostFree(c->pinned));
            c->pinned = nullptr;
            c->pinned_bytes = 0;
        }
        size_t nb = std::max<size_t>(bytes, size_t(1) << 16);
        VQC_HIP(hipHostMalloc((void**)&c->pinned, nb, hipHostMallocDefault));
        c->pinned_bytes = nb;
    }
    *out = c->pinned;
    return VQC_OK;
}

int vqc_pinned_rows(size_t bytes, double** out) {
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    if (bytes > c->rows_bytes) {
        if (c->rows) {
            VQC_HIP(hipStreamSynchronize(c->stream));
            VQC_HIP(hipHostFree(c->rows));
            c->rows = nullptr;
            c->rows_bytes = 0;
        }
        size_t nb = std::max<size_t>(bytes, size_t(1) << 14);
        VQC_HIP(hipHostMalloc((void**)&c->rows, nb, hipHostMallocCoherent));
        c->rows_bytes = nb;
    }
    *out = c->rows;
    return VQC_OK;
}

int vqc_upload_buffer(void** dev, size_t* cap, const void* host, size_t bytes) {
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    if (bytes > *cap) {
        if (*dev) {
            VQC_HIP(hipStreamSynchronize(c->stream));
            VQC_HIP(hipFree(*dev));
            *dev = nullptr;
            *cap = 0;
        }
        size_t nb = std::max<size_t>(bytes, 4096);
        VQC_HIP(hipMalloc(dev, nb));
        *cap = nb;
    }
    if (bytes) VQC_HIP(hipMemcpyAsync(*dev, host, bytes, hipMemcpyHostToDevice, c->stream));
    return VQC_OK;
}

extern "C" int vqc_device_info(int* cu_count, int* lds_bytes_per_cu, uint64_t* hbm_bytes) {
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    if (cu_count) *cu_count = c->cu_count;
    if (lds_bytes_per_cu) *lds_bytes_per_cu = c->lds_per_cu;
    if (hbm_bytes) *hbm_bytes = c->hbm_bytes;
    return VQC_OK;
}

extern "C" int vqc_sync(void) {
    VQC_API_LOCK;
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    VQC_HIP(hipStreamSynchronize(c->stream));
    return VQC_OK;
}

extern "C" int vqc_get_stream(void** stream) {
    VQC_API_LOCK;
    if (!stream) return vqc_fail(VQC_EINVAL, "null output");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    *stream = (void*)c->stream;
    return VQC_OK;
}

int vqc_check_positions(int nbits, int k, const int32_t* pos, int kmax) {
    if (!pos) return vqc_fail(VQC_EINVAL, "null position array");
    if (k < 1 || k > kmax) return vqc_fail(VQC_EINVAL, "only 1..%d-qubit operators are supported (got %d)", kmax, k);
    for (int j = 0; j < k; ++j) {
        if (pos[j] < 0 || pos[j] >= nbits)
            return vqc_fail(VQC_EINVAL, "qubit position %d out of range 0..%d", pos[j], nbits - 1);
        if (j && pos[j] <= pos[j - 1])
            return vqc_fail(VQC_EINVAL, "positions must be strictly ascending (got %d after %d)", pos[j], pos[j - 1]);
    }
    return VQC_OK;
}

int vqc_check_state(const vqc_state* s, const char* what) {
    if (!s || !s->d) return vqc_fail(VQC_EINVAL, "%s: null state", what);
    return VQC_OK;
}

int vqc_check_qubits(const vqc_state* s, const char* what) {
    VQC_TRY(vqc_check_state(s, what));
    if (s->nbits < 1) return vqc_fail(VQC_EINVAL, "%s: the state has no qubits", what);
    return VQC_OK;
}

// ---------------------------------------------------------------------------
// elementwise kernels
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_fill_basis(double2* v, uint64_t len, uint64_t one_at) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for (; i < len; i += uint64_t(gridDim.x) * blockDim.x) v[i] = make_double2(i == one_at ? 1.0 : 0.0, 0.0);
}

// x <- a x in place: one amplitude per thread, workgroups in XCD-contiguous order,
// non-temporal loads and stores (each amplitude is touched once). On a 4 GiB state
// this reaches the streaming rate of the one-qubit sweeps; a grid-stride loop over
// 65536 workgroups with plain accesses ran at 5.4 TB/s, one amplitude per thread
// with non-temporal accesses at 5.9 TB/s.
typedef double scale_d2v __attribute__((ext_vector_type(2)));
__global__ void __launch_bounds__(256) k_scale(double2* v, uint64_t len, double2 a) {
    const uint64_t i = uint64_t(xcd_block()) * 256 + threadIdx.x;
    if (i >= len) return;
    scale_d2v* p = reinterpret_cast<scale_d2v*>(v + i);
    const scale_d2v x = __builtin_nontemporal_load(p);
    scale_d2v y;
    y.x = fma(a.x, x.x, -a.y * x.y);
    y.y = fma(a.x, x.y, a.y * x.x);
    __builtin_nontemporal_store(y, p);
}

// per-block partials of sum conj(a_i) b_i
__global__ void __launch_bounds__(256) k_dot(const double2* __restrict__ a, const double2* __restrict__ b,
                                             uint64_t len, double2* parts) {
    __shared__ double red[8];
    double sr = 0, si = 0;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < len; i += uint64_t(gridDim.x) * blockDim.x) {
        double2 x = a[i], y = b[i];
        sr = fma(x.x, y.x, fma(x.y, y.y, sr));
        si = fma(x.x, y.y, fma(-x.y, y.x, si));
    }
    sr = block_sum(sr, red);
    si = block_sum(si, red);
    if (threadIdx.x == 0) parts[blockIdx.x] = make_double2(sr, si);
}

// one block per term: out[t] = sum_p parts[t*nparts + p], summed in a fixed order
__global__ void __launch_bounds__(256) k_reduce_parts(const double2* __restrict__ parts, int nparts, double2* out) {
    __shared__ double red[8];
    const double2* p = parts + size_t(blockIdx.x) * nparts;
    double sr = 0, si = 0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        sr += p[i].x;
        si += p[i].y;
    }
    sr = block_sum(sr, red);
    si = block_sum(si, red);
    if (threadIdx.x == 0) out[blockIdx.x] = make_double2(sr, si);
}

void vqc_reduce_partials(const double2* parts, int nterms, int nparts, double2* out, hipStream_t st) {
    hipLaunchKernelGGL(k_reduce_parts, dim3(nterms), dim3(256), 0, st, parts, nparts, out);
}

// one wave per row: lane l sums columns l, l + 64, ... of the row's written
// columns, then the wave's fixed-order sum; lane 0 stores the result with
// system-scope stores (out may be host memory)
__global__ void __launch_bounds__(256) k_reduce_rows(const double2* __restrict__ parts, int stride, VqcRowSegs segs, int nrows,
                                                     double2* out) {
    const int r = int(blockIdx.x) * 4 + int(threadIdx.x >> 6);
    if (r >= nrows) return;  // the whole wave
    const int lane = threadIdx.x & 63;
    int cols = 0;
    for (int k = 0; k < segs.n; ++k)
        if (r < segs.end[k]) {
            cols = segs.cols[k];
            break;
        }
    const double2* __restrict__ p = parts + size_t(r) * stride;
    double sr = 0.0, si = 0.0;
    for (int i = lane; i < cols; i += 64) {
        sr += p[i].x;
        si += p[i].y;
    }
    sr = wave_sum_full(sr);
    si = wave_sum_full(si);
    if (lane == 0) {
        __hip_atomic_store(&out[r].x, sr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_store(&out[r].y, si, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

void vqc_reduce_rows(const double2* parts, int stride, const VqcRowSegs& segs, int nrows, double2* out, hipStream_t st) {
    hipLaunchKernelGGL(k_reduce_rows, dim3((nrows + 3) / 4), dim3(256), 0, st, parts, stride, segs, nrows, out);
}

int vqc_wait_unlocked() {
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    // other threads' calls may plan and launch on their streams meanwhile: the
    // stream, the pinned staging and the device buffers are this thread's own
    const int depth = vqc_api_depth();
    for (int i = 0; i < depth; ++i) vqc_api_mutex().unlock();
    vqc_api_depth() = 0;
    const hipError_t e = hipStreamSynchronize(c->stream);
    for (int i = 0; i < depth; ++i) vqc_api_mutex().lock();
    vqc_api_depth() = depth;
    VQC_HIP(e);
    return VQC_OK;
}

int vqc_fetch_unlocked(const double2* dev, int n, double* host) {
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    double* pin;
    VQC_TRY(vqc_pinned(size_t(n) * 16, &pin));
    VQC_HIP(hipMemcpyAsync(pin, dev, size_t(n) * 16, hipMemcpyDeviceToHost, c->stream));
    VQC_TRY(vqc_wait_unlocked());
    for (int i = 0; i < 2 * n; ++i) host[i] = pin[i];
    return VQC_OK;
}

int vqc_fetch(const double2* dev, int n, double* host) {
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    double* pin;
    VQC_TRY(vqc_pinned(size_t(n) * 16, &pin));
    VQC_HIP(hipMemcpyAsync(pin, dev, size_t(n) * 16, hipMemcpyDeviceToHost, c->stream));
    VQC_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < 2 * n; ++i) host[i] = pin[i];
    return VQC_OK;
}

static unsigned stream_grid(uint64_t len) {
    uint64_t b = (len + 255) / 256;
    return unsigned(std::min<uint64_t>(b, 65536));
}

// ---------------------------------------------------------------------------
// state handles
// ---------------------------------------------------------------------------
// A state vector may have 0 qubits (one amplitude): what measure / post_select
// leave after the last qubit is taken out (test/algs.jl:59-74).
static int state_alloc(int nbits, int nq, int is_dm, vqc_state** out) {
    if (!out) return vqc_fail(VQC_EINVAL, "null output handle");
    if (nbits < (is_dm ? 2 : 0) || nbits > 34) return vqc_fail(VQC_EINVAL, "unsupported state size: %d index bits", nbits);
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    vqc_state* s = new vqc_state;
    s->nbits = nbits;
    s->nqubits = nq;
    s->is_dm = is_dm;
    hipError_t e = hipMalloc((void**)&s->d, s->len() * sizeof(double2));
    if (e != hipSuccess) {
        delete s;
        (void)hipGetLastError();
        return vqc_fail(VQC_ENOMEM, "cannot allocate %llu bytes for a %d-bit state",
                        (unsigned long long)((uint64_t(1) << nbits) * 16), nbits);
    }
    *out = s;
    return vqc_state_reset(s);
}

extern "C" int vqc_sv_create(int nqubits, vqc_state** out) {
    VQC_API_LOCK;
    return state_alloc(nqubits, nqubits, 0, out);
}
extern "C" int vqc_dm_create(int nqubits, vqc_state** out) {
    VQC_API_LOCK;
    if (nqubits < 1 || nqubits > 17) return vqc_fail(VQC_EINVAL, "unsupported density-matrix size: %d qubits", nqubits);
    return state_alloc(2 * nqubits, nqubits, 1, out);
}

extern "C" int vqc_state_wrap(int nbits, int is_dm, void* device_ptr, vqc_state** out) {
    VQC_API_LOCK;
    if (!out || !device_ptr) return vqc_fail(VQC_EINVAL, "null pointer");
    if (nbits < (is_dm ? 2 : 0) || nbits > 34 || (is_dm && (nbits & 1))) return vqc_fail(VQC_EINVAL, "bad state size %d", nbits);
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    vqc_state* s = new vqc_state;
    s->nbits = nbits;
    s->nqubits = is_dm ? nbits / 2 : nbits;
    s->is_dm = is_dm;
    s->owns = 0;
    s->d = (double2*)device_ptr;
    *out = s;
    return VQC_OK;
}

extern "C" int vqc_state_destroy(vqc_state* s) {
    VQC_API_LOCK;
    if (!s) return VQC_OK;
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    VQC_HIP(hipStreamSynchronize(c->stream));
    if (s->owns && s->d) VQC_HIP(hipFree(s->d));
    delete s;
    return VQC_OK;
}

extern "C" int vqc_state_nbits(const vqc_state* s, int* nbits) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_state(s, "vqc_state_nbits"));
    *nbits = s->nbits;
    return VQC_OK;
}

extern "C" int vqc_state_device_ptr(vqc_state* s, void** out) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_state(s, "vqc_state_device_ptr"));
    *out = s->d;
    return VQC_OK;
}

extern "C" int vqc_state_upload(vqc_state* s, const double* host, uint64_t n) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_state(s, "vqc_state_upload"));
    if (!host || n != s->len()) return vqc_fail(VQC_EINVAL, "upload of %llu amplitudes into a state of %llu",
                                                (unsigned long long)n, (unsigned long long)s->len());
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    VQC_HIP(hipMemcpyAsync(s->d, host, n * 16, hipMemcpyHostToDevice, c->stream));
    VQC_HIP(hipStreamSynchronize(c->stream));
    return VQC_OK;
}

extern "C" int vqc_state_download(const vqc_state* s, double* host, uint64_t n) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_state(s, "vqc_state_download"));
    if (!host || n != s->len()) return vqc_fail(VQC_EINVAL, "download of %llu amplitudes from a state of %llu",
                                                (unsigned long long)n, (unsigned long long)s->len());
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    VQC_HIP(hipMemcpyAsync(host, s->d, n * 16, hipMemcpyDeviceToHost, c->stream));
    VQC_HIP(hipStreamSynchronize(c->stream));
    return VQC_OK;
}

extern "C" int vqc_state_reset(vqc_state* s) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_state(s, "vqc_state_reset"));
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    hipLaunchKernelGGL(k_fill_basis, dim3(stream_grid(s->len())), dim3(256), 0, c->stream, s->d, s->len(), uint64_t(0));
    VQC_HIP(hipGetLastError());
    return VQC_OK;
}

extern "C" int vqc_state_copy(vqc_state* dst, const vqc_state* src) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_state(dst, "vqc_state_copy"));
    VQC_TRY(vqc_check_state(src, "vqc_state_copy"));
    if (dst->nbits != src->nbits) return vqc_fail(VQC_EINVAL, "copy between states of different size");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    VQC_HIP(hipMemcpyAsync(dst->d, src->d, src->len() * 16, hipMemcpyDeviceToDevice, c->stream));
    return VQC_OK;
}

__global__ void __launch_bounds__(256) k_axpy(double2* __restrict__ y, const double2* __restrict__ x, uint64_t len, double2 a) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for (; i < len; i += uint64_t(gridDim.x) * blockDim.x) y[i] = cfma(a, x[i], y[i]);
}

extern "C" int vqc_state_axpy(vqc_state* y, double re, double im, const vqc_state* x) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_state(y, "vqc_state_axpy"));
    VQC_TRY(vqc_check_state(x, "vqc_state_axpy"));
    if (x->nbits != y->nbits) return vqc_fail(VQC_EINVAL, "vqc_state_axpy: states differ in size");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    hipLaunchKernelGGL(k_axpy, dim3(stream_grid(y->len())), dim3(256), 0, c->stream, y->d, x->d, y->len(), make_double2(re, im));
    VQC_HIP(hipGetLastError());
    return VQC_OK;
}

// rho[i + j 2^n] = psi_i conj(psi_j): DensityMatrix(x::StateVector) = kron(conj(x), x)
// (densitymatrix.jl:27), one output entry per thread, columns j walked by the grid
__global__ void __launch_bounds__(256) k_dm_from_sv(double2* __restrict__ rho, const double2* __restrict__ psi, int n) {
    const uint64_t len = uint64_t(1) << (2 * n), mask = (uint64_t(1) << n) - 1;
    for (uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < len; e += uint64_t(gridDim.x) * blockDim.x)
        rho[e] = cmul(psi[e & mask], cconj(psi[e >> n]));
}

extern "C" int vqc_dm_from_sv(vqc_state* rho, const vqc_state* psi) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_state(rho, "vqc_dm_from_sv"));
    VQC_TRY(vqc_check_qubits(psi, "vqc_dm_from_sv"));
    if (!rho->is_dm || psi->is_dm || rho->nqubits != psi->nqubits)
        return vqc_fail(VQC_EINVAL, "vqc_dm_from_sv: needs a density matrix and a state vector of the same qubit count");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    hipLaunchKernelGGL(k_dm_from_sv, dim3(stream_grid(rho->len())), dim3(256), 0, c->stream, rho->d, psi->d, psi->nqubits);
    VQC_HIP(hipGetLastError());
    return VQC_OK;
}

// vec(I): ones at i + i 2^n (the DensityMatrix(one(storage(state)), n) that the
// expectation pullback of a density matrix applies H^dagger to, expecdiff.jl:24-25,64-65)
__global__ void __launch_bounds__(256) k_dm_identity(double2* __restrict__ rho, int n) {
    const uint64_t len = uint64_t(1) << (2 * n), mask = (uint64_t(1) << n) - 1;
    for (uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < len; e += uint64_t(gridDim.x) * blockDim.x)
        rho[e] = make_double2((e & mask) == (e >> n) ? 1.0 : 0.0, 0.0);
}

extern "C" int vqc_dm_identity(vqc_state* rho) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_state(rho, "vqc_dm_identity"));
    if (!rho->is_dm) return vqc_fail(VQC_EINVAL, "vqc_dm_identity: not a density matrix");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    hipLaunchKernelGGL(k_dm_identity, dim3(stream_grid(rho->len())), dim3(256), 0, c->stream, rho->d, rho->nqubits);
    VQC_HIP(hipGetLastError());
    return VQC_OK;
}

extern "C" int vqc_state_scale(vqc_state* s, double re, double im) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_state(s, "vqc_state_scale"));
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    hipLaunchKernelGGL(k_scale, dim3(unsigned((s->len() + 255) / 256)), dim3(256), 0, c->stream, s->d, s->len(), make_double2(re, im));
    VQC_HIP(hipGetLastError());
    return VQC_OK;
}

extern "C" int vqc_state_dot(const vqc_state* a, const vqc_state* b, double* out2) {
    VQC_API_LOCK;
    VQC_TRY(vqc_check_state(a, "vqc_state_dot"));
    VQC_TRY(vqc_check_state(b, "vqc_state_dot"));
    if (a->nbits != b->nbits) return vqc_fail(VQC_EINVAL, "dot of states of different size");
    if (!out2) return vqc_fail(VQC_EINVAL, "null output");
    VqcCtx* c;
    VQC_TRY(vqc_ctx(&c));
    const unsigned nb = unsigned(std::min<uint64_t>((a->len() + 255) / 256, 2048));
    void* scr;
    VQC_TRY(vqc_scratch((nb + 1) * 16, &scr));
    double2* parts = (double2*)scr;
    hipLaunchKernelGGL(k_dot, dim3(nb), dim3(256), 0, c->stream, a->d, b->d, a->len(), parts);
    vqc_reduce_partials(parts, 1, nb, parts + nb, c->stream);
    VQC_HIP(hipGetLastError());
    return vqc_fetch(parts + nb, 1, out2);
}

// ---------------------------------------------------------------------------
// tuning
// ---------------------------------------------------------------------------
namespace {
std::mutex g_tune_mu;
bool g_tune_init = false;
vqc_tuning_t g_tune;
bool g_low_bits_given = false;

int env_or(const char* name, int def) {
    const char* v = getenv(name);
    return v && *v ? atoi(v) : def;
}

void tune_init_locked() {
    if (g_tune_init) return;
    // 0: chosen by the size of what is evaluated
    g_tune.tile_fwd_bits = env_or("VQC_TILE_FWD", 0);
    g_tune.tile_rev_bits = env_or("VQC_TILE_REV", 0);
    g_tune.tile_small_bits = env_or("VQC_TILE_SMALL", 0);
    g_tune.tile_low_bits = env_or("VQC_TILE_MLOW", 4);
    g_low_bits_given = getenv("VQC_TILE_MLOW") != nullptr && *getenv("VQC_TILE_MLOW");
    const int r = env_or("VQC_DENSE_R", 1), nt = env_or("VQC_DENSE_NT", -1);
    for (int k = 0; k < 3; ++k) {
        g_tune.dense_groups[k] = r;
        // non-temporal accesses, taken by the sweeps whose targets are all at bit 3 or above
        g_tune.dense_nontemporal[k] = nt >= 0 ? (nt != 0) : 1;
    }
    g_tune_init = true;
}

bool valid_groups(int r) { return r == 1 || r == 2 || r == 4 || r == 8; }
}  // namespace

bool vqc_tuning_low_bits_given() {
    std::lock_guard<std::mutex> g(g_tune_mu);
    tune_init_locked();
    return g_low_bits_given;
}

vqc_tuning_t vqc_tuning_now() {
    std::lock_guard<std::mutex> g(g_tune_mu);
    tune_init_locked();
    return g_tune;
}

extern "C" int vqc_set_tuning(const vqc_tuning_t* t) {
    if (!t) return vqc_fail(VQC_EINVAL, "null tuning");
    std::lock_guard<std::mutex> g(g_tune_mu);
    tune_init_locked();
    vqc_tuning_t n = g_tune;
    auto take = [](int32_t v, int32_t& dst) {
        if (v != -1) dst = v;
    };
    take(t->tile_fwd_bits, n.tile_fwd_bits);
    take(t->tile_rev_bits, n.tile_rev_bits);
    take(t->tile_small_bits, n.tile_small_bits);
    take(t->tile_low_bits, n.tile_low_bits);
    for (int k = 0; k < 3; ++k) {
        take(t->dense_groups[k], n.dense_groups[k]);
        take(t->dense_nontemporal[k], n.dense_nontemporal[k]);
    }
    // tile bits: 0 (by size: at least 9 bits) or a fixed count
    const int fwd_min = n.tile_fwd_bits ? n.tile_fwd_bits : 9, rev_min = n.tile_rev_bits ? n.tile_rev_bits : 9;
    if ((n.tile_fwd_bits != 0 && (n.tile_fwd_bits < 6 || n.tile_fwd_bits > 13)) ||
        (n.tile_rev_bits != 0 && (n.tile_rev_bits < 6 || n.tile_rev_bits > 12)) || n.tile_small_bits < 0 ||
        n.tile_small_bits > 12 || n.tile_low_bits < 0 || n.tile_low_bits > 8 || n.tile_low_bits > std::min(fwd_min, rev_min))
        return vqc_fail(VQC_EINVAL, "vqc_set_tuning: tile bits out of range");
    for (int k = 0; k < 3; ++k)
        if (!valid_groups(n.dense_groups[k]) || (n.dense_nontemporal[k] != 0 && n.dense_nontemporal[k] != 1))
            return vqc_fail(VQC_EINVAL, "vqc_set_tuning: dense groups must be 1, 2, 4 or 8 and non-temporal 0 or 1");
    g_tune = n;
    if (t->tile_low_bits != -1) g_low_bits_given = true;
    return VQC_OK;
}

// Tile shapes by state size, set by the autotuner: entry b for states of b index
// bits (1..34), entry 0 for wide batches; zero fields keep the built-in rule.
static vqc_tile_entry_t g_tile_table[35];

extern "C" int vqc_set_tile_table(int nentries, const int32_t* nbits, const vqc_tile_entry_t* entries) {
    if (nentries < 0 || (nentries && (!nbits || !entries))) return vqc_fail(VQC_EINVAL, "vqc_set_tile_table: bad arrays");
    for (int i = 0; i < nentries; ++i) {
        const int b = nbits[i] < 0 ? 0 : nbits[i];
        const vqc_tile_entry_t& e = entries[i];
        if (b > 34 || (e.fwd && (e.fwd < 6 || e.fwd > 13)) || (e.rev && (e.rev < 6 || e.rev > 12)) || e.low < 0 || e.low > 8 ||
            (e.fwd && e.low > e.fwd) || (e.rev && e.low > e.rev))
            return vqc_fail(VQC_EINVAL, "vqc_set_tile_table: entry %d out of range", i);
    }
    std::lock_guard<std::mutex> g(g_tune_mu);
    for (int i = 0; i < nentries; ++i) g_tile_table[nbits[i] < 0 ? 0 : nbits[i]] = entries[i];
    return VQC_OK;
}

extern "C" int vqc_get_tile_table(vqc_tile_entry_t* out35) {
    if (!out35) return vqc_fail(VQC_EINVAL, "null table");
    std::lock_guard<std::mutex> g(g_tune_mu);
    for (int i = 0; i < 35; ++i) out35[i] = g_tile_table[i];
    return VQC_OK;
}

vqc_tile_entry_t vqc_tile_entry(int nbits) {
    std::lock_guard<std::mutex> g(g_tune_mu);
    return g_tile_table[nbits < 0 ? 0 : std::min(nbits, 34)];
}

extern "C" int vqc_get_tuning(vqc_tuning_t* t) {
    if (!t) return vqc_fail(VQC_EINVAL, "null tuning");
    *t = vqc_tuning_now();
    return VQC_OK;
}
