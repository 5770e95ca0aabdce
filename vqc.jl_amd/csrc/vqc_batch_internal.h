// vqc_batch_internal.h -- the batch engine's build / enqueue / collect steps,
// shared by the public vqc_batch_* calls and the serving loop (vqc_serve.hip).
#pragma once
#include "vqc_internal.h"

struct BatchOpts {
    hipStream_t stream = nullptr;  // nullptr: the library stream
    // device memory for the batch (one block); nullptr: hipMalloc / hipFree
    void* (*dev_alloc)(void* ctx, size_t bytes) = nullptr;
    void (*dev_free)(void* ctx, void* p) = nullptr;
    void* alloc_ctx = nullptr;
    // share of the resident-workgroup capacity the persistent grid may take, so
    // that batches on other streams find free CUs
    double grid_frac = 1.0;
    bool graph = true;    // capture a replayable HIP graph
    bool pinned = true;   // allocate pinned staging; else the caller sets it per enqueue
};

int vqc_batch_build(int ninst, const vqc_instance_t* inst, const vqc_op_t* ops, int64_t nops_total,
                    const vqc_term_t* terms, int64_t nterms_total, int64_t pool_len, int64_t hpool_len,
                    const BatchOpts& opts, vqc_batch** out);
// bytes of the batch's device block and of its two host staging areas
void vqc_batch_sizes(const vqc_batch* b, size_t* device_bytes, size_t* pool_staging_bytes, size_t* rows_staging_bytes);
// Enqueue one run on the batch's stream. h_pool / h_rows: pinned staging of at
// least the sizes above (nullptr: the batch's own). The pools are copied into
// the staging before the call returns.
int vqc_batch_enqueue_run(vqc_batch* b, const double* pool, const double* hpool, double* h_pool, double* h_rows);
// After the stream has passed the run: energies and gradients from h_rows.
int vqc_batch_collect(const vqc_batch* b, const double* h_rows, double* energies2, double* grads);
// Release without synchronising (the caller knows the stream is idle).
int vqc_batch_release(vqc_batch* b);

// One state-vector energy + gradient evaluation of |0..0> (more qubits than a
// tile) through the tiled phase kernels, on a caller's stream and device block.
struct vqc_solo;
int vqc_solo_build(int n, int nops, const vqc_op_t* ops, const double* pool, int64_t pool_len, int nterms,
                   const vqc_term_t* terms, const double* hpool, int64_t hpool_len, int nparams, vqc_solo** out);
size_t vqc_solo_device_bytes(const vqc_solo* s);
size_t vqc_solo_rows_bytes(const vqc_solo* s);
// Enqueue the evaluation; the rows land in pinned h_rows when the stream passes it.
int vqc_solo_enqueue(vqc_solo* s, void* block, hipStream_t st, double* h_rows);
int vqc_solo_collect(const vqc_solo* s, const double* h_rows, double* energy2, double* grad);
void vqc_solo_free(vqc_solo* s);
