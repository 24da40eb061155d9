/*
 * Host-side internals shared by the C sources of the engine (engine.c,
 * covariance.c): the per-device engine state and small HIP helpers.
 */
#ifndef COFFE_ENGINE_HOST_H
#define COFFE_ENGINE_HOST_H

#include <pthread.h>
#include <stddef.h>

#include <hip/hip_runtime_api.h>

#include "coffe_engine.h"
#include "ce_internal.h"

#ifdef __cplusplus
extern "C" {
#endif

struct ce_grid;

struct coffe_engine {
    int device;
    hipStream_t stream;
    int threads;
    struct ce_grid *grids;
    pthread_mutex_t lock;
    /* scratch of the drop-in path */
    void *d_scratch;
    size_t scratch_bytes;
    /* pinned host staging for uploads of many small host arrays in one copy, two
       buffers used in turn, so that a call can be staged while the previous one's
       copy is still queued */
    void *h_pinned[2];
    size_t pinned_bytes[2];
    int stage_next;
    /* asynchronous users of the scratch: ev_scratch follows the last kernel that reads
       it, ev_stage[i] the last copy out of staging buffer i (valid when *_pending) */
    hipEvent_t ev_scratch, ev_stage[2];
    int scratch_pending, stage_pending[2];
};

/* COFFE_OK or COFFE_EDEVICE */
int ce_hip_ok(hipError_t e);
/* hipMalloc of at least 8 bytes: COFFE_OK or COFFE_ENOMEM */
int ce_dev_alloc(void **p, size_t bytes);
/* interleaved e^{-2 pi i k/N}, k < N, from long-double arguments */
void ce_twiddles(int N, double *out);
/* the thread's last error of the void drop-in entry points */
void ce_set_last_error(int code);
/* radix schedule of a length-M complex FFT (radices 8, 4, 2, 5, 3) */
int ce_radix_schedule(int M, ce_radix_plan *rp);
/* the same with radix-10 passes for the pairs of factors 2 and 5 (fewer passes for
   lengths such as 8000 = 8 * 10^3); for kernels that use fft_lds10 */
int ce_radix_schedule_ex(int M, ce_radix_plan *rp, int pair10);
/* q of twofast_selectqnu (twofast.c:240-250) */
double ce_selectqnu(double l, double nu);

#ifdef __cplusplus
}
#endif
#endif
