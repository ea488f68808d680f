/* farrsight_hip.h - C ABI of the MI355X (gfx950) implementation of the
 * FARRSIGHT hot path: the periodic regularized E-field (direct sum and
 * barycentric treecode) and the remesh interpolation of f onto the panel
 * tree, plus the device-resident step loop that calls them.
 *
 * Conventions (those of the reference's ElectricField::operator(),
 * inc/FieldStructure.hpp:87-93): the caller owns every buffer, outputs are
 * overwritten, counts are element counts. Every entry point returns 0 on
 * success or a negative FS_E_* code; fs_last_error() describes the last
 * failure on the calling thread. Entry points named *_dev take device
 * pointers and enqueue on the given HIP stream (NULL = the context's own
 * stream) without synchronizing; the others take host pointers and are
 * synchronous.
 */
#ifndef FARRSIGHT_HIP_H
#define FARRSIGHT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    FS_OK = 0,
    FS_E_INVALID = -1,   /* bad argument (null pointer, negative size, L <= 0) */
    FS_E_NOMEM = -2,     /* device allocation failed */
    FS_E_HIP = -3,       /* a HIP runtime call failed */
    FS_E_RANGE = -4,     /* input outside what the algorithm accepts
                            ("too many shifts!" of AMRStructure_interp.cpp:469) */
    FS_E_INTERNAL = -5   /* a device-side bounded search did not finish */
};

typedef struct fs_ctx fs_ctx;

/* Context: one HIP device, one default-priority stream, grow-only device
 * scratch. Not thread-safe; use one context per host thread. */
int fs_ctx_create(fs_ctx** out, int device);
int fs_ctx_destroy(fs_ctx* ctx);
/* The context's HIP stream (a hipStream_t). */
void* fs_ctx_stream(fs_ctx* ctx);
int fs_ctx_synchronize(fs_ctx* ctx);
const char* fs_last_error(void);

/* ---- Part A: periodic regularized E-field ------------------------------- */

/* Replaces E_MQ_DirectSum::operator() (src/FieldStructure.cpp:106-156):
 *   es[i] = sum_j q_ws[j] * (-z L) (1/sqrt(L^2+4eps^2) - 1/sqrt(4 z^2 L^2 + 4 eps^2)),
 *   z = (targets[i]-sources[j])/L wrapped to [-1/2, 1/2].
 * FP64 throughout, compensated accumulation. Host pointers, synchronous. */
int fs_emq_periodic_f64(fs_ctx* ctx, double* es, const double* targets, int64_t nt,
                        const double* sources, const double* q_ws, int64_t ns,
                        double L, double epsilon);

/* Same operator on device-resident arrays, enqueued on `stream`. */
int fs_emq_periodic_f64_dev(fs_ctx* ctx, double* d_es, const double* d_targets,
                            int64_t nt, const double* d_sources, const double* d_q_ws,
                            int64_t ns, double L, double epsilon, void* stream);

/* Tuning knobs of the direct-sum kernel (0 = automatic). targets_per_thread
 * in {1,2,4}; source_splits >= 1. Applies to later calls on this context. */
int fs_emq_set_tuning(fs_ctx* ctx, int targets_per_thread, int source_splits);

#ifdef __cplusplus
}
#endif
#endif /* FARRSIGHT_HIP_H */
