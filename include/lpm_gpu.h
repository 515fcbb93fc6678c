/*
 * lpm_gpu.h -- C ABI of the MI355X (gfx950) spherical Biot-Savart / Poisson
 * sum.  Plain pointers and sizes only; no exceptions cross this boundary;
 * every entry point returns an lpm_status (0 = OK, < 0 = error) and leaves a
 * message retrievable with lpm_last_error().
 *
 * Particle data is SoA FP64 (x[], y[], z[]) exactly as the reference's Field
 * and Coords containers expose it (src/LpmField.h:34-51 getPtrToData).
 * Entry points without the _dev suffix take HOST pointers, copy to device,
 * run, copy back and synchronize.  The _dev variants take DEVICE pointers,
 * enqueue on the context's stream and return without synchronizing.
 *
 * A context owns one HIP stream, one growable device workspace and at most
 * one source octree.  Use one context per host thread.
 */
#ifndef LPM_GPU_H
#define LPM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lpm_ctx lpm_ctx;

enum lpm_status {
    LPM_OK = 0,
    LPM_ERR_INVALID_ARG = -1,  /* sizes, enums or parameters out of range */
    LPM_ERR_NULL_PTR = -2,     /* a required pointer is NULL */
    LPM_ERR_HIP = -3,          /* a HIP runtime call failed */
    LPM_ERR_ALLOC = -4,        /* device allocation failed (reference: std::bad_alloc) */
    LPM_ERR_STATE = -5,        /* call order violated (e.g. eval before build) */
    LPM_ERR_TREE_DEPTH = -6,   /* octree would exceed LPM_TREE_MAX_DEPTH levels */
    LPM_ERR_KERNEL = -7        /* unknown kernel id (reference: std::logic_error("invalid kernel")) */
};

/* kernel ids (bit mask for the direct sum) */
#define LPM_K_BIOT_SAVART_SPHERE 1 /* velocity, src/LpmVectorKernel.h:19-39 */
#define LPM_K_GREENS_SPHERE 2      /* streamfunction, src/LpmScalarKernel.h:31-45 */

/* arithmetic forms */
#define LPM_F64 0  /* FP64, the reference's per-pair formula */
#define LPM_F32C 1 /* FP32 pair math on tile-local offsets, |x-y|^2/2 denominator,
                      compensated FP32 tile sums, FP64 final reduction */

#define LPM_TREE_MAX_DEPTH 21

/* ------------------------------------------------------------------ context */
int lpm_create(lpm_ctx** ctx, int device);
void lpm_destroy(lpm_ctx* ctx);
const char* lpm_status_string(int status);
const char* lpm_last_error(const lpm_ctx* ctx);
int lpm_synchronize(lpm_ctx* ctx);
/* hipStream_t of the context (as void*), for callers that enqueue around it */
void* lpm_stream(lpm_ctx* ctx);

/* ------------------------------------------------------------------ direct sum
 * Replaces the pair loops of DirectSum::meshSolve (src/LpmDirectSum.cpp:25-173):
 *   velocity  u_i   = -1/(4 pi R) sum_j gamma_j (x_i x y_j) / (R^2 - x_i.y_j + delta^2)
 *             with a pair contributing 0 when |denominator| <= 1e-13
 *             (src/LpmVectorKernel.h:24-32);
 *   potential psi_i = -1/(4 pi R) sum_{j != skip_i} gamma_j log(R^2 - x_i.y_j)
 *             (src/LpmScalarKernel.h:35-38).
 * kernels  : LPM_K_BIOT_SAVART_SPHERE and/or LPM_K_GREENS_SPHERE (bit mask).
 * prec     : LPM_F64 or LPM_F32C.
 * gamma_j  : source strength (vorticity x area in the mesh solve).
 * tgt_self_src : per-target source index excluded from the potential sum
 *             (the face-target self pair of src/LpmDirectSum.cpp:52-75), -1 for
 *             none; NULL means no exclusions.  The velocity needs none: its
 *             self pair is exactly zero.
 * ux,uy,uz : required when the velocity kernel is requested, psi when the
 *             potential kernel is requested; others may be NULL.
 * nt or ns may be 0 (outputs are then zero / untouched respectively).
 */
int lpm_direct_sum(lpm_ctx* ctx, int kernels, int prec, double radius, double delta,
                   const double* tx, const double* ty, const double* tz, int64_t nt,
                   const double* sx, const double* sy, const double* sz, const double* gamma, int64_t ns,
                   const int64_t* tgt_self_src,
                   double* ux, double* uy, double* uz, double* psi);
int lpm_direct_sum_dev(lpm_ctx* ctx, int kernels, int prec, double radius, double delta,
                       const double* tx, const double* ty, const double* tz, int64_t nt,
                       const double* sx, const double* sy, const double* sz, const double* gamma, int64_t ns,
                       const int64_t* tgt_self_src,
                       double* ux, double* uy, double* uz, double* psi);

#ifdef __cplusplus
}
#endif

#endif /* LPM_GPU_H */
