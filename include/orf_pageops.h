/*
 * orf_pageops.h -- page-batched FP64 complex primitives of the gfx950 RCWA
 * library, exported so they can be checked one by one.
 *
 * They replace the MATLAB page builtins the reference's RCWA code is built
 * on (host buffers in, host buffers out; column-major interleaved complex128,
 * pages stacked along the last dimension as in docs/CONVENTIONS.md:25):
 *   orf_pagemtimes   -> pagemtimes   (smatrixStar.m:7-10, smatrixLayer.m:13-14)
 *   orf_pagemldivide -> pagemldivide (smatrixLayer.m:6-7,11,14)
 *   orf_pagemrdivide -> pagemrdivide (smatrixStar.m:5-6)
 *   orf_pageeig      -> pageeig / eig per page (modesLayer.m:2-5)
 * A stride of 0 for an operand broadcasts one matrix over every page
 * (smatrixStar.m:3 "either operand may be a single 2-D S-matrix").
 */
#ifndef ORF_PAGEOPS_H
#define ORF_PAGEOPS_H
#include <stdint.h>
#include "orf_rcwa.h"
#ifdef __cplusplus
extern "C" {
#endif
/* C(:,:,p) = op(A(:,:,p)) * op(B(:,:,p)); op is 'N', 'T' or 'C'. */
int orf_pagemtimes(orf_ctx* ctx, char opA, char opB, int32_t m, int32_t n, int32_t k,
                   const double* A, int64_t strideA, const double* B, int64_t strideB, double* C,
                   int32_t batch);
/* The same product with FP32 MFMA products and FP32 accumulation (build
 * extension, the mixed-precision form's bulk GEMM): operands are rounded to
 * complex64, results returned as complex128.  No counterpart in the reference. */
int orf_pagemtimes_f32(orf_ctx* ctx, char opA, char opB, int32_t m, int32_t n, int32_t k,
                       const double* A, int64_t strideA, const double* B, int64_t strideB, double* C,
                       int32_t batch);
/* X(:,:,p) = A(:,:,p) \ B(:,:,p), A n x n, B n x nrhs (LU, partial pivoting). */
int orf_pagemldivide(orf_ctx* ctx, int32_t n, int32_t nrhs, const double* A, int64_t strideA,
                     const double* B, double* X, int32_t batch);
/* orf_pagemldivide on DEVICE buffers without allocating.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  dA (strideA 0 = one matrix for every page, or 2 n^2 doubles),
 * dB and dX (n x nrhs per page, dX may equal dB) are device pointers.  The LU
 * copies, pivots and diagonal-block inverses live in a fixed 64 MiB scratch
 * array per GPU (orf_init sets it up): batch * n^2 * 16 B plus the pivots and
 * inverses must fit, or ORF_INVALID_ARG.  stream NULL: runs on the context's
 * stream and returns when done; otherwise it is only enqueued on `stream`.
 * Calls that overlap on one GPU share the scratch and race. */
int orf_pagemldivide_noalloc(orf_ctx* ctx, int32_t n, int32_t nrhs, const double* dA, int64_t strideA,
                             const double* dB, double* dX, int32_t batch, void* stream);
/* X(:,:,p) = B(:,:,p) / A(:,:,p), B nrows x n. */
int orf_pagemrdivide(orf_ctx* ctx, int32_t n, int32_t nrows, const double* B, const double* A,
                     int64_t strideA, double* X, int32_t batch);
/* Eigenvalues w (n x batch) and unit-2-norm right eigenvectors V of each page. */
int orf_pageeig(orf_ctx* ctx, int32_t n, const double* A, double* w, double* V, int32_t batch);
#ifdef __cplusplus
}
#endif
#endif
