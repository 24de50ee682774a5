// Internal launch wrappers shared by the kernel files and the plan layer.
#pragma once
#include <hip/hip_runtime.h>
#include "orf_common.h"
#include <functional>

typedef std::function<void(const char*, int)> OrfTimeHook;  // (name, 0 = begin / 1 = end)

void orf_set_lds_limit(int bytes);
int orf_lds_limit();

// C = alpha op(A) op(B) + beta C, batched; op in {'N','T','C'} (opB also 'U': B as is and
// upper triangular, its lower part neither read nor used); stride 0 broadcasts.
void orf_zgemm(hipStream_t st, char opA, char opB, int M, int N, int K, cplx alpha, const cplx* A,
               int lda, long long sA, const cplx* B, int ldb, long long sB, cplx beta, cplx* C,
               int ldc, long long sC, int batch);
// The same product with FP32 MFMA products and FP32 accumulation (operands rounded
// to complex64 on load; alpha, beta and C in FP64); split: hi + lo FP32 operands,
// three FP32 products, FP64 accumulation across 4-deep MFMA steps.
void orf_zgemm_f32(hipStream_t st, char opA, char opB, int M, int N, int K, cplx alpha, const cplx* A,
                   int lda, long long sA, const cplx* B, int ldb, long long sB, cplx beta, cplx* C,
                   int ldc, long long sC, int batch, bool split = false);
// LU with partial pivoting, in place; ipiv [n x batch] 0-based; info [batch] (0 ok, k+1 zero pivot).
// dinv (optional, orf_dinv_stride(n) complex per page) receives the inverses of the 16x16
// diagonal blocks of L and U, which orf_zgetrs uses for its MFMA block solves.
void orf_zgetrf(hipStream_t st, int n, cplx* A, int lda, long long sA, int* ipiv, int* info,
                int batch, cplx* dinv);
void orf_zgetrf_only(hipStream_t st, int n, cplx* A, int lda, long long sA, int* ipiv, int* info,
                     int batch);
long long orf_dinv_stride(int n);
// side 'L': A X = B (B n x nrhs); side 'R': X A = B (B nrhs x n); B overwritten with X.
void orf_zgetrs(hipStream_t st, char side, int n, int nrhs, const cplx* LU, int lda, long long sA,
                const int* ipiv, const cplx* dinv, cplx* B, int ldb, long long sB, int batch);
// Eigen-decomposition; A (overwritten), w [n x batch], V (unit-norm eigenvectors),
// Zwork/Ywork n*n per page, Hwork (n*n + n) per page; info [batch].
void orf_zgeev(hipStream_t st, int n, cplx* A, int lda, long long sA, cplx* w, cplx* V, int ldv,
               long long sV, cplx* Zwork, cplx* Ywork, cplx* Hwork, int* info, int batch);
void orf_zgeev_timed(hipStream_t st, int n, cplx* A, int lda, long long sA, cplx* w, cplx* V,
                     int ldv, long long sV, cplx* Zwork, cplx* Ywork, cplx* Hwork, int* info,
                     int batch, const OrfTimeHook& hook, double* vnorm = nullptr);
// whether orf_eig_inverse handles size n
bool orf_eig_inverse_ok(int n);
bool orf_eig_inverse(hipStream_t st, int n, const cplx* Z, const cplx* Y, const double* nrm, cplx* X,
                     cplx* dinv, int batch);
// Inverses of many small matrices (n <= 22, one wave each); false (nothing
// launched) when n is larger.
bool orf_small_inverse(hipStream_t st, int n, const cplx* A, long long sA, cplx* X, long long sX,
                       long long count);

// Eigensolver kernel choices the autotuner may set (values: QR 0 per-rotation,
// 1 multishift; Hessenberg 0 rank-1 updates, 8 / 12 panel width; eigenvectors
// 16 / 32 rows per lane, 0 one wave per vector).
enum { ORF_TUNE_QR = 0, ORF_TUNE_HESS = 1, ORF_TUNE_TREVC = 2, ORF_TUNE_KINDS = 3 };
int orf_tune_value(int kind, int n, int batch);
void orf_tune_put(int kind, int n, int large, int value);
void orf_tune_reset();
void orf_tune_override(int kind, int value);
