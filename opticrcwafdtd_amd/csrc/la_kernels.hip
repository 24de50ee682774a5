// Batched FP64 complex linear algebra for the RCWA solve on gfx950.
//
// These kernels replace the page-batched MATLAB builtins the reference's
// solver is written with: pagemtimes (smatrixStar.m:5-10, smatrixLayer.m:13-14),
// pagemldivide / pagemrdivide (smatrixLayer.m:6-14, smatrixStar.m:5-6) and
// pageeig (modesLayer.m:2-5).  Every matrix is column-major interleaved
// complex128 (double2), one "page" per batch entry at a fixed stride; a stride
// of 0 broadcasts one matrix over the batch.
//
//  * zgemm_mfma:  64x64 output tile per 256-thread workgroup, 4 waves of 32x32,
//                 v_mfma_f64_16x16x4f64 with separate real/imag accumulators.
//  * zgetrf:      one workgroup per page, partial pivoting (max |re|+|im|, first
//                 index on ties, as LAPACK izamax), matrix staged in LDS when it fits.
//  * zgetrs:      left (A X = B) and right (X A = B) solves from the LU factors.
//  * eig:         Householder Hessenberg reduction (in LDS one column at a
//                 time; otherwise blocked panels with Level-3 trailing
//                 updates), its unitary factor formed as the starting Schur
//                 basis, complex QR to Schur form (Wilkinson / exceptional
//                 shifts, Ahues-Tisseur deflation; small matrices one
//                 rotation at a time in LDS, larger ones with several bulges
//                 chased in lockstep through an LDS window and aggressive
//                 early deflation on a small trailing window), eigenvectors of
//                 the triangular factor by back substitution (LAPACK ztrevc's
//                 small-pivot rule), back-transformed by one ZGEMM and
//                 normalised to unit 2-norm.
#include <mutex>
#include <vector>

#include "orf_common.h"
#include "orf_internal.h"

typedef double v4d __attribute__((ext_vector_type(4)));

// LDS visibility and execution barrier within one wavefront.
ORF_DEV void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Workgroups are dealt to the 8 XCDs round robin by linear id (x fastest).  For a
// grid of T tiles per page and P pages, page and tile of linear id L such that
// the T tiles of page p all run on XCD p % 8 and share its L2 (pages after the
// last full group of 8 keep the plain order).
ORF_DEV void xcd_page_tile(long long L, int T, int P, int& page, int& tile) {
  const long long full = (long long)(P / 8) * 8 * T;
  if (L < full) {
    const long long slot = L >> 3;
    page = (int)((slot / T) * 8 + (L & 7));
    tile = (int)(slot % T);
  } else {
    const long long r = L - full;
    page = (P / 8) * 8 + (int)(r / T);
    tile = (int)(r % T);
  }
}
// 2-D grid (x: tiles of a page, y: pages).
ORF_DEV void xcd_page_tile2(int& page, int& tile) {
  xcd_page_tile((long long)blockIdx.x + (long long)gridDim.x * blockIdx.y, gridDim.x, gridDim.y, page, tile);
}

// Copy of count elements from global memory into LDS (element e: src(e) into
// dst(e)) by threads t = 0..nt-1, each thread's loads in batches of 8 that are
// all in flight together: a plain loop that stores each element before loading
// the next waits out one memory latency per element.  Indices are clamped into
// range, so the loads need no branches.
template <class FS, class FD>
ORF_DEV void lds_fill(int t, int nt, int count, FS src, FD dst) {
  for (int e0 = t; e0 < count; e0 += 8 * nt) {
    cplx v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = src(min(e0 + u * nt, count - 1));
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (e0 + u * nt < count) dst(e0 + u * nt) = v[u];
  }
}

// ============================================================== ZGEMM (MFMA)
namespace {
constexpr int GT_M = 64, GT_N = 64, GT_K = 16, GT_PAD = 2;

template <int OP>
ORF_DEV cplx load_op(const cplx* __restrict__ X, int ld, int r, int c) {
  // element (r, c) of op(X)
  if (OP == 0) return X[(size_t)r + (size_t)c * ld];
  cplx v = X[(size_t)c + (size_t)r * ld];
  if (OP == 2) v.y = -v.y;
  return v;
}

// OPB = 3: B used as is (like 0) and upper triangular -- entries below its diagonal
// are taken as zero and never read, and a tile column stops its k loop at the
// diagonal (about half the products of a square triangular factor).
template <int OPA, int OPB>
__global__ __launch_bounds__(256, 2) void zgemm_mfma_kernel(
    int M, int N, int K, cplx alpha, const cplx* __restrict__ A, int lda, long long sA,
    const cplx* __restrict__ B, int ldb, long long sB, cplx beta, cplx* __restrict__ C, int ldc,
    long long sC) {
  __shared__ double Asr[GT_K][GT_M + GT_PAD], Asi[GT_K][GT_M + GT_PAD];
  __shared__ double Bsr[GT_K][GT_N + GT_PAD], Bsi[GT_K][GT_N + GT_PAD];
  int page, tile;
  xcd_page_tile((long long)blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z),
                gridDim.x * gridDim.y, gridDim.z, page, tile);
  A += (size_t)page * sA;
  B += (size_t)page * sB;
  C += (size_t)page * sC;
  const int m0 = (tile / gridDim.x) * GT_M, n0 = (tile % gridDim.x) * GT_N;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = (w & 1) * 32, wn = (w >> 1) * 32;
  // three real products per complex one (Gauss): P1 = Ar Br, P2 = Ai Bi,
  // P3 = (Ar + Ai)(Br + Bi); Re = P1 - P2, Im = P3 - P1 - P2
  v4d acc1[2][2], acc2[2][2], acc3[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      acc1[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
      acc2[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
      acc3[i][j] = (v4d){0.0, 0.0, 0.0, 0.0};
    }
  // this thread's staging slots: op(A)[ra_r[q], ra_k[q]] and op(B)[rb_k[q], rb_c[q]]
  int ra_r[4], ra_k[4], rb_k[4], rb_c[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (OPA == 0) { ra_r[q] = t & 63; ra_k[q] = (t >> 6) + 4 * q; }
    else { ra_k[q] = t & 15; ra_r[q] = (t >> 4) + 16 * q; }
    if (OPB == 0 || OPB == 3) { rb_k[q] = t & 15; rb_c[q] = (t >> 4) + 16 * q; }
    else { rb_c[q] = t & 63; rb_k[q] = (t >> 6) + 4 * q; }
  }
  constexpr int OB = OPB == 3 ? 0 : OPB;
  const int Kend = OPB == 3 ? min(K, n0 + GT_N) : K;
  // the next stage is fetched into registers while the current one feeds the MFMAs
  cplx pa[4], pb[4];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int gr = m0 + ra_r[q], gk = k0 + ra_k[q];
      pa[q] = (gr < M && gk < K) ? load_op<OPA>(A, lda, gr, gk) : cmk(0.0, 0.0);
      const int gk2 = k0 + rb_k[q], gc = n0 + rb_c[q];
      pb[q] = (gk2 < K && gc < N && (OPB != 3 || gk2 <= gc)) ? load_op<OB>(B, ldb, gk2, gc) : cmk(0.0, 0.0);
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < Kend; k0 += GT_K) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      Asr[ra_k[q]][ra_r[q]] = pa[q].x;
      Asi[ra_k[q]][ra_r[q]] = pa[q].y;
      Bsr[rb_k[q]][rb_c[q]] = pb[q].x;
      Bsi[rb_k[q]][rb_c[q]] = pb[q].y;
    }
    __syncthreads();
    if (k0 + GT_K < Kend) fetch(k0 + GT_K);
#pragma unroll
    for (int kk = 0; kk < GT_K; kk += 4) {
      const int kr = kk + (lane >> 4), rr = lane & 15;
      double ar[2], ai[2], br[2], bi[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ar[i] = Asr[kr][wm + 16 * i + rr];
        ai[i] = Asi[kr][wm + 16 * i + rr];
        br[i] = Bsr[kr][wn + 16 * i + rr];
        bi[i] = Bsi[kr][wn + 16 * i + rr];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          acc1[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], acc1[i][j], 0, 0, 0);
          acc2[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], bi[j], acc2[i][j], 0, 0, 0);
          acc3[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i] + ai[i], br[j] + bi[j], acc3[i][j], 0, 0, 0);
        }
    }
    __syncthreads();
  }
  // ---- epilogue: lane holds rows lane/16 + 4 v, column lane%16 of each 16x16 block
  const bool beta0 = (beta.x == 0.0 && beta.y == 0.0);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int gr = m0 + wm + 16 * i + (lane >> 4) + 4 * v;
        int gc = n0 + wn + 16 * j + (lane & 15);
        if (gr < M && gc < N) {
          const double p1 = acc1[i][j][v], p2 = acc2[i][j][v];
          cplx acc = cmk(p1 - p2, acc3[i][j][v] - p1 - p2);
          cplx out = cmul(alpha, acc);
          size_t idx = (size_t)gr + (size_t)gc * ldc;
          if (!beta0) out = cadd(out, cmul(beta, C[idx]));
          C[idx] = out;
        }
      }
}

// Same tiling with FP32 products and FP32 accumulation (v_mfma_f32_16x16x4f32):
// operands are read as complex128 and rounded to complex64 while staged, the
// 64x64 tile is accumulated in FP32, and alpha, beta and C stay FP64.  The
// FP32 MFMA output layout differs from the FP64 one: lane l, register v holds
// row 4 (l / 16) + v, column l % 16 of each 16x16 block.
// SPLIT: each operand is split into hi = fl32(x) and lo = fl32(x - hi), the
// product is Ah Bh + Ah Bl + Al Bh (three FP32 MFMA products, the lo * lo term
// dropped), accumulated in FP32 over one 4-deep MFMA step and in FP64 across
// steps: the operands' rounding to FP32 no longer enters, only the FP32
// rounding of each step's products and sums.
typedef float v4f __attribute__((ext_vector_type(4)));
template <int OPA, int OPB, bool SPLIT>
__global__ __launch_bounds__(256) void zgemm32_mfma_kernel(
    int M, int N, int K, cplx alpha, const cplx* __restrict__ A, int lda, long long sA,
    const cplx* __restrict__ B, int ldb, long long sB, cplx beta, cplx* __restrict__ C, int ldc,
    long long sC) {
  constexpr int PAD = 4;
  __shared__ float Asr[GT_K][GT_M + PAD], Asi[GT_K][GT_M + PAD];
  __shared__ float Bsr[GT_K][GT_N + PAD], Bsi[GT_K][GT_N + PAD];
  __shared__ float Alr[SPLIT ? GT_K : 1][GT_M + PAD], Ali[SPLIT ? GT_K : 1][GT_M + PAD];
  __shared__ float Blr[SPLIT ? GT_K : 1][GT_N + PAD], Bli[SPLIT ? GT_K : 1][GT_N + PAD];
  int page, tile;
  xcd_page_tile((long long)blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z),
                gridDim.x * gridDim.y, gridDim.z, page, tile);
  A += (size_t)page * sA;
  B += (size_t)page * sB;
  C += (size_t)page * sC;
  const int m0 = (tile / gridDim.x) * GT_M, n0 = (tile % gridDim.x) * GT_N;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = (w & 1) * 32, wn = (w >> 1) * 32;
  v4f accR[2][2], accI[2][2];
  double dR[SPLIT ? 2 : 1][SPLIT ? 2 : 1][4], dI[SPLIT ? 2 : 1][SPLIT ? 2 : 1][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      accR[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
      accI[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
      if (SPLIT)
#pragma unroll
        for (int v = 0; v < 4; ++v) dR[SPLIT ? i : 0][SPLIT ? j : 0][v] = dI[SPLIT ? i : 0][SPLIT ? j : 0][v] = 0.0;
    }
  for (int k0 = 0; k0 < K; k0 += GT_K) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int r, k;
      if (OPA == 0) { r = t & 63; k = (t >> 6) + 4 * q; }
      else { k = t & 15; r = (t >> 4) + 16 * q; }
      int gr = m0 + r, gk = k0 + k;
      cplx v = (gr < M && gk < K) ? load_op<OPA>(A, lda, gr, gk) : cmk(0.0, 0.0);
      Asr[k][r] = (float)v.x;
      Asi[k][r] = (float)v.y;
      if (SPLIT) {
        Alr[SPLIT ? k : 0][r] = (float)(v.x - (double)(float)v.x);
        Ali[SPLIT ? k : 0][r] = (float)(v.y - (double)(float)v.y);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int k, c;
      if (OPB == 0) { k = t & 15; c = (t >> 4) + 16 * q; }
      else { c = t & 63; k = (t >> 6) + 4 * q; }
      int gk = k0 + k, gc = n0 + c;
      cplx v = (gk < K && gc < N) ? load_op<OPB>(B, ldb, gk, gc) : cmk(0.0, 0.0);
      Bsr[k][c] = (float)v.x;
      Bsi[k][c] = (float)v.y;
      if (SPLIT) {
        Blr[SPLIT ? k : 0][c] = (float)(v.x - (double)(float)v.x);
        Bli[SPLIT ? k : 0][c] = (float)(v.y - (double)(float)v.y);
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GT_K; kk += 4) {
      const int kr = kk + (lane >> 4), rr = lane & 15;
      float ar[2], ai[2], br[2], bi[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ar[i] = Asr[kr][wm + 16 * i + rr];
        ai[i] = Asi[kr][wm + 16 * i + rr];
        br[i] = Bsr[kr][wn + 16 * i + rr];
        bi[i] = Bsi[kr][wn + 16 * i + rr];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[i], br[j], accR[i][j], 0, 0, 0);
          accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai[i], bi[j], accR[i][j], 0, 0, 0);
          accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[i], bi[j], accI[i][j], 0, 0, 0);
          accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[i], br[j], accI[i][j], 0, 0, 0);
        }
      if (SPLIT) {  // cross terms hi * lo + lo * hi
        float alr[2], ali[2], blr[2], bli[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          alr[i] = Alr[SPLIT ? kr : 0][wm + 16 * i + rr];
          ali[i] = Ali[SPLIT ? kr : 0][wm + 16 * i + rr];
          blr[i] = Blr[SPLIT ? kr : 0][wn + 16 * i + rr];
          bli[i] = Bli[SPLIT ? kr : 0][wn + 16 * i + rr];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[i], blr[j], accR[i][j], 0, 0, 0);
            accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(alr[i], br[j], accR[i][j], 0, 0, 0);
            accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ai[i], bli[j], accR[i][j], 0, 0, 0);
            accR[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(-ali[i], bi[j], accR[i][j], 0, 0, 0);
            accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[i], bli[j], accI[i][j], 0, 0, 0);
            accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(alr[i], bi[j], accI[i][j], 0, 0, 0);
            accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[i], blr[j], accI[i][j], 0, 0, 0);
            accI[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(ali[i], br[j], accI[i][j], 0, 0, 0);
          }
        // this 4-deep step's FP32 sums into the FP64 accumulators
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              dR[SPLIT ? i : 0][SPLIT ? j : 0][v] += (double)accR[i][j][v];
              dI[SPLIT ? i : 0][SPLIT ? j : 0][v] += (double)accI[i][j][v];
              accR[i][j][v] = 0.f;
              accI[i][j][v] = 0.f;
            }
      }
    }
    __syncthreads();
  }
  const bool beta0 = (beta.x == 0.0 && beta.y == 0.0);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int gr = m0 + wm + 16 * i + 4 * (lane >> 4) + v;
        int gc = n0 + wn + 16 * j + (lane & 15);
        if (gr < M && gc < N) {
          cplx acc = SPLIT ? cmk(dR[SPLIT ? i : 0][SPLIT ? j : 0][v], dI[SPLIT ? i : 0][SPLIT ? j : 0][v])
                           : cmk((double)accR[i][j][v], (double)accI[i][j][v]);
          cplx out = cmul(alpha, acc);
          size_t idx = (size_t)gr + (size_t)gc * ldc;
          if (!beta0) out = cadd(out, cmul(beta, C[idx]));
          C[idx] = out;
        }
      }
}
}  // namespace

void orf_zgemm_f32(hipStream_t st, char opA, char opB, int M, int N, int K, cplx alpha, const cplx* A,
                   int lda, long long sA, const cplx* B, int ldb, long long sB, cplx beta, cplx* C,
                   int ldc, long long sC, int batch, bool split) {
  if (M <= 0 || N <= 0 || batch <= 0) return;
  while (batch > 65535) {  // grid z limit
    orf_zgemm_f32(st, opA, opB, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, 65535, split);
    A += (size_t)sA * 65535;
    B += (size_t)sB * 65535;
    C += (size_t)sC * 65535;
    batch -= 65535;
  }
  dim3 grid((N + GT_N - 1) / GT_N, (M + GT_M - 1) / GT_M, batch);
  int oa = opA == 'N' ? 0 : (opA == 'T' ? 1 : 2);
  int ob = opB == 'N' ? 0 : (opB == 'T' ? 1 : 2);
#define ORF_GEMM_CASE(a, b)                                                                                 \
  if (oa == a && ob == b) {                                                                                 \
    if (split)                                                                                              \
      zgemm32_mfma_kernel<a, b, true><<<grid, 256, 0, st>>>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, \
                                                            ldc, sC);                                       \
    else                                                                                                    \
      zgemm32_mfma_kernel<a, b, false><<<grid, 256, 0, st>>>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta,  \
                                                             C, ldc, sC);                                   \
    return;                                                                                                 \
  }
  ORF_GEMM_CASE(0, 0) ORF_GEMM_CASE(0, 1) ORF_GEMM_CASE(0, 2)
  ORF_GEMM_CASE(1, 0) ORF_GEMM_CASE(1, 1) ORF_GEMM_CASE(1, 2)
  ORF_GEMM_CASE(2, 0) ORF_GEMM_CASE(2, 1) ORF_GEMM_CASE(2, 2)
#undef ORF_GEMM_CASE
}

// C = alpha A B + beta C for a few columns (N <= NC): one thread per row of A streams
// the row against B held in LDS (K x NC), sixteen loads in flight; A is read once.
// The 64 x 64 MFMA tiles would leave most of each tile empty and wait out a memory
// latency per 16-deep step.
namespace {
template <int NC>
__global__ __launch_bounds__(256) void zgemm_skinny_kernel(int M, int N, int K, cplx alpha, const cplx* __restrict__ A,
                                                           int lda, long long sA, const cplx* __restrict__ B,
                                                           int ldb, long long sB, cplx beta, cplx* __restrict__ C,
                                                           int ldc, long long sC) {
  extern __shared__ cplx Bs[];  // B (k, c) at k NC + c, zero for c >= N
  const int page = blockIdx.y, t = threadIdx.x, r = blockIdx.x * 256 + t;
  A += (size_t)page * sA;
  B += (size_t)page * sB;
  C += (size_t)page * sC;
  lds_fill(t, 256, K * NC,
           [&](int e) { return (e % NC) < N ? B[(size_t)(e / NC) + (size_t)(e % NC) * ldb] : cmk(0.0, 0.0); },
           [&](int e) -> cplx& { return Bs[e]; });
  __syncthreads();
  if (r >= M) return;
  cplx acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = cmk(0.0, 0.0);
  const cplx* ar = A + r;
  for (int k0 = 0; k0 < K; k0 += 16) {
    cplx a16[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) a16[u] = ar[(size_t)min(k0 + u, K - 1) * lda];
#pragma unroll
    for (int u = 0; u < 16; ++u)
      if (k0 + u < K)
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] = cadd(acc[c], cmul(a16[u], Bs[(k0 + u) * NC + c]));
  }
  const bool beta0 = beta.x == 0.0 && beta.y == 0.0;
#pragma unroll
  for (int c = 0; c < NC; ++c)
    if (c < N) {
      cplx out = cmul(alpha, acc[c]);
      if (!beta0) out = cadd(out, cmul(beta, C[(size_t)r + (size_t)c * ldc]));
      C[(size_t)r + (size_t)c * ldc] = out;
    }
}
}  // namespace

void orf_zgemm(hipStream_t st, char opA, char opB, int M, int N, int K, cplx alpha, const cplx* A,
               int lda, long long sA, const cplx* B, int ldb, long long sB, cplx beta, cplx* C,
               int ldc, long long sC, int batch) {
  if (M <= 0 || N <= 0 || batch <= 0) return;
  while (batch > 65535) {  // grid z limit
    orf_zgemm(st, opA, opB, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, 65535);
    A += (size_t)sA * 65535;
    B += (size_t)sB * 65535;
    C += (size_t)sC * 65535;
    batch -= 65535;
  }
  if (opA == 'N' && opB == 'N' && N <= 8 && K >= 1 && (size_t)K * 8 * sizeof(cplx) <= 65536) {
    dim3 g((M + 255) / 256, batch);
    if (N <= 4)
      zgemm_skinny_kernel<4><<<g, 256, (size_t)K * 4 * sizeof(cplx), st>>>(M, N, K, alpha, A, lda, sA, B, ldb, sB,
                                                                           beta, C, ldc, sC);
    else
      zgemm_skinny_kernel<8><<<g, 256, (size_t)K * 8 * sizeof(cplx), st>>>(M, N, K, alpha, A, lda, sA, B, ldb, sB,
                                                                           beta, C, ldc, sC);
    return;
  }
  dim3 grid((N + GT_N - 1) / GT_N, (M + GT_M - 1) / GT_M, batch);
  int oa = opA == 'N' ? 0 : (opA == 'T' ? 1 : 2);
  int ob = opB == 'N' ? 0 : (opB == 'T' ? 1 : (opB == 'U' ? 3 : 2));  // 'U': B upper triangular
#define ORF_GEMM_CASE(a, b)                                                                       \
  if (oa == a && ob == b) {                                                                       \
    zgemm_mfma_kernel<a, b><<<grid, 256, 0, st>>>(M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, \
                                                  ldc, sC);                                       \
    return;                                                                                       \
  }
  ORF_GEMM_CASE(0, 0) ORF_GEMM_CASE(0, 1) ORF_GEMM_CASE(0, 2)
  ORF_GEMM_CASE(1, 0) ORF_GEMM_CASE(1, 1) ORF_GEMM_CASE(1, 2)
  ORF_GEMM_CASE(2, 0) ORF_GEMM_CASE(2, 1) ORF_GEMM_CASE(2, 2)
  ORF_GEMM_CASE(0, 3) ORF_GEMM_CASE(1, 3) ORF_GEMM_CASE(2, 3)
#undef ORF_GEMM_CASE
}

// ============================================================== LU (getrf)
namespace {
// Block argmax of (value, index); ties resolve to the smaller index.
ORF_DEV void block_argmax(double& v, int& idx, double* rv, int* ri) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int o = 32; o > 0; o >>= 1) {
    double ov = __shfl_xor(v, o, 64);
    int oi = __shfl_xor(idx, o, 64);
    if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
  }
  __syncthreads();
  if (lane == 0) { rv[w] = v; ri[w] = idx; }
  __syncthreads();
  v = rv[0];
  idx = ri[0];
  for (int i = 1; i < nw; ++i)
    if (rv[i] > v || (rv[i] == v && ri[i] < idx)) { v = rv[i]; idx = ri[i]; }
}

__global__ __launch_bounds__(256) void zgetrf_kernel(int n, cplx* __restrict__ Ag, int lda,
                                                     long long sA, int* __restrict__ ipiv,
                                                     int* __restrict__ info, int use_lds) {
  extern __shared__ cplx lds_mat[];
  __shared__ double rv[4];
  __shared__ int ri[4];
  const int page = blockIdx.x, t = threadIdx.x, nt = blockDim.x;
  Ag += (size_t)page * sA;
  ipiv += (size_t)page * n;
  cplx* A = Ag;
  int ld = lda;
  if (use_lds) {
    A = lds_mat;
    ld = n;
    lds_fill(t, nt, n * n, [&](int e) { return Ag[(size_t)(e % n) + (size_t)(e / n) * lda]; },
             [&](int e) -> cplx& { return A[e]; });
    __syncthreads();
  }
  int sing = 0;
  for (int k = 0; k < n; ++k) {
    double best = -1.0;
    int bi = n;
    for (int i = k + t; i < n; i += nt) {
      double v = cabs1(A[i + (size_t)k * ld]);
      if (v > best) { best = v; bi = i; }
    }
    block_argmax(best, bi, rv, ri);
    const int p = bi;
    if (t == 0) ipiv[k] = p;
    if (best == 0.0) {
      if (!sing) sing = k + 1;
      __syncthreads();
      continue;
    }
    if (p != k) {
      for (int j = t; j < n; j += nt) {
        cplx a = A[k + (size_t)j * ld], b = A[p + (size_t)j * ld];
        A[k + (size_t)j * ld] = b;
        A[p + (size_t)j * ld] = a;
      }
    }
    __syncthreads();
    const cplx rp = crecip(A[k + (size_t)k * ld]);
    for (int i = k + 1 + t; i < n; i += nt) A[i + (size_t)k * ld] = cmul(A[i + (size_t)k * ld], rp);
    __syncthreads();
    // trailing rank-1 update, 64 rows x 4 columns per pass
    const int m = n - k - 1;
    const int ti = t & 63, tj = t >> 6;
    for (int j = tj; j < m; j += 4) {
      const cplx u = A[k + (size_t)(k + 1 + j) * ld];
      cplx* col = A + (size_t)(k + 1 + j) * ld;
      const cplx* lcol = A + (size_t)k * ld;
      for (int i = ti; i < m; i += 64) {
        cplx l = lcol[k + 1 + i];
        cplx a = col[k + 1 + i];
        col[k + 1 + i] = csub(a, cmul(l, u));
      }
    }
    __syncthreads();
  }
  if (t == 0 && info) info[page] = sing;
  if (use_lds) {
    for (int e = t; e < n * n; e += nt) {
      int r = e % n, c = e / n;
      Ag[(size_t)r + (size_t)c * lda] = A[e];
    }
  }
}

// Left solve A X = B with A = P L U (ipiv 0-based, LAPACK order).
// One workgroup per (column block of CB columns, page); the block lives in LDS.
__global__ __launch_bounds__(256) void zgetrs_left_kernel(int n, int nrhs, int CB,
                                                          const cplx* __restrict__ LU, int lda,
                                                          long long sA, const int* __restrict__ ipiv,
                                                          cplx* __restrict__ Bg, int ldb,
                                                          long long sB) {
  extern __shared__ cplx X[];  // n x CB, column-major with ld n
  const int page = blockIdx.y, c0 = blockIdx.x * CB, t = threadIdx.x, nt = blockDim.x;
  LU += (size_t)page * sA;
  ipiv += (size_t)page * n;
  Bg += (size_t)page * sB;
  const int nc = min(CB, nrhs - c0);
  lds_fill(t, nt, n * nc, [&](int e) { return Bg[(size_t)(e % n) + (size_t)(c0 + e / n) * ldb]; },
           [&](int e) -> cplx& { return X[e]; });
  __syncthreads();
  // row interchanges in order (one wave, CB <= 64 lanes)
  if (t < 64) {
    for (int k = 0; k < n; ++k) {
      int p = ipiv[k];
      if (p != k && t < nc) {
        cplx a = X[k + t * n];
        X[k + t * n] = X[p + t * n];
        X[p + t * n] = a;
      }
    }
  }
  __syncthreads();
  // forward: unit lower triangular
  for (int k = 0; k < n - 1; ++k) {
    const int m = n - k - 1;
    for (int e = t; e < m * nc; e += nt) {
      int i = k + 1 + e % m, c = e / m;
      X[i + c * n] = csub(X[i + c * n], cmul(LU[i + (size_t)k * lda], X[k + c * n]));
    }
    __syncthreads();
  }
  // backward: upper triangular
  for (int k = n - 1; k >= 0; --k) {
    if (t < nc) X[k + t * n] = cdiv(X[k + t * n], LU[k + (size_t)k * lda]);
    __syncthreads();
    for (int e = t; e < k * nc; e += nt) {
      int i = e % k, c = e / k;
      X[i + c * n] = csub(X[i + c * n], cmul(LU[i + (size_t)k * lda], X[k + c * n]));
    }
    __syncthreads();
  }
  for (int e = t; e < n * nc; e += nt) {
    int r = e % n, c = e / n;
    Bg[(size_t)r + (size_t)(c0 + c) * ldb] = X[r + c * n];
  }
}

// Right solve X A = B (B is nrows x n) with A = P L U.  One workgroup per
// (block of RB rows, page); the row block lives in LDS as RB x n column-major.
__global__ __launch_bounds__(256) void zgetrs_right_kernel(int n, int nrows, int RB,
                                                           const cplx* __restrict__ LU, int lda,
                                                           long long sA,
                                                           const int* __restrict__ ipiv,
                                                           cplx* __restrict__ Bg, int ldb,
                                                           long long sB) {
  extern __shared__ cplx X[];  // RB x n, ld RB
  const int page = blockIdx.y, r0 = blockIdx.x * RB, t = threadIdx.x, nt = blockDim.x;
  LU += (size_t)page * sA;
  ipiv += (size_t)page * n;
  Bg += (size_t)page * sB;
  const int nr = min(RB, nrows - r0);
  lds_fill(t, nt, nr * n, [&](int e) { return Bg[(size_t)(r0 + e % nr) + (size_t)(e / nr) * ldb]; },
           [&](int e) -> cplx& { return X[e % nr + (e / nr) * RB]; });
  __syncthreads();
  // Z U = B: column j final after dividing by U[j,j]; then B[:, j'>j] -= Z[:, j] U[j, j']
  for (int j = 0; j < n; ++j) {
    if (t < nr) X[t + j * RB] = cdiv(X[t + j * RB], LU[j + (size_t)j * lda]);
    __syncthreads();
    const int m = n - j - 1;
    for (int e = t; e < nr * m; e += nt) {
      int r = e % nr, c = j + 1 + e / nr;
      X[r + c * RB] = csub(X[r + c * RB], cmul(X[r + j * RB], LU[j + (size_t)c * lda]));
    }
    __syncthreads();
  }
  // Y L = Z (L unit lower): for j = n-1..0: Y[:, j] final; Z[:, j'<j] -= Y[:, j] L[j, j']
  for (int j = n - 1; j > 0; --j) {
    for (int e = t; e < nr * j; e += nt) {
      int r = e % nr, c = e / nr;
      X[r + c * RB] = csub(X[r + c * RB], cmul(X[r + j * RB], LU[j + (size_t)c * lda]));
    }
    __syncthreads();
  }
  // X = Y P^-1: column swaps in reverse order
  if (t < 64) {
    for (int k = n - 1; k >= 0; --k) {
      int p = ipiv[k];
      if (p != k) {
        for (int r = t; r < nr; r += 64) {
          cplx a = X[r + k * RB];
          X[r + k * RB] = X[r + p * RB];
          X[r + p * RB] = a;
        }
      }
    }
  }
  __syncthreads();
  for (int e = t; e < nr * n; e += nt) {
    int r = e % nr, c = e / nr;
    Bg[(size_t)(r0 + r) + (size_t)c * ldb] = X[r + c * RB];
  }
}

// ---------------------------------------------------------------- blocked kernels
// One 16x16 complex tile on MFMA: acc += A (16 x K) * B (K x 16), K = 4 * ksteps.
// fa(row, k) and fb(k, col) return operand entries (zero outside the matrix).
// Lane layout of v_mfma_f64_16x16x4f64: A[l%16][l/16], B[l/16][l%16],
// C/D rows l/16 + 4v (v = 0..3), column l%16.
// acc += A B on one 16x16 tile from KS fragments per operand, three real MFMA
// products per complex one (Gauss): P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi),
// Re += P1 - P2, Im += P3 - P1 - P2 (three short independent chains).
template <int KS>
ORF_DEV void ztile_mfma_frag3(const cplx* av, const cplx* bv, v4d& accR, v4d& accI) {
  v4d p1 = {0, 0, 0, 0}, p2 = {0, 0, 0, 0}, p3 = {0, 0, 0, 0};
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s].x, bv[s].x, p1, 0, 0, 0);
    p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s].y, bv[s].y, p2, 0, 0, 0);
    p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s].x + av[s].y, bv[s].x + bv[s].y, p3, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    accR[v] += p1[v] - p2[v];
    accI[v] += p3[v] - p1[v] - p2[v];
  }
}

template <int KS, class FA, class FB>
ORF_DEV void ztile_mfma(FA fa, FB fb, v4d& accR, v4d& accI) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, kq = lane >> 4;
  cplx av[KS], bv[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    av[s] = fa(r, 4 * s + kq);
    bv[s] = fb(4 * s + kq, r);
  }
  ztile_mfma_frag3<KS>(av, bv, accR, accI);
}

// The MFMA chain of ztile_mfma<4> on fragments already in registers.
ORF_DEV void ztile_mfma_frag(const cplx* av, const cplx* bv, v4d& accR, v4d& accI) {
  ztile_mfma_frag3<4>(av, bv, accR, accI);
}
// Right-looking blocked LU with partial pivoting, one workgroup per page.
// The LB-wide panel and the LB rows of U12 live in LDS; the trailing update
// A22 -= L21 U12 runs on MFMA tiles; row interchanges reach the other
// columns in one pass per panel (LAPACK zgetrf order: ipiv[k] is the row
// swapped with row k when column k is factored).
template <int LB>
__global__ __launch_bounds__(256) void zgetrf_blocked_kernel(int n, cplx* __restrict__ Ag, int lda,
                                                             long long sA, int* __restrict__ ipiv,
                                                             int* __restrict__ info) {
  extern __shared__ cplx sm[];
  __shared__ double rv[4];
  __shared__ int ri[4];
  __shared__ int spiv[LB];
  cplx* Pn = sm;                  // panel, (n - kb) x LB, ld n
  cplx* U = sm + (size_t)n * LB;  // LB x (n - kb - LB), row-major with ld n
  const int page = blockIdx.x, t = threadIdx.x, nt = blockDim.x;
  const int lane = t & 63, wv = t >> 6;
  cplx* A = Ag + (size_t)page * sA;
  ipiv += (size_t)page * n;
  int sing = 0;
  for (int kb = 0; kb < n; kb += LB) {
    const int nb = min(LB, n - kb), m = n - kb;
    lds_fill(t, nt, m * nb, [&](int e) { return A[(size_t)(kb + e % m) + (size_t)(kb + e / m) * lda]; },
             [&](int e) -> cplx& { return Pn[e % m + (size_t)(e / m) * n]; });
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
      double best = -1.0;
      int bi = m;
      for (int r = j + t; r < m; r += nt) {
        const double v = cabs1(Pn[r + (size_t)j * n]);
        if (v > best) { best = v; bi = r; }
      }
      block_argmax(best, bi, rv, ri);
      if (t == 0) { spiv[j] = bi; ipiv[kb + j] = kb + bi; }
      if (best == 0.0) {
        if (!sing) sing = kb + j + 1;
        __syncthreads();
        continue;
      }
      if (bi != j && t < nb) {
        const cplx x = Pn[j + (size_t)t * n];
        Pn[j + (size_t)t * n] = Pn[bi + (size_t)t * n];
        Pn[bi + (size_t)t * n] = x;
      }
      __syncthreads();
      // each thread scales its rows' multiplier and updates the rest of those rows
      // (row j is read only, so no barrier between the two)
      const cplx rp = crecip(Pn[j + (size_t)j * n]);
      for (int r = j + 1 + t; r < m; r += nt) {
        const cplx l = cmul(Pn[r + (size_t)j * n], rp);
        Pn[r + (size_t)j * n] = l;
        for (int c = j + 1; c < nb; ++c)
          Pn[r + (size_t)c * n] = csub(Pn[r + (size_t)c * n], cmul(l, Pn[j + (size_t)c * n]));
      }
      __syncthreads();
    }
    // panel back to global
    for (int e = t; e < m * nb; e += nt) {
      const int r = e % m, c = e / m;
      A[(size_t)(kb + r) + (size_t)(kb + c) * lda] = Pn[r + (size_t)c * n];
    }
    // other columns: apply the panel's interchanges; right of the panel also U12 = L11^-1 A12
    const int nr = n - kb - nb;
    for (int c = t; c < n - nb; c += nt) {
      const int col = c < kb ? c : c + nb;
      cplx* pc = A + (size_t)col * lda;
      for (int j = 0; j < nb; ++j) {
        const int p = kb + spiv[j];
        if (p != kb + j) {
          const cplx x = pc[kb + j];
          pc[kb + j] = pc[p];
          pc[p] = x;
        }
      }
      if (col >= kb + nb) {
        cplx x[LB];
#pragma unroll
        for (int j = 0; j < LB; ++j) x[j] = (j < nb) ? pc[kb + j] : cmk(0.0, 0.0);
#pragma unroll
        for (int j = 1; j < LB; ++j)
#pragma unroll
          for (int q = 0; q < j; ++q)
            if (j < nb) x[j] = csub(x[j], cmul(Pn[j + (size_t)q * n], x[q]));
#pragma unroll
        for (int j = 0; j < LB; ++j)
          if (j < nb) {
            pc[kb + j] = x[j];
            U[(size_t)j * n + (col - kb - nb)] = x[j];
          }
      }
    }
    __syncthreads();
    // trailing update on MFMA tiles; a wave's next tile of A is loaded while it
    // works on the current one, so the memory latency overlaps the MFMAs
    if (nr > 0) {
      const int tr = (nr + 15) / 16, ntl = tr * tr, step = nt >> 6;
      auto load_tile = [&](int tile, cplx* cv) {
        const int r0 = (tile % tr) * 16, cc = min((tile / tr) * 16 + (lane & 15), nr - 1);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int rr = min(r0 + (lane >> 4) + 4 * v, nr - 1);
          cv[v] = A[(size_t)(kb + nb + rr) + (size_t)(kb + nb + cc) * lda];
        }
      };
      cplx cv[4];
      if (wv < ntl) load_tile(wv, cv);
      for (int tile = wv; tile < ntl; tile += step) {
        cplx cn[4];
        load_tile(min(tile + step, ntl - 1), cn);
        const int r0 = (tile % tr) * 16, c0 = (tile / tr) * 16;
        const int cc = c0 + (lane & 15);
        v4d aR = {0, 0, 0, 0}, aI = {0, 0, 0, 0};
        ztile_mfma<LB / 4>(
                   [&](int r, int k) -> cplx {
                     return (r0 + r < nr && k < nb) ? Pn[(nb + r0 + r) + (size_t)k * n] : cmk(0.0, 0.0);
                   },
                   [&](int k, int c) -> cplx {
                     return (k < nb && c0 + c < nr) ? U[(size_t)k * n + c0 + c] : cmk(0.0, 0.0);
                   },
                   aR, aI);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int rr = r0 + (lane >> 4) + 4 * v;
          if (rr < nr && cc < nr)
            A[(size_t)(kb + nb + rr) + (size_t)(kb + nb + cc) * lda] = csub(cv[v], cmk(aR[v], aI[v]));
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) cv[v] = cn[v];
      }
    }
    __syncthreads();
  }
  if (t == 0 && info) info[page] = sing;
}

// Inverses of many small matrices (n <= SMALLINV_MAX): one wave per matrix,
// Gauss-Jordan with partial pivoting (largest |re|+|im|, first row on ties)
// on the augmented [A | I] held in LDS, four matrices per workgroup.  For the
// inverse rule's Toeplitz blocks (one per page and grid line) this replaces a
// 256-thread workgroup per matrix and a separate solve.
constexpr int SMALLINV_MAX = 22;
__global__ __launch_bounds__(256) void small_inverse_kernel(int n, const cplx* __restrict__ A, long long sA,
                                                            cplx* __restrict__ X, long long sX, long long count,
                                                            int* __restrict__ sing) {
  extern __shared__ cplx gjs[];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long long mat = (long long)blockIdx.x * 4 + wv;
  const int n2 = 2 * n;
  cplx* M = gjs + (size_t)wv * (n * n2 + n);  // row-major n x 2n, then the column of factors
  cplx* fcol = M + n * n2;
  if (mat >= count) return;  // whole waves only; no block barrier below
  A += (size_t)mat * sA;
  X += (size_t)mat * sX;
  for (int e = lane; e < n * n2; e += 64) {
    const int r = e / n2, c = e % n2;
    M[e] = c < n ? A[r + (size_t)c * n] : cmk(c - n == r ? 1.0 : 0.0, 0.0);
  }
  wave_sync_lds();
  // lanes cover rpi rows of 2n columns per pass
  const int rpi = 64 / n2, lr = lane / n2, lc = lane % n2;
  bool singular = false;
  for (int k = 0; k < n; ++k) {
    // pivot: lane r - k looks at row r
    const int r = k + lane;
    double v = (r < n) ? cabs1(M[r * n2 + k]) : -1.0;
    int idx = r;
    for (int o = 32; o > 0; o >>= 1) {
      const double v2 = __shfl_xor(v, o, 64);
      const int i2 = __shfl_xor(idx, o, 64);
      if (v2 > v || (v2 == v && i2 < idx)) { v = v2; idx = i2; }
    }
    if (v == 0.0) singular = true;
    if (idx != k)
      for (int c = lane; c < n2; c += 64) {
        const cplx x = M[k * n2 + c];
        M[k * n2 + c] = M[idx * n2 + c];
        M[idx * n2 + c] = x;
      }
    wave_sync_lds();
    const cplx rp = singular ? cmk(0.0, 0.0) : crecip(M[k * n2 + k]);
    // factors of every other row, read before row k or any other row changes
    if (lane < n) fcol[lane] = lane == k ? cmk(0.0, 0.0) : M[lane * n2 + k];
    wave_sync_lds();
    for (int c = lane; c < n2; c += 64) M[k * n2 + c] = cmul(M[k * n2 + c], rp);
    wave_sync_lds();
    if (lr < rpi)
      for (int rr = lr; rr < n; rr += rpi)
        if (rr != k) M[rr * n2 + lc] = csub(M[rr * n2 + lc], cmul(fcol[rr], M[k * n2 + lc]));
    wave_sync_lds();
  }
  for (int e = lane; e < n * n; e += 64) {
    const int r = e % n, c = e / n;
    X[r + (size_t)c * n] = M[r * n2 + n + c];
  }
  if (sing && lane == 0) sing[mat] = singular ? 1 : 0;
}

// Inverses of the 16x16 diagonal blocks of the LU factors, once per page:
// Linv (unit lower) and Uinv (upper), padded with the identity beyond n.
// Layout per page: [Linv blk 0..nblk-1][Uinv blk 0..nblk-1], 256 complex each.
__global__ __launch_bounds__(256) void diag_inv_kernel(int n, const cplx* __restrict__ LU, int lda,
                                                       long long sA, cplx* __restrict__ dinv, long long sD) {
  __shared__ cplx T[256], X[256];
  int page, blk;
  xcd_page_tile2(page, blk);
  const int t = threadIdx.x;
  const int nblk = (n + 15) / 16, kb = blk * 16, nb = min(16, n - kb);
  LU += (size_t)page * sA;
  cplx* out = dinv + (size_t)page * sD;
  const int r = t & 15, c = t >> 4;  // element (r, c)
  T[t] = (r < nb && c < nb) ? LU[(size_t)(kb + r) + (size_t)(kb + c) * lda] : cmk(r == c ? 1.0 : 0.0, 0.0);
  __syncthreads();
  // lower unit: column c of L^-1 by forward substitution, 16 lanes per column
  X[t] = cmk(r == c ? 1.0 : 0.0, 0.0);
  __syncthreads();
  for (int j = 1; j < 16; ++j) {
    // x[j] -= sum_{q<j} L[j][q] x[q]; lane r = q contributes
    cplx v = (r < j) ? cmul(T[j + 16 * r], X[r + 16 * c]) : cmk(0.0, 0.0);
    for (int o = 8; o > 0; o >>= 1) { v.x += __shfl_xor(v.x, o, 16); v.y += __shfl_xor(v.y, o, 16); }
    __syncthreads();
    if (r == j) X[j + 16 * c] = csub(X[j + 16 * c], v);
    __syncthreads();
  }
  out[(size_t)blk * 256 + t] = X[t];
  __syncthreads();
  // upper: column c of U^-1 by backward substitution
  X[t] = cmk(r == c ? 1.0 : 0.0, 0.0);
  __syncthreads();
  for (int j = 15; j >= 0; --j) {
    cplx v = (r > j) ? cmul(T[j + 16 * r], X[r + 16 * c]) : cmk(0.0, 0.0);
    for (int o = 8; o > 0; o >>= 1) { v.x += __shfl_xor(v.x, o, 16); v.y += __shfl_xor(v.y, o, 16); }
    __syncthreads();
    if (r == j) X[j + 16 * c] = cdiv(csub(X[j + 16 * c], v), T[j + 16 * j]);
    __syncthreads();
  }
  out[(size_t)(nblk + blk) * 256 + t] = X[t];
}

// Left solve A X = B with A = P L U: per 16-row block, X_blk = D^-1 X_blk then
// the rows below (above) are updated; both steps are MFMA tile products on an
// LDS block of CB right-hand sides.
__global__ __launch_bounds__(256, 2) void ztrsm_left_kernel(int n, int nrhs, int CB, const cplx* __restrict__ LU,
                                                         int lda, long long sA, const int* __restrict__ ipiv,
                                                         const cplx* __restrict__ dinv, long long sD,
                                                         cplx* __restrict__ Bg, int ldb, long long sB,
                                                         int upper_only, const cplx* __restrict__ Zt,
                                                         const double* __restrict__ rsc) {
  // Zt (optional): B = Zt^H (n x n pages at stride sB) instead of Bg's contents;
  // rsc (optional): rows of the result scaled by rsc (n per page)
  extern __shared__ cplx X[];  // n x CB, ld n
  int page, cbk;
  xcd_page_tile2(page, cbk);
  const int c0 = cbk * CB, t = threadIdx.x, nt = blockDim.x;
  const int lane = t & 63, wv = t >> 6, nw = nt >> 6;
  const int nblk = (n + 15) / 16;
  LU += (size_t)page * sA;
  Bg += (size_t)page * sB;
  const cplx* Linv = dinv + (size_t)page * sD;
  const cplx* Uinv = Linv + (size_t)nblk * 256;
  const int nc = min(CB, nrhs - c0);
  if (Zt) {  // X(r, c) = conj(Zt(c0 + c, r)): consecutive lanes read along a column of Zt
    const cplx* Zp = Zt + (size_t)page * sB;
    lds_fill(t, nt, n * nc, [&](int e) { return cconj(Zp[(size_t)(c0 + e % nc) + (size_t)(e / nc) * n]); },
             [&](int e) -> cplx& { return X[e / nc + (e % nc) * n]; });
  } else {
    lds_fill(t, nt, n * nc, [&](int e) { return Bg[(size_t)(e % n) + (size_t)(c0 + e / n) * ldb]; },
             [&](int e) -> cplx& { return X[e]; });
  }
  __syncthreads();
  if (!upper_only && t < 64) {
    ipiv += (size_t)page * n;
    for (int k = 0; k < n; ++k) {
      const int p = ipiv[k];
      if (p != k)
        for (int c = t; c < nc; c += 64) {
          const cplx a = X[k + c * n];
          X[k + c * n] = X[p + c * n];
          X[p + c * n] = a;
        }
    }
  }
  __syncthreads();
  const int tcn = (nc + 15) / 16;
  const int fr = lane & 15, kq = lane >> 4;
  // Block steps: forward with L (pass 0), then backward with U (pass 1), 16 rows each.
  // A step's inputs that do not depend on X -- the inverse of its diagonal block and
  // the LU rows of this wave's first four update tiles -- are fetched one step ahead,
  // while the step before runs (each step would otherwise wait out two memory
  // latencies before its products).
  const int p0 = upper_only ? 1 : 0;
  const int nsteps = (2 - p0) * nblk;
  cplx dnx[4], anx[4][4];
  auto fetch = [&](int step) {
    const int pass = p0 + step / nblk, bi = step % nblk;
    const int bk = pass == 0 ? bi : nblk - 1 - bi;
    const int kb = bk * 16, nb = min(16, n - kb);
    const cplx* D = (pass == 0 ? Linv : Uinv) + (size_t)bk * 256;
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) dnx[s4] = D[fr + 16 * (4 * s4 + kq)];
    const int rlo = pass == 0 ? kb + nb : 0, rhi = pass == 0 ? n : kb;
    const int trn = rhi > rlo ? (rhi - rlo + 15) / 16 : 1, ntile = rhi > rlo ? trn * tcn : 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int tile = wv + q * nw;
      const int r0 = rlo + (tile % trn) * 16;
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        const int k = 4 * s4 + kq;
        const bool ok = tile < ntile && r0 + fr < rhi && k < nb;
        const cplx a = LU[(size_t)min(r0 + fr, n - 1) + (size_t)min(kb + k, n - 1) * lda];
        anx[q][s4] = ok ? a : cmk(0.0, 0.0);
      }
    }
  };
  if (nsteps > 0) fetch(0);
  for (int step = 0; step < nsteps; ++step) {
    {
      const int pass = p0 + step / nblk, bi = step % nblk;
      const int bk = pass == 0 ? bi : nblk - 1 - bi;
      const int kb = bk * 16, nb = min(16, n - kb);
      cplx dcur[4], acur[4][4];
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        dcur[s4] = dnx[s4];
#pragma unroll
        for (int q = 0; q < 4; ++q) acur[q][s4] = anx[q][s4];
      }
      if (step + 1 < nsteps) fetch(step + 1);
      // X_blk = D X_blk (in place: each wave owns whole column tiles)
      for (int ct = wv; ct < tcn; ct += nw) {
        const int cc0 = ct * 16;
        cplx bv[4];
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
          const int k = 4 * s4 + kq;
          bv[s4] = (k < nb && cc0 + fr < nc) ? X[kb + k + (cc0 + fr) * n] : cmk(0.0, 0.0);
        }
        v4d aR = {0, 0, 0, 0}, aI = {0, 0, 0, 0};
        ztile_mfma_frag(dcur, bv, aR, aI);
        const int cc = cc0 + (lane & 15);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int rr = (lane >> 4) + 4 * v;
          if (rr < nb && cc < nc) X[kb + rr + cc * n] = cmk(aR[v], aI[v]);
        }
      }
      __syncthreads();
      // rows below (forward) or above (backward)
      const int rlo = pass == 0 ? kb + nb : 0, rhi = pass == 0 ? n : kb;
      const int rows = rhi - rlo;
      if (rows > 0) {
        // a wave's tiles four at a time (the first four with the fragments fetched
        // ahead, any further ones -- column blocks wider than 16 -- fetched here)
        const int trn = (rows + 15) / 16, ntile = trn * tcn;
        for (int tb = wv; tb < ntile; tb += 4 * nw) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int tile = tb + q * nw;
            if (tile >= ntile) break;
            const int r0 = rlo + (tile % trn) * 16, cc0 = (tile / trn) * 16;
            // LU rows: fetched ahead for the first four tiles, here for any further ones;
            // the block's rows of X per tile (the updates write other rows)
            cplx aq[4], bq[4];
            if (tb == wv) {
#pragma unroll
              for (int s4 = 0; s4 < 4; ++s4) aq[s4] = acur[q][s4];
            } else {
#pragma unroll
              for (int s4 = 0; s4 < 4; ++s4) {
                const int k = 4 * s4 + kq;
                aq[s4] = (r0 + fr < rhi && k < nb) ? LU[(size_t)(r0 + fr) + (size_t)(kb + k) * lda] : cmk(0.0, 0.0);
              }
            }
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
              const int k = 4 * s4 + kq;
              bq[s4] = (k < nb && cc0 + fr < nc) ? X[kb + k + (cc0 + fr) * n] : cmk(0.0, 0.0);
            }
            v4d aR = {0, 0, 0, 0}, aI = {0, 0, 0, 0};
            ztile_mfma_frag(aq, bq, aR, aI);
            const int cc = cc0 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int rr = r0 + (lane >> 4) + 4 * v;
              if (rr < rhi && cc < nc) X[rr + cc * n] = csub(X[rr + cc * n], cmk(aR[v], aI[v]));
            }
          }
        }
      }
      __syncthreads();
    }
  }
  const double* rs = rsc ? rsc + (size_t)page * n : nullptr;
  for (int e = t; e < n * nc; e += nt) {
    const int r = e % n, c = e / n;
    const cplx x = X[r + c * n];
    Bg[(size_t)r + (size_t)(c0 + c) * ldb] = rs ? cscale(x, rs[r]) : x;
  }
}

// Right solve X A = B (B nrows x n) with A = P L U; LDS block of RB rows.
// Z U = B left to right, then Y L = Z right to left, then the column swaps.
__global__ __launch_bounds__(256) void ztrsm_right_kernel(int n, int nrows, int RB, const cplx* __restrict__ LU,
                                                          int lda, long long sA, const int* __restrict__ ipiv,
                                                          const cplx* __restrict__ dinv, long long sD,
                                                          cplx* __restrict__ Bg, int ldb, long long sB) {
  extern __shared__ cplx X[];  // RB x n, ld RB
  int page, rbk;
  xcd_page_tile2(page, rbk);
  const int r0g = rbk * RB, t = threadIdx.x, nt = blockDim.x;
  const int lane = t & 63, wv = t >> 6, nw = nt >> 6;
  const int nblk = (n + 15) / 16;
  LU += (size_t)page * sA;
  ipiv += (size_t)page * n;
  Bg += (size_t)page * sB;
  const cplx* Linv = dinv + (size_t)page * sD;
  const cplx* Uinv = Linv + (size_t)nblk * 256;
  const int nr = min(RB, nrows - r0g);
  lds_fill(t, nt, nr * n, [&](int e) { return Bg[(size_t)(r0g + e % nr) + (size_t)(e / nr) * ldb]; },
           [&](int e) -> cplx& { return X[e % nr + (e / nr) * RB]; });
  __syncthreads();
  const int trr = (nr + 15) / 16;
  for (int pass = 0; pass < 2; ++pass) {  // 0: U left to right, 1: L right to left
    for (int bi = 0; bi < nblk; ++bi) {
      const int bk = pass == 0 ? bi : nblk - 1 - bi;
      const int kb = bk * 16, nb = min(16, n - kb);
      const cplx* D = (pass == 0 ? Uinv : Linv) + (size_t)bk * 256;
      // X_blk = X_blk D (in place: each wave owns whole row tiles)
      for (int rt = wv; rt < trr; rt += nw) {
        const int rr0 = rt * 16;
        v4d aR = {0, 0, 0, 0}, aI = {0, 0, 0, 0};
        ztile_mfma<4>(
            [&](int r, int k) -> cplx {
              return (rr0 + r < nr && k < nb) ? X[rr0 + r + (kb + k) * RB] : cmk(0.0, 0.0);
            },
            [&](int k, int c) -> cplx { return D[k + 16 * c]; }, aR, aI);
        const int cc = lane & 15;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int rr = rr0 + (lane >> 4) + 4 * v;
          if (rr < nr && cc < nb) X[rr + (kb + cc) * RB] = cmk(aR[v], aI[v]);
        }
      }
      __syncthreads();
      const int clo = pass == 0 ? kb + nb : 0, chi = pass == 0 ? n : kb;
      const int cols = chi - clo;
      if (cols > 0) {
        // a wave's tiles four at a time, every LU fragment fetched before the first MFMA
        const int tcc = (cols + 15) / 16, ntile = trr * tcc;
        const int fr = lane & 15, kq = lane >> 4;
        for (int tb = wv; tb < ntile; tb += 4 * nw) {
          cplx af[4][4], bf[4][4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int tile = tb + q * nw;
            const int rr0 = (tile % trr) * 16, cc0 = clo + (tile / trr) * 16;
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
              const int k = 4 * s4 + kq;
              af[q][s4] = (tile < ntile && rr0 + fr < nr && k < nb) ? X[rr0 + fr + (kb + k) * RB] : cmk(0.0, 0.0);
              bf[q][s4] = (tile < ntile && k < nb && cc0 + fr < chi) ? LU[(size_t)(kb + k) + (size_t)(cc0 + fr) * lda]
                                                                     : cmk(0.0, 0.0);
            }
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int tile = tb + q * nw;
            if (tile >= ntile) break;
            const int rr0 = (tile % trr) * 16, cc0 = clo + (tile / trr) * 16;
            v4d aR = {0, 0, 0, 0}, aI = {0, 0, 0, 0};
            ztile_mfma_frag(af[q], bf[q], aR, aI);
            const int cc = cc0 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int rr = rr0 + (lane >> 4) + 4 * v;
              if (rr < nr && cc < chi) X[rr + cc * RB] = csub(X[rr + cc * RB], cmk(aR[v], aI[v]));
            }
          }
        }
      }
      __syncthreads();
    }
  }
  if (t < 64) {
    for (int k = n - 1; k >= 0; --k) {
      const int p = ipiv[k];
      if (p != k)
        for (int r = t; r < nr; r += 64) {
          const cplx a = X[r + k * RB];
          X[r + k * RB] = X[r + p * RB];
          X[r + p * RB] = a;
        }
    }
  }
  __syncthreads();
  for (int e = t; e < nr * n; e += nt) {
    const int r = e % nr, c = e / nr;
    Bg[(size_t)(r0g + r) + (size_t)c * ldb] = X[r + c * RB];
  }
}
}  // namespace

static int g_lds_limit = 65536;
// QR kernel choice (fixed by measurement on the +-5 batch): up to this size the
// per-rotation LDS kernel serves small batches; early deflation from the bottom
// of the window
static constexpr int g_qr_small_n = 96;
static constexpr int g_qr_aed = 2;         // 0 off, 1 reordering, 2 bottom only, 3 whole blocks only

void orf_set_lds_limit(int bytes) { g_lds_limit = bytes; }
int orf_lds_limit() { return g_lds_limit; }

// Tuning table of the eigensolver's kernel choices (written by the autotuner,
// loaded by the host): per choice kind, entries (n, batch class, value); a
// lookup takes the entry of the largest tuned n <= n in the batch class
// (pages >= 512 or fewer), else the smallest tuned n; -1 = no entry, the
// built-in rule applies.  A forced value overrides the table (tuning runs).
namespace {
struct TuneEntry {
  int n, large, value;
};
std::mutex g_tune_mu;
std::vector<TuneEntry> g_tune[ORF_TUNE_KINDS];
int g_tune_force[ORF_TUNE_KINDS] = {-1, -1, -1};
}  // namespace

int orf_tune_value(int kind, int n, int batch) {
  if (kind < 0 || kind >= ORF_TUNE_KINDS) return -1;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (g_tune_force[kind] >= 0) return g_tune_force[kind];
  const int large = batch >= 512;
  int below = -1, belowN = -1, smallest = -1, smallestN = 1 << 30;
  for (const TuneEntry& e : g_tune[kind]) {
    if (e.large != large) continue;
    if (e.n <= n && e.n > belowN) { belowN = e.n; below = e.value; }
    if (e.n < smallestN) { smallestN = e.n; smallest = e.value; }
  }
  return below >= 0 ? below : smallest;
}

void orf_tune_put(int kind, int n, int large, int value) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (TuneEntry& e : g_tune[kind])
    if (e.n == n && e.large == large) { e.value = value; return; }
  g_tune[kind].push_back(TuneEntry{n, large, value});
}

void orf_tune_reset() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (auto& v : g_tune) v.clear();
  for (int& f : g_tune_force) f = -1;
}

void orf_tune_override(int kind, int value) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_force[kind] = value;
}

long long orf_dinv_stride(int n) { return 2LL * ((n + 15) / 16) * 256; }

static void launch_dinv(hipStream_t st, int n, const cplx* A, int lda, long long sA, cplx* dinv, int batch) {
  while (batch > 0) {
    const int b = batch > 65535 ? 65535 : batch;
    diag_inv_kernel<<<dim3((n + 15) / 16, b), 256, 0, st>>>(n, A, lda, sA, dinv, orf_dinv_stride(n));
    A += (size_t)sA * b;
    dinv += (size_t)orf_dinv_stride(n) * b;
    batch -= b;
  }
}

void orf_zgetrf(hipStream_t st, int n, cplx* A, int lda, long long sA, int* ipiv, int* info,
                int batch, cplx* dinv) {
  if (n <= 0 || batch <= 0) return;
  orf_zgetrf_only(st, n, A, lda, sA, ipiv, info, batch);
  if (dinv && n >= 32) launch_dinv(st, n, A, lda, sA, dinv, batch);
}

void orf_zgetrf_only(hipStream_t st, int n, cplx* A, int lda, long long sA, int* ipiv, int* info,
                     int batch) {
  if (n <= 0 || batch <= 0) return;
  size_t bytes = (size_t)n * n * sizeof(cplx);
  int use_lds = bytes <= (size_t)g_lds_limit;
  if (use_lds || n < 32) {
    zgetrf_kernel<<<batch, 256, use_lds ? bytes : 0, st>>>(n, A, lda, sA, ipiv, info, use_lds);
    return;
  }
  // panel + U rows in LDS: 2 * LB * n complex (measured: a 16-wide panel at one
  // workgroup per CU beats an 8-wide one at two)
  const size_t per = 2 * (size_t)n * sizeof(cplx);
  if (16 * per <= (size_t)g_lds_limit)
    zgetrf_blocked_kernel<16><<<batch, 256, 16 * per, st>>>(n, A, lda, sA, ipiv, info);
  else if (8 * per <= (size_t)g_lds_limit)
    zgetrf_blocked_kernel<8><<<batch, 256, 8 * per, st>>>(n, A, lda, sA, ipiv, info);
  else
    zgetrf_blocked_kernel<4><<<batch, 256, 4 * per, st>>>(n, A, lda, sA, ipiv, info);
}

// X_i = A_i^-1 for count small n x n matrices (column-major, strides sA, sX);
// false when n is too large for the one-wave kernel (nothing launched).
bool orf_small_inverse(hipStream_t st, int n, const cplx* A, long long sA, cplx* X, long long sX,
                       long long count) {
  if (n <= 0 || n > SMALLINV_MAX || count <= 0) return false;
  const size_t lds = (size_t)4 * (n * 2 * n + n) * sizeof(cplx);
  for (long long b0 = 0; b0 < count; b0 += 4LL * 65535 * 4) {
    const long long cnt = std::min<long long>(count - b0, 4LL * 65535 * 4);
    small_inverse_kernel<<<(unsigned)((cnt + 3) / 4), 256, lds, st>>>(n, A + (size_t)b0 * sA, sA,
                                                                      X + (size_t)b0 * sX, sX, cnt, nullptr);
  }
  return true;
}

static int rhs_block(int n) {
  int cb = g_lds_limit / (n * (int)sizeof(cplx));
  if (cb > 64) cb = 64;
  int p = 1;
  while (p * 2 <= cb) p *= 2;
  return p < 1 ? 1 : p;
}

void orf_zgetrs(hipStream_t st, char side, int n, int nrhs, const cplx* LU, int lda, long long sA,
                const int* ipiv, const cplx* dinv, cplx* B, int ldb, long long sB, int batch) {
  if (n <= 0 || nrhs <= 0 || batch <= 0) return;
  while (batch > 65535) {  // grid y limit
    orf_zgetrs(st, side, n, nrhs, LU, lda, sA, ipiv, dinv, B, ldb, sB, 65535);
    LU += (size_t)sA * 65535;
    ipiv += (size_t)n * 65535;
    if (dinv) dinv += (size_t)orf_dinv_stride(n) * 65535;
    B += (size_t)sB * 65535;
    batch -= 65535;
  }
  int cb = rhs_block(n);
  if (n < 32 || !dinv) {  // tiny systems: per-column substitution in LDS
    dim3 grid((nrhs + cb - 1) / cb, batch);
    if (side == 'L')
      zgetrs_left_kernel<<<grid, 256, (size_t)n * cb * sizeof(cplx), st>>>(n, nrhs, cb, LU, lda, sA,
                                                                            ipiv, B, ldb, sB);
    else
      zgetrs_right_kernel<<<grid, 256, (size_t)n * cb * sizeof(cplx), st>>>(n, nrhs, cb, LU, lda,
                                                                             sA, ipiv, B, ldb, sB);
    return;
  }
  if (cb > 32) cb = 32;  // 16-wide MFMA tiles; moderate LDS footprint keeps several blocks per CU
  if (cb > 16 && (nrhs <= 16 || (size_t)n * cb * sizeof(cplx) > (size_t)g_lds_limit / 2)) cb = 16;
  dim3 grid((nrhs + cb - 1) / cb, batch);
  const long long sD = orf_dinv_stride(n);
  if (side == 'L')
    ztrsm_left_kernel<<<grid, 256, (size_t)n * cb * sizeof(cplx), st>>>(n, nrhs, cb, LU, lda, sA, ipiv,
                                                                         dinv, sD, B, ldb, sB, 0, nullptr, nullptr);
  else
    ztrsm_right_kernel<<<grid, 256, (size_t)n * cb * sizeof(cplx), st>>>(n, nrhs, cb, LU, lda, sA, ipiv,
                                                                          dinv, sD, B, ldb, sB);
}

// W^-1 of eigenvector matrices W = Z Y diag(1 / nrm) from orf_zgeev (Z unitary
// Schur vectors, Y upper-triangular eigenvectors of the Schur form, nrm the
// column norms colnorm removed): X = diag(nrm) Y^-1 Z^H, one triangular solve
// instead of an LU and a full solve.  false (nothing launched) when n is too
// small for the blocked solve (orf_eig_inverse_ok).
bool orf_eig_inverse_ok(int n) { return n >= 32; }
bool orf_eig_inverse(hipStream_t st, int n, const cplx* Z, const cplx* Y, const double* nrm, cplx* X,
                     cplx* dinv, int batch) {
  if (!orf_eig_inverse_ok(n) || batch <= 0) return false;
  const long long n2 = (long long)n * n;
  for (int b0 = 0; b0 < batch; b0 += 65535) {
    const int bb = std::min(65535, batch - b0);
    const cplx* Zb = Z + (size_t)b0 * n2;
    const cplx* Yb = Y + (size_t)b0 * n2;
    cplx* Xb = X + (size_t)b0 * n2;
    cplx* Db = dinv + (size_t)b0 * orf_dinv_stride(n);
    launch_dinv(st, n, Yb, n, n2, Db, bb);
    int cb = rhs_block(n);
    if (cb > 32) cb = 32;
    if (cb > 16 && (size_t)n * cb * sizeof(cplx) > (size_t)g_lds_limit / 2) cb = 16;
    // the solve reads Z^H straight from Z and scales its rows by nrm on the way out
    ztrsm_left_kernel<<<dim3((n + cb - 1) / cb, bb), 256, (size_t)n * cb * sizeof(cplx), st>>>(
        n, n, cb, Yb, n, n2, nullptr, Db, orf_dinv_stride(n), Xb, n, n2, 1, Zb, nrm + (size_t)b0 * n);
  }
  return true;
}

// ============================================================== eigen-decomposition
namespace {
constexpr double ULP = 2.220446049250313e-16;   // LAPACK dlamch('P') = eps * base
constexpr double SAFMIN = 2.2250738585072014e-308;
constexpr int TREVC_SLOTS_H = 12;  // columns up to 768 rows in registers

// Householder reduction to Hessenberg form without forming Q: the reflectors
// H(k) = I - tau_k v_k v_k^H (v_k[k+1] = 1) go to V (column k, rows k+1..n-1)
// and tau; Q = H(0) H(1) ... is formed afterwards as the QR's starting Schur
// basis (hess_q_apply_kernel, form = 1).  The left update keeps each column in registers
// (one read and one write per column and step).
__global__ __launch_bounds__(256) void hessenberg_noq_kernel(int n, cplx* __restrict__ Ag, int lda,
                                                             long long sA, cplx* __restrict__ Vg,
                                                             cplx* __restrict__ taug, int use_lds) {
  extern __shared__ cplx lds_mat[];  // [the matrix when use_lds] + the current reflector (n)
  __shared__ double red[4];
  const int page = blockIdx.x, t = threadIdx.x, nt = blockDim.x;
  const int lane = t & 63, w = t >> 6;
  Ag += (size_t)page * sA;
  cplx* V = Vg + (size_t)page * n * n;
  cplx* tau = taug + (size_t)page * n;
  cplx* A = Ag;
  int ld = lda;
  cplx* v = lds_mat + (use_lds ? (size_t)n * n : 0);
  if (use_lds) {
    A = lds_mat;
    ld = n;
    lds_fill(t, nt, n * n, [&](int e) { return Ag[(size_t)(e % n) + (size_t)(e / n) * lda]; },
             [&](int e) -> cplx& { return A[e]; });
  }
  __syncthreads();
  for (int k = 0; k < n - 2; ++k) {
    const int m = n - k - 1;
    cplx* colk = A + (size_t)k * ld;
    double s2 = 0.0;
    for (int i = k + 2 + t; i < n; i += nt) s2 += cnorm2(colk[i]);
    s2 = block_sum(s2, red);
    const cplx alpha = colk[k + 1];
    if (s2 == 0.0 && alpha.y == 0.0) {
      if (t == 0) tau[k] = cmk(0.0, 0.0);
      __syncthreads();
      continue;
    }
    const double beta = -copysign(sqrt(cnorm2(alpha) + s2), alpha.x);
    const cplx tk = cmk((beta - alpha.x) / beta, -alpha.y / beta);
    const cplx scal = crecip(csub(alpha, cmk(beta, 0.0)));
    for (int i = t; i < m; i += nt) {
      const cplx vi = (i == 0) ? cmk(1.0, 0.0) : cmul(colk[k + 1 + i], scal);
      v[i] = vi;
      V[(size_t)k * n + k + 1 + i] = vi;
    }
    if (t == 0) tau[k] = tk;
    __syncthreads();
    for (int i = k + 1 + t; i < n; i += nt) colk[i] = (i == k + 1) ? cmk(beta, 0.0) : cmk(0.0, 0.0);
    // right: rows 0..n-1, columns k+1..n-1:  A -= tau (A v) v^H
    for (int r = t; r < n; r += nt) {
      cplx wa = cmk(0.0, 0.0);
      for (int j = 0; j < m; ++j) wa = cadd(wa, cmul(A[r + (size_t)(k + 1 + j) * ld], v[j]));
      wa = cmul(tk, wa);
      for (int j = 0; j < m; ++j) {
        cplx* pa = A + r + (size_t)(k + 1 + j) * ld;
        *pa = csub(*pa, cmul(wa, cconj(v[j])));
      }
    }
    __syncthreads();
    // left: rows k+1..n-1, columns k+1..n-1: A -= conj(tau) v (v^H A); a wave per column
    const cplx taub = cconj(tk);
    for (int j = k + 1 + w; j < n; j += nt >> 6) {
      cplx* col = A + (size_t)j * ld + k + 1;
      cplx x[TREVC_SLOTS_H];
      double zr = 0.0, zi = 0.0;
#pragma unroll
      for (int q = 0; q < TREVC_SLOTS_H; ++q) {
        const int i = lane + 64 * q;
        x[q] = (i < m) ? col[i] : cmk(0.0, 0.0);
        if (i < m) {
          const cplx p = cmulc(v[i], x[q]);
          zr += p.x;
          zi += p.y;
        }
      }
      zr = wave_sum(zr);
      zi = wave_sum(zi);
      const cplx f = cmul(taub, cmk(zr, zi));
#pragma unroll
      for (int q = 0; q < TREVC_SLOTS_H; ++q) {
        const int i = lane + 64 * q;
        if (i < m) col[i] = csub(x[q], cmul(v[i], f));
      }
    }
    __syncthreads();
  }
  if (use_lds) {
    for (int e = t; e < n * n; e += nt) Ag[(size_t)(e % n) + (size_t)(e / n) * lda] = A[e];
  }
}

// W = Q W with Q = H(0) H(1) ... H(n-3) from hessenberg_noq_kernel.  Columns
// of W are independent: 16 columns per workgroup in LDS, 16 lanes per column,
// reflectors applied last to first with no block barrier.
// Blocked Householder reduction to Hessenberg form for matrices that do not
// fit in LDS: the published panel algorithm behind LAPACK's zgehrd/zlahr2
// (Quintana-Orti & van de Geijn).  A panel of NB reflectors is generated with
// the trailing matrix left untouched except for one matrix-vector product
// per reflector (Y = A V T accumulated in Yg); the trailing matrix is then
// updated once per panel, from the right (A -= Y V^H) and from the left
// (A -= V T^H V^H A).  Compared with one rank-1 update pair per column this
// reads and writes the trailing matrix NB times less often.  The panel's
// reflectors are staged in LDS (dynamic, n x NB) next to the current column.
// Output layout as hessenberg_noq_kernel: H in A (zeros below the
// subdiagonal), reflector k in column k of Vg (v[k+1] = 1), tau[k].
// x - sum_{j < i} Y(r, j) sw[j] (Y column-major with ld n): the loads of Y go
// out eight at a time, all in flight together.
ORF_DEV cplx ydot_sub(const cplx* Y, int n, int r, int i, cplx x, const cplx* sw) {
  for (int j0 = 0; j0 < i; j0 += 8) {
    cplx y8[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) y8[u] = Y[(size_t)r + (size_t)min(j0 + u, i - 1) * n];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (j0 + u < i) x = csub(x, cmul(y8[u], sw[j0 + u]));
  }
  return x;
}

// Three workgroups per CU (at most 168 VGPRs; a 12-wide panel keeps the LDS of
// three within the CU for n up to 249): measured 4 % faster than 16-wide
// panels at two per CU, whose extra resident pages hide more of the HBM latency
// of the per-reflector mat-vec.
// 32 sums over a wave at once: each butterfly step a lane keeps half of the values it
// carries and adds its partner's copy of that half (CNT values left, partner lane
// l ^ O); once one value is left, plain xor steps.  The pairs added, and their order,
// are those of wave_sum, so each sum equals wave_sum's; lane l ends with value l >> 1.
template <int CNT, int O>
ORF_DEV void wave_sum_32(double (&v)[32], int lane) {
  if constexpr (O > 0) {
    if constexpr (CNT > 1) {
      constexpr int H = CNT / 2;
      const bool hi = (lane & O) != 0;
#pragma unroll
      for (int i = 0; i < H; ++i) {
        const double keep = hi ? v[H + i] : v[i], send = hi ? v[i] : v[H + i];
        v[i] = keep + __shfl_xor(send, O, 64);
      }
      wave_sum_32<H, O / 2>(v, lane);
    } else {
      v[0] += __shfl_xor(v[0], O, 64);
      wave_sum_32<1, O / 2>(v, lane);
    }
  }
}
// x of lane ln (ln uniform), as a uniform value
ORF_DEV double lane_value(double x, int ln) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, ln);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), ln);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int NB, int SL>  // SL: 64-row slots of a column in the left update (n <= 64 SL + 1)
__global__ __launch_bounds__(256, 3) void hessenberg_blocked_kernel(int n, cplx* __restrict__ Ag, int lda,
                                                                 long long sA, cplx* __restrict__ Vg,
                                                                 cplx* __restrict__ taug,
                                                                 cplx* __restrict__ Yg, long long sY) {
  extern __shared__ cplx lds_dyn[];
  cplx* Vs = lds_dyn;             // Vs[r + n j] = V(r, k0 + j), rows > k0 + j valid
  cplx* bv = lds_dyn + (size_t)n * NB;  // current column / reflector, indexed by row
  __shared__ cplx Ts[NB * NB];    // T(r, c) = Ts[r + NB c], upper triangular
  __shared__ cplx sw[NB], sw2[NB];
  __shared__ double red[4];
  const int page = blockIdx.x, t = threadIdx.x, nt = blockDim.x;
  const int lane = t & 63, w = t >> 6, nw = nt >> 6;
  cplx* A = Ag + (size_t)page * sA;
  cplx* V = Vg + (size_t)page * n * n;
  cplx* tau = taug + (size_t)page * n;
  cplx* Y = Yg + (size_t)page * sY;
#define AA(r, c) A[(size_t)(r) + (size_t)(c) * lda]
#define YY(r, j) Y[(size_t)(r) + (size_t)(j) * n]
#define VS(r, j) Vs[(size_t)(r) + (size_t)(j) * n]
  const int nref = n - 2;
  for (int k0 = 0; k0 < nref; k0 += NB) {
    const int ib = min(NB, nref - k0);
    for (int i = 0; i < ib; ++i) {
      const int c = k0 + i;
      // (a) column c with the panel's earlier reflectors: right (A Q: b -= Y V(c,:)^H), then left (Q^H b)
      if (t < i) sw[t] = cconj(VS(c, t));
      __syncthreads();
      for (int r = k0 + 1 + t; r < n; r += nt) bv[r] = ydot_sub(Y, n, r, i, AA(r, c), sw);
      __syncthreads();
      if (i > 0) {
        for (int j = w; j < i; j += nw) {  // w_j = V(:, j)^H b, a wave per reflector
          double sr = 0.0, si = 0.0;
          for (int r = k0 + j + 1 + lane; r < n; r += 64) {
            const cplx p = cmulc(VS(r, j), bv[r]);
            sr += p.x;
            si += p.y;
          }
          sr = wave_sum(sr);
          si = wave_sum(si);
          if (lane == 0) sw2[j] = cmk(sr, si);
        }
        __syncthreads();
        if (t < i) {  // sw = T^H w
          cplx s = cmk(0.0, 0.0);
          for (int l = 0; l <= t; ++l) s = cadd(s, cmulc(Ts[l + NB * t], sw2[l]));
          sw[t] = s;
        }
        __syncthreads();
        for (int r = k0 + 1 + t; r < n; r += nt) {
          cplx s = cmk(0.0, 0.0);
          const int jmax = min(i, r - k0);  // V(r, j) = 0 for r <= k0 + j
          for (int j = 0; j < jmax; ++j) s = cadd(s, cmul(VS(r, j), sw[j]));
          bv[r] = csub(bv[r], s);
        }
        __syncthreads();
      }
      // (b) reflector annihilating rows c+2.. of column c
      double s2 = 0.0;
      for (int r = c + 2 + t; r < n; r += nt) s2 += cnorm2(bv[r]);
      s2 = block_sum(s2, red);
      const cplx alpha = bv[c + 1];
      const bool triv = (s2 == 0.0 && alpha.y == 0.0);
      double beta = 0.0;
      cplx tk = cmk(0.0, 0.0), scal = cmk(0.0, 0.0);
      if (!triv) {
        beta = -copysign(sqrt(cnorm2(alpha) + s2), alpha.x);
        tk = cmk((beta - alpha.x) / beta, -alpha.y / beta);
        scal = crecip(csub(alpha, cmk(beta, 0.0)));
      }
      __syncthreads();  // every thread has read alpha before bv is overwritten
      for (int r = c + 1 + t; r < n; r += nt) {
        const cplx b = bv[r];
        const cplx vr = (r == c + 1) ? cmk(1.0, 0.0) : (triv ? cmk(0.0, 0.0) : cmul(b, scal));
        V[(size_t)c * n + r] = vr;
        VS(r, i) = vr;
        AA(r, c) = triv ? b : (r == c + 1 ? cmk(beta, 0.0) : cmk(0.0, 0.0));
        bv[r] = vr;
      }
      for (int r = k0 + 1 + t; r <= c; r += nt) AA(r, c) = bv[r];
      if (t == 0) tau[c] = tk;
      __syncthreads();
      // (c) Y(:, i) = tau (A v - Y (V^H v)) on rows k0+1..; T(:, i) = -tau T (V^H v), T(i, i) = tau
      for (int j = w; j < i; j += nw) {
        double sr = 0.0, si = 0.0;
        for (int r = c + 1 + lane; r < n; r += 64) {
          const cplx p = cmulc(VS(r, j), bv[r]);
          sr += p.x;
          si += p.y;
        }
        sr = wave_sum(sr);
        si = wave_sum(si);
        if (lane == 0) sw[j] = cmk(sr, si);
      }
      __syncthreads();
      for (int r = k0 + 1 + t; r < n; r += nt) {
        // the panel's HBM stream: sixteen columns per step keep sixteen loads in flight
        cplx y8[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) y8[u] = cmk(0.0, 0.0);
        int q = c + 1;
        for (; q + 15 < n; q += 16) {
          cplx a8[16];
#pragma unroll
          for (int u = 0; u < 16; ++u) a8[u] = AA(r, q + u);
#pragma unroll
          for (int u = 0; u < 16; ++u) y8[u & 7] = cadd(y8[u & 7], cmul(a8[u], bv[q + u]));
        }
        for (; q < n; ++q) y8[0] = cadd(y8[0], cmul(AA(r, q), bv[q]));
        cplx y = cadd(cadd(cadd(y8[0], y8[1]), cadd(y8[2], y8[3])), cadd(cadd(y8[4], y8[5]), cadd(y8[6], y8[7])));
        YY(r, i) = cmul(tk, ydot_sub(Y, n, r, i, y, sw));
      }
      if (t < i) {
        cplx s = cmk(0.0, 0.0);
        for (int l = t; l < i; ++l) s = cadd(s, cmul(Ts[t + NB * l], sw[l]));
        Ts[t + NB * i] = cneg(cmul(tk, s));
      }
      if (t == 0) Ts[i + NB * i] = tk;
      for (int l = i + 1 + t; l < NB; l += nt) Ts[l + NB * i] = cmk(0.0, 0.0);
      __syncthreads();
    }
    const int q0 = k0 + ib;
    // T of this panel into the unused rows 0..c of V's columns (for forming Q by panels)
    for (int e = t; e < ib * ib; e += nt) {
      const int l = e % ib, j = e / ib;
      if (l <= j) V[(size_t)(k0 + j) * n + l] = Ts[l + NB * j];
    }
    // (e) rows 0..k0 of Y: P = A(0:k0, k0+1:) V in Y's columns NB.., then Y = P T; a
    // work item takes a row and three reflectors, so the block of A is read ib / 3
    // times per panel instead of ib
    {
      constexpr int JG = 3;
      const int ngr = (ib + JG - 1) / JG;
      for (int e = t; e < (k0 + 1) * ngr; e += nt) {
        const int r = e % (k0 + 1), jb = (e / (k0 + 1)) * JG;
        cplx acc[JG];
#pragma unroll
        for (int jj = 0; jj < JG; ++jj) acc[jj] = cmk(0.0, 0.0);
        for (int q = k0 + jb + 1; q < n; q += 8) {
          cplx a8[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) a8[u] = AA(r, min(q + u, n - 1));
#pragma unroll
          for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int jj = 0; jj < JG; ++jj) {
              const int j = jb + jj;
              if (q + u < n && j < ib && q + u >= k0 + j + 1) acc[jj] = cadd(acc[jj], cmul(a8[u], VS(q + u, j)));
            }
        }
#pragma unroll
        for (int jj = 0; jj < JG; ++jj)
          if (jb + jj < ib) YY(r, NB + jb + jj) = acc[jj];
      }
    }
    __syncthreads();
    for (int e = t; e < (k0 + 1) * ib; e += nt) {
      const int r = e % (k0 + 1), j = e / (k0 + 1);
      cplx s = cmk(0.0, 0.0);
      for (int l = 0; l <= j; ++l) s = cadd(s, cmul(YY(r, NB + l), Ts[l + NB * j]));
      YY(r, j) = s;
    }
    __syncthreads();
    // (f) trailing columns from the right, every row: A(r, q) -= Y(r, :) V(q, :)^H
    if (q0 < n) {
      for (int r = t; r < n; r += nt) {
        cplx yr[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) yr[j] = j < ib ? YY(r, j) : cmk(0.0, 0.0);
        // eight columns' loads in flight before their stores (a store may alias a
        // later column's load, so a plain loop waits out one latency per column)
        for (int qb = q0; qb < n; qb += 8) {
          cplx a8[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) a8[u] = AA(r, min(qb + u, n - 1));
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int q = qb + u;
            if (q < n) {
              cplx s = cmk(0.0, 0.0);
#pragma unroll
              for (int j = 0; j < NB; ++j)
                if (j < ib) s = cadd(s, cmulc(VS(q, j), yr[j]));
              AA(r, q) = csub(a8[u], s);
            }
          }
        }
      }
    }
    // (g) rows 0..k0 of the panel's columns k0+1..k0+ib-1
    for (int e = t; e < (k0 + 1) * (ib - 1); e += nt) {
      const int r = e % (k0 + 1), cc = k0 + 1 + e / (k0 + 1);
      cplx s = cmk(0.0, 0.0);
      for (int l = 0; l < cc - k0; ++l) s = cadd(s, cmulc(VS(cc, l), YY(r, l)));
      AA(r, cc) = csub(AA(r, cc), s);
    }
    __syncthreads();
    // (h) trailing columns from the left, rows k0+1..: A -= V T^H (V^H A), a wave per
    // column held in SL 64-row slots (SL sized to n: idle slots cost registers)
    {
      cplx x[SL];
      for (int q = q0 + w; q < n; q += nw) {
#pragma unroll
        for (int sl = 0; sl < SL; ++sl) {
          const int r = k0 + 1 + lane + 64 * sl;
          x[sl] = (r < n) ? AA(r, q) : cmk(0.0, 0.0);
        }
        cplx part[NB];
#pragma unroll
        for (int j = 0; j < NB; ++j) part[j] = cmk(0.0, 0.0);
#pragma unroll
        for (int sl = 0; sl < SL; ++sl) {
          const int r = k0 + 1 + lane + 64 * sl;
          if (r < n) {
#pragma unroll
            for (int j = 0; j < NB; ++j)
              if (j < ib && r >= k0 + j + 1) part[j] = cadd(part[j], cmulc(VS(r, j), x[sl]));
          }
        }
        // the 2 NB sums over the wave by a butterfly (wave_sum_32): 62 shuffles for
        // all of them instead of 6 per value, the same pairs added in the same order
        {
          double v[32];
#pragma unroll
          for (int j = 0; j < 16; ++j) {
            v[2 * j] = j < NB ? part[j].x : 0.0;
            v[2 * j + 1] = j < NB ? part[j].y : 0.0;
          }
          wave_sum_32<32, 32>(v, lane);
#pragma unroll
          for (int j = 0; j < NB; ++j) part[j] = cmk(lane_value(v[0], 4 * j), lane_value(v[0], 4 * j + 2));
        }
        // W := T^H W: lane j < NB forms entry j (sum over l <= j, l ascending), then the
        // entries are read out uniform again
        {
          const int jl = lane < NB ? lane : NB - 1;
          cplx s = cmk(0.0, 0.0);
#pragma unroll
          for (int l = 0; l < NB; ++l)
            if (l <= jl) s = cadd(s, cmulc(Ts[l + NB * jl], part[l]));
          if (jl >= ib) s = cmk(0.0, 0.0);
#pragma unroll
          for (int j = 0; j < NB; ++j) part[j] = cmk(lane_value(s.x, j), lane_value(s.y, j));
        }
#pragma unroll
        for (int sl = 0; sl < SL; ++sl) {
          const int r = k0 + 1 + lane + 64 * sl;
          if (r < n) {
            cplx s = cmk(0.0, 0.0);
#pragma unroll
            for (int j = 0; j < NB; ++j)
              if (j < ib && r >= k0 + j + 1) s = cadd(s, cmul(VS(r, j), part[j]));
            AA(r, q) = csub(x[sl], s);
          }
        }
      }
    }
    __syncthreads();
  }
#undef AA
#undef YY
#undef VS
}

// Z = I on every page.
__global__ void eye_pages_kernel(int n, cplx* __restrict__ Z, long long sZ) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * n) return;
  Z[(size_t)blockIdx.y * sZ + e] = cmk((e % n) == (e / n) ? 1.0 : 0.0, 0.0);
}

// Clean copies of one panel of the blocked reduction for the GEMM formation of
// Q: Vc (m x ib, rows k0+1.., unit lower trapezoid) and Tc (ib x ib upper).
__global__ void panel_extract_kernel(int n, int k0, int ib, const cplx* __restrict__ Vg, cplx* __restrict__ Vc,
                                     long long sVc, cplx* __restrict__ Tc, long long sTc) {
  const int page = blockIdx.y;
  const int m = n - k0 - 1;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const cplx* V = Vg + (size_t)page * n * n;
  if (e < (long long)m * ib) {
    const int r = (int)(e % m), j = (int)(e / m);
    cplx v = cmk(0.0, 0.0);
    if (r > j) v = V[(size_t)(k0 + j) * n + k0 + 1 + r];
    else if (r == j) v = cmk(1.0, 0.0);
    Vc[(size_t)page * sVc + e] = v;
  } else if (Tc && e < (long long)m * ib + ib * ib) {
    const int q = (int)(e - (long long)m * ib), l = q % ib, j = q / ib;
    Tc[(size_t)page * sTc + q] = (l <= j) ? V[(size_t)(k0 + j) * n + l] : cmk(0.0, 0.0);
  }
}

// T of two consecutive panels (a at ka, b at ka + iba) placed on the diagonal of
// the (iba + ibb)^2 block T2 (ld iba + ibb), zero below; the upper-right block
// -Ta (Va^H Vb) Tb is filled in afterwards by two small products.
__global__ void panel_pair_t_kernel(int n, int ka, int iba, int ibb, const cplx* __restrict__ Vg,
                                    cplx* __restrict__ T2, long long sT2) {
  const int page = blockIdx.y, ib = iba + ibb;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ib * ib) return;
  const int l = q % ib, j = q / ib;
  const cplx* V = Vg + (size_t)page * n * n;
  cplx v = cmk(0.0, 0.0);
  if (j < iba && l <= j) v = V[(size_t)(ka + j) * n + l];
  else if (j >= iba && l >= iba && l <= j) v = V[(size_t)(ka + j) * n + (l - iba)];
  T2[(size_t)page * sT2 + q] = v;
}

// T2 (ib x ib, ld ib, ib = iba + ibb) of the block reflector of panel a (iba
// reflectors from column ka) and panel b (ibb from ka + iba; ibb = 0: panel a alone):
// T2 = [Ta, -Ta (Va^H Vb) Tb; 0, Tb], one workgroup per page.  Ta and Tb come from the
// rows above the panels' reflectors, where the reduction keeps them, Va and Vb from
// its reflector storage (unit lower trapezoids from row ka + 1).
__global__ __launch_bounds__(256) void hess_pair_t_kernel(int n, int ka, int iba, int ibb, const cplx* __restrict__ Vg,
                                                          cplx* __restrict__ T2g, long long sT2) {
  __shared__ cplx Ts[32 * 32];  // T2 (l, j) at l + 32 j
  __shared__ cplx Ss[16 * 16];  // S = Va^H Vb, then W = S Tb, (i, j) at i + 16 j
  const int page = blockIdx.x, t = threadIdx.x, ib = iba + ibb, m = n - ka - 1;
  const cplx* V = Vg + (size_t)page * n * n;
  cplx* T2 = T2g + (size_t)page * sT2;
  for (int e = t; e < 32 * 32; e += 256) {
    const int l = e & 31, j = e >> 5;
    cplx v = cmk(0.0, 0.0);
    if (j < iba && l <= j) v = V[(size_t)(ka + j) * n + l];
    else if (j >= iba && j < ib && l >= iba && l <= j) v = V[(size_t)(ka + j) * n + (l - iba)];
    Ts[e] = v;
  }
  if (ibb > 0) {
    // S(i, j) = sum_r conj(Va(r, i)) Vb(r, j): Vb is zero above row iba + j and one
    // there, where Va is below its diagonal (i < iba)
    if (t < iba * ibb) {
      const int i = t % iba, j = t / iba;
      const cplx* va = V + (size_t)(ka + i) * n + ka + 1;
      const cplx* vb = V + (size_t)(ka + iba + j) * n + ka + 1;
      const int r0 = iba + j;
      cplx acc = cconj(va[r0]);
      cplx acc8[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc8[u] = cmk(0.0, 0.0);
      for (int r = r0 + 1; r < m; r += 8) {
        cplx a8[8], b8[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          a8[u] = va[min(r + u, m - 1)];
          b8[u] = vb[min(r + u, m - 1)];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (r + u < m) acc8[u] = cadd(acc8[u], cmulc(a8[u], b8[u]));
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) acc = cadd(acc, acc8[u]);
      Ss[i + 16 * j] = acc;
    }
    __syncthreads();
    // W = S Tb
    cplx w = cmk(0.0, 0.0);
    const int i = t % 16, j = t / 16;
    if (i < iba && j < ibb)
      for (int l = 0; l <= j; ++l) w = cadd(w, cmul(Ss[i + 16 * l], Ts[(iba + l) + 32 * (iba + j)]));
    __syncthreads();
    if (i < iba && j < ibb) Ss[i + 16 * j] = w;
    __syncthreads();
    // T2(0:iba, iba:) = -Ta W
    if (i < iba && j < ibb) {
      cplx x = cmk(0.0, 0.0);
      for (int l = i; l < iba; ++l) x = cadd(x, cmul(Ts[i + 32 * l], Ss[l + 16 * j]));
      Ts[i + 32 * (iba + j)] = cneg(x);
    }
  }
  __syncthreads();
  for (int e = t; e < ib * ib; e += 256) T2[e] = Ts[(e % ib) + 32 * (e / ib)];
}

// One pass of the GEMM formation of Q: Q22 := (I - V T V^H) Q22 (Q22 m x m, ld ldq;
// V m x ib unit lower trapezoid, ld ldv, read from the rows below its diagonal only --
// the reduction's own storage of the reflectors serves; T ib x ib upper, ld ldt;
// ib <= 32) with one read and one write of Q22.  A workgroup takes a 16-column
// block; wave w holds its rows 16 (RTW w + i) + 0..15, i < RTW, in registers.  X = V^H Q_blk by MFMA over each
// wave's rows, the four waves' parts summed in LDS, Y = T X, then Q_blk -= V Y by
// MFMA.  The lane that feeds Q(r, c) to the X products as B operand (row 4 v + lane/16
// of a 16-row tile, column lane % 16) holds the same element in the update's output
// layout, so the block stays in registers between the two products.
template <int RTW, int OCC>
__global__ __launch_bounds__(256, OCC) void hess_q_pass_kernel(int m, int ib, const cplx* __restrict__ Vg, int ldv,
                                                             long long sV, const cplx* __restrict__ Tg, int ldt,
                                                             long long sT, cplx* __restrict__ Qg, int ldq,
                                                             long long sQ, int pages) {
  // per-wave parts of X (j, c) at j + XL c (odd stride: the 16 columns of one row fall
  // in different banks); their sum goes to Xw[0], Y = T X to Xw[1]
  constexpr int XL = 33;
  __shared__ cplx Xw[4][XL * 16];
  __shared__ cplx Ts[32 * 32];  // T (j, l) at j + 32 l, zero outside ib x ib
  cplx* Ys = Xw[1];
  const int nblk = (m + 15) / 16;
  int page, blk;
  xcd_page_tile(blockIdx.x, nblk, pages, page, blk);
  const cplx* V = Vg + (size_t)page * sV;
  const cplx* T = Tg + (size_t)page * sT;
  cplx* Q = Qg + (size_t)page * sQ;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int cl = lane & 15, rq = lane >> 4;
  const int col = blk * 16 + cl;
  const int colc = min(col, m - 1);
  for (int e = t; e < 32 * 32; e += 256) {
    const int j = e & 31, l = e >> 5;
    Ts[e] = (j <= l && l < ib) ? T[(size_t)j + (size_t)l * ldt] : cmk(0.0, 0.0);
  }
  cplx q[RTW][4];
#pragma unroll
  for (int i = 0; i < RTW; ++i)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = 16 * (RTW * w + i) + 4 * v + rq;
      const cplx x = Q[(size_t)min(r, m - 1) + (size_t)colc * ldq];
      q[i][v] = (r < m && col < m) ? x : cmk(0.0, 0.0);
    }
  // ---- X = V^H Q_blk over this wave's rows (three real products per complex one)
  const int njt = ib > 16 ? 2 : 1;
  v4d a1[2], a2[2], a3[2];
#pragma unroll
  for (int jt = 0; jt < 2; ++jt) {
    a1[jt] = (v4d){0.0, 0.0, 0.0, 0.0};
    a2[jt] = (v4d){0.0, 0.0, 0.0, 0.0};
    a3[jt] = (v4d){0.0, 0.0, 0.0, 0.0};
  }
#pragma unroll
  for (int i = 0; i < RTW; ++i) {
    if (16 * (RTW * w + i) >= m) break;
    asm volatile("" ::: "memory");  // one tile's loads in flight at a time (registers)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = 16 * (RTW * w + i) + 4 * v + rq;
      const double br = q[i][v].x, bi = q[i][v].y;
#pragma unroll
      for (int jt = 0; jt < 2; ++jt) {
        if (jt >= njt) break;
        const int j = 16 * jt + cl;
        const cplx x = V[(size_t)min(r, m - 1) + (size_t)min(j, ib - 1) * ldv];
        const bool ok = r < m && j < ib && r >= j;
        const cplx xv = r == j ? cmk(1.0, 0.0) : x;
        const double ar = ok ? xv.x : 0.0, ai = ok ? -xv.y : 0.0;  // conj(V)
        a1[jt] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, a1[jt], 0, 0, 0);
        a2[jt] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bi, a2[jt], 0, 0, 0);
        a3[jt] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar + ai, br + bi, a3[jt], 0, 0, 0);
      }
    }
  }
  // output layout: lane holds X (16 jt + lane/16 + 4 u, lane % 16)
#pragma unroll
  for (int jt = 0; jt < 2; ++jt)
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double p1 = a1[jt][u], p2 = a2[jt][u];
      Xw[w][16 * jt + rq + 4 * u + XL * cl] = cmk(p1 - p2, a3[jt][u] - p1 - p2);
    }
  __syncthreads();
  for (int e = t; e < XL * 16; e += 256) Xw[0][e] = cadd(cadd(Xw[0][e], Xw[1][e]), cadd(Xw[2][e], Xw[3][e]));
  __syncthreads();
  // ---- Y = T X (T upper triangular; Ts and the rows of X from ib on are zero, so all
  // 32 terms go in unconditionally and the LDS loads pipeline)
  for (int e = t; e < 32 * 16; e += 256) {
    const int j = e & 31, c = e >> 5;
    cplx y = cmk(0.0, 0.0);
#pragma unroll 8
    for (int l = 0; l < 32; ++l) y = cadd(y, cmul(Ts[j + 32 * l], Xw[0][l + XL * c]));
    Ys[j + XL * c] = y;
  }
  __syncthreads();
  // ---- Q_blk -= V Y; output layout: lane holds row 16 tile + lane/16 + 4 v, column lane % 16
#pragma unroll
  for (int i = 0; i < RTW; ++i) {
    const int rb = 16 * (RTW * w + i);
    if (rb >= m) break;
    asm volatile("" ::: "memory");
    v4d u1 = (v4d){0.0, 0.0, 0.0, 0.0}, u2 = u1, u3 = u1;
    const int r = rb + cl;
    // all eight 4-deep steps with clamped loads, issued together (Y is zero from row
    // ib on): a loop bounded by ib waits out one memory latency per step
    cplx vx[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) vx[s] = V[(size_t)min(r, m - 1) + (size_t)min(4 * s + rq, ib - 1) * ldv];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int k = 4 * s + rq;
      const bool ok = r < m && k < ib && r >= k;
      const cplx xv = r == k ? cmk(1.0, 0.0) : vx[s];
      const double ar = ok ? xv.x : 0.0, ai = ok ? xv.y : 0.0;
      const cplx y = Ys[k + XL * cl];
      u1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, y.x, u1, 0, 0, 0);
      u2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, y.y, u2, 0, 0, 0);
      u3 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar + ai, y.x + y.y, u3, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int rr = rb + rq + 4 * v;
      const double p1 = u1[v], p2 = u2[v];
      if (rr < m && col < m)
        Q[(size_t)rr + (size_t)col * ldq] = csub(q[i][v], cmk(p1 - p2, u3[v] - p1 - p2));
    }
  }
}

// W := Q W with Q = H_0 H_1 ... H_{n-3} the Hessenberg reduction's reflectors,
// 16 columns per workgroup held in LDS.  form = 1 writes Q itself (W = I on
// entry is implied): column j of Q only needs H_0 .. H_{j-1}, a third of the
// work of applying Q to a dense matrix.
__global__ __launch_bounds__(256) void hess_q_apply_kernel(int n, const cplx* __restrict__ Vg,
                                                           const cplx* __restrict__ taug,
                                                           cplx* __restrict__ Wg, int ldw,
                                                           long long sW, int form) {
  extern __shared__ cplx Wc[];  // n x 16
  const int page = blockIdx.y, c0 = blockIdx.x * 16, t = threadIdx.x;
  const cplx* V = Vg + (size_t)page * n * n;
  const cplx* tau = taug + (size_t)page * n;
  cplx* W = Wg + (size_t)page * sW;
  const int nc = min(16, n - c0);
  for (int e = t; e < n * nc; e += blockDim.x) {
    const int r = e % n, c = e / n;
    Wc[r + c * n] = form ? cmk(r == c0 + c ? 1.0 : 0.0, 0.0) : W[(size_t)r + (size_t)(c0 + c) * ldw];
  }
  __syncthreads();
  const int cl = t >> 4, sub = t & 15;
  if (cl < nc) {
    cplx* col = Wc + cl * n;
    const int kstart = form ? min(n - 3, c0 + cl - 1) : n - 3;
    for (int k = kstart; k >= 0; --k) {
      const cplx tk = tau[k];
      if (tk.x == 0.0 && tk.y == 0.0) continue;
      const int m = n - k - 1;
      const cplx* vk = V + (size_t)k * n + k + 1;
      double sr = 0.0, si = 0.0;
      for (int i = sub; i < m; i += 16) {
        const cplx p = cmulc(vk[i], col[k + 1 + i]);
        sr += p.x;
        si += p.y;
      }
      for (int o = 8; o > 0; o >>= 1) {
        sr += __shfl_xor(sr, o, 16);
        si += __shfl_xor(si, o, 16);
      }
      const cplx f = cmul(tk, cmk(sr, si));
      for (int i = sub; i < m; i += 16) col[k + 1 + i] = csub(col[k + 1 + i], cmul(vk[i], f));
    }
  }
  __syncthreads();
  for (int e = t; e < n * nc; e += blockDim.x) {
    const int r = e % n, c = e / n;
    W[(size_t)r + (size_t)(c0 + c) * ldw] = Wc[r + c * n];
  }
}

// Complex Givens rotation G = [c s; -conj(s) c] with G [a; b] = [r; 0], c real.
// Inputs are scaled by max(|a|_1, |b|_1) so the squared norms neither
// overflow nor underflow; two square roots and three reciprocals.
ORF_DEV void givens_scaled(cplx a, cplx b, double& c, cplx& s, cplx& r) {
  const double sc = fmax(cabs1(a), cabs1(b));
  if (cabs1(b) == 0.0 || sc == 0.0) { c = 1.0; s = cmk(0.0, 0.0); r = a; return; }
  const double isc = 1.0 / sc;
  const cplx as = cscale(a, isc), bs = cscale(b, isc);
  const double na2 = cnorm2(as), nb2 = cnorm2(bs);
  const double rho = sqrt(na2 + nb2);
  const double irho = 1.0 / rho;
  if (na2 == 0.0) {
    c = 0.0;
    s = cscale(cconj(bs), irho);
    r = cmk(rho * sc, 0.0);
    return;
  }
  const double aa = sqrt(na2);
  const cplx ph = cscale(as, 1.0 / aa);
  c = aa * irho;
  s = cscale(cmul(ph, cconj(bs)), irho);
  r = cscale(ph, rho * sc);
}

// 1 / sqrt(x) with one Newton step on the device reciprocal square root, so the
// rotations stay unitary to rounding (without it the QR stalls on some pages).
ORF_DEV double rsqrt_nr(double x) {
  const double y = rsqrt(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(0.5 * y, e, y);
}

// The same rotation with two reciprocal square roots and no divisions when the
// squared norms are safely inside the double range (the QR's rotation chains
// are latency bound); otherwise the scaled form.
ORF_DEV void givens(cplx a, cplx b, double& c, cplx& s, cplx& r) {
  const double na2 = cnorm2(a), nb2 = cnorm2(b), t = na2 + nb2;
  if (!(na2 > 1e-250 && t < 1e250) || nb2 == 0.0) {
    givens_scaled(a, b, c, s, r);
    return;
  }
  const double irho = rsqrt_nr(t), ia = rsqrt_nr(na2);
  const cplx ph = cscale(a, ia);  // a / |a|
  c = na2 * ia * irho;            // |a| / rho
  s = cscale(cmul(ph, cconj(b)), irho);
  r = cscale(ph, t * irho);
}

// Complex Schur form by single-shift QR on a Hessenberg matrix (in place);
// Schur vectors accumulated into Z; eigenvalues to wout.  info[page] = 0 on
// success, i+1 if eigenvalue i did not converge within the iteration limit.
__global__ __launch_bounds__(256) void schur_qr_kernel(int n, cplx* __restrict__ Hg, int ldh,
                                                       long long sH, cplx* __restrict__ Zg,
                                                       long long sZ, cplx* __restrict__ wout,
                                                       int* __restrict__ info, int use_lds,
                                                       int zInLds) {
  extern __shared__ cplx lds_mat[];
  __shared__ int redk[4];
  const int page = blockIdx.x, t = threadIdx.x, nt = blockDim.x;
  const int lane = t & 63, wv = t >> 6, nw = nt >> 6;
  Hg += (size_t)page * sH;
  cplx* Zp = Zg + (size_t)page * sZ;
  cplx* Z = Zp;
  wout += (size_t)page * n;
  cplx* H = Hg;
  int ld = ldh;
  if (use_lds) {
    H = lds_mat;
    ld = n;
    lds_fill(t, nt, n * n, [&](int e) { return Hg[(size_t)(e % n) + (size_t)(e / n) * ldh]; },
             [&](int e) -> cplx& { return H[e]; });
  }
  if (zInLds) {
    Z = lds_mat + (size_t)n * n;
    lds_fill(t, nt, n * n, [&](int e) { return Zp[e]; }, [&](int e) -> cplx& { return Z[e]; });
  }
  __syncthreads();
#define HH(r, c) H[(size_t)(r) + (size_t)(c) * ld]
  const double smlnum = SAFMIN * ((double)n / ULP);
  const int itmax = 100;  // sweeps per eigenvalue before giving up (reported via info)
  int failed = 0;
  int i = n - 1;
  while (i >= 0) {
    int l = 0;
    int its;
    for (its = 0; its <= itmax; ++its) {
      // ---- deflation search: largest k in (l, i] with negligible H[k,k-1]
      int kfound = l;
      for (int k = i - t; k > l; k -= nt) {
        const cplx hkk1 = HH(k, k - 1);
        bool small = false;
        if (cabs1(hkk1) <= smlnum) small = true;
        else {
          double tst = cabs1(HH(k - 1, k - 1)) + cabs1(HH(k, k));
          if (tst == 0.0) {
            if (k - 2 >= 0) tst += cabs1(HH(k - 1, k - 2));
            if (k + 1 <= n - 1) tst += cabs1(HH(k + 1, k));
          }
          if (cabs1(hkk1) <= ULP * tst) {
            const double ab = fmax(cabs1(hkk1), cabs1(HH(k - 1, k)));
            const double ba = fmin(cabs1(hkk1), cabs1(HH(k - 1, k)));
            const cplx dd = csub(HH(k - 1, k - 1), HH(k, k));
            const double aa = fmax(cabs1(HH(k, k)), cabs1(dd));
            const double bb = fmin(cabs1(HH(k, k)), cabs1(dd));
            const double s = aa + ab;
            if (ba * (ab / s) <= fmax(smlnum, ULP * (bb * (aa / s)))) small = true;
          }
        }
        if (small) { kfound = k; break; }
      }
      // block max of kfound
      int kv = kfound;
      for (int o = 32; o > 0; o >>= 1) kv = max(kv, __shfl_xor(kv, o, 64));
      __syncthreads();
      if (lane == 0) redk[wv] = kv;
      __syncthreads();
      int kmax = redk[0];
      for (int q = 1; q < nw; ++q) kmax = max(kmax, redk[q]);
      l = kmax;
      if (l > 0) {
        __syncthreads();
        if (t == 0) HH(l, l - 1) = cmk(0.0, 0.0);
        __syncthreads();
      }
      if (l >= i) break;
      // ---- shift
      cplx sig;
      if (its % 20 == 10) {
        sig = cadd(cmk(0.75 * cabs1(HH(l + 1, l)), 0.0), HH(l, l));
      } else if (its > 0 && its % 20 == 0) {
        sig = cadd(cmk(0.75 * cabs1(HH(i, i - 1)), 0.0), HH(i, i));
      } else {
        cplx tt = HH(i, i);
        const cplx u = cmul(csqrt(HH(i - 1, i)), csqrt(HH(i, i - 1)));
        double s = cabs1(u);
        if (s != 0.0) {
          const cplx x = cscale(csub(HH(i - 1, i - 1), tt), 0.5);
          const double sx = cabs1(x);
          s = fmax(s, cabs1(x));
          const cplx xs = cscale(x, 1.0 / s), us = cscale(u, 1.0 / s);
          cplx y = cscale(csqrt(cadd(cmul(xs, xs), cmul(us, us))), s);
          if (sx > 0.0) {
            const cplx xx = cscale(x, 1.0 / sx);
            if (xx.x * y.x + xx.y * y.y < 0.0) y = cneg(y);
          }
          tt = csub(tt, cmul(u, cdiv(u, cadd(x, y))));
        }
        sig = tt;
      }
      if (!(isfinite(sig.x) && isfinite(sig.y))) {  // non-finite input: stop at once
        its = itmax + 1;
        break;
      }
      // ---- single-shift QR sweep on rows/cols l..i
      for (int k = l; k < i; ++k) {
        cplx a, b;
        if (k == l) { a = csub(HH(l, l), sig); b = HH(l + 1, l); }
        else { a = HH(k, k - 1); b = HH(k + 1, k - 1); }
        double c;
        cplx s, r;
        givens(a, b, c, s, r);
        // the first rotation reads entries the row pass rewrites
        if (k == l) __syncthreads();
        const cplx sc = cconj(s);
        // rows k, k+1 for columns k..n-1 (column k-1 is written below, after the barrier)
        for (int j = k + t; j < n; j += nt) {
          const cplx h1 = HH(k, j), h2 = HH(k + 1, j);
          HH(k, j) = cadd(cscale(h1, c), cmul(s, h2));
          HH(k + 1, j) = csub(cscale(h2, c), cmul(sc, h1));
        }
        __syncthreads();
        if (k > l && t == 0) {
          HH(k, k - 1) = r;
          HH(k + 1, k - 1) = cmk(0.0, 0.0);
        }
        // columns k, k+1 for rows 0..min(k+2, i); Z rows 0..n-1
        const int rmax = min(k + 2, i);
        for (int q = t; q <= rmax; q += nt) {
          const cplx h1 = HH(q, k), h2 = HH(q, k + 1);
          HH(q, k) = cadd(cscale(h1, c), cmul(sc, h2));
          HH(q, k + 1) = csub(cscale(h2, c), cmul(s, h1));
        }
        for (int q = t; q < n; q += nt) {
          cplx* z1 = Z + q + (size_t)k * n;
          cplx* z2 = Z + q + (size_t)(k + 1) * n;
          const cplx h1 = *z1, h2 = *z2;
          *z1 = cadd(cscale(h1, c), cmul(sc, h2));
          *z2 = csub(cscale(h2, c), cmul(s, h1));
        }
        __syncthreads();
      }
    }
    if (its > itmax) {
      if (!failed) failed = i + 1;
      l = i;  // give up on this eigenvalue; keep going so every wave exits
    }
    if (t == 0) wout[i] = HH(i, i);
    // deflated block of size 1 at i
    i = l - 1;
    __syncthreads();
  }
  if (t == 0 && info) info[page] = failed;
  if (use_lds) {
    for (int e = t; e < n * n; e += nt) Hg[(size_t)(e % n) + (size_t)(e / n) * ldh] = H[e];
  }
  if (zInLds)
    for (int e = t; e < n * n; e += nt) Zp[e] = Z[e];
#undef HH
}


// Multishift configuration: bulges per pass, bulge positions per segment, the
// waves per SIMD the register budget is cut for (two per SIMD: 256 VGPRs, so two
// four-wave or four two-wave workgroups per CU hide the latency of the
// dependent rotation chains; batches that fill the chip take the two-wave
// launch, 139-141 vs 158-167 ms on the batch below), rotations between compiler
// barriers in the strip updates, and the early-deflation window.  Measured on
// the +-5 phase-library batch (n = 242, 2048 pages): 8 bulges (8 lanes per
// bulge in the chase) with a 16-wide window: 187 ms, against 208 ms for 4
// bulges with a 12-wide window (half the passes, early-deflation windows and
// chase steps per shift); windows of 12 / 20 / 24: 199 / 187 / 198 ms;
// 8 / 24 / 32 positions per segment: 221 / 208 / 212 ms; whole strips loaded
// at once (lookahead 16) instead of 8 steps ahead: 172 vs 182 ms.  The strips
// take all 8 bulges' rotations in one sweep with 8 rotations between compiler
// barriers: 170-172 ms, against 177 ms for two sweeps of 4 bulges (each strip
// element loaded and stored twice per segment).  Deferring the rows of Z and
// above the windows over two segments (one sweep of 32 steps) was 5 % slower.
constexpr int MBQ_NB = 8, MBQ_SEG = 16, MBQ_OCC = 2, MBQ_GRP = 8, MBQ_NW = 16;
// Small matrices: 4 bulges and a 12-wide window.  The 8-bulge kernel's static
// LDS (40 KB) next to an H held in LDS leaves one workgroup per CU where this
// one (20 KB) keeps two: 8,192 pages of n = 61 take 42 ms instead of 94.
// At small n the early-deflation window's serial rotation chain on one wave is
// the critical path, so pages per CU count more than waves per page: the same
// kernel compiled for 168 VGPRs and launched as two-wave workgroups (a chase
// wave and one strip wave; H and Z in global memory, where the pages in flight
// stay cache resident) runs six pages per CU: 20.5 ms for those 8,192 pages
// (four-wave workgroups at three per CU: 33 ms; two-wave at 256 VGPRs, four per
// CU: 24.6 ms; two-wave at 128 VGPRs: 41 ms, spills).  With six pages per CU
// a narrower early-deflation window shortens that chain more than the extra
// windows cost: 8 wide, 17.5 ms (6: 16.9, 10: 19.3).  Tuning choice 3; choice 4
// runs the 8-bulge kernel the same way at four per CU (its 16-wide window stays:
// 12 / 14 wide, 153 / 148 vs 135-146 ms).
constexpr int MBS_NB = 4, MBS_NW = 12, MBS_GRP = 16, MBS_OCC3 = 3, MBS_NW2 = 8;

// Eigenvalues of a small m x m (m <= 8) complex matrix held in LDS (ld 8), by
// single-shift QR without vectors; run by one thread.
ORF_DEV void small_eigvals(int m, cplx* A, cplx* ev) {
  const double smlnum = SAFMIN * ((double)m / ULP);
  int i = m - 1;
  int guard = 0;
  while (i >= 0 && guard < 60 * m) {
    int l = 0;
    for (int k = i; k > 0; --k) {
      const double tst = cabs1(A[(k - 1) + 8 * (k - 1)]) + cabs1(A[k + 8 * k]);
      if (cabs1(A[k + 8 * (k - 1)]) <= fmax(smlnum, ULP * tst)) { l = k; break; }
    }
    if (l >= i) { ev[i] = A[i + 8 * i]; --i; guard = 0; continue; }
    ++guard;
    cplx tt = A[i + 8 * i];
    if (guard % 10 == 0) {
      tt = cadd(tt, cmk(0.75 * cabs1(A[i + 8 * (i - 1)]), 0.0));
    } else {
      const cplx u = cmul(csqrt(A[(i - 1) + 8 * i]), csqrt(A[i + 8 * (i - 1)]));
      double sm = cabs1(u);
      if (sm != 0.0) {
        const cplx x = cscale(csub(A[(i - 1) + 8 * (i - 1)], tt), 0.5);
        const double sx = cabs1(x);
        sm = fmax(sm, sx);
        const cplx xs = cscale(x, 1.0 / sm), us = cscale(u, 1.0 / sm);
        cplx y = cscale(csqrt(cadd(cmul(xs, xs), cmul(us, us))), sm);
        if (sx > 0.0) {
          const cplx xx = cscale(x, 1.0 / sx);
          if (xx.x * y.x + xx.y * y.y < 0.0) y = cneg(y);
        }
        tt = csub(tt, cmul(u, cdiv(u, cadd(x, y))));
      }
    }
    for (int k = l; k < i; ++k) {
      cplx a, b;
      if (k == l) { a = csub(A[l + 8 * l], tt); b = A[(l + 1) + 8 * l]; }
      else { a = A[k + 8 * (k - 1)]; b = A[(k + 1) + 8 * (k - 1)]; }
      double c;
      cplx sn, r;
      givens(a, b, c, sn, r);
      const cplx sc = cconj(sn);
      if (k > l) { A[k + 8 * (k - 1)] = r; A[(k + 1) + 8 * (k - 1)] = cmk(0.0, 0.0); }
      for (int j = k; j <= i; ++j) {
        const cplx h1 = A[k + 8 * j], h2 = A[(k + 1) + 8 * j];
        A[k + 8 * j] = cadd(cscale(h1, c), cmul(sn, h2));
        A[(k + 1) + 8 * j] = csub(cscale(h2, c), cmul(sc, h1));
      }
      for (int q = l; q <= min(k + 2, i); ++q) {
        const cplx h1 = A[q + 8 * k], h2 = A[q + 8 * (k + 1)];
        A[q + 8 * k] = cadd(cscale(h1, c), cmul(sc, h2));
        A[q + 8 * (k + 1)] = csub(cscale(h2, c), cmul(sn, h1));
      }
    }
  }
  for (; i >= 0; --i) ev[i] = A[i + 8 * i];  // not converged: use the diagonal
}

// ------------------------------------------------- aggressive early deflation
// Helpers run by ONE wave on an m x m window (m <= 32) held in LDS with
// leading dimension LDW.  Every lane computes the scalar quantities (shifts,
// rotations, reflectors) from the same LDS words; element updates are spread
// over the lanes, one phase per wave_sync_lds.

// Wilkinson shift of the trailing 2x2 block (rows/cols i-1, i) of W.
template <int LDW>
ORF_DEV cplx win_wilkinson(const cplx* W, int i) {
  cplx tt = W[i + LDW * i];
  const cplx u = cmul(csqrt(W[(i - 1) + LDW * i]), csqrt(W[i + LDW * (i - 1)]));
  double sm = cabs1(u);
  if (sm != 0.0) {
    const cplx x = cscale(csub(W[(i - 1) + LDW * (i - 1)], tt), 0.5);
    const double sx = cabs1(x);
    sm = fmax(sm, sx);
    const cplx xs = cscale(x, 1.0 / sm), us = cscale(u, 1.0 / sm);
    cplx y = cscale(csqrt(cadd(cmul(xs, xs), cmul(us, us))), sm);
    if (sx > 0.0) {
      const cplx xx = cscale(x, 1.0 / sx);
      if (xx.x * y.x + xx.y * y.y < 0.0) y = cneg(y);
    }
    tt = csub(tt, cmul(u, cdiv(u, cadd(x, y))));
  }
  return tt;
}

// Complex Schur form of the m x m Hessenberg window T by single-shift QR,
// rotations accumulated into U (U <- U G^H).  0 on success, 1 when an
// eigenvalue needs more than 30 sweeps (T and U then hold a partial, still
// unitary, reduction).
template <int LDW>
ORF_DEV int win_schur(int m, cplx* T, cplx* U, int lane) {
  const double smlnum = SAFMIN * ((double)m / ULP);
  int hi = m - 1, its = 0;
  while (hi > 0) {
    bool small = false;
    const int k = lane + 1;
    if (k <= hi) {
      const cplx hk = T[k + LDW * (k - 1)];
      if (cabs1(hk) <= smlnum) small = true;
      else {
        double tst = cabs1(T[(k - 1) + LDW * (k - 1)]) + cabs1(T[k + LDW * k]);
        if (tst == 0.0) {
          if (k - 2 >= 0) tst += cabs1(T[(k - 1) + LDW * (k - 2)]);
          if (k + 1 <= m - 1) tst += cabs1(T[(k + 1) + LDW * k]);
        }
        if (cabs1(hk) <= ULP * tst) {
          const double ab = fmax(cabs1(hk), cabs1(T[(k - 1) + LDW * k]));
          const double ba = fmin(cabs1(hk), cabs1(T[(k - 1) + LDW * k]));
          const cplx dd = csub(T[(k - 1) + LDW * (k - 1)], T[k + LDW * k]);
          const double aa = fmax(cabs1(T[k + LDW * k]), cabs1(dd));
          const double bb = fmin(cabs1(T[k + LDW * k]), cabs1(dd));
          const double s = aa + ab;
          if (ba * (ab / s) <= fmax(smlnum, ULP * (bb * (aa / s)))) small = true;
        }
      }
    }
    const unsigned long long msk = __ballot(small);
    const int lo = msk ? 64 - __clzll(msk) : 0;  // largest k with a negligible T[k, k-1]
    if (lo > 0 && lane == 0) T[lo + LDW * (lo - 1)] = cmk(0.0, 0.0);
    wave_sync_lds();
    if (lo >= hi) {
      hi -= 1;
      its = 0;
      continue;
    }
    if (++its > 30) return 1;
    cplx sig;
    if (its == 10) sig = cadd(cmk(0.75 * cabs1(T[(lo + 1) + LDW * lo]), 0.0), T[lo + LDW * lo]);
    else if (its == 20) sig = cadd(cmk(0.75 * cabs1(T[hi + LDW * (hi - 1)]), 0.0), T[hi + LDW * hi]);
    else sig = win_wilkinson<LDW>(T, hi);
    if (!(isfinite(sig.x) && isfinite(sig.y))) return 1;
    double c;
    cplx sn, r;
    givens(csub(T[lo + LDW * lo], sig), T[(lo + 1) + LDW * lo], c, sn, r);
    for (int kk = lo; kk < hi; ++kk) {
      // rows kk, kk+1 (columns kk..m-1) on lanes 0..31, U columns kk, kk+1 on lanes
      // 32..63, as one instruction stream: x1 <- c x1 + g x2, x2 <- c x2 - conj(g) x1
      // with g = s on T's rows and conj(s) on U's columns
      {
        const bool onU = lane >= 32;
        const int jq = onU ? lane - 32 : kk + lane;
        cplx* const p1 = onU ? U + jq + LDW * kk : T + kk + LDW * jq;
        const int d = onU ? LDW : 1;
        const cplx gs = onU ? cconj(sn) : sn;
        if (jq < m) {
          const cplx h1 = p1[0], h2 = p1[d];
          p1[0] = cadd(cscale(h1, c), cmul(gs, h2));
          p1[d] = csub(cscale(h2, c), cmulc(gs, h1));
        }
        if (lane == 0 && kk > lo) {
          T[kk + LDW * (kk - 1)] = r;
          T[(kk + 1) + LDW * (kk - 1)] = cmk(0.0, 0.0);
        }
      }
      wave_sync_lds();
      // columns kk, kk+1 for rows 0..min(kk+2, hi)
      cplx n1 = cmk(0.0, 0.0);  // this lane's new T(q, kk)
      if (lane <= min(kk + 2, hi)) {
        const int q = lane;
        const cplx h1 = T[q + LDW * kk], h2 = T[q + LDW * (kk + 1)];
        n1 = cadd(cscale(h1, c), cmulc(sn, h2));
        T[q + LDW * kk] = n1;
        T[q + LDW * (kk + 1)] = csub(cscale(h2, c), cmul(sn, h1));
      }
      // the next rotation acts on T(kk+1, kk) and T(kk+2, kk): read out of the lanes
      // that just formed them, so its Givens overlaps this phase's LDS round trip
      if (kk + 1 < hi)
        givens(cmk(lane_value(n1.x, kk + 1), lane_value(n1.y, kk + 1)),
               cmk(lane_value(n1.x, kk + 2), lane_value(n1.y, kk + 2)), c, sn, r);
      wave_sync_lds();
    }
  }
  return 0;
}

// Swap the diagonal entries q, q+1 of the upper-triangular window T by one
// rotation (the complex step of LAPACK ztrexc), carrying U along.
template <int LDW>
ORF_DEV void win_swap(int m, cplx* T, cplx* U, int q, int lane) {
  const cplx t11 = T[q + LDW * q], t22 = T[(q + 1) + LDW * (q + 1)];
  double c;
  cplx sn, r;
  givens(T[q + LDW * (q + 1)], csub(t22, t11), c, sn, r);
  if (lane < 32) {
    const int j = q + 2 + lane;  // rows q, q+1 right of the 2x2 block
    if (j < m) {
      const cplx h1 = T[q + LDW * j], h2 = T[(q + 1) + LDW * j];
      T[q + LDW * j] = cadd(cscale(h1, c), cmul(sn, h2));
      T[(q + 1) + LDW * j] = csub(cscale(h2, c), cmulc(sn, h1));
    }
    if (lane < q) {  // columns q, q+1 above the 2x2 block
      const cplx h1 = T[lane + LDW * q], h2 = T[lane + LDW * (q + 1)];
      T[lane + LDW * q] = cadd(cscale(h1, c), cmulc(sn, h2));
      T[lane + LDW * (q + 1)] = csub(cscale(h2, c), cmul(sn, h1));
    }
    if (lane == 0) {
      T[q + LDW * q] = t22;
      T[(q + 1) + LDW * (q + 1)] = t11;
    }
  } else if (lane - 32 < m) {
    const int p = lane - 32;
    const cplx u1 = U[p + LDW * q], u2 = U[p + LDW * (q + 1)];
    U[p + LDW * q] = cadd(cscale(u1, c), cmulc(sn, u2));
    U[p + LDW * (q + 1)] = csub(cscale(u2, c), cmul(sn, u1));
  }
  wave_sync_lds();
}

// Householder reflector (LAPACK zlarfg convention): for x = [alpha; rest],
// (I - conj(tau) v v^H) x = beta e1 with v = [1; rest / (alpha - beta)].
// Returns tau; beta and the scale of the tail through the references.
ORF_DEV cplx win_reflector(cplx alpha, double xnorm2, double& beta, cplx& scl) {
  if (xnorm2 == 0.0 && alpha.y == 0.0) {
    beta = alpha.x;
    scl = cmk(0.0, 0.0);
    return cmk(0.0, 0.0);
  }
  const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xnorm2);
  beta = alpha.x >= 0.0 ? -nrm : nrm;
  scl = crecip(csub(alpha, cmk(beta, 0.0)));
  return cmk((beta - alpha.x) / beta, -alpha.y / beta);
}

// After deflation with ns > 1 undeflated eigenvalues and a nonzero spike
// s * conj(U[0, :]): reflect the spike onto e1 and bring T[0:ns, 0:ns] back to
// Hessenberg form (LAPACK zlaqr3's restore), updating T and U.  vv is LDS
// scratch of LDW entries.
template <int LDW>
ORF_DEV void win_restore(int m, int ns, cplx* T, cplx* U, cplx* vv, int lane) {
  // spike reflector on w = conj(U[0, 0:ns])
  {
    double xn2 = 0.0;
    for (int k = 1; k < ns; ++k) xn2 += cnorm2(U[LDW * k]);
    double beta;
    cplx scl;
    const cplx tau = win_reflector(cconj(U[0]), xn2, beta, scl);
    if (lane < ns) vv[lane] = lane == 0 ? cmk(1.0, 0.0) : cmul(cconj(U[LDW * lane]), scl);
    wave_sync_lds();
    if (tau.x != 0.0 || tau.y != 0.0) {
      // T[0:ns, :] <- (I - conj(tau) v v^H) T[0:ns, :]
      if (lane < m) {
        const int cc = lane;
        cplx y = cmk(0.0, 0.0);
        for (int k = 0; k < ns; ++k) y = cadd(y, cmulc(vv[k], T[k + LDW * cc]));
        const cplx f = cmulc(tau, y);
        for (int k = 0; k < ns; ++k) T[k + LDW * cc] = csub(T[k + LDW * cc], cmul(vv[k], f));
      }
      wave_sync_lds();
      // T[0:ns, 0:ns] <- T (I - tau v v^H); U[:, 0:ns] likewise
      if (lane < 32) {
        if (lane < ns) {
          const int rr = lane;
          cplx y = cmk(0.0, 0.0);
          for (int k = 0; k < ns; ++k) y = cadd(y, cmul(T[rr + LDW * k], vv[k]));
          const cplx f = cmul(tau, y);
          for (int k = 0; k < ns; ++k) T[rr + LDW * k] = csub(T[rr + LDW * k], cmul(f, cconj(vv[k])));
        }
      } else if (lane - 32 < m) {
        const int rr = lane - 32;
        cplx y = cmk(0.0, 0.0);
        for (int k = 0; k < ns; ++k) y = cadd(y, cmul(U[rr + LDW * k], vv[k]));
        const cplx f = cmul(tau, y);
        for (int k = 0; k < ns; ++k) U[rr + LDW * k] = csub(U[rr + LDW * k], cmul(f, cconj(vv[k])));
      }
      wave_sync_lds();
    }
  }
  // Hessenberg reduction of T[0:ns, 0:ns]; the reflectors act on indices >= 1,
  // so the spike (now on e1) is untouched
  for (int j = 0; j + 2 < ns; ++j) {
    double xn2 = 0.0;
    for (int k = j + 2; k < ns; ++k) xn2 += cnorm2(T[k + LDW * j]);
    double beta;
    cplx scl;
    const cplx alpha = T[(j + 1) + LDW * j];
    const cplx tau = win_reflector(alpha, xn2, beta, scl);
    if (lane > j && lane < ns) vv[lane] = lane == j + 1 ? cmk(1.0, 0.0) : cmul(T[lane + LDW * j], scl);
    wave_sync_lds();
    if (tau.x == 0.0 && tau.y == 0.0) continue;
    // left: rows j+1..ns-1 of columns j+1..m-1; column j becomes beta e1
    if (lane < 32) {
      const int cc = j + 1 + lane;
      if (cc < m) {
        cplx y = cmk(0.0, 0.0);
        for (int k = j + 1; k < ns; ++k) y = cadd(y, cmulc(vv[k], T[k + LDW * cc]));
        const cplx f = cmulc(tau, y);
        for (int k = j + 1; k < ns; ++k) T[k + LDW * cc] = csub(T[k + LDW * cc], cmul(vv[k], f));
      }
    } else if (lane == 32) {
      T[(j + 1) + LDW * j] = cmk(beta, 0.0);
      for (int k = j + 2; k < ns; ++k) T[k + LDW * j] = cmk(0.0, 0.0);
    }
    wave_sync_lds();
    // right: columns j+1..ns-1 of T rows 0..ns-1 and of U rows 0..m-1
    if (lane < 32) {
      if (lane < ns) {
        const int rr = lane;
        cplx y = cmk(0.0, 0.0);
        for (int k = j + 1; k < ns; ++k) y = cadd(y, cmul(T[rr + LDW * k], vv[k]));
        const cplx f = cmul(tau, y);
        for (int k = j + 1; k < ns; ++k) T[rr + LDW * k] = csub(T[rr + LDW * k], cmul(f, cconj(vv[k])));
      }
    } else if (lane - 32 < m) {
      const int rr = lane - 32;
      cplx y = cmk(0.0, 0.0);
      for (int k = j + 1; k < ns; ++k) y = cadd(y, cmul(U[rr + LDW * k], vv[k]));
      const cplx f = cmul(tau, y);
      for (int k = j + 1; k < ns; ++k) U[rr + LDW * k] = csub(U[rr + LDW * k], cmul(f, cconj(vv[k])));
    }
    wave_sync_lds();
  }
}

// Deferred rotations of one multishift segment on one strip (a row or column
// of MB_SEG + 3*MB_NB positions starting at matrix index a0, element stride
// `stride`).  Rotations go in step order (step j: every bulge's rotation j, on
// disjoint position pairs); positions 1+j .. 3*MB_NB-1+j are live at step j,
// so a ring of 3*MB_NB + LA registers streams the strip: position 1+j is
// stored after step j and position 3*MB_NB+LA+j fetched LA steps before it is
// needed.  Compiler barriers keep the loads and stores where they are and only
// a few rotations live.  CHECK = true when positions may fall outside the
// matrix: their loads read zero and their stores go to a sink, by selects, so
// the loop stays one basic block (branches around each access make the
// register allocator spill).  Row pairs rotate with s, column pairs with conj(s).
// Write sink for strip positions outside the matrix: a select between the real
// address and this one keeps the strip loop free of branches.
__device__ cplx g_strip_sink[1];

template <int MB_NB, int MB_SEG, int GRP, int LA, bool CHECK>
ORF_DEV void strip_rotations(cplx* base, ptrdiff_t stride, int a0, int n, bool rowop,
                             const double* rc, const cplx* rs) {
  constexpr int RING = 3 * MB_NB + LA;
  constexpr int PLAST = MB_SEG + 3 * MB_NB - 2;  // last position a rotation touches
  cplx* const sink = g_strip_sink;
  auto inside = [&](int p) { return !CHECK || (a0 + p >= 0 && a0 + p < n); };
  cplx v[RING];
#pragma unroll
  for (int p = 1; p < 3 * MB_NB + LA && p <= PLAST; ++p) {
    const cplx x = *(inside(p) ? base + p * stride : sink);
    v[p % RING] = inside(p) ? x : cmk(0.0, 0.0);
  }
  cplx* pst = base + stride;
  cplx* pld = base + (ptrdiff_t)(3 * MB_NB + LA) * stride;
#pragma unroll
  for (int j = 0; j < MB_SEG; ++j) {
    if ((j * MB_NB) % GRP == 0) asm volatile("" : "+v"(pst), "+v"(pld) : : "memory");
#pragma unroll
    for (int b = 0; b < MB_NB; ++b) {
      const int q = 3 * (MB_NB - 1 - b) + 1 + j;
      const double c = rc[b * MB_SEG + j];
      const cplx sr = rs[b * MB_SEG + j];
      const cplx sn = rowop ? sr : cconj(sr);
      const cplx h1 = v[q % RING], h2 = v[(q + 1) % RING];
      v[q % RING] = cadd(cscale(h1, c), cmul(sn, h2));
      v[(q + 1) % RING] = csub(cscale(h2, c), cmulc(sn, h1));
    }
    {
      const int p = 1 + j;
      *(inside(p) ? pst : sink) = v[p % RING];
      pst += stride;
    }
    {
      const int p = 3 * MB_NB + LA + j;
      if (p <= PLAST) {
        const cplx x = *(inside(p) ? pld : sink);
        v[p % RING] = inside(p) ? x : cmk(0.0, 0.0);
        pld += stride;
      }
    }
  }
#pragma unroll
  for (int p = MB_SEG + 1; p <= PLAST; ++p) {
    *(inside(p) ? pst : sink) = v[p % RING];
    pst += stride;
  }
}

// Multishift QR with aggressive early deflation (Braman, Byers & Mathias 2002,
// the scheme of LAPACK's zlaqr0/zlaqr3).
//  * Passes: MB_NB single-shift bulges chased in lockstep through an LDS
//    window by one wave (bulge b trails bulge b-1 by three positions, so the
//    rotations of one step act on disjoint pairs), MB_SEG positions per
//    segment; the rows above the window, the columns right of it and Z take
//    the segment's rotations afterwards, streamed through a register ring
//    (strip_rotations): one load and one store per element per segment serve
//    every shift of the pass.  Wave 0 gives the next window's first MB_SEG
//    columns right their rotations after its chase, the other waves do the
//    rest one segment behind, with one barrier per segment.
//  * Early deflation (aedMode 0 and 1): before each pass the trailing NW x NW
//    window is brought to Schur form in LDS (win_schur); eigenvalues whose
//    spike entry is negligible deflate at once, from the bottom (mode 1) or
//    with the undeflatable ones moved up (mode 0, LAPACK's reordering), and
//    the remaining ones are the next pass's shifts.  An active block of at most
//    NW + 4 rows is finished in the window outright (all modes but 3; mode 2
//    does only that, mode 3 no early deflation at all).  The window's unitary
//    factor reaches the rows above, the columns right and Z in one strip
//    update instead of one per pass: on the +-5 batch this cuts the QR's HBM
//    traffic by about 45 %.
template <int MB_NB, int MB_SEG, int OCC, int GRP, int NW>
__global__ __launch_bounds__(256, OCC) void schur_qr_aed_kernel(int n, cplx* __restrict__ Hg, int ldh,
                                                           long long sH, cplx* __restrict__ Zg,
                                                           long long sZ, cplx* __restrict__ wout,
                                                           int* __restrict__ info, int hInLds,
                                                           int zInLds, int aedMode) {
  constexpr int MB_SW = MB_SEG + 3 * (MB_NB - 1) + 3;
  constexpr int NWX = NW + 4;       // largest window (a whole small block)
  constexpr int LDW = NWX + 1;
  // deflations that make another window worth it before a pass: a quarter of the
  // window (LAPACK uses 14 %; measured on the +-5 batch, 3 of 12 is ~2 % faster)
  constexpr int NIBBLE = (NW * 25) / 100;
  static_assert(NWX <= 32, "window lanes");
  constexpr int WINSZ = MB_SW * MB_SW > LDW * NWX ? MB_SW * MB_SW : LDW * NWX;
  static_assert(64 % MB_NB == 0, "lanes per bulge in the chase");
  constexpr int LA = 16;  // strip fetch lookahead (steps): a whole 4-bulge strip in flight
  extern __shared__ cplx lds_mat[];
  __shared__ cplx win[WINSZ];       // chase window, or the AED window T
  __shared__ cplx Uw[LDW * NWX];
  __shared__ cplx vv[LDW];
  __shared__ double rc[2 * MB_NB * MB_SEG];  // rotations of two segments (pipelined)
  __shared__ cplx rs[2 * MB_NB * MB_SEG];
  __shared__ cplx sm8[64];
  __shared__ cplx shifts[MB_NB];
  __shared__ int redk[4];
  __shared__ int aedres[2];
  __shared__ int prid;  // segment sequence number whose priority strips are done
  const int page = blockIdx.x, t = threadIdx.x, nt = blockDim.x;
  const int lane = t & 63, wv = t >> 6, nw = nt >> 6;
  Hg += (size_t)page * sH;
  cplx* Zp = Zg + (size_t)page * sZ;
  wout += (size_t)page * n;
  cplx* H = Hg;
  int ld = ldh;
  cplx* Z = Zp;
  cplx* ldsNext = lds_mat;
  if (hInLds) {
    H = ldsNext;
    ld = n;
    ldsNext += (size_t)n * n;
    lds_fill(t, nt, n * n, [&](int e) { return Hg[(size_t)(e % n) + (size_t)(e / n) * ldh]; },
             [&](int e) -> cplx& { return H[e]; });
  }
  if (zInLds) {
    Z = ldsNext;
    lds_fill(t, nt, n * n, [&](int e) { return Zp[e]; }, [&](int e) -> cplx& { return Z[e]; });
  }
  if (t == 0) prid = 0;
  __syncthreads();
#define HH(r, c) H[(size_t)(r) + (size_t)(c) * ld]
#define WN(r, c) win[(r) + (c) * MB_SW]
#define TW(r, c) win[(r) + (c) * LDW]
#define UW(r, c) Uw[(r) + (c) * LDW]
  const double smlnum = SAFMIN * ((double)n / ULP);
  const int itmax = 60;
  int failed = 0;
  int i = n - 1;
  int its = 0;
  int segseq = 0;  // segments started (the same count in every wave)
  while (i >= 0) {
    // ---- deflation search: largest k in (0, i] with negligible H[k, k-1]
    int kfound = 0;
    for (int k = i - t; k > 0; k -= nt) {
      const cplx hkk1 = HH(k, k - 1);
      bool small = false;
      if (cabs1(hkk1) <= smlnum) small = true;
      else {
        double tst = cabs1(HH(k - 1, k - 1)) + cabs1(HH(k, k));
        if (tst == 0.0) {
          if (k - 2 >= 0) tst += cabs1(HH(k - 1, k - 2));
          if (k + 1 <= n - 1) tst += cabs1(HH(k + 1, k));
        }
        if (cabs1(hkk1) <= ULP * tst) {
          const double ab = fmax(cabs1(hkk1), cabs1(HH(k - 1, k)));
          const double ba = fmin(cabs1(hkk1), cabs1(HH(k - 1, k)));
          const cplx dd = csub(HH(k - 1, k - 1), HH(k, k));
          const double aa = fmax(cabs1(HH(k, k)), cabs1(dd));
          const double bb = fmin(cabs1(HH(k, k)), cabs1(dd));
          const double s = aa + ab;
          if (ba * (ab / s) <= fmax(smlnum, ULP * (bb * (aa / s)))) small = true;
        }
      }
      if (small) { kfound = k; break; }
    }
    int kv = kfound;
    for (int o = 32; o > 0; o >>= 1) kv = max(kv, __shfl_xor(kv, o, 64));
    __syncthreads();
    if (lane == 0) redk[wv] = kv;
    __syncthreads();
    int l = redk[0];
    for (int q = 1; q < nw; ++q) l = max(l, redk[q]);
    if (l > 0) {
      __syncthreads();
      if (t == 0) HH(l, l - 1) = cmk(0.0, 0.0);
      __syncthreads();
    }
    if (l >= i) {  // 1x1 block at the bottom
      if (t == 0) wout[i] = HH(i, i);
      i -= 1;
      its = 0;
      __syncthreads();
      continue;
    }
    if (its > itmax) {  // give up on eigenvalue i; keep going so every wave exits
      if (!failed) failed = i + 1;
      if (t == 0) wout[i] = HH(i, i);
      i -= 1;
      its = 0;
      __syncthreads();
      continue;
    }
    int na = i - l + 1;
    // ---- aggressive early deflation on the window [kw, i] (or the whole block)
    int nsh = 0;  // AED shifts available in shifts[]
    if (aedMode != 3 && (na <= NWX || aedMode != 2)) {
      const bool whole = na <= NWX;
      const int m = whole ? na : NW;
      const int kw = i - m + 1;
      const cplx s = whole ? cmk(0.0, 0.0) : HH(kw, kw - 1);
      for (int q = t; q < LDW * NWX; q += nt) {
        const int r = q % LDW, c = q / LDW;
        win[q] = (r < m && c < m && r <= c + 1) ? HH(kw + r, kw + c) : cmk(0.0, 0.0);
        Uw[q] = cmk(r == c ? 1.0 : 0.0, 0.0);
      }
      __syncthreads();
      if (wv == 0) {
        int ns = -1;
        if (win_schur<LDW>(m, win, Uw, lane) == 0) {
          ns = m;
          if (s.x == 0.0 && s.y == 0.0) {
            ns = 0;
          } else {
            // deflation with reordering: negligible spike entries deflate from the
            // bottom; an undeflatable eigenvalue is moved to the top (LAPACK zlaqr3)
            int ilst = 0;
            const double as = cabs1(s);
            while (ilst < ns) {
              const int k = ns - 1;
              double foo = cabs1(TW(k, k));
              if (foo == 0.0) foo = as;
              if (as * cabs1(UW(0, k)) <= fmax(smlnum, ULP * foo)) {
                ns -= 1;
              } else if (aedMode == 1) {
                break;
              } else {
                for (int q = k - 1; q >= ilst; --q) win_swap<LDW>(m, win, Uw, q, lane);
                ilst += 1;
              }
            }
          }
          if (ns >= MB_NB && lane < MB_NB) shifts[lane] = TW(ns - MB_NB + lane, ns - MB_NB + lane);
          if (ns < m && ns > 1) win_restore<LDW>(m, ns, win, Uw, vv, lane);
        }
        if (lane == 0) aedres[0] = ns;
      }
      __syncthreads();
      const int ns = aedres[0];
      if (ns >= 0 && ns < m) {
        // write the window back, then the strip update with Uw
        for (int q = t; q < m * m; q += nt) {
          const int r = q % m, c = q / m;
          HH(kw + r, kw + c) = r <= c + 1 ? TW(r, c) : cmk(0.0, 0.0);
        }
        if (t == 0 && !whole) HH(kw, kw - 1) = cmul(s, cconj(UW(0, 0)));
        for (int q = ns + t; q < m; q += nt) wout[kw + q] = TW(q, q);
        const int e0 = kw + m;  // first column right of the window
        const int ncR = n - e0, nrA = kw;
        const int tasks = ncR + nrA + n;
        for (int task = t; task < tasks; task += nt) {
          cplx v[NWX];
          cplx* base;
          size_t stride;
          bool left;
          if (task < ncR) {
            base = H + (size_t)(e0 + task) * ld + kw;  // column e0+task, rows kw..kw+m-1
            stride = 1;
            left = true;
          } else if (task < ncR + nrA) {
            base = H + (task - ncR) + (size_t)kw * ld;  // row, columns kw..kw+m-1
            stride = ld;
            left = false;
          } else {
            base = Z + (task - ncR - nrA) + (size_t)kw * n;
            stride = n;
            left = false;
          }
#pragma unroll
          for (int q = 0; q < NWX; ++q) v[q] = q < m ? base[(size_t)q * stride] : cmk(0.0, 0.0);
          if (left) {  // U^H x
            for (int j = 0; j < m; ++j) {
              cplx acc = cmk(0.0, 0.0);
#pragma unroll
              for (int q = 0; q < NWX; ++q) acc = cadd(acc, cmulc(UW(q, j), v[q]));
              base[(size_t)j * stride] = acc;
            }
          } else {  // x^T U
            for (int j = 0; j < m; ++j) {
              cplx acc = cmk(0.0, 0.0);
#pragma unroll
              for (int q = 0; q < NWX; ++q) acc = cadd(acc, cmul(v[q], UW(q, j)));
              base[(size_t)j * stride] = acc;
            }
          }
        }
        __syncthreads();
        its = 0;
        if (ns == 0) {  // the window deflated completely
          i = kw - 1;
          continue;
        }
        i = kw + ns - 1;
        na = i - l + 1;
        if (m - ns > NIBBLE || ns < MB_NB || na <= NWX) continue;
      }
      if (ns >= MB_NB) nsh = MB_NB;
    }
    // ---- one multishift pass on [l, i]
    its += 1;
    const bool exc = its % 10 == 0;
    const int nbul = (exc || na < 3 * MB_NB + 4) ? 1 : MB_NB;
    if (nbul == 1) {
      cplx sig;
      if (its % 20 == 10) {
        sig = cadd(cmk(0.75 * cabs1(HH(l + 1, l)), 0.0), HH(l, l));
      } else if (its % 20 == 0) {
        sig = cadd(cmk(0.75 * cabs1(HH(i, i - 1)), 0.0), HH(i, i));
      } else {
        cplx tt = HH(i, i);
        const cplx u = cmul(csqrt(HH(i - 1, i)), csqrt(HH(i, i - 1)));
        double s = cabs1(u);
        if (s != 0.0) {
          const cplx x = cscale(csub(HH(i - 1, i - 1), tt), 0.5);
          const double sx = cabs1(x);
          s = fmax(s, cabs1(x));
          const cplx xs = cscale(x, 1.0 / s), us = cscale(u, 1.0 / s);
          cplx y = cscale(csqrt(cadd(cmul(xs, xs), cmul(us, us))), s);
          if (sx > 0.0) {
            const cplx xx = cscale(x, 1.0 / sx);
            if (xx.x * y.x + xx.y * y.y < 0.0) y = cneg(y);
          }
          tt = csub(tt, cmul(u, cdiv(u, cadd(x, y))));
        }
        sig = tt;
      }
      __syncthreads();
      if (t == 0) shifts[0] = sig;
      __syncthreads();
    } else if (nsh < MB_NB) {
      __syncthreads();
      if (t < MB_NB * MB_NB) {
        const int r = t % MB_NB, c = t / MB_NB;
        sm8[r + 8 * c] = HH(i - MB_NB + 1 + r, i - MB_NB + 1 + c);
      }
      __syncthreads();
      if (t == 0) small_eigvals(MB_NB, sm8, shifts);
      __syncthreads();
    }
    bool bad = false;
    for (int b = 0; b < nbul; ++b) bad |= !(isfinite(shifts[b].x) && isfinite(shifts[b].y));
    if (bad) {  // non-finite input: give up on this block
      if (!failed) failed = i + 1;
      for (int q = l + t; q <= i; q += nt) wout[q] = HH(q, q);
      i = l - 1;
      its = 0;
      __syncthreads();
      continue;
    }
    // ---- one pass: segments of MB_SEG positions of bulge 0
    // Segments are pipelined, one barrier per segment: while wave 0 loads and
    // chases segment `it`, writes its window back and gives the next window's
    // first MB_SEG columns right segment it's rotations, the other waves apply
    // segment it-1's rotations to everything else outside the window (the
    // columns right beyond its first MB_SEG, the rows above its window, Z).  The
    // one overlap, the first MB_SEG columns right of window it, which take
    // segment it-1's rotations first, wave 1 does as its first tasks and then
    // raises a flag that wave 0 waits for.  Rotations are double buffered.
    const int nseg = (i - 1 + 3 * (nbul - 1) - l) / MB_SEG + 1;
    for (int it = 0; it <= nseg; ++it) {
      ++segseq;
      const int cs = it;                                     // segment chased in this iteration
      const int a0 = l + it * MB_SEG - 3 * (MB_NB - 1) - 1;  // window origin of segment it
      const bool chasing = it < nseg;
      if (chasing && wv == 0) {
        double* rcb = rc + (cs & 1) * MB_NB * MB_SEG;
        cplx* rsb = rs + (cs & 1) * MB_NB * MB_SEG;
        const int p0 = l + cs * MB_SEG;
        // the window's loads go out in batches of up to 13 per lane, all in flight
        // together (a loop that stores each element before loading the next waits
        // out one memory latency per trip)
        constexpr int WLOADS = (MB_SW * MB_SW + 63) / 64;
        constexpr int WB = WLOADS < 13 ? WLOADS : 13;
#pragma unroll
        for (int u0 = 0; u0 < WLOADS; u0 += WB) {
          cplx wl[WB];
#pragma unroll
          for (int u = 0; u < WB; ++u) {
            int q = lane + 64 * (u0 + u);
            asm volatile("" : "+v"(q));  // recomputed per segment, not held in registers
            const int r = a0 + q % MB_SW, c = a0 + q / MB_SW;
            const bool ok = q < MB_SW * MB_SW && r >= 0 && r < n && c >= 0 && c < n;
            const cplx x = *(ok ? &HH(r, c) : g_strip_sink);
            wl[u] = ok ? x : cmk(0.0, 0.0);
          }
#pragma unroll
          for (int u = 0; u < WB; ++u)
            if (lane + 64 * (u0 + u) < MB_SW * MB_SW) win[lane + 64 * (u0 + u)] = wl[u];
        }
        wave_sync_lds();
        // lockstep: at step j bulge b sits at position p0 - 3b + j; the rotations
        // of one step act on disjoint row and column pairs
        for (int j = 0; j < MB_SEG; ++j) {
          // LPB lanes per bulge: lane (b, sub) computes bulge b's rotation itself
          // from the same two LDS words as the bulge's other lanes (no lane waits
          // for a single computing lane, one rotation per lane whatever the bulge
          // count) and applies it to window columns (row ops) or rows (column ops)
          // sub, sub + LPB, ...; the column ops reuse the rotation from registers.
          // All loads of a phase are issued before its stores (the pairs of one
          // step are disjoint).
          constexpr int LPB = 64 / MB_NB;
          constexpr int EPB = (MB_SW + LPB - 1) / LPB;
          const int b = lane / LPB, sub = lane % LPB;
          const int k = p0 - 3 * b + j;
          const bool act = b < nbul && k >= l && k <= i - 1;
          const int kk = act ? k - a0 : 1;
          double cr = 1.0;
          cplx sr = cmk(0.0, 0.0);
          {
            const cplx xv = k == l ? WN(kk, kk) : WN(kk, kk - 1);
            const cplx yv = k == l ? WN(kk + 1, kk) : WN(kk + 1, kk - 1);
            cplx h1v[EPB], h2v[EPB];
            bool rowop[EPB];
#pragma unroll
            for (int u = 0; u < EPB; ++u) {
              const int jc = sub + LPB * u;
              rowop[u] = act && jc < MB_SW && jc >= kk && a0 + jc < n;
              const int jr = rowop[u] ? jc : kk;
              h1v[u] = WN(kk, jr);
              h2v[u] = WN(kk + 1, jr);
            }
            cplx r = cmk(0.0, 0.0);
            if (act) {
              const cplx x = k == l ? csub(xv, shifts[b]) : xv;
              givens(x, yv, cr, sr, r);
            }
#pragma unroll
            for (int u = 0; u < EPB; ++u) {
              const int jc = sub + LPB * u;
              if (rowop[u]) {
                WN(kk, jc) = cadd(cscale(h1v[u], cr), cmul(sr, h2v[u]));
                WN(kk + 1, jc) = csub(cscale(h2v[u], cr), cmul(cconj(sr), h1v[u]));
              }
            }
            if (sub == 0) {  // one lane per bulge: record and fix
              rcb[b * MB_SEG + j] = cr;
              rsb[b * MB_SEG + j] = sr;
              if (act && k > l) {
                WN(kk, kk - 1) = r;
                WN(kk + 1, kk - 1) = cmk(0.0, 0.0);
              }
            }
          }
          wave_sync_lds();
          // column ops of every bulge on window rows 0..min(kk+2, i-a0)
          {
            cplx h1v[EPB], h2v[EPB];
            bool colop[EPB];
#pragma unroll
            for (int u = 0; u < EPB; ++u) {
              const int q = sub + LPB * u;
              colop[u] = act && q < MB_SW && q <= min(kk + 2, i - a0) && a0 + q >= 0;
              const int qr = colop[u] ? q : kk;
              h1v[u] = WN(qr, kk);
              h2v[u] = WN(qr, kk + 1);
            }
#pragma unroll
            for (int u = 0; u < EPB; ++u) {
              const int q = sub + LPB * u;
              if (colop[u]) {
                WN(q, kk) = cadd(cscale(h1v[u], cr), cmul(cconj(sr), h2v[u]));
                WN(q, kk + 1) = csub(cscale(h2v[u], cr), cmul(sr, h1v[u]));
              }
            }
          }
          wave_sync_lds();
        }
        // window back to H: nothing else touches it in this iteration
        for (int q = lane; q < MB_SW * MB_SW; q += 64) {
          const int r = a0 + q % MB_SW, c = a0 + q / MB_SW;
          if (r >= 0 && r < n && c >= 0 && c < n) HH(r, c) = win[q];
        }
        // the next window's first columns right have segment it-1's rotations once
        // wave 1 has done its first tasks
        // (wave 1 raises it on every path; the bound only keeps a broken invariant
        // from hanging the device: the page is then reported as failed)
        if (it > 0) {
          int spins = 0;
          while (__hip_atomic_load(&prid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < segseq) {
            if (++spins > (1 << 22)) {
              if (!failed) failed = i + 1;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        }
      }
      {
        // strips: wave 0 (after its chase) the first MB_SEG columns right of window it
        // with segment it's rotations; the other threads (all of them after the last
        // chase) segment it-1's rotations on the rest
        const bool w0 = chasing && wv == 0;
        const int ss = w0 ? it : it - 1;
        if (ss >= 0 && ss < nseg) {
          const int sa0 = l + ss * MB_SEG - 3 * (MB_NB - 1) - 1;
          const int e0 = sa0 + MB_SW;
          const int ncR = e0 < n ? n - e0 : 0;
          const int nrA = sa0 > 0 ? sa0 : 0;
          const int nfirst = min(MB_SEG, ncR);
          const int tlo = w0 ? 0 : nfirst, thi = w0 ? nfirst : ncR + nrA + n;
          const int t0 = w0 ? lane : (chasing ? t - 64 : t), tstep = w0 ? 64 : (chasing ? nt - 64 : nt);
          const double* rcs = rc + (ss & 1) * MB_NB * MB_SEG;
          const cplx* rss = rs + (ss & 1) * MB_NB * MB_SEG;
          bool flag = chasing && wv == 1;  // wave 1 raises the flag after its first tasks
          for (int task = tlo + t0; task < thi; task += tstep) {
            cplx* base;
            ptrdiff_t stride;
            bool rowop;
            if (task < ncR) {
              base = H + (size_t)(e0 + task) * ld + sa0;  // column e0+task, rows sa0..
              stride = 1;
              rowop = true;
            } else if (task < ncR + nrA) {
              base = H + (task - ncR) + (size_t)sa0 * ld;  // row task-ncR, columns sa0..
              stride = ld;
              rowop = false;
            } else {
              base = Z + (task - ncR - nrA) + (ptrdiff_t)sa0 * n;
              stride = n;
              rowop = false;
            }
            // one instance only: with a checked and an unchecked copy behind a branch
            // the compiler hoists every rotation's LDS loads above it and spills
            strip_rotations<MB_NB, MB_SEG, GRP, LA, true>(base, stride, sa0, n, rowop, rcs, rss);
            if (flag) {
              __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
              if (lane == 0) __hip_atomic_store(&prid, segseq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
              flag = false;
            }
          }
          if (flag) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            if (lane == 0) __hip_atomic_store(&prid, segseq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          }
        } else if (chasing && wv == 1 && lane == 0) {
          __hip_atomic_store(&prid, segseq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
      __syncthreads();
    }
  }

  if (t == 0 && info) info[page] = failed;
  if (hInLds)
    for (int e = t; e < n * n; e += nt) Hg[(size_t)(e % n) + (size_t)(e / n) * ldh] = H[e];
  if (zInLds)
    for (int e = t; e < n * n; e += nt) Zp[e] = Z[e];
#undef UW
#undef TW
#undef WN
#undef HH
}

// Eigenvectors of upper-triangular T (ztrevc, side R, no back-transform):
// Y[:, k] = x with (T - T[k,k]) x = 0, x[k] = 1, x[j>k] = 0, scaled to max |.|_1 = 1.
// One wave per eigenvector; x distributed over lanes, 64 rows per slot.
constexpr int TREVC_SLOTS = 12;  // n <= 768
__global__ __launch_bounds__(256) void trevc_kernel(int n, const cplx* __restrict__ T, int ldt,
                                                    long long sT, cplx* __restrict__ Y, int ldy,
                                                    long long sY) {
  const int page = blockIdx.y, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int k = blockIdx.x * (blockDim.x >> 6) + wv;
  if (k >= n) return;
  T += (size_t)page * sT;
  Y += (size_t)page * sY;
  const double smlnum = SAFMIN * ((double)n / ULP);
  const cplx tkk = T[k + (size_t)k * ldt];
  const double smin = fmax(ULP * cabs1(tkk), smlnum);
  cplx x[TREVC_SLOTS];
#pragma unroll
  for (int s = 0; s < TREVC_SLOTS; ++s) {
    const int r = lane + 64 * s;
    x[s] = (r < k) ? cneg(T[r + (size_t)k * ldt]) : cmk(r == k ? 1.0 : 0.0, 0.0);
  }
  for (int j = k - 1; j >= 0; --j) {
    const int sj = j >> 6, lj = j & 63;
    cplx d = csub(T[j + (size_t)j * ldt], tkk);
    if (cabs1(d) < smin) d = cmk(smin, 0.0);
    cplx xj = cmk(0.0, 0.0);
#pragma unroll
    for (int s = 0; s < TREVC_SLOTS; ++s)
      if (s == sj) xj = x[s];
    xj.x = __shfl(xj.x, lj, 64);
    xj.y = __shfl(xj.y, lj, 64);
    xj = cdiv(xj, d);
    const cplx* colj = T + (size_t)j * ldt;
#pragma unroll
    for (int s = 0; s < TREVC_SLOTS; ++s) {
      const int r = lane + 64 * s;
      if (s == sj && lane == lj) x[s] = xj;
      else if (r < j) x[s] = csub(x[s], cmul(colj[r], xj));
    }
  }
  double mx = 0.0;
#pragma unroll
  for (int s = 0; s < TREVC_SLOTS; ++s)
    if (lane + 64 * s <= k) mx = fmax(mx, cabs1(x[s]));
  mx = wave_max(mx);
  const double sc = mx > 0.0 ? 1.0 / mx : 1.0;
#pragma unroll
  for (int s = 0; s < TREVC_SLOTS; ++s) {
    const int r = lane + 64 * s;
    if (r < n) Y[r + (size_t)k * ldy] = (r <= k) ? cscale(x[s], sc) : cmk(0.0, 0.0);
  }
}

// Blocked form of trevc_kernel: TV_EB eigenvectors of one page per workgroup.
// Thread (e, g) = (t / 16, t % 16) holds rows g, g + 16, ... of eigenvector
// k0 + e in registers, so the 16 rows of each 16-row block of an eigenvector
// sit in the 16 lanes of one group.  Back substitution runs block by block from
// the top one down: the 16 x 16 diagonal block of T (staged in LDS) is solved
// across the group's lanes with shuffles, then the rows above take the block's
// 16 columns at once (T read through the cache, shared by the 16 groups): two
// workgroup barriers per block instead of one per column.  Per eigenvector the
// divisions and the small-pivot rule are those of trevc_kernel; the updates of
// the rows above a block are summed over its columns.
constexpr int TV_EB = 16;
template <int RPT>
__global__ __launch_bounds__(256, RPT == 16 ? 3 : 1) void trevc_blocked_kernel(int n, const cplx* __restrict__ T, int ldt,
                                                            long long sT, cplx* __restrict__ Y, int ldy,
                                                            long long sY) {
  __shared__ cplx tdiag[256];  // diagonal block, column-major 16 x 16
  __shared__ cplx xsh[256];    // the block's solved values, 16 per eigenvector
  int page, kbk;
  xcd_page_tile2(page, kbk);
  const int k0 = kbk * TV_EB, t = threadIdx.x;
  const int e = t >> 4, g = t & 15;
  const int k = k0 + e;
  const bool act = k < n;
  T += (size_t)page * sT;
  Y += (size_t)page * sY;
  const double smlnum = SAFMIN * ((double)n / ULP);
  const cplx tkk = act ? T[k + (size_t)k * ldt] : cmk(0.0, 0.0);
  const double smin = fmax(ULP * cabs1(tkk), smlnum);
  cplx x[RPT];
#pragma unroll
  for (int sl = 0; sl < RPT; ++sl) {
    const int r = g + 16 * sl;
    x[sl] = (act && r < k) ? cneg(T[r + (size_t)k * ldt]) : cmk(act && r == k ? 1.0 : 0.0, 0.0);
  }
  const int kmax = min(k0 + TV_EB, n) - 1;
  const int jtop = kmax - 1;  // columns jtop .. 0 carry the back substitution
  for (int sb = jtop >= 0 ? (jtop >> 4) : -1; sb >= 0; --sb) {
    const int j0 = 16 * sb, jhi = min(j0 + 15, jtop);
    __syncthreads();  // the previous block's readers of tdiag are done
    {
      const int r = t & 15, c = t >> 4;
      tdiag[t] = (j0 + r < n && j0 + c <= jhi) ? T[(j0 + r) + (size_t)(j0 + c) * ldt] : cmk(0.0, 0.0);
    }
    __syncthreads();
    cplx xb = cmk(0.0, 0.0);  // row j0 + g of this lane's eigenvector
#pragma unroll
    for (int sl = 0; sl < RPT; ++sl)
      if (sl == sb) xb = x[sl];
    // the pivot of this lane's row, T(j0+g, j0+g) - T(k, k) (small-pivot rule applied),
    // as Smith's quotient x / d = (x.x + x.y r, x.y - x.x r) / (d.x + d.y r) (|d.x| >= |d.y|,
    // else mirrored) with r and the reciprocal of the denominator formed here, off the
    // serial chain below, which then multiplies instead of dividing
    bool swp;
    double sr, sid;
    {
      cplx d = csub(tdiag[g + 16 * g], tkk);
      if (cabs1(d) < smin) d = cmk(smin, 0.0);
      swp = fabs(d.x) < fabs(d.y);
      sr = swp ? d.x / d.y : d.y / d.x;
      sid = 1.0 / (swp ? d.y + d.x * sr : d.x + d.y * sr);
    }
    for (int j = jhi; j >= j0; --j) {
      const int gj = j - j0;
      cplx xj = cmk(0.0, 0.0);
      if (g == gj && act && k > j) {
        xb = swp ? cmk((xb.x * sr + xb.y) * sid, (xb.y * sr - xb.x) * sid)
                 : cmk((xb.x + xb.y * sr) * sid, (xb.y - xb.x * sr) * sid);
        xj = xb;
      }
      xj.x = __shfl(xj.x, gj, 16);
      xj.y = __shfl(xj.y, gj, 16);
      if (g < gj) xb = csub(xb, cmul(tdiag[g + 16 * gj], xj));
    }
#pragma unroll
    for (int sl = 0; sl < RPT; ++sl)
      if (sl == sb) x[sl] = xb;
    if (sb == 0) continue;
    // the block's solved values (zero where column j0 + c does not act on this
    // eigenvector); the 16 lanes of a group are in one wave
    xsh[t] = (act && k > j0 + g && j0 + g <= jhi) ? xb : cmk(0.0, 0.0);
    wave_sync_lds();
    const cplx* xs = xsh + 16 * e;
    const int nc = jhi - j0 + 1;
#pragma unroll
    for (int sl = 0; sl < RPT; ++sl) {
      if (sl < sb) {
        const cplx* tr = T + (g + 16 * sl) + (size_t)j0 * ldt;
        cplx acc = x[sl];
#pragma unroll
        for (int h = 0; h < 16; h += 8) {  // eight loads in flight
          cplx tv[8];
#pragma unroll
          for (int c = 0; c < 8; ++c) tv[c] = tr[(size_t)min(h + c, nc - 1) * ldt];
#pragma unroll
          for (int c = 0; c < 8; ++c)
            if (h + c < nc) acc = csub(acc, cmul(tv[c], xs[h + c]));
        }
        x[sl] = acc;
      }
    }
  }
  double mx = 0.0;
#pragma unroll
  for (int sl = 0; sl < RPT; ++sl)
    if (act && g + 16 * sl <= k) mx = fmax(mx, cabs1(x[sl]));
  for (int o = 8; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  const double sc = mx > 0.0 ? 1.0 / mx : 1.0;
  if (act) {
#pragma unroll
    for (int sl = 0; sl < RPT; ++sl) {
      const int r = g + 16 * sl;
      if (r < n) Y[r + (size_t)k * ldy] = (r <= k) ? cscale(x[sl], sc) : cmk(0.0, 0.0);
    }
  }
}

// Scale every column of W to unit 2-norm; nrm (optional) receives 1 / scale.
__global__ __launch_bounds__(256) void colnorm_kernel(int n, cplx* __restrict__ W, int ld,
                                                      long long sW, double* __restrict__ nrm) {
  const int page = blockIdx.y, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int k = blockIdx.x * (blockDim.x >> 6) + wv;
  if (k >= n) return;
  cplx* col = W + (size_t)page * sW + (size_t)k * ld;
  double s = 0.0;
  for (int r = lane; r < n; r += 64) s += cnorm2(col[r]);
  s = wave_sum(s);
  const double f = s > 0.0 ? 1.0 / sqrt(s) : 1.0;
  for (int r = lane; r < n; r += 64) col[r] = cscale(col[r], f);
  if (nrm && lane == 0) nrm[(size_t)page * n + k] = s > 0.0 ? sqrt(s) : 1.0;
}

}  // namespace

// A is overwritten by the Schur factor T; Zwork and Ywork hold n x n per page.
void orf_zgeev(hipStream_t st, int n, cplx* A, int lda, long long sA, cplx* w, cplx* V, int ldv,
               long long sV, cplx* Zwork, cplx* Ywork, cplx* Hwork, int* info, int batch) {
  orf_zgeev_timed(st, n, A, lda, sA, w, V, ldv, sV, Zwork, Ywork, Hwork, info, batch, OrfTimeHook(), nullptr);
}

void orf_zgeev_timed(hipStream_t st, int n, cplx* A, int lda, long long sA, cplx* w, cplx* V,
                     int ldv, long long sV, cplx* Zwork, cplx* Ywork, cplx* Hwork, int* info,
                     int batch, const OrfTimeHook& hook, double* vnorm) {
  if (n <= 0 || batch <= 0) return;
  auto tb = [&](const char* nm) { if (hook) hook(nm, 0); };
  auto te = [&](const char* nm) { if (hook) hook(nm, 1); };
  const long long sZ = (long long)n * n;
  const size_t bytes = (size_t)n * n * sizeof(cplx);
  // the whole LDS of a workgroup; the one-page-per-workgroup kernels below keep at
  // most a few dozen bytes of static LDS beside the page (64 reserved)
  const size_t ldsMax = (size_t)g_lds_limit + 20 * 1024;
  const int use_lds = bytes + (size_t)n * sizeof(cplx) + 64 <= ldsMax;  // matrix + reflector
  // Householder vectors and tau live in Vh (n*n + n per page), the first
  // n*n of which double as scratch for nothing else until Q is applied.
  cplx* Vh = Hwork;
  cplx* tau = Hwork + (size_t)n * n * batch;
  tb("eig_hessenberg");
  // in LDS: one rank-1 update pair per column; otherwise panels of reflectors with
  // Level-3 trailing updates (panel reflectors and the current column in LDS, Y in Zwork)
  const size_t pan12 = (size_t)n * 13 * sizeof(cplx), pan8 = (size_t)n * 9 * sizeof(cplx);
  const bool ok12 = n >= 32 && pan12 <= (size_t)g_lds_limit, ok8 = n >= 32 && pan8 <= (size_t)g_lds_limit;
  // built-in rule: rank-1 updates in LDS when the page fits (or n < 32), else 12- or 8-wide panels
  int blockedNB = (use_lds || n < 32) ? 0 : (ok12 ? 12 : (ok8 ? 8 : 0));
  {
    const int tv = orf_tune_value(ORF_TUNE_HESS, n, batch);
    if (tv == 0 || (tv == 12 && ok12) || (tv == 8 && ok8)) blockedNB = tv;
  }
  // batches that fill the chip at small n: two-wave workgroups, six pages per CU
  // instead of three (8,192 pages of n = 61, reduction and Q: 5.2 vs 6.6 ms)
  if (blockedNB == 12 && n <= 128 && batch >= 512)
    hessenberg_blocked_kernel<12, 4><<<batch, 128, pan12, st>>>(n, A, lda, sA, Vh, tau, Zwork, sZ);
  else if (blockedNB == 12 && n <= 257)
    hessenberg_blocked_kernel<12, 4><<<batch, 256, pan12, st>>>(n, A, lda, sA, Vh, tau, Zwork, sZ);
  else if (blockedNB == 12)
    hessenberg_blocked_kernel<12, 12><<<batch, 256, pan12, st>>>(n, A, lda, sA, Vh, tau, Zwork, sZ);
  else if (blockedNB == 8 && n <= 257)
    hessenberg_blocked_kernel<8, 4><<<batch, 256, pan8, st>>>(n, A, lda, sA, Vh, tau, Zwork, sZ);
  else if (blockedNB == 8)
    hessenberg_blocked_kernel<8, 12><<<batch, 256, pan8, st>>>(n, A, lda, sA, Vh, tau, Zwork, sZ);
  else
    hessenberg_noq_kernel<<<batch, 256, (use_lds ? bytes : 0) + (size_t)n * sizeof(cplx), st>>>(n, A, lda, sA, Vh,
                                                                                               tau, use_lds);
  // Z = Q of the reduction, so that the QR's accumulated Z is the full Schur basis
  if (blockedNB && sZ >= (long long)n * n && (long long)3 * n * blockedNB + blockedNB * blockedNB <= (long long)n * n) {
    // Q = P_0 P_1 ... with P_p = I - V_p T_p V_p^H, applied to I from the last panel on;
    // Q stays the identity outside its trailing (n-k0-1) block, so each panel is three
    // MFMA GEMMs on that block (clean V_p, T_p and the products in Ywork)
    const int NB = blockedNB, nref = n - 2;
    const long long sY = sZ;
    for (int b0 = 0; b0 < batch; b0 += 65535) {
      const int bb = batch - b0 < 65535 ? batch - b0 : 65535;
      eye_pages_kernel<<<dim3((unsigned)(((long long)n * n + 255) / 256), bb), 256, 0, st>>>(n, Zwork + (size_t)b0 * sZ, sZ);
    }
    // two panels per pass where the scratch allows: P_a P_b = I - [Va Vb] T2 [Va Vb]^H with
    // T2 = [Ta, -Ta Va^H Vb Tb; 0, Tb] (one pass over the trailing block instead of two)
    const bool pairs = NB == 12 && (long long)3 * n * 2 * NB + 4 * NB * NB + 2 * NB * NB <= (long long)n * n;
    int k0 = ((nref - 1) / NB) * NB;
    while (k0 >= 0) {
      const bool pr = pairs && k0 >= NB;
      const int ka = pr ? k0 - NB : k0;
      const int iba = pr ? NB : min(NB, nref - k0), ibb = pr ? min(NB, nref - k0) : 0;
      const int ib = iba + ibb, m = n - ka - 1;
      const int W = pairs ? 2 * NB : NB;  // widest panel this reduction uses
      cplx* Vp = Ywork;                    // m x ib
      cplx* Xp = Ywork + (size_t)n * W;    // ib x m
      cplx* Ypp = Ywork + (size_t)2 * n * W;
      cplx* Tp = Ywork + (size_t)3 * n * W;  // ib x ib
      cplx* Sp = Tp + W * W;                 // NB x NB
      cplx* Wp = Sp + NB * NB;
      cplx* Q22 = Zwork + (size_t)(ka + 1) * n + (ka + 1);
      if (m <= 256 && ib <= 32 && iba <= 16 && ibb <= 16) {
        // T2 by one small kernel per page, then the pass with V read in place and the
        // block of Q22 held in registers (one read and one write of Q22)
        hess_pair_t_kernel<<<batch, 256, 0, st>>>(n, ka, iba, ibb, Vh, Tp, sY);
        hess_q_pass_kernel<4, 3><<<(unsigned)((long long)((m + 15) / 16) * batch), 256, 0, st>>>(
            m, ib, Vh + (size_t)ka * n + ka + 1, n, (long long)n * n, Tp, ib, sY, Q22, n, sZ, batch);
        k0 = ka - NB;
        continue;
      }
      for (int b0 = 0; b0 < batch; b0 += 65535) {
        const int bb = batch - b0 < 65535 ? batch - b0 : 65535;
        panel_extract_kernel<<<dim3((unsigned)(((long long)m * ib + ib * ib + 255) / 256), bb), 256, 0, st>>>(
            n, ka, ib, Vh + (size_t)b0 * n * n, Vp + (size_t)b0 * sY, sY, pr ? nullptr : Tp + (size_t)b0 * sY, sY);
        if (pr)
          panel_pair_t_kernel<<<dim3((unsigned)((ib * ib + 255) / 256), bb), 256, 0, st>>>(
              n, ka, iba, ibb, Vh + (size_t)b0 * n * n, Tp + (size_t)b0 * sY, sY);
      }
      if (pr) {
        // S = Va^H Vb, W = S Tb, T2(0:iba, iba:) = -Ta W
        orf_zgemm(st, 'C', 'N', iba, ibb, m, cmk(1.0, 0.0), Vp, m, sY, Vp + (size_t)m * iba, m, sY, cmk(0.0, 0.0), Sp,
                  iba, sY, batch);
        orf_zgemm(st, 'N', 'N', iba, ibb, ibb, cmk(1.0, 0.0), Sp, iba, sY, Tp + (size_t)iba * ib + iba, ib, sY,
                  cmk(0.0, 0.0), Wp, iba, sY, batch);
        orf_zgemm(st, 'N', 'N', iba, ibb, iba, cmk(-1.0, 0.0), Tp, ib, sY, Wp, iba, sY, cmk(0.0, 0.0),
                  Tp + (size_t)iba * ib, ib, sY, batch);
      }
      orf_zgemm(st, 'C', 'N', ib, m, m, cmk(1.0, 0.0), Vp, m, sY, Q22, n, sZ, cmk(0.0, 0.0), Xp, ib, sY, batch);
      orf_zgemm(st, 'N', 'N', ib, m, ib, cmk(1.0, 0.0), Tp, ib, sY, Xp, ib, sY, cmk(0.0, 0.0), Ypp, ib, sY, batch);
      orf_zgemm(st, 'N', 'N', m, m, ib, cmk(-1.0, 0.0), Vp, m, sY, Ypp, ib, sY, cmk(1.0, 0.0), Q22, n, sZ, batch);
      k0 = ka - NB;
    }
  } else {
    for (int b0 = 0; b0 < batch; b0 += 65535) {
      const int bb = batch - b0 < 65535 ? batch - b0 : 65535;
      hess_q_apply_kernel<<<dim3((n + 15) / 16, bb), 256, (size_t)n * 16 * sizeof(cplx), st>>>(
          n, Vh + (size_t)b0 * n * n, tau + (size_t)b0 * n, Zwork + (size_t)b0 * sZ, n, sZ, 1);
    }
  }
  te("eig_hessenberg");
  tb("eig_schur_qr");
  {
    // Small matrices in small batches: one rotation at a time with H (and Z when
    // it fits) in LDS, the shortest chain per page.  Larger ones, and small ones in
    // batches that fill the chip (two pages per CU and more), where total work
    // decides: multi-bulge passes over a window, off-window work deferred
    // (measured at n = 61: 47 vs 69 ms for 8192 pages; 5.2 vs 4.2 ms for 4).
    const size_t lim = (size_t)g_lds_limit;
    const bool few = batch < 512;
    const bool hFits = bytes + 64 <= ldsMax;  // the per-rotation kernel keeps 16 bytes of static LDS
    bool perRotation = hFits && (n < 32 || (n <= g_qr_small_n && few));
    // multishift passes of 8 bulges with a 16-wide window from n = 128 (H no
    // longer fits LDS there), else 4 bulges with a 12-wide window
    bool wide = n >= 128;
    // two-wave workgroups (H and Z in global memory): batches that fill the chip
    bool twoWave = !wide && !few;
    {
      // 0 per-rotation (H in LDS), 1 multishift 4 bulges, 2 multishift 8 bulges,
      // 3 / 4 the same 4 / 8 bulges in two-wave workgroups, six / four per CU
      const int tv = orf_tune_value(ORF_TUNE_QR, n, batch);
      if (tv >= 1 && tv <= 4) {
        perRotation = false;
        wide = tv == 2 || tv == 4;
        twoWave = tv >= 3;
      } else if (tv == 0 && hFits) {
        perRotation = true;
      }
    }
    if (perRotation) {
      const int zIn = 2 * bytes + 64 <= ldsMax;
      schur_qr_kernel<<<batch, 256, (1 + zIn) * bytes, st>>>(n, A, lda, sA, Zwork, sZ, w, info, 1, zIn);
    } else {
      const size_t stat = 64 * 1024;
      const size_t lim2 = lim + 20 * 1024 - stat;
      const int hIn = bytes <= lim2, zIn = 2 * bytes <= lim2;
      const int mode = g_qr_aed == 1 ? 0 : (g_qr_aed == 2 ? 1 : (g_qr_aed == 3 ? 2 : 3));
      if (wide && twoWave)  // 40 KB of static LDS: four workgroups per CU
        schur_qr_aed_kernel<MBQ_NB, MBQ_SEG, MBQ_OCC, MBQ_GRP, MBQ_NW><<<batch, 128, 0, st>>>(
            n, A, lda, sA, Zwork, sZ, w, info, 0, 0, mode);
      else if (wide)
        schur_qr_aed_kernel<MBQ_NB, MBQ_SEG, MBQ_OCC, MBQ_GRP, MBQ_NW><<<batch, 256, (hIn + zIn) * bytes, st>>>(
            n, A, lda, sA, Zwork, sZ, w, info, hIn, zIn, mode);
      else if (twoWave)  // 168 VGPRs and 21 KB of LDS: six workgroups per CU
        schur_qr_aed_kernel<MBS_NB, MBQ_SEG, MBS_OCC3, MBS_GRP, MBS_NW2><<<batch, 128, 0, st>>>(
            n, A, lda, sA, Zwork, sZ, w, info, 0, 0, mode);
      else
        schur_qr_aed_kernel<MBS_NB, MBQ_SEG, MBQ_OCC, MBS_GRP, MBS_NW><<<batch, 256, (hIn + zIn) * bytes, st>>>(
            n, A, lda, sA, Zwork, sZ, w, info, hIn, zIn, mode);
    }
  }
  te("eig_schur_qr");
  dim3 g2((n + 3) / 4, batch);
  tb("eig_vectors");
  {
    const dim3 gb((n + TV_EB - 1) / TV_EB, batch);
    // 16 or 32 rows per lane (n <= 256 / 512), or one wave per vector (n <= 768)
    int rpt = n <= 256 ? 16 : (n <= 512 ? 32 : 0);
    {
      const int tv = orf_tune_value(ORF_TUNE_TREVC, n, batch);
      if ((tv == 16 && n <= 256) || (tv == 32 && n <= 512) || tv == 0) rpt = tv;
    }
    if (rpt == 16)
      trevc_blocked_kernel<16><<<gb, 256, 0, st>>>(n, A, lda, sA, Ywork, n, sZ);
    else if (rpt == 32)
      trevc_blocked_kernel<32><<<gb, 256, 0, st>>>(n, A, lda, sA, Ywork, n, sZ);
    else
      trevc_kernel<<<g2, 256, 0, st>>>(n, A, lda, sA, Ywork, n, sZ);
  }
  // V = Z Y with Y upper triangular (the eigenvectors of T)
  orf_zgemm(st, 'N', 'U', n, n, n, cmk(1.0, 0.0), Zwork, n, sZ, Ywork, n, sZ, cmk(0.0, 0.0), V,
            ldv, sV, batch);
  colnorm_kernel<<<g2, 256, 0, st>>>(n, V, ldv, sV, vnorm);
  te("eig_vectors");
}
