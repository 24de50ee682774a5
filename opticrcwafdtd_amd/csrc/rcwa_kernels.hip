// RCWA-specific kernels for gfx950: permittivity painting, exact staircase
// Fourier transforms, Laurent / Li convolution matrices, Omega^2 = P Q
// assembly, mode matrices, S-matrix helpers and the efficiency reduction.
//
// Harmonic index p = (m + Nx) + (2Nx + 1)(n + Ny), m fastest (harmonics2D.m:4-6).
// Pages: a "mode page" is one (structure b, wavelength l) pair, p = b + B l;
// a "solve page" is a mode page times a parity sector (one sector when the
// full basis is solved).
#include "orf_common.h"
#include "orf_rcwa_dev.h"

// ------------------------------------------------------------------ painting
// Inside tests of shapeCircle.m:12, shapeRect.m:15, shapePolygon.m:17 and the
// periodic wrap of layerEps.m:40-45, with every product and sum rounded
// separately (no contraction) so samples match a plain IEEE evaluation.
__device__ double wrap_rn(double d, double L) {
  if (!isfinite(L)) return d;
  double t = __dadd_rn(d, L / 2);
  t = __dsub_rn(t, __dmul_rn(floor(t / L), L));
  return __dsub_rn(t, L / 2);
}

__device__ bool inpolygon_rn(double px, double py, const double* vx, const double* vy, int k) {
  bool inside = false, on = false;
  for (int e = 0; e < k; ++e) {
    const double x1 = vx[e], y1 = vy[e], x2 = vx[(e + 1) % k], y2 = vy[(e + 1) % k];
    const double cr = __dsub_rn(__dmul_rn(__dsub_rn(x2, x1), __dsub_rn(py, y1)),
                                __dmul_rn(__dsub_rn(y2, y1), __dsub_rn(px, x1)));
    const bool within = fmin(x1, x2) <= px && px <= fmax(x1, x2) && fmin(y1, y2) <= py &&
                        py <= fmax(y1, y2);
    const double scale = fmax(1.0, __dadd_rn(fabs(__dsub_rn(x2, x1)), fabs(__dsub_rn(y2, y1))));
    if (fabs(cr) <= 1e-12 * scale && within) on = true;
    if ((y1 > py) != (y2 > py)) {
      const double xint =
          __dadd_rn(x1, __ddiv_rn(__dmul_rn(__dsub_rn(py, y1), __dsub_rn(x2, x1)), __dsub_rn(y2, y1)));
      if (px < xint) inside = !inside;
    }
  }
  return inside || on;
}

__global__ void raster_kernel(int Nsx, int Nsy, const DevGrid* __restrict__ grids,
                              const DevShape* __restrict__ shapes, const double* __restrict__ verts,
                              uint8_t* __restrict__ labels) {
  const int g = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= Nsx * Nsy) return;
  const int i = e % Nsx, j = e / Nsx;
  const DevGrid G = grids[g];
  const double Lx = G.Lx, Ly = G.Ly;
  const double x = isfinite(Lx) ? __dadd_rn(-Lx / 2, __dmul_rn(Lx / Nsx, i + 0.5)) : 0.0;
  const double y = isfinite(Ly) ? __dadd_rn(-Ly / 2, __dmul_rn(Ly / Nsy, j + 0.5)) : 0.0;
  int lab = 0;
  for (int s = 0; s < G.nshape; ++s) {
    const DevShape S = shapes[G.shape0 + s];
    const double dx = wrap_rn(__dsub_rn(x, S.cx), Lx);
    const double dy = wrap_rn(__dsub_rn(y, S.cy), Ly);
    bool in;
    if (S.type == 0) in = __dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)) < __dmul_rn(S.a, S.a);
    else if (S.type == 1) in = fabs(dx) < S.a / 2 && fabs(dy) < S.b / 2;
    else in = inpolygon_rn(dx, dy, verts + S.vert0, verts + S.vert0 + S.nvert, S.nvert);
    if (in) lab = S.slot;
  }
  labels[(size_t)g * Nsx * Nsy + e] = (uint8_t)lab;
}

// x -> -x and y -> -y mirror symmetry of every label grid; flags[g] = 1 if symmetric.
__global__ void grid_symmetry_kernel(int Nsx, int Nsy, const uint8_t* __restrict__ labels,
                                     int* __restrict__ flags) {
  const int g = blockIdx.y;
  const uint8_t* L = labels + (size_t)g * Nsx * Nsy;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < Nsx * Nsy; e += gridDim.x * blockDim.x) {
    const int i = e % Nsx, j = e / Nsx;
    const uint8_t v = L[e];
    if (v != L[(Nsx - 1 - i) + (size_t)j * Nsx] || v != L[i + (size_t)(Nsy - 1 - j) * Nsx])
      flags[g] = 0;
  }
}

// ------------------------------------------------------------------ Fourier transforms
// chi_x[g][s][h][j] = sum_i Fx(h, i) [label(i, j) == s], h in [-2Nx, 2Nx]
__global__ void chi_x_kernel(int Nsx, int Nsy, int Hx, int nslot, const uint8_t* __restrict__ labels,
                             const cplx* __restrict__ Fx, cplx* __restrict__ chix) {
  const int g = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;  // j fastest
  if (e >= Nsy * Hx) return;
  const int j = e % Nsy, h = e / Nsy;
  const uint8_t* L = labels + (size_t)g * Nsx * Nsy + (size_t)j * Nsx;
  cplx acc[ORF_MAX_SLOTS];
#pragma unroll
  for (int s = 0; s < ORF_MAX_SLOTS; ++s) acc[s] = cmk(0.0, 0.0);
  for (int i = 0; i < Nsx; ++i) {
    const cplx f = Fx[h + (size_t)Hx * i];
    const int lab = L[i];
#pragma unroll
    for (int s = 0; s < ORF_MAX_SLOTS; ++s)
      if (s == lab) acc[s] = cadd(acc[s], f);
  }
  cplx* out = chix + (size_t)g * nslot * Hx * Nsy;
  for (int s = 0; s < nslot; ++s) out[(size_t)s * Hx * Nsy + h + (size_t)Hx * j] = acc[s];
}

// chi_y[g][s][k][i] = sum_j Fy(k, j) [label(i, j) == s], k in [-2Ny, 2Ny]
__global__ void chi_y_kernel(int Nsx, int Nsy, int Hy, int nslot, const uint8_t* __restrict__ labels,
                             const cplx* __restrict__ Fy, cplx* __restrict__ chiy) {
  const int g = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;  // i fastest
  if (e >= Nsx * Hy) return;
  const int i = e % Nsx, k = e / Nsx;
  const uint8_t* L = labels + (size_t)g * Nsx * Nsy + i;
  cplx acc[ORF_MAX_SLOTS];
#pragma unroll
  for (int s = 0; s < ORF_MAX_SLOTS; ++s) acc[s] = cmk(0.0, 0.0);
  for (int j = 0; j < Nsy; ++j) {
    const cplx f = Fy[k + (size_t)Hy * j];
    const int lab = L[(size_t)j * Nsx];
#pragma unroll
    for (int s = 0; s < ORF_MAX_SLOTS; ++s)
      if (s == lab) acc[s] = cadd(acc[s], f);
  }
  cplx* out = chiy + (size_t)g * nslot * Hy * Nsx;
  for (int s = 0; s < nslot; ++s) out[(size_t)s * Hy * Nsx + k + (size_t)Hy * i] = acc[s];
}

// chi_2d[g][s][h][k] = sum_j Fy(k, j) chi_x[g][s][h][j]
__global__ void chi_2d_kernel(int Nsy, int Hx, int Hy, int nslot, const cplx* __restrict__ chix,
                              const cplx* __restrict__ Fy, cplx* __restrict__ chi2) {
  const int g = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= Hx * Hy * nslot) return;
  const int h = e % Hx, k = (e / Hx) % Hy, s = e / (Hx * Hy);
  const cplx* cx = chix + (size_t)g * nslot * Hx * Nsy + (size_t)s * Hx * Nsy;
  cplx acc = cmk(0.0, 0.0);
  for (int j = 0; j < Nsy; ++j) acc = cadd(acc, cmul(Fy[k + (size_t)Hy * j], cx[h + (size_t)Hx * j]));
  chi2[(size_t)g * nslot * Hx * Hy + (size_t)s * Hx * Hy + h + (size_t)Hx * k] = acc;
}

// ------------------------------------------------------------------ convolution matrices
// Laurent [eps](p, q) = f(m_p - m_q, n_p - n_q), f = sum_s eps_s chi_2d[s]  (convMatrix2D.m:2-4)
__global__ void laurent_kernel(int Nx, int Ny, int M, const DevPage* __restrict__ pages,
                               const cplx* __restrict__ chi2, const cplx* __restrict__ slotEps,
                               int nslot, cplx* __restrict__ E, long long sE) {
  const int pg = blockIdx.y;
  const DevPage P = pages[pg];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= M * M) return;
  const int p = e % M, q = e / M;
  const int Hx = 4 * Nx + 1, Hy = 4 * Ny + 1, W = 2 * Nx + 1;
  const int mp = p % W, np_ = p / W, mq = q % W, nq = q / W;
  const int h = mp - mq + 2 * Nx, k = np_ - nq + 2 * Ny;
  const cplx* c2 = chi2 + (size_t)P.grid * nslot * Hx * Hy;
  const cplx* se = slotEps + (size_t)pg * nslot;
  cplx acc = cmk(0.0, 0.0);
  for (int s = 0; s < P.nslot; ++s) acc = cadd(acc, cmul(se[s], c2[(size_t)s * Hx * Hy + h + (size_t)Hx * k]));
  E[(size_t)pg * sE + e] = acc;
}

// Toeplitz matrices of 1/eps along one axis at every sample of the other
// (convMatricesLi2D.m:2-5): dir 0 -> T_j (W x W, W = 2Nx+1) for every y-sample j;
// dir 1 -> T_i (2Ny+1 square) for every x-sample i.  Out: [page][sample][W*W].
__global__ void li_toeplitz_kernel(int dir, int Nh, int Nsamp, int Hlen, const DevPage* __restrict__ pages,
                                   const cplx* __restrict__ chi, long long chiGridStride,
                                   const cplx* __restrict__ slotInvEps, int nslot,
                                   cplx* __restrict__ T) {
  const int pg = blockIdx.y;
  const DevPage P = pages[pg];
  const int W = 2 * Nh + 1;
  const long long per = (long long)W * W;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= per * Nsamp) return;
  const int j = (int)(e / per);
  const int r = (int)(e % per);
  const int a = r % W, b = r / W;
  const int h = a - b + 2 * Nh;
  const cplx* c = chi + (size_t)P.grid * chiGridStride;
  const cplx* se = slotInvEps + (size_t)pg * nslot;
  cplx acc = cmk(0.0, 0.0);
  for (int s = 0; s < P.nslot; ++s)
    acc = cadd(acc, cmul(se[s], c[(size_t)s * Hlen * Nsamp + h + (size_t)Hlen * j]));
  T[(size_t)pg * per * Nsamp + e] = acc;
  (void)dir;
}

// Exx(p,q) = G(n_p - n_q, m_p, m_q); Eyy(p,q) = H(m_p - m_q, n_p, n_q), where
// G[page] is (4Ny+1) x (2Nx+1)^2 and H[page] is (4Nx+1) x (2Ny+1)^2 (column-major).
__global__ void li_gather_kernel(int Nx, int Ny, int M, const cplx* __restrict__ G, const cplx* __restrict__ Hm,
                                 cplx* __restrict__ Exx, cplx* __restrict__ Eyy, long long sE) {
  const int pg = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= M * M) return;
  const int p = e % M, q = e / M;
  const int Wx = 2 * Nx + 1, Wy = 2 * Ny + 1, Hx = 4 * Nx + 1, Hy = 4 * Ny + 1;
  const int mp = p % Wx, np_ = p / Wx, mq = q % Wx, nq = q / Wx;
  const cplx* g = G + (size_t)pg * Hy * Wx * Wx;
  const cplx* hh = Hm + (size_t)pg * Hx * Wy * Wy;
  Exx[(size_t)pg * sE + e] = g[(np_ - nq + 2 * Ny) + (size_t)Hy * (mp + Wx * mq)];
  Eyy[(size_t)pg * sE + e] = hh[(mp - mq + 2 * Nx) + (size_t)Hx * (np_ + Wy * nq)];
}

// ------------------------------------------------------------------ Omega^2 = P Q
// RHS for G = [eps]^-1 [Kx Exx, Ky Eyy]  (M x 2M)
__global__ void omega_rhs_kernel(int M, const DevPage* __restrict__ pages, const double* __restrict__ kxy,
                                 const cplx* __restrict__ Exx, const cplx* __restrict__ Eyy, long long sE,
                                 cplx* __restrict__ R, long long sR) {
  const int pg = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2 * M * M) return;
  const int p = e % M, q = e / M;
  const double* kx = kxy + (size_t)pages[pg].lam * 2 * M;
  const double* ky = kx + M;
  cplx v = (q < M) ? cscale(Exx[(size_t)pg * sE + p + (size_t)M * q], kx[p])
                   : cscale(Eyy[(size_t)pg * sE + p + (size_t)M * (q - M)], ky[p]);
  R[(size_t)pg * sR + e] = v;
}

// Omega^2 = [[Ky^2 - Exx + Kx G1, Kx G2 - KxKy], [Ky G1 - KxKy, Kx^2 - Eyy + Ky G2]],
// G1 = [eps]^-1 Kx Exx, G2 = [eps]^-1 Ky Eyy; P, Q as in pqMatrices.m:2-4 (the
// [eps]^-1-times-diagonal terms of P Q cancel exactly).
__global__ void omega_assemble_kernel(int M, const DevPage* __restrict__ pages, const double* __restrict__ kxy,
                                      const cplx* __restrict__ G, long long sG, const cplx* __restrict__ Exx,
                                      const cplx* __restrict__ Eyy, long long sE, cplx* __restrict__ O,
                                      long long sO) {
  const int pg = blockIdx.y;
  const int n = 2 * M;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const int r = e % n, c = e / n;
  const int p = r % M, q = c % M, br = r / M, bc = c / M;
  const double* kx = kxy + (size_t)pages[pg].lam * 2 * M;
  const double* ky = kx + M;
  const cplx* g1 = G + (size_t)pg * sG;
  const cplx* g2 = g1 + (size_t)M * M;
  const double d = (p == q) ? 1.0 : 0.0;
  cplx v;
  if (br == 0 && bc == 0)
    v = cadd(csub(cmk(d * ky[p] * ky[p], 0.0), Exx[(size_t)pg * sE + p + (size_t)M * q]), cscale(g1[p + (size_t)M * q], kx[p]));
  else if (br == 0 && bc == 1)
    v = csub(cscale(g2[p + (size_t)M * q], kx[p]), cmk(d * kx[p] * ky[p], 0.0));
  else if (br == 1 && bc == 0)
    v = csub(cscale(g1[p + (size_t)M * q], ky[p]), cmk(d * kx[p] * ky[p], 0.0));
  else
    v = cadd(csub(cmk(d * kx[p] * kx[p], 0.0), Eyy[(size_t)pg * sE + p + (size_t)M * q]), cscale(g2[p + (size_t)M * q], ky[p]));
  O[(size_t)pg * sO + e] = v;
}

// ------------------------------------------------------------------ parity sectors
// Out[a,b] = sum UL[i,a]^* X[i,j] UR[j,b] with column lists (<= 4 entries each).
// X has page stride sX and is indexed by the MODE page of the solve page.
__global__ void sector_project_kernel(int nf, int nr, const DevSolve* __restrict__ solves,
                                      const DevSector* __restrict__ sectors, int left_is_h,
                                      int right_is_h, const cplx* __restrict__ X, long long sX,
                                      int x_by_solve, cplx* __restrict__ Out, long long sO) {
  const int sp = blockIdx.y;
  const DevSolve S = solves[sp];
  const DevSector& sec = sectors[S.sector];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nr * nr) return;
  const int a = e % nr, b = e / nr;
  const int* lrow = left_is_h ? sec.uh_row : sec.ue_row;
  const double* lco = left_is_h ? sec.uh_coef : sec.ue_coef;
  const int* rrow = right_is_h ? sec.uh_row : sec.ue_row;
  const double* rco = right_is_h ? sec.uh_coef : sec.ue_coef;
  const cplx* Xp = X + (size_t)(x_by_solve ? sp : S.mode) * sX;
  cplx acc = cmk(0.0, 0.0);
  for (int u = 0; u < 4; ++u) {
    const int i = lrow[4 * a + u];
    if (i < 0) continue;
    for (int v = 0; v < 4; ++v) {
      const int j = rrow[4 * b + v];
      if (j < 0) continue;
      acc = cadd(acc, cscale(Xp[i + (size_t)nf * j], lco[4 * a + u] * rco[4 * b + v]));
    }
  }
  Out[(size_t)sp * sO + e] = acc;
  (void)nf;
}

// Wf = UE Wr : Wf[i, b] = ue_rowcoef[i] * Wr[ue_rowmap[i], b]   (nf x nr)
__global__ void sector_expand_kernel(int nf, int nr, const DevSolve* __restrict__ solves,
                                     const DevSector* __restrict__ sectors, const cplx* __restrict__ Wr,
                                     long long sWr, cplx* __restrict__ Wf, long long sWf) {
  const int sp = blockIdx.y;
  const DevSector& sec = sectors[solves[sp].sector];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nf * nr) return;
  const int i = e % nf, b = e / nf;
  const int a = sec.ue_rowmap[i];
  Wf[(size_t)sp * sWf + e] = (a < 0) ? cmk(0.0, 0.0) : cscale(Wr[(size_t)sp * sWr + a + (size_t)nr * b], sec.ue_rowcoef[i]);
}

// Rows-only projection: Out[a, b] = sum_i UH[i,a] X[i, b]  (nr x ncols), X per solve page
__global__ void sector_rows_kernel(int nf, int nr, int ncols, const DevSolve* __restrict__ solves,
                                   const DevSector* __restrict__ sectors, const cplx* __restrict__ X,
                                   long long sX, cplx* __restrict__ Out, long long sO) {
  const int sp = blockIdx.y;
  const DevSector& sec = sectors[solves[sp].sector];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nr * ncols) return;
  const int a = e % nr, b = e / nr;
  cplx acc = cmk(0.0, 0.0);
  for (int u = 0; u < 4; ++u) {
    const int i = sec.uh_row[4 * a + u];
    if (i >= 0) acc = cadd(acc, cscale(X[(size_t)sp * sX + i + (size_t)nf * b], sec.uh_coef[4 * a + u]));
  }
  Out[(size_t)sp * sO + e] = acc;
}

// ------------------------------------------------------------------ modes
// lam = -i kzBranch(-mu)  (modesLayer.m:3-4)
__global__ void lambda_kernel(int n, int npages, const cplx* __restrict__ mu, cplx* __restrict__ lam) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * npages) return;
  const cplx kz = kz_branch(cneg(mu[e]));
  lam[e] = cmk(kz.y, -kz.x);  // -i * kz
}

// Q W for full W (nf x ncol) with Q = [KxKy, Eyy - Kx^2; Ky^2 - Exx, -KyKx]:
// top = kx ky W1 + T1 - kx^2 W2,  bottom = ky^2 W1 - T2 - ky kx W2,
// T1 = Eyy W2, T2 = Exx W1 (precomputed GEMMs).  Result written to Out (nf x ncol).
__global__ void qw_assemble_kernel(int M, int ncol, const DevSolve* __restrict__ solves,
                                   const DevPage* __restrict__ pages, const double* __restrict__ kxy,
                                   const cplx* __restrict__ Wf, long long sW, const cplx* __restrict__ T12,
                                   long long sT, cplx* __restrict__ Out, long long sO) {
  const int sp = blockIdx.y;
  const int nf = 2 * M;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nf * ncol) return;
  const int r = e % nf, c = e / nf, p = r % M;
  const double* kx = kxy + (size_t)pages[solves[sp].mode].lam * 2 * M;
  const double* ky = kx + M;
  const cplx* W = Wf + (size_t)sp * sW;
  const cplx w1 = W[p + (size_t)nf * c], w2 = W[M + p + (size_t)nf * c];
  const cplx* T = T12 + (size_t)sp * sT;  // [T1 (M x ncol); T2 (M x ncol)] stacked as nf x ncol
  cplx v;
  if (r < M) v = csub(cadd(cscale(w1, kx[p] * ky[p]), T[p + (size_t)nf * c]), cscale(w2, kx[p] * kx[p]));
  else v = csub(csub(cscale(w1, ky[p] * ky[p]), T[M + p + (size_t)nf * c]), cscale(w2, ky[p] * kx[p]));
  Out[(size_t)sp * sO + e] = v;
}

// X = P V0 per page in the full basis, with P = [[Kx e Ky, I - Kx e Kx],
// [Ky e Ky - I, -Ky e Kx]] (e = [eps]^-1, M x M, page stride sE; pqMatrices.m:2-4)
// and V0 the V of a uniform gap, whose four blocks are diagonal (row k of
// column c is nonzero only for k = c mod M): two terms per entry.
__global__ void pv0_kernel(int M, const DevPage* __restrict__ pages, const double* __restrict__ kxy,
                           const cplx* __restrict__ Einv, long long sE, const cplx* __restrict__ V0,
                           long long sV, cplx* __restrict__ X, long long sX) {
  const int pg = blockIdx.y;
  const int n = 2 * M;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const int r = e % n, c = e / n;
  const int p = r % M, q = c % M;
  const double* kx = kxy + (size_t)pages[pg].lam * 2 * M;
  const double* ky = kx + M;
  const cplx ein = Einv[(size_t)pg * sE + p + (size_t)M * q];
  const cplx* v0 = V0 + (size_t)pg * sV + (size_t)n * c;
  const double d = p == q ? 1.0 : 0.0;
  cplx a, b;  // P[r, q], P[r, q + M]
  if (r < M) {
    a = cscale(ein, kx[p] * ky[q]);
    b = csub(cmk(d, 0.0), cscale(ein, kx[p] * kx[q]));
  } else {
    a = csub(cscale(ein, ky[p] * ky[q]), cmk(d, 0.0));
    b = cscale(ein, -ky[p] * kx[q]);
  }
  X[(size_t)pg * sX + e] = cadd(cmul(a, v0[q]), cmul(b, v0[q + M]));
}

// X[:, c] /= lam[c]
__global__ void colscale_inv_kernel(int nr, int ncol, const cplx* __restrict__ lam, cplx* __restrict__ X,
                                    long long sX) {
  const int sp = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nr * ncol) return;
  const int c = e / nr;
  X[(size_t)sp * sX + e] = cdiv(X[(size_t)sp * sX + e], lam[(size_t)sp * nr + c]);
}

// Uniform medium of permittivity eps[pg] in the solve basis (modesHomogeneous.m:6-12):
// W = I, lam = -i [kz; kz], V = Q ./ lam.'; reduced: V_r = UH' Q UE / lam_r.
__global__ void homog_modes_kernel(int M, int nr, int reduced, const DevSolve* __restrict__ solves,
                                   const DevPage* __restrict__ pages, const DevSector* __restrict__ sectors,
                                   const double* __restrict__ kxy, const cplx* __restrict__ epsPage,
                                   cplx* __restrict__ W, cplx* __restrict__ V, long long sM,
                                   cplx* __restrict__ lamOut) {
  const int sp = blockIdx.y;
  const DevSolve S = solves[sp];
  const double* kx = kxy + (size_t)pages[S.mode].lam * 2 * M;
  const double* ky = kx + M;
  const cplx eps = epsPage[sp];
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nr * nr) return;
  const int a = e % nr, b = e / nr;
  // full-basis Q entry and lam
  auto Qf = [&](int i, int j) -> cplx {
    const int p = i % M, q = j % M;
    if (p != q) return cmk(0.0, 0.0);
    const int bi = i / M, bj = j / M;
    if (bi == 0 && bj == 0) return cmk(kx[p] * ky[p], 0.0);
    if (bi == 0 && bj == 1) return cmk(eps.x - kx[p] * kx[p], eps.y);
    if (bi == 1 && bj == 0) return cmk(ky[p] * ky[p] - eps.x, -eps.y);
    return cmk(-ky[p] * kx[p], 0.0);
  };
  auto lamf = [&](int j) -> cplx {
    const int q = j % M;
    const cplx kz = kz_nonzero(kz_branch(cmk(eps.x - kx[q] * kx[q] - ky[q] * ky[q], eps.y)));
    return cmk(kz.y, -kz.x);
  };
  cplx vq;
  cplx lb;
  if (!reduced) {
    vq = Qf(a, b);
    lb = lamf(b);
  } else {
    const DevSector& sec = sectors[S.sector];
    vq = cmk(0.0, 0.0);
    for (int u = 0; u < 4; ++u) {
      const int i = sec.uh_row[4 * a + u];
      if (i < 0) continue;
      for (int v = 0; v < 4; ++v) {
        const int j = sec.ue_row[4 * b + v];
        if (j < 0) continue;
        vq = cadd(vq, cscale(Qf(i, j), sec.uh_coef[4 * a + u] * sec.ue_coef[4 * b + v]));
      }
    }
    lb = lamf(sec.ue_row[4 * b]);
  }
  V[(size_t)sp * sM + e] = cdiv(vq, lb);
  W[(size_t)sp * sM + e] = cmk(a == b ? 1.0 : 0.0, 0.0);
  if (a == 0) lamOut[(size_t)sp * nr + b] = lb;
}

// ------------------------------------------------------------------ S-matrix helpers
// Layer S inputs in one pass: with X1 = W \ W0 and X2 = V \ V0 (rows of the
// first nScaled pages still to be divided by lam), A = X1 + X2 and B = X1 - X2
// overwrite X1 and X2, Ac receives a copy of A (for its LU), and with
// X = exp(-lam kL), R3 = [X.*B, X.*A, B] (n x 3n); Xd keeps X (n) per page.
__global__ void smat_abx_kernel(int n, const cplx* __restrict__ lam, const double* __restrict__ kL,
                                cplx* __restrict__ X1, cplx* __restrict__ X2, long long sM, int nScaled,
                                cplx* __restrict__ Ac, cplx* __restrict__ R3, long long sR, cplx* __restrict__ Xd) {
  const int sp = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const int r = e % n;
  const cplx l = lam[(size_t)sp * n + r];
  const size_t o = (size_t)sp * sM + e;
  const cplx x1 = X1[o];
  cplx x2 = X2[o];
  if (sp < nScaled) x2 = cdiv(x2, l);
  const cplx a = cadd(x1, x2), b = csub(x1, x2);
  X1[o] = a;
  X2[o] = b;
  Ac[o] = a;
  const cplx x = cexp(cscale(l, -kL[sp]));
  cplx* R = R3 + (size_t)sp * sR;
  R[e] = cmul(x, b);
  R[e + (size_t)n * n] = cmul(x, a);
  R[e + (size_t)2 * n * n] = b;
  if (e < n) Xd[(size_t)sp * n + e] = cexp(cscale(lam[(size_t)sp * n + e], -kL[sp]));
}

// One patterned layer between the two half-spaces, full basis, solved as one
// linear system on the layer's own mode amplitudes instead of through its
// S-matrix.  Layer fields [W W; -V V] (V = Q W / lam), gap fields [I I; -V0 V0]
// at both faces, X = exp(-lam kL).  With U = V0^-1 V, Wp = W + U, Wm = W - U the
// gap amplitudes at the faces are
//   c1+ = Wp h+ + Wm X h-,  c1- = Wm h+ + Wp X h-   (z = 0)
//   c2+ = Wp X h+ + Wm h-,  c2- = Wm X h+ + Wp h-   (z = d)
// for mode amplitudes h = c / 2, and c1+ = a21 C + a22 c1-, c2- = b11 c2+ give
//   [A1 B1; B2 A2] [h+; h-] = [a21 C; 0],
//   A1 = Wp - a22 Wm, B1 = (Wm - a22 Wp) X, B2 = (Wm - b11 Wp) X, A2 = Wp - b11 Wm.
// This kernel writes U and the four blocks.  V0, a22 and b11 couple harmonic p
// only with p + M (2 x 2 per harmonic) and are read from the per-context tables
// (page grp[sp], stride n^2); Q W comes from W and T12 = [Eyy W2; Exx W1]
// (pqMatrices.m:5-7).  One thread per harmonic p and group of kDirectCols
// columns: rows p and p + M of every output.
constexpr int kDirectCols = 8;
__global__ void direct_blocks_kernel(int M, const DevSolve* __restrict__ solves, const DevPage* __restrict__ pages,
                                     const double* __restrict__ kxy, const cplx* __restrict__ W,
                                     const cplx* __restrict__ T12, const cplx* __restrict__ lam,
                                     const double* __restrict__ kL, const int* __restrict__ grp,
                                     const cplx* __restrict__ V0, const cplx* __restrict__ a22,
                                     const cplx* __restrict__ b11, long long s, cplx* __restrict__ U,
                                     cplx* __restrict__ A1, cplx* __restrict__ B1, cplx* __restrict__ B2,
                                     cplx* __restrict__ A2) {
  const int sp = blockIdx.y;
  const int n = 2 * M, nb = (n + kDirectCols - 1) / kDirectCols;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= M * nb) return;
  const int p = e % M, q = p + M, j0 = (e / M) * kDirectCols, j1 = min(j0 + kDirectCols, n);
  const size_t o = (size_t)sp * s;
  const double* kx = kxy + (size_t)pages[solves[sp].mode].lam * 2 * M;
  const double kxp = kx[p], kyp = kx[M + p];
  // 2 x 2 blocks (p, p), (p, q), (q, p), (q, q) of the context's tables
  const size_t og = (size_t)grp[sp] * n * n;
  const size_t ipp = og + p + (size_t)n * p, ipq = og + p + (size_t)n * q, iqp = og + q + (size_t)n * p,
               iqq = og + q + (size_t)n * q;
  const cplx a0 = a22[ipp], a1 = a22[ipq], a2 = a22[iqp], a3 = a22[iqq];
  const cplx b0 = b11[ipp], b1 = b11[ipq], b2 = b11[iqp], b3 = b11[iqq];
  const cplx v0 = V0[ipp], v1 = V0[ipq], v2 = V0[iqp], v3 = V0[iqq];
  const cplx idet = crecip(csub(cmul(v0, v3), cmul(v1, v2)));
  const cplx g0 = cmul(v3, idet), g1 = cneg(cmul(v1, idet)), g2 = cneg(cmul(v2, idet)), g3 = cmul(v0, idet);
  const double kl = kL[sp];
  for (int j = j0; j < j1; ++j) {
    const size_t c = o + (size_t)n * j;
    const cplx w1 = W[c + p], w2 = W[c + q];
    const cplx q1 = csub(cadd(cscale(w1, kxp * kyp), T12[c + p]), cscale(w2, kxp * kxp));
    const cplx q2 = csub(csub(cscale(w1, kyp * kyp), T12[c + q]), cscale(w2, kyp * kxp));
    const cplx l = lam[(size_t)sp * n + j];
    const cplx il = crecip(l);
    const cplx u1 = cmul(cadd(cmul(g0, q1), cmul(g1, q2)), il);
    const cplx u2 = cmul(cadd(cmul(g2, q1), cmul(g3, q2)), il);
    const cplx p1 = cadd(w1, u1), p2 = cadd(w2, u2), m1 = csub(w1, u1), m2 = csub(w2, u2);
    const cplx x = cexp(cscale(l, -kl));
    U[c + p] = u1;
    U[c + q] = u2;
    A1[c + p] = csub(p1, cadd(cmul(a0, m1), cmul(a1, m2)));
    A1[c + q] = csub(p2, cadd(cmul(a2, m1), cmul(a3, m2)));
    B1[c + p] = cmul(x, csub(m1, cadd(cmul(a0, p1), cmul(a1, p2))));
    B1[c + q] = cmul(x, csub(m2, cadd(cmul(a2, p1), cmul(a3, p2))));
    B2[c + p] = cmul(x, csub(m1, cadd(cmul(b0, p1), cmul(b1, p2))));
    B2[c + q] = cmul(x, csub(m2, cadd(cmul(b2, p1), cmul(b3, p2))));
    A2[c + p] = csub(p1, cadd(cmul(b0, m1), cmul(b1, m2)));
    A2[c + q] = csub(p2, cadd(cmul(b2, m1), cmul(b3, m2)));
  }
}

// Right-hand sides of the face amplitudes from the mode amplitudes h+ (Hp) and
// h- (Hm), n x ncol each: c1- = W S1 + U D1 and c2+ = W S2 + U D2 with
// S = [h+ + X h-, X h+ + h-] and D = [X h- - h+, X h+ - h-] (n x 2 ncol each).
__global__ void direct_vec_kernel(int n, int ncol, const cplx* __restrict__ lam, const double* __restrict__ kL,
                                  const cplx* __restrict__ Hp, const cplx* __restrict__ Hm, cplx* __restrict__ S,
                                  cplx* __restrict__ D, long long s) {
  const int sp = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * ncol) return;
  const int i = e % n;
  const size_t o = (size_t)sp * s;
  const cplx x = cexp(cscale(lam[(size_t)sp * n + i], -kL[sp]));
  const cplx hp = Hp[o + e], hm = Hm[o + e];
  const cplx xp = cmul(x, hp), xm = cmul(x, hm);
  const size_t e2 = e + (size_t)n * ncol;
  S[o + e] = cadd(hp, xm);
  S[o + e2] = cadd(xp, hm);
  D[o + e] = csub(xm, hp);
  D[o + e2] = csub(xp, hm);
}

// [W + U, W - U] (n x 2n per page, stride sO) for the sector-basis form of the
// single-layer system (the full basis forms its blocks in direct_blocks_kernel).
__global__ void wpm_kernel(int n, const cplx* __restrict__ W, const cplx* __restrict__ U, long long s,
                           cplx* __restrict__ out, long long sO) {
  const int sp = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const cplx w = W[(size_t)sp * s + e], u = U[(size_t)sp * s + e];
  cplx* o = out + (size_t)sp * sO;
  o[e] = cadd(w, u);
  o[e + (size_t)n * n] = csub(w, u);
}

// The four blocks of the single-layer system from dense products (sector bases):
// Wpm = [W+ W-] (stride 2 n^2), Pa = a22 [W+ W-] (stride 3 n^2), Pbp = b11 W+,
// Pbm = b11 W- (stride s); X = exp(-lam kL) scales columns:
//   A1 = W+ - a22 W-, B1 = (W- - a22 W+) X, B2 = (W- - b11 W+) X, A2 = W+ - b11 W-.
__global__ void dense_blocks_kernel(int n, const cplx* __restrict__ lam, const double* __restrict__ kL,
                                    const cplx* __restrict__ Wpm, const cplx* __restrict__ Pa,
                                    const cplx* __restrict__ Pbp, const cplx* __restrict__ Pbm, long long s,
                                    cplx* __restrict__ A1, cplx* __restrict__ B1, cplx* __restrict__ B2,
                                    cplx* __restrict__ A2) {
  const int sp = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const size_t nn = (size_t)n * n, o = (size_t)sp * s + e;
  const cplx x = cexp(cscale(lam[(size_t)sp * n + e / n], -kL[sp]));
  const cplx* wpm = Wpm + (size_t)sp * 2 * nn;
  const cplx* pa = Pa + (size_t)sp * 3 * nn;
  const cplx wp = wpm[e], wm = wpm[e + nn];
  A1[o] = csub(wp, pa[e + nn]);
  B1[o] = cmul(x, csub(wm, pa[e]));
  B2[o] = cmul(x, csub(wm, Pbp[o]));
  A2[o] = csub(wp, Pbm[o]);
}

// h+ right-hand side a21 C: columns colIdx of the context's a21 (page grp[sp]).
__global__ void direct_rhs_kernel(int n, int ncol, const int* __restrict__ colIdx, const int* __restrict__ grp,
                                  const cplx* __restrict__ a21, cplx* __restrict__ out, long long sO) {
  const int sp = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * ncol) return;
  const int r = e % n, q = e / n;
  out[(size_t)sp * sO + e] = a21[(size_t)grp[sp] * n * n + r + (size_t)n * colIdx[sp * ncol + q]];
}

// R = a11 C + a12 c1-, T = b21 c2+ with Y = [c1- c2+] (n x 2 ncol, page stride sY);
// a11, a12, b21 per context (page grp[sp]), pair structure as in rx_pair_kernel.
__global__ void direct_out_kernel(int M, int ncol, const int* __restrict__ colIdx, const int* __restrict__ grp,
                                  const cplx* __restrict__ a11, const cplx* __restrict__ a12,
                                  const cplx* __restrict__ b21, const cplx* __restrict__ Y, long long sY,
                                  cplx* __restrict__ R, cplx* __restrict__ T) {
  const int n = 2 * M;
  const int sp = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * ncol) return;
  const int i = e % n, j = e / n, i1 = i % M, i2 = i1 + M;
  const size_t og = (size_t)grp[sp] * n * n;
  const cplx* y = Y + (size_t)sp * sY;
  const cplx* c1 = y + (size_t)n * j;
  const cplx* c2 = y + (size_t)n * (ncol + j);
  const cplx r = cadd(a11[og + i + (size_t)n * colIdx[sp * ncol + j]],
                      cadd(cmul(a12[og + i + (size_t)n * i1], c1[i1]), cmul(a12[og + i + (size_t)n * i2], c1[i2])));
  const cplx t = cadd(cmul(b21[og + i + (size_t)n * i1], c2[i1]), cmul(b21[og + i + (size_t)n * i2], c2[i2]));
  R[(size_t)sp * n * ncol + e] = r;
  T[(size_t)sp * n * ncol + e] = t;
}

// out = s1 diag(X) P + s2 C  (n x n per page; out may alias P or C)
__global__ void rowscale_add_kernel(int n, const cplx* __restrict__ Xd, const cplx* P, long long sP, double s1,
                                    const cplx* C, long long sC, double s2, cplx* out, long long sO) {
  const int sp = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const int r = e % n;
  const cplx xp = cmul(Xd[(size_t)sp * n + r], P[(size_t)sp * sP + e]);
  out[(size_t)sp * sO + e] = cadd(cscale(xp, s1), cscale(C[(size_t)sp * sC + e], s2));
}

// Y = diag(X) * Y (rows scaled)
__global__ void rowscale_kernel(int n, int ncol, const cplx* __restrict__ Xd, cplx* __restrict__ Y,
                                long long sY) {
  const int sp = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * ncol) return;
  const int r = e % n;
  Y[(size_t)sp * sY + e] = cmul(Xd[(size_t)sp * n + r], Y[(size_t)sp * sY + e]);
}

// Y = alpha * I + beta * X  (n x n per page; X may be null for alpha I)
__global__ void eye_axpby_kernel(int n, cplx alpha, cplx beta, const cplx* __restrict__ X, long long sX,
                                 cplx* __restrict__ Y, long long sY) {
  const int sp = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const int r = e % n, c = e / n;
  cplx v = (r == c) ? alpha : cmk(0.0, 0.0);
  if (X) v = cadd(v, cmul(beta, X[(size_t)sp * sX + e]));
  Y[(size_t)sp * sY + e] = v;
}

// Copy a per-group matrix (group index from a table) to every page.
__global__ void gather_pages_kernel(long long elems, const int* __restrict__ groupOf,
                                    const cplx* __restrict__ src, long long sSrc, cplx* __restrict__ dst,
                                    long long sDst) {
  const int sp = blockIdx.y;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= elems) return;
  dst[(size_t)sp * sDst + e] = src[(size_t)groupOf[sp] * sSrc + e];
}

// dst page p = src page p (sSrc = 0: the one src matrix to every page), pages along blockIdx.y
__global__ void page_copy_kernel(long long elems, const cplx* __restrict__ src, long long sSrc,
                                 cplx* __restrict__ dst, long long sDst) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= elems) return;
  dst[(size_t)blockIdx.y * sDst + e] = src[(size_t)blockIdx.y * sSrc + e];
}

// ------------------------------------------------------------------ outputs
// One block per (solve page, source).  r_xy = S11 c, t_xy = S21 c (full basis
// or UE (S c_r) in a parity sector); z from transversality; R, T per order;
// Rtot, Ttot.  Out index of (b, l, pol): b + B (l + L pol).
__global__ void output_kernel(int M, int nr, int reduced, int page0, int npages, const DevSolve* __restrict__ solves,
                              const DevPage* __restrict__ pages, const DevSector* __restrict__ sectors,
                              const DevSource* __restrict__ src, const double* __restrict__ kxy,
                              const double* __restrict__ kzIncSub, const double* __restrict__ kzinc0s,
                              const cplx* __restrict__ S11,
                              const cplx* __restrict__ S21, long long sS, double* __restrict__ R,
                              double* __restrict__ T, double* __restrict__ Rtot, double* __restrict__ Ttot,
                              cplx* __restrict__ rOut, cplx* __restrict__ tOut, int i0) {
  __shared__ double red[8];
  const int sp = blockIdx.x;
  const int srcIdx = blockIdx.y;
  const DevSolve S = solves[sp];
  const DevSource so = src[srcIdx];
  if (so.sector != S.sector) return;  // this solve page does not carry that source
  const DevPage P = pages[S.mode];
  const double* kx = kxy + (size_t)P.lam * 2 * M;
  const double* ky = kx + M;
  const cplx* kzi = (const cplx*)kzIncSub + (size_t)P.lam * 2 * M;
  const cplx* kzs = kzi + M;
  const double kzi0 = kzinc0s[P.lam];
  const cplx* s11 = S11 + (size_t)sp * sS;
  const cplx* s21 = S21 + (size_t)sp * sS;
  const long long outIdx = (long long)(page0 + S.mode) + (long long)npages * so.pol;
  double rs = 0.0, ts = 0.0;
  for (int p = threadIdx.x; p < M; p += blockDim.x) {
    cplx rxy[2], txy[2];
    for (int comp = 0; comp < 2; ++comp) {
      const int i = comp * M + p;
      cplx rv = cmk(0.0, 0.0), tv = cmk(0.0, 0.0);
      // s11 / s21 hold the columns of S11 / S21 selected for the sources:
      // full basis: columns (i0, M + i0); parity sector: the source column srcCol
      if (!reduced) {
        rv = cadd(cmul(s11[i], so.ex), cmul(s11[i + (size_t)nr], so.ey));
        tv = cadd(cmul(s21[i], so.ex), cmul(s21[i + (size_t)nr], so.ey));
      } else {
        const DevSector& sec = sectors[S.sector];
        const int a = sec.ue_rowmap[i];
        if (a >= 0) {
          const double co = sec.ue_rowcoef[i];
          // c_r = UE' c has one nonzero, at the selected column, of value srcVal
          rv = cscale(cmul(s11[a], so.srcVal), co);
          tv = cscale(cmul(s21[a], so.srcVal), co);
        }
      }
      rxy[comp] = rv;
      txy[comp] = tv;
    }
    const cplx rz = cdiv(cadd(cscale(rxy[0], kx[p]), cscale(rxy[1], ky[p])), kzi[p]);
    const cplx tz = cneg(cdiv(cadd(cscale(txy[0], kx[p]), cscale(txy[1], ky[p])), kzs[p]));
    const double Rp = kzi[p].x / kzi0 * (cnorm2(rxy[0]) + cnorm2(rxy[1]) + cnorm2(rz));
    const double Tp = kzs[p].x / kzi0 * (cnorm2(txy[0]) + cnorm2(txy[1]) + cnorm2(tz));
    R[p + (size_t)M * outIdx] = Rp;
    T[p + (size_t)M * outIdx] = Tp;
    rs += Rp;
    ts += Tp;
    if (rOut) {
      cplx* ro = rOut + 3 * (p + (size_t)M * outIdx);
      ro[0] = rxy[0]; ro[1] = rxy[1]; ro[2] = rz;
    }
    if (tOut) {
      cplx* to = tOut + 3 * (p + (size_t)M * outIdx);
      to[0] = txy[0]; to[1] = txy[1]; to[2] = tz;
    }
  }
  rs = block_sum(rs, red);
  ts = block_sum(ts, red + 4);
  if (threadIdx.x == 0) {
    Rtot[outIdx] = rs;
    Ttot[outIdx] = ts;
  }
}

// count identity matrices of size n x n, stacked
__global__ void eye_stack_kernel(int n, long long count, cplx* __restrict__ Y) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n2 = (long long)n * n;
  if (e >= n2 * count) return;
  const long long r = e % n2;
  Y[e] = cmk((r % n) == (r / n) ? 1.0 : 0.0, 0.0);
}

// out[:, q] = S[:, colIdx[page * ncol + q]]  (n x ncol per page)
// Products with a half-space S-matrix block in the full (m, n, x/y) basis:
// with W = W0 = I every block couples harmonic j only to itself and to its
// other polarisation j +- M, so R(:, j) has nonzeros in rows j mod M and
// j mod M + M only.  Out = alpha X R + beta Cin (xr) or alpha R X + beta Cin
// (rx), n = 2M, pages along blockIdx.y; Out may be Cin.
__global__ void xr_pair_kernel(int M, const cplx* __restrict__ X, long long sX, const cplx* __restrict__ R,
                               long long sR, cplx alpha, cplx beta, const cplx* Cin, long long sC, cplx* Out,
                               long long sO) {
  const int n = 2 * M;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * n) return;
  const int i = (int)(e % n), j = (int)(e / n), j1 = j % M, j2 = j1 + M;
  X += (size_t)blockIdx.y * sX;
  R += (size_t)blockIdx.y * sR;
  const cplx v = cadd(cmul(X[i + (size_t)j1 * n], R[j1 + (size_t)j * n]), cmul(X[i + (size_t)j2 * n], R[j2 + (size_t)j * n]));
  cplx o = cmul(alpha, v);
  if (beta.x != 0.0 || beta.y != 0.0) o = cadd(o, cmul(beta, Cin[(size_t)blockIdx.y * sC + e]));
  Out[(size_t)blockIdx.y * sO + e] = o;
}

__global__ void rx_pair_kernel(int M, int ncol, const cplx* __restrict__ R, long long sR, const cplx* __restrict__ X,
                               long long sX, cplx alpha, cplx beta, const cplx* Cin, long long sC, cplx* Out,
                               long long sO) {
  const int n = 2 * M;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * ncol) return;
  const int i = (int)(e % n), j = (int)(e / n), i1 = i % M, i2 = i1 + M;
  X += (size_t)blockIdx.y * sX;
  R += (size_t)blockIdx.y * sR;
  const cplx v = cadd(cmul(R[i + (size_t)i1 * n], X[i1 + (size_t)j * n]), cmul(R[i + (size_t)i2 * n], X[i2 + (size_t)j * n]));
  cplx o = cmul(alpha, v);
  if (beta.x != 0.0 || beta.y != 0.0) o = cadd(o, cmul(beta, Cin[(size_t)blockIdx.y * sC + e]));
  Out[(size_t)blockIdx.y * sO + e] = o;
}

__global__ void select_cols_kernel(int n, int ncol, const int* __restrict__ colIdx, const cplx* __restrict__ S,
                                   long long sS, cplx* __restrict__ out, long long sO) {
  const int sp = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * ncol) return;
  const int r = e % n, q = e / n;
  out[(size_t)sp * sO + e] = S[(size_t)sp * sS + r + (size_t)n * colIdx[sp * ncol + q]];
}

// ====================================================================== fields (rcwaFields.m)
// Mode amplitudes at depths zl[k] inside one region:
//   Pp[:, k] = fp .* ap + fm .* am,  Pm[:, k] = -fp .* ap + fm .* am,
//   fp = exp(-lam k0 zl), fm = exp(-lam k0 (d - zl))   (am may be null)
__global__ void field_amp_kernel(int n, int nz, const cplx* __restrict__ lam, const cplx* __restrict__ ap,
                                 const cplx* __restrict__ am, double k0, double d, const double* __restrict__ zl,
                                 cplx* __restrict__ Pp, cplx* __restrict__ Pm) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * nz) return;
  const int r = e % n, k = e / n;
  const cplx l = lam[r];
  const cplx fp = cexp(cscale(l, -k0 * zl[k]));
  const cplx a = cmul(fp, ap[r]);
  cplx b = cmk(0.0, 0.0);
  if (am) b = cmul(cexp(cscale(l, -k0 * (d - zl[k]))), am[r]);
  Pp[e] = cadd(a, b);
  Pm[e] = csub(b, a);
}

// out[:, k] = fac (Kx v_y - Ky v_x) for the 2M-row columns v of Vv (z components of the curls)
__global__ void curl_z_kernel(int M, int nz, const double* __restrict__ kx, const double* __restrict__ ky,
                              const cplx* __restrict__ Vv, cplx fac, cplx* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= M * nz) return;
  const int p = e % M, k = e / M;
  const cplx* v = Vv + (size_t)k * 2 * M;
  out[e] = cmul(fac, csub(cscale(v[M + p], kx[p]), cscale(v[p], ky[p])));
}

// C(:, col[k]) = src(:, k), n rows
__global__ void scatter_cols_kernel(int n, int ncol, const int* __restrict__ col, const cplx* __restrict__ src,
                                    int lds, cplx* __restrict__ dst, int ldd) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * ncol) return;
  const int r = e % n, k = e / n;
  dst[(size_t)r + (size_t)col[k] * ldd] = src[(size_t)r + (size_t)k * lds];
}

// Phi[i + nx j, p] = exp(i k0 (kx_p x_i + ky_p y_j))
__global__ void phase_kernel(int M, int nx, int ny, const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ kx, const double* __restrict__ ky, double k0,
                             cplx* __restrict__ Phi) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nxy = (long long)nx * ny;
  if (e >= nxy * M) return;
  const int pt = (int)(e % nxy), p = (int)(e / nxy);
  const int i = pt % nx, j = pt / nx;
  double sn, cs;
  sincos(k0 * (kx[p] * x[i] + ky[p] * y[j]), &sn, &cs);
  Phi[e] = cmk(cs, sn);
}

// flux[k] = -Im sum_p (Ex_p conj(H~y_p) - Ey_p conj(H~x_p)) / kzi0, from the 2M-row columns of S and U
__global__ void flux_kernel(int M, const cplx* __restrict__ S, const cplx* __restrict__ U, double inv_kzi0,
                            double* __restrict__ flux) {
  __shared__ double red[8];
  const int k = blockIdx.x;
  const cplx* s = S + (size_t)k * 2 * M;
  const cplx* u = U + (size_t)k * 2 * M;
  double acc = 0.0;
  for (int p = threadIdx.x; p < M; p += blockDim.x) {
    const cplx a = cmulc(u[M + p], s[p]);   // Ex conj(Hy)
    const cplx b = cmulc(u[p], s[M + p]);   // Ey conj(Hx)
    acc -= a.y - b.y;
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) flux[k] = acc * inv_kzi0;
}

// s = W c for W = I or a diagonal-free copy: out = c (used for the half-spaces, W = I)
__global__ void copy_cols_kernel(long long count, const cplx* __restrict__ src, cplx* __restrict__ dst) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < count) dst[e] = src[e];
}

// columns divided by a scalar permittivity (homogeneous [eps]^-1)
__global__ void scale_kernel(long long count, cplx f, cplx* __restrict__ v) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < count) v[e] = cmul(f, v[e]);
}
