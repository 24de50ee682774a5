// Host runtime of the gfx950 RCWA solver: validation, device-resident plans,
// the per-layer stage chain and the C ABI declared in include/orf_rcwa.h.
//
// The stage chain restates the reference's rcwa2D pipeline (rcwa2D.m:1-35):
// harmonics and k-vectors (harmonics2D.m, A.2 sign convention), geometry
// sampling (cellGrid.m, layerEps.m), Laurent / Li convolution matrices
// (fourierWeights.m, convMatrix2D.m, convMatricesLi2D.m), Omega^2 = P Q
// (pqMatrices.m), eigenmodes (modesLayer.m, modesHomogeneous.m), layer and
// half-space S-matrices (smatrixLayer.m, smatrixRegion.m), Redheffer star
// products (smatrixStar.m) and the efficiency reduction (rcwa2D.m:31-34).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <map>
#include <mutex>
#include <unordered_map>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/orf_pageops.h"
#include "../../include/orf_rcwa.h"
#include "orf_internal.h"
#include "orf_rcwa_dev.h"

typedef std::complex<double> zc;

// kernels from rcwa_kernels.hip
__global__ void raster_kernel(int, int, const DevGrid*, const DevShape*, const double*, uint8_t*);
__global__ void grid_symmetry_kernel(int, int, const uint8_t*, int*);
__global__ void chi_x_kernel(int, int, int, int, const uint8_t*, const cplx*, cplx*);
__global__ void chi_y_kernel(int, int, int, int, const uint8_t*, const cplx*, cplx*);
__global__ void chi_2d_kernel(int, int, int, int, const cplx*, const cplx*, cplx*);
__global__ void laurent_kernel(int, int, int, const DevPage*, const cplx*, const cplx*, int, cplx*, long long);
__global__ void li_toeplitz_kernel(int, int, int, int, const DevPage*, const cplx*, long long, const cplx*, int, cplx*);
__global__ void li_gather_kernel(int, int, int, const cplx*, const cplx*, cplx*, cplx*, long long);
__global__ void omega_rhs_kernel(int, const DevPage*, const double*, const cplx*, const cplx*, long long, cplx*, long long);
__global__ void omega_assemble_kernel(int, const DevPage*, const double*, const cplx*, long long, const cplx*,
                                      const cplx*, long long, cplx*, long long);
__global__ void sector_project_kernel(int, int, const DevSolve*, const DevSector*, int, int, const cplx*, long long,
                                      int, cplx*, long long);
__global__ void sector_expand_kernel(int, int, const DevSolve*, const DevSector*, const cplx*, long long, cplx*,
                                     long long);
__global__ void sector_rows_kernel(int, int, int, const DevSolve*, const DevSector*, const cplx*, long long, cplx*,
                                   long long);
__global__ void lambda_kernel(int, int, const cplx*, cplx*);
__global__ void qw_assemble_kernel(int, int, const DevSolve*, const DevPage*, const double*, const cplx*, long long,
                                   const cplx*, long long, cplx*, long long);
__global__ void colscale_inv_kernel(int, int, const cplx*, cplx*, long long);
__global__ void pv0_kernel(int, const DevPage*, const double*, const cplx*, long long, const cplx*, long long, cplx*,
                           long long);
__global__ void homog_modes_kernel(int, int, int, const DevSolve*, const DevPage*, const DevSector*, const double*,
                                   const cplx*, cplx*, cplx*, long long, cplx*);
__global__ void smat_abx_kernel(int, const cplx*, const double*, cplx*, cplx*, long long, int, cplx*, cplx*, long long,
                                cplx*);
__global__ void direct_blocks_kernel(int, const DevSolve*, const DevPage*, const double*, const cplx*, const cplx*,
                                     const cplx*, const double*, const int*, const cplx*, const cplx*, const cplx*,
                                     long long, cplx*, cplx*, cplx*, cplx*, cplx*);
__global__ void direct_rhs_kernel(int, int, const int*, const int*, const cplx*, cplx*, long long);
__global__ void wpm_kernel(int, const cplx*, const cplx*, long long, cplx*, long long);
__global__ void dense_blocks_kernel(int, const cplx*, const double*, const cplx*, const cplx*, const cplx*,
                                    const cplx*, long long, cplx*, cplx*, cplx*, cplx*);
__global__ void direct_out_kernel(int, int, const int*, const int*, const cplx*, const cplx*, const cplx*, const cplx*,
                                  long long, cplx*, cplx*);
__global__ void direct_vec_kernel(int, int, const cplx*, const double*, const cplx*, const cplx*, cplx*, cplx*,
                                  long long);
__global__ void rowscale_kernel(int, int, const cplx*, cplx*, long long);
__global__ void rowscale_add_kernel(int, const cplx*, const cplx*, long long, double, const cplx*, long long, double,
                                    cplx*, long long);
__global__ void eye_axpby_kernel(int, cplx, cplx, const cplx*, long long, cplx*, long long);
__global__ void gather_pages_kernel(long long, const int*, const cplx*, long long, cplx*, long long);
__global__ void eye_stack_kernel(int, long long, cplx*);
__global__ void select_cols_kernel(int, int, const int*, const cplx*, long long, cplx*, long long);
__global__ void xr_pair_kernel(int, const cplx*, long long, const cplx*, long long, cplx, cplx, const cplx*, long long,
                               cplx*, long long);
__global__ void rx_pair_kernel(int, int, const cplx*, long long, const cplx*, long long, cplx, cplx, const cplx*,
                               long long, cplx*, long long);
__global__ void field_amp_kernel(int, int, const cplx*, const cplx*, const cplx*, double, double, const double*, cplx*,
                                 cplx*);
__global__ void curl_z_kernel(int, int, const double*, const double*, const cplx*, cplx, cplx*);
__global__ void scatter_cols_kernel(int, int, const int*, const cplx*, int, cplx*, int);
__global__ void phase_kernel(int, int, int, const double*, const double*, const double*, const double*, double, cplx*);
__global__ void flux_kernel(int, const cplx*, const cplx*, double, double*);
__global__ void copy_cols_kernel(long long, const cplx*, cplx*);
__global__ void page_copy_kernel(long long, const cplx*, long long, cplx*, long long);
__global__ void scale_kernel(long long, cplx, cplx*);
__global__ void output_kernel(int, int, int, int, int, const DevSolve*, const DevPage*, const DevSector*,
                              const DevSource*, const double*, const double*, const double*, const cplx*,
                              const cplx*, long long, double*, double*, double*, double*, cplx*, cplx*, int);

// ====================================================================== errors
static thread_local std::string g_err;
struct OrfError {
  int code;
  std::string msg;
};
static void fail(int code, const std::string& m) { throw OrfError{code, m}; }
#define HIPCHECK(x)                                                                     \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) {                                                             \
      (void)hipGetLastError(); /* the error is reported here, not again by a later call */ \
      if (e_ == hipErrorOutOfMemory) fail(ORF_OOM, std::string("out of device memory: ") + #x); \
      fail(ORF_DEVICE_ERROR, std::string(hipGetErrorString(e_)) + " at " + #x);         \
    }                                                                                   \
  } while (0)
#define KCHECK() HIPCHECK(hipGetLastError())

static inline cplx C(zc z) { return make_double2(z.real(), z.imag()); }
static inline dim3 g1(long long n, int y = 1) { return dim3((unsigned)((n + 255) / 256), y); }

// ====================================================================== timing
struct StageTimer {
  bool on = false;
  struct Rec { const char* name; hipEvent_t a, b; };
  std::vector<Rec> pending;
  std::vector<size_t> open;  // indices of begun, not yet ended records (stages nest)
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<double, int>> acc;
  std::vector<std::string> order;
  hipEvent_t get() {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e;
    HIPCHECK(hipEventCreate(&e));
    return e;
  }
  void begin(const char* name, hipStream_t st) {
    if (!on) return;
    Rec r{name, get(), get()};
    HIPCHECK(hipEventRecord(r.a, st));
    open.push_back(pending.size());
    pending.push_back(r);
  }
  void end(hipStream_t st) {
    if (!on || open.empty()) return;
    HIPCHECK(hipEventRecord(pending[open.back()].b, st));
    open.pop_back();
  }
  void collect() {
    for (auto& r : pending) {
      float ms = 0.f;
      HIPCHECK(hipEventSynchronize(r.b));
      HIPCHECK(hipEventElapsedTime(&ms, r.a, r.b));
      auto it = acc.find(r.name);
      if (it == acc.end()) { order.push_back(r.name); acc[r.name] = {ms, 1}; }
      else { it->second.first += ms; it->second.second += 1; }
      pool.push_back(r.a);
      pool.push_back(r.b);
    }
    pending.clear();
  }
  void reset() { acc.clear(); order.clear(); open.clear(); }
  ~StageTimer() {
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

// ====================================================================== device memory
// Caching allocator.  hipFree waits for the whole device, so a serving loop
// that creates and destroys plans while other plans run on other streams
// would serialise on it; freed blocks are kept here instead and handed to the
// next request of a similar size (best fit, at most 25 % larger).  Every
// owner synchronises its stream before it returns a block.
struct DevicePool {
  std::mutex mu;
  std::multimap<size_t, void*> freeBlocks;
  std::unordered_map<void*, size_t> sizeOf;
  size_t cachedBytes = 0, liveBytes = 0;
  static size_t round(size_t b) {
    const size_t g = b >= (size_t(1) << 20) ? (size_t(2) << 20) : 512;
    return (b + g - 1) / g * g;
  }
  void* get(size_t bytes) {
    const size_t want = round(bytes);
    {
      std::lock_guard<std::mutex> lk(mu);
      auto it = freeBlocks.lower_bound(want);
      if (it != freeBlocks.end() && it->first <= want + want / 4) {
        void* p = it->second;
        cachedBytes -= it->first;
        liveBytes += it->first;
        freeBlocks.erase(it);
        return p;
      }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e == hipErrorOutOfMemory) {
      (void)hipGetLastError();
      trim();
      e = hipMalloc(&p, want);
    }
    if (e != hipSuccess) (void)hipGetLastError();  // reported here, not again by the next call's checks
    if (e == hipErrorOutOfMemory) fail(ORF_OOM, "out of device memory (" + std::to_string(want >> 20) + " MiB)");
    if (e != hipSuccess) fail(ORF_DEVICE_ERROR, std::string(hipGetErrorString(e)) + " in hipMalloc");
    std::lock_guard<std::mutex> lk(mu);
    sizeOf[p] = want;
    liveBytes += want;
    return p;
  }
  void put(void* p) {
    std::lock_guard<std::mutex> lk(mu);
    auto it = sizeOf.find(p);
    if (it == sizeOf.end()) return;
    freeBlocks.emplace(it->second, p);
    cachedBytes += it->second;
    liveBytes -= it->second;
  }
  void trim() {
    std::lock_guard<std::mutex> lk(mu);
    for (auto& kv : freeBlocks) {
      (void)hipFree(kv.second);
      sizeOf.erase(kv.second);
    }
    freeBlocks.clear();
    cachedBytes = 0;
  }
};
static DevicePool g_pool;

struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DBuf() {}
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  // Blocks go back to the pool only when no kernel can still write them.  Every
  // owner synchronises its stream before a normal return; while an exception
  // unwinds (an allocation failing half-way through a stage), kernels already
  // queued on the owner's stream may still run, so the device is drained first.
  ~DBuf() {
    if (!p) return;
    if (std::uncaught_exceptions() > 0) (void)hipDeviceSynchronize();
    g_pool.put(p);
  }
  void alloc(size_t b) {
    if (p) { g_pool.put(p); p = nullptr; }
    bytes = b;
    if (b) p = g_pool.get(b);
  }
  template <class T> T* as() const { return (T*)p; }
  template <class T> void upload(const std::vector<T>& v) {
    alloc(v.size() * sizeof(T));
    if (!v.empty()) HIPCHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  }
};

struct orf_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
};

// ====================================================================== host helpers
static zc kz_branch_h(zc kz2) {
  zc kz = std::sqrt(kz2);
  if (kz.imag() < 0 || (kz.imag() == 0 && kz.real() < 0)) kz = -kz;
  return kz;
}
static zc kz_nonzero_h(zc kz) { return std::abs(kz) < 1e-8 ? zc(1e-8, 1e-8) : kz; }

// polarizationVector.m:4-20
static void pol_vector(const double khat[3], zc jte, zc jtm, zc E[3]) {
  double aTE[3];
  if (std::fabs(khat[0]) < 1e-12 && std::fabs(khat[1]) < 1e-12) {
    aTE[0] = 0; aTE[1] = 1; aTE[2] = 0;
  } else {
    aTE[0] = -khat[1]; aTE[1] = khat[0]; aTE[2] = 0;  // z x khat
    double nn = std::sqrt(aTE[0] * aTE[0] + aTE[1] * aTE[1]);
    aTE[0] /= nn; aTE[1] /= nn;
  }
  double aTM[3] = {aTE[1] * khat[2] - aTE[2] * khat[1], aTE[2] * khat[0] - aTE[0] * khat[2],
                   aTE[0] * khat[1] - aTE[1] * khat[0]};
  double nrm = 0;
  for (int i = 0; i < 3; ++i) {
    E[i] = jte * aTE[i] + jtm * aTM[i];
    nrm += std::norm(E[i]);
  }
  nrm = std::sqrt(nrm);
  for (int i = 0; i < 3; ++i) E[i] /= nrm;
}

// exact staircase Fourier weights (fourierWeights.m:7-12), (4N+1) x Ns, h fastest
static std::vector<cplx> fourier_weights_h(int Ns, int N) {
  const int H = 4 * N + 1;
  std::vector<cplx> F((size_t)H * Ns);
  for (int j = 0; j < Ns; ++j) {
    const double xj = (j + 0.5) / Ns - 0.5;
    for (int hi = 0; hi < H; ++hi) {
      const double h = hi - 2 * N;
      const double u = M_PI * h / Ns;
      const double s = (u != 0) ? std::sin(u) / u : 1.0;
      zc v = s * std::exp(zc(0, -2 * M_PI * h * xj)) / (double)Ns;
      F[hi + (size_t)H * j] = C(v);
    }
  }
  return F;
}

struct HostSector {
  std::vector<int> ue_row, uh_row, ue_rowmap;
  std::vector<double> ue_coef, uh_coef, ue_rowcoef;
};

// symmetryBasis.m:2-15: columns = even-even orbits then odd-odd orbits, each
// ordered by (n >= 0 outer, m >= 0 inner).
static HostSector build_sector(int Nx, int Ny, bool polP, int& nr) {
  const int W = 2 * Nx + 1, M = W * (2 * Ny + 1), nf = 2 * M;
  auto idx = [&](int m, int n) { return (m + Nx) + W * (n + Ny); };
  struct Col { int rows[4]; double coef[4]; int cnt; };
  std::vector<Col> ee, oo;
  for (int b = 0; b <= Ny; ++b)
    for (int a = 0; a <= Nx; ++a) {
      int pts[4][2] = {{a, b}, {-a, b}, {a, -b}, {-a, -b}};
      Col ce{{-1, -1, -1, -1}, {0, 0, 0, 0}, 0};
      std::vector<int> seen;
      for (auto& q : pts) {
        int id = idx(q[0], q[1]);
        if (std::find(seen.begin(), seen.end(), id) != seen.end()) continue;
        seen.push_back(id);
      }
      double nrm = 1.0 / std::sqrt((double)seen.size());
      for (size_t k = 0; k < seen.size(); ++k) { ce.rows[k] = seen[k]; ce.coef[k] = nrm; ce.cnt++; }
      ee.push_back(ce);
      if (a > 0 && b > 0) {
        Col co{{-1, -1, -1, -1}, {0, 0, 0, 0}, 0};
        for (int k = 0; k < 4; ++k) {
          co.rows[k] = idx(pts[k][0], pts[k][1]);
          co.coef[k] = 0.5 * ((pts[k][0] > 0) ? 1 : -1) * ((pts[k][1] > 0) ? 1 : -1);
          co.cnt++;
        }
        oo.push_back(co);
      }
    }
  nr = (int)(ee.size() + oo.size());
  HostSector s;
  s.ue_row.assign(4 * nr, -1); s.uh_row.assign(4 * nr, -1);
  s.ue_coef.assign(4 * nr, 0); s.uh_coef.assign(4 * nr, 0);
  s.ue_rowmap.assign(nf, -1); s.ue_rowcoef.assign(nf, 0);
  // offsets of the field blocks: E = [Ex; Ey], H = [Hx; Hy]
  // p: UE = [Ex ee | Ey oo], UH = [Hy ee | Hx oo];  s: UE = [Ey ee | Ex oo], UH = [Hx ee | Hy oo]
  const int eeE = polP ? 0 : M, ooE = polP ? M : 0;
  const int eeH = polP ? M : 0, ooH = polP ? 0 : M;
  int a = 0;
  auto put = [&](const Col& c, int offE, int offH) {
    for (int k = 0; k < c.cnt; ++k) {
      s.ue_row[4 * a + k] = offE + c.rows[k];
      s.ue_coef[4 * a + k] = c.coef[k];
      s.uh_row[4 * a + k] = offH + c.rows[k];
      s.uh_coef[4 * a + k] = c.coef[k];
      s.ue_rowmap[offE + c.rows[k]] = a;
      s.ue_rowcoef[offE + c.rows[k]] = c.coef[k];
    }
    ++a;
  };
  for (auto& c : ee) put(c, eeE, eeH);
  for (auto& c : oo) put(c, ooE, ooH);
  return s;
}

// y-invariant grating lit in the x-z plane (Ny = 0, every k_y = 0): P Q is
// block diagonal, blkdiag(-(I - Kx [eps]^-1 Kx) Exx, Kx^2 - Eyy), so E_x / H_y
// (TM, p) and E_y / H_x (TE, s) are solved apart, M unknowns each.  The same
// sparse-list form as the parity sectors, one entry per column:
// p: UE = [Ex], UH = [Hy];  s: UE = [Ey], UH = [Hx].
static HostSector build_sector_te_tm(int M, bool polP) {
  const int nf = 2 * M;
  HostSector s;
  s.ue_row.assign(4 * M, -1); s.uh_row.assign(4 * M, -1);
  s.ue_coef.assign(4 * M, 0); s.uh_coef.assign(4 * M, 0);
  s.ue_rowmap.assign(nf, -1); s.ue_rowcoef.assign(nf, 0);
  const int offE = polP ? 0 : M, offH = polP ? M : 0;
  for (int a = 0; a < M; ++a) {
    s.ue_row[4 * a] = offE + a;
    s.ue_coef[4 * a] = 1.0;
    s.uh_row[4 * a] = offH + a;
    s.uh_coef[4 * a] = 1.0;
    s.ue_rowmap[offE + a] = a;
    s.ue_rowcoef[offE + a] = 1.0;
  }
  return s;
}

// ---------------------------------------------------------------- workspace set
struct Work {
  // mode-page buffers (per chunk of Cp mode pages)
  cplx *E, *Exx, *Eyy, *G, *Omega, *Toe, *Inv, *GH, *HH;
  int* ipivT;
  // solve-page buffers (Cs = Cp * nSec)
  cplx *A0, *Wm, *Vm, *Zw, *Yw, *Hv, *mu, *lam, *Xd, *Wf, *QW, *T12;
  cplx *X1, *X2, *XB, *R3, *Dm, *RHS2, *LUb;
  cplx *Sa[4], *Sn[4], *Sb[4], *Tm, *M1, *Fm;
  double* kL;
  double* vnorm;  // column norms the eigensolver removed from W (W = Z Y diag(1 / vnorm))
  cplx* Vv;       // four n x 2 vectors per solve page (final assembly), page stride 8 n
  cplx* dinv;  // diagonal-block inverses of the most recent LU
  int *ipiv, *info;
  DevSolve* solves;
  int* groupOf;
  int* colIdx;
  cplx* epsSolve;
  // layer-parallel execution: mode tables of the chunk and page maps
  DevPage *pagesA, *pagesH;
  cplx *seA, *siA;
  int *idxA, *idxB;
};

// ====================================================================== plan
// Size limits: the eigensolver keeps a column of the solved matrix in 12
// registers per lane (n <= 768); the convolution-matrix LU keeps a 4-wide panel
// of M rows plus 4 rows of U in the 140 KiB of LDS it uses (M <= 1024).
static const int kMaxSolved = 768;
static const int kMaxHarmonics = 1024;
// P.reduced: 0 full basis, 1 mirror-parity sectors, 2 TE / TM of a y-invariant grating
static const int kReducedTeTm = 2;

struct orf_plan {
  orf_ctx* ctx = nullptr;
  hipStream_t st = nullptr;
  // problem
  double Lx = 0, Ly = 0, theta = 0, phi = 0;
  int Nx = 0, Ny = 0, M = 0, nf = 0, nr = 0, Nsx = 0, Nsy = 0;
  int B = 0, L = 0, nPol = 0, nLayers = 0, nSec = 1, reduced = 0;
  int factorization = 0, precision = 0;
  int nGrids = 0, nslotMax = 1, i0 = 0;
  // pages: (structure, context); a context is one (wavelength, period) pair
  int nPages = 0, nCtx = 0;
  std::vector<int> pageB, pageCtx;
  std::vector<double> lambda, k0;             // k0 per context
  std::vector<int> mIdx, nIdx;
  std::vector<std::vector<int>> layerGrid;     // [layer][b] -> grid or -1
  std::vector<std::vector<int>> layerMixed;    // unused placeholder
  std::vector<std::vector<zc>> layerBgEps;     // [layer][page]
  std::vector<int> layerAllPat, layerAnyPat;
  std::vector<double> thick;                   // [layer * B + b]
  // device-resident inputs
  DBuf dShapes, dVerts, dGrids, dLabels, dFx, dFy, dChiX, dChiY, dChi2, dKxy, dKz, dKzi0;
  std::vector<DBuf*> dPagesByLayer;            // [layer] DevPage[nPages]
  std::vector<DBuf*> dSlotEps, dSlotInv;       // [layer] cplx[nPages*nslot]
  DBuf dSectorData, dSectors, dSources;
  std::vector<DBuf*> secArrays;
  std::vector<DevSource> sources;
  std::vector<int> sectorSrcCol;               // reduced basis: source column per sector
  // shared per (context, sector) region data, nr x nr each
  DBuf dV0, dSref, dStrn;                      // V0 [nCtx*nSec], S [4][nCtx*nSec]
  DBuf dV0inv;                                 // V0^-1 [nCtx*nSec] (sector bases)
  // chunking
  int chunkPages = 0;
  // results (device)
  DBuf dR, dT, dRtot, dTtot, dr, dt;
  bool haveResult = false;
  // workspace of one chunk, allocated at creation and reused by every execute
  std::vector<DBuf*> ws;
  Work work{};
  bool ownStream = false;
  // modes and S-matrix of every layer of the single page, kept for field reconstruction
  struct Kept {
    DBuf W, V, lam, E, S11, S12;
    bool patterned = false;
  };
  bool keepModes = false;
  // one patterned layer: solved on its mode amplitudes (finish_direct)
  // without forming the layer S-matrix
  bool directSingle = false;
  std::vector<Kept*> kept;
  // layer-parallel execution (several layers): every layer's modes and S-matrix in one
  // batch, then the star products as a tree; host copies of the per-layer tables
  bool layerParallel = false;
  int workModes = 0;
  std::vector<std::vector<DevPage>> hPages;
  std::vector<std::vector<cplx>> hSe, hSi;
  double kzi0 = 1.0;
  zc epsInc0 = 1.0, epsSub0 = 1.0;  // half-space permittivities of context 0
  StageTimer timer;
  // Index and parameter tables of the execution (solve-page lists, context maps,
  // thicknesses, distinct-page maps, star-tree pairings): the same on every
  // execute, so each is uploaded once (on the first execute) and the k-th table
  // request of later executes returns the k-th device copy.  Execution then
  // issues only kernels and device-to-device copies, which is what a captured
  // graph replays.
  struct Table {
    DBuf* buf = nullptr;
    std::vector<unsigned char> host;
  };
  std::vector<Table> tables;
  size_t tableCursor = 0;
  // eigensolver status words of the current chunk (one region per eig stage),
  // checked after the chunk has run
  int infoCursor = 0, infoCap = 0;
  std::vector<std::pair<int, int>> pendingInfo;  // (offset, count)
  // captured chunks (orf_plan_enable_graphs): one graph per chunk
  bool useGraphs = false, capturing = false, graphsStale = false;
  std::vector<hipGraphExec_t> graphExec;
  std::vector<std::vector<std::pair<int, int>>> graphInfo;  // per chunk: its status regions
  std::vector<int> chunkCursor;                               // per chunk: its first table
  ~orf_plan() {
    if (st) (void)hipStreamSynchronize(st);  // no block returns to the pool while in use
    for (auto g : graphExec)
      if (g) (void)hipGraphExecDestroy(g);
    for (auto& t : tables) delete t.buf;
    for (auto* b : dPagesByLayer) delete b;
    for (auto* b : dSlotEps) delete b;
    for (auto* b : dSlotInv) delete b;
    for (auto* b : secArrays) delete b;
    for (auto* b : ws) delete b;
    for (auto* k : kept) delete k;
    if (ownStream && st) (void)hipStreamDestroy(st);
  }
};

struct Timed {
  StageTimer& t;
  hipStream_t s;
  Timed(StageTimer& t_, const char* n, hipStream_t s_) : t(t_), s(s_) { t.begin(n, s); }
  ~Timed() { t.end(s); }
};
#define STAGE(name) Timed _tm_##__LINE__(P.timer, name, P.st)


static size_t per_mode_bytes(const orf_plan& P) {
  const size_t M2 = (size_t)P.M * P.M;
  size_t b = 0;
  b += 3 * M2 + 2 * M2 + (size_t)P.nf * P.nf;  // E, Exx, Eyy, G, Omega
  if (P.factorization == ORF_FACT_LI) {
    const size_t Wx = 2 * P.Nx + 1, Wy = 2 * P.Ny + 1;
    b += 2 * std::max((size_t)P.Nsy * Wx * Wx, (size_t)P.Nsx * Wy * Wy);                 // Toe, Inv
    b += (size_t)(4 * P.Ny + 1) * Wx * Wx + (size_t)(4 * P.Nx + 1) * Wy * Wy;            // GH, HH
  }
  b *= sizeof(cplx);
  if (P.factorization == ORF_FACT_LI)
    b += sizeof(int) * std::max((size_t)P.Nsy * (2 * P.Nx + 1), (size_t)P.Nsx * (2 * P.Ny + 1));
  return b;
}
static size_t per_solve_bytes(const orf_plan& P) {
  const size_t n2 = (size_t)P.nr * P.nr, nfr = (size_t)P.nf * P.nr;
  size_t c = 0;
  c += n2 * 6 + P.nr;        // A0, Wm, Vm, Zw, Yw, Hv (+tau)
  c += 3 * nfr;              // Wf, QW, T12
  c += n2 * (2 + 1 + 3 + 1 + 2 + 1);  // X1 X2 XB R3 Dm RHS2 LUb
  c += n2 * (12 + 3);        // Sa Sn Sb Tm M1 Fm
  c += 3 * (size_t)P.nr + 1; // mu lam Xd eps
  c += (size_t)orf_dinv_stride(P.nr); // diagonal-block inverses
  c += 8 * (size_t)P.nr;              // Vv
  return c * sizeof(cplx) + sizeof(double) * (1 + (size_t)P.nr) + sizeof(int) * (P.nr + 2) + sizeof(DevSolve);
}

static Work make_work(orf_plan& P, int Cp) {
  for (auto* b : P.ws) delete b;
  P.ws.clear();
  auto get = [&](size_t bytes) -> void* {
    DBuf* b = new DBuf();
    P.ws.push_back(b);
    b->alloc(std::max<size_t>(bytes, 16));
    return b->p;
  };
  const int Cs = Cp * P.nSec;
  const size_t M2 = (size_t)P.M * P.M, n2 = (size_t)P.nr * P.nr, nfr = (size_t)P.nf * P.nr;
  const size_t cs = sizeof(cplx);
  Work w{};
  w.E = (cplx*)get(cs * M2 * Cp);
  w.Exx = (cplx*)get(cs * M2 * Cp);
  w.Eyy = (cplx*)get(cs * M2 * Cp);
  w.G = (cplx*)get(cs * 2 * M2 * Cp);
  w.Omega = (cplx*)get(cs * (size_t)P.nf * P.nf * Cp);
  if (P.factorization == ORF_FACT_LI) {
    const size_t Wx = 2 * P.Nx + 1, Wy = 2 * P.Ny + 1;
    w.Toe = (cplx*)get(cs * std::max((size_t)P.Nsy * Wx * Wx, (size_t)P.Nsx * Wy * Wy) * Cp);
    w.Inv = (cplx*)get(cs * std::max((size_t)P.Nsy * Wx * Wx, (size_t)P.Nsx * Wy * Wy) * Cp);
    w.GH = (cplx*)get(cs * (size_t)(4 * P.Ny + 1) * Wx * Wx * Cp);
    w.HH = (cplx*)get(cs * (size_t)(4 * P.Nx + 1) * Wy * Wy * Cp);
    w.ipivT = (int*)get(sizeof(int) * std::max((size_t)P.Nsy * Wx, (size_t)P.Nsx * Wy) * Cp);
  }
  w.A0 = (cplx*)get(cs * n2 * Cs);
  w.Wm = (cplx*)get(cs * n2 * Cs);
  w.Vm = (cplx*)get(cs * n2 * Cs);
  w.Zw = (cplx*)get(cs * n2 * Cs);
  w.Yw = (cplx*)get(cs * n2 * Cs);
  w.Hv = (cplx*)get(cs * (n2 + P.nr) * Cs);
  w.mu = (cplx*)get(cs * P.nr * Cs);
  w.lam = (cplx*)get(cs * P.nr * Cs);
  w.Xd = (cplx*)get(cs * P.nr * Cs);
  w.Wf = (cplx*)get(cs * nfr * Cs);
  w.QW = (cplx*)get(cs * nfr * Cs);
  w.T12 = (cplx*)get(cs * nfr * Cs);
  w.X1 = (cplx*)get(cs * n2 * Cs);
  w.X2 = (cplx*)get(cs * n2 * Cs);
  w.XB = (cplx*)get(cs * n2 * Cs);
  w.R3 = (cplx*)get(cs * 3 * n2 * Cs);
  w.Dm = (cplx*)get(cs * n2 * Cs);
  w.RHS2 = (cplx*)get(cs * 2 * n2 * Cs);
  w.LUb = (cplx*)get(cs * n2 * Cs);
  for (int k = 0; k < 4; ++k) {
    w.Sa[k] = (cplx*)get(cs * n2 * Cs);
    w.Sn[k] = (cplx*)get(cs * n2 * Cs);
    w.Sb[k] = (cplx*)get(cs * n2 * Cs);
  }
  w.Tm = (cplx*)get(cs * n2 * Cs);
  w.M1 = (cplx*)get(cs * n2 * Cs);
  w.Fm = (cplx*)get(cs * n2 * Cs);
  w.kL = (double*)get(sizeof(double) * Cs);
  w.vnorm = (double*)get(sizeof(double) * P.nr * Cs);
  w.Vv = (cplx*)get(cs * 8 * P.nr * Cs);
  w.dinv = (cplx*)get(cs * std::max((size_t)orf_dinv_stride(P.nr) * Cs, (size_t)orf_dinv_stride(P.M) * Cp));
  w.ipiv = (int*)get(sizeof(int) * (size_t)std::max(P.nr, P.M) * std::max(Cs, Cp));
  // one status region per eigensolver stage of a chunk (one per layer at most)
  P.infoCap = std::max(Cs, Cp) * std::max(64, P.nLayers + 1);
  w.info = (int*)get(sizeof(int) * (size_t)P.infoCap);
  w.solves = (DevSolve*)get(sizeof(DevSolve) * Cs);
  w.groupOf = (int*)get(sizeof(int) * Cs);
  w.colIdx = (int*)get(sizeof(int) * 2 * Cs);
  w.epsSolve = (cplx*)get(cs * Cs);
  w.pagesA = (DevPage*)get(sizeof(DevPage) * Cp);
  w.pagesH = (DevPage*)get(sizeof(DevPage) * Cp);
  w.seA = (cplx*)get(cs * (size_t)Cp * P.nslotMax);
  w.siA = (cplx*)get(cs * (size_t)Cp * P.nslotMax);
  w.idxA = (int*)get(sizeof(int) * Cs);
  w.idxB = (int*)get(sizeof(int) * Cs);
  return w;
}

// Device bytes held by a plan (resident tables, results and workspace).
static size_t plan_bytes(const orf_plan& P) {
  size_t b = 0;
  for (const DBuf* d : {&P.dShapes, &P.dVerts, &P.dGrids, &P.dLabels, &P.dFx, &P.dFy, &P.dChiX, &P.dChiY, &P.dChi2,
                        &P.dKxy, &P.dKz, &P.dKzi0, &P.dSectorData, &P.dSectors, &P.dSources, &P.dV0, &P.dV0inv, &P.dSref,
                        &P.dStrn, &P.dR, &P.dT, &P.dRtot, &P.dTtot, &P.dr, &P.dt})
    b += d->bytes;
  for (auto* v : {&P.dPagesByLayer, &P.dSlotEps, &P.dSlotInv, &P.secArrays, &P.ws})
    for (const DBuf* d : *v) b += d->bytes;
  return b;
}

// ---------------------------------------------------------------- dense S-matrix algebra
static const cplx ONE = {1.0, 0.0}, ZERO = {0.0, 0.0}, MONE = {-1.0, 0.0};

// Bulk products that the mixed-precision form runs on FP32 MFMA: the
// convolution-matrix products and the per-layer star products.  Everything
// else (eigensolve, propagators, every LU and solve, the last star product
// and the layer S-matrix) stays FP64.
static void bulk_gemm(const orf_plan& P, char opA, char opB, int M, int N, int K, cplx alpha, const cplx* A,
                      int lda, long long sA, const cplx* B, int ldb, long long sB, cplx beta, cplx* C, int ldc,
                      long long sC, int batch) {
  if (P.precision == ORF_PREC_MIXED || P.precision == ORF_PREC_SPLIT)
    orf_zgemm_f32(P.st, opA, opB, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch,
                  P.precision == ORF_PREC_SPLIT);
  else
    orf_zgemm(P.st, opA, opB, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, batch);
}

// Device copy of an execution table (see orf_plan::Table): uploaded on first use,
// reused while the host contents are unchanged.
template <class T>
static T* table(orf_plan& P, const std::vector<T>& v) {
  const size_t bytes = v.size() * sizeof(T);
  const size_t k = P.tableCursor++;
  if (k >= P.tables.size()) P.tables.emplace_back();
  orf_plan::Table& e = P.tables[k];
  if (e.buf && e.host.size() == bytes && (bytes == 0 || std::memcmp(e.host.data(), v.data(), bytes) == 0))
    return e.buf->as<T>();
  if (P.capturing) fail(ORF_DEVICE_ERROR, "execution tables changed while a graph was captured");
  if (e.buf) P.graphsStale = true;  // a captured graph may read the old copy
  if (!e.buf) e.buf = new DBuf();
  e.buf->alloc(std::max<size_t>(bytes, 16));
  e.host.assign((const unsigned char*)v.data(), (const unsigned char*)v.data() + bytes);
  if (bytes) HIPCHECK(hipMemcpyAsync(e.buf->p, e.host.data(), bytes, hipMemcpyHostToDevice, P.st));
  return e.buf->as<T>();
}

// Next status region of the chunk's eigensolver stages (checked by check_pending_info).
static int* next_info(orf_plan& P, Work& w, int count) {
  if (P.infoCursor + count > P.infoCap) fail(ORF_DEVICE_ERROR, "status buffer too small");
  int* p = w.info + P.infoCursor;
  P.pendingInfo.emplace_back(P.infoCursor, count);
  P.infoCursor += count;
  return p;
}

static void check_pending_info(orf_plan& P, Work& w) {
  for (auto& pi : P.pendingInfo) {
    std::vector<int> h(pi.second);
    HIPCHECK(hipMemcpyAsync(h.data(), w.info + pi.first, sizeof(int) * pi.second, hipMemcpyDeviceToHost, P.st));
    HIPCHECK(hipStreamSynchronize(P.st));
    for (int i = 0; i < pi.second; ++i)
      if (h[i] != 0)
        fail(ORF_DEVICE_ERROR, "eigensolver (QR iteration did not converge) failed on solve page " + std::to_string(i));
  }
  P.pendingInfo.clear();
  P.infoCursor = 0;
}

// Layer S-matrix (smatrixLayer.m:6-18): W, V, lam per solve page; W0 = I; V0 per
// solve page in X2 on entry.  Out: S11 -> S11o, S12 -> S12o.
// The first CsEig solve pages hold eigensolver output (W = Z Y diag(1 / vnorm)
// with Z, Y still in w.Zw, w.Yw): their W^-1 comes from one triangular solve.
// unsolved: stop at D (w.Dm) and [E1 | E2] (w.RHS2, page stride 2 n^2), with
// S11 = D^-1 E1 and S12 = D^-1 E2, for finish_single.
static void smatrix_layer(orf_plan& P, Work& w, int Cs, cplx* Wm, cplx* Vm, cplx* lam, cplx* S11o, cplx* S12o,
                          int CsEig, bool unsolved = false, const DevPage* eigPages = nullptr) {
  const int n = P.nr;
  const long long n2 = (long long)n * n;
  hipStream_t st = P.st;
  if (CsEig > 0) {
    STAGE("smat_solve");
    if (!orf_eig_inverse(st, n, w.Zw, w.Yw, w.vnorm, w.X1, w.dinv, CsEig)) CsEig = 0;
  }
  // Leading pages whose V \ V0 comes from the eigensolver: with V = Q W lam^-1 and
  // P Q = W lam^2 W^-1, V^-1 = lam^-1 W^-1 P, so V \ V0 = lam^-1 W^-1 (P V0) — one
  // product instead of an LU and an n-column solve.  Full basis only ([eps]^-1 is
  // kept in Omega's pages by the mode stage; eigPages are the mode pages of those solves).
  const int CsV = (eigPages && !P.reduced) ? CsEig : 0;
  if (CsV > 0) {
    STAGE("smat_gemm");
    const long long M2 = (long long)P.M * P.M;
    const long long nf2 = (long long)P.nf * P.nf;
    pv0_kernel<<<g1(n2, CsV), 256, 0, st>>>(P.M, eigPages, P.dKxy.as<double>(), w.Omega, nf2, w.X2, n2, w.R3, n2);
    orf_zgemm(st, 'N', 'N', n, n, n, ONE, w.X1, n, n2, w.R3, n, n2, ZERO, w.X2, n, n2, CsV);  // lam scaling below
    KCHECK();
  }
  if (CsEig < Cs) {  // W \ W0 by LU for the rest
    const int Cr = Cs - CsEig;
    cplx* Wr = Wm + (size_t)CsEig * n2;
    int* ipr = w.ipiv + (size_t)CsEig * n;
    cplx* dr = w.dinv + (size_t)CsEig * orf_dinv_stride(n);
    cplx* Xr = w.X1 + (size_t)CsEig * n2;
    {
      STAGE("smat_lu");
      orf_zgetrf(st, n, Wr, n, n2, ipr, nullptr, Cr, dr);
      KCHECK();
    }
    eye_axpby_kernel<<<g1(n2, Cr), 256, 0, st>>>(n, ONE, ZERO, nullptr, 0, Xr, n2);
    {
      STAGE("smat_solve");
      orf_zgetrs(st, 'L', n, n, Wr, n, n2, ipr, dr, Xr, n, n2, Cr);
    }
  }
  if (CsV < Cs) {  // V \ V0 by LU for the rest
    const int Cr = Cs - CsV;
    cplx* Vr = Vm + (size_t)CsV * n2;
    int* ipr = w.ipiv + (size_t)CsV * n;
    cplx* dr = w.dinv + (size_t)CsV * orf_dinv_stride(n);
    {
      STAGE("smat_lu");
      orf_zgetrf(st, n, Vr, n, n2, ipr, nullptr, Cr, dr);
    }
    {
      STAGE("smat_solve");
      orf_zgetrs(st, 'L', n, n, Vr, n, n2, ipr, dr, w.X2 + (size_t)CsV * n2, n, n2, Cr);
    }
  }
  // X1 = A, X2 = B, LUb = A, R3 = [XB XA B]
  smat_abx_kernel<<<g1(n2, Cs), 256, 0, st>>>(n, lam, w.kL, w.X1, w.X2, n2, CsV, w.LUb, w.R3, 3 * n2, w.Xd);
  cplx* XBm = w.R3;
  cplx* XAm = w.R3 + n2;
  cplx* F = w.R3 + 2 * n2;
  // F = B A^-1 (one right solve; X B A^-1 X A = X F X A and B A^-1 B = F B)
  {
    STAGE("smat_lu");
    orf_zgetrf(st, n, w.LUb, n, n2, w.ipiv, nullptr, Cs, w.dinv);
  }
  {
    STAGE("smat_solve");
    orf_zgetrs(st, 'R', n, n, w.LUb, n, n2, w.ipiv, w.dinv, F, n, 3 * n2, Cs);
  }
  {
    STAGE("smat_gemm");
    // D = A - X (F X B)
    orf_zgemm(st, 'N', 'N', n, n, n, ONE, F, n, 3 * n2, XBm, n, 3 * n2, ZERO, w.XB, n, n2, Cs);
    rowscale_add_kernel<<<g1(n2, Cs), 256, 0, st>>>(n, w.Xd, w.XB, n2, -1.0, w.X1, n2, 1.0, w.Dm, n2);
    // RHS2 = [X F X A - B, X (A - F B)]
    orf_zgemm(st, 'N', 'N', n, n, n, ONE, F, n, 3 * n2, XAm, n, 3 * n2, ZERO, w.RHS2, n, 2 * n2, Cs);
    rowscale_add_kernel<<<g1(n2, Cs), 256, 0, st>>>(n, w.Xd, w.RHS2, 2 * n2, 1.0, w.X2, n2, -1.0, w.RHS2, 2 * n2);
    eye_axpby_kernel<<<g1(n2, Cs), 256, 0, st>>>(n, ZERO, ONE, w.X1, n2, w.RHS2 + n2, 2 * n2);
    orf_zgemm(st, 'N', 'N', n, n, n, MONE, F, n, 3 * n2, w.X2, n, n2, ONE, w.RHS2 + n2, n, 2 * n2, Cs);
  }
  rowscale_kernel<<<g1(n2, Cs), 256, 0, st>>>(n, n, w.Xd, w.RHS2 + n2, 2 * n2);
  if (unsolved) return;
  {
    STAGE("smat_lu");
    orf_zgetrf(st, n, w.Dm, n, n2, w.ipiv, nullptr, Cs, w.dinv);
  }
  {
    STAGE("smat_solve");
    orf_zgetrs(st, 'L', n, 2 * n, w.Dm, n, n2, w.ipiv, w.dinv, w.RHS2, n, 2 * n2, Cs);
  }
  // split [S11 | S12] per page into separate page arrays
  for (int k = 0; k < 2; ++k) {
    cplx* dst = k == 0 ? S11o : S12o;
    HIPCHECK(hipMemcpy2DAsync(dst, sizeof(cplx) * n2, w.RHS2 + k * n2, sizeof(cplx) * 2 * n2, sizeof(cplx) * n2,
                              Cs, hipMemcpyDeviceToDevice, st));
  }
}

// Redheffer star (smatrixStar.m:4-10): SA (4 page arrays) * SB -> out (4 page arrays).
// One LU serves both inverses: with M = I - SB11 SA22, (I - SA22 SB11)^-1 =
// I + SA22 M^-1 SB11, so D = SA12 M^-1 and F = SB21 + SB21 SA22 (M^-1 SB11).
static void smatrix_star(orf_plan& P, Work& w, int Cs, cplx* const SA[4], cplx* const SB[4], cplx* const out[4]) {
  const int n = P.nr;
  const long long n2 = (long long)n * n;
  hipStream_t st = P.st;
  STAGE("star");
  eye_axpby_kernel<<<g1(n2, Cs), 256, 0, st>>>(n, ONE, ZERO, nullptr, 0, w.M1, n2);
  bulk_gemm(P, 'N', 'N', n, n, n, MONE, SB[0], n, n2, SA[3], n, n2, ONE, w.M1, n, n2, Cs);
  orf_zgetrf(st, n, w.M1, n, n2, w.ipiv, nullptr, Cs, w.dinv);
  // D = SA12 M^-1
  HIPCHECK(hipMemcpyAsync(w.Dm, SA[1], sizeof(cplx) * n2 * Cs, hipMemcpyDeviceToDevice, st));
  orf_zgetrs(st, 'R', n, n, w.M1, n, n2, w.ipiv, w.dinv, w.Dm, n, n2, Cs);
  // G = M^-1 SB11;  F = SB21 + (SB21 SA22) G
  HIPCHECK(hipMemcpyAsync(w.Tm, SB[0], sizeof(cplx) * n2 * Cs, hipMemcpyDeviceToDevice, st));
  orf_zgetrs(st, 'L', n, n, w.M1, n, n2, w.ipiv, w.dinv, w.Tm, n, n2, Cs);
  bulk_gemm(P, 'N', 'N', n, n, n, ONE, SB[2], n, n2, SA[3], n, n2, ZERO, w.M1, n, n2, Cs);
  HIPCHECK(hipMemcpyAsync(w.Fm, SB[2], sizeof(cplx) * n2 * Cs, hipMemcpyDeviceToDevice, st));
  bulk_gemm(P, 'N', 'N', n, n, n, ONE, w.M1, n, n2, w.Tm, n, n2, ONE, w.Fm, n, n2, Cs);
  // S11 = SA11 + D SB11 SA21
  bulk_gemm(P, 'N', 'N', n, n, n, ONE, w.Dm, n, n2, SB[0], n, n2, ZERO, w.Tm, n, n2, Cs);
  HIPCHECK(hipMemcpyAsync(out[0], SA[0], sizeof(cplx) * n2 * Cs, hipMemcpyDeviceToDevice, st));
  bulk_gemm(P, 'N', 'N', n, n, n, ONE, w.Tm, n, n2, SA[2], n, n2, ONE, out[0], n, n2, Cs);
  // S12 = D SB12
  bulk_gemm(P, 'N', 'N', n, n, n, ONE, w.Dm, n, n2, SB[1], n, n2, ZERO, out[1], n, n2, Cs);
  // S21 = F SA21
  bulk_gemm(P, 'N', 'N', n, n, n, ONE, w.Fm, n, n2, SA[2], n, n2, ZERO, out[2], n, n2, Cs);
  // S22 = SB22 + F SA22 SB12
  bulk_gemm(P, 'N', 'N', n, n, n, ONE, w.Fm, n, n2, SA[3], n, n2, ZERO, w.Tm, n, n2, Cs);
  HIPCHECK(hipMemcpyAsync(out[3], SB[3], sizeof(cplx) * n2 * Cs, hipMemcpyDeviceToDevice, st));
  bulk_gemm(P, 'N', 'N', n, n, n, ONE, w.Tm, n, n2, SB[1], n, n2, ONE, out[3], n, n2, Cs);
}

// Last star product S = SA * S_trn restricted to the source columns (rcwa2D
// needs r = S11 c and t = S21 c only, smatrixStar.m:5-10 applied to vectors).
// With M = I - SB11 SA22, U = SA21 C and X = M^-1 SB11 U:
// R = SA11 C + SA12 X,  T = SB21 (I - SA22 SB11)^-1 U = SB21 (U + SA22 X).
static void smatrix_star_cols(orf_plan& P, Work& w, int Cs, cplx* const SA[4], cplx* const SB[4], int ncol,
                              const int* colIdx, cplx* R, cplx* T) {
  const int n = P.nr;
  const long long n2 = (long long)n * n, nc2 = (long long)n * ncol;
  hipStream_t st = P.st;
  STAGE("star_final");
  cplx* U = w.Tm;   // SA21 C, then U + SA22 X
  cplx* Xv = w.Fm;  // SB11 U, then X = M^-1 SB11 U
  select_cols_kernel<<<g1(nc2, Cs), 256, 0, st>>>(n, ncol, colIdx, SA[2], n2, U, nc2);
  select_cols_kernel<<<g1(nc2, Cs), 256, 0, st>>>(n, ncol, colIdx, SA[0], n2, R, nc2);
  orf_zgemm(st, 'N', 'N', n, ncol, n, ONE, SB[0], n, n2, U, n, nc2, ZERO, Xv, n, nc2, Cs);
  eye_axpby_kernel<<<g1(n2, Cs), 256, 0, st>>>(n, ONE, ZERO, nullptr, 0, w.M1, n2);
  orf_zgemm(st, 'N', 'N', n, n, n, MONE, SB[0], n, n2, SA[3], n, n2, ONE, w.M1, n, n2, Cs);
  orf_zgetrf(st, n, w.M1, n, n2, w.ipiv, nullptr, Cs, w.dinv);
  orf_zgetrs(st, 'L', n, ncol, w.M1, n, n2, w.ipiv, w.dinv, Xv, n, nc2, Cs);
  orf_zgemm(st, 'N', 'N', n, ncol, n, ONE, SA[1], n, n2, Xv, n, nc2, ONE, R, n, nc2, Cs);
  orf_zgemm(st, 'N', 'N', n, ncol, n, ONE, SA[3], n, n2, Xv, n, nc2, ONE, U, n, nc2, Cs);
  orf_zgemm(st, 'N', 'N', n, ncol, n, ONE, SB[2], n, n2, U, n, nc2, ZERO, T, n, nc2, Cs);
  KCHECK();
}

// Source-column result of S_ref * L * S_trn, with L the composite S-matrix of
// the layer stack (rcwa2D.m:31-35; smatrixStar.m:4-10 applied to vectors
// instead of forming S_ref * L).  With a = S_ref, b = S_trn and the gap
// amplitudes below the incident region c1+ (down) and c1- (up):
//   K = (I - L22 b11)^-1 L21,   G = L11 + L12 b11 K   (c1- = G c1+),
//   (I - a22 G) c1+ = a21 C,   R = a11 C + a12 G c1+,   T = b21 K c1+.
// Two LUs and one n-column solve instead of a full star product (one LU, two
// full solves, eight products) followed by the column star.
static void finish_stack(orf_plan& P, Work& w, int Cs, cplx* const a[4], cplx* const L[4], const cplx* b11,
                         const cplx* b21, int ncol, const int* colIdx, cplx* R, cplx* T) {
  const int n = P.nr;
  const long long n2 = (long long)n * n, nc2 = (long long)n * ncol;
  hipStream_t st = P.st;
  STAGE("star_final");
  cplx* Nm = w.M1;   // I - L22 b11, then I - a22 G
  cplx* Km = w.Tm;   // K
  cplx* Pm = w.Fm;   // L12 b11
  cplx* Gm = w.Dm;   // G
  const long long sv = 8LL * n;  // vector page stride in Vv
  cplx* U = w.Vv;             // a21 C, then c1+
  cplx* gv = w.Vv + 2 * n;    // G c1+
  cplx* kv = w.Vv + 4 * n;    // K c1+
  // in the full basis the half-space blocks pair each harmonic with its other
  // polarisation only: their products are O(n^2) (xr_pair / rx_pair)
  const bool pair = !P.reduced;
  eye_axpby_kernel<<<g1(n2, Cs), 256, 0, st>>>(n, ONE, ZERO, nullptr, 0, Nm, n2);
  if (pair) xr_pair_kernel<<<g1(n2, Cs), 256, 0, st>>>(P.M, L[3], n2, b11, n2, MONE, ONE, Nm, n2, Nm, n2);
  else bulk_gemm(P, 'N', 'N', n, n, n, MONE, L[3], n, n2, b11, n, n2, ONE, Nm, n, n2, Cs);
  orf_zgetrf(st, n, Nm, n, n2, w.ipiv, nullptr, Cs, w.dinv);
  HIPCHECK(hipMemcpyAsync(Km, L[2], sizeof(cplx) * n2 * Cs, hipMemcpyDeviceToDevice, st));
  orf_zgetrs(st, 'L', n, n, Nm, n, n2, w.ipiv, w.dinv, Km, n, n2, Cs);
  if (pair) xr_pair_kernel<<<g1(n2, Cs), 256, 0, st>>>(P.M, L[1], n2, b11, n2, ONE, ZERO, nullptr, 0, Pm, n2);
  else bulk_gemm(P, 'N', 'N', n, n, n, ONE, L[1], n, n2, b11, n, n2, ZERO, Pm, n, n2, Cs);
  HIPCHECK(hipMemcpyAsync(Gm, L[0], sizeof(cplx) * n2 * Cs, hipMemcpyDeviceToDevice, st));
  bulk_gemm(P, 'N', 'N', n, n, n, ONE, Pm, n, n2, Km, n, n2, ONE, Gm, n, n2, Cs);
  eye_axpby_kernel<<<g1(n2, Cs), 256, 0, st>>>(n, ONE, ZERO, nullptr, 0, Nm, n2);
  if (pair) rx_pair_kernel<<<g1(n2, Cs), 256, 0, st>>>(P.M, n, a[3], n2, Gm, n2, MONE, ONE, Nm, n2, Nm, n2);
  else bulk_gemm(P, 'N', 'N', n, n, n, MONE, a[3], n, n2, Gm, n, n2, ONE, Nm, n, n2, Cs);
  orf_zgetrf(st, n, Nm, n, n2, w.ipiv, nullptr, Cs, w.dinv);
  select_cols_kernel<<<g1(nc2, Cs), 256, 0, st>>>(n, ncol, colIdx, a[2], n2, U, sv);
  select_cols_kernel<<<g1(nc2, Cs), 256, 0, st>>>(n, ncol, colIdx, a[0], n2, R, nc2);
  orf_zgetrs(st, 'L', n, ncol, Nm, n, n2, w.ipiv, w.dinv, U, n, sv, Cs);
  orf_zgemm(st, 'N', 'N', n, ncol, n, ONE, Gm, n, n2, U, n, sv, ZERO, gv, n, sv, Cs);
  orf_zgemm(st, 'N', 'N', n, ncol, n, ONE, Km, n, n2, U, n, sv, ZERO, kv, n, sv, Cs);
  if (pair) {
    rx_pair_kernel<<<g1(nc2, Cs), 256, 0, st>>>(P.M, ncol, a[1], n2, gv, sv, ONE, ONE, R, nc2, R, nc2);
    rx_pair_kernel<<<g1(nc2, Cs), 256, 0, st>>>(P.M, ncol, b21, n2, kv, sv, ONE, ZERO, nullptr, 0, T, nc2);
  } else {
    orf_zgemm(st, 'N', 'N', n, ncol, n, ONE, a[1], n, n2, gv, n, sv, ONE, R, n, nc2, Cs);
    orf_zgemm(st, 'N', 'N', n, ncol, n, ONE, b21, n, n2, kv, n, sv, ZERO, T, n, nc2, Cs);
  }
  KCHECK();
}

// Source-column result for a single layer whose S-matrix is left unsolved,
// S11 = D^-1 E1, S12 = D^-1 E2 (S22 = S11, S21 = S12): substituting into
// finish_stack's relations removes D^-1 (one LU and a 2n-column solve):
//   K = (D - E1 b11)^-1 E2,   H = E1 + E2 b11 K   (G = D^-1 H),
//   (D - H a22) g = H a21 C,  c1+ = a21 C + a22 g,
//   R = a11 C + a12 g,        T = b21 K c1+.
static void finish_single(orf_plan& P, Work& w, int Cs, cplx* const a[4], const cplx* b11, const cplx* b21,
                          int ncol, const int* colIdx, cplx* R, cplx* T) {
  const int n = P.nr;
  const long long n2 = (long long)n * n, nc2 = (long long)n * ncol;
  hipStream_t st = P.st;
  STAGE("star_final");
  const cplx* Dm = w.Dm;
  cplx* E1 = w.RHS2;       // E1, then H in place
  cplx* E2 = w.RHS2 + n2;  // E2, then K in place (E1, E2 page stride 2 n2)
  const long long s2 = 2 * n2;
  cplx* Nm = w.M1;   // D - E1 b11, then D - H a22
  cplx* Pm = w.Fm;   // E2 b11
  const long long sv = 8LL * n;  // vector page stride in Vv
  cplx* u0 = w.Vv;            // a21 C
  cplx* gv = w.Vv + 2 * n;    // H a21 C, then g
  cplx* c1 = w.Vv + 4 * n;    // c1+
  cplx* kv = w.Vv + 6 * n;    // K c1+
  const bool pair = !P.reduced;
  if (pair) {
    xr_pair_kernel<<<g1(n2, Cs), 256, 0, st>>>(P.M, E1, s2, b11, n2, MONE, ONE, Dm, n2, Nm, n2);
  } else {
    HIPCHECK(hipMemcpyAsync(Nm, Dm, sizeof(cplx) * n2 * Cs, hipMemcpyDeviceToDevice, st));
    bulk_gemm(P, 'N', 'N', n, n, n, MONE, E1, n, s2, b11, n, n2, ONE, Nm, n, n2, Cs);
  }
  orf_zgetrf(st, n, Nm, n, n2, w.ipiv, nullptr, Cs, w.dinv);
  // E2 b11 before E2 turns into K
  if (pair) xr_pair_kernel<<<g1(n2, Cs), 256, 0, st>>>(P.M, E2, s2, b11, n2, ONE, ZERO, nullptr, 0, Pm, n2);
  else bulk_gemm(P, 'N', 'N', n, n, n, ONE, E2, n, s2, b11, n, n2, ZERO, Pm, n, n2, Cs);
  cplx* Km = E2;
  orf_zgetrs(st, 'L', n, n, Nm, n, n2, w.ipiv, w.dinv, Km, n, s2, Cs);
  cplx* Hm = E1;
  bulk_gemm(P, 'N', 'N', n, n, n, ONE, Pm, n, n2, Km, n, s2, ONE, Hm, n, s2, Cs);
  if (pair) {
    xr_pair_kernel<<<g1(n2, Cs), 256, 0, st>>>(P.M, Hm, s2, a[3], n2, MONE, ONE, Dm, n2, Nm, n2);
  } else {
    HIPCHECK(hipMemcpyAsync(Nm, Dm, sizeof(cplx) * n2 * Cs, hipMemcpyDeviceToDevice, st));
    bulk_gemm(P, 'N', 'N', n, n, n, MONE, Hm, n, s2, a[3], n, n2, ONE, Nm, n, n2, Cs);
  }
  orf_zgetrf(st, n, Nm, n, n2, w.ipiv, nullptr, Cs, w.dinv);
  select_cols_kernel<<<g1(nc2, Cs), 256, 0, st>>>(n, ncol, colIdx, a[2], n2, u0, sv);
  select_cols_kernel<<<g1(nc2, Cs), 256, 0, st>>>(n, ncol, colIdx, a[0], n2, R, nc2);
  orf_zgemm(st, 'N', 'N', n, ncol, n, ONE, Hm, n, s2, u0, n, sv, ZERO, gv, n, sv, Cs);
  orf_zgetrs(st, 'L', n, ncol, Nm, n, n2, w.ipiv, w.dinv, gv, n, sv, Cs);
  if (pair) {  // products with half-space blocks: two terms per entry
    rx_pair_kernel<<<g1(nc2, Cs), 256, 0, st>>>(P.M, ncol, a[3], n2, gv, sv, ONE, ONE, u0, sv, c1, sv);
    rx_pair_kernel<<<g1(nc2, Cs), 256, 0, st>>>(P.M, ncol, a[1], n2, gv, sv, ONE, ONE, R, nc2, R, nc2);
    orf_zgemm(st, 'N', 'N', n, ncol, n, ONE, Km, n, s2, c1, n, sv, ZERO, kv, n, sv, Cs);
    rx_pair_kernel<<<g1(nc2, Cs), 256, 0, st>>>(P.M, ncol, b21, n2, kv, sv, ONE, ZERO, nullptr, 0, T, nc2);
  } else {
    HIPCHECK(hipMemcpy2DAsync(c1, sizeof(cplx) * sv, u0, sizeof(cplx) * sv, sizeof(cplx) * nc2, Cs,
                              hipMemcpyDeviceToDevice, st));
    orf_zgemm(st, 'N', 'N', n, ncol, n, ONE, a[3], n, n2, gv, n, sv, ONE, c1, n, sv, Cs);
    orf_zgemm(st, 'N', 'N', n, ncol, n, ONE, a[1], n, n2, gv, n, sv, ONE, R, n, nc2, Cs);
    orf_zgemm(st, 'N', 'N', n, ncol, n, ONE, Km, n, s2, c1, n, sv, ZERO, kv, n, sv, Cs);
    orf_zgemm(st, 'N', 'N', n, ncol, n, ONE, b21, n, n2, kv, n, sv, ZERO, T, n, nc2, Cs);
  }
  KCHECK();
}

// One patterned layer: R and T on the source columns straight from the layer's
// modes (W in w.Wm, lam, kL; full basis: T12 = [Eyy W2; Exx W1] in w.T12 and the
// per-context V0 and half-space tables; sector bases: V in w.Vm, V0^-1 in w.X2
// and the half-space blocks gathered per solve in SA, SB), without the layer
// S-matrix (smatrixLayer.m:6-18 and the two star products of rcwa2D.m give the
// same R, T).  direct_blocks_kernel (full basis) or dense_blocks_kernel (sector
// bases) states the system [A1 B1; B2 A2] [h+; h-] = [a21 C; 0]; its Schur
// complement on h+:
//   K = A2^-1 B2,  (A1 - B1 K) h+ = a21 C,  h- = -K h+,
// then c1- = W S1 + U D1, c2+ = W S2 + U D2 (direct_vec_kernel), R = a11 C + a12 c1-,
// T = b21 c2+.  Two n x n LUs, one n-column solve and one product.
static void finish_direct(orf_plan& P, Work& w, int Cs, const DevPage* pages, cplx* const SA[4], cplx* const SB[4],
                          int ncol, const int* colIdx, cplx* R, cplx* T) {
  const int n = P.nr, M = P.M;
  const long long n2 = (long long)n * n, nc2 = (long long)n * ncol, G = (long long)P.nCtx * P.nSec;
  // half-space blocks and the gap's V0 per context (page w.groupOf[sp])
  const cplx* a = P.dSref.as<cplx>();
  const cplx* b = P.dStrn.as<cplx>();
  const cplx *a11 = a, *a12 = a + n2 * G, *a21 = a + 2 * n2 * G, *a22 = a + 3 * n2 * G;
  const cplx *b11 = b, *b21 = b + 2 * n2 * G;
  hipStream_t st = P.st;
  STAGE("star_final");
  cplx* U = w.XB;
  cplx* A1 = w.Dm;   // then the Schur complement A1 - B1 K and its LU
  cplx* B1 = w.LUb;
  cplx* B2 = w.Fm;   // then K
  cplx* A2 = w.M1;   // then its LU
  if (!P.reduced) {
    const int nb = (n + 7) / 8;  // kDirectCols columns per thread
    direct_blocks_kernel<<<g1((long long)M * nb, Cs), 256, 0, st>>>(
        M, w.solves, pages, P.dKxy.as<double>(), w.Wm, w.T12, w.lam, w.kL, w.groupOf, P.dV0.as<cplx>(), a22, b11, n2,
        U, A1, B1, B2, A2);
  } else {
    // sector bases: V in w.Vm, V0^-1 per solve in w.X2, dense half-space blocks
    // (gathered per solve in SA, SB): U = V0^-1 V, [W+ W-] in RHS2, a22 [W+ W-]
    // in R3, b11 W+ and b11 W- in Sa[0], Sa[1]
    orf_zgemm(st, 'N', 'N', n, n, n, ONE, w.X2, n, n2, w.Vm, n, n2, ZERO, U, n, n2, Cs);
    wpm_kernel<<<g1(n2, Cs), 256, 0, st>>>(n, w.Wm, U, n2, w.RHS2, 2 * n2);
    orf_zgemm(st, 'N', 'N', n, 2 * n, n, ONE, SA[3], n, n2, w.RHS2, n, 2 * n2, ZERO, w.R3, n, 3 * n2, Cs);
    orf_zgemm(st, 'N', 'N', n, n, n, ONE, SB[0], n, n2, w.RHS2, n, 2 * n2, ZERO, w.Sa[0], n, n2, Cs);
    orf_zgemm(st, 'N', 'N', n, n, n, ONE, SB[0], n, n2, w.RHS2 + n2, n, 2 * n2, ZERO, w.Sa[1], n, n2, Cs);
    dense_blocks_kernel<<<g1(n2, Cs), 256, 0, st>>>(n, w.lam, w.kL, w.RHS2, w.R3, w.Sa[0], w.Sa[1], n2, A1, B1, B2,
                                                     A2);
  }
  KCHECK();
  orf_zgetrf(st, n, A2, n, n2, w.ipiv, nullptr, Cs, w.dinv);
  orf_zgetrs(st, 'L', n, n, A2, n, n2, w.ipiv, w.dinv, B2, n, n2, Cs);
  orf_zgemm(st, 'N', 'N', n, n, n, MONE, B1, n, n2, B2, n, n2, ONE, A1, n, n2, Cs);
  orf_zgetrf(st, n, A1, n, n2, w.ipiv, nullptr, Cs, w.dinv);
  // vectors in w.Tm (page stride n2 >= 16 n): h+, h-, S (2 ncol), D (2 ncol), Y (2 ncol)
  cplx* hp = w.Tm;
  cplx* hm = w.Tm + nc2;
  cplx* Sv = w.Tm + 2 * nc2;
  cplx* Dv = w.Tm + 4 * nc2;
  cplx* Yv = w.Tm + 6 * nc2;
  if (!P.reduced) direct_rhs_kernel<<<g1(nc2, Cs), 256, 0, st>>>(n, ncol, colIdx, w.groupOf, a21, hp, n2);
  else select_cols_kernel<<<g1(nc2, Cs), 256, 0, st>>>(n, ncol, colIdx, SA[2], n2, hp, n2);
  orf_zgetrs(st, 'L', n, ncol, A1, n, n2, w.ipiv, w.dinv, hp, n, n2, Cs);
  orf_zgemm(st, 'N', 'N', n, ncol, n, MONE, B2, n, n2, hp, n, n2, ZERO, hm, n, n2, Cs);
  direct_vec_kernel<<<g1(nc2, Cs), 256, 0, st>>>(n, ncol, w.lam, w.kL, hp, hm, Sv, Dv, n2);
  orf_zgemm(st, 'N', 'N', n, 2 * ncol, n, ONE, w.Wm, n, n2, Sv, n, n2, ZERO, Yv, n, n2, Cs);
  orf_zgemm(st, 'N', 'N', n, 2 * ncol, n, ONE, U, n, n2, Dv, n, n2, ONE, Yv, n, n2, Cs);
  if (!P.reduced) {
    direct_out_kernel<<<g1(nc2, Cs), 256, 0, st>>>(M, ncol, colIdx, w.groupOf, a11, a12, b21, Yv, n2, R, T);
  } else {
    select_cols_kernel<<<g1(nc2, Cs), 256, 0, st>>>(n, ncol, colIdx, SA[0], n2, R, nc2);
    orf_zgemm(st, 'N', 'N', n, ncol, n, ONE, SA[1], n, n2, Yv, n, n2, ONE, R, n, nc2, Cs);
    orf_zgemm(st, 'N', 'N', n, ncol, n, ONE, SB[2], n, n2, Yv + nc2, n, n2, ZERO, T, n, nc2, Cs);
  }
  KCHECK();
}

// ---------------------------------------------------------------- region S-matrices
// smatrixRegion.m:8-20 with W = W0 = I for every (context, sector) group.
static void build_regions(orf_plan& P, const std::vector<zc>& epsInc, const std::vector<zc>& epsSub,
                          const std::vector<zc>& epsGap) {
  const int n = P.nr, G = P.nCtx * P.nSec;
  const long long n2 = (long long)n * n;
  hipStream_t st = P.st;
  // region pages: one DevPage per context, DevSolve per group
  std::vector<DevPage> pg(P.nCtx);
  for (int c = 0; c < P.nCtx; ++c) pg[c] = DevPage{0, c, -1, 0};
  std::vector<DevSolve> sv(3 * G);
  std::vector<cplx> eps(3 * G);
  for (int k = 0; k < 3; ++k)
    for (int c = 0; c < P.nCtx; ++c)
      for (int s = 0; s < P.nSec; ++s) {
        int g = k * G + c * P.nSec + s;
        sv[g] = DevSolve{c, s};
        eps[g] = C(k == 0 ? epsGap[c] : (k == 1 ? epsInc[c] : epsSub[c]));
      }
  DBuf dpg, dsv, deps, dW, dV, dlam, dX, dLU, dA, dB, dAi, dT, dip;
  dpg.upload(pg);
  dsv.upload(sv);
  deps.upload(eps);
  dW.alloc(sizeof(cplx) * n2 * 3 * G);
  dV.alloc(sizeof(cplx) * n2 * 3 * G);
  dlam.alloc(sizeof(cplx) * n * 3 * G);
  homog_modes_kernel<<<g1(n2, 3 * G), 256, 0, st>>>(P.M, n, P.reduced, dsv.as<DevSolve>(), dpg.as<DevPage>(),
                                                    P.dSectors.as<DevSector>(), P.dKxy.as<double>(),
                                                    deps.as<cplx>(), dW.as<cplx>(), dV.as<cplx>(), n2,
                                                    dlam.as<cplx>());
  KCHECK();
  P.dV0.alloc(sizeof(cplx) * n2 * G);
  HIPCHECK(hipMemcpyAsync(P.dV0.p, dV.p, sizeof(cplx) * n2 * G, hipMemcpyDeviceToDevice, st));
  // X = V0 \ V for inc and sub (2G pages); V0 broadcast per group
  dLU.alloc(sizeof(cplx) * n2 * 2 * G);
  for (int k = 0; k < 2; ++k)
    HIPCHECK(hipMemcpyAsync(dLU.as<cplx>() + k * n2 * G, dV.p, sizeof(cplx) * n2 * G, hipMemcpyDeviceToDevice, st));
  dip.alloc(sizeof(int) * n * 2 * G);
  DBuf dDinv;
  dDinv.alloc(sizeof(cplx) * orf_dinv_stride(n) * 2 * G);
  orf_zgetrf(st, n, dLU.as<cplx>(), n, n2, dip.as<int>(), nullptr, 2 * G, dDinv.as<cplx>());
  if (P.reduced) {  // V0^-1 per group for the direct single-layer solve (finish_direct)
    P.dV0inv.alloc(sizeof(cplx) * n2 * G);
    eye_axpby_kernel<<<g1(n2, G), 256, 0, st>>>(n, ONE, ZERO, nullptr, 0, P.dV0inv.as<cplx>(), n2);
    orf_zgetrs(st, 'L', n, n, dLU.as<cplx>(), n, n2, dip.as<int>(), dDinv.as<cplx>(), P.dV0inv.as<cplx>(), n, n2, G);
  }
  dX.alloc(sizeof(cplx) * n2 * 2 * G);
  HIPCHECK(hipMemcpyAsync(dX.p, dV.as<cplx>() + n2 * G, sizeof(cplx) * n2 * 2 * G, hipMemcpyDeviceToDevice, st));
  orf_zgetrs(st, 'L', n, n, dLU.as<cplx>(), n, n2, dip.as<int>(), dDinv.as<cplx>(), dX.as<cplx>(), n, n2, 2 * G);
  // A = I + X, B = I - X
  dA.alloc(sizeof(cplx) * n2 * 2 * G);
  dB.alloc(sizeof(cplx) * n2 * 2 * G);
  eye_axpby_kernel<<<g1(n2, 2 * G), 256, 0, st>>>(n, ONE, ONE, dX.as<cplx>(), n2, dA.as<cplx>(), n2);
  eye_axpby_kernel<<<g1(n2, 2 * G), 256, 0, st>>>(n, ONE, MONE, dX.as<cplx>(), n2, dB.as<cplx>(), n2);
  // LU(A); AiB = A \ B; Ai = A \ I; T = 0.5 (A - B AiB); BA = B / A
  HIPCHECK(hipMemcpyAsync(dLU.p, dA.p, sizeof(cplx) * n2 * 2 * G, hipMemcpyDeviceToDevice, st));
  orf_zgetrf(st, n, dLU.as<cplx>(), n, n2, dip.as<int>(), nullptr, 2 * G, dDinv.as<cplx>());
  dAi.alloc(sizeof(cplx) * n2 * 2 * G * 3);
  cplx* AiB = dAi.as<cplx>();
  cplx* Ai = AiB + n2 * 2 * G;
  cplx* BA = Ai + n2 * 2 * G;
  HIPCHECK(hipMemcpyAsync(AiB, dB.p, sizeof(cplx) * n2 * 2 * G, hipMemcpyDeviceToDevice, st));
  orf_zgetrs(st, 'L', n, n, dLU.as<cplx>(), n, n2, dip.as<int>(), dDinv.as<cplx>(), AiB, n, n2, 2 * G);
  eye_axpby_kernel<<<g1(n2, 2 * G), 256, 0, st>>>(n, ONE, ZERO, nullptr, 0, Ai, n2);
  orf_zgetrs(st, 'L', n, n, dLU.as<cplx>(), n, n2, dip.as<int>(), dDinv.as<cplx>(), Ai, n, n2, 2 * G);
  HIPCHECK(hipMemcpyAsync(BA, dB.p, sizeof(cplx) * n2 * 2 * G, hipMemcpyDeviceToDevice, st));
  orf_zgetrs(st, 'R', n, n, dLU.as<cplx>(), n, n2, dip.as<int>(), dDinv.as<cplx>(), BA, n, n2, 2 * G);
  dT.alloc(sizeof(cplx) * n2 * 2 * G);
  eye_axpby_kernel<<<g1(n2, 2 * G), 256, 0, st>>>(n, ZERO, {0.5, 0.0}, dA.as<cplx>(), n2, dT.as<cplx>(), n2);
  orf_zgemm(st, 'N', 'N', n, n, n, {-0.5, 0.0}, dB.as<cplx>(), n, n2, AiB, n, n2, ONE, dT.as<cplx>(), n, n2, 2 * G);
  // assemble: ref uses the incident half (k = 0), trn the substrate half (k = 1)
  P.dSref.alloc(sizeof(cplx) * n2 * G * 4);
  P.dStrn.alloc(sizeof(cplx) * n2 * G * 4);
  cplx* R = P.dSref.as<cplx>();
  cplx* T = P.dStrn.as<cplx>();
  const long long blk = n2 * G;
  // ref: S11 = -AiB, S12 = 2 Ai, S21 = T, S22 = B / A
  eye_axpby_kernel<<<g1(n2, G), 256, 0, st>>>(n, ZERO, MONE, AiB, n2, R + 0 * blk, n2);
  eye_axpby_kernel<<<g1(n2, G), 256, 0, st>>>(n, ZERO, {2.0, 0.0}, Ai, n2, R + 1 * blk, n2);
  HIPCHECK(hipMemcpyAsync(R + 2 * blk, dT.p, sizeof(cplx) * blk, hipMemcpyDeviceToDevice, st));
  HIPCHECK(hipMemcpyAsync(R + 3 * blk, BA, sizeof(cplx) * blk, hipMemcpyDeviceToDevice, st));
  // trn: S11 = B / A, S12 = T, S21 = 2 Ai, S22 = -AiB
  HIPCHECK(hipMemcpyAsync(T + 0 * blk, BA + blk, sizeof(cplx) * blk, hipMemcpyDeviceToDevice, st));
  HIPCHECK(hipMemcpyAsync(T + 1 * blk, dT.as<cplx>() + blk, sizeof(cplx) * blk, hipMemcpyDeviceToDevice, st));
  eye_axpby_kernel<<<g1(n2, G), 256, 0, st>>>(n, ZERO, {2.0, 0.0}, Ai + blk, n2, T + 2 * blk, n2);
  eye_axpby_kernel<<<g1(n2, G), 256, 0, st>>>(n, ZERO, MONE, AiB + blk, n2, T + 3 * blk, n2);
  KCHECK();
  HIPCHECK(hipStreamSynchronize(st));
}

// Solve pages per chunk in the all-layers-at-once execution.
static const int kLayerChunkSolves = 4096;

// ---------------------------------------------------------------- plan creation
static void plan_create(orf_ctx* ctx, const orf_problem* pr, orf_plan& P) {
  if (!pr) fail(ORF_INVALID_ARG, "problem is NULL");
  P.ctx = ctx;
  // one stream per plan (default priority) so that plans on different host
  // threads overlap on the device
  HIPCHECK(hipStreamCreateWithFlags(&P.st, hipStreamNonBlocking));
  P.ownStream = true;
  P.Lx = pr->Lx;
  P.Ly = pr->Ly;
  if (!(P.Lx > 0) || !std::isfinite(P.Lx) || !(P.Ly > 0)) fail(ORF_INVALID_ARG, "period must be positive (Lx finite)");
  P.Nx = pr->Nx;
  P.Ny = std::isfinite(P.Ly) ? pr->Ny : 0;
  if (P.Nx < 0 || P.Ny < 0) fail(ORF_INVALID_ARG, "Nx, Ny must be non-negative");
  P.Nsx = pr->Nsx;
  P.Nsy = std::isfinite(P.Ly) ? pr->Nsy : 1;
  if (P.Nsx < 4 * P.Nx + 1 || P.Nsy < 4 * P.Ny + 1)
    fail(ORF_SAMPLING, "Need at least 4N+1 samples per period (convMatrix1D:sampling)");
  const int Wx = 2 * P.Nx + 1, Wy = 2 * P.Ny + 1;
  P.M = Wx * Wy;
  P.nf = 2 * P.M;
  // the M x M convolution-matrix LU keeps a 4-column panel of M rows in LDS
  if (P.M > kMaxHarmonics)
    fail(ORF_INVALID_ARG, "M = " + std::to_string(P.M) + " harmonics: at most " + std::to_string(kMaxHarmonics) +
                              " are supported");
  P.B = pr->nStruct;
  P.L = pr->nLambda;
  P.nPol = pr->nPol;
  P.nLayers = pr->nLayers;
  if (P.B <= 0 || P.L <= 0 || P.nPol <= 0 || P.nLayers < 0) fail(ORF_INVALID_ARG, "empty batch");
  P.theta = pr->theta;
  P.phi = pr->phi;
  P.factorization = pr->factorization;
  P.precision = pr->precision;
  if (P.precision != ORF_PREC_FP64 && P.precision != ORF_PREC_MIXED && P.precision != ORF_PREC_SPLIT)
    fail(ORF_INVALID_ARG, "unknown precision");
  if (pr->nMaterials <= 0 || !pr->matEps) fail(ORF_INVALID_ARG, "material table missing");
  auto matEps = [&](int mat, int l) -> zc {
    if (mat < 0 || mat >= pr->nMaterials) fail(ORF_INVALID_ARG, "material index out of range");
    const double* e = pr->matEps + 2 * ((size_t)mat + (size_t)pr->nMaterials * l);
    return zc(e[0], e[1]);
  };
  P.lambda.assign(pr->lambda0, pr->lambda0 + P.L);
  for (double lm : P.lambda)
    if (!(lm > 0)) fail(ORF_INVALID_ARG, "lambda0 must be positive");
  // harmonics (harmonics2D.m:4-6)
  P.mIdx.resize(P.M);
  P.nIdx.resize(P.M);
  for (int p = 0; p < P.M; ++p) {
    P.mIdx[p] = p % Wx - P.Nx;
    P.nIdx[p] = p / Wx - P.Ny;
  }
  P.i0 = P.Nx + Wx * P.Ny;
  // period of every structure (rcwa2D.m:4-7 shares one; uneven batches give one per structure)
  std::vector<double> perX(P.B, P.Lx), perY(P.B, P.Ly);
  if (pr->structPeriod) {
    for (int b = 0; b < P.B; ++b) {
      perX[b] = pr->structPeriod[2 * (size_t)b];
      perY[b] = pr->structPeriod[2 * (size_t)b + 1];
      if (!(perX[b] > 0) || !std::isfinite(perX[b]) || !(perY[b] > 0) ||
          std::isfinite(perY[b]) != std::isfinite(P.Ly))
        fail(ORF_INVALID_ARG, "structPeriod must be positive, with Ly finite exactly when the problem's Ly is");
    }
  }
  // pages: every (structure, wavelength) pair, or the caller's explicit list
  if (pr->nPages > 0) {
    if (!pr->pageStruct || !pr->pageLambda) fail(ORF_INVALID_ARG, "pageStruct and pageLambda are required");
    P.nPages = pr->nPages;
    for (int q = 0; q < P.nPages; ++q)
      if (pr->pageStruct[q] < 0 || pr->pageStruct[q] >= P.B || pr->pageLambda[q] < 0 || pr->pageLambda[q] >= P.L)
        fail(ORF_INVALID_ARG, "page index out of range");
  } else if (pr->nPages == 0) {
    P.nPages = P.B * P.L;
  } else {
    fail(ORF_INVALID_ARG, "nPages must be non-negative");
  }
  // contexts: distinct (wavelength, period) pairs in page order
  P.pageB.resize(P.nPages);
  P.pageCtx.resize(P.nPages);
  std::map<std::tuple<int, double, double>, int> ctxOf;
  std::vector<int> ctxLam;
  std::vector<double> ctxLx, ctxLy;
  for (int q = 0; q < P.nPages; ++q) {
    const int b = pr->nPages > 0 ? pr->pageStruct[q] : q % P.B;
    const int l = pr->nPages > 0 ? pr->pageLambda[q] : q / P.B;
    auto key = std::make_tuple(l, perX[b], std::isfinite(perY[b]) ? perY[b] : -1.0);
    auto it = ctxOf.find(key);
    int c;
    if (it == ctxOf.end()) {
      c = (int)ctxLam.size();
      ctxOf[key] = c;
      ctxLam.push_back(l);
      ctxLx.push_back(perX[b]);
      ctxLy.push_back(perY[b]);
    } else {
      c = it->second;
    }
    P.pageB[q] = b;
    P.pageCtx[q] = c;
  }
  P.nCtx = (int)ctxLam.size();
  // k-vectors per context (sign convention: +m lambda0 / Lx)
  const double st_ = std::sin(P.theta), ct = std::cos(P.theta);
  const double khat[3] = {st_ * std::cos(P.phi), st_ * std::sin(P.phi), ct};
  std::vector<double> kxy((size_t)2 * P.M * P.nCtx), kzi0(P.nCtx);
  std::vector<cplx> kz((size_t)2 * P.M * P.nCtx);
  std::vector<zc> eInc(P.nCtx), eSub(P.nCtx), eGap(P.nCtx);
  P.k0.resize(P.nCtx);
  for (int c = 0; c < P.nCtx; ++c) {
    const int l = ctxLam[c];
    const double lm = P.lambda[l];
    P.k0[c] = 2 * M_PI / lm;
    eInc[c] = matEps(pr->incMaterial, l);
    eSub[c] = matEps(pr->subMaterial, l);
    const zc ninc = std::sqrt(eInc[c]);
    if (std::fabs(ninc.imag()) > 1e-12) fail(ORF_LOSSY_INCIDENT, "Incident medium must be lossless.");
    const double nInc = ninc.real();
    kzi0[c] = nInc * ct;
    if (c == 0) {
      P.kzi0 = kzi0[0];
      P.epsInc0 = eInc[0];
      P.epsSub0 = eSub[0];
    }
    double kmax = 0;
    for (int p = 0; p < P.M; ++p) {
      const double kx = nInc * khat[0] + P.mIdx[p] * lm / ctxLx[c];
      const double ky = nInc * khat[1] + (std::isfinite(ctxLy[c]) ? P.nIdx[p] * lm / ctxLy[c] : 0.0);
      kxy[(size_t)c * 2 * P.M + p] = kx;
      kxy[(size_t)c * 2 * P.M + P.M + p] = ky;
      kz[(size_t)c * 2 * P.M + p] = C(kz_nonzero_h(kz_branch_h(eInc[c] - kx * kx - ky * ky)));
      kz[(size_t)c * 2 * P.M + P.M + p] = C(kz_nonzero_h(kz_branch_h(eSub[c] - kx * kx - ky * ky)));
      kmax = std::max(kmax, kx * kx + ky * ky);
    }
    eGap[c] = std::isnan(pr->gapEps) ? zc(1.0 + kmax, 0.0) : zc(pr->gapEps, 0.0);
  }
  P.dKxy.upload(kxy);
  P.dKz.upload(kz);
  P.dKzi0.upload(kzi0);
  // sources
  bool allSP = true;
  std::vector<zc> Ex(P.nPol), Ey(P.nPol);
  for (int i = 0; i < P.nPol; ++i) {
    const int k = pr->polKind[i];
    zc jte, jtm;
    if (k == ORF_POL_S) { jte = 1; jtm = 0; }
    else if (k == ORF_POL_P) { jte = 0; jtm = 1; }
    else if (k == ORF_POL_JONES) {
      if (!pr->jones) fail(ORF_INVALID_POL, "Jones vector must be [aTE aTM].");
      jte = zc(pr->jones[4 * i], pr->jones[4 * i + 1]);
      jtm = zc(pr->jones[4 * i + 2], pr->jones[4 * i + 3]);
      allSP = false;
    } else fail(ORF_INVALID_POL, "pol must be \"s\", \"p\" or [aTE aTM].");
    if (std::abs(jte) == 0 && std::abs(jtm) == 0) fail(ORF_INVALID_POL, "Jones vector must be nonzero.");
    zc E[3];
    pol_vector(khat, jte, jtm, E);
    Ex[i] = E[0];
    Ey[i] = E[1];
  }
  // layers and grids
  P.layerGrid.assign(P.nLayers, std::vector<int>(P.B, -1));
  P.layerAllPat.assign(P.nLayers, 0);
  P.layerAnyPat.assign(P.nLayers, 0);
  P.thick.resize((size_t)P.nLayers * P.B);
  std::vector<DevGrid> grids;
  std::vector<DevShape> shapes;
  std::vector<std::vector<int>> gridSlotMat;  // slot -> material
  for (int j = 0; j < P.nLayers; ++j) {
    int npat = 0;
    for (int b = 0; b < P.B; ++b) npat += pr->shapeCount[j + (size_t)P.nLayers * b] > 0;
    P.layerAnyPat[j] = npat > 0;
    P.layerAllPat[j] = npat == P.B;
    for (int b = 0; b < P.B; ++b) {
      const size_t jb = j + (size_t)P.nLayers * b;
      P.thick[(size_t)j * P.B + b] = pr->thickness[jb];
      if (pr->thickness[jb] < 0) fail(ORF_INVALID_ARG, "thickness must be non-negative");
      const int cnt = pr->shapeCount[jb];
      if (cnt == 0 && npat == 0) continue;
      // patterned (or uniform inside a layer that is patterned elsewhere in the batch)
      DevGrid g{(int)shapes.size(), cnt, perX[b], perY[b]};
      std::vector<int> slots{pr->background[jb]};
      for (int s = 0; s < cnt; ++s) {
        const orf_shape& sh = pr->shapes[pr->shapeStart[jb] + s];
        int slot = (int)(std::find(slots.begin(), slots.end(), sh.material) - slots.begin());
        if (slot == (int)slots.size()) slots.push_back(sh.material);
        if ((int)slots.size() > ORF_MAX_SLOTS) fail(ORF_INVALID_ARG, "more than 8 materials in one layer");
        DevShape d{sh.type, slot, sh.cx, sh.cy, sh.a, sh.b, 0, 0};
        if (sh.type == ORF_SHAPE_POLYGON) {
          if (sh.nvert < 3 || sh.vert0 < 0 || sh.vert0 + sh.nvert > pr->nVertices)
            fail(ORF_INVALID_ARG, "A polygon needs at least 3 vertices.");
          d.vert0 = sh.vert0;
          d.nvert = sh.nvert;
        } else if (sh.type != ORF_SHAPE_CIRCLE && sh.type != ORF_SHAPE_RECT) {
          fail(ORF_INVALID_ARG, "unknown shape type");
        }
        (void)matEps(sh.material, 0);
        shapes.push_back(d);
      }
      P.layerGrid[j][b] = (int)grids.size();
      grids.push_back(g);
      gridSlotMat.push_back(slots);
      P.nslotMax = std::max(P.nslotMax, (int)slots.size());
    }
  }
  P.nGrids = (int)grids.size();
  // polygon vertices as separate x and y runs: [x(vert0..), y(vert0..)] per polygon
  std::vector<double> verts;
  for (auto& d : shapes)
    if (d.type == ORF_SHAPE_POLYGON) {
      int off = (int)verts.size();
      for (int v = 0; v < d.nvert; ++v) verts.push_back(pr->vertices[2 * (size_t)(d.vert0 + v)]);
      for (int v = 0; v < d.nvert; ++v) verts.push_back(pr->vertices[2 * (size_t)(d.vert0 + v) + 1]);
      d.vert0 = off;
    }
  if (verts.empty()) verts.push_back(0.0);
  hipStream_t st = P.st;
  if (P.nGrids > 0) {
    P.dShapes.upload(shapes.empty() ? std::vector<DevShape>(1) : shapes);
    P.dVerts.upload(verts);
    P.dGrids.upload(grids);
    const long long cells = (long long)P.Nsx * P.Nsy;
    P.dLabels.alloc((size_t)cells * P.nGrids);
    raster_kernel<<<dim3((unsigned)((cells + 255) / 256), P.nGrids), 256, 0, st>>>(
        P.Nsx, P.Nsy, P.dGrids.as<DevGrid>(), P.dShapes.as<DevShape>(), P.dVerts.as<double>(),
        P.dLabels.as<uint8_t>());
    KCHECK();
  }
  // symmetry decision (rcwa2D.m:24-27): theta = 0, pol "s"/"p", mirror-symmetric grids
  P.reduced = 0;
  P.keepModes = pr->keepModes != 0;
  if (P.keepModes && P.nPages != 1) fail(ORF_INVALID_ARG, "keepModes needs exactly one page");
  if (pr->symmetry == ORF_SYM_AUTO && P.theta == 0.0 && allSP && !P.keepModes) {
    int sym = 1;
    if (P.nGrids > 0) {
      DBuf flags;
      flags.upload(std::vector<int>(P.nGrids, 1));
      const long long cells = (long long)P.Nsx * P.Nsy;
      grid_symmetry_kernel<<<dim3((unsigned)std::min<long long>((cells + 255) / 256, 1024), P.nGrids), 256, 0, st>>>(
          P.Nsx, P.Nsy, P.dLabels.as<uint8_t>(), flags.as<int>());
      KCHECK();
      std::vector<int> h(P.nGrids);
      HIPCHECK(hipMemcpyAsync(h.data(), flags.p, sizeof(int) * P.nGrids, hipMemcpyDeviceToHost, st));
      HIPCHECK(hipStreamSynchronize(st));
      for (int v : h) sym &= v;
    }
    P.reduced = sym;
  }
  // otherwise a y-invariant grating lit in the x-z plane splits into TE and TM
  // (exact, like the mirror reduction; same sources "s" / "p" only)
  if (!P.reduced && pr->symmetry == ORF_SYM_AUTO && P.Ny == 0 && std::sin(P.theta) * std::sin(P.phi) == 0.0 &&
      allSP && !P.keepModes)
    P.reduced = kReducedTeTm;
  // sectors and sources
  std::vector<HostSector> hs;
  std::vector<int> srcSector(P.nPol, 0);
  if (P.reduced) {
    int secOf[2] = {-1, -1};  // [s, p]
    for (int i = 0; i < P.nPol; ++i) {
      const int k = pr->polKind[i];
      if (secOf[k] < 0) {
        int nr = P.M;
        if (P.reduced == kReducedTeTm) hs.push_back(build_sector_te_tm(P.M, k == ORF_POL_P));
        else hs.push_back(build_sector(P.Nx, P.Ny, k == ORF_POL_P, nr));
        secOf[k] = (int)hs.size() - 1;
        P.nr = nr;
      }
      srcSector[i] = secOf[k];
    }
    P.nSec = (int)hs.size();
  } else {
    P.nr = P.nf;
    P.nSec = 1;
  }
  if (P.nr > kMaxSolved)
    fail(ORF_INVALID_ARG, "solved matrix size " + std::to_string(P.nr) + " (2M = " + std::to_string(P.nf) +
                              (P.reduced ? ", reduced basis" : ", full basis") + "): at most " +
                              std::to_string(kMaxSolved) + " is supported");
  {
    std::vector<DevSector> ds(std::max(1, P.nSec));
    for (int s = 0; s < P.nSec && P.reduced; ++s) {
      auto up = [&](auto& vec) {
        DBuf* b = new DBuf();
        P.secArrays.push_back(b);
        b->upload(vec);
        return b->p;
      };
      ds[s].ue_row = (const int*)up(hs[s].ue_row);
      ds[s].ue_coef = (const double*)up(hs[s].ue_coef);
      ds[s].uh_row = (const int*)up(hs[s].uh_row);
      ds[s].uh_coef = (const double*)up(hs[s].uh_coef);
      ds[s].ue_rowmap = (const int*)up(hs[s].ue_rowmap);
      ds[s].ue_rowcoef = (const double*)up(hs[s].ue_rowcoef);
    }
    P.dSectors.upload(ds);
  }
  P.sources.resize(P.nPol);
  for (int i = 0; i < P.nPol; ++i) {
    DevSource so{};
    so.pol = i;
    so.sector = srcSector[i];
    so.ex = C(Ex[i]);
    so.ey = C(Ey[i]);
    so.srcCol = 0;
    so.srcVal = C(0);
    if (P.reduced) {
      const HostSector& s = hs[srcSector[i]];
      int a1 = s.ue_rowmap[P.i0], a2 = s.ue_rowmap[P.M + P.i0];
      zc v1 = a1 >= 0 ? s.ue_rowcoef[P.i0] * Ex[i] : 0.0;
      zc v2 = a2 >= 0 ? s.ue_rowcoef[P.M + P.i0] * Ey[i] : 0.0;
      if (a1 >= 0 && std::abs(v1) > 0) { so.srcCol = a1; so.srcVal = C(v1); }
      else { so.srcCol = std::max(a2, 0); so.srcVal = C(v2); }
    }
    P.sources[i] = so;
  }
  P.dSources.upload(P.sources);
  P.sectorSrcCol.assign(std::max(1, P.nSec), 0);
  for (auto& so : P.sources) P.sectorSrcCol[so.sector] = so.srcCol;
  // Fourier tables and geometry transforms
  if (P.nGrids > 0) {
    const int Hx = 4 * P.Nx + 1, Hy = 4 * P.Ny + 1;
    P.dFx.upload(fourier_weights_h(P.Nsx, P.Nx));
    P.dFy.upload(fourier_weights_h(P.Nsy, P.Ny));
    const int ns = P.nslotMax;
    P.dChiX.alloc(sizeof(cplx) * (size_t)P.nGrids * ns * Hx * P.Nsy);
    P.dChiY.alloc(sizeof(cplx) * (size_t)P.nGrids * ns * Hy * P.Nsx);
    P.dChi2.alloc(sizeof(cplx) * (size_t)P.nGrids * ns * Hx * Hy);
    HIPCHECK(hipMemsetAsync(P.dChiX.p, 0, P.dChiX.bytes, st));
    HIPCHECK(hipMemsetAsync(P.dChiY.p, 0, P.dChiY.bytes, st));
    chi_x_kernel<<<dim3((unsigned)(((long long)P.Nsy * Hx + 255) / 256), P.nGrids), 256, 0, st>>>(
        P.Nsx, P.Nsy, Hx, ns, P.dLabels.as<uint8_t>(), P.dFx.as<cplx>(), P.dChiX.as<cplx>());
    chi_y_kernel<<<dim3((unsigned)(((long long)P.Nsx * Hy + 255) / 256), P.nGrids), 256, 0, st>>>(
        P.Nsx, P.Nsy, Hy, ns, P.dLabels.as<uint8_t>(), P.dFy.as<cplx>(), P.dChiY.as<cplx>());
    chi_2d_kernel<<<dim3((unsigned)(((long long)Hx * Hy * ns + 255) / 256), P.nGrids), 256, 0, st>>>(
        P.Nsy, Hx, Hy, ns, P.dChiX.as<cplx>(), P.dFy.as<cplx>(), P.dChi2.as<cplx>());
    KCHECK();
  }
  // per-layer page tables and slot permittivities
  const int npages = P.nPages;
  P.layerBgEps.assign(P.nLayers, std::vector<zc>(npages));
  for (int j = 0; j < P.nLayers; ++j) {
    std::vector<DevPage> pages(npages);
    std::vector<cplx> se((size_t)npages * P.nslotMax, C(0)), si((size_t)npages * P.nslotMax, C(0));
    for (int pg = 0; pg < npages; ++pg) {
      const int b = P.pageB[pg], c = P.pageCtx[pg], l = ctxLam[c];
      const size_t jb = j + (size_t)P.nLayers * b;
      const int g = P.layerGrid[j][b];
      pages[pg] = DevPage{b, c, g, g >= 0 ? (int)gridSlotMat[g].size() : 0};
      P.layerBgEps[j][pg] = matEps(pr->background[jb], l);
      if (g >= 0)
        for (size_t s = 0; s < gridSlotMat[g].size(); ++s) {
          zc e = matEps(gridSlotMat[g][s], l);
          se[(size_t)pg * P.nslotMax + s] = C(e);
          si[(size_t)pg * P.nslotMax + s] = C(1.0 / e);
        }
    }
    DBuf* dp = new DBuf();
    dp->upload(pages);
    P.dPagesByLayer.push_back(dp);
    P.hPages.push_back(pages);
    P.hSe.push_back(se);
    P.hSi.push_back(si);
    DBuf* de = new DBuf();
    de->upload(se);
    P.dSlotEps.push_back(de);
    DBuf* di = new DBuf();
    di->upload(si);
    P.dSlotInv.push_back(di);
  }
  {
    std::vector<DevPage> base(npages);
    for (int pg = 0; pg < npages; ++pg) base[pg] = DevPage{P.pageB[pg], P.pageCtx[pg], -1, 0};
    DBuf* dp = new DBuf();
    dp->upload(base);
    P.dPagesByLayer.push_back(dp);  // index nLayers: base pages
  }
  // shared half-space S-matrices per (context, sector)
  build_regions(P, eInc, eSub, eGap);
  // chunk size by device memory
  size_t freeB = 0, totB = 0;
  HIPCHECK(hipMemGetInfo(&freeB, &totB));
  const size_t perMode = per_mode_bytes(P) + (size_t)P.nSec * per_solve_bytes(P);
  size_t budget = (size_t)(0.6 * (double)(freeB + g_pool.cachedBytes));
  int cp = (int)std::max<size_t>(1, budget / std::max<size_t>(perMode, 1));
  cp = std::min(cp, 32768 / P.nSec);  // grid-dimension limit of the page-batched launches
  P.layerParallel = P.nLayers >= 2 && !P.keepModes;
  P.directSingle = P.nLayers == 1 && !P.keepModes && P.layerAnyPat[0] && P.nr >= (P.reduced ? 8 : 16);
  if (P.layerParallel) {
    // every layer of a page is in the chunk at once: cap the chunk at kLayerChunkSolves
    // solve pages (enough to fill the device) so that its workspace stays moderate
    const int capPages = std::max(1, kLayerChunkSolves / (P.nLayers * P.nSec));
    P.chunkPages = std::max(1, std::min(std::min(cp / P.nLayers, npages), capPages));
    if (pr->batchSize > 0) P.chunkPages = std::min(P.chunkPages, pr->nPages > 0 ? pr->batchSize : pr->batchSize * P.L);
    P.workModes = P.chunkPages * P.nLayers;
  } else {
    if (pr->batchSize > 0) cp = std::min(cp, pr->nPages > 0 ? pr->batchSize : pr->batchSize * P.L);
    P.chunkPages = std::min(cp, npages);
    P.workModes = P.chunkPages;
  }
  P.work = make_work(P, P.workModes);
  // results
  const size_t nout = (size_t)P.nPages * P.nPol;
  P.dR.alloc(sizeof(double) * P.M * nout);
  P.dT.alloc(sizeof(double) * P.M * nout);
  P.dRtot.alloc(sizeof(double) * nout);
  P.dTtot.alloc(sizeof(double) * nout);
  P.dr.alloc(sizeof(cplx) * 3 * P.M * nout);
  P.dt.alloc(sizeof(cplx) * 3 * P.M * nout);
  HIPCHECK(hipStreamSynchronize(st));
}

// Modes of Cp patterned mode pages (pages[m], slot permittivities se/si at
// m * nslotMax), solve pages w.solves[0 .. Cs): convolution matrices
// (Laurent / Li), Omega^2 = P Q, eigen-decomposition, W, V = Q W lam^-1 and
// lam into w.Wm, w.Vm, w.lam (V only for sector solves, kept modes and sizes
// the eigensolver's W^-1 does not cover; full-basis solves leave [eps]^-1 in
// w.Omega instead).  keep (single-page field solves) receives [eps].
// Pages of a chunk that share a label grid and slot permittivities (the same
// structure at wavelengths where its materials do not disperse) share [eps],
// Exx and Eyy: the convolution stage runs once per distinct pair (Cu pages,
// tables uPages / uSe / uSi) and uMap[m] names the distinct page of mode page m.
struct ConvUnique {
  int Cu = 0;
  const DevPage* uPages = nullptr;
  const cplx *uSe = nullptr, *uSi = nullptr;
  const int* uMap = nullptr;
};

static void patterned_modes(orf_plan& P, Work& w, const DevPage* pages, const cplx* se, const cplx* si, int Cp, int Cs,
                            orf_plan::Kept* keep, const ConvUnique* cu = nullptr) {
  hipStream_t st = P.st;
  const int M = P.M, nf = P.nf, nr = P.nr, nSec = P.nSec;
  const long long M2 = (long long)M * M, nf2 = (long long)nf * nf, n2 = (long long)nr * nr;
  const long long nfr = (long long)nf * nr;
  const int Wx = 2 * P.Nx + 1, Wy = 2 * P.Ny + 1, Hx = 4 * P.Nx + 1, Hy = 4 * P.Ny + 1;
  // with distinct pages: [eps] into G and Exx, Eyy into Omega (free until the
  // omega stage), Exx and Eyy then copied out to every mode page
  const bool uq = cu && cu->Cu < Cp;
  const int Cc = uq ? cu->Cu : Cp;
  {
    STAGE("conv");
    const DevPage* cpages = uq ? cu->uPages : pages;
    const cplx* cse = uq ? cu->uSe : se;
    const cplx* csi = uq ? cu->uSi : si;
    cplx* cE = uq ? w.G : w.E;
    cplx* cExx = uq ? w.Omega : w.Exx;
    cplx* cEyy = uq ? w.Omega + (size_t)Cc * M2 : w.Eyy;
    laurent_kernel<<<g1(M2, Cc), 256, 0, st>>>(P.Nx, P.Ny, M, cpages, P.dChi2.as<cplx>(), cse, P.nslotMax, cE, M2);
    if (P.factorization == ORF_FACT_LI) {
      // Exx: inverse rule along x at each y-sample, Laurent along y
      const long long perx = (long long)Wx * Wx * P.Nsy;
      li_toeplitz_kernel<<<g1(perx, Cc), 256, 0, st>>>(0, P.Nx, P.Nsy, Hx, cpages, P.dChiX.as<cplx>(),
                                                        (long long)P.nslotMax * Hx * P.Nsy, csi, P.nslotMax, w.Toe);
      KCHECK();
      // invert every Toeplitz block: LU then solve against the identity (in Yw scratch)
      const int nb = Cc * P.Nsy;
      cplx* inv = w.Inv;
      if (!orf_small_inverse(st, Wx, w.Toe, (long long)Wx * Wx, inv, (long long)Wx * Wx, nb)) {
        eye_stack_kernel<<<g1((long long)Wx * Wx * nb), 256, 0, st>>>(Wx, (long long)nb, inv);
        orf_zgetrf(st, Wx, w.Toe, Wx, (long long)Wx * Wx, w.ipivT, nullptr, nb, nullptr);
        orf_zgetrs(st, 'L', Wx, Wx, w.Toe, Wx, (long long)Wx * Wx, w.ipivT, nullptr, inv, Wx, (long long)Wx * Wx, nb);
      }
      // G = Fy * [A_j]^T : (Hy x Nsy) * (Nsy x Wx^2)
      bulk_gemm(P, 'N', 'T', Hy, Wx * Wx, P.Nsy, ONE, P.dFy.as<cplx>(), Hy, 0, inv, Wx * Wx,
                (long long)Wx * Wx * P.Nsy, ZERO, w.GH, Hy, (long long)Hy * Wx * Wx, Cc);
      // Eyy: inverse rule along y at each x-sample, Laurent along x
      const long long pery = (long long)Wy * Wy * P.Nsx;
      li_toeplitz_kernel<<<g1(pery, Cc), 256, 0, st>>>(1, P.Ny, P.Nsx, Hy, cpages, P.dChiY.as<cplx>(),
                                                        (long long)P.nslotMax * Hy * P.Nsx, csi, P.nslotMax, w.Toe);
      const int nb2 = Cc * P.Nsx;
      if (!orf_small_inverse(st, Wy, w.Toe, (long long)Wy * Wy, inv, (long long)Wy * Wy, nb2)) {
        eye_stack_kernel<<<g1((long long)Wy * Wy * nb2), 256, 0, st>>>(Wy, (long long)nb2, inv);
        orf_zgetrf(st, Wy, w.Toe, Wy, (long long)Wy * Wy, w.ipivT, nullptr, nb2, nullptr);
        orf_zgetrs(st, 'L', Wy, Wy, w.Toe, Wy, (long long)Wy * Wy, w.ipivT, nullptr, inv, Wy, (long long)Wy * Wy, nb2);
      }
      bulk_gemm(P, 'N', 'T', Hx, Wy * Wy, P.Nsx, ONE, P.dFx.as<cplx>(), Hx, 0, inv, Wy * Wy,
                (long long)Wy * Wy * P.Nsx, ZERO, w.HH, Hx, (long long)Hx * Wy * Wy, Cc);
      li_gather_kernel<<<g1(M2, Cc), 256, 0, st>>>(P.Nx, P.Ny, M, w.GH, w.HH, cExx, cEyy, M2);
    } else {
      HIPCHECK(hipMemcpyAsync(cExx, cE, sizeof(cplx) * M2 * Cc, hipMemcpyDeviceToDevice, st));
      HIPCHECK(hipMemcpyAsync(cEyy, cE, sizeof(cplx) * M2 * Cc, hipMemcpyDeviceToDevice, st));
    }
    if (uq) {  // [eps] stays compact: the omega stage inverts it there
      gather_pages_kernel<<<g1(M2, Cp), 256, 0, st>>>(M2, cu->uMap, cExx, M2, w.Exx, M2);
      gather_pages_kernel<<<g1(M2, Cp), 256, 0, st>>>(M2, cu->uMap, cEyy, M2, w.Eyy, M2);
    }
    if (keep) {
      keep->E.alloc(sizeof(cplx) * M2);
      keep->patterned = true;
      HIPCHECK(hipMemcpyAsync(keep->E.p, w.E, sizeof(cplx) * M2, hipMemcpyDeviceToDevice, st));
    }
    KCHECK();
  }
  {
    STAGE("omega");
    // [eps]^-1 explicitly, once per distinct page, into the first M^2 of each page
    // of Omega (the full basis's layer S-matrix uses it: V^-1 = lam^-1 W^-1 P), then
    // G = [eps]^-1 [Kx Exx, Ky Eyy] by one product
    if (uq) {  // compact [eps] in G; its compact inverse in E (not expanded, so free)
      orf_zgetrf(st, M, w.G, M, M2, w.ipiv, nullptr, Cc, w.dinv);
      eye_axpby_kernel<<<g1(M2, Cc), 256, 0, st>>>(M, ONE, ZERO, nullptr, 0, w.E, M2);
      orf_zgetrs(st, 'L', M, M, w.G, M, M2, w.ipiv, w.dinv, w.E, M, M2, Cc);
      gather_pages_kernel<<<g1(M2, Cp), 256, 0, st>>>(M2, cu->uMap, w.E, M2, w.Omega, nf2);
    } else {
      orf_zgetrf(st, M, w.E, M, M2, w.ipiv, nullptr, Cp, w.dinv);
      eye_axpby_kernel<<<g1(M2, Cp), 256, 0, st>>>(M, ONE, ZERO, nullptr, 0, w.Omega, nf2);
      orf_zgetrs(st, 'L', M, M, w.E, M, M2, w.ipiv, w.dinv, w.Omega, M, nf2, Cp);
    }
    // full basis: right-hand side [Kx Exx, Ky Eyy] in G, G beside [eps]^-1 in Omega
    // (Omega^2 goes straight into A0); sector bases assemble Omega^2 in Omega and
    // project it, so there the right-hand side sits beside [eps]^-1 and G in G
    cplx* rhs = P.reduced ? w.Omega + M2 : w.G;
    cplx* Gm = P.reduced ? w.G : w.Omega + M2;
    const long long sR = P.reduced ? nf2 : 2 * M2, sG = P.reduced ? 2 * M2 : nf2;
    omega_rhs_kernel<<<g1(2 * M2, Cp), 256, 0, st>>>(M, pages, P.dKxy.as<double>(), w.Exx, w.Eyy, M2, rhs, sR);
    orf_zgemm(st, 'N', 'N', M, 2 * M, M, ONE, w.Omega, M, nf2, rhs, M, sR, ZERO, Gm, M, sG, Cp);
    omega_assemble_kernel<<<g1(nf2, Cp), 256, 0, st>>>(M, pages, P.dKxy.as<double>(), Gm, sG, w.Exx, w.Eyy, M2,
                                                       P.reduced ? w.Omega : w.A0, nf2);
    KCHECK();
    if (P.reduced)
      sector_project_kernel<<<g1(n2, Cs), 256, 0, st>>>(nf, nr, w.solves, P.dSectors.as<DevSector>(), 0, 0,
                                                         w.Omega, nf2, 0, w.A0, n2);
    KCHECK();
  }
  {
    STAGE("eig");
    int* info = next_info(P, w, Cs);
    orf_zgeev_timed(st, nr, w.A0, nr, n2, w.mu, w.Wm, nr, n2, w.Zw, w.Yw, w.Hv, info, Cs,
                    [&](const char* nm, int phase) {
                      if (phase == 0) P.timer.begin(nm, st); else P.timer.end(st);
                    }, w.vnorm);
    KCHECK();
  }
  {
    STAGE("modes");
    lambda_kernel<<<g1((long long)nr * Cs), 256, 0, st>>>(nr, Cs, w.mu, w.lam);
    cplx* Wf = w.Wm;
    if (P.reduced) {
      sector_expand_kernel<<<g1(nfr, Cs), 256, 0, st>>>(nf, nr, w.solves, P.dSectors.as<DevSector>(), w.Wm, n2,
                                                         w.Wf, nfr);
      Wf = w.Wf;
    }
    // V = Q W / lam; the layer S-matrix of a full-basis solve does without it
    // (V^-1 = lam^-1 W^-1 P), so it is formed only for sector solves and kept modes
    if (!P.reduced && !keep && !P.directSingle && orf_eig_inverse_ok(nr)) {
      KCHECK();
      return;
    }
    // T1 = Eyy W2, T2 = Exx W1 per sector
    for (int s = 0; s < nSec; ++s) {
      bulk_gemm(P, 'N', 'N', M, nr, M, ONE, w.Eyy, M, M2, Wf + s * nfr + M, nf, nSec * nfr, ZERO,
                w.T12 + s * nfr, nf, nSec * nfr, Cp);
      bulk_gemm(P, 'N', 'N', M, nr, M, ONE, w.Exx, M, M2, Wf + s * nfr, nf, nSec * nfr, ZERO,
                w.T12 + s * nfr + M, nf, nSec * nfr, Cp);
    }
    if (P.directSingle && !P.reduced) {  // finish_direct forms Q W itself from W and T12
      KCHECK();
      return;
    }
    qw_assemble_kernel<<<g1(nfr, Cs), 256, 0, st>>>(M, nr, w.solves, pages, P.dKxy.as<double>(), Wf, nfr, w.T12,
                                                    nfr, w.QW, nfr);
    if (P.reduced)
      sector_rows_kernel<<<g1(n2, Cs), 256, 0, st>>>(nf, nr, nr, w.solves, P.dSectors.as<DevSector>(), w.QW, nfr,
                                                     w.Vm, n2);
    else
      HIPCHECK(hipMemcpyAsync(w.Vm, w.QW, sizeof(cplx) * n2 * Cs, hipMemcpyDeviceToDevice, st));
    colscale_inv_kernel<<<g1(n2, Cs), 256, 0, st>>>(nr, nr, w.lam, w.Vm, n2);
    KCHECK();
  }
}

// Last stage of a chunk: S = S_acc * S_trn on the source columns, then the
// efficiencies.  Expects w.solves / w.groupOf in the standard order
// (solve page m * nSec + s for mode page m of the chunk).  SB[0], SB[2] receive
// S_trn; R, T (n x ncol per solve page) are scratch distinct from SA and SB.
// Stack form (L != nullptr): SA receives S_ref, L is the layers' composite
// S-matrix and finish_stack gives the result; otherwise SA is the accumulated
// S_ref * L and the column star with S_trn gives it.
static void finish_chunk(orf_plan& P, Work& w, int p0, int Cp, cplx* const SA[4], cplx* const SB[4], cplx* R, cplx* T,
                         cplx* const* L = nullptr, bool single = false) {
  hipStream_t st = P.st;
  const int nr = P.nr, nSec = P.nSec, Cs = Cp * nSec;
  const long long n2 = (long long)nr * nr, G = (long long)P.nCtx * nSec;
  const bool direct = single && P.directSingle;
  const bool gath = !direct || P.reduced;  // the full-basis direct solve reads the per-context tables itself
  for (int k = 0; k < 4 && gath; k += 2)
    gather_pages_kernel<<<g1(n2, Cs), 256, 0, st>>>(n2, w.groupOf, P.dStrn.as<cplx>() + k * n2 * G, n2, SB[k], n2);
  if ((L || single) && gath)
    for (int k = 0; k < 4; ++k)
      gather_pages_kernel<<<g1(n2, Cs), 256, 0, st>>>(n2, w.groupOf, P.dSref.as<cplx>() + k * n2 * G, n2, SA[k], n2);
  const int ncol = P.reduced ? 1 : 2;
  {
    std::vector<int> cols((size_t)Cs * ncol);
    for (int sp = 0; sp < Cs; ++sp) {
      if (P.reduced) cols[sp] = P.sectorSrcCol[sp % nSec];
      else { cols[2 * sp] = P.i0; cols[2 * sp + 1] = P.M + P.i0; }
    }
    w.colIdx = table(P, cols);
  }
  if (direct) finish_direct(P, w, Cs, P.dPagesByLayer[0]->as<DevPage>() + p0, SA, SB, ncol, w.colIdx, R, T);
  else if (single) finish_single(P, w, Cs, SA, SB[0], SB[2], ncol, w.colIdx, R, T);
  else if (L) finish_stack(P, w, Cs, SA, L, SB[0], SB[2], ncol, w.colIdx, R, T);
  else smatrix_star_cols(P, w, Cs, SA, SB, ncol, w.colIdx, R, T);
  {
    STAGE("output");
    const DevPage* basePages = P.dPagesByLayer[P.nLayers]->as<DevPage>() + p0;
    output_kernel<<<dim3(Cs, P.nPol), 256, 0, st>>>(
        P.M, nr, P.reduced, p0, P.nPages, w.solves, basePages, P.dSectors.as<DevSector>(), P.dSources.as<DevSource>(),
        P.dKxy.as<double>(), (const double*)P.dKz.p, P.dKzi0.as<double>(), R, T, (long long)nr * ncol,
        P.dR.as<double>(), P.dT.as<double>(), P.dRtot.as<double>(), P.dTtot.as<double>(), P.dr.as<cplx>(),
        P.dt.as<cplx>(), P.i0);
    KCHECK();
  }
}

// One chunk with every layer at once: the modes and layer S-matrices of all
// (layer, page, sector) triples in one batch (patterned layers through the
// eigen-solver, homogeneous ones analytically), then S_1 * ... * S_L as a
// binary tree of batched star products (log2 L levels instead of L sequential
// products), S_ref * that, and the last star product on the source columns.
// Used when no progress callback is given (the per-layer path reports after
// every layer).
static void execute_chunk_layers(orf_plan& P, Work& w, int p0, int Cp) {
  hipStream_t st = P.st;
  const int L = P.nLayers, nSec = P.nSec, nr = P.nr, M = P.M, ns = P.nslotMax;
  const long long n2 = (long long)nr * nr, G = (long long)P.nCtx * nSec;
  const int CpS = Cp * nSec;
  std::vector<int> patL, homL;
  for (int j = 0; j < L; ++j) (P.layerAnyPat[j] ? patL : homL).push_back(j);
  const int Lp = (int)patL.size(), Lh = (int)homL.size();
  const int CpP = Lp * Cp, CsP = CpP * nSec, CsH = Lh * CpS, CsAll = CsP + CsH;
  std::vector<DevPage> pgA(std::max(CpP, 1)), pgH(std::max(Lh * Cp, 1));
  std::vector<cplx> seA((size_t)std::max(CpP, 1) * ns), siA((size_t)std::max(CpP, 1) * ns);
  std::vector<DevSolve> sv(CsAll);
  std::vector<int> grp(CsAll), unified((size_t)L * CpS);
  std::vector<double> kL(CsAll);
  std::vector<cplx> epsH(std::max(CsH, 1));
  auto fill = [&](int j, int a, int base, bool pat) {
    for (int m = 0; m < Cp; ++m) {
      const int pg = p0 + m, mode = a * Cp + m;
      if (pat) {
        pgA[mode] = P.hPages[j][pg];
        for (int q = 0; q < ns; ++q) {
          seA[(size_t)mode * ns + q] = P.hSe[j][(size_t)pg * ns + q];
          siA[(size_t)mode * ns + q] = P.hSi[j][(size_t)pg * ns + q];
        }
      } else {
        pgH[mode] = P.hPages[j][pg];
      }
      for (int sc = 0; sc < nSec; ++sc) {
        const int sp = base + mode * nSec + sc;
        sv[sp] = DevSolve{mode, sc};
        grp[sp] = P.pageCtx[pg] * nSec + sc;
        kL[sp] = P.k0[P.pageCtx[pg]] * P.thick[(size_t)j * P.B + P.pageB[pg]];
        if (!pat) epsH[sp - base] = C(P.layerBgEps[j][pg]);
        unified[(size_t)j * CpS + m * nSec + sc] = sp;
      }
    }
  };
  for (int a = 0; a < Lp; ++a) fill(patL[a], a, 0, true);
  for (int a = 0; a < Lh; ++a) fill(homL[a], a, CsP, false);
  w.solves = table(P, sv);
  w.groupOf = table(P, grp);
  w.kL = table(P, kL);
  w.idxA = table(P, unified);
  if (CpP) {
    w.pagesA = table(P, pgA);
    w.seA = table(P, seA);
    w.siA = table(P, siA);
    patterned_modes(P, w, w.pagesA, w.seA, w.siA, CpP, CsP, nullptr);
  }
  if (CsH) {
    STAGE("modes");
    w.pagesH = table(P, pgH);
    w.epsSolve = table(P, epsH);
    homog_modes_kernel<<<g1(n2, CsH), 256, 0, st>>>(M, nr, P.reduced, w.solves + CsP, w.pagesH,
                                                    P.dSectors.as<DevSector>(), P.dKxy.as<double>(), w.epsSolve,
                                                    w.Wm + (size_t)CsP * n2, w.Vm + (size_t)CsP * n2, n2,
                                                    w.lam + (size_t)CsP * nr);
    KCHECK();
  }
  gather_pages_kernel<<<g1(n2, CsAll), 256, 0, st>>>(n2, w.groupOf, P.dV0.as<cplx>(), n2, w.X2, n2);
  smatrix_layer(P, w, CsAll, w.Wm, w.Vm, w.lam, w.Sb[0], w.Sb[1], CsP, false, w.pagesA);
  // layer order: Sa[0] = S11 (= S22), Sa[1] = S12 (= S21) of page (j, m, s) at j * CpS + m * nSec + s
  gather_pages_kernel<<<g1(n2, CsAll), 256, 0, st>>>(n2, w.idxA, w.Sb[0], n2, w.Sa[0], n2);
  gather_pages_kernel<<<g1(n2, CsAll), 256, 0, st>>>(n2, w.idxA, w.Sb[1], n2, w.Sa[1], n2);
  KCHECK();
  // binary tree of star products over the layers, in order
  cplx* cur[4] = {w.Sa[0], w.Sa[1], w.Sa[1], w.Sa[0]};
  bool inSa = true;
  int Kc = L;
  const size_t half = (size_t)(CsAll / 2) * n2;
  while (Kc > 1) {
    const int K = Kc / 2;
    std::vector<int> ev((size_t)K * CpS), od((size_t)K * CpS);
    for (int q = 0; q < K * CpS; ++q) {
      const int k = q / CpS, r = q % CpS;
      ev[q] = 2 * k * CpS + r;
      od[q] = (2 * k + 1) * CpS + r;
    }
    w.idxA = table(P, ev);
    w.idxB = table(P, od);
    cplx* A[4];
    cplx* B[4];
    cplx* out[4];
    for (int k = 0; k < 4; ++k) {
      A[k] = w.Sn[k];
      B[k] = w.Sn[k] + half;
      out[k] = inSa ? w.Sb[k] : w.Sa[k];
      gather_pages_kernel<<<g1(n2, K * CpS), 256, 0, st>>>(n2, w.idxA, cur[k], n2, A[k], n2);
      gather_pages_kernel<<<g1(n2, K * CpS), 256, 0, st>>>(n2, w.idxB, cur[k], n2, B[k], n2);
    }
    smatrix_star(P, w, K * CpS, A, B, out);
    if (Kc & 1)
      for (int k = 0; k < 4; ++k)
        HIPCHECK(hipMemcpyAsync(out[k] + (size_t)K * CpS * n2, cur[k] + (size_t)(Kc - 1) * CpS * n2,
                                sizeof(cplx) * CpS * n2, hipMemcpyDeviceToDevice, st));
    for (int k = 0; k < 4; ++k) cur[k] = out[k];
    inSa = !inSa;
    Kc = K + (Kc & 1);
  }
  // standard solve order for the chunk's last stages
  std::vector<DevSolve> sv0(CpS);
  std::vector<int> grp0(CpS);
  for (int m = 0; m < Cp; ++m)
    for (int sc = 0; sc < nSec; ++sc) {
      sv0[m * nSec + sc] = DevSolve{m, sc};
      grp0[m * nSec + sc] = P.pageCtx[p0 + m] * nSec + sc;
    }
  w.solves = table(P, sv0);
  w.groupOf = table(P, grp0);
  // result of S_ref * (S_1 * ... * S_L) * S_trn on the source columns; cur holds the
  // layers' composite (in one of the Sa / Sb sets), the regions go to Sn and the other set
  cplx* Ref[4];
  cplx* Trn[4];
  for (int k = 0; k < 4; ++k) {
    Ref[k] = w.Sn[k];
    Trn[k] = inSa ? w.Sb[k] : w.Sa[k];
  }
  (void)G;
  finish_chunk(P, w, p0, Cp, Ref, Trn, w.X1, w.X2, cur);
}

// ---------------------------------------------------------------- execution
// The GPU work of one chunk: kernels and device-to-device copies only (tables
// come from the plan's table cache), so that it can be captured into a graph.
// With a progress callback it waits for the device after every layer and calls
// back; a nonzero return cancels.
static void execute_chunk(orf_plan& P, int ch, orf_progress_fn progress, void* user) {
  hipStream_t st = P.st;
  const int npages = P.nPages;
  const int Cmax = P.chunkPages;
  Work& w = P.work;
  const int nChunks = (npages + Cmax - 1) / Cmax;
  const int M = P.M, nr = P.nr, nSec = P.nSec;
  const long long n2 = (long long)nr * nr;
  {
    const int p0 = ch * Cmax, Cp = std::min(Cmax, npages - p0), Cs = Cp * nSec;
    if (P.layerParallel && !progress) {
      execute_chunk_layers(P, w, p0, Cp);
      return;
    }
    // solve pages: sp = mode * nSec + sector
    std::vector<DevSolve> sv(Cs);
    std::vector<int> grp(Cs);
    for (int m = 0; m < Cp; ++m)
      for (int s = 0; s < nSec; ++s) {
        sv[m * nSec + s] = DevSolve{m, s};
        grp[m * nSec + s] = P.pageCtx[p0 + m] * nSec + s;
      }
    w.solves = table(P, sv);
    w.groupOf = table(P, grp);
    const long long G = (long long)P.nCtx * nSec;
    // with layers: L = S_1 * ... * S_L accumulates in Sa and the regions join at the
    // end (finish_stack); without: Sa = S_ref and the column star with S_trn
    if (P.nLayers == 0)
      for (int k = 0; k < 4; ++k)
        gather_pages_kernel<<<g1(n2, Cs), 256, 0, st>>>(n2, w.groupOf, P.dSref.as<cplx>() + k * n2 * G, n2, w.Sa[k], n2);
    KCHECK();
    for (int j = 0; j < P.nLayers; ++j) {
      const DevPage* pages = P.dPagesByLayer[j]->as<DevPage>() + p0;
      if (P.layerAnyPat[j]) {
        const cplx* se = P.dSlotEps[j]->as<cplx>() + (size_t)p0 * P.nslotMax;
        const cplx* si = P.dSlotInv[j]->as<cplx>() + (size_t)p0 * P.nslotMax;
        // distinct (label grid, slot permittivities) pairs of the chunk, first occurrence order
        const int ns = P.nslotMax;
        std::vector<DevPage> uPg;
        std::vector<cplx> uSe, uSi;
        std::vector<int> uMap(Cp);
        {
          std::map<std::vector<double>, int> seen;
          for (int m = 0; m < Cp; ++m) {
            const DevPage& pg = P.hPages[j][p0 + m];
            std::vector<double> key{(double)pg.grid, (double)pg.nslot};
            for (int q = 0; q < pg.nslot; ++q) {
              const cplx e = P.hSe[j][(size_t)(p0 + m) * ns + q];
              key.push_back(e.x);
              key.push_back(e.y);
            }
            auto it = seen.find(key);
            if (it == seen.end()) {
              it = seen.emplace(key, (int)uPg.size()).first;
              uPg.push_back(pg);
              for (int q = 0; q < ns; ++q) {
                uSe.push_back(P.hSe[j][(size_t)(p0 + m) * ns + q]);
                uSi.push_back(P.hSi[j][(size_t)(p0 + m) * ns + q]);
              }
            }
            uMap[m] = it->second;
          }
        }
        ConvUnique cu;
        cu.Cu = (int)uPg.size();
        if (cu.Cu < Cp) {
          w.pagesA = table(P, uPg);
          w.seA = table(P, uSe);
          w.siA = table(P, uSi);
          w.idxA = table(P, uMap);
          cu.uPages = w.pagesA;
          cu.uSe = w.seA;
          cu.uSi = w.siA;
          cu.uMap = w.idxA;
        }
        patterned_modes(P, w, pages, se, si, Cp, Cs, P.keepModes ? P.kept[j] : nullptr, &cu);
      } else {
        STAGE("modes");
        std::vector<cplx> eps(Cs);
        for (int sp = 0; sp < Cs; ++sp) eps[sp] = C(P.layerBgEps[j][p0 + sp / nSec]);
        w.epsSolve = table(P, eps);
        homog_modes_kernel<<<g1(n2, Cs), 256, 0, st>>>(M, nr, P.reduced, w.solves, pages, P.dSectors.as<DevSector>(),
                                                       P.dKxy.as<double>(), w.epsSolve, w.Wm, w.Vm, n2, w.lam);
        KCHECK();
      }
      // layer S-matrix against the gap (W0 = I, V0 per context and sector)
      std::vector<double> kL(Cs);
      for (int sp = 0; sp < Cs; ++sp) {
        const int pg = p0 + sp / nSec;
        kL[sp] = P.k0[P.pageCtx[pg]] * P.thick[(size_t)j * P.B + P.pageB[pg]];
      }
      w.kL = table(P, kL);
      if (!P.directSingle)
        gather_pages_kernel<<<g1(n2, Cs), 256, 0, st>>>(n2, w.groupOf, P.dV0.as<cplx>(), n2, w.X2, n2);
      else if (P.reduced)  // finish_direct: U = V0^-1 V
        gather_pages_kernel<<<g1(n2, Cs), 256, 0, st>>>(n2, w.groupOf, P.dV0inv.as<cplx>(), n2, w.X2, n2);
      if (P.keepModes) {
        auto* K = P.kept[j];
        K->W.alloc(sizeof(cplx) * n2);
        K->V.alloc(sizeof(cplx) * n2);
        K->lam.alloc(sizeof(cplx) * nr);
        HIPCHECK(hipMemcpyAsync(K->W.p, w.Wm, sizeof(cplx) * n2, hipMemcpyDeviceToDevice, st));
        HIPCHECK(hipMemcpyAsync(K->V.p, w.Vm, sizeof(cplx) * n2, hipMemcpyDeviceToDevice, st));
        HIPCHECK(hipMemcpyAsync(K->lam.p, w.lam, sizeof(cplx) * nr, hipMemcpyDeviceToDevice, st));
      }
      const bool single = P.nLayers == 1 && !P.keepModes;  // finish_single takes D, E1, E2
      if (!P.directSingle)  // finish_direct needs no layer S-matrix
        smatrix_layer(P, w, Cs, w.Wm, w.Vm, w.lam, w.Sb[0], w.Sb[1], P.layerAnyPat[j] ? Cs : 0, single,
                      P.layerAnyPat[j] ? pages : nullptr);
      if (!single) {
        if (P.keepModes) {
          auto* K = P.kept[j];
          K->S11.alloc(sizeof(cplx) * n2);
          K->S12.alloc(sizeof(cplx) * n2);
          HIPCHECK(hipMemcpyAsync(K->S11.p, w.Sb[0], sizeof(cplx) * n2, hipMemcpyDeviceToDevice, st));
          HIPCHECK(hipMemcpyAsync(K->S12.p, w.Sb[1], sizeof(cplx) * n2, hipMemcpyDeviceToDevice, st));
        }
        cplx* SB[4] = {w.Sb[0], w.Sb[1], w.Sb[1], w.Sb[0]};
        if (j == 0) {
          for (int k = 0; k < 4; ++k)
            HIPCHECK(hipMemcpyAsync(w.Sa[k], SB[k], sizeof(cplx) * n2 * Cs, hipMemcpyDeviceToDevice, st));
        } else {
          smatrix_star(P, w, Cs, w.Sa, SB, w.Sn);
          for (int k = 0; k < 4; ++k) std::swap(w.Sa[k], w.Sn[k]);
        }
      }
      if (progress) {
        HIPCHECK(hipStreamSynchronize(st));
        const double frac = (double)(ch * P.nLayers + j + 1) / (double)(nChunks * P.nLayers);
        char msg[64];
        snprintf(msg, sizeof msg, "layer %d/%d", j + 1, P.nLayers);
        if (progress(frac, msg, user)) {
          HIPCHECK(hipStreamSynchronize(st));
          fail(ORF_CANCELLED, "cancelled by the progress callback");
        }
      }
    }
    if (P.nLayers == 0) finish_chunk(P, w, p0, Cp, w.Sa, w.Sb, w.Sn[0], w.Sn[2]);
    else if (P.nLayers == 1 && !P.keepModes) finish_chunk(P, w, p0, Cp, w.Sn, w.Sb, w.X1, w.X2, nullptr, true);
    else finish_chunk(P, w, p0, Cp, w.Sn, w.Sb, w.X1, w.X2, w.Sa);
    if (progress && P.nLayers == 0) {
      HIPCHECK(hipStreamSynchronize(st));
      if (progress((double)(ch + 1) / nChunks, "done", user)) fail(ORF_CANCELLED, "cancelled by the progress callback");
    }
  }
}

static void plan_execute(orf_plan& P, orf_progress_fn progress, void* user) {
  hipStream_t st = P.st;
  const int nChunks = (P.nPages + P.chunkPages - 1) / P.chunkPages;
  if (P.timer.on) P.timer.reset();
  if (P.keepModes && P.kept.empty())
    for (int j = 0; j < P.nLayers; ++j) P.kept.push_back(new orf_plan::Kept());
  P.tableCursor = 0;
  P.pendingInfo.clear();
  P.infoCursor = 0;
  // graphs: replay the captured chunks; capture them on the first execute after the
  // tables exist (progress callbacks, stage timing and kept modes run eagerly)
  if (P.graphsStale) {
    HIPCHECK(hipStreamSynchronize(st));
    for (auto& g : P.graphExec)
      if (g) (void)hipGraphExecDestroy(g);
    P.graphExec.clear();
    P.graphInfo.clear();
    P.graphsStale = false;
  }
  const bool graphs = P.useGraphs && !progress && !P.timer.on && !P.keepModes && !P.tables.empty();
  if (graphs && (int)P.graphExec.size() < nChunks) {
    P.graphExec.assign(nChunks, nullptr);
    P.graphInfo.assign(nChunks, {});
  }
  if ((int)P.chunkCursor.size() < nChunks) P.chunkCursor.assign(nChunks, -1);
  for (int ch = 0; ch < nChunks; ++ch) {
    // each chunk walks its own run of the table cache
    if (P.chunkCursor[ch] >= 0) P.tableCursor = (size_t)P.chunkCursor[ch];
    else P.chunkCursor[ch] = (int)P.tableCursor;
    if (graphs) {
      if (!P.graphExec[ch]) {
        hipGraph_t g = nullptr;
        HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        P.capturing = true;
        try {
          execute_chunk(P, ch, nullptr, nullptr);
        } catch (...) {
          P.capturing = false;
          hipGraph_t dead = nullptr;
          (void)hipStreamEndCapture(st, &dead);
          if (dead) (void)hipGraphDestroy(dead);
          throw;
        }
        P.capturing = false;
        HIPCHECK(hipStreamEndCapture(st, &g));
        hipGraphExec_t ge = nullptr;
        const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHECK(e);
        P.graphExec[ch] = ge;
        P.graphInfo[ch] = P.pendingInfo;
      } else {
        P.pendingInfo = P.graphInfo[ch];  // status regions the captured chunk writes
      }
      HIPCHECK(hipGraphLaunch(P.graphExec[ch], st));
    } else {
      execute_chunk(P, ch, progress, user);
    }
    HIPCHECK(hipStreamSynchronize(st));
    check_pending_info(P, P.work);
  }
  HIPCHECK(hipStreamSynchronize(st));
  P.timer.collect();
  // a singular interface system (or non-finite input) shows as a non-finite total
  {
    const size_t nout = (size_t)P.nPages * P.nPol;
    std::vector<double> rt(nout), tt(nout);
    HIPCHECK(hipMemcpyAsync(rt.data(), P.dRtot.p, sizeof(double) * nout, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(tt.data(), P.dTtot.p, sizeof(double) * nout, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    for (size_t q = 0; q < nout; ++q)
      if (!std::isfinite(rt[q]) || !std::isfinite(tt[q]))
        fail(ORF_DEVICE_ERROR, "non-finite efficiencies on page " + std::to_string(q % P.nPages) + ", source " +
                                   std::to_string(q / P.nPages) + " (singular interface system or non-finite input)");
  }
  P.haveResult = true;
}

static void plan_fetch(orf_plan& P, orf_result* out) {
  if (!P.haveResult) fail(ORF_INVALID_ARG, "plan has not been executed");
  if (!out) fail(ORF_INVALID_ARG, "result is NULL");
  const size_t nout = (size_t)P.nPages * P.nPol;
  if (out->m) for (int p = 0; p < P.M; ++p) out->m[p] = P.mIdx[p];
  if (out->n) for (int p = 0; p < P.M; ++p) out->n[p] = P.nIdx[p];
  if (out->R) HIPCHECK(hipMemcpy(out->R, P.dR.p, sizeof(double) * P.M * nout, hipMemcpyDeviceToHost));
  if (out->T) HIPCHECK(hipMemcpy(out->T, P.dT.p, sizeof(double) * P.M * nout, hipMemcpyDeviceToHost));
  if (out->Rtot) HIPCHECK(hipMemcpy(out->Rtot, P.dRtot.p, sizeof(double) * nout, hipMemcpyDeviceToHost));
  if (out->Ttot) HIPCHECK(hipMemcpy(out->Ttot, P.dTtot.p, sizeof(double) * nout, hipMemcpyDeviceToHost));
  if (out->r) HIPCHECK(hipMemcpy(out->r, P.dr.p, sizeof(cplx) * 3 * P.M * nout, hipMemcpyDeviceToHost));
  if (out->t) HIPCHECK(hipMemcpy(out->t, P.dt.p, sizeof(cplx) * 3 * P.M * nout, hipMemcpyDeviceToHost));
}

// ---------------------------------------------------------------- fields (rcwaFields.m:1-21)
// Every region is expanded in its own modes: s = [Ex; Ey] = W (fp a+ + fm a-),
// u = [H~x; H~y] = V (-fp a+ + fm a-), fp = e^{-lam k0 z}, fm = e^{-lam k0 (d - z)},
// z from the top of the region (the convention of smatrixLayer.m:6-18).  The
// amplitudes come from the zero-thickness gaps between layers: with S_up the
// stack above a gap and S_down the stack below, c+ = (I - S_up22 S_down11)^-1
// S_up21 c_inc and c- = S_down11 c+; a layer's a+ follows from the gap above
// and a- from the gap below, so no growing exponential is formed.
// Ez = i [eps]^-1 (Kx H~y - Ky H~x) (pqMatrices.m:4), H~z = i (Kx Ey - Ky Ex);
// the (x, y) synthesis is one ZGEMM with the phase matrix e^{i k0 (kx x + ky y)}.
static void plan_fields(orf_plan& P, int nx, const double* x, int ny, const double* y, int nz, const double* z,
                        double* Eout, double* Hout, double* fluxOut) {
  if (!P.keepModes || !P.haveResult) fail(ORF_INVALID_ARG, "fields need an executed plan with keepModes");
  if (nx <= 0 || ny <= 0 || nz <= 0 || !x || !y || !z) fail(ORF_INVALID_ARG, "empty field grid");
  const int M = P.M, n = P.nf, nL = P.nLayers;
  const long long n2 = (long long)n * n;
  hipStream_t st = P.st;
  Work& w = P.work;
  const double k0 = P.k0[0];
  const double* kx = P.dKxy.as<double>();
  const double* ky = kx + M;
  // half-space modes (V, lam) of the incident and substrate media
  DBuf dsv, dpg, deps, dWr, dVr, dLr;
  dsv.upload(std::vector<DevSolve>{DevSolve{0, 0}, DevSolve{0, 0}});
  dpg.upload(std::vector<DevPage>{DevPage{0, 0, -1, 0}});
  deps.upload(std::vector<cplx>{C(P.epsInc0), C(P.epsSub0)});
  dWr.alloc(sizeof(cplx) * n2 * 2);
  dVr.alloc(sizeof(cplx) * n2 * 2);
  dLr.alloc(sizeof(cplx) * n * 2);
  homog_modes_kernel<<<g1(n2, 2), 256, 0, st>>>(M, n, 0, dsv.as<DevSolve>(), dpg.as<DevPage>(),
                                                P.dSectors.as<DevSector>(), kx, deps.as<cplx>(), dWr.as<cplx>(),
                                                dVr.as<cplx>(), n2, dLr.as<cplx>());
  KCHECK();
  // prefix (up) and suffix (down) stacks around every gap
  DBuf chain;
  chain.alloc(sizeof(cplx) * n2 * 8 * (nL + 1));
  auto blk = [&](int which, int j, int k) { return chain.as<cplx>() + ((size_t)(which * (nL + 1) + j) * 4 + k) * n2; };
  for (int k = 0; k < 4; ++k) {
    HIPCHECK(hipMemcpyAsync(blk(0, 0, k), P.dSref.as<cplx>() + k * n2, sizeof(cplx) * n2, hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipMemcpyAsync(blk(1, nL, k), P.dStrn.as<cplx>() + k * n2, sizeof(cplx) * n2, hipMemcpyDeviceToDevice, st));
  }
  for (int j = 0; j < nL; ++j) {
    cplx* SA[4] = {blk(0, j, 0), blk(0, j, 1), blk(0, j, 2), blk(0, j, 3)};
    cplx* SL[4] = {P.kept[j]->S11.as<cplx>(), P.kept[j]->S12.as<cplx>(), P.kept[j]->S12.as<cplx>(),
                   P.kept[j]->S11.as<cplx>()};
    cplx* O[4] = {blk(0, j + 1, 0), blk(0, j + 1, 1), blk(0, j + 1, 2), blk(0, j + 1, 3)};
    smatrix_star(P, w, 1, SA, SL, O);
  }
  for (int j = nL - 1; j >= 0; --j) {
    cplx* SL[4] = {P.kept[j]->S11.as<cplx>(), P.kept[j]->S12.as<cplx>(), P.kept[j]->S12.as<cplx>(),
                   P.kept[j]->S11.as<cplx>()};
    cplx* SB[4] = {blk(1, j + 1, 0), blk(1, j + 1, 1), blk(1, j + 1, 2), blk(1, j + 1, 3)};
    cplx* O[4] = {blk(1, j, 0), blk(1, j, 1), blk(1, j, 2), blk(1, j, 3)};
    smatrix_star(P, w, 1, SL, SB, O);
  }
  // gap amplitudes: c+ = (I - up22 down11)^-1 up21 c_inc, c- = down11 c+
  std::vector<cplx> cinc(n, C(0));
  cinc[P.i0] = P.sources[0].ex;
  cinc[M + P.i0] = P.sources[0].ey;
  DBuf dcinc, gap;
  dcinc.upload(cinc);
  gap.alloc(sizeof(cplx) * n * 2 * (nL + 1));  // [c+ | c-] per gap
  for (int j = 0; j <= nL; ++j) {
    cplx* cp = gap.as<cplx>() + (size_t)2 * j * n;
    cplx* cm = cp + n;
    eye_axpby_kernel<<<g1(n2, 1), 256, 0, st>>>(n, ONE, ZERO, nullptr, 0, w.M1, n2);
    orf_zgemm(st, 'N', 'N', n, n, n, MONE, blk(0, j, 3), n, n2, blk(1, j, 0), n, n2, ONE, w.M1, n, n2, 1);
    orf_zgetrf(st, n, w.M1, n, n2, w.ipiv, nullptr, 1, w.dinv);
    orf_zgemm(st, 'N', 'N', n, 1, n, ONE, blk(0, j, 2), n, n2, dcinc.as<cplx>(), n, n, ZERO, cp, n, n, 1);
    orf_zgetrs(st, 'L', n, 1, w.M1, n, n2, w.ipiv, w.dinv, cp, n, n, 1);
    orf_zgemm(st, 'N', 'N', n, 1, n, ONE, blk(1, j, 0), n, n2, cp, n, n, ZERO, cm, n, n, 1);
  }
  // reflected and transmitted amplitudes: c_ref = Sref11 c_inc + Sref12 c-_0, c_trn = Strn21 c+_nL
  DBuf cref, ctrn;
  cref.alloc(sizeof(cplx) * n);
  ctrn.alloc(sizeof(cplx) * n);
  orf_zgemm(st, 'N', 'N', n, 1, n, ONE, P.dSref.as<cplx>(), n, n2, dcinc.as<cplx>(), n, n, ZERO, cref.as<cplx>(), n, n, 1);
  orf_zgemm(st, 'N', 'N', n, 1, n, ONE, P.dSref.as<cplx>() + n2, n, n2, gap.as<cplx>() + n, n, n, ONE, cref.as<cplx>(), n,
            n, 1);
  orf_zgemm(st, 'N', 'N', n, 1, n, ONE, P.dStrn.as<cplx>() + 2 * n2, n, n2, gap.as<cplx>() + (size_t)2 * nL * n, n, n, ZERO,
            ctrn.as<cplx>(), n, n, 1);
  // layer amplitudes: a+ = (W^-1 (c+ + c-) + V^-1 V0 (c+ - c-)) / 2 from the gap above,
  //                   a- = (W^-1 (d+ + d-) + V^-1 V0 (d- - d+)) / 2 from the gap below
  std::vector<cplx> hgap((size_t)2 * n * (nL + 1));
  HIPCHECK(hipMemcpyAsync(hgap.data(), gap.p, sizeof(cplx) * hgap.size(), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  auto Z = [](cplx a) { return zc(a.x, a.y); };
  DBuf amps, rhs, rhs2;
  amps.alloc(sizeof(cplx) * n * 2 * std::max(nL, 1));  // [a+ | a-] per layer
  rhs.alloc(sizeof(cplx) * n * 2);
  rhs2.alloc(sizeof(cplx) * n * 2);
  for (int j = 0; j < nL; ++j) {
    const cplx* cp = hgap.data() + (size_t)2 * j * n;
    const cplx* cm = cp + n;
    const cplx* dp = hgap.data() + (size_t)2 * (j + 1) * n;
    const cplx* dm = dp + n;
    std::vector<cplx> sum(2 * n), dif(2 * n);
    for (int r = 0; r < n; ++r) {
      sum[r] = C(Z(cp[r]) + Z(cm[r]));
      sum[n + r] = C(Z(dp[r]) + Z(dm[r]));
      dif[r] = C(Z(cp[r]) - Z(cm[r]));
      dif[n + r] = C(Z(dm[r]) - Z(dp[r]));
    }
    HIPCHECK(hipMemcpyAsync(rhs.p, sum.data(), sizeof(cplx) * 2 * n, hipMemcpyHostToDevice, st));
    HIPCHECK(hipMemcpyAsync(w.LUb, P.kept[j]->W.p, sizeof(cplx) * n2, hipMemcpyDeviceToDevice, st));
    orf_zgetrf(st, n, w.LUb, n, n2, w.ipiv, nullptr, 1, w.dinv);
    orf_zgetrs(st, 'L', n, 2, w.LUb, n, n2, w.ipiv, w.dinv, rhs.as<cplx>(), n, 2 * n, 1);
    HIPCHECK(hipMemcpyAsync(w.Dm, dif.data(), sizeof(cplx) * 2 * n, hipMemcpyHostToDevice, st));
    orf_zgemm(st, 'N', 'N', n, 2, n, ONE, P.dV0.as<cplx>(), n, n2, w.Dm, n, 2 * n, ZERO, rhs2.as<cplx>(), n, 2 * n, 1);
    HIPCHECK(hipMemcpyAsync(w.LUb, P.kept[j]->V.p, sizeof(cplx) * n2, hipMemcpyDeviceToDevice, st));
    orf_zgetrf(st, n, w.LUb, n, n2, w.ipiv, nullptr, 1, w.dinv);
    orf_zgetrs(st, 'L', n, 2, w.LUb, n, n2, w.ipiv, w.dinv, rhs2.as<cplx>(), n, 2 * n, 1);
    std::vector<cplx> h1(2 * n), h2(2 * n), a(2 * n);
    HIPCHECK(hipMemcpyAsync(h1.data(), rhs.p, sizeof(cplx) * 2 * n, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(h2.data(), rhs2.p, sizeof(cplx) * 2 * n, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    for (int r = 0; r < 2 * n; ++r) a[r] = C(0.5 * (Z(h1[r]) + Z(h2[r])));
    HIPCHECK(hipMemcpyAsync(amps.as<cplx>() + (size_t)2 * j * n, a.data(), sizeof(cplx) * 2 * n, hipMemcpyHostToDevice,
                            st));
    HIPCHECK(hipStreamSynchronize(st));
  }
  // regions of the field points: 0 incident, 1 + j layer j, nL + 1 substrate
  std::vector<double> ztop(nL + 1, 0.0);
  for (int j = 0; j < nL; ++j) ztop[j + 1] = ztop[j] + P.thick[(size_t)j * P.B + P.pageB[0]];
  std::vector<std::vector<int>> idx(nL + 2);
  std::vector<std::vector<double>> zl(nL + 2);
  for (int k = 0; k < nz; ++k) {
    int reg;
    double zz;
    if (z[k] < 0) { reg = 0; zz = z[k]; }
    else if (z[k] >= ztop[nL]) { reg = nL + 1; zz = z[k] - ztop[nL]; }
    else {
      const int j = (int)(std::upper_bound(ztop.begin(), ztop.end(), z[k]) - ztop.begin()) - 1;
      reg = 1 + j;
      zz = z[k] - ztop[j];
    }
    idx[reg].push_back(k);
    zl[reg].push_back(zz);
  }
  DBuf Sv, Uv, Ez, Pp, Pm, Tq, dz, di;
  Sv.alloc(sizeof(cplx) * n * nz);
  Uv.alloc(sizeof(cplx) * n * nz);
  Ez.alloc(sizeof(cplx) * M * nz);
  Pp.alloc(sizeof(cplx) * n * nz);
  Pm.alloc(sizeof(cplx) * n * nz);
  Tq.alloc(sizeof(cplx) * n * nz);
  const cplx I1 = {0.0, 1.0};
  for (int reg = 0; reg < nL + 2; ++reg) {
    const int cnt = (int)idx[reg].size();
    if (!cnt) continue;
    dz.upload(zl[reg]);
    di.upload(idx[reg]);
    const cplx *lam, *ap, *am, *Wr = nullptr, *Vr;
    double d = 0.0;
    zc epsH = 0.0;
    const cplx* Econv = nullptr;
    if (reg == 0) {
      lam = dLr.as<cplx>(); ap = dcinc.as<cplx>(); am = cref.as<cplx>(); Vr = dVr.as<cplx>(); epsH = P.epsInc0;
    } else if (reg == nL + 1) {
      lam = dLr.as<cplx>() + n; ap = ctrn.as<cplx>(); am = nullptr; Vr = dVr.as<cplx>() + n2; epsH = P.epsSub0;
    } else {
      const int j = reg - 1;
      auto* K = P.kept[j];
      lam = K->lam.as<cplx>(); ap = amps.as<cplx>() + (size_t)2 * j * n; am = ap + n;
      Wr = K->W.as<cplx>(); Vr = K->V.as<cplx>();
      d = ztop[j + 1] - ztop[j];
      if (K->patterned) Econv = K->E.as<cplx>();
      else epsH = P.layerBgEps[j][0];
    }
    field_amp_kernel<<<g1((long long)n * cnt), 256, 0, st>>>(n, cnt, lam, ap, am, k0, d, dz.as<double>(), Pp.as<cplx>(),
                                                             Pm.as<cplx>());
    // s = W Pp (W = I in the half-spaces), u = V Pm
    const cplx* sv = Pp.as<cplx>();
    if (Wr) {
      orf_zgemm(st, 'N', 'N', n, cnt, n, ONE, Wr, n, 0, Pp.as<cplx>(), n, 0, ZERO, Tq.as<cplx>(), n, 0, 1);
      sv = Tq.as<cplx>();
    }
    scatter_cols_kernel<<<g1((long long)n * cnt), 256, 0, st>>>(n, cnt, di.as<int>(), sv, n, Sv.as<cplx>(), n);
    orf_zgemm(st, 'N', 'N', n, cnt, n, ONE, Vr, n, 0, Pm.as<cplx>(), n, 0, ZERO, Tq.as<cplx>(), n, 0, 1);
    scatter_cols_kernel<<<g1((long long)n * cnt), 256, 0, st>>>(n, cnt, di.as<int>(), Tq.as<cplx>(), n, Uv.as<cplx>(), n);
    // Ez = i [eps]^-1 (Kx H~y - Ky H~x)
    curl_z_kernel<<<g1((long long)M * cnt), 256, 0, st>>>(M, cnt, kx, ky, Tq.as<cplx>(), I1, Pm.as<cplx>());
    if (Econv) {
      HIPCHECK(hipMemcpyAsync(w.LUb, Econv, sizeof(cplx) * M * M, hipMemcpyDeviceToDevice, st));
      orf_zgetrf(st, M, w.LUb, M, (long long)M * M, w.ipiv, nullptr, 1, w.dinv);
      orf_zgetrs(st, 'L', M, cnt, w.LUb, M, (long long)M * M, w.ipiv, w.dinv, Pm.as<cplx>(), M, (long long)M * cnt, 1);
    } else {
      scale_kernel<<<g1((long long)M * cnt), 256, 0, st>>>((long long)M * cnt, C(1.0 / epsH), Pm.as<cplx>());
    }
    scatter_cols_kernel<<<g1((long long)M * cnt), 256, 0, st>>>(M, cnt, di.as<int>(), Pm.as<cplx>(), M, Ez.as<cplx>(), M);
    KCHECK();
    HIPCHECK(hipStreamSynchronize(st));  // dz / di are reused by the next region
  }
  // coefficient matrix C (M x 6 nz): columns c * nz + k for Ex, Ey, Ez, H~x, H~y, H~z
  DBuf Cf, Phi, F, dx, dy, seq;
  Cf.alloc(sizeof(cplx) * M * 6 * nz);
  std::vector<int> ks(nz);
  for (int k = 0; k < nz; ++k) ks[k] = k;
  seq.upload(ks);
  const long long Mz = (long long)M * nz;
  scatter_cols_kernel<<<g1(Mz), 256, 0, st>>>(M, nz, seq.as<int>(), Sv.as<cplx>(), n, Cf.as<cplx>() + 0 * Mz, M);
  scatter_cols_kernel<<<g1(Mz), 256, 0, st>>>(M, nz, seq.as<int>(), Sv.as<cplx>() + M, n, Cf.as<cplx>() + 1 * Mz, M);
  copy_cols_kernel<<<g1(Mz), 256, 0, st>>>(Mz, Ez.as<cplx>(), Cf.as<cplx>() + 2 * Mz);
  scatter_cols_kernel<<<g1(Mz), 256, 0, st>>>(M, nz, seq.as<int>(), Uv.as<cplx>(), n, Cf.as<cplx>() + 3 * Mz, M);
  scatter_cols_kernel<<<g1(Mz), 256, 0, st>>>(M, nz, seq.as<int>(), Uv.as<cplx>() + M, n, Cf.as<cplx>() + 4 * Mz, M);
  curl_z_kernel<<<g1(Mz), 256, 0, st>>>(M, nz, kx, ky, Sv.as<cplx>(), I1, Cf.as<cplx>() + 5 * Mz);  // H~z
  // synthesis on the grid
  const long long nxy = (long long)nx * ny;
  dx.upload(std::vector<double>(x, x + nx));
  dy.upload(std::vector<double>(y, y + ny));
  Phi.alloc(sizeof(cplx) * nxy * M);
  phase_kernel<<<g1(nxy * M), 256, 0, st>>>(M, nx, ny, dx.as<double>(), dy.as<double>(), kx, ky, k0, Phi.as<cplx>());
  F.alloc(sizeof(cplx) * nxy * 6 * nz);
  orf_zgemm(st, 'N', 'N', (int)nxy, 6 * nz, M, ONE, Phi.as<cplx>(), (int)nxy, 0, Cf.as<cplx>(), M, 0, ZERO,
            F.as<cplx>(), (int)nxy, 0, 1);
  // flux through the cell at every z
  DBuf fl;
  fl.alloc(sizeof(double) * nz);
  flux_kernel<<<nz, 256, 0, st>>>(M, Sv.as<cplx>(), Uv.as<cplx>(), 1.0 / P.kzi0, fl.as<double>());
  KCHECK();
  const size_t half = sizeof(cplx) * nxy * 3 * nz;
  if (Eout) HIPCHECK(hipMemcpyAsync(Eout, F.p, half, hipMemcpyDeviceToHost, st));
  if (Hout) HIPCHECK(hipMemcpyAsync(Hout, (char*)F.p + half, half, hipMemcpyDeviceToHost, st));
  if (fluxOut) HIPCHECK(hipMemcpyAsync(fluxOut, fl.p, sizeof(double) * nz, hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
}

// ====================================================================== allocation-free solves
// Device scratch of the allocation-free entry point: each GPU the library runs
// on holds its own copy of this array until the process ends.  Host code never
// passes `scratch` itself to a kernel or a copy (in host code the name is a
// stand-in, not a device address): orf_init looks up each device's address
// once, with that device current, and the entry point uses that address.
static constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static std::mutex g_scratch_mu;
static std::vector<void*> g_scratch_ptr;  // [device] -> address of that device's scratch

static void scratch_setup(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
  if (g_scratch_ptr[device]) return;
  void* ptr = nullptr;
  HIPCHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) fail(ORF_DEVICE_ERROR, "no device address for the scratch array");
  g_scratch_ptr[device] = ptr;
}

static void* scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= (int)g_scratch_ptr.size() || !g_scratch_ptr[device])
    fail(ORF_INVALID_ARG, "context not initialised by orf_init");
  return g_scratch_ptr[device];
}

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// ====================================================================== C ABI
#define ORF_TRY(...)                         \
  try {                                      \
    __VA_ARGS__;                             \
    return ORF_OK;                           \
  } catch (const OrfError& e) {              \
    g_err = e.msg;                           \
    return e.code;                           \
  } catch (const std::bad_alloc&) {          \
    g_err = "host out of memory";            \
    return ORF_OOM;                          \
  } catch (const std::exception& e) {        \
    g_err = e.what();                        \
    return ORF_DEVICE_ERROR;                 \
  }

extern "C" {

const char* orf_last_error(void) { return g_err.c_str(); }

int orf_init(int device, orf_ctx** out) {
  ORF_TRY({
    if (!out) fail(ORF_INVALID_ARG, "ctx is NULL");
    int count = 0;
    HIPCHECK(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) fail(ORF_DEVICE_ERROR, "no such HIP device");
    hipDeviceProp_t prop;
    HIPCHECK(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      fail(ORF_DEVICE_ERROR, std::string("device is ") + prop.gcnArchName + ", the library is compiled for gfx950 only");
    HIPCHECK(hipSetDevice(device));
    scratch_setup(device);
    orf_ctx* c = new orf_ctx();
    c->device = device;
    HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    // leave room for each kernel's static LDS
    const int lim = (int)std::min<size_t>(prop.sharedMemPerBlock, 160 * 1024) - 20 * 1024;
    orf_set_lds_limit(std::max(lim, 32 * 1024));
    *out = c;
  })
}

void orf_release(orf_ctx* ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int orf_harmonic_count(const orf_problem* p) {
  if (!p) return 0;
  const int Ny = std::isfinite(p->Ly) ? p->Ny : 0;
  return (2 * p->Nx + 1) * (2 * Ny + 1);
}

int orf_plan_create(orf_ctx* ctx, const orf_problem* p, orf_plan** out) {
  orf_plan* P = nullptr;
  try {
    if (!ctx || !out) fail(ORF_INVALID_ARG, "ctx or plan is NULL");
    HIPCHECK(hipSetDevice(ctx->device));
    P = new orf_plan();
    plan_create(ctx, p, *P);
    *out = P;
    return ORF_OK;
  } catch (const OrfError& e) {
    delete P;
    g_err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    delete P;
    g_err = e.what();
    return ORF_DEVICE_ERROR;
  }
}

int orf_plan_execute(orf_plan* plan, orf_progress_fn progress, void* user) {
  ORF_TRY({
    if (!plan) fail(ORF_INVALID_ARG, "plan is NULL");
    HIPCHECK(hipSetDevice(plan->ctx->device));
    plan_execute(*plan, progress, user);
  })
}

int orf_plan_fetch(orf_plan* plan, orf_result* out) {
  ORF_TRY({
    if (!plan) fail(ORF_INVALID_ARG, "plan is NULL");
    plan_fetch(*plan, out);
  })
}

void orf_plan_destroy(orf_plan* plan) { delete plan; }

int orf_plan_info(const orf_plan* plan, int32_t* solved_size, int32_t* symmetric, int32_t* n_solve_pages) {
  if (!plan) return ORF_INVALID_ARG;
  if (solved_size) *solved_size = plan->nr;
  if (symmetric) *symmetric = plan->reduced;
  if (n_solve_pages) *n_solve_pages = plan->nPages * plan->nSec;
  return ORF_OK;
}

int orf_plan_memory(const orf_plan* plan, int64_t* bytes) {
  if (!plan || !bytes) return ORF_INVALID_ARG;
  *bytes = (int64_t)plan_bytes(*plan);
  return ORF_OK;
}

int orf_plan_estimate(const orf_problem* p, int32_t reduced, int64_t* bytes) {
  ORF_TRY({
    if (!p || !bytes) fail(ORF_INVALID_ARG, "problem or bytes is NULL");
    orf_plan P;
    P.Nx = p->Nx;
    P.Ny = std::isfinite(p->Ly) ? p->Ny : 0;
    P.Nsx = p->Nsx;
    P.Nsy = std::isfinite(p->Ly) ? p->Nsy : 1;
    P.M = (2 * P.Nx + 1) * (2 * P.Ny + 1);
    P.nf = 2 * P.M;
    P.factorization = p->factorization;
    if (reduced) {
      // mirror-parity sectors; a y-invariant grating at most its TE / TM split
      P.nr = P.Ny == 0 ? P.M : (P.Nx + 1) * (P.Ny + 1) + P.Nx * P.Ny;
      bool s = false, pp = false;
      for (int i = 0; i < p->nPol; ++i) { s |= p->polKind[i] == ORF_POL_S; pp |= p->polKind[i] == ORF_POL_P; }
      P.nSec = (int)s + (int)pp;
      if (P.nSec == 0) P.nSec = 1;
    } else {
      P.nr = P.nf;
      P.nSec = 1;
    }
    const size_t pages = p->nPages > 0 ? (size_t)p->nPages : (size_t)p->nStruct * p->nLambda;
    const size_t n2 = (size_t)P.nr * P.nr;
    // workspace: one chunk (all pages; with several layers, every layer of at most
    // kLayerChunkSolves solve pages' worth of pages)
    size_t work = pages;
    if (p->nLayers >= 2 && !p->keepModes) {
      const size_t cap = (size_t)std::max(1, kLayerChunkSolves / (p->nLayers * P.nSec));
      work = std::min(pages, cap) * p->nLayers;
    }
    size_t b = work * (per_mode_bytes(P) + (size_t)P.nSec * per_solve_bytes(P));
    b += pages * P.nSec * 9 * n2 * sizeof(cplx);          // half-space S-matrices and gap modes (upper bound)
    b += (size_t)p->nStruct * p->nLayers * ((size_t)P.Nsx * P.Nsy +
         sizeof(cplx) * ORF_MAX_SLOTS * ((size_t)(4 * P.Nx + 1) * (P.Nsy + 4 * P.Ny + 1) + (size_t)(4 * P.Ny + 1) * P.Nsx));
    b += pages * p->nPol * (size_t)P.M * (2 * sizeof(double) + 6 * sizeof(cplx));
    *bytes = (int64_t)b;
  })
}

void orf_release_cached_memory(void) { g_pool.trim(); }

int orf_fields(orf_ctx* ctx, const orf_problem* p, int32_t nx, const double* x, int32_t ny, const double* y,
               int32_t nz, const double* z, double* E, double* H, double* flux, orf_result* out) {
  orf_plan* P = nullptr;
  try {
    if (!ctx || !p) fail(ORF_INVALID_ARG, "ctx or problem is NULL");
    if (p->nStruct != 1 || p->nLambda != 1 || p->nPol != 1 || p->nPages > 1)
      fail(ORF_INVALID_ARG, "fields take one structure, one wavelength and one source");
    HIPCHECK(hipSetDevice(ctx->device));
    orf_problem q = *p;
    q.keepModes = 1;
    q.nPages = 0;
    P = new orf_plan();
    plan_create(ctx, &q, *P);
    plan_execute(*P, nullptr, nullptr);
    if (out) plan_fetch(*P, out);
    plan_fields(*P, nx, x, ny, y, nz, z, E, H, flux);
    delete P;
    return ORF_OK;
  } catch (const OrfError& e) {
    delete P;
    g_err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    delete P;
    g_err = e.what();
    return ORF_DEVICE_ERROR;
  }
}

int orf_tune_set(int32_t kind, int32_t n, int32_t large_batch, int32_t value) {
  if (kind < 0 || kind >= ORF_TUNE_KINDS || n <= 0) return ORF_INVALID_ARG;
  orf_tune_put(kind, n, large_batch ? 1 : 0, value);
  return ORF_OK;
}

int orf_tune_force(int32_t kind, int32_t value) {
  if (kind < 0 || kind >= ORF_TUNE_KINDS) return ORF_INVALID_ARG;
  orf_tune_override(kind, value);
  return ORF_OK;
}

void orf_tune_clear(void) { orf_tune_reset(); }

int orf_tune_load(const char* path) {
  ORF_TRY({
    if (!path) fail(ORF_INVALID_ARG, "path is NULL");
    FILE* f = std::fopen(path, "r");
    if (!f) fail(ORF_INVALID_ARG, std::string("cannot open ") + path);
    char line[256];
    int lineNo = 0;
    while (std::fgets(line, sizeof line, f)) {
      ++lineNo;
      if (line[0] == '#' || line[0] == '\n') continue;
      int kind, n, large, value;
      if (std::sscanf(line, "%d %d %d %d", &kind, &n, &large, &value) != 4 || kind < 0 || kind >= ORF_TUNE_KINDS ||
          n <= 0) {
        std::fclose(f);
        fail(ORF_INVALID_ARG, std::string(path) + ": bad entry on line " + std::to_string(lineNo));
      }
      orf_tune_put(kind, n, large ? 1 : 0, value);
    }
    std::fclose(f);
  })
}

void orf_plan_enable_graphs(orf_plan* plan, int32_t on) {
  if (plan) plan->useGraphs = on != 0;
}

void orf_plan_enable_timing(orf_plan* plan, int32_t on) {
  if (plan) plan->timer.on = on != 0;
}

int orf_plan_stage_times(const orf_plan* plan, int32_t max_stages, const char** names, double* total_ms,
                         int32_t* launches, int32_t* n_stages) {
  if (!plan) return ORF_INVALID_ARG;
  int k = 0;
  for (auto& nm : plan->timer.order) {
    if (k >= max_stages) break;
    auto it = plan->timer.acc.find(nm);
    if (names) names[k] = it->first.c_str();
    if (total_ms) total_ms[k] = it->second.first;
    if (launches) launches[k] = it->second.second;
    ++k;
  }
  if (n_stages) *n_stages = k;
  return ORF_OK;
}

int orf_rcwa2d(orf_ctx* ctx, const orf_problem* p, orf_result* out, orf_progress_fn progress, void* user) {
  orf_plan* plan = nullptr;
  int rc = orf_plan_create(ctx, p, &plan);
  if (rc != ORF_OK) return rc;
  rc = orf_plan_execute(plan, progress, user);
  if (rc == ORF_OK) rc = orf_plan_fetch(plan, out);
  orf_plan_destroy(plan);
  return rc;
}

// ---------------------------------------------------------------- page primitives
int orf_pagemtimes(orf_ctx* ctx, char opA, char opB, int32_t m, int32_t n, int32_t k, const double* A,
                   int64_t strideA, const double* B, int64_t strideB, double* Cout, int32_t batch) {
  ORF_TRY({
    if (!ctx || m <= 0 || n <= 0 || k <= 0 || batch <= 0) fail(ORF_INVALID_ARG, "bad sizes");
    HIPCHECK(hipSetDevice(ctx->device));
    const int lda = (opA == 'N') ? m : k, ldb = (opB == 'N') ? k : n;
    const size_t na = (size_t)(strideA ? batch : 1) * (size_t)m * k, nb = (size_t)(strideB ? batch : 1) * (size_t)k * n;
    DBuf a, b, c;
    a.alloc(na * 16);
    b.alloc(nb * 16);
    c.alloc((size_t)batch * m * n * 16);
    HIPCHECK(hipMemcpy(a.p, A, na * 16, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(b.p, B, nb * 16, hipMemcpyHostToDevice));
    orf_zgemm(ctx->stream, opA, opB, m, n, k, ONE, a.as<cplx>(), lda, strideA ? (long long)m * k : 0, b.as<cplx>(), ldb,
              strideB ? (long long)k * n : 0, ZERO, c.as<cplx>(), m, (long long)m * n, batch);
    KCHECK();
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    HIPCHECK(hipMemcpy(Cout, c.p, (size_t)batch * m * n * 16, hipMemcpyDeviceToHost));
  })
}

int orf_pagemtimes_f32(orf_ctx* ctx, char opA, char opB, int32_t m, int32_t n, int32_t k, const double* A,
                       int64_t strideA, const double* B, int64_t strideB, double* Cout, int32_t batch) {
  ORF_TRY({
    if (!ctx || m <= 0 || n <= 0 || k <= 0 || batch <= 0) fail(ORF_INVALID_ARG, "bad sizes");
    HIPCHECK(hipSetDevice(ctx->device));
    const int lda = (opA == 'N') ? m : k, ldb = (opB == 'N') ? k : n;
    const size_t na = (size_t)(strideA ? batch : 1) * (size_t)m * k, nb = (size_t)(strideB ? batch : 1) * (size_t)k * n;
    DBuf a, b, c;
    a.alloc(na * 16);
    b.alloc(nb * 16);
    c.alloc((size_t)batch * m * n * 16);
    HIPCHECK(hipMemcpy(a.p, A, na * 16, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(b.p, B, nb * 16, hipMemcpyHostToDevice));
    orf_zgemm_f32(ctx->stream, opA, opB, m, n, k, ONE, a.as<cplx>(), lda, strideA ? (long long)m * k : 0, b.as<cplx>(),
                  ldb, strideB ? (long long)k * n : 0, ZERO, c.as<cplx>(), m, (long long)m * n, batch);
    KCHECK();
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    HIPCHECK(hipMemcpy(Cout, c.p, (size_t)batch * m * n * 16, hipMemcpyDeviceToHost));
  })
}

int orf_pagemldivide(orf_ctx* ctx, int32_t n, int32_t nrhs, const double* A, int64_t strideA, const double* B,
                     double* X, int32_t batch) {
  ORF_TRY({
    if (!ctx || n <= 0 || nrhs <= 0 || batch <= 0) fail(ORF_INVALID_ARG, "bad sizes");
    HIPCHECK(hipSetDevice(ctx->device));
    const size_t n2 = (size_t)n * n;
    DBuf a, b, ip, dv;
    dv.alloc(16 * (size_t)orf_dinv_stride(n) * batch);
    a.alloc((size_t)batch * n2 * 16);
    b.alloc((size_t)batch * n * nrhs * 16);
    ip.alloc(sizeof(int) * (size_t)n * batch);
    for (int q = 0; q < batch; ++q)
      HIPCHECK(hipMemcpy(a.as<cplx>() + q * n2, A + (strideA ? 2 * n2 * q : 0), n2 * 16, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(b.p, B, (size_t)batch * n * nrhs * 16, hipMemcpyHostToDevice));
    orf_zgetrf(ctx->stream, n, a.as<cplx>(), n, n2, ip.as<int>(), nullptr, batch, dv.as<cplx>());
    orf_zgetrs(ctx->stream, 'L', n, nrhs, a.as<cplx>(), n, n2, ip.as<int>(), dv.as<cplx>(), b.as<cplx>(), n, (long long)n * nrhs, batch);
    KCHECK();
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    HIPCHECK(hipMemcpy(X, b.p, (size_t)batch * n * nrhs * 16, hipMemcpyDeviceToHost));
  })
}

int orf_pagemldivide_noalloc(orf_ctx* ctx, int32_t n, int32_t nrhs, const double* dA, int64_t strideA,
                             const double* dB, double* dX, int32_t batch, void* stream) {
  ORF_TRY({
    if (!ctx || n <= 0 || nrhs <= 0 || batch <= 0) fail(ORF_INVALID_ARG, "bad sizes");
    if (!dA || !dB || !dX) fail(ORF_INVALID_ARG, "device pointer is NULL");
    if (strideA != 0 && strideA != 2 * (int64_t)n * n) fail(ORF_INVALID_ARG, "strideA must be 0 or 2 n^2 doubles");
    const size_t n2 = (size_t)n * n;
    // scratch layout: LU copies of A, diagonal-block inverses, pivots
    const size_t luB = align256((size_t)batch * n2 * sizeof(cplx));
    const size_t dvB = align256(sizeof(cplx) * (size_t)orf_dinv_stride(n) * batch);
    const size_t ipB = align256(sizeof(int) * (size_t)n * batch);
    if (luB + dvB + ipB > kScratchBytes)
      fail(ORF_INVALID_ARG, "batch x n^2 too large for the allocation-free solve: needs " +
                                std::to_string((luB + dvB + ipB) >> 10) + " KiB of scratch, " +
                                std::to_string(kScratchBytes >> 10) + " KiB available");
    HIPCHECK(hipSetDevice(ctx->device));
    unsigned char* base = (unsigned char*)scratch_of(ctx->device);
    cplx* lu = (cplx*)base;
    cplx* dv = (cplx*)(base + luB);
    int* ip = (int*)(base + luB + dvB);
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    // every scratch word the solve reads is written first (LU copy, pivots and
    // inverses by the factorisation), so nothing needs clearing
    for (int b0 = 0; b0 < batch; b0 += 65535) {
      const int bb = std::min(65535, batch - b0);
      page_copy_kernel<<<dim3((unsigned)((n2 + 255) / 256), bb), 256, 0, st>>>(
          (long long)n2, (const cplx*)dA + (strideA ? (size_t)b0 * n2 : 0), strideA ? (long long)n2 : 0,
          lu + (size_t)b0 * n2, (long long)n2);
    }
    if (dX != dB)
      HIPCHECK(hipMemcpyAsync(dX, dB, (size_t)batch * n * nrhs * sizeof(cplx), hipMemcpyDeviceToDevice, st));
    orf_zgetrf(st, n, lu, n, n2, ip, nullptr, batch, dv);
    orf_zgetrs(st, 'L', n, nrhs, lu, n, n2, ip, dv, (cplx*)dX, n, (long long)n * nrhs, batch);
    KCHECK();
    if (!stream) HIPCHECK(hipStreamSynchronize(st));
  })
}

int orf_pagemrdivide(orf_ctx* ctx, int32_t n, int32_t nrows, const double* B, const double* A, int64_t strideA,
                     double* X, int32_t batch) {
  ORF_TRY({
    if (!ctx || n <= 0 || nrows <= 0 || batch <= 0) fail(ORF_INVALID_ARG, "bad sizes");
    HIPCHECK(hipSetDevice(ctx->device));
    const size_t n2 = (size_t)n * n;
    DBuf a, b, ip, dv;
    dv.alloc(16 * (size_t)orf_dinv_stride(n) * batch);
    a.alloc((size_t)batch * n2 * 16);
    b.alloc((size_t)batch * n * nrows * 16);
    ip.alloc(sizeof(int) * (size_t)n * batch);
    for (int q = 0; q < batch; ++q)
      HIPCHECK(hipMemcpy(a.as<cplx>() + q * n2, A + (strideA ? 2 * n2 * q : 0), n2 * 16, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(b.p, B, (size_t)batch * n * nrows * 16, hipMemcpyHostToDevice));
    orf_zgetrf(ctx->stream, n, a.as<cplx>(), n, n2, ip.as<int>(), nullptr, batch, dv.as<cplx>());
    orf_zgetrs(ctx->stream, 'R', n, nrows, a.as<cplx>(), n, n2, ip.as<int>(), dv.as<cplx>(), b.as<cplx>(), nrows,
               (long long)n * nrows, batch);
    KCHECK();
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    HIPCHECK(hipMemcpy(X, b.p, (size_t)batch * n * nrows * 16, hipMemcpyDeviceToHost));
  })
}

int orf_pageeig(orf_ctx* ctx, int32_t n, const double* A, double* wout, double* Vout, int32_t batch) {
  ORF_TRY({
    if (!ctx || n <= 0 || batch <= 0) fail(ORF_INVALID_ARG, "bad sizes");
    if (2 * ((n + 63) / 64) > 24 || n > 768) fail(ORF_INVALID_ARG, "n > 768 not supported");
    HIPCHECK(hipSetDevice(ctx->device));
    const size_t n2 = (size_t)n * n;
    DBuf a, wv, v, z, y, hv, info;
    hv.alloc((n2 + n) * 16 * batch);
    a.alloc(batch * n2 * 16);
    wv.alloc((size_t)batch * n * 16);
    v.alloc(batch * n2 * 16);
    z.alloc(batch * n2 * 16);
    y.alloc(batch * n2 * 16);
    info.alloc(sizeof(int) * batch);
    HIPCHECK(hipMemcpy(a.p, A, batch * n2 * 16, hipMemcpyHostToDevice));
    orf_zgeev(ctx->stream, n, a.as<cplx>(), n, n2, wv.as<cplx>(), v.as<cplx>(), n, n2, z.as<cplx>(), y.as<cplx>(),
              hv.as<cplx>(), info.as<int>(), batch);
    KCHECK();
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    std::vector<int> h(batch);
    HIPCHECK(hipMemcpy(h.data(), info.p, sizeof(int) * batch, hipMemcpyDeviceToHost));
    for (int q = 0; q < batch; ++q)
      if (h[q]) fail(ORF_DEVICE_ERROR, "QR iteration did not converge on page " + std::to_string(q));
    HIPCHECK(hipMemcpy(wout, wv.p, (size_t)batch * n * 16, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(Vout, v.p, batch * n2 * 16, hipMemcpyDeviceToHost));
  })
}

}  // extern "C"
