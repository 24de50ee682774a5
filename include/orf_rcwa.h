/*
 * orf_rcwa.h -- C ABI of the MI355X (gfx950) 2D RCWA unit-cell solver.
 *
 * Drop-in boundary for the reference's RCWA path (MoonUsagi/OpticRCWAFDTD,
 * MATLAB).  A MATLAB MEX gateway (opticrcwafdtd_amd/matlab/orf_rcwa_mex.c) and the Python host
 * package (opticrcwafdtd_amd) bind exactly these entry points.
 *
 * Conventions (docs/CONVENTIONS.md:9-19 of the reference): e^{-i w t},
 * lengths in um, z is the stacking axis, light travels +z from the incident
 * medium (first) through the layers to the substrate (last).  All arrays are
 * column-major; complex numbers are interleaved (re, im) doubles, the MATLAB
 * R2018a+ "interleaved complex" layout, so a MEX gateway copies nothing.
 * The caller owns every host buffer; the library copies in and out and keeps
 * no host pointer after a call returns.
 */
#ifndef ORF_RCWA_H
#define ORF_RCWA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes; the MEX gateway maps them to the reference's error IDs */
enum {
  ORF_OK = 0,
  ORF_INVALID_ARG = 1,    /* arguments-block validation failures                  */
  ORF_INVALID_POL = 2,    /* 'rcwa:pol'        polarizationVector.m:6,10           */
  ORF_LOSSY_INCIDENT = 3, /* lossy incident medium, docs/CONVENTIONS.md:13         */
  ORF_OOM = 4,            /* device memory exhausted                               */
  ORF_CANCELLED = 5,      /* 'orf:cancelled'   thrown from Progress, rcwa2D.m:28-29 */
  ORF_DEVICE_ERROR = 6,   /* HIP error, eigensolver non-convergence, no gfx950     */
  ORF_SAMPLING = 7        /* 'convMatrix1D:sampling' convMatrix1D.m:14-15          */
};

/* Progress/cancel (rcwa2D.m:28-29): called on the CALLING thread after every
 * layer of every batch chunk with a monotone fraction in (0, 1]; a nonzero
 * return cancels the solve with ORF_CANCELLED. */
typedef int (*orf_progress_fn)(double fraction, const char* message, void* user);

typedef struct orf_ctx orf_ctx;
typedef struct orf_plan orf_plan;

/* Shape records (shapeCircle.m / shapeRect.m / shapePolygon.m). */
enum { ORF_SHAPE_CIRCLE = 0, ORF_SHAPE_RECT = 1, ORF_SHAPE_POLYGON = 2 };
typedef struct orf_shape {
  int32_t type;       /* ORF_SHAPE_*                                           */
  int32_t material;   /* index into the material table                         */
  double cx, cy;      /* centre [um]                                           */
  double a, b;        /* circle: a = radius; rect: a = wx, b = wy (may be inf) */
  int32_t vert0;      /* polygon: first vertex in orf_problem.vertices         */
  int32_t nvert;      /* polygon: vertex count (>= 3)                          */
} orf_shape;

/* Polarisation of one source (polarizationVector.m:4-20). */
enum { ORF_POL_S = 0, ORF_POL_P = 1, ORF_POL_JONES = 2 };

enum { ORF_FACT_LI = 0, ORF_FACT_LAURENT = 1 };          /* Factorization         */
enum { ORF_SYM_AUTO = 0, ORF_SYM_NONE = 1 };             /* Symmetry              */
/* build extension: the bulk products (convolution matrices, star products) in
 * FP32 MFMA (MIXED), or as hi + lo FP32 split products with FP64 accumulation
 * across K stages (SPLIT); eigensolver, propagators and LU solves stay FP64 */
enum { ORF_PREC_FP64 = 0, ORF_PREC_MIXED = 1, ORF_PREC_SPLIT = 2 };

/*
 * One batched solve: B structures sharing the period, the incident and
 * substrate materials and the layer count (rcwa2D.m:4-7), at nLambda
 * wavelengths (a build extension; the reference takes one) and nPol sources.
 */
typedef struct orf_problem {
  double Lx, Ly;          /* period [um]; Ly = +inf gives a 1D grating, Ny := 0 */
  int32_t Nx, Ny;         /* harmonics m = -Nx..Nx, n = -Ny..Ny                  */
  int32_t Nsx, Nsy;       /* permittivity samples per period (cellGrid.m)        */
  int32_t nLambda;
  const double* lambda0;  /* [nLambda] vacuum wavelengths [um]                   */
  double theta, phi;      /* incidence [rad]                                     */
  int32_t nPol;
  const int32_t* polKind; /* [nPol] ORF_POL_*                                    */
  const double* jones;    /* [4 x nPol] (aTE re, im, aTM re, im), used for JONES */
  int32_t nMaterials;
  const double* matEps;   /* complex [nMaterials x nLambda]: eps = n^2           */
  int32_t incMaterial, subMaterial;
  int32_t nStruct, nLayers;
  const double* thickness;      /* [nLayers x nStruct]                          */
  const int32_t* background;    /* [nLayers x nStruct] material index           */
  const int32_t* shapeStart;    /* [nLayers x nStruct] first shape              */
  const int32_t* shapeCount;    /* [nLayers x nStruct] 0 = homogeneous layer    */
  int32_t nShapes;
  const orf_shape* shapes;      /* painted in order (layerEps.m:31-38)          */
  int32_t nVertices;
  const double* vertices;       /* [2 x nVertices] polygon vertices (x; y)      */
  int32_t factorization;        /* ORF_FACT_*  (default li)                     */
  int32_t symmetry;             /* ORF_SYM_*   (default auto)                   */
  double gapEps;                /* NaN: every gap harmonic propagates           */
  int32_t precision;            /* ORF_PREC_*                                    */
  int32_t batchSize;            /* structures per chunk; 0 = auto by memory     */
  /* ---- uneven batches (build extension; zero / NULL = reference behaviour) */
  const double* structPeriod;   /* [2 x nStruct] period of each structure, NULL: Lx, Ly */
  int32_t nPages;               /* explicit (structure, wavelength) pages; 0: all pairs */
  const int32_t* pageStruct;    /* [nPages] structure index                     */
  const int32_t* pageLambda;    /* [nPages] wavelength index                    */
  int32_t keepModes;            /* 1: keep every layer's modes and S-matrix (one page only; orf_fields) */
} orf_problem;

/* Results, column-major like the reference's out struct (rcwa2D.m:31-34),
 * with wavelength and polarisation appended as trailing dimensions.  With
 * explicit pages the (structure, wavelength) pair is replaced by the page:
 * R is [M x nPages x nPol], Rtot [nPages x nPol], r [3 x M x nPages x nPol]. */
typedef struct orf_result {
  int32_t* m;     /* [M]                                                     */
  int32_t* n;     /* [M]                                                     */
  double* R;      /* [M x nStruct x nLambda x nPol]                          */
  double* T;      /* [M x nStruct x nLambda x nPol]                          */
  double* Rtot;   /* [nStruct x nLambda x nPol]                              */
  double* Ttot;   /* [nStruct x nLambda x nPol]                              */
  double* r;      /* complex [3 x M x nStruct x nLambda x nPol]; may be NULL */
  double* t;      /* complex [3 x M x nStruct x nLambda x nPol]; may be NULL */
} orf_result;

/* ------------------------------------------------------------------ context */
/* Open device `device` (must be gfx950). */
int orf_init(int device, orf_ctx** ctx);
void orf_release(orf_ctx* ctx);
/* Thread-local message of the last failing call on this thread. */
const char* orf_last_error(void);
/* Harmonic count M = (2Nx+1)(2Ny+1) for a problem (Ny forced 0 when Ly=inf). */
int orf_harmonic_count(const orf_problem* p);

/* ------------------------------------------------------------------ one call
 * orf.rcwa2D(devs, lambda0, theta, phi, pol, Name=Value) -- rcwa2D.m:2-35.
 * Copies the problem to the device, solves it, copies the results back. */
int orf_rcwa2d(orf_ctx* ctx, const orf_problem* p, orf_result* out, orf_progress_fn progress,
               void* user);

/* ------------------------------------------------------------------ resident plans
 * The same solve split so that the inputs stay in HBM between executions
 * (phase-library sweeps, serving): create uploads geometry and tables and
 * allocates the workspace; execute runs the whole GPU pipeline; fetch copies
 * the results of the last execute. */
int orf_plan_create(orf_ctx* ctx, const orf_problem* p, orf_plan** plan);
int orf_plan_execute(orf_plan* plan, orf_progress_fn progress, void* user);
int orf_plan_fetch(orf_plan* plan, orf_result* out);
void orf_plan_destroy(orf_plan* plan);
/* Solved matrix size (2M full, Mr reduced); symmetric = 0 for the full basis,
 * 1 for the mirror-parity sectors (Symmetry "auto", symmetryBasis.m:2-15),
 * 2 for the TE / TM split of a y-invariant grating lit in the x-z plane
 * (Symmetry "auto", Ny = 0, every k_y = 0; M unknowns per polarisation). */
int orf_plan_info(const orf_plan* plan, int32_t* solved_size, int32_t* symmetric,
                  int32_t* n_solve_pages);
/* Per-stage device time of the last execute, measured with HIP events on the
 * plan's stream: names[i] (static strings), total ms and launch count. */
int orf_plan_stage_times(const orf_plan* plan, int32_t max_stages, const char** names,
                         double* total_ms, int32_t* launches, int32_t* n_stages);
void orf_plan_enable_timing(orf_plan* plan, int32_t on);
/* Build extension: with on = 1, each batch chunk's stage chain is captured into a
 * HIP graph on the next execute and replayed on the executes after it (one
 * graph launch per chunk instead of one launch per kernel).  Executes with a
 * progress callback or stage timing, and kept-mode (field) plans, run eagerly. */
void orf_plan_enable_graphs(orf_plan* plan, int32_t on);

/* ------------------------------------------------------------------ fields
 * orf.rcwaFields(dev, lambda0, theta, phi, pol, x, y, z) -- rcwaFields.m:1-21.
 * One structure, one wavelength, one source (nStruct = nLambda = nPol = 1),
 * solved in the full basis.  E and H~ = i eta0 H are complex
 * [nx x ny x nz x 3] (x fastest, components x, y, z), flux [nz] is the power
 * through the unit cell normalised to the incident flux; out receives the
 * rcwa2D-style result (may have NULL members).  z = 0 is the top of the
 * first layer. */
int orf_fields(orf_ctx* ctx, const orf_problem* p, int32_t nx, const double* x, int32_t ny, const double* y,
               int32_t nz, const double* z, double* E, double* H, double* flux, orf_result* out);

/* ------------------------------------------------------------------ tuning (build extension)
 * Kernel choices of the batched eigensolver per matrix size and batch class
 * (pages >= 512 or fewer), as measured by opticrcwafdtd_amd/autotune.py and
 * stored in opticrcwafdtd_amd/tuning_table.txt ("kind n large_batch value"
 * per line).  Kinds: 0 Schur QR (0 per-rotation with H in LDS, 1 multishift
 * with 4 bulges per pass, 2 multishift with 8 bulges per pass, 3 / 4 the
 * same 4 / 8 bulges in two-wave workgroups, six / four per CU),
 * 1 Hessenberg reduction (0 rank-1 updates, 8 or 12 panel width), 2
 * eigenvectors (16 or 32 rows per lane, 0 one wave per vector).  A lookup
 * uses the entry of the largest tuned size <= n; choices a size cannot take
 * (LDS or register limits) fall back to the built-in rule.  orf_tune_force
 * (value -1 clears) overrides the table for every size (tuning runs).
 * Process-wide, not per context. */
int orf_tune_set(int32_t kind, int32_t n, int32_t large_batch, int32_t value);
int orf_tune_force(int32_t kind, int32_t value);
void orf_tune_clear(void);
int orf_tune_load(const char* path);

/* ------------------------------------------------------------------ memory (build extension)
 * Admission by device footprint for request loops.  orf_plan_memory: bytes a
 * created plan holds.  orf_plan_estimate: an upper bound before creation,
 * for the full basis (reduced = 0) or the parity-sector basis (reduced = 1).
 * Freed device blocks are cached by the library for reuse (hipFree would
 * wait for the whole device); orf_release_cached_memory returns them. */
int orf_plan_memory(const orf_plan* plan, int64_t* bytes);
int orf_plan_estimate(const orf_problem* p, int32_t reduced, int64_t* bytes);
void orf_release_cached_memory(void);

#ifdef __cplusplus
}
#endif
#endif /* ORF_RCWA_H */
