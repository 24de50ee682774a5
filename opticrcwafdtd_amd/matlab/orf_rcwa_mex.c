/*
 * orf_rcwa_mex.c -- MATLAB MEX gateway to liborf_rcwa.so (include/orf_rcwa.h).
 *
 * Called by +orfgpu/rcwa2D.m, which has the signature and output struct of
 * orf.rcwa2D (toolbox/+orf/rcwa2D.m:2-35) and passes the batch flattened into
 * one struct whose fields mirror orf_problem.  Arrays are handed to the
 * library in place: MATLAB's R2018a interleaved-complex layout is the ABI's
 * complex layout, so nothing is copied or reordered.
 *
 *   out = orf_rcwa_mex(p, progressFcn)     % progressFcn may be []
 *
 * Build (MATLAB R2018a or newer, from the repository root):
 *   mex -R2018a -Iinclude opticrcwafdtd_amd/matlab/orf_rcwa_mex.c ...
 *       -Lopticrcwafdtd_amd -lorf_rcwa -outdir opticrcwafdtd_amd/matlab
 *
 * Status codes become the reference's error identifiers (rcwa:pol,
 * orf:cancelled, convMatrix1D:sampling, rcwa:lossyIncident); an error thrown
 * by the progress function cancels the solve and is rethrown unchanged
 * (tests/TestGUISupport.m:15-18).
 */
#include <math.h>
#include <string.h>

#include "mex.h"
#include "orf_rcwa.h"

static orf_ctx* g_ctx = NULL;
/* MException thrown by the progress function.  It is a temporary of the current
 * MEX call (the callback runs inside orf_rcwa2d, inside mexFunction) and is
 * rethrown before mexFunction returns, so it is not made persistent: MATLAB
 * frees it with the call's other temporaries. */
static mxArray* g_exception = NULL;

static void release_ctx(void) {
  if (g_ctx) orf_release(g_ctx);
  g_ctx = NULL;
}

static const mxArray* field(const mxArray* p, const char* name) {
  const mxArray* f = mxGetField(p, 0, name);
  if (!f) mexErrMsgIdAndTxt("orf:invalidArgument", "problem field '%s' is missing", name);
  return f;
}
static double scalar(const mxArray* p, const char* name) { return mxGetScalar(field(p, name)); }
static const double* doubles(const mxArray* p, const char* name) {
  const mxArray* f = field(p, name);
  if (!mxIsDouble(f) || mxIsComplex(f)) mexErrMsgIdAndTxt("orf:invalidArgument", "'%s' must be real double", name);
  return mxGetDoubles(f);
}
static const int32_t* ints(const mxArray* p, const char* name) {
  const mxArray* f = field(p, name);
  if (!mxIsInt32(f)) mexErrMsgIdAndTxt("orf:invalidArgument", "'%s' must be int32", name);
  return mxGetInt32s(f);
}

/* Progress on the calling (MATLAB) thread: orf_rcwa2d calls it between layers. */
static int progress_cb(double fraction, const char* message, void* user) {
  const mxArray* fcn = (const mxArray*)user;
  mxArray* args[3];
  args[0] = (mxArray*)fcn;
  args[1] = mxCreateDoubleScalar(fraction);
  args[2] = mxCreateString(message);
  mxArray* err = mexCallMATLABWithTrap(0, NULL, 3, args, "feval");
  mxDestroyArray(args[1]);
  mxDestroyArray(args[2]);
  if (err) {
    g_exception = err;
    return 1;
  }
  return 0;
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  if (nrhs < 1 || !mxIsStruct(prhs[0])) mexErrMsgIdAndTxt("orf:invalidArgument", "usage: orf_rcwa_mex(p, progressFcn)");
  const mxArray* p = prhs[0];
  const mxArray* fcn = (nrhs > 1 && !mxIsEmpty(prhs[1])) ? prhs[1] : NULL;
  if (!g_ctx) {
    if (orf_init(0, &g_ctx) != ORF_OK) mexErrMsgIdAndTxt("orf:device", "%s", orf_last_error());
    mexAtExit(release_ctx);
  }
  orf_problem pr;
  memset(&pr, 0, sizeof pr);
  pr.Lx = scalar(p, "Lx");
  pr.Ly = scalar(p, "Ly");
  pr.Nx = (int32_t)scalar(p, "Nx");
  pr.Ny = (int32_t)scalar(p, "Ny");
  pr.Nsx = (int32_t)scalar(p, "Nsx");
  pr.Nsy = (int32_t)scalar(p, "Nsy");
  pr.nLambda = (int32_t)mxGetNumberOfElements(field(p, "lambda0"));
  pr.lambda0 = doubles(p, "lambda0");
  pr.theta = scalar(p, "theta");
  pr.phi = scalar(p, "phi");
  pr.nPol = (int32_t)mxGetNumberOfElements(field(p, "polKind"));
  pr.polKind = ints(p, "polKind");
  pr.jones = doubles(p, "jones");
  const mxArray* eps = field(p, "matEps");                 /* complex nMaterials x nLambda */
  if (!mxIsComplex(eps)) mexErrMsgIdAndTxt("orf:invalidArgument", "matEps must be complex");
  pr.nMaterials = (int32_t)mxGetM(eps);
  pr.matEps = (const double*)mxGetComplexDoubles(eps);
  pr.incMaterial = (int32_t)scalar(p, "incMaterial");
  pr.subMaterial = (int32_t)scalar(p, "subMaterial");
  pr.nStruct = (int32_t)scalar(p, "nStruct");
  pr.nLayers = (int32_t)scalar(p, "nLayers");
  pr.thickness = doubles(p, "thickness");
  pr.background = ints(p, "background");
  pr.shapeStart = ints(p, "shapeStart");
  pr.shapeCount = ints(p, "shapeCount");
  /* shapes: [type material cx cy a b vert0 nvert] rows, nShapes x 8 double */
  const mxArray* sh = field(p, "shapes");
  pr.nShapes = (int32_t)mxGetM(sh);
  orf_shape* shapes = (orf_shape*)mxCalloc(pr.nShapes > 0 ? pr.nShapes : 1, sizeof(orf_shape));
  const double* sd = mxGetDoubles(sh);
  for (int i = 0; i < pr.nShapes; ++i) {
    const int m = pr.nShapes;
    shapes[i].type = (int32_t)sd[i];
    shapes[i].material = (int32_t)sd[i + m];
    shapes[i].cx = sd[i + 2 * m];
    shapes[i].cy = sd[i + 3 * m];
    shapes[i].a = sd[i + 4 * m];
    shapes[i].b = sd[i + 5 * m];
    shapes[i].vert0 = (int32_t)sd[i + 6 * m];
    shapes[i].nvert = (int32_t)sd[i + 7 * m];
  }
  pr.shapes = shapes;
  pr.nVertices = (int32_t)mxGetN(field(p, "vertices"));  /* 2 x nVertices */
  pr.vertices = doubles(p, "vertices");
  pr.factorization = (int32_t)scalar(p, "factorization");
  pr.symmetry = (int32_t)scalar(p, "symmetry");
  pr.gapEps = scalar(p, "gapEps");
  pr.precision = (int32_t)scalar(p, "precision");
  pr.batchSize = (int32_t)scalar(p, "batchSize");

  const int M = orf_harmonic_count(&pr);
  const mwSize nout = (mwSize)pr.nStruct * pr.nLambda * pr.nPol;
  const char* names[] = {"m", "n", "R", "T", "Rtot", "Ttot", "r", "t"};
  mxArray* out = mxCreateStructMatrix(1, 1, 8, names);
  mxArray* mm = mxCreateNumericMatrix(M, 1, mxINT32_CLASS, mxREAL);
  mxArray* nn = mxCreateNumericMatrix(M, 1, mxINT32_CLASS, mxREAL);
  mxArray* R = mxCreateDoubleMatrix(M, nout, mxREAL);
  mxArray* T = mxCreateDoubleMatrix(M, nout, mxREAL);
  mxArray* Rt = mxCreateDoubleMatrix(1, nout, mxREAL);
  mxArray* Tt = mxCreateDoubleMatrix(1, nout, mxREAL);
  const mwSize dims[2] = {3, (mwSize)M * nout};
  mxArray* r = mxCreateNumericArray(2, dims, mxDOUBLE_CLASS, mxCOMPLEX);
  mxArray* t = mxCreateNumericArray(2, dims, mxDOUBLE_CLASS, mxCOMPLEX);
  orf_result res = {mxGetInt32s(mm), mxGetInt32s(nn), mxGetDoubles(R), mxGetDoubles(T), mxGetDoubles(Rt),
                    mxGetDoubles(Tt), (double*)mxGetComplexDoubles(r), (double*)mxGetComplexDoubles(t)};
  g_exception = NULL;
  const int rc = orf_rcwa2d(g_ctx, &pr, &res, fcn ? progress_cb : NULL, (void*)fcn);
  mxFree(shapes);
  if (rc != ORF_OK) {
    if (rc == ORF_CANCELLED && g_exception) mexCallMATLAB(0, NULL, 1, &g_exception, "throw");
    const char* id = rc == ORF_INVALID_POL ? "rcwa:pol"
                   : rc == ORF_CANCELLED ? "orf:cancelled"
                   : rc == ORF_SAMPLING ? "convMatrix1D:sampling"
                   : rc == ORF_LOSSY_INCIDENT ? "rcwa:lossyIncident"
                   : rc == ORF_OOM ? "orf:outOfMemory"
                   : rc == ORF_INVALID_ARG ? "orf:invalidArgument" : "orf:device";
    mexErrMsgIdAndTxt(id, "%s", orf_last_error());
  }
  mxSetField(out, 0, "m", mm);
  mxSetField(out, 0, "n", nn);
  mxSetField(out, 0, "R", R);
  mxSetField(out, 0, "T", T);
  mxSetField(out, 0, "Rtot", Rt);
  mxSetField(out, 0, "Ttot", Tt);
  mxSetField(out, 0, "r", r);
  mxSetField(out, 0, "t", t);
  plhs[0] = out;
  (void)nlhs;
}
