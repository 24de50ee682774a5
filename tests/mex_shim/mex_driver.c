/* Drives orf_rcwa_mex.c's mexFunction through the stand-in MEX API (test
 * scaffolding).  argv[1]: solve | cancel | missing | badpol.  Prints one JSON
 * line: the outputs of a solve, or the MATLAB error identifier raised. */
#include <math.h>
#include <setjmp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

extern jmp_buf fake_mex_jmp;
extern char fake_mex_err_id[128];
extern char fake_mex_err_msg[512];
extern int fake_mex_persistent;
void fake_mex_unload(void);

static mxArray* dbl(const double* v, int m, int n) {
  mxArray* a = mxCreateDoubleMatrix(m, n, mxREAL);
  memcpy(mxGetDoubles(a), v, sizeof(double) * m * n);
  return a;
}
static mxArray* i32(const int* v, int m, int n) {
  mxArray* a = mxCreateNumericMatrix(m, n, mxINT32_CLASS, mxREAL);
  for (int k = 0; k < m * n; ++k) mxGetInt32s(a)[k] = v[k];
  return a;
}
static mxArray* sc(double v) { return mxCreateDoubleScalar(v); }

static int stop_at_first(double f, const char* msg, char* id, size_t idlen) {
  (void)f; (void)msg;
  snprintf(id, idlen, "user:stop");
  return 1;
}
static int log_progress(double f, const char* msg, char* id, size_t idlen) {
  (void)id; (void)idlen;
  fprintf(stderr, "progress %.3f %s\n", f, msg);
  return 0;
}

/* two a-Si pillars (r = 0.06, 0.12 um, h = 0.6 um) in air on SiO2, lit from
 * the SiO2 side, period 0.4 um, lambda 0.94 um, Nx = Ny = 2, 64 x 64, pol s */
static mxArray* problem(int polKind) {
  const char* names[] = {"Lx", "Ly", "Nx", "Ny", "Nsx", "Nsy", "lambda0", "theta", "phi", "polKind", "jones",
                         "matEps", "incMaterial", "subMaterial", "nStruct", "nLayers", "thickness", "background",
                         "shapeStart", "shapeCount", "shapes", "vertices", "factorization", "symmetry", "gapEps",
                         "precision", "batchSize"};
  mxArray* p = mxCreateStructMatrix(1, 1, 27, names);
  const double lam = 0.94, jones[4] = {0, 0, 0, 0}, thick[2] = {0.6, 0.6};
  const int pk = polKind, bg[2] = {1, 1}, s0[2] = {0, 1}, scnt[2] = {1, 1};
  mxSetField(p, 0, "Lx", sc(0.4));
  mxSetField(p, 0, "Ly", sc(0.4));
  mxSetField(p, 0, "Nx", sc(2));
  mxSetField(p, 0, "Ny", sc(2));
  mxSetField(p, 0, "Nsx", sc(64));
  mxSetField(p, 0, "Nsy", sc(64));
  mxSetField(p, 0, "lambda0", dbl(&lam, 1, 1));
  mxSetField(p, 0, "theta", sc(0));
  mxSetField(p, 0, "phi", sc(0));
  mxSetField(p, 0, "polKind", i32(&pk, 1, 1));
  mxSetField(p, 0, "jones", dbl(jones, 4, 1));
  mxArray* eps = mxCreateDoubleMatrix(3, 1, mxCOMPLEX);
  mxComplexDouble* e = mxGetComplexDoubles(eps);
  e[0].real = 1.45 * 1.45;                       /* SiO2 */
  e[1].real = 1.0;                               /* air  */
  e[2].real = 3.6 * 3.6 - 0.01 * 0.01;           /* a-Si: (3.6 + 0.01i)^2 */
  e[2].imag = 2 * 3.6 * 0.01;
  mxSetField(p, 0, "matEps", eps);
  mxSetField(p, 0, "incMaterial", sc(0));
  mxSetField(p, 0, "subMaterial", sc(1));
  mxSetField(p, 0, "nStruct", sc(2));
  mxSetField(p, 0, "nLayers", sc(1));
  mxSetField(p, 0, "thickness", dbl(thick, 1, 2));
  mxSetField(p, 0, "background", i32(bg, 1, 2));
  mxSetField(p, 0, "shapeStart", i32(s0, 1, 2));
  mxSetField(p, 0, "shapeCount", i32(scnt, 1, 2));
  /* [type material cx cy a b vert0 nvert] rows, column-major 2 x 8 */
  const double shp[16] = {0, 0, 2, 2, 0, 0, 0, 0, 0.06, 0.12, 0, 0, 0, 0, 0, 0};
  mxSetField(p, 0, "shapes", dbl(shp, 2, 8));
  const double v[2] = {0, 0};
  mxSetField(p, 0, "vertices", dbl(v, 2, 0));
  mxSetField(p, 0, "factorization", sc(0));
  mxSetField(p, 0, "symmetry", sc(1));
  mxSetField(p, 0, "gapEps", sc(NAN));
  mxSetField(p, 0, "precision", sc(0));
  mxSetField(p, 0, "batchSize", sc(0));
  return p;
}

static void print_arr(const char* name, const double* v, int n) {
  printf("\"%s\": [", name);
  for (int k = 0; k < n; ++k) printf("%s%.17g", k ? ", " : "", v[k]);
  printf("]");
}

int main(int argc, char** argv) {
  const char* mode = argc > 1 ? argv[1] : "solve";
  mxArray* p = problem(strcmp(mode, "badpol") == 0 ? 7 : 0);
  mxArray* fcn = NULL;
  if (strcmp(mode, "cancel") == 0 || strcmp(mode, "progress") == 0) {
    fcn = mxCreateNumericMatrix(1, 1, mxFUNCTION_CLASS, mxREAL);
    fcn->cls = mxFUNCTION_CLASS;
    fcn->fcn = strcmp(mode, "cancel") == 0 ? stop_at_first : log_progress;
  }
  if (strcmp(mode, "missing") == 0) p->fieldNames[0][0] = 'X';  /* no "Lx" field */
  const mxArray* prhs[2] = {p, fcn};
  mxArray* plhs[1] = {NULL};
  int rc = setjmp(fake_mex_jmp);
  if (rc == 0) {
    mexFunction(1, plhs, fcn ? 2 : 1, prhs);
    const mxArray* out = plhs[0];
    const mxArray* R = mxGetField(out, 0, "R");
    const mxArray* T = mxGetField(out, 0, "T");
    const mxArray* t = mxGetField(out, 0, "t");
    printf("{\"status\": \"ok\", \"M\": %d, ", (int)mxGetM(R));
    print_arr("R", mxGetDoubles(R), (int)mxGetNumberOfElements(R));
    printf(", ");
    print_arr("T", mxGetDoubles(T), (int)mxGetNumberOfElements(T));
    printf(", ");
    print_arr("t", (const double*)mxGetComplexDoubles(t), 2 * (int)mxGetNumberOfElements(t));
    printf(", \"persistent\": %d}\n", fake_mex_persistent);
  } else {
    char esc[1024];
    size_t k = 0;
    for (const char* c = fake_mex_err_msg; *c && k + 2 < sizeof esc; ++c) {  /* JSON string escape */
      if (*c == '"' || *c == '\\') esc[k++] = '\\';
      esc[k++] = *c;
    }
    esc[k] = 0;
    printf("{\"status\": \"error\", \"error_id\": \"%s\", \"message\": \"%s\", \"rethrown\": %s, \"persistent\": %d}\n",
           fake_mex_err_id, esc, rc == 2 ? "true" : "false", fake_mex_persistent);
  }
  fake_mex_unload();
  return 0;
}
