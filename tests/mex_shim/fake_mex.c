/* Stand-in implementation of the MEX / matrix API declared in mex.h (test
 * scaffolding; see mex.h).  Errors and rethrows longjmp to fake_mex_jmp. */
#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

jmp_buf fake_mex_jmp;
char fake_mex_err_id[128];
char fake_mex_err_msg[512];
int fake_mex_persistent = 0;     /* persistent arrays alive */
static void (*g_atexit)(void) = NULL;

static mxArray* newarr(mxClassID cls, mwSize m, mwSize n, bool cplx) {
  mxArray* a = (mxArray*)calloc(1, sizeof(mxArray));
  a->cls = cls;
  a->m = m;
  a->n = n;
  a->complex = cplx;
  size_t el = cls == mxDOUBLE_CLASS ? (cplx ? 16 : 8) : cls == mxINT32_CLASS ? 4 : cls == mxCHAR_CLASS ? 1 : 0;
  if (el) a->data = calloc(m * n + 1, el);
  return a;
}

mxArray* mxGetField(const mxArray* a, mwSize i, const char* name) {
  (void)i;
  for (int k = 0; k < a->nfields; ++k)
    if (strcmp(a->fieldNames[k], name) == 0) return a->fields[k];
  return NULL;
}
void mxSetField(mxArray* a, mwSize i, const char* name, mxArray* v) {
  (void)i;
  for (int k = 0; k < a->nfields; ++k)
    if (strcmp(a->fieldNames[k], name) == 0) { a->fields[k] = v; return; }
}
double mxGetScalar(const mxArray* a) {
  if (a->cls == mxDOUBLE_CLASS) return ((double*)a->data)[0];
  if (a->cls == mxINT32_CLASS) return ((int32_t*)a->data)[0];
  return 0.0;
}
bool mxIsDouble(const mxArray* a) { return a->cls == mxDOUBLE_CLASS; }
bool mxIsComplex(const mxArray* a) { return a->complex; }
bool mxIsInt32(const mxArray* a) { return a->cls == mxINT32_CLASS; }
bool mxIsStruct(const mxArray* a) { return a->cls == mxSTRUCT_CLASS; }
bool mxIsEmpty(const mxArray* a) { return a->m * a->n == 0 && a->cls != mxFUNCTION_CLASS; }
double* mxGetDoubles(const mxArray* a) { return (double*)a->data; }
int32_t* mxGetInt32s(const mxArray* a) { return (int32_t*)a->data; }
mxComplexDouble* mxGetComplexDoubles(const mxArray* a) { return (mxComplexDouble*)a->data; }
mwSize mxGetNumberOfElements(const mxArray* a) { return a->m * a->n; }
mwSize mxGetM(const mxArray* a) { return a->m; }
mwSize mxGetN(const mxArray* a) { return a->n; }
mxArray* mxCreateDoubleScalar(double v) {
  mxArray* a = newarr(mxDOUBLE_CLASS, 1, 1, false);
  ((double*)a->data)[0] = v;
  return a;
}
mxArray* mxCreateString(const char* s) {
  mxArray* a = newarr(mxCHAR_CLASS, 1, strlen(s), false);
  memcpy(a->data, s, strlen(s));
  return a;
}
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c) { return newarr(mxDOUBLE_CLASS, m, n, c == mxCOMPLEX); }
mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity c) { return newarr(cls, m, n, c == mxCOMPLEX); }
mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID cls, mxComplexity c) {
  mwSize n = 1;
  for (mwSize k = 1; k < ndim; ++k) n *= dims[k];
  return newarr(cls, dims[0], n, c == mxCOMPLEX);
}
mxArray* mxCreateStructMatrix(mwSize m, mwSize n, int nfields, const char** names) {
  mxArray* a = newarr(mxSTRUCT_CLASS, m, n, false);
  a->nfields = nfields;
  a->fieldNames = (char**)calloc(nfields, sizeof(char*));
  a->fields = (mxArray**)calloc(nfields, sizeof(mxArray*));
  for (int k = 0; k < nfields; ++k) a->fieldNames[k] = strdup(names[k]);
  return a;
}
void mxDestroyArray(mxArray* a) {
  if (!a) return;
  if (a->ident[0] == '\x01') fake_mex_persistent -= 1;
  free(a->data);
  free(a);
}
void* mxCalloc(size_t n, size_t size) { return calloc(n, size); }
void mxFree(void* p) { free(p); }
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
  snprintf(fake_mex_err_id, sizeof fake_mex_err_id, "%s", id);
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(fake_mex_err_msg, sizeof fake_mex_err_msg, fmt, ap);
  va_end(ap);
  longjmp(fake_mex_jmp, 1);
}
int mexAtExit(void (*fn)(void)) {
  g_atexit = fn;
  return 0;
}
void fake_mex_unload(void) {
  if (g_atexit) g_atexit();
  g_atexit = NULL;
}
void mexMakeArrayPersistent(mxArray* a) {
  /* marks the array so that a later mxDestroyArray is counted */
  fake_mex_persistent += 1;
  a->ident[0] = '\x01';
}
mxArray* mexCallMATLABWithTrap(int nlhs, mxArray* plhs[], int nrhs, mxArray* prhs[], const char* name) {
  (void)nlhs; (void)plhs;
  if (strcmp(name, "feval") != 0 || nrhs < 3 || prhs[0]->cls != mxFUNCTION_CLASS) return NULL;
  char id[128] = {0};
  char msg[256] = {0};
  const mxArray* m = prhs[2];
  if (m->cls == mxCHAR_CLASS) memcpy(msg, m->data, m->n < sizeof msg - 1 ? m->n : sizeof msg - 1);
  if (prhs[0]->fcn(mxGetScalar(prhs[1]), msg, id, sizeof id) == 0) return NULL;
  mxArray* e = newarr(mxEXCEPTION_CLASS, 1, 1, false);
  snprintf(e->ident + 1, sizeof e->ident - 1, "%.100s", id);
  e->ident[0] = '\x02';
  return e;
}
int mexCallMATLAB(int nlhs, mxArray* plhs[], int nrhs, mxArray* prhs[], const char* name) {
  (void)nlhs; (void)plhs;
  if (strcmp(name, "throw") == 0 && nrhs == 1 && prhs[0]->cls == mxEXCEPTION_CLASS) {
    snprintf(fake_mex_err_id, sizeof fake_mex_err_id, "%s", prhs[0]->ident + 1);
    snprintf(fake_mex_err_msg, sizeof fake_mex_err_msg, "rethrown from the progress function");
    longjmp(fake_mex_jmp, 2);
  }
  return 1;
}
