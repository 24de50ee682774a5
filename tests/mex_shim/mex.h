/*
 * Minimal stand-in for MATLAB's MEX / matrix API (R2018a interleaved complex),
 * enough to compile and drive opticrcwafdtd_amd/matlab/orf_rcwa_mex.c without
 * MATLAB.  Test scaffolding only: arrays are plain heap blocks, errors and
 * rethrown exceptions longjmp back to the driver (mex_driver.c).
 */
#ifndef FAKE_MEX_H
#define FAKE_MEX_H
#include <stddef.h>
#include <stdint.h>
#include <stdbool.h>

typedef size_t mwSize;
typedef enum { mxUNKNOWN_CLASS = 0, mxSTRUCT_CLASS, mxDOUBLE_CLASS, mxINT32_CLASS, mxCHAR_CLASS,
               mxFUNCTION_CLASS, mxEXCEPTION_CLASS } mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;
typedef struct { double real, imag; } mxComplexDouble;

/* a "function handle": C callback standing in for a MATLAB function; returns
 * nonzero to throw an MException with identifier id */
typedef int (*fake_fcn)(double fraction, const char* message, char* id, size_t idlen);

typedef struct mxArray_tag {
  mxClassID cls;
  bool complex;
  mwSize m, n;
  void* data;          /* doubles (interleaved if complex), int32s, chars */
  int nfields;
  char** fieldNames;
  struct mxArray_tag** fields;
  fake_fcn fcn;
  char ident[128];     /* exception identifier */
} mxArray;

mxArray* mxGetField(const mxArray* a, mwSize i, const char* name);
void mxSetField(mxArray* a, mwSize i, const char* name, mxArray* v);
double mxGetScalar(const mxArray* a);
bool mxIsDouble(const mxArray* a);
bool mxIsComplex(const mxArray* a);
bool mxIsInt32(const mxArray* a);
bool mxIsStruct(const mxArray* a);
bool mxIsEmpty(const mxArray* a);
double* mxGetDoubles(const mxArray* a);
int32_t* mxGetInt32s(const mxArray* a);
mxComplexDouble* mxGetComplexDoubles(const mxArray* a);
mwSize mxGetNumberOfElements(const mxArray* a);
mwSize mxGetM(const mxArray* a);
mwSize mxGetN(const mxArray* a);
mxArray* mxCreateDoubleScalar(double v);
mxArray* mxCreateString(const char* s);
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity c);
mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID cls, mxComplexity c);
mxArray* mxCreateStructMatrix(mwSize m, mwSize n, int nfields, const char** names);
void mxDestroyArray(mxArray* a);
void* mxCalloc(size_t n, size_t size);
void mxFree(void* p);
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...);
int mexAtExit(void (*fn)(void));
void mexMakeArrayPersistent(mxArray* a);
mxArray* mexCallMATLABWithTrap(int nlhs, mxArray* plhs[], int nrhs, mxArray* prhs[], const char* name);
int mexCallMATLAB(int nlhs, mxArray* plhs[], int nrhs, mxArray* prhs[], const char* name);

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);
#endif
