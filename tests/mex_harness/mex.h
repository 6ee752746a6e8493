/* Minimal stand-in of the MATLAB MEX/mx C API for compiling and exercising
 * matlab/pk_decon_mex.c in tests without MATLAB (test harness only). */
#ifndef PK_FAKE_MEX_H
#define PK_FAKE_MEX_H
#include <stddef.h>
typedef size_t mwSize;
typedef enum { mxUNKNOWN_CLASS, mxCHAR_CLASS, mxSTRUCT_CLASS, mxDOUBLE_CLASS, mxSINGLE_CLASS, mxUINT16_CLASS } mxClassID;
typedef enum { mxREAL, mxCOMPLEX } mxComplexity;
typedef struct mxArray_tag mxArray;
mwSize mxGetNumberOfDimensions(const mxArray *a);
const mwSize *mxGetDimensions(const mxArray *a);
size_t mxGetNumberOfElements(const mxArray *a);
mxClassID mxGetClassID(const mxArray *a);
int mxIsStruct(const mxArray *a);
int mxIsChar(const mxArray *a);
int mxIsSingle(const mxArray *a);
int mxIsEmpty(const mxArray *a);
mxArray *mxGetField(const mxArray *a, size_t i, const char *name);
double mxGetScalar(const mxArray *a);
double *mxGetPr(const mxArray *a);
void *mxGetData(const mxArray *a);
char *mxArrayToString(const mxArray *a);
void *mxMalloc(size_t n);
void mxFree(void *p);
mxArray *mxCreateNumericArray(mwSize nd, const mwSize *dims, mxClassID c, mxComplexity cx);
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity cx);
mxArray *mxCreateDoubleScalar(double v);
void mxDestroyArray(mxArray *a);
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...);
int mexAtExit(void (*fn)(void));
#endif
