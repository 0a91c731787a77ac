/* Declarations-only stand-in for MATLAB's mex.h, used by the CPU test suite
 * to compile-check gbm3d_amd/matlab/gbm3d_mex.cpp where no MATLAB is
 * installed.  It holds the prototypes of the C Matrix / MEX API functions the
 * gateway calls, as MATLAB's External API documentation gives them; nothing
 * here is defined, so it checks types and signatures only (no link, no run). */
#ifndef GBM3D_TEST_MEX_STUB_H
#define GBM3D_TEST_MEX_STUB_H

#include <stddef.h>

typedef struct mxArray_tag mxArray;
typedef size_t mwSize;
typedef size_t mwIndex;
typedef bool mxLogical;
typedef enum { mxREAL, mxCOMPLEX } mxComplexity;

#ifdef __cplusplus
extern "C" {
#endif

mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity flag);
mxArray* mxCreateDoubleScalar(double value);
mxArray* mxCreateLogicalScalar(mxLogical value);
mxArray* mxCreateString(const char* str);
mxArray* mxCreateStructMatrix(mwSize m, mwSize n, int nfields, const char** fieldnames);
void* mxGetData(const mxArray* pa);
double* mxGetPr(const mxArray* pa);
mxArray* mxGetField(const mxArray* pm, mwIndex index, const char* fieldname);
void mxSetField(mxArray* pm, mwIndex index, const char* fieldname, mxArray* pvalue);
size_t mxGetM(const mxArray* pm);
size_t mxGetN(const mxArray* pm);
mwSize mxGetNumberOfDimensions(const mxArray* pm);
size_t mxGetNumberOfElements(const mxArray* pm);
double mxGetScalar(const mxArray* pm);
int mxGetString(const mxArray* pm, char* str, mwSize strlen);
bool mxIsChar(const mxArray* pm);
bool mxIsComplex(const mxArray* pm);
bool mxIsDouble(const mxArray* pm);
bool mxIsEmpty(const mxArray* pm);
bool mxIsSingle(const mxArray* pm);
bool mxIsStruct(const mxArray* pm);
void mexErrMsgIdAndTxt(const char* errorid, const char* errormsg, ...);
int mexAtExit(void (*exit_fcn)(void));
void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);

#ifdef __cplusplus
}
#endif

#endif
