// Drives gbm3d_mex.cpp's mexFunction through the fake mxArray runtime and
// prints one JSON line per case: the returned opts struct, or the MATLAB
// error identifier the gateway raised.  Only calls that need no GPU.
#include <cstdio>
#include <string>
#include <vector>

#include "fake_mx.hpp"

static void run(const char* name, std::vector<const mxArray*> in, int nlhs = 1) {
  mxArray* out[2] = {nullptr, nullptr};
  try {
    mexFunction(nlhs, out, (int)in.size(), in.data());
    std::printf("{\"case\": \"%s\", \"opts\": %s}\n", name, out[0] ? fake_struct_json(out[0]).c_str() : "null");
  } catch (const MexError& e) {
    std::printf("{\"case\": \"%s\", \"error_id\": \"%s\"}\n", name, e.id.c_str());
  }
}

int main() {
  for (const char* p : {"default", "fast", "superFast", "accuracy", "unknown"})
    run((std::string("set:") + p).c_str(), {mxCreateString("setBM3Dopts"), mxCreateString(p)});
  run("set:none", {mxCreateString("setBM3Dopts")});
  {
    const char* f[] = {"blockSize"};
    mxArray* s = mxCreateStructMatrix(1, 1, 1, f);
    mxSetField(s, 0, "blockSize", mxCreateDoubleScalar(8));
    run("check:blockSize8", {mxCreateString("checkBM3Dopts"), s});
  }
  {
    const char* f[] = {"matchSpins", "Wiener", "filtType", "numMax", "numMin"};
    mxArray* s = mxCreateStructMatrix(1, 1, 5, f);
    mxArray* sp = mxCreateDoubleMatrix(1, 3, mxREAL);
    mxGetPr(sp)[0] = 0;
    mxGetPr(sp)[1] = 2;
    mxGetPr(sp)[2] = 4;
    mxSetField(s, 0, "matchSpins", sp);
    mxSetField(s, 0, "Wiener", mxCreateLogicalScalar(false));
    mxSetField(s, 0, "filtType", mxCreateString("median"));
    mxSetField(s, 0, "numMax", mxCreateDoubleScalar(8));
    mxSetField(s, 0, "numMin", mxCreateDoubleScalar(8));
    run("check:spins_median", {mxCreateString("checkBM3Dopts"), s});
  }
  run("check:missing", {mxCreateString("checkBM3Dopts")});
  {
    const char* f[] = {"filtType"};
    mxArray* s = mxCreateStructMatrix(1, 1, 1, f);
    mxSetField(s, 0, "filtType", mxCreateString("gauss"));
    run("check:badfilt", {mxCreateString("checkBM3Dopts"), s});
  }
  run("unknown_command", {mxCreateString("noSuchCommand")});
  run("denoise:matrix_sigma", {mxCreateDoubleMatrix(64, 64, mxREAL), mxCreateDoubleMatrix(2, 2, mxREAL)}, 2);
  run("denoise:complex", {mxCreateDoubleMatrix(64, 64, mxCOMPLEX), mxCreateDoubleScalar(25)}, 2);
  run("denoise:too_few", {fake_single_matrix(64, 64)}, 2);
  return 0;
}
