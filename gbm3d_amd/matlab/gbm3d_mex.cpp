// gbm3d_mex.cpp -- MATLAB MEX gateway onto libgbm3d.so (include/gbm3d.h).
//
// Builds only where MATLAB's mex.h exists:
//     mex -R2018a gbm3d_mex.cpp -I<repo>/include -L<repo>/gbm3d_amd -lgbm3d
// Calls (used by GBM3D.m / setBM3Dopts.m / checkBM3Dopts.m in this folder):
//     [U, out] = gbm3d_mex(I, sigma)            % nargin < 3: profile 'default'
//     [U, out] = gbm3d_mex(I, sigma, opts)      % source/GBM3D.m:1
//     opts     = gbm3d_mex('setBM3Dopts', profile)
//     opts     = gbm3d_mex('checkBM3Dopts', opts)
//     [U, out] = gbm3d_mex('GBM3D_distributed', I, sigma, opts)   % source/GBM3D_distributed.m:1
//     [U, out] = gbm3d_mex('GBM3D_colored', I, sigmaPSD, sigmaWie, opts)   % source/GBM3D.m:33-35
// I may be double or single (m x n, real); U and out.U1 come back as double,
// as GBM3D.m returns them.  Library errors become MATLAB errors with id
// gbm3d:<status>.  One library context (device 0, its own stream) lives for
// the lifetime of the MEX file and is released at mexAtExit.
#include <cstring>
#include <string>
#include <vector>

#include "gbm3d.h"
#include "mex.h"

namespace {

gbm3d_ctx* g_ctx = nullptr;

void release() {
  if (g_ctx) gbm3d_destroy(g_ctx);
  g_ctx = nullptr;
}

void fail(int status) {
  std::string id = "gbm3d:" + std::to_string(status);
  mexErrMsgIdAndTxt(id.c_str(), "%s", gbm3d_last_error());
}

double scalar_field(const mxArray* s, const char* name, bool* present) {
  const mxArray* f = mxGetField(s, 0, name);
  *present = f != nullptr && !mxIsEmpty(f);
  return *present ? mxGetScalar(f) : 0.0;
}

void int_field(const mxArray* s, const char* name, int32_t* dst) {
  bool present = false;
  double v = scalar_field(s, name, &present);
  *dst = present ? (int32_t)v : GBM3D_UNSET;
}

void spin_field(const mxArray* s, const char* name, int32_t* n, int32_t* dst) {
  const mxArray* f = mxGetField(s, 0, name);
  if (!f) {
    *n = GBM3D_UNSET;
    return;
  }
  const size_t cnt = mxGetNumberOfElements(f);
  if (cnt > GBM3D_MAX_SPINS) mexErrMsgIdAndTxt("gbm3d:arg", "%s: at most %d entries", name, GBM3D_MAX_SPINS);
  std::vector<double> v(cnt);
  if (!mxIsDouble(f)) mexErrMsgIdAndTxt("gbm3d:arg", "%s must be a double vector", name);
  std::memcpy(v.data(), mxGetPr(f), cnt * sizeof(double));
  *n = (int32_t)cnt;
  for (size_t i = 0; i < cnt; ++i) dst[i] = (int32_t)v[i];
}

// MATLAB opts struct -> gbm3d_opts (absent fields stay absent).
void to_opts(const mxArray* s, gbm3d_opts* o) {
  gbm3d_opts_clear(o);
  if (!mxIsStruct(s)) mexErrMsgIdAndTxt("gbm3d:arg", "opts must be a struct");
  const mxArray* prof = mxGetField(s, 0, "profile");
  if (prof && mxIsChar(prof)) mxGetString(prof, o->profile, sizeof(o->profile));
  int_field(s, "numMax", &o->numMax);
  int_field(s, "numMin", &o->numMin);
  int_field(s, "levels", &o->levels);
  int_field(s, "cycleSpin", &o->cycleSpin);
  int_field(s, "tauMode", &o->tauMode);
  int_field(s, "blockSize", &o->blockSize);
  int_field(s, "blockSizeWie", &o->blockSizeWie);
  int_field(s, "matchSize", &o->matchSize);
  int_field(s, "Wiener", &o->Wiener);
  spin_field(s, "matchSpins", &o->nMatchSpins, o->matchSpins);
  spin_field(s, "matchSpinsWie", &o->nMatchSpinsWie, o->matchSpinsWie);
  const mxArray* ft = mxGetField(s, 0, "filtType");
  if (ft && mxIsChar(ft)) {
    char buf[16] = {0};
    mxGetString(ft, buf, sizeof(buf));
    if (!std::strcmp(buf, "ht"))
      o->filtType = GBM3D_FILT_HT;
    else if (!std::strcmp(buf, "median"))
      o->filtType = GBM3D_FILT_MEDIAN;
    else
      mexErrMsgIdAndTxt("gbm3d:filt", "unrecognized filtering");  // denoiseAndAggregateGPU.m:54
  }
}

// gbm3d_opts -> MATLAB struct with the present fields (setBM3Dopts.m order).
mxArray* from_opts(const gbm3d_opts& o) {
  const char* names[] = {"numMax", "numMin", "levels", "cycleSpin", "matchSpins", "matchSpinsWie", "tauMode",
                         "blockSize", "blockSizeWie", "matchSize", "Wiener", "filtType"};
  mxArray* s = mxCreateStructMatrix(1, 1, 12, names);
  auto put = [&](const char* n, int32_t v) {
    if (v != GBM3D_UNSET) mxSetField(s, 0, n, mxCreateDoubleScalar(v));
  };
  put("numMax", o.numMax);
  put("numMin", o.numMin);
  put("levels", o.levels);
  put("cycleSpin", o.cycleSpin);
  put("tauMode", o.tauMode);
  put("blockSize", o.blockSize);
  put("blockSizeWie", o.blockSizeWie);
  put("matchSize", o.matchSize);
  if (o.Wiener != GBM3D_UNSET) mxSetField(s, 0, "Wiener", mxCreateLogicalScalar(o.Wiener != 0));
  if (o.filtType != GBM3D_UNSET)
    mxSetField(s, 0, "filtType", mxCreateString(o.filtType == GBM3D_FILT_MEDIAN ? "median" : "ht"));
  auto spins = [&](const char* n, int32_t cnt, const int32_t* v) {
    if (cnt == GBM3D_UNSET) return;
    mxArray* a = mxCreateDoubleMatrix(1, cnt, mxREAL);
    for (int i = 0; i < cnt; ++i) mxGetPr(a)[i] = v[i];
    mxSetField(s, 0, n, a);
  };
  spins("matchSpins", o.nMatchSpins, o.matchSpins);
  spins("matchSpinsWie", o.nMatchSpinsWie, o.matchSpinsWie);
  return s;
}

}  // namespace

// Image argument -> float column-major copy (double or single input).
std::vector<float> to_image(const mxArray* I, int* m, int* n) {
  if (mxIsComplex(I) || mxGetNumberOfDimensions(I) != 2 || !(mxIsDouble(I) || mxIsSingle(I)))
    mexErrMsgIdAndTxt("gbm3d:arg", "I must be a real 2-D double or single image");
  *m = (int)mxGetM(I);
  *n = (int)mxGetN(I);
  std::vector<float> in((size_t)*m * *n);
  if (mxIsDouble(I)) {
    const double* p = mxGetPr(I);
    for (size_t i = 0; i < in.size(); ++i) in[i] = (float)p[i];
  } else {
    std::memcpy(in.data(), mxGetData(I), in.size() * sizeof(float));
  }
  return in;
}

mxArray* to_double(const std::vector<float>& v, int m, int n) {
  mxArray* a = mxCreateDoubleMatrix(m, n, mxREAL);
  double* p = mxGetPr(a);
  for (size_t i = 0; i < v.size(); ++i) p[i] = v[i];
  return a;
}

void ensure_ctx() {
  if (!g_ctx) {
    int st = gbm3d_create(0, &g_ctx);
    if (st) fail(st);
    mexAtExit(release);
  }
}

void distributed(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  if (nrhs < 3) mexErrMsgIdAndTxt("gbm3d:arg", "usage: gbm3d_mex('GBM3D_distributed', I, sigma, opts)");
  int m, n;
  std::vector<float> in = to_image(prhs[1], &m, &n);
  if (mxGetNumberOfElements(prhs[2]) != 1)
    mexErrMsgIdAndTxt("gbm3d:unsupported",
                      "sigma must be a scalar here; a matrix sigma (noise PSD) goes through "
                      "gbm3d_mex('GBM3D_colored', I, sigmaPSD, sigmaWie, opts)");
  const float sigma = (float)mxGetScalar(prhs[2]);
  gbm3d_opts o;
  const gbm3d_opts* op = nullptr;
  if (nrhs >= 4) {
    to_opts(prhs[3], &o);
    op = &o;
  }
  ensure_ctx();
  std::vector<float> U(in.size()), U2(in.size());
  int st = gbm3d_denoise_distributed(g_ctx, in.data(), m, n, sigma, op, U.data(), U2.data());
  if (st) fail(st);
  plhs[0] = to_double(U, m, n);
  if (nlhs > 1) {
    const char* f[] = {"U2"};
    plhs[1] = mxCreateStructMatrix(1, 1, 1, f);
    mxSetField(plhs[1], 0, "U2", to_double(U2, m, n));
  }
}

void colored(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  if (nrhs < 4) mexErrMsgIdAndTxt("gbm3d:arg", "usage: gbm3d_mex('GBM3D_colored', I, sigmaPSD, sigmaWie, opts)");
  int m, n;
  std::vector<float> in = to_image(prhs[1], &m, &n);
  if (!mxIsDouble(prhs[2]) || !mxIsDouble(prhs[3])) mexErrMsgIdAndTxt("gbm3d:arg", "sigmaPSD, sigmaWie: double");
  const int d1 = (int)mxGetM(prhs[2]), d2 = (int)mxGetN(prhs[2]);
  gbm3d_opts o;
  const gbm3d_opts* op = nullptr;
  if (nrhs >= 5) {
    to_opts(prhs[4], &o);
    op = &o;
  }
  ensure_ctx();
  std::vector<float> U(in.size()), U1(in.size());
  // the library rejects a sigmaWie that is not blockSizeWie x blockSizeWie;
  // an empty one ([]) is computed by the library (getColoredWie)
  const bool have_wie = !mxIsEmpty(prhs[3]);
  const int wm = have_wie ? (int)mxGetM(prhs[3]) : 0, wn = have_wie ? (int)mxGetN(prhs[3]) : 0;
  int st = gbm3d_denoise_colored(g_ctx, in.data(), m, n, mxGetPr(prhs[2]), d1, d2,
                                 have_wie ? mxGetPr(prhs[3]) : nullptr, wm, wn, op, U.data(), U1.data());
  if (st) fail(st);
  plhs[0] = to_double(U, m, n);
  if (nlhs > 1) {
    const char* f[] = {"U1"};
    plhs[1] = mxCreateStructMatrix(1, 1, 1, f);
    mxSetField(plhs[1], 0, "U1", to_double(U1, m, n));
  }
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  if (nrhs >= 1 && mxIsChar(prhs[0])) {
    char cmd[32] = {0};
    mxGetString(prhs[0], cmd, sizeof(cmd));
    gbm3d_opts o;
    if (!std::strcmp(cmd, "GBM3D_colored")) {
      colored(nlhs, plhs, nrhs, prhs);
      return;
    }
    if (!std::strcmp(cmd, "GBM3D_distributed")) {
      distributed(nlhs, plhs, nrhs, prhs);
      return;
    }
    if (!std::strcmp(cmd, "setBM3Dopts")) {
      char prof[32] = "default";
      if (nrhs >= 2 && mxIsChar(prhs[1])) mxGetString(prhs[1], prof, sizeof(prof));
      int st = gbm3d_set_opts(prof, &o);
      if (st) fail(st);
    } else if (!std::strcmp(cmd, "checkBM3Dopts")) {
      if (nrhs < 2) mexErrMsgIdAndTxt("gbm3d:arg", "checkBM3Dopts needs opts");
      to_opts(prhs[1], &o);
      int st = gbm3d_check_opts(&o);
      if (st) fail(st);
    } else {
      mexErrMsgIdAndTxt("gbm3d:arg", "unknown command %s", cmd);
    }
    plhs[0] = from_opts(o);
    return;
  }
  if (nrhs < 2) mexErrMsgIdAndTxt("gbm3d:arg", "usage: [U, out] = gbm3d_mex(I, sigma, opts)");
  const mxArray* I = prhs[0];
  if (mxIsComplex(I) || mxGetNumberOfDimensions(I) != 2 || !(mxIsDouble(I) || mxIsSingle(I)))
    mexErrMsgIdAndTxt("gbm3d:arg", "I must be a real 2-D double or single image");
  if (mxGetNumberOfElements(prhs[1]) != 1)
    mexErrMsgIdAndTxt("gbm3d:unsupported",
                      "sigma must be a scalar here; a matrix sigma (noise PSD) goes through "
                      "gbm3d_mex('GBM3D_colored', I, sigmaPSD, sigmaWie, opts)");
  const int m = (int)mxGetM(I), n = (int)mxGetN(I);
  const float sigma = (float)mxGetScalar(prhs[1]);
  std::vector<float> in((size_t)m * n), U((size_t)m * n), U1((size_t)m * n);
  if (mxIsDouble(I)) {
    const double* p = mxGetPr(I);
    for (size_t i = 0; i < in.size(); ++i) in[i] = (float)p[i];
  } else {
    std::memcpy(in.data(), mxGetData(I), in.size() * sizeof(float));
  }
  gbm3d_opts o;
  const gbm3d_opts* op = nullptr;
  if (nrhs >= 3) {
    to_opts(prhs[2], &o);
    op = &o;
  }
  if (!g_ctx) {
    int st = gbm3d_create(0, &g_ctx);
    if (st) fail(st);
    mexAtExit(release);
  }
  int st = gbm3d_denoise(g_ctx, in.data(), m, n, sigma, op, U.data(), U1.data());
  if (st) fail(st);
  plhs[0] = mxCreateDoubleMatrix(m, n, mxREAL);
  double* pu = mxGetPr(plhs[0]);
  for (size_t i = 0; i < U.size(); ++i) pu[i] = U[i];
  if (nlhs > 1) {
    const char* f[] = {"U1"};
    plhs[1] = mxCreateStructMatrix(1, 1, 1, f);
    mxArray* u1 = mxCreateDoubleMatrix(m, n, mxREAL);
    double* p1 = mxGetPr(u1);
    for (size_t i = 0; i < U1.size(); ++i) p1[i] = U1[i];
    mxSetField(plhs[1], 0, "U1", u1);
  }
}
