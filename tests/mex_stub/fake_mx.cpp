// A minimal in-process stand-in for MATLAB's mxArray runtime (test only):
// real double / single matrices, char rows, logical scalars and 1 x 1
// structs -- what gbm3d_mex.cpp's host-side marshalling touches -- so that
// the gateway's option commands and argument checks run on the CPU without
// MATLAB.  mexErrMsgIdAndTxt throws MexError (MATLAB unwinds the MEX call).
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fake_mx.hpp"

struct mxArray_tag {
  enum Kind { DOUBLE, SINGLE, CHAR, LOGICAL, STRUCT } kind = DOUBLE;
  size_t m = 0, n = 0;
  bool complex = false;
  std::vector<double> d;
  std::vector<float> f;
  std::string s;
  std::vector<std::string> names;
  std::vector<mxArray*> vals;
};

extern "C" {

mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity flag) {
  mxArray* a = new mxArray;
  a->m = m;
  a->n = n;
  a->complex = flag == mxCOMPLEX;
  a->d.assign(m * n, 0.0);
  return a;
}
mxArray* mxCreateDoubleScalar(double v) {
  mxArray* a = mxCreateDoubleMatrix(1, 1, mxREAL);
  a->d[0] = v;
  return a;
}
mxArray* mxCreateLogicalScalar(mxLogical v) {
  mxArray* a = mxCreateDoubleScalar(v ? 1.0 : 0.0);
  a->kind = mxArray::LOGICAL;
  return a;
}
mxArray* mxCreateString(const char* str) {
  mxArray* a = new mxArray;
  a->kind = mxArray::CHAR;
  a->s = str;
  a->m = 1;
  a->n = a->s.size();
  return a;
}
mxArray* mxCreateStructMatrix(mwSize m, mwSize n, int nfields, const char** fieldnames) {
  mxArray* a = new mxArray;
  a->kind = mxArray::STRUCT;
  a->m = m;
  a->n = n;
  for (int i = 0; i < nfields; ++i) {
    a->names.push_back(fieldnames[i]);
    a->vals.push_back(nullptr);
  }
  return a;
}
void* mxGetData(const mxArray* a) {
  return a->kind == mxArray::SINGLE ? (void*)a->f.data() : (void*)a->d.data();
}
double* mxGetPr(const mxArray* a) { return const_cast<double*>(a->d.data()); }
mxArray* mxGetField(const mxArray* a, mwIndex, const char* name) {
  for (size_t i = 0; i < a->names.size(); ++i)
    if (a->names[i] == name) return a->vals[i];
  return nullptr;
}
void mxSetField(mxArray* a, mwIndex, const char* name, mxArray* v) {
  for (size_t i = 0; i < a->names.size(); ++i)
    if (a->names[i] == name) {
      a->vals[i] = v;
      return;
    }
  a->names.push_back(name);
  a->vals.push_back(v);
}
size_t mxGetM(const mxArray* a) { return a->m; }
size_t mxGetN(const mxArray* a) { return a->n; }
mwSize mxGetNumberOfDimensions(const mxArray*) { return 2; }
size_t mxGetNumberOfElements(const mxArray* a) { return a->m * a->n; }
double mxGetScalar(const mxArray* a) {
  if (a->kind == mxArray::SINGLE) return a->f.empty() ? 0.0 : a->f[0];
  return a->d.empty() ? 0.0 : a->d[0];
}
int mxGetString(const mxArray* a, char* str, mwSize len) {
  if (a->kind != mxArray::CHAR || len == 0) return 1;
  std::snprintf(str, len, "%s", a->s.c_str());
  return a->s.size() + 1 > len ? 1 : 0;
}
bool mxIsChar(const mxArray* a) { return a->kind == mxArray::CHAR; }
bool mxIsComplex(const mxArray* a) { return a->complex; }
bool mxIsDouble(const mxArray* a) { return a->kind == mxArray::DOUBLE; }
bool mxIsEmpty(const mxArray* a) { return a->m * a->n == 0; }
bool mxIsSingle(const mxArray* a) { return a->kind == mxArray::SINGLE; }
bool mxIsStruct(const mxArray* a) { return a->kind == mxArray::STRUCT; }
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw MexError{id, buf};
}
int mexAtExit(void (*)(void)) { return 0; }

}  // extern "C"

mxArray* fake_single_matrix(size_t m, size_t n) {
  mxArray* a = mxCreateDoubleMatrix(m, n, mxREAL);
  a->kind = mxArray::SINGLE;
  a->d.clear();
  a->f.assign(m * n, 0.f);
  return a;
}

// The struct's present fields as a JSON object (numbers, vectors, strings, logicals).
std::string fake_struct_json(const mxArray* s) {
  std::string out = "{";
  for (size_t i = 0; i < s->names.size(); ++i) {
    const mxArray* v = s->vals[i];
    if (!v) continue;
    if (out.size() > 1) out += ", ";
    out += "\"" + s->names[i] + "\": ";
    char num[64];
    if (v->kind == mxArray::CHAR) {
      out += "\"" + v->s + "\"";
    } else if (v->kind == mxArray::LOGICAL) {
      out += v->d[0] != 0 ? "true" : "false";
    } else if (v->m * v->n == 1 && v->names.empty() && v->n == 1 && s->names[i].find("Spins") == std::string::npos) {
      std::snprintf(num, sizeof(num), "%.17g", v->d[0]);
      out += num;
    } else {
      out += "[";
      for (size_t k = 0; k < v->d.size(); ++k) {
        std::snprintf(num, sizeof(num), "%s%.17g", k ? ", " : "", v->d[k]);
        out += num;
      }
      out += "]";
    }
  }
  return out + "}";
}
