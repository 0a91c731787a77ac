// Test-only helpers of the fake mxArray runtime (fake_mx.cpp).
#pragma once
#include <string>

#include "mex.h"

struct MexError {
  std::string id, msg;
};
mxArray* fake_single_matrix(size_t m, size_t n);
std::string fake_struct_json(const mxArray* s);
