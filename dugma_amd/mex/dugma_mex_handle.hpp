// dugma_mex_handle.hpp -- MATLAB handle convention of the gateways: a 1x1
// uint64 holding a pointer to a small signed wrapper of the dugma_ctx
// (the convention of class_handle.hpp:24-45: signature word checked on every
// use).  init returns the same handle six times so that DUGMA.m's
// ad0..ad5 bookkeeping works unchanged; fr_m frees it once.
#pragma once
#include <cstdint>
#include <cstring>

#include "mex.h"
#include "../../include/dugma.h"

namespace dugma_mex {

constexpr uint32_t kSignature = 0xD06A5EC7u;

struct Handle {
  uint32_t signature;
  dugma_ctx *ctx;
};

inline mxArray *to_mat(Handle *h) {
  mxArray *out = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
  uint64_t v = reinterpret_cast<uint64_t>(h);
  std::memcpy(mxGetData(out), &v, sizeof(v));
  return out;
}

inline Handle *from_mat(const mxArray *in) {
  if (mxGetNumberOfElements(in) != 1 || mxGetClassID(in) != mxUINT64_CLASS || mxIsComplex(in))
    mexErrMsgTxt("Input must be a real uint64 scalar.");
  uint64_t v;
  std::memcpy(&v, mxGetData(in), sizeof(v));
  Handle *h = reinterpret_cast<Handle *>(v);
  if (!h || h->signature != kSignature || !h->ctx) mexErrMsgTxt("Handle not valid.");
  return h;
}

inline void check(int rc) {
  if (rc != DUGMA_OK) mexErrMsgTxt(dugma_last_error());
}

// [out0, h, h, h, h, h, h]
inline void return_handles(int nlhs, mxArray *plhs[], Handle *h) {
  for (int k = 1; k < nlhs && k <= 6; ++k) plhs[k] = to_mat(h);
}

inline double scalar(const mxArray *a) {
  if (!mxIsDouble(a) || mxGetNumberOfElements(a) < 1) mexErrMsgTxt("expected a double scalar");
  return mxGetPr(a)[0];
}

}  // namespace dugma_mex
