// Minimal mxArray layer for exercising the MEX gateways without MATLAB
// (test scaffolding for this project's own gateways; only what they use).
#pragma once
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <vector>

typedef size_t mwSize;
typedef enum { mxDOUBLE_CLASS = 6, mxUINT64_CLASS = 13 } mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

struct mxArray {
  mxClassID cls;
  size_t m, n;
  std::vector<double> d;
  std::vector<uint64_t> u;
};

inline mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity) {
  mxArray *a = new mxArray{mxDOUBLE_CLASS, m, n, std::vector<double>(m * n), {}};
  return a;
}
inline mxArray *mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID c, mxComplexity) {
  mxArray *a = new mxArray{c, m, n, {}, std::vector<uint64_t>(m * n)};
  return a;
}
inline double *mxGetPr(const mxArray *a) { return const_cast<double *>(a->d.data()); }
inline void *mxGetData(const mxArray *a) {
  return a->cls == mxUINT64_CLASS ? (void *)a->u.data() : (void *)a->d.data();
}
inline size_t mxGetM(const mxArray *a) { return a->m; }
inline size_t mxGetN(const mxArray *a) { return a->n; }
inline size_t mxGetNumberOfElements(const mxArray *a) { return a->m * a->n; }
inline mxClassID mxGetClassID(const mxArray *a) { return a->cls; }
inline bool mxIsComplex(const mxArray *) { return false; }
inline bool mxIsDouble(const mxArray *a) { return a->cls == mxDOUBLE_CLASS; }
inline void mexErrMsgTxt(const char *msg) { throw std::runtime_error(msg); }
inline void mexLock() {}
inline void mexUnlock() {}
