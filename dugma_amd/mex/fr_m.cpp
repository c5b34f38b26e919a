// fr_m.cpp -- MEX gateway: ok = fr_m(ad0..ad5)   (reference gateway
// fr_m.cpp:8-35; DUGMA.m:126).  Frees the context once (the six handles
// are the same object); unlike the reference it does not reset the device.
#include "dugma_mex_handle.hpp"

using namespace dugma_mex;

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
  if (nrhs != 6) mexErrMsgTxt("fr_m: expected ad0..ad5");
  Handle *h = from_mat(prhs[0]);
  double ok = 0.0;
  check(dugma_fr_m(h->ctx, &ok));
  h->ctx = nullptr;
  h->signature = 0;
  delete h;
  mexUnlock();
  plhs[0] = mxCreateDoubleMatrix(1, 1, mxREAL);
  mxGetPr(plhs[0])[0] = ok;
  (void)nlhs;
}
