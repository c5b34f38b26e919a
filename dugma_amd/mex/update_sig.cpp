// update_sig.cpp -- MEX gateway: [sig, ad0..ad5] = update_sig(ad0..ad5, R, t, N, M, D)
// (reference gateway update_sig.cpp:8-66; DUGMA.m:47,105).
#include "dugma_mex_handle.hpp"

using namespace dugma_mex;

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
  if (nrhs != 11) mexErrMsgTxt("update_sig: expected ad0..ad5, R, t, N, M, D");
  Handle *h = from_mat(prhs[0]);
  double sig = 0.0;
  check(dugma_update_sig(h->ctx, mxGetPr(prhs[6]), mxGetPr(prhs[7]), scalar(prhs[8]), scalar(prhs[9]),
                         scalar(prhs[10]), &sig));
  plhs[0] = mxCreateDoubleMatrix(1, 1, mxREAL);
  mxGetPr(plhs[0])[0] = sig;
  return_handles(nlhs, plhs, h);
}
