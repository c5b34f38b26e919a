// init.cpp -- MEX gateway: [ok, ad0..ad5] = init(X, X_C, Y, Y_C)
// (reference gateway init.cpp:36-83; DUGMA.m:44).
#include "dugma_mex_handle.hpp"

using namespace dugma_mex;

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
  if (nrhs != 4) mexErrMsgTxt("init: expected X, X_C, Y, Y_C");
  for (int k = 0; k < 4; ++k)
    if (!mxIsDouble(prhs[k])) mexErrMsgTxt("init: arguments must be double");
  const int N = (int)mxGetN(prhs[0]), M = (int)mxGetN(prhs[2]), D = (int)mxGetM(prhs[0]);
  if ((int)mxGetM(prhs[2]) != D || (int)mxGetM(prhs[1]) != D || (int)mxGetN(prhs[1]) != D * N ||
      (int)mxGetM(prhs[3]) != D || (int)mxGetN(prhs[3]) != D * M)
    mexErrMsgTxt("init: X is D-by-N, X_C D-by-(D*N), Y D-by-M, Y_C D-by-(D*M)");
  dugma_ctx *ctx = nullptr;
  double ok = 0.0;
  check(dugma_init(mxGetPr(prhs[0]), mxGetPr(prhs[1]), mxGetPr(prhs[2]), mxGetPr(prhs[3]), N, M, D, &ctx, &ok));
  Handle *h = new Handle{kSignature, ctx};
  mexLock();
  plhs[0] = mxCreateDoubleMatrix(1, 1, mxREAL);
  mxGetPr(plhs[0])[0] = ok;
  return_handles(nlhs, plhs, h);
}
