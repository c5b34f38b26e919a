// compi.cpp -- MEX gateway: [coef, ad0..ad5] = compi(ad0..ad5, R, t, N, M, D, sig)
// (reference gateway compi.cpp:8-74; DUGMA.m:60).  coef is 637x1 (3-D) or
// 45x1 (2-D) in energy3D.m / energy2D.m monomial order.
#include "dugma_mex_handle.hpp"

using namespace dugma_mex;

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
  if (nrhs != 12) mexErrMsgTxt("compi: expected ad0..ad5, R, t, N, M, D, sig");
  Handle *h = from_mat(prhs[0]);
  const double D = scalar(prhs[10]);
  const int num = D == 2 ? 45 : 637;
  plhs[0] = mxCreateDoubleMatrix(num, 1, mxREAL);
  check(dugma_compi(h->ctx, mxGetPr(prhs[6]), mxGetPr(prhs[7]), scalar(prhs[8]), scalar(prhs[9]), D,
                    scalar(prhs[11]), mxGetPr(plhs[0])));
  return_handles(nlhs, plhs, h);
}
