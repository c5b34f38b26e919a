// dugma_compi.hip -- compi (compi.cpp / compi.cu): surrogate coefficients.
#include "dugma_internal.h"

using namespace dugma;

extern "C" int dugma_compi(dugma_ctx *ctx, const double *R, const double *t,
                           double N, double M, double D, double sig,
                           double *coef) {
  (void)N; (void)M; (void)D; (void)sig;
  if (!ctx || !R || !t || !coef) return fail(DUGMA_ERR_ARG, "dugma_compi: null argument");
  return fail(DUGMA_ERR_STATE, "dugma_compi: not built yet");
}
