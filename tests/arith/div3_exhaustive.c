/* Exhaustive check, over every non-negative finite float x, that
 *   q = x * RN(1/3);  q + RN(1/3) * fma(-3, q, x)   (both fused)
 * equals the IEEE round-to-nearest quotient x / 3.  The update_sig pair
 * kernel divides the squared distance by D = 3 this way (update_sig.cu:255
 * divides in IEEE single precision); exit status 0 when all agree. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
int main(void) {
  const float c = 1.0f / 3.0f;
  unsigned long long bad = 0, first = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
  for (long long u = 0; u < 0x7f800000LL; ++u) {
    uint32_t b = (uint32_t)u;
    float x;
    memcpy(&x, &b, 4);
    const float ref = x / 3.0f;
    const float q = x * c;
    const float r = fmaf(-3.0f, q, x);
    const float q2 = fmaf(r, c, q);
    uint32_t rb, qb;
    memcpy(&rb, &ref, 4);
    memcpy(&qb, &q2, 4);
    if (rb != qb) bad++;
  }
  printf("mismatches: %llu of %lld\n", bad, 0x7f800000LL);
  return bad != 0;
}
