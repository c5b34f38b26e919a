// Drives the MEX gateways the way DUGMA.m:44-126 calls them.
//   harness errors            argument and handle checks (no device work)
//   harness run <in> <out>    init, update_sig, compi, fr_m on a problem file
//   harness grad <in> <out>   init, energy_grad, fr_m; the problem file ends
//                             with dT (12 or 4 doubles) and a step h
// Problem file: int32 N, M, D, then doubles X (D*N), X_C (D*D*N), Y, Y_C,
// R (D*D column-major), t (D).  Output of run: doubles sig, coef (637 or 45),
// ok.  Output of grad: -E, the gateway's gradient, and central differences
// of the gateway's own value with step h (each of length 12 or 4).
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "mex.h"

void mex_init(int, mxArray **, int, const mxArray **);
void mex_update_sig(int, mxArray **, int, const mxArray **);
void mex_compi(int, mxArray **, int, const mxArray **);
void mex_fr_m(int, mxArray **, int, const mxArray **);
void mex_energy_grad(int, mxArray **, int, const mxArray **);

static mxArray *mat(size_t m, size_t n, const double *v) {
  mxArray *a = mxCreateDoubleMatrix(m, n, mxREAL);
  std::memcpy(mxGetPr(a), v, sizeof(double) * m * n);
  return a;
}

static mxArray *scal(double v) { return mat(1, 1, &v); }

static int expect_error(const char *what, void (*f)(int, mxArray **, int, const mxArray **), int nrhs,
                        const mxArray **prhs, const char *needle) {
  mxArray *out[7] = {nullptr};
  try {
    f(7, out, nrhs, prhs);
  } catch (const std::exception &e) {
    if (std::string(e.what()).find(needle) != std::string::npos) return 0;
    std::printf("%s: wrong message '%s'\n", what, e.what());
    return 1;
  }
  std::printf("%s: no error\n", what);
  return 1;
}

int main(int argc, char **argv) {
  if (argc >= 2 && std::string(argv[1]) == "errors") {
    int bad = 0;
    double x[6] = {0, 0, 0, 1, 1, 1};
    const mxArray *wrong[4] = {mat(3, 2, x), mat(3, 2, x), mat(3, 2, x), mat(3, 2, x)};
    bad += expect_error("init shapes", mex_init, 4, wrong, "init: X is D-by-N");
    bad += expect_error("init nrhs", mex_init, 3, wrong, "expected X, X_C, Y, Y_C");
    mxArray *fake = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
    const mxArray *notuint[11] = {scal(1.0)};
    for (int k = 1; k < 11; ++k) notuint[k] = scal(1.0);
    bad += expect_error("update_sig handle type", mex_update_sig, 11, notuint, "uint64 scalar");
    const mxArray *nullh[6] = {fake, fake, fake, fake, fake, fake};
    bad += expect_error("fr_m null handle", mex_fr_m, 6, nullh, "Handle not valid");
    std::printf(bad ? "FAIL\n" : "OK\n");
    return bad;
  }
  const bool grad = argc >= 4 && std::string(argv[1]) == "grad";
  if (argc >= 4 && (grad || std::string(argv[1]) == "run")) {
    FILE *f = std::fopen(argv[2], "rb");
    if (!f) return 2;
    int hdr[3];
    if (std::fread(hdr, sizeof(int), 3, f) != 3) return 2;
    const int N = hdr[0], M = hdr[1], D = hdr[2];
    std::vector<double> X(D * N), XC(D * D * N), Y(D * M), YC(D * D * M), R(D * D), t(D);
    std::vector<double> dT(grad ? (D == 3 ? 13 : 5) : 0);  // dT, then h
    for (auto *v : {&X, &XC, &Y, &YC, &R, &t, &dT})
      if (std::fread(v->data(), sizeof(double), v->size(), f) != v->size()) return 2;
    std::fclose(f);
    const mxArray *in[4] = {mat(D, N, X.data()), mat(D, D * N, XC.data()), mat(D, M, Y.data()),
                            mat(D, D * M, YC.data())};
    mxArray *o[7] = {nullptr};
    mex_init(7, o, 4, in);  // [ok, ad0..ad5] = init(X, X_C, Y, Y_C)
    if (grad) {
      const int nv = D == 3 ? 12 : 4;
      const double h = dT[nv];
      auto call = [&](const std::vector<double> &v, int nl, mxArray **out) {
        const mxArray *eg[4] = {o[1], mat(D, D, R.data()), mat(D, 1, t.data()), mat(nv, 1, v.data())};
        mex_energy_grad(nl, out, 4, eg);  // [negE, g] = energy_grad(ad0, R, t, dT)
      };
      std::vector<double> v(dT.begin(), dT.begin() + nv), res;
      mxArray *og[2] = {nullptr, nullptr};
      call(v, 2, og);
      res.push_back(mxGetPr(og[0])[0]);
      for (int k = 0; k < nv; ++k) res.push_back(mxGetPr(og[1])[k]);
      for (int k = 0; k < nv; ++k) {
        mxArray *op[1] = {nullptr}, *om[1] = {nullptr};
        std::vector<double> vp = v, vm = v;
        vp[k] += h;
        vm[k] -= h;
        call(vp, 1, op);
        call(vm, 1, om);
        res.push_back((mxGetPr(op[0])[0] - mxGetPr(om[0])[0]) / (2 * h));
      }
      const mxArray *fr[6] = {o[1], o[2], o[3], o[4], o[5], o[6]};
      mxArray *o4[1] = {nullptr};
      mex_fr_m(1, o4, 6, fr);
      FILE *g = std::fopen(argv[3], "wb");
      std::fwrite(res.data(), sizeof(double), res.size(), g);
      std::fclose(g);
      std::printf("OK\n");
      return 0;
    }
    const mxArray *us[11] = {o[1], o[2], o[3], o[4], o[5], o[6], mat(D, D, R.data()), mat(D, 1, t.data()),
                             scal(N), scal(M), scal(D)};
    mxArray *o2[7] = {nullptr};
    mex_update_sig(7, o2, 11, us);  // [sig, ad0..ad5] = update_sig(...)
    const double sig = mxGetPr(o2[0])[0];
    const mxArray *cp[12] = {o2[1], o2[2], o2[3], o2[4], o2[5], o2[6], mat(D, D, R.data()), mat(D, 1, t.data()),
                             scal(N), scal(M), scal(D), scal(sig)};
    mxArray *o3[7] = {nullptr};
    mex_compi(7, o3, 12, cp);  // [coef, ad0..ad5] = compi(...)
    const mxArray *fr[6] = {o3[1], o3[2], o3[3], o3[4], o3[5], o3[6]};
    mxArray *o4[1] = {nullptr};
    mex_fr_m(1, o4, 6, fr);  // ok = fr_m(ad0..ad5)
    FILE *g = std::fopen(argv[3], "wb");
    std::fwrite(&sig, sizeof(double), 1, g);
    std::fwrite(mxGetPr(o3[0]), sizeof(double), mxGetM(o3[0]), g);
    const double ok = mxGetPr(o4[0])[0];
    std::fwrite(&ok, sizeof(double), 1, g);
    std::fclose(g);
    std::printf("OK\n");
    return 0;
  }
  std::printf("usage: harness errors | run <in> <out> | grad <in> <out>\n");
  return 2;
}
