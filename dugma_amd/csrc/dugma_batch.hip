// dugma_batch.hip -- thousands of independent rigid registrations of uneven
// sizes on one GPU.
//
// Each trial maximises the convolved-uncertainty energy
//   E(R,t) = sum_ij N(y_i - R x_j - t; (g (s_i + s_j) + a) I)
// (isotropic covariances s) by majorise-minimise steps: with e_ij at the
// current pose, exp(-u) >= e0 (1 - (u - u0)) makes E grow whenever
// sum_ij (e_ij / s_ij) |y_i - R x_j - t|^2 shrinks, and that weighted
// least-squares problem has the closed-form (Horn) solution.
//
// The variance scale g follows DUGMA's sigma update: DUGMA.m:47,105 measure
// sig = mean_i min_j |y_i - R x_j - t|^2 / D (update_sig.cu:238-260, 384-385)
// and compi divides every inverse covariance by sig (compi.cu:642-661), i.e.
// the kernels use sig * Sigma.  Here the point variances act as a unitless
// shape, g = sig / sbar with sbar the trial's mean target plus mean source
// variance, so the kernel width tracks the current mean nearest distance.
// The nearest distances come out of the same all-pairs pass as the
// statistics (one fminf per pair, one atomicMin per target and work item);
// a first pass before the iterations measures the starting sig, as DUGMA.m
// calls update_sig before its loop.  Optionally an extra variance a -> 0 is
// annealed on top.
//
// One iteration for all trials is four kernels:
//   k_batch_transform  the targets in each trial's source frame,
//                      z_i = R_k^T (y_i - (t_k - c_k)) (FP64, rounded to
//                      FP32), and g s_i + a; the sources stay as given (FP32)
//   k_batch_pairs      persistent; workgroups pop (trial, target tile, source
//                      chunk) items from a queue ordered longest first and
//                      write FP64 sufficient statistics per item, plus the
//                      per-target nearest squared distance
//   k_batch_fold       one workgroup per trial: its nearest-distance sum
//                      and a fixed-order sum of its items; re-arms the queue
//   k_batch_update     one thread per trial: Horn's quaternion solution
//                      (4x4 Jacobi, FP64), pose, sigma, annealing,
//                      convergence
// captured once into a hipGraph and replayed.  The same kernels run one
// trial per launch for the baseline.
//
// Two resolutions: the first coarse_iters iterations run on every 4th target
// and source point of each trial (1/16 of the pairs; same centroid and
// variances), never declaring convergence; then the full clouds take over,
// after one pass that measures sig again on them.  Each resolution has its
// own work items and its own captured iteration.
//
// Refinement (dugma_batch_set_covariances + refine_iters): after the
// isotropic iterations, each trial maximises the convolved energy with
// full per-point covariances (for surfaces sampled independently, covariances
// flat along the surface -- e.g. the sample covariance of a point's nearest
// neighbours -- make the energy's maximum sit at the true pose) by BFGS on
// -log E over left pose increments, one fused energy + pose-gradient pass per
// trial pose (the logic of dugma_register_energy, one thread per trial):
//   k_batchf_transform  targets and their covariances in each trial's source
//                       frame at the candidate pose (FP64 -> FP32 hi/lo)
//   k_batchf_pairs      the same persistent item queue; per target lane E,
//                       EW and B (full covariance), then the pose gradient
//                       by the target-side identity, in the source frame;
//                       a wave skips each block of 64 sources whose terms
//                       are all below 7e-13 of their peaks (kFarExp)
//   k_batchf_fold       sums each running trial's items (one workgroup each)
//   k_batchf_update     one thread per trial: rotates the gradient back, Armijo
//                       test, BFGS update, next candidate pose; at an
//                       accepted pose the direction comes from the pass's
//                       own approximate Hessian (Newton) when that is
//                       positive definite, else from the BFGS matrix
// Optionally the refinement runs first on sparser levels' points, with
// their covariances times the level's point step: the coarse level
// (refine_coarse_iters; a quarter of the sampling density, so a surface
// covariance spans twice the distance), then every 2nd point
// (refine_mid_iters), each restarting BFGS from the previous level's pose
// and scaled inverse Hessian, before the full clouds.
//
// Each trial's points are kept in spatial order (spatial_order): the
// sparser levels are then evenly spread subsamples, and 64 consecutive
// points lie close together, which the far-field skip needs.
//
// Launch shape: waves per workgroup (target tile = 64 x waves), the number
// of source chunks per target tile for each size class of trials (by source
// count), and the resident waves of the persistent kernel are chosen by
// dugma_batch_autotune, which times whole iterations on the batch's own data.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "dugma_internal.h"
#include "dugma_pairs.h"

namespace dugma {

constexpr int kBSub = 64;
constexpr int kStat = 28;         // stride of the per-item statistics (doubles; 17 / 28 used)
constexpr int kUpd = 256;         // threads of the fold kernel (one workgroup per trial)
constexpr int kClasses = 4;       // size classes by source count: < 1000, < 2000, < 4000, more
constexpr unsigned kInfBits = 0x7f800000u;  // +inf as u32: start value of the nearest distance
constexpr double kGMin = 1e-6;    // smallest variance scale g
// variances are held times ln 4 = 2 ln 2, so exp2(-|d|^2 / s') = exp(-|d|^2 / 2s)
// (kLn4 and kLn4Pow1p5, the factor e is then short by, are shared with the
// isotropic pair records)

static inline int size_class(int ns) { return ns < 1000 ? 0 : ns < 2000 ? 1 : ns < 4000 ? 2 : 3; }

// A work item: targets [t0, t0 + tn) of a trial against its sources
// [s0, s0 + sn); s0 sits a multiple of 64 sources into the trial's, and blk
// is the index of its first 64-source block in the level's block bounds.
struct BItem {
  int trial, t0, tn, s0, sn, blk, pad1, pad2;
};

struct BTrial {
  double R[9];  // column-major
  double t[3];
  double c[3];  // target centroid (targets are stored relative to it)
  double a;     // annealing variance
  double a_stop;
  double E;     // energy at the returned pose, from the energy pass (without the (2 pi)^-3/2 factor)
  double dang, dtr;  // last rotation / translation step
  double g;     // variance scale: sig / sbar with the sigma update, else 1
  double sig;   // mean squared nearest distance / 3 at the last pass
  double sbar;  // mean target variance + mean source variance
  int iters;
  int status;   // 0 running, 1 converged, 2 iteration limit, 3 degenerate
  int sig_ready;
  int pad;
};

struct BParams {
  double beta;       // annealing decay per iteration
  double tol_rot;    // radians
  double tol_trans;  // in the trial's own units
  double relax;      // over-relaxation of the pose step once annealing has ended (1: none)
  int max_iter;
  int sigma;         // 1: DUGMA sigma update
};

// A streamed source point: the input coordinates (FP32) and its variance
// times ln 4.  Static: the pose moves the targets into the source frame.
struct alignas(16) BSrc {
  float x, y, z, s;
};

// A target in its trial's source frame at the current pose: z (FP32),
// s = (g s_i + a) ln 4 and the trial's variance scale g.
struct alignas(16) BZ {
  float x, y, z, s, g, pad0, pad1, pad2;
};

// Per iteration, for every target of a running trial: z = R^T (y - (t - c))
// in FP64 (y relative to the trial's target centroid c), rounded to FP32,
// with (g s_i + a) ln 4 and g.  z - x_j = R^T (y - R x_j - t) has the
// length of the pair difference in the target frame.  Rounding z moves the
// target by at most half an FP32 ulp of its coordinates (~3e-8 of the
// trial's extent), a millionth of the kernel width at the smallest sigma
// the trials have.
__global__ __launch_bounds__(256) void k_batch_transform(
    const Rec3I *__restrict__ T, const int *__restrict__ tgt_trial, int i0,
    int i1, const BTrial *__restrict__ trials, const int *__restrict__ done,
    BZ *__restrict__ Z) {
  const int i = i0 + blockIdx.x * 256 + threadIdx.x;
  if (i >= i1) return;
  const int k = tgt_trial[i];
  if (done[k]) return;
  const BTrial &tr = trials[k];
  const Rec3I y = T[i];
  const double u[3] = {(double)y.hx - (tr.t[0] - tr.c[0]), (double)y.hy - (tr.t[1] - tr.c[1]),
                       (double)y.hz - (tr.t[2] - tr.c[2])};
  double v[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < 3; ++b) s = fma(tr.R[b + 3 * a], u[b], s);  // (R^T u)_a
    v[a] = s;
  }
  BZ r;
  r.x = (float)v[0];
  r.y = (float)v[1];
  r.z = (float)v[2];
  r.s = (float)fma(tr.g, (double)y.s, tr.a * kLn4);  // y.s holds s_i ln 4
  r.g = (float)tr.g;
  r.pad0 = r.pad1 = r.pad2 = 0.0f;
  Z[i] = r;
}

__global__ __launch_bounds__(256) void k_batch_fill(unsigned *__restrict__ p, int n, unsigned v) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

// Marks every running trial's sigma as unmeasured, so the next pass (on the
// other level) measures it again before moving the pose.
__global__ __launch_bounds__(256) void k_batch_resig(BTrial *__restrict__ trials, int K) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < K) trials[k].sig_ready = 0;
}

// Statistics of one item, per target i (lane), in the trial's source frame:
// W_i = sum_j w_ij, B'_i = sum_j w_ij x_j, E_i = sum_j e_ij with
// w = e / s_ij, s_ij = g (s_i + s_j) + a.  Item totals: [E, W,
// sum W_i y_i (3), sum B'_i (3), sum y_i B'_i^T (9, row a = y_a)], y_i in the
// target frame relative to the centroid; the update kernel rotates the B'
// parts into the target frame.
// Every operand of the pair arithmetic that is not a streamed value sits in
// a VGPR (VALU instructions reading an SGPR or a literal issue at half rate):
//  - the pair difference is d' = z - x: one streamed operand per axis
//    (exact for nearby points, so d' carries the FP32 z unchanged);
//  - B'_i comes from D'_i = sum_j w_ij d'_ij as B'_i = W_i z_i - D'_i, with
//    the same FP32 z_i;
//  - s_ij = fma(s_j, g, g s_i + a), g and g s_i + a in VGPRs;
//  - variances are held times ln 4, so that exp2(-|d|^2 / s') is the
//    Gaussian exp(-|d|^2 / 2s) with no constant factor per pair; e and w
//    then carry the constant factors (ln 4)^-3/2 and (ln 4)^-5/2, which the
//    pose step does not see (ratios of sums) and the update kernel takes out
//    of E.
// With SIG the item's smallest |d_ij|^2 per target is folded into nn[i]
// (atomicMin on the bits of a non-negative float, which order like the
// values).  The source stream is software-pipelined two deep: record j + 2
// is requested while record j is used (index clamped to the item), so the
// scalar-load latency overlaps the arithmetic.
// EN: also sum E_i.  The pose step does not need it, so the iterations
// leave it out (one VALU operation per pair fewer) and one pass with EN
// after the iterations measures the energy at the returned poses.
constexpr int kItemStat = 17;
template <int W, bool SIG, bool EN>
__global__ __launch_bounds__(64 * W) void k_batch_pairs(
    const Rec3I *__restrict__ T, const BZ *__restrict__ Z, const BSrc *__restrict__ S,
    const BItem *__restrict__ items, const int *__restrict__ order, int nitems,
    unsigned *__restrict__ queue, const int *__restrict__ done,
    unsigned *__restrict__ nn, double *__restrict__ part) {
  __shared__ int s_p;
  __shared__ double red[W][kItemStat];
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (;;) {
    if (threadIdx.x == 0) s_p = (int)atomicAdd(queue, 1u);
    __syncthreads();
    const int p = __builtin_amdgcn_readfirstlane(s_p);
    __syncthreads();
    if (p >= nitems) break;
    const int idx = __builtin_amdgcn_readfirstlane(order[p]);
    const BItem it = items[idx];
    if (done[it.trial]) continue;
    const int ti = it.t0 + wave * 64 + lane;
    const bool valid = (wave * 64 + lane) < it.tn;
    const BZ pz = Z[valid ? ti : it.t0];
    const float g = pz.g, cs = pz.s;
    double dE = 0.0, dW = 0.0, dD[3] = {0.0, 0.0, 0.0};
    float dmin = __builtin_huge_valf();
    const int s_end = it.s0 + it.sn;
    // a wave past the tile's last target skips the pair loop (its sums stay
    // zero, as its masked lanes would have made them)
    const int s_run = wave * 64 < it.tn ? s_end : it.s0;
    BSrc qa = S[it.s0], qb = S[min(it.s0 + 1, s_end - 1)];  // it.sn >= 1
    for (int j0 = it.s0; j0 < s_run; j0 += kBSub) {
      float fE = 0.0f, fW = 0.0f, fD0 = 0.0f, fD1 = 0.0f, fD2 = 0.0f;
      const int j1 = min(j0 + kBSub, s_end);
#pragma unroll 2
      for (int j = j0; j < j1; ++j) {
        const BSrc q = qa;
        qa = qb;
        qb = S[min(j + 2, s_end - 1)];
        const float dx = pz.x - q.x;
        const float dy = pz.y - q.y;
        const float dz = pz.z - q.z;
        const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        if (SIG) dmin = fminf(dmin, d2);
        const float rs = frsq(fmaf(q.s, g, cs));
        const float inv = rs * rs;
        const float e = fexp2(-(d2 * inv)) * (rs * inv);
        const float w = e * inv;
        if (EN) fE += e;
        fW += w;
        fD0 = fmaf(w, dx, fD0);
        fD1 = fmaf(w, dy, fD1);
        fD2 = fmaf(w, dz, fD2);
      }
      dE += fE;
      dW += fW;
      dD[0] += fD0;
      dD[1] += fD1;
      dD[2] += fD2;
    }
    if (SIG && valid) atomicMin(nn + ti, __float_as_uint(dmin));
    // item totals, one statistic at a time (few registers live): [E, W,
    // W y (3), B' (3), y B'^T (9)]
    const Rec3I pt = T[valid ? ti : it.t0];
    const double y[3] = {(double)pt.hx, (double)pt.hy, (double)pt.hz};
    const double z[3] = {(double)pz.x, (double)pz.y, (double)pz.z};
    const double dB[3] = {fma(dW, z[0], -dD[0]), fma(dW, z[1], -dD[1]), fma(dW, z[2], -dD[2])};
#pragma unroll
    for (int a = 0; a < kItemStat; ++a) {
      double v;
      if (a == 0)
        v = dE;
      else if (a == 1)
        v = dW;
      else if (a < 5)
        v = dW * y[a - 2];
      else if (a < 8)
        v = dB[a - 5];
      else
        v = y[(a - 8) / 3] * dB[(a - 8) % 3];
      v = wave_sum(valid ? v : 0.0);
      if (lane == 0) red[wave][a] = v;
    }
    __syncthreads();
    if (threadIdx.x < kItemStat) {
      double s = red[0][threadIdx.x];
#pragma unroll
      for (int w = 1; w < W; ++w) s += red[w][threadIdx.x];
      part[(size_t)idx * kStat + threadIdx.x] = s;
    }
    __syncthreads();
  }
}

// Largest-eigenvalue eigenvector of a symmetric 4x4 (cyclic Jacobi, FP64).
// Every index is a compile-time constant (loops unrolled), so A and V stay
// in registers.
__device__ void eig4_max(double A[4][4], double q[4]) {
  double V[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) V[a][b] = a == b ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = 0.0, tot = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        tot += A[a][b] * A[a][b];
        if (a != b) off += A[a][b] * A[a][b];
      }
    if (off <= 1e-30 * tot) break;
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int r = p + 1; r < 4; ++r) {
        if (A[p][r] == 0.0) continue;
        const double th = 0.5 * (A[r][r] - A[p][p]) / A[p][r];
        const double tt = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        const double c = 1.0 / sqrt(tt * tt + 1.0), s = tt * c;
#pragma unroll
        for (int k = 0; k < 4; ++k) {  // A <- A J
          const double akp = A[k][p], akr = A[k][r];
          A[k][p] = c * akp - s * akr;
          A[k][r] = s * akp + c * akr;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {  // A <- J^T A
          const double apk = A[p][k], ark = A[r][k];
          A[p][k] = c * apk - s * ark;
          A[r][k] = s * apk + c * ark;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const double vkp = V[k][p], vkr = V[k][r];
          V[k][p] = c * vkp - s * vkr;
          V[k][r] = s * vkp + c * vkr;
        }
      }
  }
  double best = A[0][0];
#pragma unroll
  for (int a = 0; a < 4; ++a) q[a] = V[a][0];
#pragma unroll
  for (int b = 1; b < 4; ++b)
    if (A[b][b] > best) {
      best = A[b][b];
#pragma unroll
      for (int a = 0; a < 4; ++a) q[a] = V[a][b];
    }
}

// One workgroup per trial.  All threads fold the trial's nearest distances
// (and reset them for the next pass) into sig = mean / 3 (-1: none), and
// stage the item statistics in LDS a block of items at a time, from which
// thread a < 17 sums statistic a in item order.  fold[k] receives the 17
// sums and, at index 17, sig.
constexpr int kUpdItems = 64;  // items staged per block
__global__ __launch_bounds__(kUpd) void k_batch_fold(
    const double *__restrict__ part, const int *__restrict__ titem, const int *__restrict__ ttgt, int k0,
    int k1, const BTrial *__restrict__ trials, const int *__restrict__ done, unsigned *__restrict__ queue,
    unsigned *__restrict__ nn, double *__restrict__ fold, int sigma) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *queue = 0u;
  const int k = k0 + blockIdx.x;
  if (k >= k1 || done[k]) return;
  double *f = fold + (size_t)k * kStat;
  if (sigma) {
    __shared__ double s_nn[kUpd / 64];
    double snn = 0.0;
#pragma unroll 4
    for (int i = ttgt[k] + threadIdx.x; i < ttgt[k + 1]; i += kUpd) {
      snn += (double)__uint_as_float(nn[i]);
      nn[i] = kInfBits;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) snn += __shfl_xor(snn, off);
    if ((threadIdx.x & 63) == 0) s_nn[threadIdx.x >> 6] = snn;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0.0;
      for (int w = 0; w < kUpd / 64; ++w) s += s_nn[w];
      const int nt = ttgt[k + 1] - ttgt[k];
      f[kItemStat] = (nt > 0 && isfinite(s)) ? s / (3.0 * nt) : -1.0;
    }
  }
  if (sigma && !trials[k].sig_ready) return;  // this pass only measures sig
  __shared__ double s_part[kUpdItems * kStat];
  const int i0 = titem[k], i1 = titem[k + 1];
  double acc = 0.0;
  for (int ib = i0; ib < i1; ib += kUpdItems) {
    const int nb = min(kUpdItems, i1 - ib);
    for (int e = threadIdx.x; e < nb * kStat; e += kUpd) s_part[e] = part[(size_t)ib * kStat + e];
    __syncthreads();
    if (threadIdx.x < kItemStat)
      for (int i = 0; i < nb; ++i) acc += s_part[i * kStat + threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x < kItemStat) f[threadIdx.x] = acc;
}

// One thread per trial, after k_batch_fold: the pose step from the folded
// sums, the sigma update, annealing and convergence.
__global__ __launch_bounds__(64) void k_batch_update(const double *__restrict__ fold, int k0, int k1,
                                                     BTrial *__restrict__ trials, int *__restrict__ done,
                                                     BParams bp) {
  const int k = k0 + blockIdx.x * 64 + threadIdx.x;
  if (k >= k1 || done[k]) return;
  const double *f = fold + (size_t)k * kStat;
  BTrial tr = trials[k];
  const double sig = bp.sigma ? f[kItemStat] : -1.0;  // < 0: not measured
  if (bp.sigma && !tr.sig_ready) {  // the pass before the first iteration only measures sig
    if (sig >= 0.0) {
      tr.sig = sig;
      tr.g = fmax(sig / tr.sbar, kGMin);
    }
    tr.sig_ready = 1;
    trials[k] = tr;
    return;
  }
  double st[kStat];
  for (int a = 0; a < kStat; ++a) st[a] = a < kItemStat ? f[a] : 0.0;
  const double W = st[1];
  {
    // B' parts from the source frame to the target frame (relative to c):
    // B = R B' + W (t - c);  sum y B^T = (sum y B'^T) R^T + (sum W y)(t - c)^T
    double tc[3], Bp[3], YB[9];
    for (int a = 0; a < 3; ++a) {
      tc[a] = tr.t[a] - tr.c[a];
      Bp[a] = st[5 + a];
    }
    for (int a = 0; a < 9; ++a) YB[a] = st[8 + a];
    for (int a = 0; a < 3; ++a)
      st[5 + a] = tr.R[a] * Bp[0] + tr.R[a + 3] * Bp[1] + tr.R[a + 6] * Bp[2] + W * tc[a];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        st[8 + 3 * a + b] = YB[3 * a] * tr.R[b] + YB[3 * a + 1] * tr.R[b + 3] + YB[3 * a + 2] * tr.R[b + 6] +
                            st[2 + a] * tc[b];
  }
  tr.iters += 1;
  if (!(W > 0.0) || !isfinite(W)) {
    tr.status = 3;  // no pair has any weight: nothing to align
    trials[k] = tr;
    done[k] = 1;
    return;
  }
  double my[3], mx[3], H[3][3];
  for (int a = 0; a < 3; ++a) {
    my[a] = st[2 + a] / W;
    mx[a] = st[5 + a] / W;
  }
  // H_ab = sum w (x'_a - mx_a)(y_b - my_b); st[8+3a+b] = sum y_a x'_b
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) H[a][b] = st[8 + 3 * b + a] - W * mx[a] * my[b];
  const double Sxx = H[0][0], Sxy = H[0][1], Sxz = H[0][2];
  const double Syx = H[1][0], Syy = H[1][1], Syz = H[1][2];
  const double Szx = H[2][0], Szy = H[2][1], Szz = H[2][2];
  double N4[4][4] = {{Sxx + Syy + Szz, Syz - Szy, Szx - Sxz, Sxy - Syx},
                     {Syz - Szy, Sxx - Syy - Szz, Sxy + Syx, Szx + Sxz},
                     {Szx - Sxz, Sxy + Syx, -Sxx + Syy - Szz, Syz + Szy},
                     {Sxy - Syx, Szx + Sxz, Syz + Szy, -Sxx - Syy + Szz}};
  double q[4];
  eig4_max(N4, q);
  const double qn = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  for (int a = 0; a < 4; ++a) q[a] /= qn;
  const bool annealing = tr.a > 0.0;
  // over-relaxed step (after annealing): the rotation angle times relax,
  // and below the weighted source centroid moved relax times as far
  const double om = (!annealing && bp.relax > 1.0) ? bp.relax : 1.0;
  if (om != 1.0) {
    if (q[0] < 0.0)
      for (int a = 0; a < 4; ++a) q[a] = -q[a];
    const double vn = sqrt(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (vn > 0.0) {
      const double ph = om * atan2(vn, q[0]);
      const double sv = sin(ph) / vn;
      q[0] = cos(ph);
      for (int a = 1; a < 4; ++a) q[a] *= sv;
    }
  }
  const double w0 = q[0], x = q[1], y = q[2], z = q[3];
  double dR[3][3] = {{w0 * w0 + x * x - y * y - z * z, 2 * (x * y - w0 * z), 2 * (x * z + w0 * y)},
                     {2 * (x * y + w0 * z), w0 * w0 - x * x + y * y - z * z, 2 * (y * z - w0 * x)},
                     {2 * (x * z - w0 * y), 2 * (y * z + w0 * x), w0 * w0 - x * x - y * y + z * z}};
  // dt takes the source centroid mx to mx + om (my - mx) (to my when om = 1)
  double dt[3];
  for (int a = 0; a < 3; ++a)
    dt[a] = mx[a] + om * (my[a] - mx[a]) - (dR[a][0] * mx[0] + dR[a][1] * mx[1] + dR[a][2] * mx[2]);
  // x'_new = dR x' + dt with x' = R x + t - c
  double Rn[9], tc[3];
  for (int a = 0; a < 3; ++a) {
    for (int b = 0; b < 3; ++b)
      Rn[a + 3 * b] = dR[a][0] * tr.R[0 + 3 * b] + dR[a][1] * tr.R[1 + 3 * b] + dR[a][2] * tr.R[2 + 3 * b];
    tc[a] = tr.t[a] - tr.c[a];
  }
  double dtr2 = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double v = dR[a][0] * tc[0] + dR[a][1] * tc[1] + dR[a][2] * tc[2] + dt[a];
    const double tn = tr.c[a] + v;
    dtr2 += (tn - tr.t[a]) * (tn - tr.t[a]);
    tr.t[a] = tn;
  }
  for (int a = 0; a < 9; ++a) tr.R[a] = Rn[a];
  const double ca = fmin(1.0, fmax(-1.0, 0.5 * (dR[0][0] + dR[1][1] + dR[2][2] - 1.0)));
  tr.dang = acos(ca);
  tr.dtr = sqrt(dtr2);
  // sig measured in this pass (at the pose before this step) sets the
  // variance scale of the next pass
  if (sig >= 0.0) {
    tr.sig = sig;
    tr.g = fmax(sig / tr.sbar, kGMin);
  }
  tr.a *= bp.beta;
  if (tr.a < tr.a_stop) tr.a = 0.0;
  if (!annealing && tr.dang < bp.tol_rot && tr.dtr < bp.tol_trans) {
    tr.status = 1;
    done[k] = 1;
  } else if (tr.iters >= bp.max_iter) {
    tr.status = 2;
    done[k] = 1;
  }
  trials[k] = tr;
}

// After the isotropic iterations: each trial's energy at its pose, from the
// item sums of one energy pass (in item order); the queue is re-armed.
__global__ __launch_bounds__(256) void k_batch_energy(const double *__restrict__ part,
                                                      const int *__restrict__ titem, int K,
                                                      BTrial *__restrict__ trials, unsigned *__restrict__ queue) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *queue = 0u;
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= K) return;
  double e = 0.0;
  for (int i = titem[k]; i < titem[k + 1]; ++i) e += part[(size_t)i * kStat];
  trials[k].E = e * kLn4Pow1p5;  // e was computed with variances times ln 4
}

// ---------------------------------------------------------------------------
// Refinement: BFGS on -log E with full covariances, one pass per trial pose
// ---------------------------------------------------------------------------
struct alignas(16) BCov {  // a target's covariance (FP32, symmetric)
  float s11, s12, s13, s22, s23, s33;
  float lam;  // upper bound of its largest eigenvalue
  float pad;
};

// A block of 64 consecutive sources of one trial (the last one may be
// shorter): their bounding box and the largest eigenvalue bound of their
// covariances.
struct alignas(16) BBound {
  float lo[3], lam;
  float hi[3], pad;
};

// A wave skips a block of sources when, for every pair, the exponent
// d^T S^-1 d >= |d|^2 / lambda_max(S) >= gap^2 / (lam_target + lam_source)
// exceeds kFarExp: each skipped term is below exp(-kFarExp / 2) = 7e-13 of
// that pair's own peak, far below the FP32 rounding of the terms summed.
// DUGMA_FAR_SKIP=0 in the environment when the covariances are set turns
// the skip off for that batch (for comparisons).
constexpr float kFarExp = 56.0f;

// Per-trial optimiser state (all FP64).  The pose parameters are
// p = [w L, dt] for R <- exp([w]x) R, t <- t + dt (L: the trial's RMS source
// radius, so that both blocks are lengths); f = -log E without (2 pi)^-3/2.
struct BRef {
  double R[9], t[3];    // accepted pose
  double cR[9], ct[3];  // candidate pose: the next pass evaluates it
  double gq[6], d[6];   // gradient of f at the accepted pose; search direction
  double H[36];         // inverse Hessian approximation
  double f, slope, alpha, L, w0, smax, E;
  int state;            // 0: the start pose is being evaluated; 1: line search
  int scaled;           // the first curvature pair has scaled H
  int bt;               // halvings in the current line search
  int it;               // accepted steps
  int status;           // 0 running, 1 converged, 2 step limit, 3 no weight
  int warm;             // 1: H (and scaled) carried over from the coarse level
  int pad[2];
};

struct RParams {
  double tol_rot, tol_trans;
  int max_iter;
  int newton;  // 1: Newton direction at each level's first pose, 2: at every accepted pose
};

constexpr int kRefStat = 28;  // per refinement item: E, gradient (6), Hessian (21)

// Candidate pose as every pass sees it: z = cR^T (y - (ct - c)) and
// S' = cR^T S cR, FP64 -> FP32 (position hi + lo).
__global__ __launch_bounds__(256) void k_batchf_transform(
    const Rec3I *__restrict__ T, const BCov *__restrict__ TC, const int *__restrict__ tgt_trial, int i0,
    int i1, const BTrial *__restrict__ trials, const BRef *__restrict__ ref, const int *__restrict__ done,
    Rec3F *__restrict__ Z) {
  const int i = i0 + blockIdx.x * 256 + threadIdx.x;
  if (i >= i1) return;
  const int k = tgt_trial[i];
  if (done[k]) return;
  const BRef &r = ref[k];
  const double *c = trials[k].c;
  const Rec3I y = T[i];
  const double u[3] = {(double)y.hx - (r.ct[0] - c[0]), (double)y.hy - (r.ct[1] - c[1]),
                       (double)y.hz - (r.ct[2] - c[2])};
  double z[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) z[a] = r.cR[3 * a] * u[0] + r.cR[3 * a + 1] * u[1] + r.cR[3 * a + 2] * u[2];
  const BCov cv = TC[i];
  const double S[9] = {cv.s11, cv.s12, cv.s13, cv.s12, cv.s22, cv.s23, cv.s13, cv.s23, cv.s33};
  double SR[9], Sp[9];  // S cR, then cR^T S cR (column-major)
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
      SR[a + 3 * b] = S[a] * r.cR[3 * b] + S[a + 3] * r.cR[1 + 3 * b] + S[a + 6] * r.cR[2 + 3 * b];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
      Sp[a + 3 * b] = r.cR[3 * a] * SR[3 * b] + r.cR[3 * a + 1] * SR[1 + 3 * b] + r.cR[3 * a + 2] * SR[2 + 3 * b];
  Rec3F o;
  o.hx = (float)z[0];
  o.hy = (float)z[1];
  o.hz = (float)z[2];
  o.lx = (float)(z[0] - (double)o.hx);
  o.ly = (float)(z[1] - (double)o.hy);
  o.lz = (float)(z[2] - (double)o.hz);
  o.s11 = (float)Sp[0];
  o.s22 = (float)Sp[4];
  o.s33 = (float)Sp[8];
  o.s12 = (float)(0.5 * (Sp[3] + Sp[1]));
  o.s13 = (float)(0.5 * (Sp[6] + Sp[2]));
  o.s23 = (float)(0.5 * (Sp[7] + Sp[5]));
  o.pad0 = cv.lam;  // the largest eigenvalue does not change with the rotation
  o.pad1 = o.pad2 = o.pad3 = 0.0f;
  Z[i] = o;
}

// Per item: the targets of a tile (lanes, hi + lo, in the source frame)
// against a chunk of sources (static records, streamed two deep through
// SGPRs): per target E, EW = sum e S^-1 d and B = sum e (S^-1 d d^T S^-1 -
// S^-1), FP32 over sub-tiles of kBSub sources, then FP64; the pose
// gradient follows from these target-side sums (pose_contrib: rotation
// about the origin of the frame, i.e. about t, then translation).  Item
// totals: [E, rotation (3), translation (3), Hessian (21, pose_hessian_entry,
// FP32 wave sums)], in the source frame.
__device__ __forceinline__ float wave_min_f(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off));
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ float wave_max_f(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off));
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// Second-order pose statistics of one target lane, from the sums the pair
// loop leaves (a = [E, EW (3), B (6)], EW = sum e S^-1 d, B = sum e (S^-1 d
// d^T S^-1 - S^-1), d = z - x): the Hessian of E for moving the target by
// z -> z + theta x z + delta, keeping its covariance's turn only in the
// gradient.  With J = [-[z]x | I]:  H = J^T B J, plus, from the second-order
// turn of z, -1/2 (EW z^T + z EW^T) + (EW . z) I in the theta block.  To
// first order this is also the Hessian for moving the sources the opposite
// way (the pose step's parameters).  Entry (i, j) of the 6 x 6 (theta,
// delta) matrix, i <= j, in FP32 (it only shapes the search direction); one
// entry at a time keeps few registers live.
__device__ __forceinline__ float pose_hessian_entry(int i, int j, const float *z, const float *a) {
  auto Bm = [&](int u, int v) -> float {  // B, symmetric, from a[4..9]
    const int lo = u < v ? u : v, hi = u < v ? v : u;
    return a[4 + (lo == 0 ? hi : lo == 1 ? 2 + hi : 5)];
  };
  auto Mm = [&](int u, int v) -> float {  // [z]x
    if (u == v) return 0.0f;
    const int w = 3 - u - v;
    return ((u + 1) % 3 == v) ? -z[w] : z[w];
  };
  if (i >= 3) return Bm(i - 3, j - 3);
  if (j >= 3) {  // theta delta: M B
    const int c = j - 3;
    return Mm(i, 0) * Bm(0, c) + Mm(i, 1) * Bm(1, c) + Mm(i, 2) * Bm(2, c);
  }
  float v = 0.0f;  // theta theta: M^T B M
  for (int u = 0; u < 3; ++u)
    for (int w = 0; w < 3; ++w) v += Mm(u, i) * Bm(u, w) * Mm(w, j);
  const float ez = a[1] * z[0] + a[2] * z[1] + a[3] * z[2];
  return v - 0.5f * (a[1 + i] * z[j] + z[i] * a[1 + j]) + (i == j ? ez : 0.0f);
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

template <int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(7, 8))) void k_batchf_pairs(
    const Rec3F *__restrict__ Z, const Rec3F *__restrict__ S, const BBound *__restrict__ SB, float far_exp,
    const BItem *__restrict__ items, const int *__restrict__ order, int nitems, unsigned *__restrict__ queue,
    const int *__restrict__ done, double *__restrict__ part) {
  __shared__ int s_p;
  __shared__ double red[W][kRefStat];
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (;;) {
    if (threadIdx.x == 0) s_p = (int)atomicAdd(queue, 1u);
    __syncthreads();
    const int p = __builtin_amdgcn_readfirstlane(s_p);
    __syncthreads();
    if (p >= nitems) break;
    const int idx = __builtin_amdgcn_readfirstlane(order[p]);
    const BItem it = items[idx];
    if (done[it.trial]) continue;
    const int ti = it.t0 + wave * 64 + lane;
    const bool valid = (wave * 64 + lane) < it.tn;
    const Rec3F pz = Z[valid ? ti : it.t0];
    double dacc[10];
#pragma unroll
    for (int a = 0; a < 10; ++a) dacc[a] = 0.0;
    const int s_end = it.s0 + it.sn;
    const int s_run = wave * 64 < it.tn ? s_end : it.s0;  // idle waves skip the pairs
    // the wave's targets: bounding box and largest eigenvalue bound (lanes
    // past the tile hold the tile's first target)
    const float zlo0 = wave_min_f(pz.hx), zlo1 = wave_min_f(pz.hy), zlo2 = wave_min_f(pz.hz);
    const float zhi0 = wave_max_f(pz.hx), zhi1 = wave_max_f(pz.hy), zhi2 = wave_max_f(pz.hz);
    const float zlam = wave_max_f(pz.pad0);
    Rec3F qa = S[it.s0], qb = S[min(it.s0 + 1, s_end - 1)];  // it.sn >= 1
    bool primed = true;  // qa, qb hold records j0 and j0 + 1
    for (int j0 = it.s0, blk = it.blk; j0 < s_run; j0 += kBSub, ++blk) {
      const BBound bd = SB[blk];
      const float g0 = fmaxf(0.0f, fmaxf(bd.lo[0] - zhi0, zlo0 - bd.hi[0]));
      const float g1 = fmaxf(0.0f, fmaxf(bd.lo[1] - zhi1, zlo1 - bd.hi[1]));
      const float g2 = fmaxf(0.0f, fmaxf(bd.lo[2] - zhi2, zlo2 - bd.hi[2]));
      const float gap2 = fmaf(g0, g0, fmaf(g1, g1, g2 * g2));
      if (__builtin_amdgcn_readfirstlane(gap2 > far_exp * (zlam + bd.lam) ? 1 : 0)) {
        primed = false;  // far from every target of the wave
        continue;
      }
      if (!primed) {
        qa = S[j0];
        qb = S[min(j0 + 1, s_end - 1)];
        primed = true;
      }
      float acc[10];
#pragma unroll
      for (int a = 0; a < 10; ++a) acc[a] = 0.0f;
      const int j1 = min(j0 + kBSub, s_end);
      for (int j = j0; j < j1; ++j) {
        const Rec3F q = qa;
        qa = qb;
        qb = S[min(j + 2, s_end - 1)];
        pair_eval<1, 2>(pz, q, acc);
      }
#pragma unroll
      for (int a = 0; a < 10; ++a) dacc[a] += (double)acc[a];
    }
    {
      double r[7] = {0, 0, 0, 0, 0, 0, 0};
      const double zero[3] = {0.0, 0.0, 0.0};
      if (valid) pose_contrib(pz, dacc, zero, r);
#pragma unroll
      for (int a = 0; a < 7; ++a) {
        const double v = wave_sum(r[a]);
        if (lane == 0) red[wave][a] = v;
      }
    }
    {
      const float z[3] = {pz.hx, pz.hy, pz.hz};
      float af[10];
#pragma unroll
      for (int a = 0; a < 10; ++a) af[a] = (float)dacc[a];
      int q = 7;
#pragma unroll
      for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = i; j < 6; ++j) {
          const float v = wave_sum_f(valid ? pose_hessian_entry(i, j, z, af) : 0.0f);
          if (lane == 0) red[wave][q] = (double)v;
          ++q;
        }
    }
    __syncthreads();
    if (threadIdx.x < kRefStat) {
      double v = red[0][threadIdx.x];
#pragma unroll
      for (int w = 1; w < W; ++w) v += red[w][threadIdx.x];
      part[(size_t)idx * kStat + threadIdx.x] = v;
    }
    __syncthreads();
  }
}

// R <- exp([w]x) R, t <- t + dt for p = [w L, dt] (Rodrigues, FP64).
__device__ void ref_step(const double *R, const double *t, const double *s, double L, double *Rn,
                         double *tn) {
  const double w[3] = {s[0] / L, s[1] / L, s[2] / L};
  const double th = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  const double A = th < 1e-12 ? 1.0 : sin(th) / th, B = th < 1e-12 ? 0.5 : (1.0 - cos(th)) / (th * th);
  const double K[9] = {0, w[2], -w[1], -w[2], 0, w[0], w[1], -w[0], 0};  // [w]x, column-major
  double E[9];
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) {
      double k2 = 0.0;
      for (int b = 0; b < 3; ++b) k2 += K[a + 3 * b] * K[b + 3 * c];
      E[a + 3 * c] = (a == c ? 1.0 : 0.0) + A * K[a + 3 * c] + B * k2;
    }
  for (int a = 0; a < 3; ++a)
    for (int c = 0; c < 3; ++c) Rn[a + 3 * c] = E[a] * R[3 * c] + E[a + 3] * R[1 + 3 * c] + E[a + 6] * R[2 + 3 * c];
  for (int a = 0; a < 3; ++a) tn[a] = t[a] + s[3 + a];
}

// The next line search: d = -H g (the gradient if that is no descent
// direction), first step w0 (a kernel width) until H is scaled, at most smax;
// the candidate pose follows.  Returns false at a zero gradient.
__device__ bool ref_direction(BRef &r) {
  double slope = 0.0, dn = 0.0;
  for (int a = 0; a < 6; ++a) {
    double v = 0.0;
    for (int b = 0; b < 6; ++b) v -= r.H[6 * a + b] * r.gq[b];
    r.d[a] = v;
    slope += r.gq[a] * v;
  }
  if (!(slope < 0.0)) {  // restart from the gradient
    for (int a = 0; a < 36; ++a) r.H[a] = (a % 7 == 0) ? 1.0 : 0.0;
    r.scaled = 0;
    slope = 0.0;
    for (int a = 0; a < 6; ++a) {
      r.d[a] = -r.gq[a];
      slope -= r.gq[a] * r.gq[a];
    }
    if (!(slope < 0.0)) return false;
  }
  for (int a = 0; a < 6; ++a) dn += r.d[a] * r.d[a];
  dn = sqrt(dn);
  double alpha = r.scaled ? 1.0 : r.w0 / dn;
  if (alpha * dn > r.smax) alpha = r.smax / dn;
  r.slope = slope;
  r.alpha = alpha;
  r.bt = 0;
  double s[6];
  for (int a = 0; a < 6; ++a) s[a] = alpha * r.d[a];
  ref_step(r.R, r.t, s, r.L, r.cR, r.ct);
  return true;
}

// The step ref_direction just proposed, from an inverse Hessian scaled by a
// curvature pair (here or on the previous level), is already below the
// tolerances: the trial has converged at its accepted pose, without a pass
// to evaluate that step.
__device__ bool next_step_converged(const BRef &r, const RParams &rp) {
  if (!r.scaled) return false;
  double pa = 0.0, pt = 0.0;
  for (int a = 0; a < 3; ++a) {
    pa += r.d[a] * r.d[a];
    pt += r.d[3 + a] * r.d[3 + a];
  }
  return r.alpha * sqrt(pa) / r.L < rp.tol_rot && r.alpha * sqrt(pt) < rp.tol_trans;
}

// The Newton inverse Hessian of f = -log E in p = [w L, dt], from a pass's
// sums st = [E, gradient (6), Hessian (21, upper triangle)] in the source
// frame: both turned to the world frame by cR (g = cR g' per 3-block, H =
// Q H' Q^T), H_f = -H_E / E + g g^T / E^2, then scaled to p.  False unless
// H_f is positive definite (Cholesky pivots above 1e-12 of the largest
// diagonal entry); g: the world gradient of E.
__device__ bool newton_inverse(const double *st, const double *cR, const double *g, double E, double L,
                               double *Hinv) {
  double Hs[6][6];
  int q = 7;
  for (int i = 0; i < 6; ++i)
    for (int j = i; j < 6; ++j) Hs[i][j] = Hs[j][i] = st[q++];
  double Q[6][6];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) Q[i][j] = 0.0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) Q[a][b] = Q[3 + a][3 + b] = cR[a + 3 * b];
  double T[6][6], H[6][6];
  for (int a = 0; a < 6; ++a)
    for (int j = 0; j < 6; ++j) {
      double v = 0.0;
      for (int i = 0; i < 6; ++i) v += Q[a][i] * Hs[i][j];
      T[a][j] = v;
    }
  const double sc[6] = {1.0 / L, 1.0 / L, 1.0 / L, 1.0, 1.0, 1.0};
  double dmax = 0.0;
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) {
      double v = 0.0;
      for (int j = 0; j < 6; ++j) v += T[a][j] * Q[b][j];
      H[a][b] = (-v / E + g[a] * g[b] / (E * E)) * sc[a] * sc[b];
    }
  for (int a = 0; a < 6; ++a) dmax = fmax(dmax, H[a][a]);
  if (!(dmax > 0.0) || !isfinite(dmax)) return false;
  // Cholesky H = C C^T (lower C in place)
  for (int j = 0; j < 6; ++j) {
    double d = H[j][j];
    for (int k = 0; k < j; ++k) d -= H[j][k] * H[j][k];
    if (!(d > 1e-12 * dmax)) return false;
    d = sqrt(d);
    H[j][j] = d;
    for (int i = j + 1; i < 6; ++i) {
      double v = H[i][j];
      for (int k = 0; k < j; ++k) v -= H[i][k] * H[j][k];
      H[i][j] = v / d;
    }
  }
  // H^-1, one unit vector at a time
  for (int c = 0; c < 6; ++c) {
    double y[6], x[6];
    for (int i = 0; i < 6; ++i) {
      double v = i == c ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) v -= H[i][k] * y[k];
      y[i] = v / H[i][i];
    }
    for (int i = 5; i >= 0; --i) {
      double v = y[i];
      for (int k = i + 1; k < 6; ++k) v -= H[k][i] * x[k];
      x[i] = v / H[i][i];
    }
    for (int i = 0; i < 6; ++i) Hinv[6 * i + c] = x[i];
  }
  return true;
}

// One thread per trial: the pass's energy and gradient at the candidate
// pose, then the optimiser's next move (dugma_register_energy's logic, and
// a trial also stops when its next quasi-Newton step is below the
// tolerances).
// One workgroup per running trial: its items' kRefStat statistics staged
// in LDS a block of items at a time, thread a summing statistic a in item
// order, into fold[k]; re-arms the queue.
__global__ __launch_bounds__(kUpd) void k_batchf_fold(const double *__restrict__ part,
                                                      const int *__restrict__ titem, int k0, int k1,
                                                      const int *__restrict__ done,
                                                      unsigned *__restrict__ queue, double *__restrict__ fold) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *queue = 0u;
  const int k = k0 + blockIdx.x;
  if (k >= k1 || done[k]) return;
  __shared__ double s_part[kUpdItems * kStat];
  const int i0 = titem[k], i1 = titem[k + 1];
  double acc = 0.0;
  for (int ib = i0; ib < i1; ib += kUpdItems) {
    const int nb = min(kUpdItems, i1 - ib);
    for (int e = threadIdx.x; e < nb * kStat; e += kUpd) s_part[e] = part[(size_t)ib * kStat + e];
    __syncthreads();
    if (threadIdx.x < kRefStat)
      for (int i = 0; i < nb; ++i) acc += s_part[i * kStat + threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x < kRefStat) fold[(size_t)k * kStat + threadIdx.x] = acc;
}

__global__ __launch_bounds__(256) void k_batchf_update(const double *__restrict__ fold, int k0, int k1,
                                                       BRef *__restrict__ ref, int *__restrict__ done,
                                                       RParams rp) {
  const int k = k0 + blockIdx.x * 256 + threadIdx.x;
  if (k >= k1 || done[k]) return;
  BRef r = ref[k];
  double st[kRefStat];
  for (int a = 0; a < kRefStat; ++a) st[a] = fold[(size_t)k * kStat + a];
  const double E = st[0];
  // gradient (rotation, translation) back from the source frame: g = cR g'
  double g[6];
  for (int a = 0; a < 3; ++a) {
    g[a] = r.cR[a] * st[1] + r.cR[a + 3] * st[2] + r.cR[a + 6] * st[3];
    g[3 + a] = r.cR[a] * st[4] + r.cR[a + 3] * st[5] + r.cR[a + 6] * st[6];
  }
  const bool good = E > 0.0 && isfinite(E);
  double gq[6];  // gradient of -log E in p
  for (int a = 0; a < 6; ++a) gq[a] = a < 3 ? -g[a] / (E * r.L) : -g[a] / E;
  bool finished = false;
  if (r.state == 0) {
    if (!good) {
      r.status = 3;
      finished = true;
    } else {
      for (int a = 0; a < 9; ++a) r.R[a] = r.cR[a];
      for (int a = 0; a < 3; ++a) r.t[a] = r.ct[a];
      r.f = -log(E);
      r.E = E;
      for (int a = 0; a < 6; ++a) r.gq[a] = gq[a];
      if (!r.warm) {
        for (int a = 0; a < 36; ++a) r.H[a] = (a % 7 == 0) ? 1.0 : 0.0;
        r.scaled = 0;
      }
      if (rp.newton >= 1 && newton_inverse(st, r.cR, g, E, r.L, r.H)) r.scaled = 1;
      r.state = 1;
      if (rp.max_iter < 1 || !ref_direction(r)) {
        r.status = rp.max_iter < 1 ? 2 : 1;
        finished = true;
      } else if (next_step_converged(r, rp)) {  // a carried-over H already says so
        r.status = 1;
        finished = true;
      }
    }
  } else if (good && -log(E) <= r.f + 1e-4 * r.alpha * r.slope) {
    // accepted: BFGS update with s = alpha d, y = gq_new - gq
    r.it += 1;
    double s[6], y[6], sy = 0.0, yy = 0.0, ss = 0.0;
    for (int a = 0; a < 6; ++a) {
      s[a] = r.alpha * r.d[a];
      y[a] = gq[a] - r.gq[a];
      sy += s[a] * y[a];
      yy += y[a] * y[a];
      ss += s[a] * s[a];
    }
    if (sy > 1e-12 * sqrt(ss * yy)) {
      if (!r.scaled) {
        for (int a = 0; a < 36; ++a) r.H[a] *= sy / yy;
        r.scaled = 1;
      }
      const double rr = 1.0 / sy;
      double Hy[6], yHy = 0.0;
      for (int a = 0; a < 6; ++a) {
        double v = 0.0;
        for (int b = 0; b < 6; ++b) v += r.H[6 * a + b] * y[b];
        Hy[a] = v;
        yHy += y[a] * v;
      }
      for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b)
          r.H[6 * a + b] += -rr * (Hy[a] * s[b] + s[a] * Hy[b]) + (rr * rr * yHy + rr) * s[a] * s[b];
    }
    if (rp.newton >= 2 && newton_inverse(st, r.cR, g, E, r.L, r.H)) r.scaled = 1;
    for (int a = 0; a < 9; ++a) r.R[a] = r.cR[a];
    for (int a = 0; a < 3; ++a) r.t[a] = r.ct[a];
    r.f = -log(E);
    r.E = E;
    for (int a = 0; a < 6; ++a) r.gq[a] = gq[a];
    const double ang = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]) / r.L;
    const double tr = sqrt(s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
    if (ang < rp.tol_rot && tr < rp.tol_trans) {
      r.status = 1;
      finished = true;
    } else if (r.it >= rp.max_iter) {
      r.status = 2;
      finished = true;
    } else if (!ref_direction(r)) {
      r.status = 1;
      finished = true;
    } else if (next_step_converged(r, rp)) {
      r.status = 1;
      finished = true;
    }
  } else {
    // no sufficient decrease: halve the step
    r.alpha *= 0.5;
    r.bt += 1;
    if (r.bt >= 40) {
      r.status = 1;  // stationary to working precision
      finished = true;
    } else {
      double s[6];
      for (int a = 0; a < 6; ++a) s[a] = r.alpha * r.d[a];
      ref_step(r.R, r.t, s, r.L, r.cR, r.ct);
    }
  }
  ref[k] = r;
  if (finished) done[k] = 1;
}

// One resolution of every trial's clouds.  Level 0 holds all kept points;
// level 1 every kCoarseStep-th target and source of each trial (same
// centroid, same per-point variances), for the early, broad iterations;
// level 2 every 2nd point, between the two in the refinement only.
constexpr int kLevels = 3;
constexpr int kLevelStep[kLevels] = {1, 4, 2};
struct BLevel {
  int ntgt = 0, nsrc = 0, nitems = 0;
  Rec3I *dT = nullptr;       // targets (all trials), relative to each trial's centroid
  BZ *dZ = nullptr;          // per-iteration targets in each trial's source frame
  int *dtgt_trial = nullptr; // trial of each target
  BSrc *dS = nullptr;        // sources (all trials), as given
  BItem *ditems = nullptr;
  int *dorder = nullptr;     // queue position -> item, longest first
  int *dident = nullptr;     // identity order for one-trial launches
  int *dactive = nullptr;    // the running trials' items only, longest first (tail passes)
  int *dtitem = nullptr;     // CSR: items of trial k are [titem[k], titem[k+1])
  int *dttgt = nullptr;      // CSR: targets of trial k are [ttgt[k], ttgt[k+1])
  unsigned *dnn = nullptr;   // per target: smallest |d|^2 of the current pass (float bits)
  double *dpart = nullptr;   // per-item statistics
  size_t items_cap = 0;      // capacity of ditems / dorder / dident / dpart
  std::vector<int> tb, nt, sb, ns;  // per trial: target / source base and count
  std::vector<int> bb;              // per trial: first 64-source block (prefix of ceil(ns / 64))
  int nblocks = 0;                  // 64-source blocks of the level
  std::vector<int> titem;
  std::vector<int> horder, hitem_trial;  // host copies: queue order; trial of each item
  long long pairs = 0;              // target x source pairs of one full iteration
  // launch shape of this level's persistent pair kernels
  int waves = 4;                        // waves per workgroup: target tile = 64 * waves
  int chunks[kClasses] = {1, 1, 2, 4};  // source chunks per target tile, per size class
  int grid_waves = 8192;                // resident waves
};

constexpr int kCoarseStep = kLevelStep[1];

}  // namespace dugma

using namespace dugma;

struct dugma_batch {
  int ntrials = 0;
  hipStream_t stream = nullptr;
  BLevel lv[kLevels];
  BTrial *dtrials = nullptr;
  int *ddone = nullptr;
  double *dfold = nullptr;  // per trial: folded item sums and sig of the last pass (kStat doubles)
  unsigned *dqueue = nullptr;
  std::vector<double> a0, astop, cent, sbar;  // per trial
  // refinement with full covariances, per level (the level's points, their
  // covariances times the level's point step)
  BCov *dTC[kLevels] = {};                  // per target: covariance
  Rec3F *dSF[kLevels] = {};                 // per source: position and covariance (static)
  Rec3F *dZF[kLevels] = {};                 // per target at each pass: in the trial's source frame
  BBound *dSB[kLevels] = {};                // per 64-source block: box and eigenvalue bound
  float far_exp = kFarExp;                  // skip threshold (infinite: no skip)
  BRef *dref = nullptr;                     // per trial: optimiser state
  std::vector<int> keepT, keepS;            // kept input indices per trial (level-0 order)
  std::vector<BSrc> hsrc0;                  // host copy of the level-0 sources
  std::vector<double> refL, refW0;          // per trial: RMS source radius, kernel width
  // launch shape
  // (launch shapes: per level, BLevel::waves / chunks / grid_waves)
};

static void free_level(BLevel &L) {
  void *ptrs[] = {L.dT, L.dZ, L.dtgt_trial, L.dS, L.ditems, L.dorder, L.dident, L.dactive,
                  L.dtitem, L.dttgt, L.dnn, L.dpart};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
}

static void free_batch(dugma_batch *b) {
  if (!b) return;
  for (int l = 0; l < kLevels; ++l) {
    free_level(b->lv[l]);
    void *lp[] = {b->dTC[l], b->dSF[l], b->dZF[l], b->dSB[l]};
    for (void *p : lp)
      if (p) (void)hipFree(p);
  }
  void *ptrs[] = {b->dtrials, b->ddone, b->dfold, b->dqueue, b->dref};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

template <class T>
static int bmalloc(T **p, size_t n) {
  *p = nullptr;
  hipError_t e = hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return hip_fail(e, "hipMalloc(batch)");
  }
  return DUGMA_OK;
}

// Work items of one level for the batch's current launch shape: per trial,
// target tiles of 64 * waves points times the source chunks of its size
// class; queue order longest first.  (Re)allocates the item arrays when they
// grow.
static int make_items(dugma_batch *b, BLevel &L) {
  const int K = b->ntrials, tile_t = 64 * L.waves;
  std::vector<BItem> items;
  L.titem.assign(K + 1, 0);
  L.bb.assign(K, 0);
  L.nblocks = 0;
  for (int k = 0; k < K; ++k) {
    L.titem[k] = (int)items.size();
    L.bb[k] = L.nblocks;
    const int nt = L.nt[k], ns = L.ns[k];
    L.nblocks += (ns + kBSub - 1) / kBSub;
    if (nt == 0 || ns == 0) continue;
    const int nch = std::max(1, std::min(L.chunks[size_class(ns)], ns));
    // chunks of whole 64-source blocks
    const int tile_s = ((ns + nch - 1) / nch + kBSub - 1) / kBSub * kBSub;
    for (int t0 = 0; t0 < nt; t0 += tile_t)
      for (int s0 = 0; s0 < ns; s0 += tile_s) {
        BItem it;
        memset(&it, 0, sizeof(it));
        it.trial = k;
        it.t0 = L.tb[k] + t0;
        it.tn = std::min(tile_t, nt - t0);
        it.s0 = L.sb[k] + s0;
        it.sn = std::min(tile_s, ns - s0);
        it.blk = L.bb[k] + s0 / kBSub;
        items.push_back(it);
      }
  }
  L.titem[K] = (int)items.size();
  std::vector<int> order(items.size()), ident(items.size());
  std::iota(order.begin(), order.end(), 0);
  std::iota(ident.begin(), ident.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int u, int v) {
    return (long long)items[u].tn * items[u].sn > (long long)items[v].tn * items[v].sn;
  });
  if (items.size() > L.items_cap) {
    void *old[] = {L.ditems, L.dorder, L.dident, L.dactive, L.dpart};
    for (void *p : old)
      if (p) (void)hipFree(p);
    L.ditems = nullptr;
    L.dorder = L.dident = L.dactive = nullptr;
    L.dpart = nullptr;
    L.items_cap = 0;
    int rc;
    if ((rc = bmalloc(&L.ditems, items.size())) || (rc = bmalloc(&L.dorder, items.size())) ||
        (rc = bmalloc(&L.dident, items.size())) || (rc = bmalloc(&L.dactive, items.size())) ||
        (rc = bmalloc(&L.dpart, items.size() * kStat)))
      return rc;
    L.items_cap = items.size();
  }
  L.hitem_trial.resize(items.size());
  for (size_t i = 0; i < items.size(); ++i) L.hitem_trial[i] = items[i].trial;
  L.horder = order;
  L.nitems = (int)items.size();
  if (!items.empty()) {
    DUGMA_HIP(hipMemcpy(L.ditems, items.data(), items.size() * sizeof(BItem), hipMemcpyHostToDevice));
    DUGMA_HIP(hipMemcpy(L.dorder, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice));
    DUGMA_HIP(hipMemcpy(L.dident, ident.data(), ident.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  DUGMA_HIP(hipMemcpy(L.dtitem, L.titem.data(), L.titem.size() * sizeof(int), hipMemcpyHostToDevice));
  return DUGMA_OK;
}

static int make_all_items(dugma_batch *b) {
  int rc = DUGMA_OK;
  for (int l = 0; l < kLevels && !rc; ++l) rc = make_items(b, b->lv[l]);
  return rc;
}

// Host copies of one level's points, built per trial, then uploaded.
struct LevelHost {
  std::vector<Rec3I> tg;
  std::vector<BSrc> sr;
  std::vector<int> ttrial, ttgt;
};

static int upload_level(dugma_batch *b, BLevel &L, const LevelHost &H) {
  const int K = b->ntrials;
  L.ntgt = (int)H.tg.size();
  L.nsrc = (int)H.sr.size();
  int rc;
  if ((rc = bmalloc(&L.dT, H.tg.size())) || (rc = bmalloc(&L.dZ, H.tg.size())) ||
      (rc = bmalloc(&L.dtgt_trial, H.ttrial.size())) || (rc = bmalloc(&L.dS, H.sr.size())) ||
      (rc = bmalloc(&L.dtitem, (size_t)K + 1)) || (rc = bmalloc(&L.dttgt, (size_t)K + 1)) ||
      (rc = bmalloc(&L.dnn, H.tg.size())))
    return rc;
  struct Up {
    void *d;
    const void *h;
    size_t n;
  } ups[] = {{L.dT, H.tg.data(), H.tg.size() * sizeof(Rec3I)},
             {L.dtgt_trial, H.ttrial.data(), H.ttrial.size() * sizeof(int)},
             {L.dS, H.sr.data(), H.sr.size() * sizeof(BSrc)},
             {L.dttgt, H.ttgt.data(), H.ttgt.size() * sizeof(int)}};
  for (const Up &u : ups)
    if (u.n) DUGMA_HIP(hipMemcpy(u.d, u.h, u.n, hipMemcpyHostToDevice));
  return DUGMA_OK;
}

// Variance of point i of trial k's target (which = 0) or source (1) cloud:
// > 0 for an isotropic positive covariance, -1 for a point to leave out,
// -2 for a covariance that is not s I.
typedef double (*VarFn)(const void *ctx, int k, int which, int i);

struct BlockVars {
  const double *const *XC, *const *YC;
};
static double block_var(const void *c, int k, int which, int i) {
  const BlockVars *v = (const BlockVars *)c;
  const double *C = (which ? v->YC[k] : v->XC[k]) + (size_t)i * 9;
  for (int a = 0; a < 9; ++a)
    if (!std::isfinite(C[a])) return -1.0;
  if (C[1] != 0 || C[2] != 0 || C[3] != 0 || C[5] != 0 || C[6] != 0 || C[7] != 0 ||
      C[0] != C[4] || C[0] != C[8])
    return -2.0;
  const float c0 = (float)C[0];
  return c0 > 0 ? (double)c0 : -1.0;
}

struct IsoVars {
  const double *const *sX, *const *sY;
};
static double iso_var(const void *c, int k, int which, int i) {
  const IsoVars *v = (const IsoVars *)c;
  const double s = (which ? v->sY[k] : v->sX[k])[i];
  const float s32 = (float)s;
  return (std::isfinite(s) && s32 > 0) ? (double)s32 : -1.0;
}

// Order of one trial's points for the refinement's far-field skip: points
// whose variance exceeds 4x the trial's median (outliers, whose wide kernels
// would widen every bound they share a wave with) after the rest, each group
// along a 30-bit Morton curve over the trial's bounding box, so that 64
// consecutive points (a wave's targets, a sub-tile of sources) lie close
// together -- and so do every 2nd and every 4th, the sparser levels.
// p: n x stride doubles (x, y, z first), v: n variances.
static std::vector<int> spatial_order(const double *p, int stride, const double *v, int n) {
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  if (n < 2) return ord;
  std::vector<double> vs(v, v + n);
  std::nth_element(vs.begin(), vs.begin() + n / 2, vs.end());
  const double big = 4.0 * vs[n / 2];
  double lo[3] = {p[0], p[1], p[2]}, hi[3] = {p[0], p[1], p[2]};
  for (int i = 1; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], p[(size_t)i * stride + a]);
      hi[a] = std::max(hi[a], p[(size_t)i * stride + a]);
    }
  auto spread = [](uint64_t x) {  // 10 bits -> every 3rd bit
    x &= 0x3ff;
    x = (x | (x << 16)) & 0x30000ffull;
    x = (x | (x << 8)) & 0x300f00full;
    x = (x | (x << 4)) & 0x30c30c3ull;
    x = (x | (x << 2)) & 0x9249249ull;
    return x;
  };
  std::vector<uint64_t> key(n);
  for (int i = 0; i < n; ++i) {
    uint64_t code = 0;
    for (int a = 0; a < 3; ++a) {
      const double w = hi[a] - lo[a];
      const double u = w > 0 ? (p[(size_t)i * stride + a] - lo[a]) / w : 0.0;
      code |= spread((uint64_t)std::min(1023.0, std::max(0.0, u * 1023.0))) << a;
    }
    key[i] = ((v[i] > big ? 1ull : 0ull) << 40) | code;
  }
  std::stable_sort(ord.begin(), ord.end(), [&](int u, int w) { return key[u] < key[w]; });
  return ord;
}

static int build_batch(int ntrials, const int *N, const int *M, const double *const *X,
                       const double *const *Y, VarFn var, const void *vctx,
                       dugma_batch **out) {
  dugma_batch *b = new dugma_batch();
  b->ntrials = ntrials;
  LevelHost H[kLevels];
  for (int l = 0; l < kLevels; ++l) {
    BLevel &L = b->lv[l];
    L.tb.assign(ntrials, 0);
    L.nt.assign(ntrials, 0);
    L.sb.assign(ntrials, 0);
    L.ns.assign(ntrials, 0);
    H[l].ttgt.assign(ntrials + 1, 0);
  }
  b->a0.resize(ntrials);
  b->astop.resize(ntrials);
  b->sbar.resize(ntrials);
  b->cent.resize(3 * (size_t)ntrials);
  for (int k = 0; k < ntrials; ++k) {
    if (N[k] < 1 || M[k] < 1 || !X[k] || !Y[k]) {
      free_batch(b);
      return fail(DUGMA_ERR_ARG, "dugma_batch_create: empty trial");
    }
    std::vector<double> ty, tv, sy;
    double c[3] = {0, 0, 0};
    for (int i = 0; i < N[k]; ++i) {
      const double *p = X[k] + (size_t)i * 3;
      const double v = var(vctx, k, 0, i);
      if (v == -2.0) {
        free_batch(b);
        return fail(DUGMA_ERR_ARG, "dugma_batch_create: covariances must be isotropic (s I)");
      }
      if (v < 0 || !std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2])) continue;
      for (int a = 0; a < 3; ++a) {
        ty.push_back((double)(float)p[a]);
        c[a] += (double)(float)p[a];
      }
      tv.push_back(v);
      b->keepT.push_back(i);
    }
    for (int j = 0; j < M[k]; ++j) {
      const double *p = Y[k] + (size_t)j * 3;
      const double v = var(vctx, k, 1, j);
      if (v == -2.0) {
        free_batch(b);
        return fail(DUGMA_ERR_ARG, "dugma_batch_create: covariances must be isotropic (s I)");
      }
      if (v < 0 || !std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2])) continue;
      for (int a = 0; a < 3; ++a) sy.push_back((double)(float)p[a]);
      sy.push_back(v);
      b->keepS.push_back(j);
    }
    const int nt = (int)tv.size(), ns = (int)(sy.size() / 4);
    {
      // the kept points in spatial order (keepT / keepS follow)
      const std::vector<int> ot = spatial_order(ty.data(), 3, tv.data(), nt);
      std::vector<double> ty2(ty.size()), tv2(tv.size());
      std::vector<int> kt(b->keepT.end() - nt, b->keepT.end());
      for (int i = 0; i < nt; ++i) {
        for (int a = 0; a < 3; ++a) ty2[3 * (size_t)i + a] = ty[3 * (size_t)ot[i] + a];
        tv2[i] = tv[ot[i]];
        b->keepT[b->keepT.size() - nt + i] = kt[ot[i]];
      }
      ty.swap(ty2);
      tv.swap(tv2);
      std::vector<double> svar(ns);
      for (int j = 0; j < ns; ++j) svar[j] = sy[4 * (size_t)j + 3];
      const std::vector<int> os = spatial_order(sy.data(), 4, svar.data(), ns);
      std::vector<double> sy2(sy.size());
      std::vector<int> ks(b->keepS.end() - ns, b->keepS.end());
      for (int j = 0; j < ns; ++j) {
        for (int a = 0; a < 4; ++a) sy2[4 * (size_t)j + a] = sy[4 * (size_t)os[j] + a];
        b->keepS[b->keepS.size() - ns + j] = ks[os[j]];
      }
      sy.swap(sy2);
    }
    if (nt > 0)
      for (int a = 0; a < 3; ++a) c[a] /= nt;
    double r2 = 0.0, vt = 0.0, vs = 0.0;
    for (int l = 0; l < kLevels; ++l) {
      const int step = kLevelStep[l];
      BLevel &L = b->lv[l];
      LevelHost &h = H[l];
      L.tb[k] = (int)h.tg.size();
      for (int i = 0; i < nt; i += step) {
        Rec3I r;
        memset(&r, 0, sizeof(r));
        r.hx = (float)(ty[3 * i] - c[0]);
        r.hy = (float)(ty[3 * i + 1] - c[1]);
        r.hz = (float)(ty[3 * i + 2] - c[2]);
        r.s = (float)(tv[i] * kLn4);
        if (l == 0) {
          r2 += (double)r.hx * r.hx + (double)r.hy * r.hy + (double)r.hz * r.hz;
          vt += tv[i];
        }
        h.tg.push_back(r);
        h.ttrial.push_back(k);
      }
      L.nt[k] = (int)h.tg.size() - L.tb[k];
      L.sb[k] = (int)h.sr.size();
      for (int j = 0; j < ns; j += step) {
        BSrc q;
        q.x = (float)sy[4 * (size_t)j];
        q.y = (float)sy[4 * (size_t)j + 1];
        q.z = (float)sy[4 * (size_t)j + 2];
        q.s = (float)(sy[4 * (size_t)j + 3] * kLn4);
        h.sr.push_back(q);
        if (l == 0) vs += sy[4 * (size_t)j + 3];
      }
      L.ns[k] = (int)h.sr.size() - L.sb[k];
      h.ttgt[k + 1] = h.ttgt[k] + L.nt[k];
      L.pairs += (long long)L.nt[k] * L.ns[k];
    }
    for (int a = 0; a < 3; ++a) b->cent[3 * k + a] = c[a];
    // annealing from a quarter of the target's mean squared radius down to
    // a hundredth of the mean point variance
    b->a0[k] = nt ? 0.25 * r2 / nt : 0.0;
    b->astop[k] = (nt + ns) ? 0.01 * (vt + vs) / (nt + ns) : 0.0;
    b->sbar[k] = (nt ? vt / nt : 0.0) + (ns ? vs / ns : 0.0);
  }
  int rc = DUGMA_OK;
  {
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) rc = hip_fail(e, "hipStreamCreateWithFlags");
  }
  b->hsrc0 = H[0].sr;
  for (int l = 0; l < kLevels && !rc; ++l) rc = upload_level(b, b->lv[l], H[l]);
  if (!rc) rc = bmalloc(&b->dtrials, (size_t)ntrials);
  if (!rc) rc = bmalloc(&b->ddone, (size_t)ntrials);
  if (!rc) rc = bmalloc(&b->dfold, (size_t)ntrials * kStat);
  if (!rc) rc = bmalloc(&b->dqueue, 1);
  if (!rc) {
    hipError_t e = hipMemset(b->dqueue, 0, sizeof(unsigned));
    if (e != hipSuccess) rc = hip_fail(e, "hipMemset(queue)");
  }
  if (!rc) rc = make_all_items(b);
  if (rc) {
    free_batch(b);
    return rc;
  }
  *out = b;
  return DUGMA_OK;
}

extern "C" int dugma_batch_create(int ntrials, int D, const int *N, const int *M,
                                  const double *const *X, const double *const *X_C,
                                  const double *const *Y, const double *const *Y_C,
                                  dugma_batch **out) {
  if (out) *out = nullptr;
  if (!out || !N || !M || !X || !X_C || !Y || !Y_C || ntrials < 1)
    return fail(DUGMA_ERR_ARG, "dugma_batch_create: bad argument");
  if (D != 3) return fail(DUGMA_ERR_ARG, "dugma_batch_create: batched registration is 3-D");
  for (int k = 0; k < ntrials; ++k)
    if (!X_C[k] || !Y_C[k]) return fail(DUGMA_ERR_ARG, "dugma_batch_create: null covariance");
  BlockVars v{X_C, Y_C};
  return build_batch(ntrials, N, M, X, Y, block_var, &v, out);
}

extern "C" int dugma_batch_create_iso(int ntrials, const int *N, const int *M,
                                      const double *const *X, const double *const *sX,
                                      const double *const *Y, const double *const *sY,
                                      dugma_batch **out) {
  if (out) *out = nullptr;
  if (!out || !N || !M || !X || !sX || !Y || !sY || ntrials < 1)
    return fail(DUGMA_ERR_ARG, "dugma_batch_create_iso: bad argument");
  for (int k = 0; k < ntrials; ++k)
    if (!sX[k] || !sY[k]) return fail(DUGMA_ERR_ARG, "dugma_batch_create_iso: null variances");
  IsoVars v{sX, sY};
  return build_batch(ntrials, N, M, X, Y, iso_var, &v, out);
}

extern "C" int dugma_batch_free(dugma_batch *b) {
  if (!b) return fail(DUGMA_ERR_ARG, "dugma_batch_free: null batch");
  (void)hipStreamSynchronize(b->stream);
  free_batch(b);
  return DUGMA_OK;
}

extern "C" int dugma_batch_info(const dugma_batch *b, int *ntrials, int *nitems,
                                long long *pairs_per_iteration) {
  if (!b) return fail(DUGMA_ERR_ARG, "dugma_batch_info: null batch");
  if (ntrials) *ntrials = b->ntrials;
  if (nitems) *nitems = b->lv[0].nitems;
  if (pairs_per_iteration) *pairs_per_iteration = b->lv[0].pairs;
  return DUGMA_OK;
}

template <int W>
static void launch_pairs(dugma_batch *b, const BLevel &L, int grid, const int *order, int ni, bool sigma,
                         bool energy = false) {
  if (energy)  // the energy pass after the iterations (no sigma measurement)
    hipLaunchKernelGGL((k_batch_pairs<W, false, true>), dim3(grid), dim3(64 * W), 0, b->stream, L.dT, L.dZ,
                       L.dS, L.ditems, order, ni, b->dqueue, b->ddone, L.dnn, L.dpart);
  else if (sigma)
    hipLaunchKernelGGL((k_batch_pairs<W, true, false>), dim3(grid), dim3(64 * W), 0, b->stream, L.dT, L.dZ,
                       L.dS, L.ditems, order, ni, b->dqueue, b->ddone, L.dnn, L.dpart);
  else
    hipLaunchKernelGGL((k_batch_pairs<W, false, false>), dim3(grid), dim3(64 * W), 0, b->stream, L.dT, L.dZ,
                       L.dS, L.ditems, order, ni, b->dqueue, b->ddone, L.dnn, L.dpart);
}

// One iteration of trials [k0, k1) on level l.  grid: workgroups of the
// persistent pair kernel (0: from the launch shape).
static int launch_iteration(dugma_batch *b, int l, int k0, int k1, const BParams &bp,
                            bool one_trial, int grid, const int *order_in = nullptr, int ni_in = 0) {
  const BLevel &L = b->lv[l];
  int t0 = 0, t1 = L.ntgt;
  const int *order = order_in ? order_in : L.dorder;
  int ni = order_in ? ni_in : L.nitems;
  if (one_trial) {
    // items of trial k0 are contiguous: [titem[k0], titem[k0+1])
    const int i0 = L.titem[k0];
    ni = L.titem[k0 + 1] - i0;
    order = L.dident + i0;
    t0 = L.tb[k0];
    t1 = t0 + L.nt[k0];
  }
  if (t1 > t0)
    hipLaunchKernelGGL(k_batch_transform, dim3((t1 - t0 + 255) / 256), dim3(256), 0, b->stream, L.dT,
                       L.dtgt_trial, t0, t1, b->dtrials, b->ddone, L.dZ);
  if (ni > 0) {
    if (grid <= 0) grid = std::max(1, L.grid_waves / L.waves);
    grid = std::min(grid, ni);
    switch (L.waves) {
      case 1: launch_pairs<1>(b, L, grid, order, ni, bp.sigma != 0); break;
      case 2: launch_pairs<2>(b, L, grid, order, ni, bp.sigma != 0); break;
      default: launch_pairs<4>(b, L, grid, order, ni, bp.sigma != 0); break;
    }
  }
  hipLaunchKernelGGL(k_batch_fold, dim3(k1 - k0), dim3(kUpd), 0, b->stream, L.dpart, L.dtitem, L.dttgt, k0, k1,
                     b->dtrials, b->ddone, b->dqueue, L.dnn, b->dfold, bp.sigma);
  hipLaunchKernelGGL(k_batch_update, dim3((k1 - k0 + 63) / 64), dim3(64), 0, b->stream, b->dfold, k0, k1,
                     b->dtrials, b->ddone, bp);
  DUGMA_HIP(hipGetLastError());
  return DUGMA_OK;
}

// The tail of a run: once at most a sixteenth of the trials are still
// running, passes are launched directly (not from the captured graph) over
// a queue that holds only those trials' items, in the same longest-first
// order, rebuilt whenever the running count has halved again.  A pass then
// no longer takes every finished trial's items off the queue (on the fine
// level, 4,096 trials, a pass with every trial finished still took 1.2 ms).
// Results do not change: an item writes only its own statistics slot, and
// a finished trial's items are skipped either way.
struct TailQueue {
  int last_active = -1;  // running trials when the queue was last rebuilt
  int n = 0;             // items in the level's dactive (0: the full queue)
  // hd: the done flags just read back
  int update(dugma_batch *b, const BLevel &L, const std::vector<int> &hd) {
    const int K = b->ntrials;
    int active = 0;
    for (int v : hd) active += v == 0;
    if (active == 0 || active * 16 > K) return DUGMA_OK;
    if (last_active >= 0 && active * 2 > last_active) return DUGMA_OK;
    std::vector<int> q;
    q.reserve(L.horder.size());
    for (int p : L.horder)
      if (!hd[L.hitem_trial[p]]) q.push_back(p);
    last_active = active;
    n = (int)q.size();
    if (n > 0) {
      DUGMA_HIP(hipMemcpyAsync(L.dactive, q.data(), q.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
      DUGMA_HIP(hipStreamSynchronize(b->stream));  // q goes out of scope
    }
    return DUGMA_OK;
  }
};

// One pass over every trial at its pose after the isotropic iterations that
// sums the energy (the iterations leave it out), on the full clouds with
// the trial's last variance scale; the done flags are cleared for it.
static int energy_pass(dugma_batch *b) {
  const BLevel &L = b->lv[0];
  const int K = b->ntrials;
  hipLaunchKernelGGL(k_batch_fill, dim3((K + 255) / 256), dim3(256), 0, b->stream, (unsigned *)b->ddone, K, 0u);
  if (L.ntgt > 0)
    hipLaunchKernelGGL(k_batch_transform, dim3((L.ntgt + 255) / 256), dim3(256), 0, b->stream, L.dT,
                       L.dtgt_trial, 0, L.ntgt, b->dtrials, b->ddone, L.dZ);
  if (L.nitems > 0) {
    const int grid = std::min(std::max(1, L.grid_waves / L.waves), L.nitems);
    switch (L.waves) {
      case 1: launch_pairs<1>(b, L, grid, L.dorder, L.nitems, false, true); break;
      case 2: launch_pairs<2>(b, L, grid, L.dorder, L.nitems, false, true); break;
      default: launch_pairs<4>(b, L, grid, L.dorder, L.nitems, false, true); break;
    }
  }
  hipLaunchKernelGGL(k_batch_energy, dim3((K + 255) / 256), dim3(256), 0, b->stream, L.dpart, L.dtitem, K,
                     b->dtrials, b->dqueue);
  DUGMA_HIP(hipGetLastError());
  return DUGMA_OK;
}

// Trial states at the start of a run, uploaded with every done flag clear
// and every nearest distance at +inf.
static int start_trials(dugma_batch *b, const double *R0, const double *t0, bool anneal, bool sigma) {
  const int K = b->ntrials;
  std::vector<BTrial> tr(K);
  for (int k = 0; k < K; ++k) {
    memset(&tr[k], 0, sizeof(BTrial));
    for (int a = 0; a < 9; ++a) tr[k].R[a] = R0[9 * (size_t)k + a];
    for (int a = 0; a < 3; ++a) {
      tr[k].t[a] = t0[3 * (size_t)k + a];
      tr[k].c[a] = b->cent[3 * (size_t)k + a];
    }
    tr[k].a = anneal ? b->a0[k] : 0.0;
    tr[k].a_stop = b->astop[k];
    tr[k].g = 1.0;
    tr[k].sbar = b->sbar[k] > 0 ? b->sbar[k] : 1.0;
    tr[k].sig_ready = sigma ? 0 : 1;
  }
  std::vector<int> done(K, 0);
  DUGMA_HIP(hipMemcpyAsync(b->dtrials, tr.data(), sizeof(BTrial) * K, hipMemcpyHostToDevice, b->stream));
  DUGMA_HIP(hipMemcpyAsync(b->ddone, done.data(), sizeof(int) * K, hipMemcpyHostToDevice, b->stream));
  for (int l = 0; l < kLevels; ++l)
    if (b->lv[l].ntgt > 0)
      hipLaunchKernelGGL(k_batch_fill, dim3((b->lv[l].ntgt + 255) / 256), dim3(256), 0, b->stream, b->lv[l].dnn,
                         b->lv[l].ntgt, kInfBits);
  DUGMA_HIP(hipGetLastError());
  DUGMA_HIP(hipStreamSynchronize(b->stream));  // the host vectors go out of scope
  return DUGMA_OK;
}

template <int W>
static void launch_pairsf(dugma_batch *b, int l, int grid, const int *order, int ni) {
  const BLevel &L = b->lv[l];
  hipLaunchKernelGGL((k_batchf_pairs<W>), dim3(grid), dim3(64 * W), 0, b->stream, b->dZF[l], b->dSF[l],
                     b->dSB[l], b->far_exp, L.ditems, order, ni, b->dqueue, b->ddone, L.dpart);
}

// One refinement pass of trials [k0, k1) on level l: transform at the
// candidate poses, pairs, optimiser step.
static int launch_refine(dugma_batch *b, int l, int k0, int k1, const RParams &rp, bool one_trial, int grid,
                         const int *order_in = nullptr, int ni_in = 0) {
  const BLevel &L = b->lv[l];
  int t0 = 0, t1 = L.ntgt;
  const int *order = order_in ? order_in : L.dorder;
  int ni = order_in ? ni_in : L.nitems;
  if (one_trial) {
    const int i0 = L.titem[k0];
    ni = L.titem[k0 + 1] - i0;
    order = L.dident + i0;
    t0 = L.tb[k0];
    t1 = t0 + L.nt[k0];
  }
  if (t1 > t0)
    hipLaunchKernelGGL(k_batchf_transform, dim3((t1 - t0 + 255) / 256), dim3(256), 0, b->stream, L.dT, b->dTC[l],
                       L.dtgt_trial, t0, t1, b->dtrials, b->dref, b->ddone, b->dZF[l]);
  if (ni > 0) {
    if (grid <= 0) grid = std::max(1, L.grid_waves / L.waves);
    grid = std::min(grid, ni);
    switch (L.waves) {
      case 1: launch_pairsf<1>(b, l, grid, order, ni); break;
      case 2: launch_pairsf<2>(b, l, grid, order, ni); break;
      default: launch_pairsf<4>(b, l, grid, order, ni); break;
    }
  }
  hipLaunchKernelGGL(k_batchf_fold, dim3(k1 - k0), dim3(kUpd), 0, b->stream, L.dpart, L.dtitem, k0, k1, b->ddone,
                     b->dqueue, b->dfold);
  hipLaunchKernelGGL(k_batchf_update, dim3((k1 - k0 + 255) / 256), dim3(256), 0, b->stream, b->dfold, k0, k1,
                     b->dref, b->ddone, rp);
  DUGMA_HIP(hipGetLastError());
  return DUGMA_OK;
}

// The refinement on level l after the isotropic iterations: every trial
// starts from the pose in dtrials, with its done flag cleared; at most
// max_steps accepted BFGS steps (each line search needs one pass when its
// first step is accepted).
// Optimiser states for a refinement on level l from the poses in dtrials,
// every done flag cleared.  carry (optional): the final states of the
// previous level, carry_level; a trial whose inverse Hessian was scaled
// there starts from it times kLevelStep[l] / kLevelStep[carry_level] (the
// covariances scale with the point step, and the curvature of -log E
// inversely), so that its first step is a quasi-Newton step rather than a
// kernel width.
static int init_refine(dugma_batch *b, int l, const std::vector<BRef> *carry = nullptr, int carry_level = 0) {
  const int K = b->ntrials;
  std::vector<BTrial> tr(K);
  DUGMA_HIP(hipMemcpyAsync(tr.data(), b->dtrials, sizeof(BTrial) * K, hipMemcpyDeviceToHost, b->stream));
  DUGMA_HIP(hipStreamSynchronize(b->stream));
  std::vector<BRef> ref(K);
  for (int k = 0; k < K; ++k) {
    BRef &r = ref[k];
    memset(&r, 0, sizeof(BRef));
    for (int a = 0; a < 9; ++a) r.cR[a] = r.R[a] = tr[k].R[a];
    for (int a = 0; a < 3; ++a) r.ct[a] = r.t[a] = tr[k].t[a];
    r.L = b->refL[k];
    // the first step is a kernel width: wider on the sparser levels
    r.w0 = b->refW0[k] * std::sqrt((double)kLevelStep[l]);
    r.smax = 0.25 * r.L;
    if (carry && (*carry)[k].state == 1 && (*carry)[k].scaled) {
      const double f = (double)kLevelStep[l] / kLevelStep[carry_level];
      for (int a = 0; a < 36; ++a) r.H[a] = (*carry)[k].H[a] * f;
      r.scaled = 1;
      r.warm = 1;
    }
  }
  DUGMA_HIP(hipMemcpyAsync(b->dref, ref.data(), sizeof(BRef) * K, hipMemcpyHostToDevice, b->stream));
  hipLaunchKernelGGL(k_batch_fill, dim3((K + 255) / 256), dim3(256), 0, b->stream, (unsigned *)b->ddone, K, 0u);
  DUGMA_HIP(hipGetLastError());
  DUGMA_HIP(hipStreamSynchronize(b->stream));  // the host vectors go out of scope
  return DUGMA_OK;
}

static int run_refine(dugma_batch *b, const dugma_batch_opts *opts, int l, int max_steps,
                      const std::vector<BRef> *carry = nullptr, int carry_level = 0) {
  const int K = b->ntrials;
  int rc = init_refine(b, l, carry, carry_level);
  if (rc) return rc;
  RParams rp;
  memset(&rp, 0, sizeof(rp));
  rp.tol_rot = opts->tol_rot;
  rp.tol_trans = opts->tol_trans;
  rp.max_iter = max_steps;
  {
    // DUGMA_REF_NEWTON=0 or 1 in the environment: plain BFGS, or the Newton
    // direction at each level's first pose only (for comparisons)
    const char *nv = getenv("DUGMA_REF_NEWTON");
    rp.newton = nv && (nv[0] == '0' || nv[0] == '1') ? nv[0] - '0' : 2;
  }
  // passes: each accepted step takes one pass at least; the halvings of a
  // line search take more, so allow 4 passes per step on average
  const int passes = 1 + 4 * max_steps;
  const int grid = opts->grid;
  if (opts->one_trial_per_launch) {
    int *hdone = nullptr;
    DUGMA_HIP(hipHostMalloc((void **)&hdone, sizeof(int), hipHostMallocDefault));
    for (int k = 0; k < K && !rc; ++k)
      for (int it = 0; it < passes && !rc; ++it) {
        rc = launch_refine(b, l, k, k + 1, rp, true, grid);
        if (rc) break;
        hipError_t e = hipMemcpyAsync(hdone, b->ddone + k, sizeof(int), hipMemcpyDeviceToHost, b->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
        if (e != hipSuccess) rc = hip_fail(e, "one-trial refinement");
        if (*hdone) break;
      }
    (void)hipHostFree(hdone);
    return rc;
  }
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  if (opts->use_graph) {
    hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
    rc = launch_refine(b, l, 0, K, rp, false, grid);
    e = hipStreamEndCapture(b->stream, &g);
    if (!rc && e != hipSuccess) rc = hip_fail(e, "hipStreamEndCapture");
    if (!rc) {
      e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      if (e != hipSuccess) rc = hip_fail(e, "hipGraphInstantiate");
    }
  }
  const int check = opts->check_every > 0 ? opts->check_every : 10;
  TailQueue tail;
  for (int it = 0; it < passes && !rc; ++it) {
    if (tail.n > 0) {
      rc = launch_refine(b, l, 0, K, rp, false, grid, b->lv[l].dactive, tail.n);
    } else if (ge) {
      hipError_t e = hipGraphLaunch(ge, b->stream);
      if (e != hipSuccess) rc = hip_fail(e, "hipGraphLaunch");
    } else {
      rc = launch_refine(b, l, 0, K, rp, false, grid);
    }
    if (rc) break;
    if ((it + 1) % check == 0) {
      std::vector<int> hd(K);
      hipError_t e = hipMemcpyAsync(hd.data(), b->ddone, sizeof(int) * K, hipMemcpyDeviceToHost, b->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
      if (e != hipSuccess) {
        rc = hip_fail(e, "refinement pass");
        break;
      }
      int nd = 0;
      for (int v : hd) nd += v;
      if (nd == K) break;
      rc = tail.update(b, b->lv[l], hd);
    }
  }
  if (ge) (void)hipGraphExecDestroy(ge);
  if (g) (void)hipGraphDestroy(g);
  return rc;
}

// After a refinement on a sparser level: each trial whose start pose was
// evaluated continues from the accepted pose, its steps counted with the
// isotropic ones; ref receives that level's final optimiser states.
static int fold_refined_poses(dugma_batch *b, std::vector<BRef> &ref) {
  const int K = b->ntrials;
  std::vector<BTrial> tr(K);
  ref.assign(K, BRef());
  DUGMA_HIP(hipMemcpyAsync(tr.data(), b->dtrials, sizeof(BTrial) * K, hipMemcpyDeviceToHost, b->stream));
  DUGMA_HIP(hipMemcpyAsync(ref.data(), b->dref, sizeof(BRef) * K, hipMemcpyDeviceToHost, b->stream));
  DUGMA_HIP(hipStreamSynchronize(b->stream));
  for (int k = 0; k < K; ++k) {
    if (ref[k].state != 1) continue;
    for (int a = 0; a < 9; ++a) tr[k].R[a] = ref[k].R[a];
    for (int a = 0; a < 3; ++a) tr[k].t[a] = ref[k].t[a];
    tr[k].iters += ref[k].it;
  }
  DUGMA_HIP(hipMemcpyAsync(b->dtrials, tr.data(), sizeof(BTrial) * K, hipMemcpyHostToDevice, b->stream));
  DUGMA_HIP(hipStreamSynchronize(b->stream));  // the host vectors go out of scope
  return DUGMA_OK;
}

extern "C" int dugma_batch_register(dugma_batch *b, const double *R0,
                                    const double *t0, const dugma_batch_opts *opts,
                                    double *R, double *t, int *iters,
                                    int *status, double *E) {
  if (!b || !R0 || !t0 || !opts) return fail(DUGMA_ERR_ARG, "dugma_batch_register: bad argument");
  if (opts->max_iter < 1) return fail(DUGMA_ERR_ARG, "dugma_batch_register: max_iter must be >= 1");
  if (opts->coarse_iters < 0) return fail(DUGMA_ERR_ARG, "dugma_batch_register: coarse_iters must be >= 0");
  if (opts->refine_iters < 0) return fail(DUGMA_ERR_ARG, "dugma_batch_register: refine_iters must be >= 0");
  if (opts->refine_coarse_iters < 0)
    return fail(DUGMA_ERR_ARG, "dugma_batch_register: refine_coarse_iters must be >= 0");
  if (opts->refine_mid_iters < 0)
    return fail(DUGMA_ERR_ARG, "dugma_batch_register: refine_mid_iters must be >= 0");
  if (opts->refine_iters > 0 && !b->dref)
    return fail(DUGMA_ERR_STATE, "dugma_batch_register: refine_iters needs dugma_batch_set_covariances first");
  const int K = b->ntrials;
  const bool sigma = opts->sigma_update != 0;
  int rc = start_trials(b, R0, t0, opts->anneal != 0, sigma);
  if (rc) return rc;
  BParams bp;
  memset(&bp, 0, sizeof(bp));
  bp.beta = opts->anneal_decay > 0 ? opts->anneal_decay : 0.85;
  bp.tol_rot = opts->tol_rot;
  bp.tol_trans = opts->tol_trans;
  bp.max_iter = opts->max_iter;
  bp.sigma = sigma ? 1 : 0;
  bp.relax = opts->relax > 1.0 ? opts->relax : 1.0;
  const int grid = opts->grid;
  // pass schedule: with the sigma update a first pass only measures sig; the
  // first coarse_iters pose steps run on the coarse level, and when the fine
  // level takes over, its first pass measures sig again
  const int coarse = std::min(opts->coarse_iters, opts->max_iter);
  // no full-cloud iterations after the coarse ones: every trial has reached
  // max_iter there, and the full clouds' measuring pass would find none running
  const int extra = sigma ? (coarse > 0 && coarse < opts->max_iter ? 2 : 1) : 0;
  const int passes = opts->max_iter + extra;
  const int fine_from = coarse > 0 ? coarse + (sigma ? 1 : 0) : 0;  // first fine pass
  auto level_of = [&](int pass) { return pass < fine_from ? 1 : 0; };
  // the coarse level never declares convergence
  BParams bpc = bp;
  bpc.tol_rot = -1.0;
  bpc.tol_trans = -1.0;
  auto params = [&](int l) -> const BParams & { return l == 1 ? bpc : bp; };
  auto switch_level = [&](int k0, int n) {
    if (sigma) hipLaunchKernelGGL(k_batch_resig, dim3((n + 255) / 256), dim3(256), 0, b->stream, b->dtrials + k0, n);
  };
  if (opts->one_trial_per_launch) {
    // baseline: each trial on its own, 3 launches and a host check per iteration
    int *hdone = nullptr;
    DUGMA_HIP(hipHostMalloc((void **)&hdone, sizeof(int), hipHostMallocDefault));
    for (int k = 0; k < K && !rc; ++k) {
      for (int it = 0; it < passes && !rc; ++it) {
        if (coarse > 0 && it == fine_from) switch_level(k, 1);
        rc = launch_iteration(b, level_of(it), k, k + 1, params(level_of(it)), true, grid);
        if (rc) break;
        hipError_t e = hipMemcpyAsync(hdone, b->ddone + k, sizeof(int), hipMemcpyDeviceToHost, b->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
        if (e != hipSuccess) rc = hip_fail(e, "one-trial iteration");
        if (*hdone) break;
      }
    }
    (void)hipHostFree(hdone);
  } else if (opts->use_graph) {
    // one captured iteration per level, replayed
    hipGraph_t g[2] = {nullptr, nullptr};
    hipGraphExec_t ge[2] = {nullptr, nullptr};
    for (int l = 0; l < 2 && !rc; ++l) {
      if (l == 1 && coarse == 0) break;
      hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) {
        rc = hip_fail(e, "hipStreamBeginCapture");
        break;
      }
      rc = launch_iteration(b, l, 0, K, params(l), false, grid);
      e = hipStreamEndCapture(b->stream, &g[l]);
      if (!rc && e != hipSuccess) rc = hip_fail(e, "hipStreamEndCapture");
      if (!rc) {
        e = hipGraphInstantiate(&ge[l], g[l], nullptr, nullptr, 0);
        if (e != hipSuccess) rc = hip_fail(e, "hipGraphInstantiate");
      }
    }
    const int check = opts->check_every > 0 ? opts->check_every : 10;
    TailQueue tail;  // fine level only (the coarse level never converges)
    for (int it = 0; it < passes && !rc; ++it) {
      if (coarse > 0 && it == fine_from) switch_level(0, K);
      const int l = level_of(it);
      if (l == 0 && tail.n > 0) {
        rc = launch_iteration(b, 0, 0, K, bp, false, grid, b->lv[0].dactive, tail.n);
        if (rc) break;
      } else {
        hipError_t e = hipGraphLaunch(ge[l], b->stream);
        if (e != hipSuccess) {
          rc = hip_fail(e, "hipGraphLaunch");
          break;
        }
      }
      if ((it + 1) % check == 0) {
        // all trials finished?  (done flags summed on the host)
        std::vector<int> hd(K);
        hipError_t e = hipMemcpyAsync(hd.data(), b->ddone, sizeof(int) * K, hipMemcpyDeviceToHost, b->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
        if (e != hipSuccess) {
          rc = hip_fail(e, "graph iteration");
          break;
        }
        int nd = 0;
        for (int v : hd) nd += v;
        if (nd == K) break;
        if (l == 0) rc = tail.update(b, b->lv[0], hd);
      }
    }
    for (int l = 0; l < 2; ++l) {
      if (ge[l]) (void)hipGraphExecDestroy(ge[l]);
      if (g[l]) (void)hipGraphDestroy(g[l]);
    }
  } else {
    for (int it = 0; it < passes && !rc; ++it) {
      if (coarse > 0 && it == fine_from) switch_level(0, K);
      rc = launch_iteration(b, level_of(it), 0, K, params(level_of(it)), false, grid);
    }
  }
  if (rc) return rc;
  // without a refinement, E is the isotropic energy at the returned pose
  if (opts->refine_iters == 0) rc = energy_pass(b);
  if (rc) return rc;
  // the refinement levels in order: coarse (every 4th point), middle (every
  // 2nd), full clouds; each starts from the previous one's pose and inverse
  // Hessian
  if (opts->refine_iters > 0) {
    const int steps[kLevels] = {opts->refine_iters, opts->refine_coarse_iters, opts->refine_mid_iters};
    const int seq[3] = {1, 2, 0};
    std::vector<BRef> prev;
    int prev_level = 0;
    for (int q = 0; q < 3 && !rc; ++q) {
      const int l = seq[q];
      if (steps[l] <= 0) continue;
      rc = run_refine(b, opts, l, steps[l], prev.empty() ? nullptr : &prev, prev_level);
      if (!rc && l != 0) {
        rc = fold_refined_poses(b, prev);
        prev_level = l;
      }
    }
  }
  if (rc) return rc;
  std::vector<BTrial> tr(K);
  DUGMA_HIP(hipMemcpyAsync(tr.data(), b->dtrials, sizeof(BTrial) * K, hipMemcpyDeviceToHost, b->stream));
  std::vector<BRef> ref(opts->refine_iters > 0 ? K : 0);
  if (!ref.empty())
    DUGMA_HIP(hipMemcpyAsync(ref.data(), b->dref, sizeof(BRef) * K, hipMemcpyDeviceToHost, b->stream));
  DUGMA_HIP(hipStreamSynchronize(b->stream));
  const double cn = 1.0 / (6.283185307179586476925 * std::sqrt(6.283185307179586476925));
  for (int k = 0; k < K; ++k) {
    const bool refined = !ref.empty() && ref[k].state == 1;
    const double *Rk = refined ? ref[k].R : tr[k].R, *tk = refined ? ref[k].t : tr[k].t;
    if (R)
      for (int a = 0; a < 9; ++a) R[9 * (size_t)k + a] = Rk[a];
    if (t)
      for (int a = 0; a < 3; ++a) t[3 * (size_t)k + a] = tk[a];
    if (iters) iters[k] = tr[k].iters + (refined ? ref[k].it : 0);
    if (status) status[k] = ref.empty() ? tr[k].status : ref[k].status;
    if (E) E[k] = refined ? cn * ref[k].E : cn * tr[k].E;
  }
  return DUGMA_OK;
}

// Full covariances of every trial's points for the refinement (layout of
// dugma_init: one column-major D x D block per point).  Only the points the
// batch keeps are read; each must have a finite, positive-definite FP32
// covariance.  Also sets each trial's length scales: L, the RMS radius of
// its sources about their centroid, and w0, the per-axis kernel width
// sqrt(2 x mean source variance).
// An upper bound of the largest eigenvalue of the symmetric 3x3 [s11 s12
// s13; . s22 s23; . . s33]: the closed form (trigonometric solution of the
// characteristic cubic, O. K. Smith, Comm. ACM 4(4):168, 1961) in FP64,
// raised by 1e-4 of the spectral radius bound so that its rounding errs
// upward.
static float lam_max_bound(float s11, float s12, float s13, float s22, float s23, float s33) {
  const double a11 = s11, a12 = s12, a13 = s13, a22 = s22, a23 = s23, a33 = s33;
  const double p1 = a12 * a12 + a13 * a13 + a23 * a23;
  const double q = (a11 + a22 + a33) / 3.0;
  double lam;
  if (p1 == 0.0) {
    lam = std::max(a11, std::max(a22, a33));
  } else {
    const double p2 = (a11 - q) * (a11 - q) + (a22 - q) * (a22 - q) + (a33 - q) * (a33 - q) + 2.0 * p1;
    const double pp = std::sqrt(p2 / 6.0);
    const double b11 = (a11 - q) / pp, b22 = (a22 - q) / pp, b33 = (a33 - q) / pp;
    const double b12 = a12 / pp, b13 = a13 / pp, b23 = a23 / pp;
    const double det = b11 * (b22 * b33 - b23 * b23) - b12 * (b12 * b33 - b23 * b13) + b13 * (b12 * b23 - b22 * b13);
    const double r = std::min(1.0, std::max(-1.0, 0.5 * det));
    lam = q + 2.0 * pp * std::cos(std::acos(r) / 3.0);
  }
  const double bound = std::fabs(a11) + std::fabs(a22) + std::fabs(a33) + 2.0 * std::sqrt(3.0 * p1);
  return (float)(lam + 1e-4 * bound) * (1.0f + 1e-6f);
}

extern "C" int dugma_batch_set_covariances(dugma_batch *b, const double *const *X_C, const double *const *Y_C) {
  if (!b || !X_C || !Y_C) return fail(DUGMA_ERR_ARG, "dugma_batch_set_covariances: null argument");
  const int K = b->ntrials;
  const BLevel &L = b->lv[0];
  std::vector<BCov> tc((size_t)L.ntgt);
  std::vector<Rec3F> sf((size_t)L.nsrc);
  std::vector<double> refL(K), refW0(K);
  auto sym32 = [](const double *C, float *o) -> bool {
    // o = [s11 s12 s13 s22 s23 s33] of the symmetric part, rounded to FP32
    const float v[6] = {(float)C[0], (float)(0.5 * (C[3] + C[1])), (float)(0.5 * (C[6] + C[2])), (float)C[4],
                        (float)(0.5 * (C[7] + C[5])), (float)C[8]};
    for (int a = 0; a < 6; ++a) {
      if (!std::isfinite(v[a])) return false;
      o[a] = v[a];
    }
    const double a11 = v[0], a12 = v[1], a13 = v[2], a22 = v[3], a23 = v[4], a33 = v[5];
    const double m2 = a11 * a22 - a12 * a12;
    const double m3 = a11 * (a22 * a33 - a23 * a23) - a12 * (a12 * a33 - a23 * a13) + a13 * (a12 * a23 - a22 * a13);
    return a11 > 0 && m2 > 0 && m3 > 0;
  };
  for (int k = 0; k < K; ++k) {
    if (!X_C[k] || !Y_C[k]) return fail(DUGMA_ERR_ARG, "dugma_batch_set_covariances: null covariance array");
    for (int q = 0; q < L.nt[k]; ++q) {
      const int i = b->keepT[L.tb[k] + q];
      float o[6];
      if (!sym32(X_C[k] + (size_t)i * 9, o))
        return fail(DUGMA_ERR_ARG, "dugma_batch_set_covariances: a target covariance is not positive definite");
      BCov &c = tc[L.tb[k] + q];
      c.s11 = o[0]; c.s12 = o[1]; c.s13 = o[2]; c.s22 = o[3]; c.s23 = o[4]; c.s33 = o[5];
      c.lam = c.pad = 0.0f;
    }
    double cen[3] = {0, 0, 0}, var = 0.0, r2 = 0.0;
    const int ns = L.ns[k];
    for (int q = 0; q < ns; ++q) {
      const int j = b->keepS[L.sb[k] + q];
      float o[6];
      if (!sym32(Y_C[k] + (size_t)j * 9, o))
        return fail(DUGMA_ERR_ARG, "dugma_batch_set_covariances: a source covariance is not positive definite");
      const BSrc &x = b->hsrc0[L.sb[k] + q];
      Rec3F &r = sf[L.sb[k] + q];
      memset(&r, 0, sizeof(r));
      r.hx = x.x; r.hy = x.y; r.hz = x.z;
      r.s11 = o[0]; r.s12 = o[1]; r.s13 = o[2]; r.s22 = o[3]; r.s23 = o[4]; r.s33 = o[5];
      cen[0] += x.x; cen[1] += x.y; cen[2] += x.z;
      var += ((double)o[0] + o[3] + o[5]) / 3.0;
    }
    if (ns > 0) {
      for (int a = 0; a < 3; ++a) cen[a] /= ns;
      for (int q = 0; q < ns; ++q) {
        const BSrc &x = b->hsrc0[L.sb[k] + q];
        r2 += (x.x - cen[0]) * (x.x - cen[0]) + (x.y - cen[1]) * (x.y - cen[1]) + (x.z - cen[2]) * (x.z - cen[2]);
      }
      var /= ns;
      r2 /= ns;
    }
    refL[k] = r2 > 0 ? std::sqrt(r2) : 1.0;
    refW0[k] = var > 0 ? std::sqrt(2.0 * var) : 1e-3 * refL[k];
  }
  // the sparser levels' points are every kLevelStep[l]-th of each trial's
  // (in level-0 order); their covariances times kLevelStep[l] (exact: a
  // power of two)
  std::vector<BCov> tcl[kLevels];
  std::vector<Rec3F> sfl[kLevels];
  for (int l = 1; l < kLevels; ++l) {
    const BLevel &Ll = b->lv[l];
    const int step = kLevelStep[l];
    const float cs = (float)step;
    tcl[l].resize((size_t)Ll.ntgt);
    sfl[l].resize((size_t)Ll.nsrc);
    for (int k = 0; k < K; ++k) {
      for (int q = 0; q < Ll.nt[k]; ++q) {
        BCov c = tc[L.tb[k] + (size_t)q * step];
        c.s11 *= cs; c.s12 *= cs; c.s13 *= cs; c.s22 *= cs; c.s23 *= cs; c.s33 *= cs;
        tcl[l][Ll.tb[k] + q] = c;
      }
      for (int q = 0; q < Ll.ns[k]; ++q) {
        Rec3F r = sf[L.sb[k] + (size_t)q * step];
        r.s11 *= cs; r.s12 *= cs; r.s13 *= cs; r.s22 *= cs; r.s23 *= cs; r.s33 *= cs;
        sfl[l][Ll.sb[k] + q] = r;
      }
    }
  }
  tcl[0].swap(tc);
  sfl[0].swap(sf);
  // eigenvalue bounds per target, and per 64-source block the box and the
  // largest bound
  std::vector<BBound> sbl[kLevels];
  for (int l = 0; l < kLevels; ++l) {
    const BLevel &Ll = b->lv[l];
    for (BCov &c : tcl[l]) c.lam = lam_max_bound(c.s11, c.s12, c.s13, c.s22, c.s23, c.s33);
    sbl[l].resize((size_t)std::max(Ll.nblocks, 1));
    for (int k = 0; k < K; ++k)
      for (int q0 = 0, m = Ll.bb[k]; q0 < Ll.ns[k]; q0 += kBSub, ++m) {
        BBound bd;
        memset(&bd, 0, sizeof(bd));
        for (int a = 0; a < 3; ++a) {
          bd.lo[a] = __builtin_huge_valf();
          bd.hi[a] = -__builtin_huge_valf();
        }
        for (int q = q0; q < std::min(q0 + kBSub, Ll.ns[k]); ++q) {
          const Rec3F &r = sfl[l][Ll.sb[k] + q];
          const float x[3] = {r.hx, r.hy, r.hz};
          for (int a = 0; a < 3; ++a) {
            bd.lo[a] = std::min(bd.lo[a], x[a]);
            bd.hi[a] = std::max(bd.hi[a], x[a]);
          }
          bd.lam = std::max(bd.lam, lam_max_bound(r.s11, r.s12, r.s13, r.s22, r.s23, r.s33));
        }
        sbl[l][m] = bd;
      }
  }
  {
    // allocated once (sizes are fixed by the batch); a buffer is allocated
    // only where a previous call left none
    int rc = b->dref ? DUGMA_OK : bmalloc(&b->dref, (size_t)K);
    for (int l = 0; l < kLevels && !rc; ++l) {
      if (!b->dTC[l] && (rc = bmalloc(&b->dTC[l], tcl[l].size()))) break;
      if (!b->dSF[l] && (rc = bmalloc(&b->dSF[l], sfl[l].size()))) break;
      if (!b->dZF[l] && (rc = bmalloc(&b->dZF[l], tcl[l].size()))) break;
      if (!b->dSB[l] && (rc = bmalloc(&b->dSB[l], sbl[l].size()))) break;
    }
    if (rc) return rc;
  }
  for (int l = 0; l < kLevels; ++l) {
    if (!tcl[l].empty())
      DUGMA_HIP(hipMemcpy(b->dTC[l], tcl[l].data(), tcl[l].size() * sizeof(BCov), hipMemcpyHostToDevice));
    if (!sfl[l].empty())
      DUGMA_HIP(hipMemcpy(b->dSF[l], sfl[l].data(), sfl[l].size() * sizeof(Rec3F), hipMemcpyHostToDevice));
    DUGMA_HIP(hipMemcpy(b->dSB[l], sbl[l].data(), sbl[l].size() * sizeof(BBound), hipMemcpyHostToDevice));
  }
  b->refL = refL;
  b->refW0 = refW0;
  const char *fs = getenv("DUGMA_FAR_SKIP");
  b->far_exp = (fs && fs[0] == '0') ? __builtin_huge_valf() : kFarExp;
  return DUGMA_OK;
}

// Average time of one all-trial isotropic iteration (ms) on level l under
// the current launch shape, from the given start poses: one untimed pass,
// then `reps` timed.
static int time_iteration(dugma_batch *b, int l, const double *R0, const double *t0, bool sigma, int reps,
                          hipEvent_t ev0, hipEvent_t ev1, double *ms) {
  int rc = start_trials(b, R0, t0, false, sigma);
  if (rc) return rc;
  BParams bp;
  memset(&bp, 0, sizeof(bp));
  bp.beta = 1.0;
  bp.tol_rot = -1.0;  // nothing converges while timing
  bp.tol_trans = -1.0;
  bp.max_iter = 1 << 30;
  bp.relax = 1.0;
  bp.sigma = sigma ? 1 : 0;
  if ((rc = launch_iteration(b, l, 0, b->ntrials, bp, false, 0))) return rc;
  DUGMA_HIP(hipEventRecord(ev0, b->stream));
  for (int r = 0; r < reps; ++r)
    if ((rc = launch_iteration(b, l, 0, b->ntrials, bp, false, 0))) return rc;
  DUGMA_HIP(hipEventRecord(ev1, b->stream));
  DUGMA_HIP(hipEventSynchronize(ev1));
  float f = 0.0f;
  DUGMA_HIP(hipEventElapsedTime(&f, ev0, ev1));
  *ms = (double)f / reps;
  return DUGMA_OK;
}

// Average time of one all-trial refinement pass (ms) on level l under the
// current launch shape, from the given start poses: one untimed pass (the
// start poses), then `reps` timed (candidate poses; nothing converges).
static int time_refine_pass(dugma_batch *b, int l, const double *R0, const double *t0, int reps,
                            hipEvent_t ev0, hipEvent_t ev1, double *ms) {
  int rc = start_trials(b, R0, t0, false, false);
  if (!rc) rc = init_refine(b, l);
  if (rc) return rc;
  RParams rp;
  memset(&rp, 0, sizeof(rp));
  rp.tol_rot = -1.0;
  rp.tol_trans = -1.0;
  rp.max_iter = 1 << 30;
  if ((rc = launch_refine(b, l, 0, b->ntrials, rp, false, 0))) return rc;
  DUGMA_HIP(hipEventRecord(ev0, b->stream));
  for (int r = 0; r < reps; ++r)
    if ((rc = launch_refine(b, l, 0, b->ntrials, rp, false, 0))) return rc;
  DUGMA_HIP(hipEventRecord(ev1, b->stream));
  DUGMA_HIP(hipEventSynchronize(ev1));
  float f = 0.0f;
  DUGMA_HIP(hipEventElapsedTime(&f, ev0, ev1));
  *ms = (double)f / reps;
  return DUGMA_OK;
}

extern "C" int dugma_batch_autotune(dugma_batch *b, const double *R0, const double *t0,
                                    int sigma_update, int *waves, int *chunks, int *grid_waves,
                                    double *ms, double *refine_ms) {
  if (!b || !R0 || !t0) return fail(DUGMA_ERR_ARG, "dugma_batch_autotune: bad argument");
  const bool sigma = sigma_update != 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  DUGMA_HIP(hipEventCreate(&ev0));
  hipError_t e = hipEventCreate(&ev1);
  if (e != hipSuccess) {
    (void)hipEventDestroy(ev0);
    return hip_fail(e, "hipEventCreate");
  }
  int rc = DUGMA_OK;
  double best = 0.0, cur = 0.0;
  // the figure of merit on level l: one isotropic iteration, plus (once
  // covariances are set) one refinement pass
  double cur_it = 0.0, cur_ref = 0.0, best_it = 0.0, best_ref = 0.0;
  int lvl = 0;
  auto timed = [&](double *out) {
    cur_it = cur_ref = 0.0;
    const int reps = lvl == 0 ? 2 : 6;  // a coarse pass is ~1/10 of a fine one
    int r = DUGMA_OK;
    if (lvl != 2) r = time_iteration(b, lvl, R0, t0, sigma, reps, ev0, ev1, &cur_it);  // no iterations on level 2
    if (!r && b->dTC[lvl]) r = time_refine_pass(b, lvl, R0, t0, reps, ev0, ev1, &cur_ref);
    *out = cur_it + cur_ref;
    return r;
  };
  auto keep_best = [&]() {
    best = cur;
    best_it = cur_it;
    best_ref = cur_ref;
  };
  double fine_it = 0.0, fine_ref = 0.0;
  for (lvl = 0; lvl < kLevels && !rc; ++lvl) {
    BLevel &L = b->lv[lvl];
    if (L.nitems == 0) continue;
    if (lvl == 2 && !b->dTC[2]) {  // refinement only: keep the full clouds' shape
      L.waves = b->lv[0].waves;
      for (int c = 0; c < kClasses; ++c) L.chunks[c] = b->lv[0].chunks[c];
      L.grid_waves = b->lv[0].grid_waves;
      continue;
    }
    best = 0.0;
    bool present[kClasses] = {false, false, false, false};
    for (int k = 0; k < b->ntrials; ++k)
      if (L.nt[k] > 0 && L.ns[k] > 0) present[size_class(L.ns[k])] = true;
    // 1. waves per workgroup (the target tile).  On the full clouds the
    // largest tile is kept unless a smaller one is faster by more than 3%:
    // with all trials active (as timed here) the shapes are close, and once
    // trials converge the larger tile's fewer work items cost less to skip
    // (measured on 4,096 trials: the refinement's tail 13% faster with 4
    // waves than with 2).  The coarse level's iterations never converge,
    // and its small trials fill a large tile poorly: plain fastest.
    const double margin = lvl == 0 ? 0.97 : 1.0;
    const int wave_opts[] = {4, 2, 1};
    int best_w = L.waves;
    for (int w : wave_opts) {
      L.waves = w;
      if ((rc = make_items(b, L)) || (rc = timed(&cur))) break;
      if (best == 0.0 || cur < margin * best) {
        keep_best();
        best_w = w;
      }
    }
    L.waves = best_w;
    // 2. source chunks per target tile, one size class at a time
    const int chunk_opts[] = {1, 2, 4, 8};
    for (int c = 0; c < kClasses && !rc; ++c) {
      if (!present[c]) continue;
      const int keep = L.chunks[c];
      int best_n = keep;
      for (int n : chunk_opts) {
        if (n == keep) continue;
        L.chunks[c] = n;
        if ((rc = make_items(b, L)) || (rc = timed(&cur))) break;
        if (cur < best) {
          keep_best();
          best_n = n;
        }
      }
      L.chunks[c] = best_n;
    }
    if (!rc) rc = make_items(b, L);
    // 3. resident waves of the persistent kernels
    const int grid_opts[] = {4096, 8192, 16384, 32768};
    const int keep_g = L.grid_waves;
    int best_g = keep_g;
    for (int gw : grid_opts) {
      if (rc || gw == keep_g) continue;
      L.grid_waves = gw;
      if ((rc = timed(&cur))) break;
      if (cur < best) {
        keep_best();
        best_g = gw;
      }
    }
    L.grid_waves = best_g;
    if (lvl == 0) {
      fine_it = best_it;
      fine_ref = best_ref;
    }
  }
  {
    // both levels follow the kept shapes, also after a failed timing run
    const int rc2 = make_all_items(b);
    if (!rc) rc = rc2;
  }
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  if (rc) return rc;
  if (waves) *waves = b->lv[0].waves;
  if (chunks)
    for (int c = 0; c < kClasses; ++c) chunks[c] = b->lv[0].chunks[c];
  if (grid_waves) *grid_waves = b->lv[0].grid_waves;
  if (ms) *ms = fine_it;
  if (refine_ms) *refine_ms = fine_ref;
  return DUGMA_OK;
}

extern "C" int dugma_batch_set_shape(dugma_batch *b, int waves, const int *chunks, int grid_waves) {
  if (!b || (waves != 1 && waves != 2 && waves != 4) || !chunks || grid_waves < 1)
    return fail(DUGMA_ERR_ARG, "dugma_batch_set_shape: waves must be 1, 2 or 4, chunks >= 1, grid_waves >= 1");
  for (int c = 0; c < kClasses; ++c)
    if (chunks[c] < 1) return fail(DUGMA_ERR_ARG, "dugma_batch_set_shape: chunks must be >= 1");
  for (BLevel &L : b->lv) {
    L.waves = waves;
    for (int c = 0; c < kClasses; ++c) L.chunks[c] = chunks[c];
    L.grid_waves = grid_waves;
  }
  return make_all_items(b);
}

extern "C" int dugma_batch_set_level_shape(dugma_batch *b, int level, int waves, const int *chunks,
                                           int grid_waves) {
  if (!b || level < 0 || level >= kLevels || (waves != 1 && waves != 2 && waves != 4) || !chunks ||
      grid_waves < 1)
    return fail(DUGMA_ERR_ARG, "dugma_batch_set_level_shape: level 0, 1 or 2, waves 1, 2 or 4, chunks >= 1, "
                               "grid_waves >= 1");
  for (int c = 0; c < kClasses; ++c)
    if (chunks[c] < 1) return fail(DUGMA_ERR_ARG, "dugma_batch_set_level_shape: chunks must be >= 1");
  BLevel &L = b->lv[level];
  L.waves = waves;
  for (int c = 0; c < kClasses; ++c) L.chunks[c] = chunks[c];
  L.grid_waves = grid_waves;
  return make_items(b, L);
}

extern "C" int dugma_batch_get_shape(const dugma_batch *b, int level, int *waves, int *chunks, int *grid_waves) {
  if (!b || level < 0 || level >= kLevels)
    return fail(DUGMA_ERR_ARG, "dugma_batch_get_shape: null batch or level not 0, 1 or 2");
  const BLevel &L = b->lv[level];
  if (waves) *waves = L.waves;
  if (chunks)
    for (int c = 0; c < kClasses; ++c) chunks[c] = L.chunks[c];
  if (grid_waves) *grid_waves = L.grid_waves;
  return DUGMA_OK;
}
