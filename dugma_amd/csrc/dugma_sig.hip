// dugma_sig.hip -- update_sig (update_sig.cpp / update_sig.cu) and exact
// k-nearest neighbours with lowest-index tie-break.
//
// Both are integer/ordering results, so they are bit-exact: this file is
// compiled without FP contraction, every FP32 operation is written in a fixed
// order (explicit fmaf where a fused operation is meant), and the minimum of
// the quantised distances is order-independent.  Queries sit one per lane;
// candidates are streamed with wave-uniform addresses (scalar loads).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include <hipcub/hipcub.hpp>

#include "dugma_internal.h"
#include "dugma_pairs.h"

namespace dugma {

constexpr int kSBlock = 256;

struct PoseArgC {
  double R[9];  // column-major
  double t[3];
};

// update_sig.cu:247-252: m_j = fl32(r_a1*y1 + r_a2*y2 (+ r_a3*y3) + t_a), FP64.
// Threads below nfill also set fill[i] = fv (the start value of the
// quantised minimum), saving a launch.
template <int D>
__global__ __launch_bounds__(kSBlock) void k_transform_compat(
    const float *__restrict__ Y, int M, PoseArgC ps, float *__restrict__ out,
    unsigned *__restrict__ fill, int nfill, unsigned fv) {
  const int j = blockIdx.x * kSBlock + threadIdx.x;
  if (j < nfill) fill[j] = fv;
  if (j >= M) return;
  const float *y = Y + (size_t)j * D;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    double s = ps.R[a] * (double)y[0];
#pragma unroll
    for (int b = 1; b < D; ++b) s = s + ps.R[a + D * b] * (double)y[b];
    s = s + ps.t[a];
    out[(size_t)j * D + a] = (float)s;
  }
}

// __float2uint_rn: round half to even, saturating; NaN and negatives -> 0.
__device__ __forceinline__ unsigned f2u_rn(float s) {
  if (!(s > 0.0f)) return 0u;
  if (s >= 4294967296.0f) return 4294967295u;
  return (unsigned)rintf(s);
}

// update_sig.cu:238-265: per target the minimum over a chunk of sources of
// rn_u32(fl32(fl32(|n - m|^2) / D) * 1e6), merged with atomicMin (the
// minimum does not depend on the order).  IEEE FP32 semantics in source
// order; the reference's binaries use -use_fast_math (compile.m:2), whose
// division may differ in the last bit.
template <int D>
__global__ __launch_bounds__(kSBlock) void k_nn1_quant(
    const float *__restrict__ X, int N, const float *__restrict__ Mp, int M,
    int chunk, unsigned *__restrict__ qmin) {
  const int i = blockIdx.x * kSBlock + threadIdx.x;
  const int j0 = blockIdx.y * chunk;
  const int j1 = min(j0 + chunk, M);
  const int ii = min(i, N - 1);
  float n[3];
#pragma unroll
  for (int a = 0; a < D; ++a) n[a] = X[(size_t)ii * D + a];
  constexpr float kThird = 1.0f / 3.0f;
  unsigned best = 4000000000u;
  // candidate j + 1 is requested before candidate j is used (j0 < j1 for
  // every workgroup; index clamped to the chunk)
  float mn[3] = {Mp[(size_t)j0 * D], Mp[(size_t)j0 * D + 1], D == 3 ? Mp[(size_t)j0 * D + (D - 1)] : 0.0f};
  for (int j = j0; j < j1; ++j) {
    const float m[3] = {mn[0], mn[1], mn[2]};
    const size_t jn = (size_t)min(j + 1, j1 - 1) * D;
    mn[0] = Mp[jn];
    mn[1] = Mp[jn + 1];
    if (D == 3) mn[2] = Mp[jn + (D - 1)];
    const float e0 = n[0] - m[0], e1 = n[1] - m[1];
    float s2 = e0 * e0 + e1 * e1;
    if (D == 3) {
      const float e2 = n[2] - m[2];
      s2 = s2 + e2 * e2;
    }
    // fl32(s2 / D): for D = 3 the FMA-corrected product below is the IEEE
    // quotient for every non-negative finite float (checked exhaustively on
    // the CPU by tests/arith/div3_exhaustive.c); for D = 2 halving is exact
    float s;
    if (D == 3) {
      const float q = s2 * kThird;
      s = fmaf(fmaf(-3.0f, q, s2), kThird, q);
    } else {
      s = s2 * 0.5f;
    }
    s = s * 1000000.0f;
    best = min(best, f2u_rn(s));
  }
  if (i < N) atomicMin(qmin + i, best);
}

// update_sig.cu:79-91 then 291-307 with n_c = 1: a[i] = fl32(q[i]) / 1e6
// (IEEE division), then the pairwise tree, level by level; at each level
// element 2k*stride absorbs element (2k+1)*stride, an odd last element is
// carried.  The same tree, 6 levels at a time: each wave sums 64 aligned
// elements with lane shuffles (lane i adds lane i + s for s = 1, 2, ..., 32;
// lane 0 ends with the group's tree sum), the group sums form the next
// level's array, and so on.  The values are >= 0, so padding with +0 is
// exact and stands in for the carry of an odd last element.  One workgroup;
// a holds at least n floats (ping-pong of the group sums); writes sum / N
// (update_sig.cu:385) to out (the context's pinned result slot).
__global__ __launch_bounds__(1024) void k_tree_sum_div(const unsigned *__restrict__ q,
                                                       float *__restrict__ a, int n,
                                                       float *__restrict__ out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  const int g0 = (n + 63) / 64;
  float *cur = a, *next = a + g0;
  for (int g = wave; g < g0; g += nwaves) {
    const int i = g * 64 + lane;
    float v = i < n ? ((float)q[i]) / 1000000.0f : 0.0f;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) v = v + __shfl_down(v, s, 64);
    if (lane == 0) cur[g] = v;
  }
  __syncthreads();
  for (int m = g0; m > 1;) {
    const int groups = (m + 63) / 64;
    for (int g = wave; g < groups; g += nwaves) {
      const int i = g * 64 + lane;
      float v = i < m ? cur[i] : 0.0f;
#pragma unroll
      for (int s = 1; s < 64; s <<= 1) v = v + __shfl_down(v, s, 64);
      if (lane == 0) next[g] = v;
    }
    __syncthreads();
    float *t = cur;
    cur = next;
    next = t;
    m = groups;
  }
  if (threadIdx.x == 0) out[0] = cur[0] / (float)n;
}

// ---------------------------------------------------------------------------
// k-NN: running sorted top-K per lane; candidates of a chunk arrive in
// increasing index, so a candidate enters only when strictly closer, which
// ranks equal distances by lowest index.
// ---------------------------------------------------------------------------
template <int K>
__device__ __forceinline__ void topk_insert(float *bd, int *bi, float dd, int j) {
  if (!(dd < bd[K - 1])) return;
#pragma unroll
  for (int s = K - 1; s >= 0; --s) {
    if (bd[s] > dd) {
      if (s > 0 && bd[s - 1] > dd) {
        bd[s] = bd[s - 1];
        bi[s] = bi[s - 1];
      } else {
        bd[s] = dd;
        bi[s] = j;
      }
    }
  }
}

template <int D, int K>
__global__ __launch_bounds__(kSBlock) void k_knn_chunk(
    const float *__restrict__ Qp, int nq, const float *__restrict__ Pp, int np,
    int chunk, int self_exclude, float *__restrict__ pd, int *__restrict__ pi) {
  const int i = blockIdx.x * kSBlock + threadIdx.x;
  const int j0 = blockIdx.y * chunk;
  const int j1 = min(j0 + chunk, np);
  const int ii = min(i, nq - 1);
  float q[3];
#pragma unroll
  for (int a = 0; a < D; ++a) q[a] = Qp[(size_t)ii * D + a];
  float bd[K];
  int bi[K];
#pragma unroll
  for (int s = 0; s < K; ++s) {
    bd[s] = INFINITY;
    bi[s] = -1;
  }
  for (int j = j0; j < j1; ++j) {
    const float *p = Pp + (size_t)j * D;
    const float dx = q[0] - p[0], dy = q[1] - p[1];
    float dd;
    if (D == 3) {
      const float dz = q[2] - p[2];
      dd = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    } else {
      dd = fmaf(dy, dy, dx * dx);
    }
    if (self_exclude && j == i) dd = INFINITY;
    topk_insert<K>(bd, bi, dd, j);
  }
  if (i < nq) {
    const size_t base = ((size_t)blockIdx.y * nq + i) * K;
#pragma unroll
    for (int s = 0; s < K; ++s) {
      pd[base + s] = bd[s];
      pi[base + s] = bi[s];
    }
  }
}

// Merge the per-chunk lists in chunk order (chunk c holds larger indices
// than chunk c-1, so strict-less insertion keeps the lowest-index rule).
template <int K>
__global__ __launch_bounds__(kSBlock) void k_knn_merge(
    const float *__restrict__ pd, const int *__restrict__ pi, int nq,
    int nchunks, int k, float *__restrict__ od, int *__restrict__ oi) {
  const int i = blockIdx.x * kSBlock + threadIdx.x;
  if (i >= nq) return;
  float bd[K];
  int bi[K];
#pragma unroll
  for (int s = 0; s < K; ++s) {
    bd[s] = INFINITY;
    bi[s] = -1;
  }
  for (int c = 0; c < nchunks; ++c) {
    const size_t base = ((size_t)c * nq + i) * K;
    for (int s = 0; s < K; ++s) {
      const int j = pi[base + s];
      if (j < 0) break;
      topk_insert<K>(bd, bi, pd[base + s], j);
    }
  }
#pragma unroll
  for (int s = 0; s < K; ++s)
    if (s < k) {
      od[(size_t)i * k + s] = bd[s];
      oi[(size_t)i * k + s] = bi[s];
    }
}

// Candidate chunks so that about 8,192 waves run, each chunk at least
// min_chunk candidates long.
// ---------------------------------------------------------------------------
// k-NN over a uniform grid (exact, same results as the all-pairs kernels):
// the finite candidates are binned into cubic cells of side h (counting
// sort: counts, exclusive scan, scatter); each query visits the cells of
// Chebyshev rings r = 0, 1, 2, ... around its own cell.  A candidate not yet
// visited after ring r lies in a cell at ring distance >= r + 1, so at least
// r h away (also for a query outside the grid: its cell index is clamped
// towards the candidates).  The search stops once the list holds K entries
// and the squared distance to the unvisited cells (at least (r h)^2, more for
// a query outside the grid), times 1 - 1e-4, exceeds the K-th distance (the
// margin covers the FP32 rounding of the distances and the cell
// arithmetic); after
// kRingFallback rings it rescans all candidates from an empty list, which
// bounds the cost of sparse regions.  Cells are visited out of index order,
// so the list is kept in (distance, index) order, which is the all-pairs
// kernels' lowest-index rule.  Distances use the same FP32 expression.
// ---------------------------------------------------------------------------
constexpr int kBoxGroups = 64;     // workgroups of the bounding-box pass
constexpr int kRingFallback = 8;   // rings before the full rescan
constexpr int kGridMinPairs = 1 << 24;  // below this many pairs the all-pairs kernels run

struct GridParams {
  double ox, oy, oz;  // origin (minimum corner of the finite candidates)
  double h, inv_h;    // cell side
  int nx, ny, nz, nvalid;
};

// Per-workgroup minimum / maximum of the finite candidates and their count:
// part[g] = {min x, y, z, count (bits), max x, y, z, unused}.
template <int D>
__global__ __launch_bounds__(kSBlock) void k_grid_bbox(const float *__restrict__ P, int np,
                                                       float *__restrict__ part) {
  __shared__ float smin[D][kSBlock], smax[D][kSBlock];
  __shared__ int scnt[kSBlock];
  const int len = (np + (int)gridDim.x - 1) / (int)gridDim.x;
  const int j0 = blockIdx.x * len, j1 = min(j0 + len, np);
  float lo[D], hi[D];
  int cnt = 0;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    lo[a] = __builtin_huge_valf();
    hi[a] = -__builtin_huge_valf();
  }
  for (int j = j0 + (int)threadIdx.x; j < j1; j += kSBlock) {
    bool fin = true;
    float x[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
      x[a] = P[(size_t)j * D + a];
      fin = fin && isfinite(x[a]);
    }
    if (!fin) continue;
    ++cnt;
#pragma unroll
    for (int a = 0; a < D; ++a) {
      lo[a] = fminf(lo[a], x[a]);
      hi[a] = fmaxf(hi[a], x[a]);
    }
  }
#pragma unroll
  for (int a = 0; a < D; ++a) {
    smin[a][threadIdx.x] = lo[a];
    smax[a][threadIdx.x] = hi[a];
  }
  scnt[threadIdx.x] = cnt;
  __syncthreads();
  for (int w = kSBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int a = 0; a < D; ++a) {
        smin[a][threadIdx.x] = fminf(smin[a][threadIdx.x], smin[a][threadIdx.x + w]);
        smax[a][threadIdx.x] = fmaxf(smax[a][threadIdx.x], smax[a][threadIdx.x + w]);
      }
      scnt[threadIdx.x] += scnt[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    float *o = part + (size_t)blockIdx.x * 8;
    for (int a = 0; a < 3; ++a) {
      o[a] = a < D ? smin[a][0] : 0.0f;
      o[4 + a] = a < D ? smax[a][0] : 0.0f;
    }
    o[3] = __int_as_float(scnt[0]);
    o[7] = 0.0f;
  }
}

// Grid parameters from the partial boxes (one thread): about 2 cells per
// finite candidate over the box, extents below 1/64 of the largest counted
// as 1/64 of it (flat and thin clouds), at most 2048 cells per axis and
// `cap` cells in all (the cell side grows until both hold).
template <int D>
__global__ void k_grid_params(const float *__restrict__ part, int groups, long long cap,
                              GridParams *__restrict__ gp) {
  // one wave: each lane reads its share of the group partials (the loads
  // overlap), then the wave reduces them; lane 0 sizes the grid
  const int lane = threadIdx.x;
  float flo[3] = {__builtin_huge_valf(), __builtin_huge_valf(), __builtin_huge_valf()};
  float fhi[3] = {-__builtin_huge_valf(), -__builtin_huge_valf(), -__builtin_huge_valf()};
  int n = 0;
  for (int g = lane; g < groups; g += 64) {
    const float *o = part + (size_t)g * 8;
    const int c = __float_as_int(o[3]);
    if (c == 0) continue;
    n += c;
    for (int a = 0; a < D; ++a) {
      flo[a] = fminf(flo[a], o[a]);
      fhi[a] = fmaxf(fhi[a], o[4 + a]);
    }
  }
  for (int m = 32; m >= 1; m >>= 1) {
    n += __shfl_xor(n, m, 64);
    for (int a = 0; a < D; ++a) {
      flo[a] = fminf(flo[a], __shfl_xor(flo[a], m, 64));
      fhi[a] = fmaxf(fhi[a], __shfl_xor(fhi[a], m, 64));
    }
  }
  if (lane != 0 || blockIdx.x != 0) return;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  if (n > 0)
    for (int a = 0; a < D; ++a) {
      lo[a] = (double)flo[a];
      hi[a] = (double)fhi[a];
    }
  GridParams p;
  p.ox = lo[0];
  p.oy = lo[1];
  p.oz = lo[2];
  p.nvalid = n;
  double e[3] = {0, 0, 0}, emax = 0.0;
  for (int a = 0; a < D; ++a) {
    e[a] = hi[a] - lo[a];
    emax = fmax(emax, e[a]);
  }
  double h = 1.0;
  if (n > 0 && emax > 0.0) {
    double vol = 1.0;
    for (int a = 0; a < D; ++a) vol *= fmax(e[a], emax / 64.0);
    const double cells = 2.0 * n;
    h = D == 3 ? cbrt(vol / cells) : sqrt(vol / cells);
  }
  // h grows until both limits hold, so that every candidate lies inside its
  // own cell (no candidate cell index is clamped)
  int nn[3] = {1, 1, 1};
  for (int it = 0; it < 400; ++it) {
    double tot = 1.0, mx = 1.0;
    for (int a = 0; a < D; ++a) {
      const double na = floor(e[a] / h) + 1.0;
      tot *= na;
      mx = fmax(mx, na);
    }
    if (mx <= 2048.0 && tot <= (double)cap) break;
    h *= 1.26;
  }
  for (int a = 0; a < D; ++a) nn[a] = (int)(floor(e[a] / h) + 1.0);
  p.h = h;
  p.inv_h = 1.0 / h;
  p.nx = nn[0];
  p.ny = nn[1];
  p.nz = D == 3 ? nn[2] : 1;
  *gp = p;
}

__device__ __forceinline__ int grid_coord(float x, double o, double inv_h, int n) {
  const double c = floor(((double)x - o) * inv_h);
  return (int)fmin(fmax(c, 0.0), (double)(n - 1));
}

template <int D>
__device__ __forceinline__ int grid_cell(const float *x, const GridParams &g) {
  const int cx = grid_coord(x[0], g.ox, g.inv_h, g.nx);
  const int cy = grid_coord(x[1], g.oy, g.inv_h, g.ny);
  const int cz = D == 3 ? grid_coord(x[2], g.oz, g.inv_h, g.nz) : 0;
  return (cz * g.ny + cy) * g.nx + cx;
}

template <int D>
__global__ __launch_bounds__(kSBlock) void k_grid_count(const float *__restrict__ P, int np,
                                                        const GridParams *__restrict__ gp,
                                                        int *__restrict__ cnt) {
  const int j = blockIdx.x * kSBlock + threadIdx.x;
  if (j >= np) return;
  float x[D];
  bool fin = true;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    x[a] = P[(size_t)j * D + a];
    fin = fin && isfinite(x[a]);
  }
  if (!fin) return;
  atomicAdd(cnt + grid_cell<D>(x, *gp), 1);
}

// Candidates into cell order (order inside a cell is arbitrary: the lists
// are kept in (distance, index) order).  cnt counts down to 0.
template <int D>
__global__ __launch_bounds__(kSBlock) void k_grid_scatter(const float *__restrict__ P, int np,
                                                          const GridParams *__restrict__ gp,
                                                          const int *__restrict__ start,
                                                          int *__restrict__ cnt, float *__restrict__ spts,
                                                          int *__restrict__ sidx) {
  const int j = blockIdx.x * kSBlock + threadIdx.x;
  if (j >= np) return;
  float x[D];
  bool fin = true;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    x[a] = P[(size_t)j * D + a];
    fin = fin && isfinite(x[a]);
  }
  if (!fin) return;
  const int cell = grid_cell<D>(x, *gp);
  const int pos = start[cell] + atomicSub(cnt + cell, 1) - 1;
#pragma unroll
  for (int a = 0; a < D; ++a) spts[(size_t)pos * D + a] = x[a];
  sidx[pos] = j;
}

// Queries into cell order too, so that the lanes of a wave search
// neighbouring cells (shared ring extents, nearby candidate loads).  The key
// is the query's cell, or the number of cells for a query with a
// non-finite coordinate (answered without a search).  Each query's result
// does not depend on the order.
template <int D>
__device__ __forceinline__ int query_key(const float *__restrict__ Q, int i, const GridParams &g) {
  float x[D];
  bool fin = true;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    x[a] = Q[(size_t)i * D + a];
    fin = fin && isfinite(x[a]);
  }
  return fin ? grid_cell<D>(x, g) : g.nx * g.ny * g.nz;
}

template <int D>
__global__ __launch_bounds__(kSBlock) void k_query_count(const float *__restrict__ Q, int nq,
                                                         const GridParams *__restrict__ gp,
                                                         int *__restrict__ cnt) {
  const int i = blockIdx.x * kSBlock + threadIdx.x;
  if (i >= nq) return;
  atomicAdd(cnt + query_key<D>(Q, i, *gp), 1);
}

template <int D>
__global__ __launch_bounds__(kSBlock) void k_query_scatter(const float *__restrict__ Q, int nq,
                                                           const GridParams *__restrict__ gp,
                                                           const int *__restrict__ start,
                                                           int *__restrict__ cnt, int *__restrict__ qord) {
  const int i = blockIdx.x * kSBlock + threadIdx.x;
  if (i >= nq) return;
  const int key = query_key<D>(Q, i, *gp);
  qord[start[key] + atomicSub(cnt + key, 1) - 1] = i;
}

template <int K>
__device__ __forceinline__ void topk_insert_lex(float *bd, int *bi, float dd, int j) {
  if (!(dd < bd[K - 1] || (dd == bd[K - 1] && j < bi[K - 1]))) return;
  bool after[K];  // entry s comes after (dd, j)
#pragma unroll
  for (int s = 0; s < K; ++s) after[s] = bd[s] > dd || (bd[s] == dd && bi[s] > j);
#pragma unroll
  for (int s = K - 1; s >= 0; --s) {
    if (after[s]) {
      if (s > 0 && after[s - 1]) {
        bd[s] = bd[s - 1];
        bi[s] = bi[s - 1];
      } else {
        bd[s] = dd;
        bi[s] = j;
      }
    }
  }
}

template <int D, int K>
__device__ __forceinline__ void grid_visit(const float *q, int i, int self_exclude, int p0, int p1,
                                           const float *__restrict__ spts, const int *__restrict__ sidx,
                                           float *bd, int *bi) {
  for (int pos = p0; pos < p1; ++pos) {
    const int j = sidx[pos];
    const float *p = spts + (size_t)pos * D;
    const float dx = q[0] - p[0], dy = q[1] - p[1];
    float dd;
    if (D == 3) {
      const float dz = q[2] - p[2];
      dd = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    } else {
      dd = fmaf(dy, dy, dx * dx);
    }
    if (self_exclude && j == i) continue;
    topk_insert_lex<K>(bd, bi, dd, j);
  }
}

// Lower bound of the squared distance from q to every cell outside the
// block of rings 0..r around its cell (cx, cy, cz): per axis and side with
// cells left beyond the block, the squared distance to that face of the
// block plus the squared gaps of q to the grid's extent along the other axes
// (zero for a query inside the grid).  +inf when no cell is left.
template <int D>
__device__ __forceinline__ double unvisited_dist2(const float *q, int cx, int cy, int cz, int r,
                                                  const GridParams &g) {
  const double o[3] = {g.ox, g.oy, g.oz};
  const int c[3] = {cx, cy, cz}, n[3] = {g.nx, g.ny, g.nz};
  double gap2[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int a = 0; a < D; ++a) {
    const double x = q[a], lo = o[a], hi = o[a] + n[a] * g.h;
    const double gp = x < lo ? lo - x : (x > hi ? x - hi : 0.0);
    gap2[a] = gp * gp;
  }
  double lb = __builtin_huge_val();
#pragma unroll
  for (int a = 0; a < D; ++a) {
    double others = 0.0;
#pragma unroll
    for (int b = 0; b < D; ++b)
      if (b != a) others += gap2[b];
    const double x = q[a];
    if (c[a] - r > 0) {  // cells below the block
      const double d = fmax(x - (o[a] + (c[a] - r) * g.h), 0.0);
      lb = fmin(lb, d * d + others);
    }
    if (c[a] + r < n[a] - 1) {  // cells above the block
      const double d = fmax(o[a] + (c[a] + r + 1) * g.h - x, 0.0);
      lb = fmin(lb, d * d + others);
    }
  }
  return lb;
}

template <int D, int K>
__global__ __launch_bounds__(kSBlock) void k_knn_grid(
    const float *__restrict__ Q, int nq, const int *__restrict__ qord, const float *__restrict__ spts,
    const int *__restrict__ sidx, const int *__restrict__ start, const GridParams *__restrict__ gp,
    int self_exclude, int k, float *__restrict__ od, int *__restrict__ oi) {
  const int lane = blockIdx.x * kSBlock + threadIdx.x;
  if (lane >= nq) return;
  const int i = qord[lane];
  const GridParams g = *gp;
  float q[D];
  bool fin = true;
#pragma unroll
  for (int a = 0; a < D; ++a) {
    q[a] = Q[(size_t)i * D + a];
    fin = fin && isfinite(q[a]);
  }
  float bd[K];
  int bi[K];
#pragma unroll
  for (int s = 0; s < K; ++s) {
    bd[s] = __builtin_huge_valf();
    bi[s] = -1;
  }
  if (fin && g.nvalid > 0) {
    const int cx = grid_coord(q[0], g.ox, g.inv_h, g.nx);
    const int cy = grid_coord(q[1], g.oy, g.inv_h, g.ny);
    const int cz = D == 3 ? grid_coord(q[2], g.oz, g.inv_h, g.nz) : 0;
    // ring beyond which no cell is left
    int rall = max(max(cx, g.nx - 1 - cx), max(cy, g.ny - 1 - cy));
    if (D == 3) rall = max(rall, max(cz, g.nz - 1 - cz));
    for (int r = 0;; ++r) {
      if (r > kRingFallback) {
#pragma unroll
        for (int s = 0; s < K; ++s) {
          bd[s] = __builtin_huge_valf();
          bi[s] = -1;
        }
        grid_visit<D, K>(q, i, self_exclude, 0, g.nvalid, spts, sidx, bd, bi);
        break;
      }
      const int zr = D == 3 ? r : 0;
      for (int a = -r; a <= r; ++a) {
        const int x = cx + a;
        if (x < 0 || x >= g.nx) continue;
        for (int b = -r; b <= r; ++b) {
          const int y = cy + b;
          if (y < 0 || y >= g.ny) continue;
          const bool edge = a == -r || a == r || b == -r || b == r;
          if (D == 2) {
            if (!edge) continue;
            const int cell = y * g.nx + x;
            grid_visit<D, K>(q, i, self_exclude, start[cell], start[cell + 1], spts, sidx, bd, bi);
            continue;
          }
          // 3-D: the whole z column on the rim of the (a, b) square, else its two ends
          const int step = (edge || zr == 0) ? 1 : 2 * zr;
          for (int c = -zr; c <= zr; c += step) {
            const int z = cz + c;
            if (z < 0 || z >= g.nz) continue;
            const int cell = (z * g.ny + y) * g.nx + x;
            grid_visit<D, K>(q, i, self_exclude, start[cell], start[cell + 1], spts, sidx, bd, bi);
          }
        }
      }
      if (r >= rall) break;
      if (bi[K - 1] >= 0 && unvisited_dist2<D>(q, cx, cy, cz, r, g) * (1.0 - 1e-4) > (double)bd[K - 1])
        break;
    }
  }
#pragma unroll
  for (int s = 0; s < K; ++s)
    if (s < k) {
      od[(size_t)i * k + s] = bd[s];
      oi[(size_t)i * k + s] = bi[s];
    }
}

static int chunks_for(int nq, int np, int min_chunk = 256) {
  const int qblocks = (nq + kSBlock - 1) / kSBlock;
  int c = (8192 + qblocks * 4 - 1) / (qblocks * 4);
  c = std::max(1, std::min(c, (np + min_chunk - 1) / min_chunk));
  return c;
}

static int transform_compat(dugma_ctx *c, const double *R, const double *t,
                            float *out, unsigned *fill = nullptr, int nfill = 0,
                            unsigned fv = 0u) {
  PoseArgC ps;
  memset(&ps, 0, sizeof(ps));
  for (int a = 0; a < c->D * c->D; ++a) ps.R[a] = R[a];
  for (int a = 0; a < c->D; ++a) ps.t[a] = t[a];
  dim3 grid((std::max(c->M, nfill) + kSBlock - 1) / kSBlock);
  if (c->D == 3)
    hipLaunchKernelGGL(k_transform_compat<3>, grid, dim3(kSBlock), 0, c->stream, c->dY, c->M, ps, out, fill,
                       nfill, fv);
  else
    hipLaunchKernelGGL(k_transform_compat<2>, grid, dim3(kSBlock), 0, c->stream, c->dY, c->M, ps, out, fill,
                       nfill, fv);
  DUGMA_HIP(hipGetLastError());
  return DUGMA_OK;
}

static int check_dims(dugma_ctx *c, double N, double M, double D) {
  if ((int)N != c->N || (int)M != c->M || (int)D != c->D)
    return fail(DUGMA_ERR_STATE, "N, M, D do not match the context");
  return DUGMA_OK;
}

}  // namespace dugma

using namespace dugma;

extern "C" int dugma_update_sig(dugma_ctx *ctx, const double *R, const double *t,
                                double N, double M, double D, double *sig) {
  if (!ctx || !R || !t || !sig) return fail(DUGMA_ERR_ARG, "dugma_update_sig: null argument");
  int rc = check_dims(ctx, N, M, D);
  if (rc) return rc;
  dugma_ctx *c = ctx;
  // scratch: m (M*D floats) | q (N u32) | f (N floats)
  const size_t off_q = ((size_t)c->M * c->D * 4 + 255) & ~(size_t)255;
  const size_t off_f = off_q + (((size_t)c->N * 4 + 255) & ~(size_t)255);
  rc = c->scratch2.ensure(off_f + (((size_t)c->N * 4 + 255) & ~(size_t)255));
  if (rc) return rc;
  char *base = (char *)c->scratch2.ptr;
  float *m = (float *)base;
  unsigned *q = (unsigned *)(base + off_q);
  float *f = (float *)(base + off_f);
  // transform, with q = 4e9 (update_sig.cu:13) set in the same launch
  rc = transform_compat(c, R, t, m, q, c->N, 4000000000u);
  if (rc) return rc;
  // chunks merge through one atomicMin per target each: short chunks are cheap
  const int nch = chunks_for(c->N, c->M, 64);
  const int chunk = (c->M + nch - 1) / nch;
  const int nchunks = (c->M + chunk - 1) / chunk;
  dim3 grid((c->N + kSBlock - 1) / kSBlock, nchunks);
  if (c->D == 3)
    hipLaunchKernelGGL(k_nn1_quant<3>, grid, dim3(kSBlock), 0, c->stream, c->dX, c->N, m, c->M, chunk, q);
  else
    hipLaunchKernelGGL(k_nn1_quant<2>, grid, dim3(kSBlock), 0, c->stream, c->dX, c->N, m, c->M, chunk, q);
  float *slot = (float *)(c->hres + dugma_ctx::kResSig);
  preset_slots(slot, 1);
  hipLaunchKernelGGL(k_tree_sum_div, dim3(1), dim3(1024), 0, c->stream, q, f, c->N,
                     (float *)(c->dres + dugma_ctx::kResSig));
  DUGMA_HIP(hipGetLastError());
  rc = wait_slots(c, slot, 1);
  if (rc) return rc;
  const volatile float *hs = (const volatile float *)(c->hres + dugma_ctx::kResSig);
  *sig = (double)hs[0];
  return DUGMA_OK;
}

// The lists to the caller's arrays (pageable memory): one device-to-host
// copy of each into the context's pinned staging buffer at DMA speed, then
// host copies of 1 MiB pieces on a few threads.
static int download_lists(dugma_ctx *c, const float *od, const int *oi, size_t n, float *d2, int *idx) {
  if (!d2 && !idx) return DUGMA_OK;
  const size_t bytes = n * 4;
  int rc = c->hstage.ensure(2 * bytes);
  if (rc) return rc;
  char *h = (char *)c->hstage.ptr;
  if (d2) DUGMA_HIP(hipMemcpyAsync(h, od, bytes, hipMemcpyDeviceToHost, c->stream));
  if (idx) DUGMA_HIP(hipMemcpyAsync(h + bytes, oi, bytes, hipMemcpyDeviceToHost, c->stream));
  DUGMA_HIP(hipStreamSynchronize(c->stream));
  constexpr size_t kPiece = size_t(1) << 20;
  const size_t np = (bytes + kPiece - 1) / kPiece;
  parallel_for(2 * np, [&](size_t q) {
    const bool isd = q < np;
    char *dst = isd ? (char *)d2 : (char *)idx;
    if (!dst) return;
    const size_t off = (isd ? q : q - np) * kPiece;
    memcpy(dst + off, h + (isd ? 0 : bytes) + off, std::min(kPiece, bytes - off));
  });
  return DUGMA_OK;
}

// The grid k-NN: candidates Pp (transformed from the source cloud first when
// R is given), queries Qp.
static int knn_grid(dugma_ctx *c, int D, int K, int k, const double *R, const double *t, const float *Qp,
                    int nq, const float *Pp, int np, int self, const float **lists_d, const int **lists_i) {
  const long long cap = std::min<long long>(8LL * np + 64, 1LL << 24);
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t temp_bytes = 0;
  DUGMA_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, temp_bytes, (const int *)nullptr, (int *)nullptr,
                                             (int)(cap + 1), c->stream));
  const size_t need_m = R ? al((size_t)np * D * 4) : 0;
  const size_t b_part = al((size_t)kBoxGroups * 8 * 4), b_gp = al(sizeof(GridParams));
  const size_t b_cnt = al((size_t)(cap + 1) * 4), b_pts = al((size_t)np * D * 4), b_idx = al((size_t)np * 4);
  const size_t b_tmp = al(temp_bytes), b_out = al((size_t)nq * k * 4), b_ord = al((size_t)nq * 4);
  int rc = c->scratch2.ensure(need_m + b_part + b_gp + 3 * b_cnt + b_pts + b_idx + b_tmp + 2 * b_out + b_ord);
  if (rc) return rc;
  char *p = (char *)c->scratch2.ptr;
  if (R) {
    rc = transform_compat(c, R, t, (float *)p);
    if (rc) return rc;
    Pp = (const float *)p;
  }
  p += need_m;
  float *part = (float *)p;
  p += b_part;
  GridParams *gp = (GridParams *)p;
  p += b_gp;
  int *cnt = (int *)p;
  p += b_cnt;
  int *start = (int *)p;
  p += b_cnt;
  int *qstart = (int *)p;
  p += b_cnt;
  int *qord = (int *)p;
  p += b_ord;
  float *spts = (float *)p;
  p += b_pts;
  int *sidx = (int *)p;
  p += b_idx;
  void *tmp = p;
  p += b_tmp;
  float *od = (float *)p;
  p += b_out;
  int *oi = (int *)p;
  const dim3 blk(kSBlock), pgrid((np + kSBlock - 1) / kSBlock), qgrid((nq + kSBlock - 1) / kSBlock);
  if (D == 3) {
    hipLaunchKernelGGL(k_grid_bbox<3>, dim3(kBoxGroups), blk, 0, c->stream, Pp, np, part);
    hipLaunchKernelGGL(k_grid_params<3>, dim3(1), dim3(64), 0, c->stream, part, kBoxGroups, cap, gp);
  } else {
    hipLaunchKernelGGL(k_grid_bbox<2>, dim3(kBoxGroups), blk, 0, c->stream, Pp, np, part);
    hipLaunchKernelGGL(k_grid_params<2>, dim3(1), dim3(64), 0, c->stream, part, kBoxGroups, cap, gp);
  }
  DUGMA_HIP(hipMemsetAsync(cnt, 0, (size_t)(cap + 1) * 4, c->stream));
  if (D == 3)
    hipLaunchKernelGGL(k_grid_count<3>, pgrid, blk, 0, c->stream, Pp, np, gp, cnt);
  else
    hipLaunchKernelGGL(k_grid_count<2>, pgrid, blk, 0, c->stream, Pp, np, gp, cnt);
  DUGMA_HIP(hipGetLastError());
  DUGMA_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, temp_bytes, cnt, start, (int)(cap + 1), c->stream));
  if (D == 3)
    hipLaunchKernelGGL(k_grid_scatter<3>, pgrid, blk, 0, c->stream, Pp, np, gp, start, cnt, spts, sidx);
  else
    hipLaunchKernelGGL(k_grid_scatter<2>, pgrid, blk, 0, c->stream, Pp, np, gp, start, cnt, spts, sidx);
  // the candidate scatter has counted cnt back down to zero: reuse it
  if (D == 3)
    hipLaunchKernelGGL(k_query_count<3>, qgrid, blk, 0, c->stream, Qp, nq, gp, cnt);
  else
    hipLaunchKernelGGL(k_query_count<2>, qgrid, blk, 0, c->stream, Qp, nq, gp, cnt);
  DUGMA_HIP(hipGetLastError());
  DUGMA_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, temp_bytes, cnt, qstart, (int)(cap + 1), c->stream));
  if (D == 3)
    hipLaunchKernelGGL(k_query_scatter<3>, qgrid, blk, 0, c->stream, Qp, nq, gp, qstart, cnt, qord);
  else
    hipLaunchKernelGGL(k_query_scatter<2>, qgrid, blk, 0, c->stream, Qp, nq, gp, qstart, cnt, qord);
  DUGMA_HIP(hipGetLastError());
#define KNN_GRID(DD, KK)                                                                               \
  hipLaunchKernelGGL((k_knn_grid<DD, KK>), qgrid, blk, 0, c->stream, Qp, nq, qord, spts, sidx, start, gp, \
                     self, k, od, oi)
  if (D == 3) {
    if (K == 8) KNN_GRID(3, 8); else if (K == 16) KNN_GRID(3, 16); else KNN_GRID(3, 32);
  } else {
    if (K == 8) KNN_GRID(2, 8); else if (K == 16) KNN_GRID(2, 16); else KNN_GRID(2, 32);
  }
#undef KNN_GRID
  DUGMA_HIP(hipGetLastError());
  *lists_d = od;
  *lists_i = oi;
  return DUGMA_OK;
}

// k-NN lists on the device: nq x k distances and indices in the context's
// scratch (valid until the next call that uses it).  Candidates Pp, or the
// transformed source cloud when R is given; self: skip candidate j == i.
static int knn_device(dugma_ctx *c, int k, const double *R, const double *t, const float *Qp, int nq,
                      const float *Pp, int np, int self, const float **lists_d, const int **lists_i) {
  const int D = c->D;
  const int K = k <= 8 ? 8 : (k <= 16 ? 16 : 32);
  if ((long long)nq * np >= kGridMinPairs)
    return knn_grid(c, D, K, k, R, t, Qp, nq, Pp, np, self, lists_d, lists_i);
  // scratch layout: [transformed candidates] | partial d | partial idx | out d | out idx
  const size_t need_m = R ? ((size_t)np * D * 4 + 255) & ~(size_t)255 : 0;
  const int nch = chunks_for(nq, np);
  const int chunk = (np + nch - 1) / nch;
  const int nchunks = (np + chunk - 1) / chunk;
  const size_t part = (size_t)nchunks * nq * K * 4;
  const size_t part_al = (part + 255) & ~(size_t)255;
  const size_t outsz = (((size_t)nq * k * 4) + 255) & ~(size_t)255;
  int rc = c->scratch2.ensure(need_m + 2 * part_al + 2 * outsz);
  if (rc) return rc;
  char *base = (char *)c->scratch2.ptr;
  if (need_m) {
    rc = transform_compat(c, R, t, (float *)base);
    if (rc) return rc;
    Pp = (const float *)base;
  }
  const size_t off = need_m;
  float *pd = (float *)(base + off);
  int *pi = (int *)(base + off + part_al);
  float *od = (float *)(base + off + 2 * part_al);
  int *oi = (int *)(base + off + 2 * part_al + outsz);
  dim3 grid((nq + kSBlock - 1) / kSBlock, nchunks), block(kSBlock);
#define KNN_LAUNCH(DD, KK)                                                              \
  hipLaunchKernelGGL((k_knn_chunk<DD, KK>), grid, block, 0, c->stream, Qp, nq, Pp, np, \
                     chunk, self, pd, pi)
  if (D == 3) {
    if (K == 8) KNN_LAUNCH(3, 8); else if (K == 16) KNN_LAUNCH(3, 16); else KNN_LAUNCH(3, 32);
  } else {
    if (K == 8) KNN_LAUNCH(2, 8); else if (K == 16) KNN_LAUNCH(2, 16); else KNN_LAUNCH(2, 32);
  }
#undef KNN_LAUNCH
  dim3 mgrid((nq + kSBlock - 1) / kSBlock);
  if (K == 8)
    hipLaunchKernelGGL(k_knn_merge<8>, mgrid, block, 0, c->stream, pd, pi, nq, nchunks, k, od, oi);
  else if (K == 16)
    hipLaunchKernelGGL(k_knn_merge<16>, mgrid, block, 0, c->stream, pd, pi, nq, nchunks, k, od, oi);
  else
    hipLaunchKernelGGL(k_knn_merge<32>, mgrid, block, 0, c->stream, pd, pi, nq, nchunks, k, od, oi);
  DUGMA_HIP(hipGetLastError());
  *lists_d = od;
  *lists_i = oi;
  return DUGMA_OK;
}

extern "C" int dugma_knn(dugma_ctx *ctx, int which, int k, const double *R,
                         const double *t, int *idx, float *d2) {
  if (!ctx) return fail(DUGMA_ERR_ARG, "dugma_knn: null context");
  if (k < 1 || k > 32) return fail(DUGMA_ERR_ARG, "dugma_knn: k must be in 1..32");
  dugma_ctx *c = ctx;
  const float *Qp, *Pp = nullptr;
  int nq, np, self = 0;
  if (which == DUGMA_KNN_TARGET_SELF) {
    Qp = c->dX; nq = c->N; Pp = c->dX; np = c->N; self = 1;
    R = t = nullptr;
  } else if (which == DUGMA_KNN_SOURCE_SELF) {
    Qp = c->dY; nq = c->M; Pp = c->dY; np = c->M; self = 1;
    R = t = nullptr;
  } else if (which == DUGMA_KNN_TARGET_TO_SOURCE) {
    if (!R || !t) return fail(DUGMA_ERR_ARG, "dugma_knn: R and t are required");
    Qp = c->dX; nq = c->N; np = c->M;
  } else {
    return fail(DUGMA_ERR_ARG, "dugma_knn: unknown query set");
  }
  const float *od;
  const int *oi;
  int rc = knn_device(c, k, R, t, Qp, nq, Pp, np, self, &od, &oi);
  if (rc) return rc;
  return download_lists(c, od, oi, (size_t)nq * k, d2, idx);
}

// ---------------------------------------------------------------------------
// Per-point sigma from the k nearest neighbours within a cloud (the k-NN
// extension of update_sig.cu:181-269): the cloud is the points the energy
// keeps, gathered from init's FP32 points; the lists are those of the k-NN
// search above (self excluded, (distance, index) order).
// ---------------------------------------------------------------------------
namespace dugma {

template <int D>
__global__ __launch_bounds__(kSBlock) void k_gather_points(const float *__restrict__ Pall,
                                                           const int *__restrict__ idx, int n,
                                                           float *__restrict__ P) {
  const int q = blockIdx.x * kSBlock + threadIdx.x;
  if (q >= n) return;
  const size_t i = (size_t)idx[q];
#pragma unroll
  for (int a = 0; a < D; ++a) P[(size_t)q * D + a] = Pall[i * D + a];
}

// sigma_q = (sqrtf(d2_1) + sqrtf(d2_2) + ... + sqrtf(d2_k)) / k, nearest
// first, in FP32; var = fl32(sigma^2).  A point whose var is zero or not
// finite keeps its covariance and is counted.  Otherwise its covariance
// becomes var I everywhere the context holds it, exactly as dugma_init would
// have stored that FP32 covariance: init's inverse and g (reference
// handles, original index i), and the energy state (kept index q), scaled
// by sc2 = scale^2 in FP64 and rounded once; isotropic records hold ln 4
// times the variance.
template <int D, bool TARGET>
__global__ __launch_bounds__(kSBlock) void k_sigma_apply(
    const float *__restrict__ lists, int n, int k, const int *__restrict__ idx, double sc2,
    float *__restrict__ Cinv, float *__restrict__ g, void *__restrict__ recF, void *__restrict__ recI,
    double *__restrict__ src, float *__restrict__ sig_out, int *__restrict__ unchanged) {
  const int q = blockIdx.x * kSBlock + threadIdx.x;
  if (q >= n) return;
  const float *d2 = lists + (size_t)q * k;
  float s = sqrtf(d2[0]);
  for (int r = 1; r < k; ++r) s = s + sqrtf(d2[r]);
  const float sigma = s / (float)k;
  sig_out[q] = sigma;
  const float var = sigma * sigma;
  if (!(var > 0.0f) || !(var < INFINITY)) {
    atomicAdd(unchanged, 1);
    return;
  }
  const size_t i = (size_t)idx[q];
  float a[D * D];
#pragma unroll
  for (int e = 0; e < D * D; ++e) a[e] = (e % (D + 1) == 0) ? var : 0.0f;
  compat_inverse<D>(a, Cinv + i * D * D, g + i);
  const float sv = (float)((double)var * sc2);
  const float si = (float)((double)sv * kLn4);
  if (TARGET) {
    if (D == 3) {
      Rec3F &r = reinterpret_cast<Rec3F *>(recF)[q];
      r.s11 = r.s22 = r.s33 = sv;
      r.s12 = r.s13 = r.s23 = 0.0f;
      if (recI) {
        Rec3I ri;
        ri.hx = r.hx; ri.hy = r.hy; ri.hz = r.hz; ri.s = si;
        ri.lx = ri.ly = ri.lz = ri.pad0 = 0.0f;
        reinterpret_cast<Rec3I *>(recI)[q] = ri;
      }
    } else {
      Rec2F &r = reinterpret_cast<Rec2F *>(recF)[q];
      r.s11 = r.s22 = sv;
      r.s12 = 0.0f;
      if (recI) {
        Rec2I ri;
        ri.hx = r.hx; ri.hy = r.hy; ri.lx = ri.ly = 0.0f; ri.s = si;
        ri.pad0 = ri.pad1 = ri.pad2 = 0.0f;
        reinterpret_cast<Rec2I *>(recI)[q] = ri;
      }
    }
  } else {
    double *o = src + (size_t)q * (D + D * D) + D;
#pragma unroll
    for (int e = 0; e < D * D; ++e) o[e] = (e % (D + 1) == 0) ? (double)var * sc2 : 0.0;
  }
}

// Surface covariance from the same lists (dugma_update_cov_knn): sigma as
// above, then, nearest first in FP32, mu = (p_1 + ... + p_k) / k,
// c_ab = ((p_1a - mu_a)(p_1b - mu_b) + ...) / k and
// C_ab = scale c_ab + (a == b ? floor (sigma sigma) : 0).  A point with
// fewer than k other points, or whose C fails the energy's validity rule
// (positive leading principal minors, in FP64 of the FP32 values), keeps
// its covariance and is counted.  Otherwise C is written everywhere the
// context holds the point's covariance, as dugma_init would have stored it:
// init's inverse and g, and the full-covariance energy state (scaled by
// sc2 in FP64, rounded once).
template <int D, bool TARGET>
__global__ __launch_bounds__(kSBlock) void k_cov_apply(
    const float *__restrict__ lists, const int *__restrict__ li, const float *__restrict__ P, int n, int k,
    float scale, float floor, const int *__restrict__ idx, double sc2, float *__restrict__ Cinv,
    float *__restrict__ g, void *__restrict__ recF, double *__restrict__ src, float *__restrict__ sig_out,
    int *__restrict__ unchanged) {
  const int q = blockIdx.x * kSBlock + threadIdx.x;
  if (q >= n) return;
  const float *d2 = lists + (size_t)q * k;
  const int *nb = li + (size_t)q * k;
  float s = sqrtf(d2[0]);
  for (int r = 1; r < k; ++r) s = s + sqrtf(d2[r]);
  const float sigma = s / (float)k;
  sig_out[q] = sigma;
  if (nb[k - 1] < 0) {
    atomicAdd(unchanged, 1);
    return;
  }
  float mu[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int a = 0; a < D; ++a) {
    float m = P[(size_t)nb[0] * D + a];
    for (int r = 1; r < k; ++r) m = m + P[(size_t)nb[r] * D + a];
    mu[a] = m / (float)k;
  }
  const float fl = floor * (sigma * sigma);
  float C[D * D];
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int b = 0; b < D; ++b) {
      float c = 0.0f;
      for (int r = 0; r < k; ++r) {
        const float u = P[(size_t)nb[r] * D + a] - mu[a], v = P[(size_t)nb[r] * D + b] - mu[b];
        c = c + u * v;
      }
      float cab = scale * (c / (float)k);
      if (a == b) cab = cab + fl;
      C[a + D * b] = cab;
    }
  // the energy's validity rule (FP64 of the FP32 values)
  bool ok = true;
#pragma unroll
  for (int e = 0; e < D * D; ++e) ok = ok && isfinite(C[e]);
  const double S0 = C[0], m2 = (double)C[0] * C[D + 1] - (double)C[1] * C[D];
  ok = ok && S0 > 0 && m2 > 0;
  if (D == 3) {
    const double det = (double)C[0] * ((double)C[4] * C[8] - (double)C[7] * C[5]) -
                       (double)C[3] * ((double)C[1] * C[8] - (double)C[7] * C[2]) +
                       (double)C[6] * ((double)C[1] * C[5] - (double)C[4] * C[2]);
    ok = ok && det > 0;
  }
  if (!ok) {
    atomicAdd(unchanged, 1);
    return;
  }
  const size_t i = (size_t)idx[q];
  compat_inverse<D>(C, Cinv + i * D * D, g + i);
  auto sv = [&](int a, int b) { return (float)((double)C[a + D * b] * sc2); };
  if (TARGET) {
    if (D == 3) {
      Rec3F &r = reinterpret_cast<Rec3F *>(recF)[q];
      r.s11 = sv(0, 0); r.s22 = sv(1, 1); r.s33 = sv(2, 2);
      r.s12 = sv(0, 1); r.s13 = sv(0, 2); r.s23 = sv(1, 2);
    } else {
      Rec2F &r = reinterpret_cast<Rec2F *>(recF)[q];
      r.s11 = sv(0, 0); r.s22 = sv(1, 1); r.s12 = sv(0, 1);
    }
  } else {
    double *o = src + (size_t)q * (D + D * D) + D;
#pragma unroll
    for (int e = 0; e < D * D; ++e) o[e] = (double)C[e] * sc2;
  }
}

// Isotropic target records from the full ones (every kept covariance s I):
// s times ln 4, as dugma_init builds them.
template <int D>
__global__ __launch_bounds__(kSBlock) void k_iso_records(const void *__restrict__ recF, int n,
                                                         void *__restrict__ recI) {
  const int q = blockIdx.x * kSBlock + threadIdx.x;
  if (q >= n) return;
  if (D == 3) {
    const Rec3F r = reinterpret_cast<const Rec3F *>(recF)[q];
    Rec3I ri;
    ri.hx = r.hx; ri.hy = r.hy; ri.hz = r.hz; ri.s = (float)((double)r.s11 * kLn4);
    ri.lx = ri.ly = ri.lz = ri.pad0 = 0.0f;
    reinterpret_cast<Rec3I *>(recI)[q] = ri;
  } else {
    const Rec2F r = reinterpret_cast<const Rec2F *>(recF)[q];
    Rec2I ri;
    ri.hx = r.hx; ri.hy = r.hy; ri.lx = ri.ly = 0.0f; ri.s = (float)((double)r.s11 * kLn4);
    ri.pad0 = ri.pad1 = ri.pad2 = 0.0f;
    reinterpret_cast<Rec2I *>(recI)[q] = ri;
  }
}

}  // namespace dugma

extern "C" int dugma_update_sigma_knn(dugma_ctx *ctx, int which, int k, float *sigma_target,
                                      float *sigma_source, int *unchanged) {
  if (unchanged) *unchanged = 0;
  if (!ctx) return fail(DUGMA_ERR_ARG, "dugma_update_sigma_knn: null context");
  if (which < 1 || which > 3)
    return fail(DUGMA_ERR_ARG, "dugma_update_sigma_knn: which must be DUGMA_SIGMA_TARGET, DUGMA_SIGMA_SOURCE or both");
  if (k < 1 || k > 32) return fail(DUGMA_ERR_ARG, "dugma_update_sigma_knn: k must be in 1..32");
  dugma_ctx *c = ctx;
  const int D = c->D;
  const bool doT = (which & DUGMA_SIGMA_TARGET) != 0, doS = (which & DUGMA_SIGMA_SOURCE) != 0;
  const int nT = doT ? c->nT : 0, nS = doS ? c->nS : 0;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  // sigbuf: target points | target sigma | source points | source sigma | 2 counts
  const size_t oPT = 0, oST = al((size_t)nT * D * 4), oPS = oST + al((size_t)nT * 4),
               oSS = oPS + al((size_t)nS * D * 4), oCnt = oSS + al((size_t)nS * 4);
  int rc = c->sigbuf.ensure(oCnt + 256);
  if (rc) return rc;
  char *sb = (char *)c->sigbuf.ptr;
  int *cnt = (int *)(sb + oCnt);
  DUGMA_HIP(hipMemsetAsync(cnt, 0, 2 * sizeof(int), c->stream));
  const bool was_iso = c->iso;
  // isotropic target records are kept up to date while the context is isotropic
  void *recI = was_iso ? c->dTrecI : nullptr;
  const dim3 blk(kSBlock);
  for (int cloud = 0; cloud < 2; ++cloud) {
    const bool tgt = cloud == 0;
    const int n = tgt ? nT : nS;
    if (n == 0) continue;
    if (!tgt && nT > 0) DUGMA_HIP(hipStreamSynchronize(c->stream));  // the target's lists are consumed
    float *P = (float *)(sb + (tgt ? oPT : oPS));
    float *sig = (float *)(sb + (tgt ? oST : oSS));
    const int *idx = tgt ? c->dTidx : c->dSidx;
    const float *Pall = tgt ? c->dX : c->dY;
    const dim3 grid((n + kSBlock - 1) / kSBlock);
    if (D == 3)
      hipLaunchKernelGGL(k_gather_points<3>, grid, blk, 0, c->stream, Pall, idx, n, P);
    else
      hipLaunchKernelGGL(k_gather_points<2>, grid, blk, 0, c->stream, Pall, idx, n, P);
    DUGMA_HIP(hipGetLastError());
    const float *od;
    const int *oi;
    rc = knn_device(c, k, nullptr, nullptr, P, n, P, n, 1, &od, &oi);
    if (rc) return rc;
    const double sc2 = c->scale * c->scale;
    float *Cinv = tgt ? c->dXCinv : c->dYCinv;
    float *g = tgt ? c->dgX : c->dgY;
#define SIGMA_APPLY(DD, TT)                                                                                    \
  hipLaunchKernelGGL((k_sigma_apply<DD, TT>), grid, blk, 0, c->stream, od, n, k, idx, sc2, Cinv, g, c->dTrecF, \
                     recI, c->dSsrc, sig, cnt + cloud)
    if (D == 3) {
      if (tgt) SIGMA_APPLY(3, true); else SIGMA_APPLY(3, false);
    } else {
      if (tgt) SIGMA_APPLY(2, true); else SIGMA_APPLY(2, false);
    }
#undef SIGMA_APPLY
    DUGMA_HIP(hipGetLastError());
  }
  // counts (and the sigmas, when asked for) back to the host
  const bool wantT = sigma_target && doT, wantS = sigma_source && doS;
  const size_t hbytes = 16 + (wantT ? (size_t)nT * 4 : 0) + (wantS ? (size_t)nS * 4 : 0);
  rc = c->hstage.ensure(hbytes);
  if (rc) return rc;
  char *h = (char *)c->hstage.ptr;
  DUGMA_HIP(hipMemcpyAsync(h, cnt, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  float *hT = (float *)(h + 16), *hS = hT + (wantT ? nT : 0);
  if (wantT && nT) DUGMA_HIP(hipMemcpyAsync(hT, sb + oST, (size_t)nT * 4, hipMemcpyDeviceToHost, c->stream));
  if (wantS && nS) DUGMA_HIP(hipMemcpyAsync(hS, sb + oSS, (size_t)nS * 4, hipMemcpyDeviceToHost, c->stream));
  DUGMA_HIP(hipStreamSynchronize(c->stream));
  const int badT = ((const int *)h)[0], badS = ((const int *)h)[1];
  if (doT && nT > 0) c->isoT = c->isoT || badT == 0;
  if (doS && nS > 0) {
    c->isoS = c->isoS || badS == 0;
    c->hSsrc_stale = true;
  }
  c->iso = c->isoT && c->isoS && c->nT > 0 && c->nS > 0;
  if (c->iso && !was_iso) {
    // the context has just become isotropic: its isotropic records
    const size_t recszI = D == 3 ? sizeof(Rec3I) : sizeof(Rec2I);
    if (!c->dTrecI) DUGMA_HIP(hipMalloc(&c->dTrecI, (size_t)c->nT * recszI));
    if (!c->dSrecI) DUGMA_HIP(hipMalloc(&c->dSrecI, (size_t)c->nS * recszI));
    const dim3 grid((c->nT + kSBlock - 1) / kSBlock);
    if (D == 3)
      hipLaunchKernelGGL(k_iso_records<3>, grid, blk, 0, c->stream, c->dTrecF, c->nT, c->dTrecI);
    else
      hipLaunchKernelGGL(k_iso_records<2>, grid, blk, 0, c->stream, c->dTrecF, c->nT, c->dTrecI);
    DUGMA_HIP(hipGetLastError());
  }
  if (unchanged) *unchanged = badT + badS;
  const float qnan = std::numeric_limits<float>::quiet_NaN();
  if (wantT) {
    std::fill(sigma_target, sigma_target + c->N, qnan);
    for (int q = 0; q < nT; ++q) sigma_target[c->hTidx[q]] = hT[q];
  }
  if (wantS) {
    std::fill(sigma_source, sigma_source + c->M, qnan);
    for (int q = 0; q < nS; ++q) sigma_source[c->hSidx[q]] = hS[q];
  }
  return DUGMA_OK;
}

extern "C" int dugma_update_cov_knn(dugma_ctx *ctx, int which, int k, double scale, double floor,
                                    float *sigma_target, float *sigma_source, int *unchanged) {
  if (unchanged) *unchanged = 0;
  if (!ctx) return fail(DUGMA_ERR_ARG, "dugma_update_cov_knn: null context");
  if (which < 1 || which > 3)
    return fail(DUGMA_ERR_ARG, "dugma_update_cov_knn: which must be DUGMA_SIGMA_TARGET, DUGMA_SIGMA_SOURCE or both");
  if (k < 2 || k > 32) return fail(DUGMA_ERR_ARG, "dugma_update_cov_knn: k must be in 2..32");
  if (!(scale >= 0) || !(floor >= 0) || !std::isfinite(scale) || !std::isfinite(floor) || !(scale + floor > 0))
    return fail(DUGMA_ERR_ARG, "dugma_update_cov_knn: scale and floor must be >= 0, not both 0");
  dugma_ctx *c = ctx;
  const int D = c->D;
  const bool doT = (which & DUGMA_SIGMA_TARGET) != 0, doS = (which & DUGMA_SIGMA_SOURCE) != 0;
  const int nT = doT ? c->nT : 0, nS = doS ? c->nS : 0;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  // sigbuf: target points | target sigma | source points | source sigma | 2 counts
  const size_t oPT = 0, oST = al((size_t)nT * D * 4), oPS = oST + al((size_t)nT * 4),
               oSS = oPS + al((size_t)nS * D * 4), oCnt = oSS + al((size_t)nS * 4);
  int rc = c->sigbuf.ensure(oCnt + 256);
  if (rc) return rc;
  char *sb = (char *)c->sigbuf.ptr;
  int *cnt = (int *)(sb + oCnt);
  DUGMA_HIP(hipMemsetAsync(cnt, 0, 2 * sizeof(int), c->stream));
  const dim3 blk(kSBlock);
  const float fs = (float)scale, ff = (float)floor;
  for (int cloud = 0; cloud < 2; ++cloud) {
    const bool tgt = cloud == 0;
    const int n = tgt ? nT : nS;
    if (n == 0) continue;
    if (!tgt && nT > 0) DUGMA_HIP(hipStreamSynchronize(c->stream));  // the target's lists are consumed
    float *P = (float *)(sb + (tgt ? oPT : oPS));
    float *sig = (float *)(sb + (tgt ? oST : oSS));
    const int *idx = tgt ? c->dTidx : c->dSidx;
    const float *Pall = tgt ? c->dX : c->dY;
    const dim3 grid((n + kSBlock - 1) / kSBlock);
    if (D == 3)
      hipLaunchKernelGGL(k_gather_points<3>, grid, blk, 0, c->stream, Pall, idx, n, P);
    else
      hipLaunchKernelGGL(k_gather_points<2>, grid, blk, 0, c->stream, Pall, idx, n, P);
    DUGMA_HIP(hipGetLastError());
    const float *od;
    const int *oi;
    rc = knn_device(c, k, nullptr, nullptr, P, n, P, n, 1, &od, &oi);
    if (rc) return rc;
    const double sc2 = c->scale * c->scale;
    float *Cinv = tgt ? c->dXCinv : c->dYCinv;
    float *g = tgt ? c->dgX : c->dgY;
#define COV_APPLY(DD, TT)                                                                                       \
  hipLaunchKernelGGL((k_cov_apply<DD, TT>), grid, blk, 0, c->stream, od, oi, P, n, k, fs, ff, idx, sc2, Cinv, g, \
                     c->dTrecF, c->dSsrc, sig, cnt + cloud)
    if (D == 3) {
      if (tgt) COV_APPLY(3, true); else COV_APPLY(3, false);
    } else {
      if (tgt) COV_APPLY(2, true); else COV_APPLY(2, false);
    }
#undef COV_APPLY
    DUGMA_HIP(hipGetLastError());
  }
  const bool wantT = sigma_target && doT, wantS = sigma_source && doS;
  const size_t hbytes = 16 + (wantT ? (size_t)nT * 4 : 0) + (wantS ? (size_t)nS * 4 : 0);
  rc = c->hstage.ensure(hbytes);
  if (rc) return rc;
  char *h = (char *)c->hstage.ptr;
  DUGMA_HIP(hipMemcpyAsync(h, cnt, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
  float *hT = (float *)(h + 16), *hS = hT + (wantT ? nT : 0);
  if (wantT && nT) DUGMA_HIP(hipMemcpyAsync(hT, sb + oST, (size_t)nT * 4, hipMemcpyDeviceToHost, c->stream));
  if (wantS && nS) DUGMA_HIP(hipMemcpyAsync(hS, sb + oSS, (size_t)nS * 4, hipMemcpyDeviceToHost, c->stream));
  DUGMA_HIP(hipStreamSynchronize(c->stream));
  const int badT = ((const int *)h)[0], badS = ((const int *)h)[1];
  // full covariances from now on (unless no point changed)
  if (doT && nT > badT) c->isoT = false;
  if (doS && nS > badS) {
    c->isoS = false;
    c->hSsrc_stale = true;
  }
  c->iso = c->isoT && c->isoS && c->nT > 0 && c->nS > 0;
  if (unchanged) *unchanged = badT + badS;
  const float qnan = std::numeric_limits<float>::quiet_NaN();
  if (wantT) {
    std::fill(sigma_target, sigma_target + c->N, qnan);
    for (int q = 0; q < nT; ++q) sigma_target[c->hTidx[q]] = hT[q];
  }
  if (wantS) {
    std::fill(sigma_source, sigma_source + c->M, qnan);
    for (int q = 0; q < nS; ++q) sigma_source[c->hSidx[q]] = hS[q];
  }
  return DUGMA_OK;
}
