// MAC stencils and Schur-CG vector kernels (HBM-bound), gfx950.
//
// D (cell divergence) and D^T follow the reference's matrix-free kernels term by term
// (solver/mac/kernels.py:46-132): each face value is divided by the spacing before it
// is added, so the FP64 results are bitwise those of the reference loops.  Reductions
// are two-level with a block count fixed by the vector length, so every sum has one
// fixed order and a solve is bitwise reproducible run to run.
#include "internal.h"

namespace macs {
namespace {

// Workgroups go to the 8 XCDs in turn (block b to XCD b % 8), each XCD with its own L2.
// Renumbered so that each XCD walks one contiguous eighth of the index range, a
// stencil's neighbour rows stay in the L2 that loaded them instead of being fetched again
// for another XCD.  A bijection on [0, nb): partials indexed by the logical block keep
// every sum's order.
__device__ inline unsigned xcd_block(unsigned b, unsigned nb) {
  const unsigned per = nb / 8;
  return b < 8 * per ? (b % 8) * per + b / 8 : b;
}

__device__ inline double ld_d(const double* p, int64_t i) { return p[i]; }
__device__ inline double ld_d(const float* p, int64_t i) { return (double)p[i]; }
__device__ inline double ld_d(const __bf16* p, int64_t i) { return (double)(float)p[i]; }
__device__ inline void st_d(double* p, int64_t i, double v) { p[i] = v; }
__device__ inline void st_d(float* p, int64_t i, double v) { p[i] = (float)v; }
__device__ inline void st_d(__bf16* p, int64_t i, double v) { p[i] = (__bf16)(float)v; }
// D at one cell, term by term as the reference loop (+u_e/hx, -u_w/hx, +v_n/hy, -v_s/hy,
// each divided before it is added).  The four face loads go out first at clamped
// indices (an absent face is read from a valid address and left unused), so they
// overlap instead of each waiting behind its own branch.
template <typename TIN>
__device__ inline double cell_divergence(const TIN* u, const TIN* v, const VelLayout& L, const GridDims& g, int j,
                                         int i) {
  const int nx = g.nx, ny = g.ny;
  const double ue = ld_d(u, (int64_t)j * L.ldu + min(i, nx - 2));
  const double uw = ld_d(u, (int64_t)j * L.ldu + max(i - 1, 0));
  const double vn = ld_d(v, (int64_t)min(j, ny - 2) * L.ldv + i);
  const double vs = ld_d(v, (int64_t)max(j - 1, 0) * L.ldv + i);
  double value = 0.0;
  if (i < nx - 1) value += ue / g.hx;
  if (i > 0) value -= uw / g.hx;
  if (j < ny - 1) value += vn / g.hy;
  if (j > 0) value -= vs / g.hy;
  return value;
}

template <typename TIN, int MODE>
__global__ __launch_bounds__(kThreads) void k_div(const TIN* __restrict__ vel, VelLayout L, GridDims g,
                                                  double* __restrict__ out, int64_t out_stride, double sign,
                                                  double* partials, const CgState* st, const double* add) {
  const int c = blockIdx.y;
  if (st != nullptr && !st[c].active) return;
  __shared__ double sh[4];
  const TIN* u = vel + (int64_t)c * L.stride;
  const TIN* v = u + L.voff;
  double* o = out + (int64_t)c * out_stride;
  const int nx = g.nx, ny = g.ny;
  const int np = nx * ny;
  const SmallDiv dnx(nx);
  const unsigned bx = xcd_block(blockIdx.x, gridDim.x);
  double accum = 0.0;
  for (int cell = bx * kThreads + threadIdx.x; cell < np; cell += gridDim.x * kThreads) {
    const int j = dnx.quot(cell), i = cell - j * nx;
    const double a = add != nullptr ? add[(int64_t)c * out_stride + cell] : 0.0;
    double value = sign * cell_divergence(u, v, L, g, j, i);
    if (add != nullptr) value = a + value;
    o[cell] = value;
    if (MODE == RED_SUM) accum += value;
    if (MODE == RED_MAXABS) accum = fmax(accum, absn(value));
  }
  if (MODE != RED_NONE) {
    accum = block_reduce<MODE>(accum, sh);
    if (threadIdx.x == 0) partials[(int64_t)c * gridDim.x + bx] = accum;
  }
}

// D on the velocity columns like k_div (no add), with three partial sums per block for
// the mixed CG: sum(out), sum(out * y) and sum(y), at partials[(k * ncols + c) * gx + bx].
template <typename TIN>
__global__ __launch_bounds__(kThreads) void k_div_sums3(const TIN* __restrict__ vel, VelLayout L, GridDims g,
                                                        double* __restrict__ out, int64_t out_stride, double sign,
                                                        const double* __restrict__ y, double* partials,
                                                        const CgState* st) {
  const int c = blockIdx.y;
  if (st != nullptr && !st[c].active) return;
  __shared__ double sh[4];
  const TIN* u = vel + (int64_t)c * L.stride;
  const TIN* v = u + L.voff;
  double* o = out + (int64_t)c * out_stride;
  const double* yc = y + (int64_t)c * out_stride;
  const int nx = g.nx, ny = g.ny;
  const int np = nx * ny;
  const SmallDiv dnx(nx);
  const unsigned bx = xcd_block(blockIdx.x, gridDim.x);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int cell = bx * kThreads + threadIdx.x; cell < np; cell += gridDim.x * kThreads) {
    const int j = dnx.quot(cell), i = cell - j * nx;
    const double yv = yc[cell];
    const double value = sign * cell_divergence(u, v, L, g, j, i);
    o[cell] = value;
    s0 += value;
    s1 += value * yv;
    s2 += yv;
  }
  const int64_t gx = gridDim.x, nc = gridDim.y;
  s0 = block_reduce<RED_SUM>(s0, sh);
  __syncthreads();
  if (threadIdx.x == 0) partials[(0 * nc + c) * gx + bx] = s0;
  s1 = block_reduce<RED_SUM>(s1, sh);
  __syncthreads();
  if (threadIdx.x == 0) partials[(1 * nc + c) * gx + bx] = s1;
  s2 = block_reduce<RED_SUM>(s2, sh);
  if (threadIdx.x == 0) partials[(2 * nc + c) * gx + bx] = s2;
}

// sum(a), sum(a * b), sum(b) per column, partials as k_div_sums3.
__global__ __launch_bounds__(kThreads) void k_sums3(const double* __restrict__ a, const double* __restrict__ b,
                                                    int64_t n, int64_t stride, const CgState* st,
                                                    double* partials) {
  const int c = blockIdx.y;
  if (!st[c].active) return;
  __shared__ double sh[4];
  const int64_t off = (int64_t)c * stride;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  // Two elements per step with their loads issued first, summed in index order.
  const int64_t step = (int64_t)gridDim.x * kThreads;
  int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  for (; i + step < n; i += 2 * step) {
    const double a0 = a[off + i], b0 = b[off + i], a1 = a[off + i + step], b1 = b[off + i + step];
    s0 += a0;
    s1 += a0 * b0;
    s2 += b0;
    s0 += a1;
    s1 += a1 * b1;
    s2 += b1;
  }
  for (; i < n; i += step) {
    const double av = a[off + i], bv = b[off + i];
    s0 += av;
    s1 += av * bv;
    s2 += bv;
  }
  const int64_t gx = gridDim.x, nc = gridDim.y;
  s0 = block_reduce<RED_SUM>(s0, sh);
  __syncthreads();
  if (threadIdx.x == 0) partials[(0 * nc + c) * gx + blockIdx.x] = s0;
  s1 = block_reduce<RED_SUM>(s1, sh);
  __syncthreads();
  if (threadIdx.x == 0) partials[(1 * nc + c) * gx + blockIdx.x] = s1;
  s2 = block_reduce<RED_SUM>(s2, sh);
  if (threadIdx.x == 0) partials[(2 * nc + c) * gx + blockIdx.x] = s2;
}

template <typename TOUT>
__global__ __launch_bounds__(kThreads) void k_dt(const double* __restrict__ p, int64_t p_stride,
                                                 const double* mean_sums, const double* __restrict__ add,
                                                 int64_t add_stride, TOUT* __restrict__ out, VelLayout L,
                                                 GridDims g, const CgState* st) {
  // One thread per cell: it writes the u face east of the cell and the v face north of
  // it, so every pressure value is fetched once (per face the same arithmetic as the
  // reference loop).
  const int c = blockIdx.y;
  if (st != nullptr && !st[c].active) return;
  const int nx = g.nx, ny = g.ny, np = nx * ny;
  const int cell = (int)xcd_block(blockIdx.x, gridDim.x) * kThreads + (int)threadIdx.x;
  if (cell >= np) return;
  const int64_t nu = (int64_t)(nx - 1) * ny;
  const double* pc = p + (int64_t)c * p_stride;
  const double* ac = add != nullptr ? add + (int64_t)c * add_stride : nullptr;
  const double mean = mean_sums ? mean_sums[c] / (double)np : 0.0;
  TOUT* oc = out + (int64_t)c * L.stride;
  const int j = SmallDiv(nx).quot(cell), i = cell - j * nx;
  // Loads first, at clamped indices (unused where the face does not exist).
  const double pv = pc[cell], pe = pc[min(cell + 1, np - 1)], pn = pc[min(cell + nx, np - 1)];
  const double au = ac != nullptr ? ac[(int64_t)j * (nx - 1) + min(i, nx - 2)] : 0.0;
  const double av = ac != nullptr ? ac[nu + min(cell, np - nx - 1)] : 0.0;
  if (i < nx - 1) {
    double value = mean_sums ? ((pv - mean) - (pe - mean)) / g.hx : (pv - pe) / g.hx;
    if (ac != nullptr) value = au + value;
    st_d(oc, (int64_t)j * L.ldu + i, value);
  }
  if (j < ny - 1) {
    double value = mean_sums ? ((pv - mean) - (pn - mean)) / g.hy : (pv - pn) / g.hy;
    if (ac != nullptr) value = av + value;
    st_d(oc, L.voff + (int64_t)j * L.ldv + i, value);
  }
}

template <int MODE, bool DOT>
__global__ __launch_bounds__(kThreads) void k_reduce(const double* __restrict__ x, const double* __restrict__ y,
                                                     int64_t n, int64_t stride, double* partials,
                                                     const CgState* st) {
  const int c = blockIdx.y;
  if (st != nullptr && !st[c].active) return;
  __shared__ double sh[4];
  const double* xc = x + (int64_t)c * stride;
  const double* yc = DOT ? y + (int64_t)c * stride : nullptr;
  double accum = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    if (MODE == RED_MAXABS) accum = fmax(accum, absn(xc[i]));
    else accum += DOT ? xc[i] * yc[i] : xc[i];
  }
  accum = block_reduce<MODE>(accum, sh);
  if (threadIdx.x == 0) partials[(int64_t)c * gridDim.x + blockIdx.x] = accum;
}

// One column's partial results combined in a fixed order (thread-strided, then
// block_reduce): the value every reduction of this file reports, valid on thread 0.
// These run as one block per column with nothing else to hide memory latency, so the
// loads go out eight at a time; each thread still adds its values in index order.
template <int MODE>
__device__ inline double combine_partials(const double* part, int nblk, double* sh) {
  double accum = 0.0;
  int i = threadIdx.x;
  for (; i + 7 * kThreads < nblk; i += 8 * kThreads) {
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = part[i + k * kThreads];
#pragma unroll
    for (int k = 0; k < 8; ++k) accum = MODE == RED_MAXABS ? fmax(accum, v[k]) : accum + v[k];
  }
  for (; i < nblk; i += kThreads) {
    const double v = part[i];
    accum = MODE == RED_MAXABS ? fmax(accum, v) : accum + v;
  }
  return block_reduce<MODE>(accum, sh);
}

template <int MODE>
__global__ __launch_bounds__(kThreads) void k_finalize(const double* partials, int nblk, double* out) {
  const int c = blockIdx.x;
  __shared__ double sh[4];
  const double accum = combine_partials<MODE>(partials + (int64_t)c * nblk, nblk, sh);
  if (threadIdx.x == 0) out[c] = accum;
}

// Two reductions' final combines in one launch: blocks [0, n1) combine (p1, len1) into
// out1 as k_finalize does, the others (p2, len2) into out2.
template <int MODE1, int MODE2>
__global__ __launch_bounds__(kThreads) void k_finalize2(const double* p1, int len1, int n1, double* out1,
                                                        const double* p2, int len2, double* out2) {
  __shared__ double sh[4];
  const bool first = (int)blockIdx.x < n1;
  const int c = first ? (int)blockIdx.x : (int)blockIdx.x - n1;
  const double accum = first ? combine_partials<MODE1>(p1 + (int64_t)c * len1, len1, sh)
                             : combine_partials<MODE2>(p2 + (int64_t)c * len2, len2, sh);
  if (threadIdx.x == 0) (first ? out1 : out2)[c] = accum;
}

// Start of a mixed sweep's inner CG on b: b -= mean(b) (sums[c] = sum of b), r = b,
// x = 0, with partial (b, b) at partials[c * gx + bx] (the order of k_sub_mean_dot) and
// partial max|b| at partials[(ncols + c) * gx + bx].
__global__ __launch_bounds__(kThreads) void k_sweep_start(double* __restrict__ b, double* __restrict__ r,
                                                          double* __restrict__ x, int64_t n, const double* sums,
                                                          double* partials, float* __restrict__ r32) {
  const int c = blockIdx.y;
  __shared__ double sh[4];
  const int64_t off = (int64_t)c * n;
  const double mean = sums[c] / (double)n;
  double dot = 0.0, mx = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    const double v = b[off + i] - mean;
    b[off + i] = v;
    r[off + i] = v;
    if (r32 != nullptr) r32[off + i] = (float)v;
    x[off + i] = 0.0;
    dot += v * v;
    mx = fmax(mx, absn(v));
  }
  const int64_t gx = gridDim.x, nc = gridDim.y;
  dot = block_reduce<RED_SUM>(dot, sh);
  __syncthreads();
  if (threadIdx.x == 0) partials[c * gx + blockIdx.x] = dot;
  mx = block_reduce<RED_MAXABS>(mx, sh);
  if (threadIdx.x == 0) partials[(nc + c) * gx + blockIdx.x] = mx;
}

template <bool DOT>
__global__ __launch_bounds__(kThreads) void k_sub_mean_dot(double* x, int64_t n, int64_t stride,
                                                           const double* sums, const double* y,
                                                           double* partials, const CgState* st) {
  const int c = blockIdx.y;
  if (st != nullptr && !st[c].active) return;
  __shared__ double sh[4];
  double* xc = x + (int64_t)c * stride;
  const double* yc = DOT ? y + (int64_t)c * stride : nullptr;
  const double mean = sums[c] / (double)n;
  double accum = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    const double v = xc[i] - mean;
    xc[i] = v;
    if (DOT) accum += yc[i] * v;
  }
  if (DOT) {
    accum = block_reduce<RED_SUM>(accum, sh);
    if (threadIdx.x == 0) partials[(int64_t)c * gridDim.x + blockIdx.x] = accum;
  }
}

__global__ void k_sub_cols(const double* a, const double* b, double* out, int64_t n, int64_t stride,
                           const CgState* st) {
  const int c = blockIdx.y;
  if (st != nullptr && !st[c].active) return;
  const int64_t off = (int64_t)c * stride;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    out[off + i] = a[off + i] - b[off + i];
}

// Galerkin-scaled start: theta = (g, b) / (g, S g) per column (0 when (g, S g) <= 0);
// x = theta g, r = b - theta S g.  x holds g and q holds S g on entry.
__global__ void k_theta_start(double* __restrict__ x, double* __restrict__ r, const double* __restrict__ b,
                              const double* __restrict__ q, const double* gb, const double* gq, int64_t n,
                              int64_t stride, const CgState* st) {
  const int c = blockIdx.y;
  if (st != nullptr && !st[c].active) return;
  const double theta = gq[c] > 0.0 ? gb[c] / gq[c] : 0.0;
  const int64_t off = (int64_t)c * stride;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    x[off + i] = theta * x[off + i];
    r[off + i] = b[off + i] - theta * q[off + i];
  }
}

// Two-term start over span{g, c} (base guess g, correction c), per column: the
// coefficients minimise the starting residual ||b - a S g - beta S c||_2, the quantity
// the CG stopping test measures, so the start is never worse in it than any multiple of
// the base alone:
//   [(v1,v1) (v1,v2)] [a   ]   [(v1,b)]
//   [(v1,v2) (v2,v2)] [beta] = [(v2,b)],   v1 = S g (taken as bg), v2 = S c = q.
// A near-singular pair falls back to the base alone (beta = 0), a useless base to zero.
// dots holds (v1,v1), (v1,v2), (v2,v2), (v1,b), (v2,b), ncols each.
__global__ void k_galerkin2(double* __restrict__ x, double* __restrict__ r, const double* __restrict__ g,
                            const double* __restrict__ cc, const double* __restrict__ b,
                            const double* __restrict__ bg, const double* __restrict__ q, const double* dots,
                            int ncols, int64_t n, const CgState* st) {
  const int c = blockIdx.y;
  if (st != nullptr && !st[c].active) return;
  const double v11 = dots[c], v12 = dots[ncols + c], v22 = dots[2 * ncols + c];
  const double v1b = dots[3 * ncols + c], v2b = dots[4 * ncols + c];
  const double det = v11 * v22 - v12 * v12;
  double a = 0.0, beta = 0.0;
  if (v11 > 0.0 && v22 > 0.0 && det > 1e-12 * v11 * v22) {
    a = (v1b * v22 - v12 * v2b) / det;
    beta = (v11 * v2b - v12 * v1b) / det;
  } else if (v11 > 0.0) {
    a = v1b / v11;
  }
  const int64_t off = (int64_t)c * n;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    const double gi = g[off + i], ci = cc[off + i];
    x[off + i] = a * gi + beta * ci;
    r[off + i] = b[off + i] - a * bg[off + i] - beta * q[off + i];
  }
}

__global__ void k_fill(double* x, int64_t n, double v) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) x[i] = v;
}

// A solve's hand-out in one pass: v = src - sums[c] / n (sums null: v = src) written to up
// to three destinations (null ones skipped), column c at c * stride of each.
__global__ void k_finish_cols(const double* __restrict__ src, int64_t n, const double* __restrict__ sums,
                              double* __restrict__ d0, int64_t s0, double* __restrict__ d1, int64_t s1,
                              double* __restrict__ d2, int64_t s2) {
  const int c = blockIdx.y;
  const double mean = sums != nullptr ? sums[c] / (double)n : 0.0;
  const double* sc = src + (int64_t)c * n;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    const double v = sums != nullptr ? sc[i] - mean : sc[i];
    if (d0 != nullptr) d0[(int64_t)c * s0 + i] = v;
    if (d1 != nullptr) d1[(int64_t)c * s1 + i] = v;
    if (d2 != nullptr) d2[(int64_t)c * s2 + i] = v;
  }
}

__global__ void k_copy_cols(const double* src, int64_t ss, double* dst, int64_t ds, int64_t n) {
  const int c = blockIdx.y;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    dst[(int64_t)c * ds + i] = src[(int64_t)c * ss + i];
}

// ---- CG scalars (scipy 1.15.3 cg loop: stop when ||r|| < rtol*||b||, tested at the
// top of each iteration; rho = (r,z); beta = rho/rho_prev; alpha = rho/(p,q)).
__global__ void k_cg_init(CgState* st, const double* bb, const double* bmax, double rtol, int ncols,
                          double* rr_out, double mtol) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  if (rr_out != nullptr) rr_out[c] = bb[c];
  CgState s{};
  s.bnorm = sqrt(bb[c]);
  s.tol = rtol * s.bnorm;
  s.mtol = mtol;
  s.rr = bb[c];
  s.active = (bmax[c] != 0.0 && s.bnorm != 0.0) ? 1 : 0;
  s.converged = s.active ? 0 : 1;
  s.iters = 0;
  st[c] = s;
}

// A mixed sweep's CG start after k_sweep_start, one block per column: (b, b) and max|b|
// combined from the sweep start's partials (as k_finalize2 combines them), the column's
// state (as k_cg_init) and the first convergence check (as k_cg_check).
__global__ __launch_bounds__(kThreads) void k_sweep_cg_start(const double* partials, int gx, int ncols, double* bb,
                                                             double* bmax, CgState* st, double rtol, double* rr_out,
                                                             double mtol, int* flags) {
  __shared__ double sh[4];
  const int c = blockIdx.x;
  const double b2 = combine_partials<RED_SUM>(partials + (int64_t)c * gx, gx, sh);
  __syncthreads();  // sh is reused
  const double bm = combine_partials<RED_MAXABS>(partials + (int64_t)(ncols + c) * gx, gx, sh);
  if (threadIdx.x != 0) return;
  bb[c] = b2;
  bmax[c] = bm;
  if (rr_out != nullptr) rr_out[c] = b2;
  CgState v{};
  v.bnorm = sqrt(b2);
  v.tol = rtol * v.bnorm;
  v.mtol = mtol;
  v.rr = b2;
  v.active = (bm != 0.0 && v.bnorm != 0.0) ? 1 : 0;
  v.converged = v.active ? 0 : 1;
  v.iters = 0;
  if (v.active && sqrt(b2) < v.tol) {
    v.active = 0;
    v.converged = 1;
  }
  st[c] = v;
  flags[c] = v.active;
}

__global__ void k_cg_check(CgState* st, const double* rr, int* flags, int ncols) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  if (st[c].active) {
    st[c].rr = rr[c];
    if (sqrt(rr[c]) < st[c].tol) {
      st[c].active = 0;
      st[c].converged = 1;
    }
  }
  flags[c] = st[c].active;
}

// Polynomial extrapolation in time from equally spaced past values (newest first):
// out = 2 a - b (linear), or 3 a - 3 b + c (quadratic) when c is given.
__global__ void k_extrapolate(Extrapolation e, double* __restrict__ out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    double v[kMaxExtrapolation];
#pragma unroll
    for (int k = 0; k < kMaxExtrapolation; ++k) v[k] = k < e.count ? e.src[k][i] : 0.0;  // loads first
    double acc = e.w[0] * v[0];
#pragma unroll
    for (int k = 1; k < kMaxExtrapolation; ++k)
      if (k < e.count) acc += e.w[k] * v[k];
    out[i] = acc;
  }
}

// The same with two elements per thread and 16-byte loads (every slot and out 16-byte
// aligned; pairs i of n2), one pair per thread: each element's arithmetic and order as
// above.
__global__ __launch_bounds__(kThreads) void k_extrapolate2(Extrapolation e, double2* __restrict__ out, int64_t n2) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= n2) return;
  double2 v[kMaxExtrapolation];
#pragma unroll
  for (int k = 0; k < kMaxExtrapolation; ++k)
    v[k] = k < e.count ? reinterpret_cast<const double2*>(e.src[k])[i] : make_double2(0.0, 0.0);  // loads first
  double ax = e.w[0] * v[0].x, ay = e.w[0] * v[0].y;
#pragma unroll
  for (int k = 1; k < kMaxExtrapolation; ++k)
    if (k < e.count) {
      ax += e.w[k] * v[k].x;
      ay += e.w[k] * v[k].y;
    }
  out[i] = make_double2(ax, ay);
}

// With partials, rr is first finalized from them (one column after another, in
// k_finalize's order) by this block.
__global__ __launch_bounds__(kThreads) void k_cg_check_post(CgState* st, double* rr, const double* partials,
                                                            int nblk, int* flags, int ncols, Mailbox* mb,
                                                            int seq) {
  __shared__ double sh[4];
  __shared__ double rmax[64];
  if (partials != nullptr) {
    for (int k = 0; k < ncols; ++k) {
      const double v = combine_partials<RED_SUM>(partials + (int64_t)k * nblk, nblk, sh);
      if (threadIdx.x == 0) rr[k] = v;
      __syncthreads();  // sh is reused; rr[k] is read by thread k below
      // max|r| from the update's second set of partials, for a column with a max-norm stop
      double m = INFINITY;
      if (st[k].mtol > 0.0) m = combine_partials<RED_MAXABS>(partials + (int64_t)(ncols + k) * nblk, nblk, sh);
      if (threadIdx.x == 0) rmax[k] = m;
      __syncthreads();
    }
  }
  const int c = threadIdx.x;
  if (c < ncols) {
    if (st[c].active) {
      st[c].rr = rr[c];
      const bool small = partials != nullptr && st[c].mtol > 0.0 && rmax[c] <= st[c].mtol;
      if (sqrt(rr[c]) < st[c].tol || small) {
        st[c].active = 0;
        st[c].converged = 1;
      }
    }
    flags[c] = st[c].active;
    mb->flags[c] = st[c].active;
    mb->iters[c] = st[c].iters;
    mb->converged[c] = st[c].converged;
  }
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(&mb->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ void k_post_values(const double* a, int na, const double* b, int nb, Mailbox* mb, int seq) {
  const int t = threadIdx.x;
  if (t < na) mb->vals[t] = a[t];
  else if (t < na + nb) mb->vals[t] = b[t - na];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(&mb->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// The CG scalar steps fused with the finalize of their dot product's partials.
__global__ __launch_bounds__(kThreads) void k_cg_rho_fin(CgState* st, const double* partials, int nblk,
                                                         double* out) {
  __shared__ double sh[4];
  const int c = blockIdx.x;
  const double rz = combine_partials<RED_SUM>(partials + (int64_t)c * nblk, nblk, sh);
  if (threadIdx.x != 0) return;
  out[c] = rz;
  if (!st[c].active) return;
  st[c].rho = rz;
  st[c].beta = st[c].iters > 0 ? rz / st[c].rho_prev : 0.0;
}

__global__ __launch_bounds__(kThreads) void k_cg_alpha_fin(CgState* st, const double* partials, int nblk,
                                                           double* out) {
  __shared__ double sh[4];
  const int c = blockIdx.x;
  const double pq = combine_partials<RED_SUM>(partials + (int64_t)c * nblk, nblk, sh);
  if (threadIdx.x != 0) return;
  out[c] = pq;
  if (!st[c].active) return;
  st[c].alpha = st[c].rho / pq;
  st[c].rho_prev = st[c].rho;
  st[c].iters += 1;
}

// Mixed-CG scalar steps on the unprojected vectors: from S_a = sum(a), S_ab = sum(a b),
// S_b = sum(b) (partials of k_sums3 / k_div_sums3), mean = S_a / n and the projected dot
// (a - mean, b) = S_ab - mean S_b.  rho from (z, r); alpha from (q, p).
__device__ inline void sums3(const double* partials, int gx, int ncols, int c, double* sh, double* out3) {
  for (int k = 0; k < 3; ++k) {
    const double v = combine_partials<RED_SUM>(partials + ((int64_t)k * ncols + c) * gx, gx, sh);
    if (threadIdx.x == 0) out3[k] = v;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kThreads) void k_cg_rho_mean(CgState* st, const double* partials, int gx, int ncols,
                                                          double n, double* mean_out) {
  __shared__ double sh[4];
  __shared__ double s3[3];
  const int c = blockIdx.x;
  sums3(partials, gx, ncols, c, sh, s3);
  if (threadIdx.x != 0) return;
  const double mean = s3[0] / n;
  mean_out[c] = mean;
  if (!st[c].active) return;
  const double rz = s3[1] - mean * s3[2];
  st[c].rho = rz;
  st[c].beta = st[c].iters > 0 ? rz / st[c].rho_prev : 0.0;
}

__global__ __launch_bounds__(kThreads) void k_cg_alpha_mean(CgState* st, const double* partials, int gx,
                                                            int ncols, double n, double* mean_out) {
  __shared__ double sh[4];
  __shared__ double s3[3];
  const int c = blockIdx.x;
  sums3(partials, gx, ncols, c, sh, s3);
  if (threadIdx.x != 0) return;
  const double mean = s3[0] / n;
  mean_out[c] = mean;
  if (!st[c].active) return;
  st[c].alpha = st[c].rho / (s3[1] - mean * s3[2]);
  st[c].rho_prev = st[c].rho;
  st[c].iters += 1;
}

// p = (z - zmean) + beta p (first iteration: z - zmean).
__global__ __launch_bounds__(kThreads) void k_update_p_mean(double* __restrict__ p, const double* __restrict__ z,
                                                            const double* __restrict__ zmean, int64_t n,
                                                            int64_t stride, const CgState* st) {
  const int c = blockIdx.y;
  if (!st[c].active) return;
  const bool first = st[c].iters == 0;
  const double beta = st[c].beta, zm = zmean[c];
  double* pc = p + (int64_t)c * stride;
  const double* zc = z + (int64_t)c * stride;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    const double zv = zc[i] - zm;
    pc[i] = first ? zv : pc[i] * beta + zv;
  }
}

// x += alpha p; r -= alpha (q - qmean); partial sums of r r.
__global__ __launch_bounds__(kThreads) void k_update_xr_mean(double* __restrict__ x, double* __restrict__ r,
                                                             const double* __restrict__ p,
                                                             const double* __restrict__ q,
                                                             const double* __restrict__ qmean, int64_t n,
                                                             int64_t stride, const CgState* st, double* partials,
                                                             float* __restrict__ r32) {
  const int c = blockIdx.y;
  if (!st[c].active) return;
  __shared__ double sh[4];
  const double alpha = st[c].alpha, qm = qmean[c];
  const int64_t off = (int64_t)c * stride;
  double accum = 0.0, amax = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    x[off + i] += alpha * p[off + i];
    const double rv = r[off + i] - alpha * (q[off + i] - qm);
    r[off + i] = rv;
    if (r32 != nullptr) r32[off + i] = (float)rv;
    accum += rv * rv;
    amax = fmax(amax, absn(rv));
  }
  accum = block_reduce<RED_SUM>(accum, sh);
  __syncthreads();
  amax = block_reduce<RED_MAXABS>(amax, sh);
  // (r, r) per block, then max|r| per block in a second set after every column's sums.
  if (threadIdx.x == 0) {
    partials[(int64_t)c * gridDim.x + blockIdx.x] = accum;
    partials[(int64_t)(gridDim.y + c) * gridDim.x + blockIdx.x] = amax;
  }
}

__global__ void k_cg_rho(CgState* st, const double* rz, int ncols) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols || !st[c].active) return;
  st[c].rho = rz[c];
  st[c].beta = st[c].iters > 0 ? rz[c] / st[c].rho_prev : 0.0;
}

__global__ void k_cg_alpha(CgState* st, const double* pq, int ncols) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols || !st[c].active) return;
  st[c].alpha = st[c].rho / pq[c];
  st[c].rho_prev = st[c].rho;
  st[c].iters += 1;
}

__global__ __launch_bounds__(kThreads) void k_update_p(double* __restrict__ p, const double* __restrict__ z,
                                                       int64_t n, int64_t stride, const CgState* st) {
  const int c = blockIdx.y;
  if (!st[c].active) return;
  const bool first = st[c].iters == 0;
  const double beta = st[c].beta;
  double* pc = p + (int64_t)c * stride;
  const double* zc = z + (int64_t)c * stride;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    pc[i] = first ? zc[i] : __dadd_rn(__dmul_rn(pc[i], beta), zc[i]);
}

__global__ __launch_bounds__(kThreads) void k_update_xr(double* __restrict__ x, double* __restrict__ r,
                                                        const double* __restrict__ p, const double* __restrict__ q,
                                                        int64_t n, int64_t stride, const CgState* st,
                                                        double* partials) {
  const int c = blockIdx.y;
  if (!st[c].active) return;
  __shared__ double sh[4];
  const double alpha = st[c].alpha;
  const int64_t off = (int64_t)c * stride;
  double accum = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
    x[off + i] = __dadd_rn(x[off + i], __dmul_rn(alpha, p[off + i]));
    const double rv = __dsub_rn(r[off + i], __dmul_rn(alpha, q[off + i]));
    r[off + i] = rv;
    accum += rv * rv;
  }
  accum = block_reduce<RED_SUM>(accum, sh);
  if (threadIdx.x == 0) partials[(int64_t)c * gridDim.x + blockIdx.x] = accum;
}

// defect = mass*vel + visc*(K vel) - D^T p - f on every free face (operators.py:6-17 for
// K: 5-point, tangential wall diagonal 3, normal diagonal 2), max-abs partials of the
// defect, of f and of vel (three per block).
template <typename TR>
__global__ __launch_bounds__(kThreads) void k_defect(const double* __restrict__ vel, const double* __restrict__ f,
                                                     const double* __restrict__ p, GridDims g, int ncols,
                                                     double mass, double visc, double* partials,
                                                     TR* __restrict__ rout, VelLayout Lr, int want_div,
                                                     double* __restrict__ div_out, double div_sign) {
  // One thread per cell: the u face east of the cell and the v face north of it, so the
  // pressure and both velocity components are each walked once, row by row.  Per face
  // the arithmetic is the one of K u + G p - f term by term.  With want_div the thread
  // also forms div_sign * (D u) at its cell from the same loads, term by term as k_div.
  __shared__ double sh[4];
  const int nx = g.nx, ny = g.ny, w = nx - 1;
  const int nu = w * ny, nv = nx * (ny - 1), np = nx * ny;  // nx, ny >= 3
  const int64_t size = (int64_t)nu + nv;
  const double hx2 = g.hx * g.hx, hy2 = g.hy * g.hy;
  const double ox = -1.0 / hx2, oy = -1.0 / hy2;
  // Diagonals of K: tangential walls (the u faces of the first and last rows, the v
  // faces of the first and last columns) carry 3 / h^2 instead of 2 / h^2.
  const double diag_u_in = 2.0 / hx2 + 2.0 / hy2, diag_u_wall = 2.0 / hx2 + 3.0 / hy2;
  const double diag_v_in = 2.0 / hx2 + 2.0 / hy2, diag_v_wall = 3.0 / hx2 + 2.0 / hy2;
  double md = 0.0, mf = 0.0, mv = 0.0, mdiv = 0.0;
  // Faces of one column are < 2^31 (the grid caps nx, ny at 16384): 32-bit index math,
  // columns walked by the grid's y dimension.
  const int c = blockIdx.y;
  double* dc = div_out != nullptr ? div_out + (int64_t)c * np : nullptr;
  const double* u = vel + (int64_t)c * size;
  const double* v = u + nu;
  const double* fc = f + (int64_t)c * size;
  const double* pc = p + (int64_t)c * np;
  TR* rc = rout != nullptr ? rout + (int64_t)c * Lr.stride : nullptr;
  const SmallDiv dnx(nx);
  const unsigned bx = xcd_block(blockIdx.x, gridDim.x);
  for (int cell = bx * kThreads + threadIdx.x; cell < np; cell += gridDim.x * kThreads) {
    const int j = dnx.quot(cell), i = cell - j * nx;
    // Every load goes out first, at an index clamped into its array (a value whose face
    // or neighbour does not exist is read from a valid address and left unused), so the
    // loads overlap instead of each waiting behind its own branch.
    const bool east = i < w, west = i > 0, south = j > 0, north = j < ny - 1;
    const int fu = min(j * w + i, nu - 1);  // the u face east of the cell, when east
    const int e = min(cell, nv - 1);        // the v face north of the cell, when north
    const double uc = u[fu], uw = u[max(j * w + i - 1, 0)], ue = u[min(fu + 1, nu - 1)];
    const double us = u[max(fu - w, 0)], un = u[min(fu + w, nu - 1)];
    const double vc = v[e], vs = v[max(cell - nx, 0)], vw = v[max(e - 1, 0)];
    const double ve = v[min(e + 1, nv - 1)], vn = v[min(e + nx, nv - 1)];
    const double pv = pc[cell], pe = pc[min(cell + 1, np - 1)], pn = pc[min(cell + nx, np - 1)];
    const double fu_v = fc[fu], fv_v = fc[nu + e];
    // Both faces are evaluated on every lane and enter the maxima through selects (a
    // face that does not exist adds 0 to a maximum of magnitudes), so no load is sunk
    // into a branch; only the stores are conditional.
    FaceRow ku;
    if (south) ku.add(oy, us);
    if (west) ku.add(ox, uw);
    ku.add(south && north ? diag_u_in : diag_u_wall, uc);
    if (i < w - 1) ku.add(ox, ue);
    if (north) ku.add(oy, un);
    const double du_ = face_defect(mass, uc, visc, ku.kv, (pv - pe) / g.hx, fu_v);
    FaceRow kv;
    if (south) kv.add(oy, vs);
    if (west) kv.add(ox, vw);
    kv.add(west && east ? diag_v_in : diag_v_wall, vc);
    if (east) kv.add(ox, ve);
    if (j < ny - 2) kv.add(oy, vn);
    const double dv_ = face_defect(mass, vc, visc, kv.kv, (pv - pn) / g.hy, fv_v);
    if (rc != nullptr) {  // -defect = r_u, in the output layout and type
      if (east) st_d(rc, (int64_t)j * Lr.ldu + i, -du_);
      if (north) st_d(rc, Lr.voff + (int64_t)j * Lr.ldv + i, -dv_);
    }
    md = fmax(md, fmax(east ? absn(du_) : 0.0, north ? absn(dv_) : 0.0));
    mf = fmax(mf, fmax(east ? absn(fu_v) : 0.0, north ? absn(fv_v) : 0.0));
    mv = fmax(mv, fmax(east ? absn(uc) : 0.0, north ? absn(vc) : 0.0));
    if (want_div) {
      double value = 0.0;
      if (east) value += uc / g.hx;
      if (west) value -= uw / g.hx;
      if (north) value += vc / g.hy;
      if (south) value -= vs / g.hy;
      value = div_sign * value;
      if (dc != nullptr) dc[cell] = value;
      mdiv = fmax(mdiv, absn(value));
    }
  }
  md = block_reduce<RED_MAXABS>(md, sh);
  __syncthreads();
  mf = block_reduce<RED_MAXABS>(mf, sh);
  __syncthreads();
  mv = block_reduce<RED_MAXABS>(mv, sh);
  if (want_div) {
    __syncthreads();
    mdiv = block_reduce<RED_MAXABS>(mdiv, sh);
  }
  const int nb = gridDim.x * gridDim.y;
  if (threadIdx.x == 0) {
    const int b = blockIdx.y * gridDim.x + bx;
    partials[b] = md;
    partials[nb + b] = mf;
    partials[2 * nb + b] = mv;
    if (want_div) partials[3 * nb + b] = mdiv;  // per column: [c * gx + bx]
  }
}

inline dim3 col_grid(int64_t n, int ncols) { return dim3((unsigned)reduce_blocks(n), (unsigned)ncols); }

}  // namespace

int defect_blocks(int64_t size, int ncols) { return reduce_blocks(size) * ncols; }

// Two elements per thread up to 4096 blocks per column: enough waves in flight to cover
// HBM latency in the grid-stride loops (8 per thread at most 1024 blocks left these
// HBM-bound kernels latency-bound at ~2 TB/s).
int reduce_blocks(int64_t n) {
  int64_t b = (n + (int64_t)kThreads * 2 - 1) / ((int64_t)kThreads * 2);
  if (b < 1) b = 1;
  if (b > 4096) b = 4096;
  return (int)b;
}

void divergence(int in_type, const void* vel, const VelLayout& L, const GridDims& g, double* out,
                int64_t out_stride, double sign, int ncols, int reduce, double* partials, const CgState* st,
                hipStream_t s, const double* add, Fin fin) {
  const dim3 grid = col_grid((int64_t)g.nx * g.ny, ncols);
#define MACS_DIV(T)                                                                                          \
  switch (reduce) {                                                                                          \
    case RED_SUM: k_div<T, RED_SUM><<<grid, kThreads, 0, s>>>((const T*)vel, L, g, out, out_stride, sign,   \
                                                                partials, st, add); break;                   \
    case RED_MAXABS: k_div<T, RED_MAXABS><<<grid, kThreads, 0, s>>>((const T*)vel, L, g, out, out_stride,   \
                                                                      sign, partials, st, add); break;       \
    default: k_div<T, RED_NONE><<<grid, kThreads, 0, s>>>((const T*)vel, L, g, out, out_stride, sign,       \
                                                            partials, st, add); break;                       \
  }
  if (in_type == DT_F64) { MACS_DIV(double) }
  else if (in_type == DT_F32) { MACS_DIV(float) }
  else { MACS_DIV(__bf16) }
#undef MACS_DIV
  if (fin.out != nullptr && reduce != RED_NONE) finalize(partials, grid.x, ncols, reduce, fin.out, s);
}

void divergence_transpose(const double* p, int64_t p_stride, const double* mean_sums, const double* add,
                          int64_t add_stride, int out_type, void* out, const VelLayout& L, const GridDims& g,
                          int ncols, const CgState* st, hipStream_t s) {
  const int64_t cells = (int64_t)g.nx * g.ny;
  const dim3 grid((unsigned)((cells + kThreads - 1) / kThreads), (unsigned)ncols);
  if (out_type == DT_F64)
    k_dt<double><<<grid, kThreads, 0, s>>>(p, p_stride, mean_sums, add, add_stride, (double*)out, L, g, st);
  else if (out_type == DT_F32)
    k_dt<float><<<grid, kThreads, 0, s>>>(p, p_stride, mean_sums, add, add_stride, (float*)out, L, g, st);
  else
    k_dt<__bf16><<<grid, kThreads, 0, s>>>(p, p_stride, mean_sums, add, add_stride, (__bf16*)out, L, g, st);
}

void reduce(const double* x, const double* y, int64_t n, int64_t stride, int ncols, int mode, double* partials,
            const CgState* st, hipStream_t s, Fin fin) {
  const dim3 grid = col_grid(n, ncols);
  if (mode == RED_MAXABS) k_reduce<RED_MAXABS, false><<<grid, kThreads, 0, s>>>(x, y, n, stride, partials, st);
  else if (y != nullptr) k_reduce<RED_SUM, true><<<grid, kThreads, 0, s>>>(x, y, n, stride, partials, st);
  else k_reduce<RED_SUM, false><<<grid, kThreads, 0, s>>>(x, y, n, stride, partials, st);
  if (fin.out != nullptr) finalize(partials, grid.x, ncols, mode, fin.out, s);
}

void finalize(const double* partials, int nblk, int ncols, int mode, double* out, hipStream_t s) {
  if (mode == RED_MAXABS) k_finalize<RED_MAXABS><<<ncols, kThreads, 0, s>>>(partials, nblk, out);
  else k_finalize<RED_SUM><<<ncols, kThreads, 0, s>>>(partials, nblk, out);
}

void sub_mean_dot(double* x, int64_t n, int64_t stride, int ncols, const double* sums, const double* y,
                  double* partials, const CgState* st, hipStream_t s, Fin fin) {
  const dim3 grid = col_grid(n, ncols);
  if (y != nullptr) {
    k_sub_mean_dot<true><<<grid, kThreads, 0, s>>>(x, n, stride, sums, y, partials, st);
    if (fin.out != nullptr) finalize(partials, grid.x, ncols, RED_SUM, fin.out, s);
  } else {
    k_sub_mean_dot<false><<<grid, kThreads, 0, s>>>(x, n, stride, sums, y, partials, st);
  }
}

void sub_cols(const double* a, const double* b, double* out, int64_t n, int64_t stride, int ncols,
              const CgState* st, hipStream_t s) {
  k_sub_cols<<<col_grid(n, ncols), kThreads, 0, s>>>(a, b, out, n, stride, st);
}

void theta_start(double* x, double* r, const double* b, const double* q, const double* gb, const double* gq,
                 int64_t n, int64_t stride, int ncols, const CgState* st, hipStream_t s) {
  k_theta_start<<<col_grid(n, ncols), kThreads, 0, s>>>(x, r, b, q, gb, gq, n, stride, st);
}

void galerkin2_start(double* x, double* r, const double* g, const double* c, const double* b, const double* bg,
                     const double* q, const double* dots, int64_t n, int ncols, const CgState* st, hipStream_t s) {
  k_galerkin2<<<col_grid(n, ncols), kThreads, 0, s>>>(x, r, g, c, b, bg, q, dots, ncols, n, st);
}

void fill(double* x, int64_t n, double v, hipStream_t s) {
  k_fill<<<reduce_blocks(n), kThreads, 0, s>>>(x, n, v);
}

void finish_cols(const double* src, int64_t n, int ncols, const double* sums, double* d0, int64_t s0, double* d1,
                 int64_t s1, double* d2, int64_t s2, hipStream_t s) {
  k_finish_cols<<<col_grid(n, ncols), kThreads, 0, s>>>(src, n, sums, d0, s0, d1, s1, d2, s2);
}
void copy_cols(const double* src, int64_t ss, double* dst, int64_t ds, int64_t n, int ncols, hipStream_t s) {
  k_copy_cols<<<col_grid(n, ncols), kThreads, 0, s>>>(src, ss, dst, ds, n);
}

void cg_init(CgState* st, const double* bb, const double* bmax, double rtol, int ncols, hipStream_t s,
             double* rr_out, double mtol) {
  k_cg_init<<<(ncols + 63) / 64, 64, 0, s>>>(st, bb, bmax, rtol, ncols, rr_out, mtol);
}
void sweep_start(double* b, double* r, double* x, int64_t n, int ncols, const double* sums, double* partials,
                 double* bb, double* bmax, hipStream_t s, float* r32) {
  const dim3 grid = col_grid(n, ncols);
  k_sweep_start<<<grid, kThreads, 0, s>>>(b, r, x, n, sums, partials, r32);
  k_finalize2<RED_SUM, RED_MAXABS><<<2 * ncols, kThreads, 0, s>>>(partials, (int)grid.x, ncols, bb,
                                                                  partials + (int64_t)ncols * grid.x,
                                                                  (int)grid.x, bmax);
}
void sweep_start_cg(double* b, double* r, double* x, int64_t n, int ncols, const double* sums, double* partials,
                    double* bb, double* bmax, CgState* st, double rtol, double* rr_out, double mtol, int* flags,
                    hipStream_t s, float* r32) {
  const dim3 grid = col_grid(n, ncols);
  k_sweep_start<<<grid, kThreads, 0, s>>>(b, r, x, n, sums, partials, r32);
  k_sweep_cg_start<<<ncols, kThreads, 0, s>>>(partials, (int)grid.x, ncols, bb, bmax, st, rtol, rr_out, mtol, flags);
}
void cg_check(CgState* st, const double* rr, int* flags, int ncols, hipStream_t s) {
  k_cg_check<<<(ncols + 63) / 64, 64, 0, s>>>(st, rr, flags, ncols);
}
void cg_check_post(CgState* st, double* rr, int* flags, int ncols, Mailbox* mb, int seq, hipStream_t s,
                   const double* rr_partials, int nblk) {
  k_cg_check_post<<<1, kThreads, 0, s>>>(st, rr, rr_partials, nblk, flags, ncols, mb, seq);
}
void mixed_rho(const double* z, const double* r, int64_t n, int ncols, CgState* st, double* partials,
               double* zmean, hipStream_t s) {
  const dim3 grid = col_grid(n, ncols);
  k_sums3<<<grid, kThreads, 0, s>>>(z, r, n, n, st, partials);
  k_cg_rho_mean<<<ncols, kThreads, 0, s>>>(st, partials, (int)grid.x, ncols, (double)n, zmean);
}
void mixed_rho_partials(CgState* st, const double* partials, int gx, int64_t n, int ncols, double* zmean,
                        hipStream_t s) {
  k_cg_rho_mean<<<ncols, kThreads, 0, s>>>(st, partials, gx, ncols, (double)n, zmean);
}
void mixed_update_p(double* p, const double* z, const double* zmean, int64_t n, int ncols, const CgState* st,
                    hipStream_t s) {
  k_update_p_mean<<<col_grid(n, ncols), kThreads, 0, s>>>(p, z, zmean, n, n, st);
}
void mixed_alpha(const double* partials, int64_t n, int ncols, CgState* st, double* qmean, hipStream_t s) {
  k_cg_alpha_mean<<<ncols, kThreads, 0, s>>>(st, partials, reduce_blocks(n), ncols, (double)n, qmean);
}
void mixed_update_xr(double* x, double* r, const double* p, const double* q, const double* qmean, int64_t n,
                     int ncols, const CgState* st, double* partials, hipStream_t s, float* r32) {
  k_update_xr_mean<<<col_grid(n, ncols), kThreads, 0, s>>>(x, r, p, q, qmean, n, n, st, partials, r32);
}
void divergence_sums3(int in_type, const void* vel, const VelLayout& L, const GridDims& g, double* out,
                      int64_t out_stride, double sign, const double* y, int ncols, double* partials,
                      const CgState* st, hipStream_t s) {
  const dim3 grid = col_grid((int64_t)g.nx * g.ny, ncols);
  if (in_type == DT_F64)
    k_div_sums3<double><<<grid, kThreads, 0, s>>>((const double*)vel, L, g, out, out_stride, sign, y, partials, st);
  else if (in_type == DT_F32)
    k_div_sums3<float><<<grid, kThreads, 0, s>>>((const float*)vel, L, g, out, out_stride, sign, y, partials, st);
  else
    k_div_sums3<__bf16><<<grid, kThreads, 0, s>>>((const __bf16*)vel, L, g, out, out_stride, sign, y, partials, st);
}
void cg_rho_fin(CgState* st, const double* partials, int nblk, double* out, int ncols, hipStream_t s) {
  k_cg_rho_fin<<<ncols, kThreads, 0, s>>>(st, partials, nblk, out);
}
void cg_alpha_fin(CgState* st, const double* partials, int nblk, double* out, int ncols, hipStream_t s) {
  k_cg_alpha_fin<<<ncols, kThreads, 0, s>>>(st, partials, nblk, out);
}
void post_values(const double* a, int na, const double* b, int nb, Mailbox* mb, int seq, hipStream_t s) {
  k_post_values<<<1, 128, 0, s>>>(a, na, b, nb, mb, seq);
}
void extrapolate(const double* const* src, const double* w, int count, double* out, int64_t n, hipStream_t s) {
  Extrapolation e{};
  e.count = count;
  for (int k = 0; k < count; ++k) {
    e.src[k] = src[k];
    e.w[k] = w[k];
  }
  bool aligned = (reinterpret_cast<uintptr_t>(out) & 15) == 0;
  for (int k = 0; k < count; ++k) aligned = aligned && (reinterpret_cast<uintptr_t>(src[k]) & 15) == 0;
  const int64_t n2 = aligned ? n / 2 : 0;
  if (n2 > 0) {
    k_extrapolate2<<<(int)((n2 + kThreads - 1) / kThreads), kThreads, 0, s>>>(e, reinterpret_cast<double2*>(out), n2);
    if (n % 2 == 0) return;
    // The odd last element.
    for (int k = 0; k < count; ++k) e.src[k] += 2 * n2;
    k_extrapolate<<<1, kThreads, 0, s>>>(e, out + 2 * n2, 1);
    return;
  }
  k_extrapolate<<<reduce_blocks(n), kThreads, 0, s>>>(e, out, n);
}
void cg_rho(CgState* st, const double* rz, int ncols, hipStream_t s) { k_cg_rho<<<(ncols + 63) / 64, 64, 0, s>>>(st, rz, ncols); }
void cg_alpha(CgState* st, const double* pq, int ncols, hipStream_t s) {
  k_cg_alpha<<<(ncols + 63) / 64, 64, 0, s>>>(st, pq, ncols);
}
void cg_update_p(double* p, const double* z, int64_t n, int64_t stride, int ncols, const CgState* st,
                 hipStream_t s) {
  k_update_p<<<col_grid(n, ncols), kThreads, 0, s>>>(p, z, n, stride, st);
}
void cg_update_xr(double* x, double* r, const double* p, const double* q, int64_t n, int64_t stride, int ncols,
                  const CgState* st, double* partials, hipStream_t s, Fin fin) {
  const dim3 grid = col_grid(n, ncols);
  k_update_xr<<<grid, kThreads, 0, s>>>(x, r, p, q, n, stride, st, partials);
  if (fin.out != nullptr) finalize(partials, grid.x, ncols, RED_SUM, fin.out, s);
}

void defect_max(const double* vel, const double* f, const double* p, const GridDims& g, int ncols, double mass,
                double visc, double* partials3, hipStream_t s, void* residual_out, Fin fin, double* div_out,
                double div_sign, Fin div_fin, int residual_type, const VelLayout* residual_layout) {
  const int64_t nu = (int64_t)(g.nx - 1) * g.ny, size = nu + (int64_t)g.nx * (g.ny - 1);
  const dim3 grid = col_grid(size, ncols);
  const int want_div = div_out != nullptr || div_fin.out != nullptr;
  VelLayout Lr{};  // packed unless given
  Lr.ldu = g.nx - 1;
  Lr.ldv = g.nx;
  Lr.voff = nu;
  Lr.stride = size;
  if (residual_layout != nullptr) Lr = *residual_layout;
  if (residual_type == DT_F32)
    k_defect<float><<<grid, kThreads, 0, s>>>(vel, f, p, g, ncols, mass, visc, partials3, (float*)residual_out, Lr,
                                              want_div, div_out, div_sign);
  else if (residual_type == DT_BF16)
    k_defect<__bf16><<<grid, kThreads, 0, s>>>(vel, f, p, g, ncols, mass, visc, partials3, (__bf16*)residual_out,
                                               Lr, want_div, div_out, div_sign);
  else
    k_defect<double><<<grid, kThreads, 0, s>>>(vel, f, p, g, ncols, mass, visc, partials3, (double*)residual_out,
                                               Lr, want_div, div_out, div_sign);
  // The three maxima (defect, f, u) at partials3 + q * nb, one block each; the
  // divergence's per-column maxima at partials3 + 3 nb + c * grid.x.
  const int nb = defect_blocks(size, ncols);
  if (fin.out != nullptr && div_fin.out != nullptr)
    k_finalize2<RED_MAXABS, RED_MAXABS><<<3 + ncols, kThreads, 0, s>>>(partials3, nb, 3, fin.out,
                                                                       partials3 + 3 * nb, (int)grid.x,
                                                                       div_fin.out);
  else if (fin.out != nullptr)
    finalize(partials3, nb, 3, RED_MAXABS, fin.out, s);
  else if (div_fin.out != nullptr)
    finalize(partials3 + 3 * nb, grid.x, ncols, RED_MAXABS, div_fin.out, s);
}

__global__ void k_add_cols(double* __restrict__ x, const double* __restrict__ y, int64_t total) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < total; i += (int64_t)gridDim.x * kThreads)
    x[i] = x[i] + y[i];
}

__global__ __launch_bounds__(kThreads) void k_accum_velocity(float* __restrict__ du, const float* __restrict__ w,
                                                             int64_t n, int64_t stride, const CgState* st,
                                                             const double* gb, const double* gq) {
  const int c = blockIdx.y;
  if (!st[c].active) return;
  double a;
  if (gb != nullptr) a = gq[c] > 0.0 ? gb[c] / gq[c] : 0.0;
  else if (gq != nullptr) a = 1.0;
  else a = st[c].alpha;
  const float af = (float)a;
  const int64_t off = (int64_t)c * stride;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    du[off + i] = fmaf(af, w[off + i], du[off + i]);
}

void accum_velocity(float* du, const float* w, int64_t n, int64_t stride, int ncols, const CgState* st,
                    const double* gb, const double* gq, hipStream_t s) {
  k_accum_velocity<<<col_grid(n, ncols), kThreads, 0, s>>>(du, w, n, stride, st, gb, gq);
}

__global__ void k_add_f32(double* __restrict__ x, const float* __restrict__ y, int64_t total) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < total; i += (int64_t)gridDim.x * kThreads)
    x[i] = x[i] + (double)y[i];
}

void add_f32(double* x, const float* y, int64_t n, hipStream_t s) {
  k_add_f32<<<reduce_blocks(n), kThreads, 0, s>>>(x, y, n);
}

__global__ __launch_bounds__(kThreads) void k_add_pair(double* __restrict__ x, const double* __restrict__ y,
                                                      int64_t nx, double* __restrict__ u,
                                                      const float* __restrict__ w, int64_t nu) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < nx) {
    x[i] = x[i] + y[i];
  } else if (i < nx + nu) {
    const int64_t j = i - nx;
    u[j] = u[j] + (double)w[j];
  }
}

void add_pair(double* x, const double* y, int64_t nx, double* u, const float* w, int64_t nu, hipStream_t s) {
  const int64_t n = nx + nu;
  if (n > 0) k_add_pair<<<(unsigned)((n + kThreads - 1) / kThreads), kThreads, 0, s>>>(x, y, nx, u, w, nu);
}

void add_cols(double* x, const double* y, int64_t n, int ncols, hipStream_t s) {
  k_add_cols<<<reduce_blocks(n * ncols), kThreads, 0, s>>>(x, y, n * ncols);
}

}  // namespace macs
