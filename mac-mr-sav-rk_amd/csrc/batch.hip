// Batched MAC stage solve: many independent cavities of uneven sizes advanced together.
//
// Each cavity runs the reference's TensorStokes.solve (solver/mac/tensor_stokes.py:245-364)
// on its own grid, viscosity and columns; nothing couples them.  A (cavity, column) pair is
// a "segment".  Every transform pass is ONE grouped GEMM launch over a device table holding
// every segment's u/v (or pressure) block, and every vector step is one launch whose grid's
// y dimension walks the segments.  A segment whose Schur-CG has converged is skipped on the
// device (its GEMM tiles and vector blocks return at once), so the host only reads the
// per-segment active flags once per iteration.  One CG iteration is a fixed launch sequence,
// captured once into a hipGraph and replayed.  In the mixed form a whole stage can run as one
// captured graph instead (the stage launch shape): the outer test, the inner solve's set-up,
// the CG iterations and the update are switched on and off by device flags, and the host
// reads one control block per graph launch.
//
// Per-segment reductions use the same block partition and order as the single-cavity
// solver (reduce_blocks of the segment's own length), so every segment's iterates, CG
// iteration count and results are bitwise those of the same cavity solved alone.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "host_common.h"
#include "internal.h"
#include "macstokes.h"

namespace macs {
namespace {

struct Seg {
  int nx, ny;
  double hx, hy;
  int nu, size, np;   // u faces, all faces, cells
  int nbv, nbp;       // reduction blocks over faces / cells
  int cav;
  int64_t voff, poff; // first element of the segment's velocity / pressure column
  int64_t t32u, t32v; // mixed form: the u / v blocks' FP32 transform input, rows padded to 8
};

// Index of packed face f of segment S in the FP32 transform-input layout (the u block's
// rows padded to round8(nx - 1), the v block's to round8(nx)).
__device__ inline int64_t pad32_index(const Seg& S, int f) {
  if (f < S.nu) {
    const int w = S.nx - 1, j = f / w;
    return S.t32u + (int64_t)j * ((w + 7) & ~7) + (f - j * w);
  }
  const int e = f - S.nu, j = e / S.nx;
  return S.t32v + (int64_t)j * ((S.nx + 7) & ~7) + (e - j * S.nx);
}

enum Len : int { LEN_FACES = 0, LEN_CELLS = 1, ITEMS_DT = 2 };

// A launch whose `run` flag (device memory, may be null) reads 0 does nothing: the stage
// graph switches its outer-loop steps with it.
__device__ inline bool halted(const int* run) { return run != nullptr && *run == 0; }

// Work items of a launch over the segments: item w is block `y` of segment `x` (the block
// index within the segment's own partition, so reductions keep the single-cavity order).
// A launch walks its items in a grid-stride loop, so its grid stays bounded whatever the
// batch, and a launch whose segments are inactive costs one flag load per item.
struct Items {
  const int2* it;
  int n;
};

// out = add + sign * D vel per segment (add optional); partial sums / max of the result.
template <int MODE>
__global__ __launch_bounds__(kThreads) void k_bdiv(const Items W, const double* __restrict__ vel,
                                                   double* __restrict__ out, const Seg* segs, double sign,
                                                   double* partials, int pstride, const CgState* st,
                                                   const double* __restrict__ add, const int* run) {
  if (halted(run)) return;
  __shared__ double sh[4];
  for (int w = blockIdx.x; w < W.n; w += gridDim.x) {
    const int s = W.it[w].x, blk = W.it[w].y;
    if (st != nullptr && !st[s].active) continue;
    const Seg S = segs[s];
    const double* u = vel + S.voff;
    const double* v = u + S.nu;
    double* o = out + S.poff;
    const int nx = S.nx, ny = S.ny;
    const SmallDiv dnx(nx);
    double accum = 0.0;
    for (int cell = blk * kThreads + threadIdx.x; cell < S.np; cell += S.nbp * kThreads) {
      const int j = dnx.quot(cell), i = cell - j * nx;
      double value = 0.0;
      if (i < nx - 1) value += u[j * (nx - 1) + i] / S.hx;
      if (i > 0) value -= u[j * (nx - 1) + i - 1] / S.hx;
      if (j < ny - 1) value += v[j * nx + i] / S.hy;
      if (j > 0) value -= v[(j - 1) * nx + i] / S.hy;
      value = sign * value;
      if (add != nullptr) value = add[S.poff + cell] + value;
      o[cell] = value;
      if (MODE == RED_SUM) accum += value;
      if (MODE == RED_MAXABS) accum = fmax(accum, absn(value));
    }
    if (MODE != RED_NONE) {
      accum = block_reduce<MODE>(accum, sh);
      if (threadIdx.x == 0) partials[(int64_t)s * pstride + blk] = accum;
      __syncthreads();  // sh is reused by the next item
    }
  }
}

// out = add + D^T p per segment (packed velocity layout; add optional), or (out32) D^T p
// in the FP32 transform-input layout.  One thread per face.
__global__ __launch_bounds__(kThreads) void k_bdt(const Items W, const double* __restrict__ p,
                                                  const double* __restrict__ add, double* __restrict__ out,
                                                  const Seg* segs, const CgState* st, float* __restrict__ out32) {
  for (int w = blockIdx.x; w < W.n; w += gridDim.x) {
    const int s = W.it[w].x, blk = W.it[w].y;
    if (st != nullptr && !st[s].active) continue;
    const Seg S = segs[s];
    const int f = blk * kThreads + threadIdx.x;
    if (f >= S.size) continue;
    const double* pc = p + S.poff;
    const int nx = S.nx;
    double value;
    if (f < S.nu) {
      const int j = SmallDiv(nx - 1).quot(f), i = f - j * (nx - 1);
      const int left = j * nx + i;
      value = (pc[left] - pc[left + 1]) / S.hx;
    } else {
      const int e = f - S.nu;
      const int j = SmallDiv(nx).quot(e), i = e - j * nx;
      const int lower = j * nx + i;
      value = (pc[lower] - pc[lower + nx]) / S.hy;
    }
    if (add != nullptr) value = add[S.voff + f] + value;
    if (out32 != nullptr) out32[pad32_index(S, f)] = (float)value;
    else out[S.voff + f] = value;
  }
}

template <int MODE, bool DOT>
__global__ __launch_bounds__(kThreads) void k_breduce(const Items W, const double* __restrict__ x,
                                                      const double* __restrict__ y, const Seg* segs, int len,
                                                      double* partials, int pstride, const CgState* st) {
  __shared__ double sh[4];
  for (int w = blockIdx.x; w < W.n; w += gridDim.x) {
    const int s = W.it[w].x, blk = W.it[w].y;
    if (st != nullptr && !st[s].active) continue;
    const Seg S = segs[s];
    const int nb = len == LEN_FACES ? S.nbv : S.nbp;
    const int64_t off = len == LEN_FACES ? S.voff : S.poff;
    const int n = len == LEN_FACES ? S.size : S.np;
    double accum = 0.0;
    for (int i = blk * kThreads + threadIdx.x; i < n; i += nb * kThreads) {
      if (MODE == RED_MAXABS) accum = fmax(accum, absn(x[off + i]));
      else accum += DOT ? x[off + i] * y[off + i] : x[off + i];
    }
    accum = block_reduce<MODE>(accum, sh);
    if (threadIdx.x == 0) partials[(int64_t)s * pstride + blk] = accum;
    __syncthreads();
  }
}

// out[s] = the segment's partials reduced in the order of the single-cavity finalize.
template <int MODE>
__global__ __launch_bounds__(kThreads) void k_bfinalize(const double* partials, int pstride, const Seg* segs,
                                                        int len, double* out, const int* run) {
  if (halted(run)) return;
  const int s = blockIdx.x;
  const Seg S = segs[s];
  const int nb = len == LEN_FACES ? S.nbv : S.nbp;
  __shared__ double sh[4];
  double accum = 0.0;
  for (int i = threadIdx.x; i < nb; i += kThreads) {
    const double v = partials[(int64_t)s * pstride + i];
    accum = MODE == RED_MAXABS ? fmax(accum, v) : accum + v;
  }
  accum = block_reduce<MODE>(accum, sh);
  if (threadIdx.x == 0) out[s] = accum;
}

// x -= sums[s]/np on the cells; optional partial sums of y*x afterwards.
template <bool DOT>
__global__ __launch_bounds__(kThreads) void k_bsub_mean_dot(const Items W, double* x, const double* sums,
                                                            const double* y, const Seg* segs, double* partials,
                                                            int pstride, const CgState* st) {
  __shared__ double sh[4];
  for (int w = blockIdx.x; w < W.n; w += gridDim.x) {
    const int s = W.it[w].x, blk = W.it[w].y;
    if (st != nullptr && !st[s].active) continue;
    const Seg S = segs[s];
    double* xc = x + S.poff;
    const double* yc = DOT ? y + S.poff : nullptr;
    const double mean = sums[s] / (double)S.np;
    double accum = 0.0;
    for (int i = blk * kThreads + threadIdx.x; i < S.np; i += S.nbp * kThreads) {
      const double v = xc[i] - mean;
      xc[i] = v;
      if (DOT) accum += yc[i] * v;
    }
    if (DOT) {
      accum = block_reduce<RED_SUM>(accum, sh);
      if (threadIdx.x == 0) partials[(int64_t)s * pstride + blk] = accum;
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_bupdate_p(const Items W, double* __restrict__ p,
                                                        const double* __restrict__ z, const Seg* segs,
                                                        const CgState* st) {
  for (int w = blockIdx.x; w < W.n; w += gridDim.x) {
    const int s = W.it[w].x, blk = W.it[w].y;
    if (!st[s].active) continue;
    const Seg S = segs[s];
    const bool first = st[s].iters == 0;
    const double beta = st[s].beta;
    double* pc = p + S.poff;
    const double* zc = z + S.poff;
    for (int i = blk * kThreads + threadIdx.x; i < S.np; i += S.nbp * kThreads)
      pc[i] = first ? zc[i] : __dadd_rn(__dmul_rn(pc[i], beta), zc[i]);
  }
}

__global__ __launch_bounds__(kThreads) void k_bupdate_xr(const Items W, double* __restrict__ x,
                                                         double* __restrict__ r, const double* __restrict__ p,
                                                         const double* __restrict__ q, const Seg* segs,
                                                         const CgState* st, double* partials, int pstride) {
  __shared__ double sh[4];
  for (int w = blockIdx.x; w < W.n; w += gridDim.x) {
    const int s = W.it[w].x, blk = W.it[w].y;
    if (!st[s].active) continue;
    const Seg S = segs[s];
    const double alpha = st[s].alpha;
    const int64_t off = S.poff;
    double accum = 0.0;
    for (int i = blk * kThreads + threadIdx.x; i < S.np; i += S.nbp * kThreads) {
      x[off + i] = __dadd_rn(x[off + i], __dmul_rn(alpha, p[off + i]));
      const double rv = __dsub_rn(r[off + i], __dmul_rn(alpha, q[off + i]));
      r[off + i] = rv;
      accum += rv * rv;
    }
    accum = block_reduce<RED_SUM>(accum, sh);
    if (threadIdx.x == 0) partials[(int64_t)s * pstride + blk] = accum;
    __syncthreads();
  }
}

// Per segment: max|mass u + visc K u - D^T p - f|, max|f|, max|u| (partials at q * nseg * pstride);
// WRITE: the momentum residual r_u = f - (mass u + visc K u - D^T p) as well, in FP32 in the
// transform-input layout (the split transform keeps 16 of its bits).
template <bool WRITE>
__global__ __launch_bounds__(kThreads) void k_bdefect(const Items W, const double* __restrict__ vel,
                                                      const double* __restrict__ f, const double* __restrict__ p,
                                                      const Seg* segs, const double* __restrict__ mass_visc,
                                                      double* partials, int pstride, int nseg,
                                                      float* __restrict__ ru32, const int* run) {
  if (halted(run)) return;
  __shared__ double sh[4];
  for (int item = blockIdx.x; item < W.n; item += gridDim.x) {
  const int s = W.it[item].x, blk = W.it[item].y;
  const Seg S = segs[s];
  const int nx = S.nx, ny = S.ny, nu = S.nu;
  const double hx2 = S.hx * S.hx, hy2 = S.hy * S.hy;
  const double ox = -1.0 / hx2, oy = -1.0 / hy2;
  const double mass = mass_visc[2 * s], visc = mass_visc[2 * s + 1];
  const double* vc = vel + S.voff;
  const double* fc = f + S.voff;
  const double* pc = p + S.poff;
  const SmallDiv dw(nx - 1), dnx(nx);
  double md = 0.0, mf = 0.0, mv = 0.0;
  for (int fidx = blk * kThreads + threadIdx.x; fidx < S.size; fidx += S.nbv * kThreads) {
    FaceRow k;
    double lift;
    if (fidx < nu) {
      const int w = nx - 1;
      const int j = dw.quot(fidx), i = fidx - j * w;
      const double* u = vc;
      const double diag = 2.0 / hx2 + ((j == 0 || j == ny - 1) ? 3.0 : 2.0) / hy2;
      if (j > 0) k.add(oy, u[fidx - w]);
      if (i > 0) k.add(ox, u[fidx - 1]);
      k.add(diag, u[fidx]);
      if (i < w - 1) k.add(ox, u[fidx + 1]);
      if (j < ny - 1) k.add(oy, u[fidx + w]);
      const int left = j * nx + i;
      lift = (pc[left] - pc[left + 1]) / S.hx;
    } else {
      const int e = fidx - nu;
      const int j = dnx.quot(e), i = e - j * nx;
      const double* v = vc + nu;
      const double diag = ((i == 0 || i == nx - 1) ? 3.0 : 2.0) / hx2 + 2.0 / hy2;
      if (j > 0) k.add(oy, v[e - nx]);
      if (i > 0) k.add(ox, v[e - 1]);
      k.add(diag, v[e]);
      if (i < nx - 1) k.add(ox, v[e + 1]);
      if (j < ny - 2) k.add(oy, v[e + nx]);
      const int lower = j * nx + i;
      lift = (pc[lower] - pc[lower + nx]) / S.hy;
    }
    const double fv = fc[fidx];
    const double vv = vc[fidx];
    const double d = face_defect(mass, vv, visc, k.kv, lift, fv);
    if (WRITE) ru32[pad32_index(S, fidx)] = (float)(-d);
    md = fmax(md, absn(d));
    mf = fmax(mf, absn(fv));
    mv = fmax(mv, absn(vv));
  }
  md = block_reduce<RED_MAXABS>(md, sh);
  __syncthreads();
  mf = block_reduce<RED_MAXABS>(mf, sh);
  __syncthreads();
  mv = block_reduce<RED_MAXABS>(mv, sh);
  if (threadIdx.x == 0) {
    const int64_t set = (int64_t)nseg * pstride, b = (int64_t)s * pstride + blk;
    partials[b] = md;
    partials[set + b] = mf;
    partials[2 * set + b] = mv;
  }
  __syncthreads();
  }
}

// dst = src over n elements (kernel form, so it can sit inside a captured graph).
__global__ void k_bcopy(const double* __restrict__ src, double* __restrict__ dst, int64_t n, const int* run) {
  if (halted(run)) return;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    dst[i] = src[i];
}
__global__ void k_bfill(double* __restrict__ dst, int64_t n, double v, const int* run) {
  if (halted(run)) return;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
    dst[i] = v;
}

// ---- mixed form: FP64 defect correction around bf16x3-transform inner solves
// Per segment, its cavity's outer state from the maxima of the cavity's columns (chk:
// four sets of nseg values, max|defect|, max|f|, max|u|, max|D u|): finished when
// max|defect| / (1 + max|f|) <= tol and max|D u| <= tol, else active.
__global__ void k_bouter(const double* __restrict__ chk, int nseg, int ncols, double tol, int* __restrict__ active) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  const int first = (s / ncols) * ncols;
  double d = 0.0, f = 0.0, dv = 0.0;
  for (int c = first; c < first + ncols; ++c) {
    d = fmax(d, chk[c]);
    f = fmax(f, chk[nseg + c]);
    dv = fmax(dv, chk[3 * nseg + c]);
  }
  active[s] = (d / (1.0 + f) <= tol && dv <= tol) ? 0 : 1;
}

// Segment CG states against the outer flags.  GATE_SET: the state's active flag becomes
// the cavity's (an inner solve's first launches skip finished cavities); GATE_MASK: after
// the CG set-up, finished cavities stay inactive; GATE_POST: after the inner CG, its
// iterations are added to the segment's total and the flag returns to the cavity's.
enum Gate : int { GATE_SET = 0, GATE_MASK = 1, GATE_POST = 2 };
__global__ void k_bgate(CgState* st, const int* __restrict__ active, int nseg, int mode, int* iters_total,
                        const int* run) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg || halted(run)) return;
  if (mode == GATE_MASK) {
    if (!active[s]) {
      st[s].active = 0;
      st[s].converged = 1;
    }
    return;
  }
  if (mode == GATE_POST) iters_total[s] += st[s].iters;
  st[s].active = active[s];
}

// du += alpha L~^-1 D^T p (vout) on the active segments: the inner CG builds the velocity
// correction du = W + sum_k alpha_k L~^-1 D^T p_k = L~^-1 (r_u + D^T x) as it goes.
__global__ __launch_bounds__(kThreads) void k_bacc_du(const Items W, double* __restrict__ du,
                                                      const double* __restrict__ vout, const Seg* segs,
                                                      const CgState* st) {
  for (int w = blockIdx.x; w < W.n; w += gridDim.x) {
    const int s = W.it[w].x, blk = W.it[w].y;
    if (!st[s].active) continue;
    const Seg S = segs[s];
    const double alpha = st[s].alpha;
    for (int i = blk * kThreads + threadIdx.x; i < S.size; i += S.nbv * kThreads)
      du[S.voff + i] = fma(alpha, vout[S.voff + i], du[S.voff + i]);
  }
}

// u += du, p += dp on the segments whose state is active.
__global__ __launch_bounds__(kThreads) void k_baccum(const Items W, double* __restrict__ u,
                                                     const double* __restrict__ du, double* __restrict__ p,
                                                     const double* __restrict__ dp, const Seg* segs,
                                                     const CgState* st) {
  for (int w = blockIdx.x; w < W.n; w += gridDim.x) {
    const int s = W.it[w].x, blk = W.it[w].y;
    if (!st[s].active) continue;
    const Seg S = segs[s];
    for (int i = blk * kThreads + threadIdx.x; i < S.size; i += S.nbv * kThreads) u[S.voff + i] += du[S.voff + i];
    for (int i = blk * kThreads + threadIdx.x; i < S.np; i += S.nbv * kThreads) p[S.poff + i] += dp[S.poff + i];
  }
}

// ---- stage launch shape (mixed form): a whole stage as one captured graph
// The host-driven sweep loop's decisions, made on the device.  A graph round runs the outer
// test, the inner solve's set-up, a fixed number of CG iterations and the fold; the flags
// below switch each part on for the segments and rounds where the host-driven loop would
// run it, so every segment sees the same operations in the same order.
struct StageCtl {
  int halt;          // the stage is finished: no cavity above the bar, or the sweep cap
  int in_cg;         // the running inner solve's CG is iterating
  int fold_pending;  // the running inner solve still has to be folded into u and p
  int run_outer;     // this round makes the outer test
  int go_setup;      // this round starts an inner solve
  int go_fold;       // this round folds the finished inner solve in
  int sweeps;        // inner solves started
  int cg_iters;      // CG iterations of the running inner solve
  int max_cg;        // the most CG iterations any inner solve of the stage has taken
};

__global__ void k_stage_init(StageCtl* c) { *c = StageCtl{}; }

__global__ void k_stage_round_begin(StageCtl* c) {
  c->run_outer = (!c->halt && !c->in_cg && !c->fold_pending) ? 1 : 0;
  c->go_setup = 0;
  c->go_fold = 0;
}

// k_bouter, and a segment whose own max|defect| or max|D u| is not finite stops (the
// host-driven loop does the same once it has read the maxima).
__global__ void k_bouter_stage(const double* __restrict__ chk, int nseg, int ncols, double tol,
                               int* __restrict__ active, const int* run) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg || halted(run)) return;
  const int first = (s / ncols) * ncols;
  double d = 0.0, f = 0.0, dv = 0.0;
  for (int c = first; c < first + ncols; ++c) {
    d = fmax(d, chk[c]);
    f = fmax(f, chk[nseg + c]);
    dv = fmax(dv, chk[3 * nseg + c]);
  }
  const bool finite = isfinite(chk[s]) && isfinite(chk[3 * nseg + s]);
  active[s] = (d / (1.0 + f) <= tol && dv <= tol) ? 0 : (finite ? 1 : 0);
}

// After the outer test: the stage is finished when no segment is active or max_sweeps
// inner solves have run; otherwise an inner solve starts.
__global__ void k_stage_after_outer(StageCtl* c, const int* __restrict__ active, int nseg, int max_sweeps) {
  const int run = c->run_outer;
  int any = 0;
  if (run)
    for (int s = threadIdx.x; s < nseg; s += blockDim.x) any |= active[s];
  any = __syncthreads_or(any);
  if (threadIdx.x != 0 || !run) return;
  if (!any || c->sweeps >= max_sweeps) {
    c->halt = 1;
  } else {
    c->go_setup = 1;
    c->sweeps += 1;
  }
}

// Segment CG states for the next part of a round.
//   STAGE_SETUP: active = the cavity's outer flag when an inner solve starts, else 0 (the
//     set-up skips every segment).
//   STAGE_CG_START: after a set-up, run = the states' active flags (from the CG set-up
//     and the mask of finished cavities), and the inner CG starts.
//   STAGE_FOLD: once the CG has stopped, the iterations are added to the totals and
//     active = the outer flag (the host loop's GATE_POST); otherwise active = 0 (the fold
//     skips every segment).
enum StageGate : int { STAGE_SETUP = 0, STAGE_CG_START = 1, STAGE_FOLD = 2 };
__global__ void k_stage_gate(StageCtl* c, CgState* st, const int* __restrict__ outer, int* __restrict__ run,
                             int* __restrict__ iters_total, int nseg, int mode) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (mode == STAGE_SETUP) {
    if (s < nseg) st[s].active = c->go_setup ? outer[s] : 0;
    return;
  }
  if (mode == STAGE_CG_START) {
    if (!c->go_setup) return;
    if (s < nseg) run[s] = st[s].active;
    if (s == 0) {
      c->in_cg = 1;
      c->fold_pending = 1;
      c->cg_iters = 0;
    }
    return;
  }
  const bool fold = c->fold_pending && !c->in_cg;
  if (s < nseg) {
    if (fold) {
      iters_total[s] += st[s].iters;
      st[s].active = outer[s];
    } else {
      st[s].active = 0;
    }
  }
  if (s == 0) c->go_fold = fold ? 1 : 0;
}

// cg_init's states where `run` is set.
__global__ void k_stage_cg_init(CgState* st, const double* bb, const double* bmax, double rtol, int n,
                                const int* run) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n || halted(run)) return;
  CgState s{};
  s.bnorm = sqrt(bb[c]);
  s.tol = rtol * s.bnorm;
  s.rr = bb[c];
  s.active = (bmax[c] != 0.0 && s.bnorm != 0.0) ? 1 : 0;
  s.converged = s.active ? 0 : 1;
  s.iters = 0;
  st[c] = s;
}

// The inner CG's check at the top of an iteration (k_cg_check on the segments whose CG
// iterates), with the host loop's cap: after max_iterations iterations the CG stops without
// a check.  When no segment iterates any more, the inner CG has finished.  Outside an inner
// CG every state stays inactive, so the iteration's launches skip every segment.  A round
// ends with one more check (`counting` 0) and no iteration after it, so a CG that fits the
// round finishes there; when it does not, the next round's first check repeats it on the
// same residuals and counts the iteration that follows.
__global__ void k_stage_cg_check(StageCtl* c, CgState* st, const double* __restrict__ rr, int* __restrict__ run,
                                 int nseg, int max_iters, int counting) {
  const int in_cg = c->in_cg;
  const bool going = in_cg && c->cg_iters < max_iters;
  int any = 0;
  for (int s = threadIdx.x; s < nseg; s += blockDim.x) {
    int a = going ? run[s] : 0;
    if (a) {
      st[s].rr = rr[s];
      if (sqrt(rr[s]) < st[s].tol) {
        a = 0;
        st[s].converged = 1;
      }
    }
    st[s].active = a;
    if (in_cg) run[s] = a;
    any |= a;
  }
  any = __syncthreads_or(any);
  if (threadIdx.x == 0 && in_cg) {
    if (!any) {
      c->in_cg = 0;
      c->max_cg = max(c->max_cg, c->cg_iters);
    } else if (counting) {
      c->cg_iters += 1;
    }
  }
}

__global__ void k_stage_round_end(StageCtl* c) {
  if (c->go_fold) c->fold_pending = 0;
}

inline unsigned blocks_for(int64_t n) { return (unsigned)std::min<int64_t>((n + kThreads - 1) / kThreads, 4096); }

struct DevMatrix64 {
  int n = 0, kind = -1;
  double* q = nullptr;
  double* qt = nullptr;
};

// bf16 hi / lo parts of Q and Q^T (n x ld8, zero padded) for the mixed form.
struct DevMatrixB {
  int n = 0, kind = -1, ld8 = 0;
  __bf16 *qb = nullptr, *qb_lo = nullptr, *qtb = nullptr, *qtb_lo = nullptr;
};

struct Cavity {
  int nx, ny;
  double lx, ly, hx, hy;
  int64_t size, np;
  double p_thresh;
  const DevMatrix64 *ux, *uy, *vx, *vy, *px, *py;
  double *lux, *luy, *lvx, *lvy, *lpx, *lpy;
};

class BatchSolver {
 public:
  BatchSolver(int count, const int* nx, const int* ny, const double* lx, const double* ly, int ncols,
              const mac_stokes_opts& o)
      : opts_(o), ncols_(ncols) {
    HIP_CHECK(hipSetDevice(o.device));
    HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    cav_.resize(count);
    for (int i = 0; i < count; ++i) {
      Cavity& C = cav_[i];
      C.nx = nx[i];
      C.ny = ny[i];
      C.lx = lx[i];
      C.ly = ly[i];
      C.hx = lx[i] / nx[i];
      C.hy = ly[i] / ny[i];
      C.size = (int64_t)(C.nx - 1) * C.ny + (int64_t)C.nx * (C.ny - 1);
      C.np = (int64_t)C.nx * C.ny;
      C.ux = matrix(C.nx - 1, KIND_DST1);
      C.uy = matrix(C.ny, KIND_DST2);
      C.vx = matrix(C.nx, KIND_DST2);
      C.vy = matrix(C.ny - 1, KIND_DST1);
      C.px = matrix(C.nx, KIND_DCT2);
      C.py = matrix(C.ny, KIND_DCT2);
      C.lux = upload(eigenvalues_1d(C.nx - 1, C.hx, KIND_DST1));
      C.luy = upload(eigenvalues_1d(C.ny, C.hy, KIND_DST2));
      C.lvx = upload(eigenvalues_1d(C.nx, C.hx, KIND_DST2));
      C.lvy = upload(eigenvalues_1d(C.ny - 1, C.hy, KIND_DST1));
      const std::vector<double> lpx = eigenvalues_1d(C.nx, C.hx, KIND_DCT2);
      const std::vector<double> lpy = eigenvalues_1d(C.ny, C.hy, KIND_DCT2);
      C.lpx = upload(lpx);
      C.lpy = upload(lpy);
      double lmax = 0.0;
      for (double a : lpy)
        for (double b : lpx) lmax = std::max(lmax, a + b);
      C.p_thresh = 100.0 * 2.220446049250313e-16 * lmax;
    }
    // Segments: cavity-major, column-minor; each segment's columns contiguous.
    nseg_ = count * ncols;
    segs_host_.resize(nseg_);
    int64_t voff = 0, poff = 0;
    pstride_ = 1;
    maxfaces_ = 1;
    for (int i = 0; i < count; ++i)
      for (int c = 0; c < ncols; ++c) {
        Seg& S = segs_host_[i * ncols + c];
        const Cavity& C = cav_[i];
        S.nx = C.nx;
        S.ny = C.ny;
        S.hx = C.hx;
        S.hy = C.hy;
        S.nu = (C.nx - 1) * C.ny;
        S.size = (int)C.size;
        S.np = (int)C.np;
        S.nbv = reduce_blocks(C.size);
        S.nbp = reduce_blocks(C.np);
        S.cav = i;
        S.voff = voff;
        S.poff = poff;
        voff += C.size;
        poff += C.np;
        pstride_ = std::max(pstride_, std::max(S.nbv, S.nbp));
        nbv_max_ = std::max(nbv_max_, S.nbv);
        nbp_max_ = std::max(nbp_max_, S.nbp);
        maxfaces_ = std::max<int64_t>(maxfaces_, C.size);
      }
    V_ = voff;
    P_ = poff;
    segs_ = static_cast<Seg*>(upload_raw(segs_host_.data(), sizeof(Seg) * nseg_));
    for (int kind = 0; kind < 3; ++kind) {
      std::vector<int2> it;
      for (int q = 0; q < nseg_; ++q) {
        const Seg& S = segs_host_[q];
        const int nb = kind == LEN_FACES ? S.nbv : kind == LEN_CELLS ? S.nbp : (S.size + kThreads - 1) / kThreads;
        for (int b = 0; b < nb; ++b) it.push_back(make_int2(q, b));
      }
      items_n_[kind] = (int)it.size();
      items_[kind] = static_cast<const int2*>(upload_raw(it.data(), sizeof(int2) * it.size()));
    }
    auto take = [&](size_t bytes) { return dev_alloc(bytes); };
    const size_t vb = sizeof(double) * V_, pb = sizeof(double) * P_;
    io_f_ = (double*)take(vb);
    io_v_ = (double*)take(vb);
    io_p_ = (double*)take(pb);
    vin_ = (double*)take(vb);
    vout_ = (double*)take(vb);
    s1_ = (double*)take(vb);
    s2_ = (double*)take(vb);
    b_ = (double*)take(pb);
    x_ = (double*)take(pb);
    r_ = (double*)take(pb);
    z_ = (double*)take(pb);
    p_ = (double*)take(pb);
    q_ = (double*)take(pb);
    partials_ = (double*)take(sizeof(double) * 3 * (size_t)nseg_ * pstride_);
    sums_ = (double*)take(sizeof(double) * nseg_);
    sums2_ = (double*)take(sizeof(double) * nseg_);
    sums3_ = (double*)take(sizeof(double) * nseg_);
    rr_ = (double*)take(sizeof(double) * nseg_);
    chk_ = (double*)take(sizeof(double) * 4 * nseg_);
    mv_ = (double*)take(sizeof(double) * 2 * nseg_);
    cg_ = (CgState*)take(sizeof(CgState) * nseg_);
    flags_ = (int*)take(sizeof(int) * nseg_);
    HIP_CHECK(hipHostMalloc((void**)&host_flags_, sizeof(int) * nseg_, hipHostMallocDefault));
    build_tables();
    mixed_ = o.precision == MAC_PRECISION_MIXED;
    if (mixed_) setup_mixed();
  }

  ~BatchSolver() {
    (void)hipSetDevice(opts_.device);
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    for (StageGraph& g : stage_graphs_) {
      (void)hipGraphExecDestroy(g.exec);
      (void)hipGraphDestroy(g.graph);
    }
    if (ctl_host_) (void)hipHostFree(ctl_host_);
    for (void* p : allocs_) (void)hipFree(p);
    if (hist_u_) (void)hipFree(hist_u_);
    if (hist_p_) (void)hipFree(hist_p_);
    if (host_flags_) (void)hipHostFree(host_flags_);
    if (stream_) (void)hipStreamDestroy(stream_);
  }

  int count() const { return (int)cav_.size(); }
  int ncols() const { return ncols_; }
  int64_t total_faces() const { return V_; }
  int64_t total_cells() const { return P_; }
  const Seg& seg(int s) const { return segs_host_[s]; }
  const Cavity& cavity(int i) const { return cav_[i]; }
  hipStream_t stream() const { return stream_; }
  const mac_stokes_opts& opts() const { return opts_; }
  double* io_f() { return io_f_; }
  double* io_v() { return io_v_; }
  double* io_p() { return io_p_; }
  std::string error;
  int use_graph = 1;        // 1: replay one captured CG iteration instead of launching it;
                            // 2 (mixed form): the stage graph, one launch per stage
  int iters_per_check = 1;  // CG iterations queued between host reads of the flags
  static constexpr int kStageRounds = 4;  // stage graph: most rounds in a first launch
  int stage_iters_ = 4;     // stage graph: CG iterations per round (then the last stage's longest CG)
  int stage_rounds_ = kStageRounds;  // rounds of a stage's first launch (then the last stage's sweeps + 1)
  int last_sweeps = 0;      // outer sweeps of the last solve (1 for the FP64 form)

  bool mixed() const { return mixed_; }

  struct Outcome {
    std::vector<int> iterations;              // per segment
    std::vector<double> residual, divergence, vmax;  // per cavity
    std::vector<int> failed_cg;               // cavities whose CG did not converge
    int sweeps = 1;                           // mixed form: outer sweeps of the batch
    bool warm = false;                        // mixed form: started from the history
  };

  // Same-stage history for the mixed form (period: solves per step; steps: how many steps
  // the extrapolation reaches back, 1..kMaxExtrapolation; period 0 switches it off).  All
  // cavities step together, so one step-size sequence (cavity 0's viscosity / mass) gives
  // every cavity's weights.  The ring keeps steps + 1 steps, so the solves of an unmarked
  // trial never overwrite the committed ones the extrapolation reads.
  void set_history(int period, int steps) {
    HIP_CHECK(hipStreamSynchronize(stream_));
    const int slots = period > 0 ? (steps + 1) * period : 0;
    if (slots > hist_slots_) {
      if (hist_u_) HIP_CHECK(hipFree(hist_u_));
      if (hist_p_) HIP_CHECK(hipFree(hist_p_));
      hist_u_ = hist_p_ = nullptr;
      hist_slots_ = 0;
      HIP_CHECK(hipMalloc(&hist_u_, sizeof(double) * V_ * slots));
      HIP_CHECK(hipMalloc(&hist_p_, sizeof(double) * P_ * slots));
      hist_slots_ = slots;
    }
    hist_period_ = period;
    hist_steps_ = steps;
    hist_count_ = hist_committed_ = 0;
    hist_shift_.assign((size_t)std::max(slots, 1), 0.0);
  }
  // rollback: forget the solves since the last mark (a rejected trial), else keep them.
  void history_mark(bool rollback) {
    if (rollback) hist_count_ = hist_committed_;
    else hist_committed_ = hist_count_;
  }

  // f, vel, pres: device buffers in the batch layout.  mass, visc: per cavity.
  // Per-segment max|f| (non-finite values reported as inf).
  std::vector<double> rhs_max(const double* f) {
    k_breduce<RED_MAXABS, false><<<igrid(LEN_FACES), kThreads, 0, stream_>>>(
        items(LEN_FACES), f, nullptr, segs_, LEN_FACES, partials_, pstride_, nullptr);
    bfinalize(RED_MAXABS, LEN_FACES, sums_);
    std::vector<double> m(nseg_);
    HIP_CHECK(hipMemcpyAsync(m.data(), sums_, sizeof(double) * nseg_, hipMemcpyDeviceToHost, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
    return m;
  }

  Outcome solve(const double* f, const double* mass, const double* visc, double* vel, double* pres) {
    const int nc = count();
    set_shifts(mass, visc);
    // W = L^-1 f ; b = -D W, mean removed; ||b||^2 and max|b| per segment.
    copy(f, vin_, V_);
    velocity_inverse(false);
    bdiv(vout_, b_, -1.0, RED_SUM, nullptr);
    bfinalize(RED_SUM, LEN_CELLS, sums_);
    k_bsub_mean_dot<true><<<igrid(LEN_CELLS), kThreads, 0, stream_>>>(
        items(LEN_CELLS), b_, sums_, b_, segs_, partials_, pstride_, nullptr);
    bfinalize(RED_SUM, LEN_CELLS, sums2_);
    k_breduce<RED_MAXABS, false><<<igrid(LEN_CELLS), kThreads, 0, stream_>>>(
        items(LEN_CELLS), b_, nullptr, segs_, LEN_CELLS, partials_, pstride_, nullptr);
    bfinalize(RED_MAXABS, LEN_CELLS, sums3_);
    cg_init(cg_, sums2_, sums3_, opts_.krylov_tolerance, nseg_, stream_);
    k_bfill<<<blocks_for(P_), kThreads, 0, stream_>>>(x_, P_, 0.0, nullptr);
    copy(b_, r_, P_);
    copy(sums2_, rr_, nseg_);
    run_cg();
    std::vector<CgState> st(nseg_);
    HIP_CHECK(hipMemcpyAsync(st.data(), cg_, sizeof(CgState) * nseg_, hipMemcpyDeviceToHost, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
    Outcome out;
    out.iterations.resize(nseg_);
    for (int s = 0; s < nseg_; ++s) {
      out.iterations[s] = st[s].iters;
      const int cav = segs_host_[s].cav;
      if (!st[s].converged && (out.failed_cg.empty() || out.failed_cg.back() != cav)) out.failed_cg.push_back(cav);
    }
    if (!out.failed_cg.empty()) return out;
    // p = x - mean(x); u = L^-1 (f + D^T p).
    k_breduce<RED_SUM, false><<<igrid(LEN_CELLS), kThreads, 0, stream_>>>(
        items(LEN_CELLS), x_, nullptr, segs_, LEN_CELLS, partials_, pstride_, nullptr);
    bfinalize(RED_SUM, LEN_CELLS, sums_);
    k_bsub_mean_dot<false><<<igrid(LEN_CELLS), kThreads, 0, stream_>>>(
        items(LEN_CELLS), x_, sums_, nullptr, segs_, nullptr, pstride_, nullptr);
    k_bdt<<<igrid(ITEMS_DT), kThreads, 0, stream_>>>(items(ITEMS_DT), x_, f, vin_, segs_, nullptr, nullptr);
    velocity_inverse(false);
    copy(vout_, vel, V_);
    copy(x_, pres, P_);
    // Checks: residual and divergence per cavity (max over its columns).
    k_bdefect<false><<<igrid(LEN_FACES), kThreads, 0, stream_>>>(
        items(LEN_FACES), vel, f, pres, segs_, mv_, partials_, pstride_, nseg_, nullptr, nullptr);
    for (int q = 0; q < 3; ++q)
      k_bfinalize<RED_MAXABS><<<nseg_, kThreads, 0, stream_>>>(partials_ + (int64_t)q * nseg_ * pstride_, pstride_,
                                                              segs_, LEN_FACES, chk_ + q * nseg_, nullptr);
    bdiv(vel, q_, 1.0, RED_MAXABS, nullptr);
    bfinalize(RED_MAXABS, LEN_CELLS, chk_ + 3 * nseg_);
    std::vector<double> h(4 * (size_t)nseg_);
    HIP_CHECK(hipMemcpyAsync(h.data(), chk_, sizeof(double) * 4 * nseg_, hipMemcpyDeviceToHost, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
    out.residual.assign(nc, 0.0);
    out.divergence.assign(nc, 0.0);
    out.vmax.assign(nc, 0.0);
    std::vector<double> dmax(nc, 0.0), fmax_(nc, 0.0);
    for (int s = 0; s < nseg_; ++s) {
      const int c = segs_host_[s].cav;
      dmax[c] = std::max(dmax[c], h[s]);
      fmax_[c] = std::max(fmax_[c], h[nseg_ + s]);
      out.vmax[c] = std::max(out.vmax[c], h[2 * nseg_ + s]);
      out.divergence[c] = std::max(out.divergence[c], h[3 * nseg_ + s]);
    }
    for (int c = 0; c < nc; ++c) out.residual[c] = dmax[c] / (1.0 + fmax_[c]);
    return out;
  }

  // Mixed form, per cavity as the single-cavity mixed solve: from u = 0, p = 0, sweeps of
  //   FP64 defect r_u = f - (mass u + visc K u + G p), r_c = -D u, and their maxima;
  //   stop when max|r_u| / (1 + max|f|) <= refine_tolerance and max|D u| <= refine_tolerance;
  //   inner solve with bf16x3 transforms: W = L~^-1 r_u, b = r_c - D W (mean removed),
  //   Schur PCG to inner_tolerance, dp = x - mean(x), du = L~^-1 (r_u + D^T dp), summed
  //   during the CG as W + sum_k alpha_k L~^-1 D^T p_k from the Schur products;
  //   u += du, p += dp.
  // Finished cavities drop out on the device (their segments' states stay inactive); the
  // host reads the outer flags and maxima once per sweep.  The CG vectors, dot products,
  // defects and updates are FP64; only the transforms are split bf16 with FP32
  // accumulation.  Returns per-cavity residual and max|D u| of the last defect, and per
  // segment the inner CG iterations summed over sweeps.
  Outcome solve_mixed(const double* f, const double* mass, const double* visc, double* vel, double* pres) {
    const int nc = count();
    set_shifts(mass, visc);
    // Outer start: the same-stage history's extrapolation when one is recorded, else zero.
    const double shift = mass[0] > 0.0 ? visc[0] / mass[0] : 0.0;
    const int slots = hist_period_ > 0 ? (hist_steps_ + 1) * hist_period_ : 0;
    const int steps = hist_period_ > 0 ? (int)std::min<int64_t>(hist_count_ / hist_period_, hist_steps_) : 0;
    if (steps > 0) {
      double h[kMaxExtrapolation + 1] = {shift}, w[kMaxExtrapolation];
      const double* su[kMaxExtrapolation];
      const double* sp[kMaxExtrapolation];
      for (int k = 1; k <= steps; ++k) {
        const int slot = (int)((hist_count_ - (int64_t)k * hist_period_) % slots);
        h[k] = hist_shift_[slot];
        su[k - 1] = hist_u_ + (int64_t)slot * V_;
        sp[k - 1] = hist_p_ + (int64_t)slot * P_;
      }
      history_weights(h, steps, w);
      extrapolate(su, w, steps, uacc_, V_, stream_);
      extrapolate(sp, w, steps, pacc_, P_, stream_);
    } else {
      k_bfill<<<blocks_for(V_), kThreads, 0, stream_>>>(uacc_, V_, 0.0, nullptr);
      k_bfill<<<blocks_for(P_), kThreads, 0, stream_>>>(pacc_, P_, 0.0, nullptr);
    }
    HIP_CHECK(hipMemsetAsync(iters_total_, 0, sizeof(int) * nseg_, stream_));
    const unsigned gb = (unsigned)((nseg_ + 63) / 64);
    std::vector<double> h(4 * (size_t)nseg_);
    std::vector<int> act(nseg_);
    Outcome out;
    int sweep = 0;
    if (use_graph == 2) {
      stage_launches(f);
      sweep = ctl_host_->sweeps;
      HIP_CHECK(hipMemcpyAsync(h.data(), chk_, sizeof(double) * 4 * nseg_, hipMemcpyDeviceToHost, stream_));
      HIP_CHECK(hipStreamSynchronize(stream_));
      for (int q = 0; q < nseg_; ++q)
        if (!std::isfinite(h[nseg_ + q])) throw ArgFailure("Invalid Stokes RHS");
    }
    for (; use_graph != 2; ++sweep) {
      outer_defect(f, nullptr);
      k_bouter<<<gb, 64, 0, stream_>>>(chk_, nseg_, ncols_, opts_.refine_tolerance, outer_);
      HIP_CHECK(hipMemcpyAsync(h.data(), chk_, sizeof(double) * 4 * nseg_, hipMemcpyDeviceToHost, stream_));
      HIP_CHECK(hipMemcpyAsync(act.data(), outer_, sizeof(int) * nseg_, hipMemcpyDeviceToHost, stream_));
      HIP_CHECK(hipStreamSynchronize(stream_));
      // A non-finite right-hand side is an invalid argument (tensor_stokes.py:246-257).
      if (sweep == 0)
        for (int q = 0; q < nseg_; ++q)
          if (!std::isfinite(h[nseg_ + q])) throw ArgFailure("Invalid Stokes RHS");
      int active = 0;
      bool changed = false;
      for (int q = 0; q < nseg_; ++q) {
        // A cavity whose defect is no longer finite stops here; the gate reports it.
        const bool finite = std::isfinite(h[q]) && std::isfinite(h[3 * nseg_ + q]);
        if (act[q] && !finite) {
          act[q] = 0;
          changed = true;
        }
        active += act[q] != 0;
      }
      if (active == 0 || sweep >= opts_.max_sweeps) break;
      if (changed) HIP_CHECK(hipMemcpy(outer_, act.data(), sizeof(int) * nseg_, hipMemcpyHostToDevice));
      // Inner solve for (du, dp) on the unfinished cavities.
      k_bgate<<<gb, 64, 0, stream_>>>(cg_, outer_, nseg_, GATE_SET, nullptr, nullptr);
      inner_setup(nullptr);
      run_cg();
      k_bgate<<<gb, 64, 0, stream_>>>(cg_, outer_, nseg_, GATE_POST, iters_total_, nullptr);
      inner_fold(nullptr);
    }
    out.sweeps = sweep;
    // p mean-zero; hand out u and p.
    k_breduce<RED_SUM, false><<<igrid(LEN_CELLS), kThreads, 0, stream_>>>(
        items(LEN_CELLS), pacc_, nullptr, segs_, LEN_CELLS, partials_, pstride_, nullptr);
    bfinalize(RED_SUM, LEN_CELLS, sums_);
    k_bsub_mean_dot<false><<<igrid(LEN_CELLS), kThreads, 0, stream_>>>(
        items(LEN_CELLS), pacc_, sums_, nullptr, segs_, nullptr, pstride_, nullptr);
    copy(uacc_, vel, V_);
    copy(pacc_, pres, P_);
    if (slots > 0) {  // record this solve for the later same-stage starts
      const int slot = (int)(hist_count_ % slots);
      copy(uacc_, hist_u_ + (int64_t)slot * V_, V_);
      copy(pacc_, hist_p_ + (int64_t)slot * P_, P_);
      hist_shift_[slot] = shift;
      ++hist_count_;
    }
    out.warm = steps > 0;
    out.iterations.resize(nseg_);
    HIP_CHECK(hipMemcpyAsync(out.iterations.data(), iters_total_, sizeof(int) * nseg_, hipMemcpyDeviceToHost,
                             stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
    out.residual.assign(nc, 0.0);
    out.divergence.assign(nc, 0.0);
    out.vmax.assign(nc, 0.0);
    std::vector<double> dmax(nc, 0.0), fmax_(nc, 0.0);
    for (int q = 0; q < nseg_; ++q) {
      const int c = segs_host_[q].cav;
      dmax[c] = std::max(dmax[c], h[q]);
      fmax_[c] = std::max(fmax_[c], h[nseg_ + q]);
      out.vmax[c] = std::max(out.vmax[c], h[2 * nseg_ + q]);
      out.divergence[c] = std::max(out.divergence[c], h[3 * nseg_ + q]);
    }
    for (int c = 0; c < nc; ++c) out.residual[c] = dmax[c] / (1.0 + fmax_[c]);
    return out;
  }

 private:
  // The outer test's FP64 defect of every segment: r_u = f - (mass u + visc K u + G p) into
  // the FP32 transform input, the maxima of the defect, f and u, and max|D u| (rc = -D u)
  // into chk_.
  void outer_defect(const double* f, const int* run) {
    k_bdefect<true><<<igrid(LEN_FACES), kThreads, 0, stream_>>>(
        items(LEN_FACES), uacc_, f, pacc_, segs_, mv_, partials_, pstride_, nseg_, tf0_, run);
    for (int q = 0; q < 3; ++q)
      k_bfinalize<RED_MAXABS><<<nseg_, kThreads, 0, stream_>>>(partials_ + (int64_t)q * nseg_ * pstride_, pstride_,
                                                              segs_, LEN_FACES, chk_ + q * nseg_, run);
    bdiv(uacc_, rc_, -1.0, RED_MAXABS, nullptr, nullptr, run);
    bfinalize(RED_MAXABS, LEN_CELLS, chk_ + 3 * nseg_, run);
  }

  // The inner solve's set-up on the segments whose state is active: W = L~^-1 r_u (r_u in
  // the FP32 transform input), du = W (the CG adds alpha_k L~^-1 D^T p_k), b = r_c - D W
  // with its mean removed, the CG states from (b, b) and max|b| (finished cavities kept
  // inactive), x = 0, r = b.
  void inner_setup(const int* run) {
    const unsigned gb = (unsigned)((nseg_ + 63) / 64);
    velocity_inverse(true, run);
    copy(vout_, du_, V_, run);
    bdiv(vout_, b_, -1.0, RED_SUM, cg_, rc_, run);
    bfinalize(RED_SUM, LEN_CELLS, sums_, run);
    k_bsub_mean_dot<true><<<igrid(LEN_CELLS), kThreads, 0, stream_>>>(
        items(LEN_CELLS), b_, sums_, b_, segs_, partials_, pstride_, cg_);
    bfinalize(RED_SUM, LEN_CELLS, sums2_, run);
    k_breduce<RED_MAXABS, false><<<igrid(LEN_CELLS), kThreads, 0, stream_>>>(
        items(LEN_CELLS), b_, nullptr, segs_, LEN_CELLS, partials_, pstride_, cg_);
    bfinalize(RED_MAXABS, LEN_CELLS, sums3_, run);
    if (run == nullptr) cg_init(cg_, sums2_, sums3_, opts_.inner_tolerance, nseg_, stream_);
    else k_stage_cg_init<<<gb, 64, 0, stream_>>>(cg_, sums2_, sums3_, opts_.inner_tolerance, nseg_, run);
    k_bgate<<<gb, 64, 0, stream_>>>(cg_, outer_, nseg_, GATE_MASK, nullptr, run);
    k_bfill<<<blocks_for(P_), kThreads, 0, stream_>>>(x_, P_, 0.0, run);
    copy(b_, r_, P_, run);
    copy(sums2_, rr_, nseg_, run);
  }

  // After the inner CG, on the segments whose state is active: dp = x - mean(x) (D^T
  // annihilates the mean, so du needs no correction), u += du, p += dp.
  void inner_fold(const int* run) {
    k_breduce<RED_SUM, false><<<igrid(LEN_CELLS), kThreads, 0, stream_>>>(
        items(LEN_CELLS), x_, nullptr, segs_, LEN_CELLS, partials_, pstride_, cg_);
    bfinalize(RED_SUM, LEN_CELLS, sums_, run);
    k_bsub_mean_dot<false><<<igrid(LEN_CELLS), kThreads, 0, stream_>>>(
        items(LEN_CELLS), x_, sums_, nullptr, segs_, nullptr, pstride_, cg_);
    k_baccum<<<igrid(LEN_FACES), kThreads, 0, stream_>>>(items(LEN_FACES), uacc_, du_, pacc_, x_, segs_, cg_);
  }

  // ------------------------------------------------------------ stage graph
  // One round: the outer test (segments whose last inner CG has been folded in), the
  // inner solve's set-up when a cavity is unfinished, `iters` CG iterations, and the fold
  // once every CG has stopped.  Each segment goes through the same operations in the
  // same order as in the host-driven loop; a CG that needs more iterations than a round
  // holds continues in the next round, whose outer steps are then switched off.
  void stage_round(int iters) {
    const unsigned gb = (unsigned)((nseg_ + 63) / 64);
    k_stage_round_begin<<<1, 1, 0, stream_>>>(ctl_);
    outer_defect(io_f_, &ctl_->run_outer);
    k_bouter_stage<<<gb, 64, 0, stream_>>>(chk_, nseg_, ncols_, opts_.refine_tolerance, outer_, &ctl_->run_outer);
    k_stage_after_outer<<<1, kThreads, 0, stream_>>>(ctl_, outer_, nseg_, opts_.max_sweeps);
    k_stage_gate<<<gb, 64, 0, stream_>>>(ctl_, cg_, outer_, run_, iters_total_, nseg_, STAGE_SETUP);
    inner_setup(&ctl_->go_setup);
    k_stage_gate<<<gb, 64, 0, stream_>>>(ctl_, cg_, outer_, run_, iters_total_, nseg_, STAGE_CG_START);
    cg_run_ = &ctl_->in_cg;
    for (int k = 0; k < iters; ++k) {
      k_stage_cg_check<<<1, kThreads, 0, stream_>>>(ctl_, cg_, rr_, run_, nseg_, opts_.max_iterations, 1);
      iteration_body();
    }
    k_stage_cg_check<<<1, kThreads, 0, stream_>>>(ctl_, cg_, rr_, run_, nseg_, opts_.max_iterations, 0);
    cg_run_ = nullptr;
    k_stage_gate<<<gb, 64, 0, stream_>>>(ctl_, cg_, outer_, run_, iters_total_, nseg_, STAGE_FOLD);
    inner_fold(&ctl_->go_fold);
    k_stage_round_end<<<1, 1, 0, stream_>>>(ctl_);
  }

  // The captured graph of `rounds` rounds of `iters` CG iterations (captured once per shape).
  hipGraphExec_t stage_graph(int rounds, int iters) {
    for (const StageGraph& g : stage_graphs_)
      if (g.rounds == rounds && g.iters == iters) return g.exec;
    StageGraph g{rounds, iters, nullptr, nullptr};
    HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    for (int r = 0; r < rounds; ++r) stage_round(iters);
    HIP_CHECK(hipStreamEndCapture(stream_, &g.graph));
    HIP_CHECK(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
    stage_graphs_.push_back(g);
    return g.exec;
  }

  // The mixed solve's sweeps as stage-graph launches until the device reports the stage
  // finished (every cavity at the bar or stopped, or the sweep cap).  The first launch
  // holds as many rounds as the previous stage needed (its sweeps and the final test) of
  // as many CG iterations as its longest inner solve, so a stage like the last one takes
  // one launch; a longer one continues with one-round launches.  The shape changes the
  // time only, never the results.
  void stage_launches(const double* f) {
    if (f != io_f_) copy(f, io_f_, V_);  // the graphs read the right-hand side from io_f_
    k_stage_init<<<1, 1, 0, stream_>>>(ctl_);
    // Every round advances a CG iteration or a sweep, so this many launches always reach
    // the end; the device stops itself at the caps.
    const int64_t cap = (int64_t)(opts_.max_sweeps + 1) * (opts_.max_iterations + 2) + 1;
    for (int64_t launches = 0;; ++launches) {
      if (launches >= cap) throw HipFailure("stage graph did not finish");
      HIP_CHECK(hipGraphLaunch(stage_graph(launches == 0 ? stage_rounds_ : 1, stage_iters_), stream_));
      HIP_CHECK(hipMemcpyAsync(ctl_host_, ctl_, sizeof(StageCtl), hipMemcpyDeviceToHost, stream_));
      HIP_CHECK(hipStreamSynchronize(stream_));
      if (ctl_host_->halt) break;
    }
    stage_rounds_ = std::min(kStageRounds, std::max(1, ctl_host_->sweeps + 1));
    stage_iters_ = std::min(16, std::max(1, ctl_host_->max_cg));
  }

  // Schur PCG, every segment at once; converged segments drop out on the device.  As in
  // scipy's loop, the convergence test runs only at the top of an iteration, so none
  // follows the last of max_iterations.
  void run_cg() {
    int done_iters = 0;
    while (done_iters < opts_.max_iterations) {
      cg_check(cg_, rr_, flags_, nseg_, stream_);
      HIP_CHECK(hipMemcpyAsync(host_flags_, flags_, sizeof(int) * nseg_, hipMemcpyDeviceToHost, stream_));
      HIP_CHECK(hipStreamSynchronize(stream_));
      int active = 0;
      for (int q = 0; q < nseg_; ++q) active += host_flags_[q] != 0;
      if (active == 0) break;
      // Queue up to iters_per_check iterations (never past max_iterations); the state
      // test at the top of each queued iteration runs on the device.
      const int burst = std::min(iters_per_check, opts_.max_iterations - done_iters);
      for (int k = 0; k < burst; ++k) {
        if (k > 0) cg_check(cg_, rr_, flags_, nseg_, stream_);
        iteration();
      }
      done_iters += burst;
    }
  }

  // ------------------------------------------------------------ one CG iteration
  void iteration() {
    if (use_graph) {
      if (!graph_exec_) {
        HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        iteration_body();
        HIP_CHECK(hipStreamEndCapture(stream_, &graph_));
        HIP_CHECK(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
      }
      HIP_CHECK(hipGraphLaunch(graph_exec_, stream_));
    } else {
      iteration_body();
    }
  }

  void iteration_body() {
    // z = mass (D D^T)^+ r + visc r (per cavity), mean removed; rho = (r, z)
    pressure_precondition(cg_run_);
    k_breduce<RED_SUM, false><<<igrid(LEN_CELLS), kThreads, 0, stream_>>>(
        items(LEN_CELLS), z_, nullptr, segs_, LEN_CELLS, partials_, pstride_, cg_);
    bfinalize(RED_SUM, LEN_CELLS, sums_);
    k_bsub_mean_dot<true><<<igrid(LEN_CELLS), kThreads, 0, stream_>>>(
        items(LEN_CELLS), z_, sums_, r_, segs_, partials_, pstride_, cg_);
    bfinalize(RED_SUM, LEN_CELLS, sums2_);
    cg_rho(cg_, sums2_, nseg_, stream_);
    k_bupdate_p<<<igrid(LEN_CELLS), kThreads, 0, stream_>>>(items(LEN_CELLS), p_, z_, segs_, cg_);
    // q = S p = D L^-1 D^T p, mean removed; alpha = rho / (p, q)
    k_bdt<<<igrid(ITEMS_DT), kThreads, 0, stream_>>>(
        items(ITEMS_DT), p_, nullptr, vin_, segs_, cg_, mixed_ ? tf0_ : nullptr);
    velocity_inverse(true, cg_run_);
    bdiv(vout_, q_, 1.0, RED_SUM, cg_);
    bfinalize(RED_SUM, LEN_CELLS, sums_);
    k_bsub_mean_dot<true><<<igrid(LEN_CELLS), kThreads, 0, stream_>>>(
        items(LEN_CELLS), q_, sums_, p_, segs_, partials_, pstride_, cg_);
    bfinalize(RED_SUM, LEN_CELLS, sums2_);
    cg_alpha(cg_, sums2_, nseg_, stream_);
    if (mixed_) k_bacc_du<<<igrid(LEN_FACES), kThreads, 0, stream_>>>(items(LEN_FACES), du_, vout_, segs_, cg_);
    k_bupdate_xr<<<igrid(LEN_CELLS), kThreads, 0, stream_>>>(
        items(LEN_CELLS), x_, r_, p_, q_, segs_, cg_, partials_, pstride_);
    bfinalize(RED_SUM, LEN_CELLS, rr_);
  }

  // ------------------------------------------------------------ transforms
  // Velocity: vout = L^-1 vin for every segment (per-cavity shifts in the tables); the
  // mixed form's tables run the split bf16 GEMM.
  void velocity_inverse(bool use_state, const int* run = nullptr) {
    for (int k = 0; k < 4; ++k) {
      GemmTable t = mixed_ ? velb_tab_[k] : vel_tab_[k];
      t.use_state = use_state ? 1 : 0;
      t.run = run;
      if (mixed_) run_table(t, k == 0 ? DT_F32 : kPassTypes[k][0], kPassTypes[k][1], true);
      else gemm_f64_table(t, stream_);
    }
  }
  // Pressure: z = mass (D D^T)^+ r + visc r.
  void pressure_precondition(const int* run = nullptr) {
    for (int k = 0; k < 4; ++k) {
      GemmTable t = mixed_ ? preb_tab_[k] : pre_tab_[k];
      t.run = run;
      if (mixed_) run_table(t, kPassTypes[k][0], kPassTypes[k][1], k == 0 ? preb_avec_ : true);
      else gemm_f64_table(t, stream_);
    }
  }
  // A-operand / output types of the four passes in the mixed form: the input (FP64 for the
  // pressure, the FP32 transform input for the velocity), two FP32 intermediates, the FP64
  // output.
  static constexpr int kPassTypes[4][2] = {{DT_F64, DT_F32}, {DT_F32, DT_F32}, {DT_F32, DT_F32}, {DT_F32, DT_F64}};
  void run_table(const GemmTable& t, int a_type, int c_type, bool a_vec) {
    if (gemm_bf16_table(t, a_type, c_type, a_vec, stream_) != 0)
      throw HipFailure("unsupported split-bf16 table pass");
  }

  static void pass(std::vector<GemmEntry>& v, const double* A, int lda, const DevMatrix64* Q, double* C, int ldc,
                   int M, int N, int K, const double* ea, const double* eb, const CgState* st) {
    GemmEntry e{};
    e.A = A;
    e.Bt = Q ? Q->q : nullptr;
    e.C = C;
    e.ea = ea;
    e.eb = eb;
    e.M = M;
    e.N = N;
    e.K = K;
    e.lda = lda;
    e.ldb = Q->n;
    e.ldc = ldc;
    e.st = st;
    v.push_back(e);
  }

  // The four passes of every segment's transforms, as in the single-cavity solver:
  // x-forward (transposed store), y-forward (+ divide), y-inverse (transposed store),
  // x-inverse.  Q^T passes take the transposed matrix as Bt.
  void build_tables() {
    std::vector<GemmEntry> va, vbb, vc, vd, pa, pbb, pc, pd;
    for (int s = 0; s < nseg_; ++s) {
      const Seg& S = segs_host_[s];
      const Cavity& C = cav_[S.cav];
      const CgState* st = cg_ + s;
      struct Blk { int ny, nx; const DevMatrix64 *my, *mx; double *ly, *lx; int64_t off; };
      const Blk blocks[2] = {{C.ny, C.nx - 1, C.uy, C.ux, C.luy, C.lux, 0},
                             {C.ny - 1, C.nx, C.vy, C.vx, C.lvy, C.lvx, (int64_t)S.nu}};
      for (const Blk& B : blocks) {
        const int64_t o = S.voff + B.off;
        pass(va, vin_ + o, B.nx, B.mx, s1_ + o, B.ny, B.ny, B.nx, B.nx, nullptr, nullptr, st);
        pass(vbb, s1_ + o, B.ny, B.my, s2_ + o, B.ny, B.nx, B.ny, B.ny, B.lx, B.ly, st);
        pass(vc, s2_ + o, B.ny, B.my, s1_ + o, B.nx, B.nx, B.ny, B.ny, nullptr, nullptr, st);
        vc.back().Bt = B.my->qt;
        pass(vd, s1_ + o, B.nx, B.mx, vout_ + o, B.nx, B.ny, B.nx, B.nx, nullptr, nullptr, st);
        vd.back().Bt = B.mx->qt;
      }
      const int64_t o = S.poff;
      // Pressure scratch reuses the segment's velocity scratch: np <= size for nx, ny >= 3.
      double* t1 = s1_ + S.voff;
      double* t2 = s2_ + S.voff;
      pass(pa, r_ + o, C.nx, C.px, t1, C.ny, C.ny, C.nx, C.nx, nullptr, nullptr, st);
      pass(pbb, t1, C.ny, C.py, t2, C.ny, C.nx, C.ny, C.ny, C.lpx, C.lpy, st);
      pbb.back().thresh = C.p_thresh;
      pass(pc, t2, C.ny, C.py, t1, C.nx, C.nx, C.ny, C.ny, nullptr, nullptr, st);
      pc.back().Bt = C.py->qt;
      pass(pd, t1, C.nx, C.px, z_ + o, C.nx, C.ny, C.nx, C.nx, nullptr, nullptr, st);
      pd.back().Bt = C.px->qt;
      pd.back().aux = r_ + o;
    }
    vel_host_[0] = va; vel_host_[1] = vbb; vel_host_[2] = vc; vel_host_[3] = vd;
    pre_host_[0] = pa; pre_host_[1] = pbb; pre_host_[2] = pc; pre_host_[3] = pd;
    const int vepi[4] = {EPI_STORE, EPI_SHIFT, EPI_STORE, EPI_STORE};
    const int pepi[4] = {EPI_STORE, EPI_LAMBDA, EPI_STORE, EPI_COMBINE};
    const int trans[4] = {1, 0, 1, 0};
    for (int k = 0; k < 4; ++k) {
      vel_tab_[k] = make_table(vel_host_[k], vel_dev_[k], vel_beg_[k], vepi[k], trans[k]);
      pre_tab_[k] = make_table(pre_host_[k], pre_dev_[k], pre_beg_[k], pepi[k], trans[k]);
    }
  }

  // ------------------------------------------------------------ mixed form set-up
  // FP32 transform intermediates: per segment one region for each block (u, v, p), large
  // enough for both transposed layouts with rows padded to 8 floats, 16-byte aligned; the
  // split-bf16 tables over them; the outer iterate and residual buffers.
  void setup_mixed() {
    int64_t t = 0;
    toff_.assign(3 * (size_t)nseg_, 0);
    for (int q = 0; q < nseg_; ++q) {
      const Seg& S = segs_host_[q];
      const int dims[3][2] = {{S.ny, S.nx - 1}, {S.ny - 1, S.nx}, {S.ny, S.nx}};
      for (int b = 0; b < 3; ++b) {
        const int by = dims[b][0], bx = dims[b][1];
        const int64_t need = std::max<int64_t>((int64_t)bx * round8(by), (int64_t)by * round8(bx));
        toff_[3 * q + b] = t;
        t += (need + 3) & ~int64_t(3);
      }
    }
    tf0_ = static_cast<float*>(dev_alloc(sizeof(float) * t));
    tf1_ = static_cast<float*>(dev_alloc(sizeof(float) * t));
    tf2_ = static_cast<float*>(dev_alloc(sizeof(float) * t));
    for (int q = 0; q < nseg_; ++q) {
      segs_host_[q].t32u = toff_[3 * q];
      segs_host_[q].t32v = toff_[3 * q + 1];
    }
    HIP_CHECK(hipMemcpy(segs_, segs_host_.data(), sizeof(Seg) * nseg_, hipMemcpyHostToDevice));
    uacc_ = static_cast<double*>(dev_alloc(sizeof(double) * V_));
    pacc_ = static_cast<double*>(dev_alloc(sizeof(double) * P_));
    du_ = static_cast<double*>(dev_alloc(sizeof(double) * V_));
    rc_ = static_cast<double*>(dev_alloc(sizeof(double) * P_));
    outer_ = static_cast<int*>(dev_alloc(sizeof(int) * nseg_));
    iters_total_ = static_cast<int*>(dev_alloc(sizeof(int) * nseg_));
    run_ = static_cast<int*>(dev_alloc(sizeof(int) * nseg_));
    ctl_ = static_cast<StageCtl*>(dev_alloc(sizeof(StageCtl)));
    HIP_CHECK(hipHostMalloc((void**)&ctl_host_, sizeof(StageCtl), hipHostMallocDefault));
    build_tables_bf16();
  }

  static void passb(std::vector<GemmEntry>& v, const void* A, int lda, const DevMatrixB* Q, bool transposed, void* C,
                    int ldc, int M, int N, int K, const double* ea, const double* eb, const CgState* st) {
    GemmEntry e{};
    e.A = A;
    e.Bt = transposed ? Q->qtb : Q->qb;
    e.Bt_lo = transposed ? Q->qtb_lo : Q->qb_lo;
    e.C = C;
    e.ea = ea;
    e.eb = eb;
    e.M = M;
    e.N = N;
    e.K = K;
    e.lda = lda;
    e.ldb = Q->ld8;
    e.ldc = ldc;
    e.st = st;
    v.push_back(e);
  }

  // The same four passes per block as build_tables, FP32 intermediates in between.
  void build_tables_bf16() {
    std::vector<GemmEntry> v[4], pr[4];
    for (int q = 0; q < nseg_; ++q) {
      const Seg& S = segs_host_[q];
      const Cavity& C = cav_[S.cav];
      const CgState* st = cg_ + q;
      struct Blk { int ny, nx; const DevMatrixB *my, *mx; double *ly, *lx; int64_t off; int reg; };
      const Blk blocks[2] = {
          {C.ny, C.nx - 1, matrix_b(C.ny, KIND_DST2), matrix_b(C.nx - 1, KIND_DST1), C.luy, C.lux, 0, 0},
          {C.ny - 1, C.nx, matrix_b(C.ny - 1, KIND_DST1), matrix_b(C.nx, KIND_DST2), C.lvy, C.lvx, (int64_t)S.nu, 1}};
      for (const Blk& B : blocks) {
        const int64_t o = S.voff + B.off;
        float* t1 = tf1_ + toff_[3 * q + B.reg];
        float* t2 = tf2_ + toff_[3 * q + B.reg];
        const int ry = round8(B.ny), rx = round8(B.nx);
        passb(v[0], tf0_ + toff_[3 * q + B.reg], rx, B.mx, false, t1, ry, B.ny, B.nx, B.nx, nullptr, nullptr, st);
        passb(v[1], t1, ry, B.my, false, t2, ry, B.nx, B.ny, B.ny, B.lx, B.ly, st);
        passb(v[2], t2, ry, B.my, true, t1, rx, B.nx, B.ny, B.ny, nullptr, nullptr, st);
        passb(v[3], t1, rx, B.mx, true, vout_ + o, B.nx, B.ny, B.nx, B.nx, nullptr, nullptr, st);
      }
      const DevMatrixB* px = matrix_b(C.nx, KIND_DCT2);
      const DevMatrixB* py = matrix_b(C.ny, KIND_DCT2);
      float* t1 = tf1_ + toff_[3 * q + 2];
      float* t2 = tf2_ + toff_[3 * q + 2];
      const int ry = round8(C.ny), rx = round8(C.nx);
      passb(pr[0], r_ + S.poff, C.nx, px, false, t1, ry, C.ny, C.nx, C.nx, nullptr, nullptr, st);
      passb(pr[1], t1, ry, py, false, t2, ry, C.nx, C.ny, C.ny, C.lpx, C.lpy, st);
      pr[1].back().thresh = C.p_thresh;
      passb(pr[2], t2, ry, py, true, t1, rx, C.nx, C.ny, C.ny, nullptr, nullptr, st);
      passb(pr[3], t1, rx, px, true, z_ + S.poff, C.nx, C.ny, C.nx, C.nx, nullptr, nullptr, st);
      pr[3].back().aux = r_ + S.poff;
    }
    // Vector loads of the FP64 first pass need even rows and 16-byte aligned bases; the
    // FP32 intermediates have them by construction.
    auto vec64 = [](const std::vector<GemmEntry>& t) {
      for (const GemmEntry& e : t)
        if (e.lda % 2 != 0 || reinterpret_cast<uintptr_t>(e.A) % 16 != 0) return false;
      return true;
    };
    preb_avec_ = vec64(pr[0]);
    for (const GemmEntry& e : v[0])
      if (e.lda % 4 != 0 || reinterpret_cast<uintptr_t>(e.A) % 16 != 0)
        throw HipFailure("misaligned FP32 transform input");
    for (int k = 1; k < 4; ++k)
      for (const std::vector<GemmEntry>* t : {&v[k], &pr[k]})
        for (const GemmEntry& e : *t)
          if (e.lda % 4 != 0 || reinterpret_cast<uintptr_t>(e.A) % 16 != 0)
            throw HipFailure("misaligned FP32 transform intermediate");
    const int vepi[4] = {EPI_STORE, EPI_SHIFT, EPI_STORE, EPI_STORE};
    const int pepi[4] = {EPI_STORE, EPI_LAMBDA, EPI_STORE, EPI_COMBINE};
    const int trans[4] = {1, 0, 1, 0};
    for (int k = 0; k < 4; ++k) {
      velb_host_[k] = v[k];
      preb_host_[k] = pr[k];
      velb_tab_[k] = make_table(velb_host_[k], velb_dev_[k], velb_beg_[k], vepi[k], trans[k]);
      preb_tab_[k] = make_table(preb_host_[k], preb_dev_[k], preb_beg_[k], pepi[k], trans[k]);
    }
  }

  const DevMatrixB* matrix_b(int n, int kind) {
    for (auto& m : matsb_)
      if (m->n == n && m->kind == kind) return m.get();
    auto m = std::make_unique<DevMatrixB>();
    m->n = n;
    m->kind = kind;
    m->ld8 = round8(n);
    const std::vector<double> q = transform_matrix(n, kind);
    const size_t len = (size_t)n * m->ld8;
    std::vector<__bf16> qb(len, (__bf16)0.0f), qbl(len, (__bf16)0.0f), qtb(len, (__bf16)0.0f), qtbl(len, (__bf16)0.0f);
    for (int a = 0; a < n; ++a)
      for (int b = 0; b < n; ++b) {
        const double v = q[(size_t)a * n + b], vt = q[(size_t)b * n + a];
        const __bf16 hv = (__bf16)(float)v, ht = (__bf16)(float)vt;
        qb[(size_t)a * m->ld8 + b] = hv;
        qtb[(size_t)a * m->ld8 + b] = ht;
        qbl[(size_t)a * m->ld8 + b] = (__bf16)(float)(v - (double)(float)hv);
        qtbl[(size_t)a * m->ld8 + b] = (__bf16)(float)(vt - (double)(float)ht);
      }
    m->qb = static_cast<__bf16*>(upload_raw(qb.data(), len * sizeof(__bf16)));
    m->qb_lo = static_cast<__bf16*>(upload_raw(qbl.data(), len * sizeof(__bf16)));
    m->qtb = static_cast<__bf16*>(upload_raw(qtb.data(), len * sizeof(__bf16)));
    m->qtb_lo = static_cast<__bf16*>(upload_raw(qtbl.data(), len * sizeof(__bf16)));
    matsb_.push_back(std::move(m));
    return matsb_.back().get();
  }

  GemmTable make_table(std::vector<GemmEntry>& host, GemmEntry*& dev, int*& beg, int epi, int trans) {
    std::vector<int> begins(host.size());
    GemmTable t{};
    t.total_tiles = tile_table_f64(host.data(), (int)host.size(), begins.data());
    dev = static_cast<GemmEntry*>(upload_raw(host.data(), sizeof(GemmEntry) * host.size()));
    beg = static_cast<int*>(upload_raw(begins.data(), sizeof(int) * begins.size()));
    t.e = dev;
    t.begin = beg;
    t.count = (int)host.size();
    t.epi = epi;
    t.trans_c = trans;
    t.use_state = 1;
    return t;
  }

  // Per-cavity (mass, viscosity) into the velocity shift pass, the preconditioner
  // combination and the defect check.
  void set_shifts(const double* mass, const double* visc) {
    std::vector<double> mv(2 * (size_t)nseg_);
    for (int s = 0; s < nseg_; ++s) {
      mv[2 * s] = mass[segs_host_[s].cav];
      mv[2 * s + 1] = visc[segs_host_[s].cav];
    }
    bool changed = shifts_.size() != mv.size() || !std::equal(mv.begin(), mv.end(), shifts_.begin());
    if (!changed) return;
    shifts_ = mv;
    for (size_t i = 0; i < vel_host_[1].size(); ++i) {
      const int s = (int)(i / 2);
      vel_host_[1][i].s0 = mv[2 * s];
      vel_host_[1][i].s1 = mv[2 * s + 1];
    }
    for (size_t s = 0; s < pre_host_[3].size(); ++s) {
      pre_host_[3][s].s0 = mv[2 * s];
      pre_host_[3][s].s1 = mv[2 * s + 1];
    }
    HIP_CHECK(hipMemcpyAsync(vel_dev_[1], vel_host_[1].data(), sizeof(GemmEntry) * vel_host_[1].size(),
                             hipMemcpyHostToDevice, stream_));
    HIP_CHECK(hipMemcpyAsync(pre_dev_[3], pre_host_[3].data(), sizeof(GemmEntry) * pre_host_[3].size(),
                             hipMemcpyHostToDevice, stream_));
    if (mixed_) {
      for (size_t i = 0; i < velb_host_[1].size(); ++i) {
        const int q = (int)(i / 2);
        velb_host_[1][i].s0 = mv[2 * q];
        velb_host_[1][i].s1 = mv[2 * q + 1];
      }
      for (size_t q = 0; q < preb_host_[3].size(); ++q) {
        preb_host_[3][q].s0 = mv[2 * q];
        preb_host_[3][q].s1 = mv[2 * q + 1];
      }
      HIP_CHECK(hipMemcpyAsync(velb_dev_[1], velb_host_[1].data(), sizeof(GemmEntry) * velb_host_[1].size(),
                               hipMemcpyHostToDevice, stream_));
      HIP_CHECK(hipMemcpyAsync(preb_dev_[3], preb_host_[3].data(), sizeof(GemmEntry) * preb_host_[3].size(),
                               hipMemcpyHostToDevice, stream_));
    }
    HIP_CHECK(hipMemcpyAsync(mv_, mv.data(), sizeof(double) * mv.size(), hipMemcpyHostToDevice, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
  }

  // ------------------------------------------------------------ helpers
  // Item lists (LEN_FACES: the segments' reduction blocks over faces, LEN_CELLS: over
  // cells, ITEMS_DT: one thread per face) and the bounded grids that walk them.
  Items items(int kind) const { return Items{items_[kind], items_n_[kind]}; }
  unsigned igrid(int kind) const { return (unsigned)std::max(1, std::min(items_n_[kind], kMaxItemGrid)); }
  void bdiv(const double* vel, double* out, double sign, int mode, const CgState* st, const double* add = nullptr,
            const int* run = nullptr) {
    const unsigned g = igrid(LEN_CELLS);
    const Items w = items(LEN_CELLS);
    if (mode == RED_SUM)
      k_bdiv<RED_SUM><<<g, kThreads, 0, stream_>>>(w, vel, out, segs_, sign, partials_, pstride_, st, add, run);
    else
      k_bdiv<RED_MAXABS><<<g, kThreads, 0, stream_>>>(w, vel, out, segs_, sign, partials_, pstride_, st, add, run);
  }
  void bfinalize(int mode, int len, double* out, const int* run = nullptr) {
    if (mode == RED_SUM)
      k_bfinalize<RED_SUM><<<nseg_, kThreads, 0, stream_>>>(partials_, pstride_, segs_, len, out, run);
    else
      k_bfinalize<RED_MAXABS><<<nseg_, kThreads, 0, stream_>>>(partials_, pstride_, segs_, len, out, run);
  }
  void copy(const double* src, double* dst, int64_t n, const int* run = nullptr) {
    k_bcopy<<<blocks_for(n), kThreads, 0, stream_>>>(src, dst, n, run);
  }

  const DevMatrix64* matrix(int n, int kind) {
    for (auto& m : mats_)
      if (m->n == n && m->kind == kind) return m.get();
    auto m = std::make_unique<DevMatrix64>();
    m->n = n;
    m->kind = kind;
    std::vector<double> q = transform_matrix(n, kind);
    std::vector<double> qt((size_t)n * n);
    for (int a = 0; a < n; ++a)
      for (int b = 0; b < n; ++b) qt[(size_t)b * n + a] = q[(size_t)a * n + b];
    m->q = upload(q);
    m->qt = upload(qt);
    mats_.push_back(std::move(m));
    return mats_.back().get();
  }
  void* dev_alloc(size_t bytes) {
    void* p = nullptr;
    HIP_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 256)));
    allocs_.push_back(p);
    return p;
  }
  void* upload_raw(const void* h, size_t bytes) {
    void* d = dev_alloc(bytes);
    HIP_CHECK(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
    return d;
  }
  double* upload(const std::vector<double>& v) {
    return static_cast<double*>(upload_raw(v.data(), v.size() * sizeof(double)));
  }

  mac_stokes_opts opts_;
  static constexpr int kMaxItemGrid = 4096;  // blocks of an item launch
  int ncols_ = 0, nseg_ = 0, pstride_ = 1, nbv_max_ = 1, nbp_max_ = 1;
  const int2* items_[3] = {};
  int items_n_[3] = {};
  int64_t V_ = 0, P_ = 0, maxfaces_ = 1;
  hipStream_t stream_ = nullptr;
  std::vector<Cavity> cav_;
  std::vector<Seg> segs_host_;
  Seg* segs_ = nullptr;
  std::vector<std::unique_ptr<DevMatrix64>> mats_;
  std::vector<void*> allocs_;
  double *io_f_ = nullptr, *io_v_ = nullptr, *io_p_ = nullptr;
  double *vin_ = nullptr, *vout_ = nullptr, *s1_ = nullptr, *s2_ = nullptr;
  double *b_ = nullptr, *x_ = nullptr, *r_ = nullptr, *z_ = nullptr, *p_ = nullptr, *q_ = nullptr;
  double *partials_ = nullptr, *sums_ = nullptr, *sums2_ = nullptr, *sums3_ = nullptr, *rr_ = nullptr;
  double *chk_ = nullptr, *mv_ = nullptr;
  CgState* cg_ = nullptr;
  int* flags_ = nullptr;
  int* host_flags_ = nullptr;
  std::vector<GemmEntry> vel_host_[4], pre_host_[4];
  GemmEntry *vel_dev_[4] = {}, *pre_dev_[4] = {};
  int *vel_beg_[4] = {}, *pre_beg_[4] = {};
  GemmTable vel_tab_[4] = {}, pre_tab_[4] = {};
  std::vector<double> shifts_;
  // mixed form
  bool mixed_ = false;
  std::vector<std::unique_ptr<DevMatrixB>> matsb_;
  std::vector<int64_t> toff_;
  float *tf0_ = nullptr, *tf1_ = nullptr, *tf2_ = nullptr;  // transform input, intermediates
  double *uacc_ = nullptr, *pacc_ = nullptr, *rc_ = nullptr, *du_ = nullptr;
  int *outer_ = nullptr, *iters_total_ = nullptr;
  std::vector<GemmEntry> velb_host_[4], preb_host_[4];
  GemmEntry *velb_dev_[4] = {}, *preb_dev_[4] = {};
  int *velb_beg_[4] = {}, *preb_beg_[4] = {};
  GemmTable velb_tab_[4] = {}, preb_tab_[4] = {};
  bool preb_avec_ = false;
  int hist_period_ = 0, hist_steps_ = 0, hist_slots_ = 0;
  int64_t hist_count_ = 0, hist_committed_ = 0;
  double *hist_u_ = nullptr, *hist_p_ = nullptr;
  std::vector<double> hist_shift_;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  // stage launch shape (mixed form)
  struct StageGraph {
    int rounds, iters;
    hipGraph_t graph;
    hipGraphExec_t exec;
  };
  StageCtl* ctl_ = nullptr;       // device control block
  StageCtl* ctl_host_ = nullptr;  // pinned copy read after each launch
  int* run_ = nullptr;            // per segment: its inner CG is iterating
  const int* cg_run_ = nullptr;   // while capturing a stage graph: the inner CG's flag
  std::vector<StageGraph> stage_graphs_;
};

std::string g_batch_create_error;

int batch_fail(BatchSolver* s, int code, const std::string& msg) {
  s->error = msg;
  return code;
}

template <typename F>
int batch_guarded(BatchSolver* s, F&& body) {
  try {
    s->error.clear();
    body();
    return MAC_OK;
  } catch (const ArgFailure& e) {
    return batch_fail(s, MAC_INVALID_ARGUMENT, e.what());
  } catch (const SolveFailure& e) {
    return batch_fail(s, e.code, e.what());
  } catch (const std::exception& e) {
    return batch_fail(s, MAC_HIP_ERROR, e.what());
  }
}

}  // namespace
}  // namespace macs

using macs::BatchSolver;

extern "C" {

void* mac_batch_create(int32_t count, const int32_t* nx, const int32_t* ny, const double* lx, const double* ly,
                       int32_t ncols, const mac_stokes_opts* opts) {
  macs::g_batch_create_error.clear();
  mac_stokes_opts o;
  mac_stokes_default_opts(&o);
  if (opts) o = *opts;
  if (count < 1 || count > 4096 || ncols < 1 || ncols > 8 || !nx || !ny || !lx || !ly) {
    macs::g_batch_create_error = "Require 1 <= count <= 4096 cavities and 1 <= ncols <= 8";
    return nullptr;
  }
  for (int i = 0; i < count; ++i)
    if (nx[i] < 3 || ny[i] < 3 || nx[i] > 4096 || ny[i] > 4096 || !std::isfinite(lx[i]) || !std::isfinite(ly[i]) ||
        lx[i] <= 0 || ly[i] <= 0) {
      macs::g_batch_create_error = "Require 3 <= nx,ny <= 4096 and positive finite domain lengths (cavity " +
                                   std::to_string(i) + ")";
      return nullptr;
    }
  const bool mixed_ok = o.precision == MAC_PRECISION_MIXED && o.refine_tolerance > 0 &&
                        std::isfinite(o.refine_tolerance) && o.inner_tolerance > 0 && o.inner_tolerance < 1 &&
                        o.max_sweeps >= 1;
  if (!(o.tolerance > 0) || !(o.krylov_tolerance > 0) || !std::isfinite(o.tolerance) ||
      !std::isfinite(o.krylov_tolerance) || o.max_iterations < 1 ||
      (o.precision != MAC_PRECISION_FP64 && !mixed_ok)) {
    macs::g_batch_create_error = "Invalid solver controls";
    return nullptr;
  }
  try {
    return new BatchSolver(count, nx, ny, lx, ly, ncols, o);
  } catch (const std::exception& e) {
    macs::g_batch_create_error = e.what();
    return nullptr;
  }
}

const char* mac_batch_create_error(void) { return macs::g_batch_create_error.c_str(); }

void mac_batch_destroy(void* h) { delete static_cast<BatchSolver*>(h); }

const char* mac_batch_last_error(void* h) {
  BatchSolver* s = static_cast<BatchSolver*>(h);
  return s ? s->error.c_str() : "null handle";
}

int32_t mac_batch_layout(void* h, int64_t* velocity_offsets, int64_t* pressure_offsets) {
  BatchSolver* s = static_cast<BatchSolver*>(h);
  if (!s || !velocity_offsets || !pressure_offsets) return MAC_INVALID_ARGUMENT;
  const int n = s->count() * s->ncols();
  for (int i = 0; i < n; ++i) {
    velocity_offsets[i] = s->seg(i).voff;
    pressure_offsets[i] = s->seg(i).poff;
  }
  velocity_offsets[n] = s->total_faces();
  pressure_offsets[n] = s->total_cells();
  return MAC_OK;
}

int32_t mac_batch_set_history(void* h, int32_t period, int32_t steps) {
  BatchSolver* s = static_cast<BatchSolver*>(h);
  if (!s) return MAC_INVALID_ARGUMENT;
  return macs::batch_guarded(s, [&] {
    if (!s->mixed() || period < 0 || period > 16 || (period > 0 && (steps < 1 || steps > macs::kMaxExtrapolation)))
      throw macs::ArgFailure("Require the mixed form, 0 <= period <= 16 and 1 <= steps <= 6");
    HIP_CHECK(hipSetDevice(s->opts().device));
    s->set_history(period, period > 0 ? steps : 0);
  });
}

int32_t mac_batch_history_mark(void* h, int32_t rollback) {
  BatchSolver* s = static_cast<BatchSolver*>(h);
  if (!s) return MAC_INVALID_ARGUMENT;
  s->history_mark(rollback != 0);
  return MAC_OK;
}

int32_t mac_batch_last_sweeps(void* h) {
  BatchSolver* s = static_cast<BatchSolver*>(h);
  return s ? s->last_sweeps : -1;
}

int32_t mac_batch_set_launch(void* h, int32_t use_graph, int32_t iters_per_check) {
  BatchSolver* s = static_cast<BatchSolver*>(h);
  if (!s || iters_per_check < 1 || iters_per_check > 16 || use_graph < 0 || use_graph > 2 ||
      (use_graph == 2 && !s->mixed()))
    return MAC_INVALID_ARGUMENT;
  s->use_graph = use_graph;
  if (use_graph == 2) {
    s->stage_iters_ = iters_per_check;
    s->stage_rounds_ = BatchSolver::kStageRounds;
  }
  else s->iters_per_check = iters_per_check;
  return MAC_OK;
}

int32_t mac_stokes_solve_batched(void* h, const double* rhs, const double* mass, const double* viscosity,
                                 double* velocity, double* pressure, int32_t* iterations, double* residual,
                                 double* divergence_inf, int32_t device_pointers) {
  BatchSolver* s = static_cast<BatchSolver*>(h);
  if (!s) return MAC_INVALID_ARGUMENT;
  return macs::batch_guarded(s, [&] {
    const int nc = s->count();
    if (!rhs || !mass || !viscosity || !velocity || !pressure) throw macs::ArgFailure("Invalid Stokes RHS");
    for (int i = 0; i < nc; ++i)
      if (!std::isfinite(mass[i]) || !std::isfinite(viscosity[i]) || std::min(mass[i], viscosity[i]) < 0 ||
          mass[i] + viscosity[i] <= 0)
        throw macs::ArgFailure("Require mass,viscosity >= 0 and nonzero total (cavity " + std::to_string(i) + ")");
    HIP_CHECK(hipSetDevice(s->opts().device));
    hipStream_t st = s->stream();
    const int64_t V = s->total_faces(), P = s->total_cells();
    const double* f = rhs;
    double* vel = velocity;
    double* pr = pressure;
    if (!device_pointers) {
      // Pageable host arrays: blocking copies with the stream idle.
      HIP_CHECK(hipStreamSynchronize(st));
      HIP_CHECK(hipMemcpy(s->io_f(), rhs, sizeof(double) * V, hipMemcpyHostToDevice));
      f = s->io_f();
      vel = s->io_v();
      pr = s->io_p();
    }
    // Non-finite right-hand sides are rejected like the reference (ValueError); the mixed
    // form finds them in its first defect.
    if (!s->mixed())
      for (double m : s->rhs_max(f))
        if (!std::isfinite(m)) throw macs::ArgFailure("Invalid Stokes RHS");
    BatchSolver::Outcome out = s->mixed() ? s->solve_mixed(f, mass, viscosity, vel, pr)
                                          : s->solve(f, mass, viscosity, vel, pr);
    s->last_sweeps = out.sweeps;
    if (iterations)
      for (size_t i = 0; i < out.iterations.size(); ++i) iterations[i] = out.iterations[i];
    if (!out.failed_cg.empty())
      throw macs::SolveFailure(MAC_NOT_CONVERGED, "Pressure Schur CG failed with info=" +
                                                      std::to_string(s->opts().max_iterations) + " (cavity " +
                                                      std::to_string(out.failed_cg.front()) + ")");
    HIP_CHECK(hipStreamSynchronize(st));
    if (!device_pointers) {
      HIP_CHECK(hipMemcpy(velocity, vel, sizeof(double) * V, hipMemcpyDeviceToHost));
      HIP_CHECK(hipMemcpy(pressure, pr, sizeof(double) * P, hipMemcpyDeviceToHost));
    }
    for (int i = 0; i < nc; ++i) {
      if (residual) residual[i] = out.residual[i];
      if (divergence_inf) divergence_inf[i] = out.divergence[i];
    }
    const double tol = s->opts().tolerance;
    for (int i = 0; i < nc; ++i) {
      if (!(out.residual[i] <= tol) || !(out.divergence[i] <= tol * (1.0 + out.vmax[i]))) {
        char msg[256];
        std::snprintf(msg, sizeof(msg), "Stokes residual=%.3e, divergence=%.3e (cavity %d)", out.residual[i],
                      out.divergence[i], i);
        throw macs::SolveFailure(MAC_GATE_FAILED, msg);
      }
    }
  });
}

}  // extern "C"
