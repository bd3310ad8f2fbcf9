// Internal declarations shared by the gfx950 kernels and the C-ABI host code.
// Nothing here is part of the public interface (include/macstokes.h is).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace macs {

// ----------------------------------------------------------------- GEMM batch
// One entry computes C = epi(A . Bt^T): A is [M][K] row-major (lda), Bt is [N][K]
// row-major (ldb), so B[k][n] = Bt[n][k].  C is stored [M][N] (ldc) or, with
// trans_c, as its transpose [N][M] (ldc).  Every 2-D DST/DCT pass is one such
// product with the transform matrix as Bt (see transforms in stokes.hip).
enum Epilogue : int {
  EPI_STORE = 0,    // C = acc
  EPI_SHIFT = 1,    // C = acc / (s0 + s1 * (ea[m] + eb[n]))        velocity shift
  EPI_LAMBDA = 2,   // lam = ea[m] + eb[n]; C = lam > thresh ? acc/lam : 0   (D D^T)^+
  EPI_COMBINE = 3,  // C = s0 * acc + s1 * aux                          preconditioner
  EPI_ACCUM = 4,    // C = C + acc                                      defect correction
};
enum DType : int { DT_F64 = 0, DT_F32 = 1, DT_BF16 = 2 };

struct CgState;

struct GemmEntry {
  const void* A;
  const void* Bt;
  const void* Bt_lo;   // split mode: low bf16 part of the matrix (same layout as Bt)
  void* C;
  const void* aux;     // EPI_COMBINE: same layout/type as C
  const double* ea;    // eigenvalues indexed by m
  const double* eb;    // eigenvalues indexed by n
  int M, N, K;
  int lda, ldb, ldc;
  int tiles_m, tiles_n, tile_begin;
  double s0, s1;       // EPI_SHIFT / EPI_COMBINE scalars (the launcher copies the batch's)
  double thresh;       // EPI_LAMBDA zero-mode threshold (likewise)
  const CgState* st;   // table launches with use_state: skip the entry if !st->active
  int pair_n;          // paired forward pass (bf16 split): A's row length n, whose column k
  int pair_sign;       // enters as a[k] + pair_sign * a[n - 1 - k]; pair_sign 0: not paired
};

constexpr int kMaxEntries = 8;
struct GemmBatch {
  GemmEntry e[kMaxEntries];
  int count;
  int total_tiles;
  int epi;
  int trans_c;
  int group;        // tile-rows per visiting group (set by the launcher)
  double s0, s1, thresh;
};

// Device-resident entry table for grouped launches over many blocks (a batch of
// cavities): entries carry their own scalars and tile_begin (ascending, set by
// tile_table), `begin` mirrors tile_begin for the per-block search.
struct GemmTable {
  const GemmEntry* e;
  const int* begin;
  int count, total_tiles, epi, trans_c, use_state;
  const int* run;  // device flag (may be null): a launch that reads 0 there does nothing
};

// Tile edge used by the launchers (entries are tiled by the launcher itself).
void gemm_f64(GemmBatch& b, hipStream_t s);
// FP64 grouped launch over a device table; tile_table fills tiles_m/n and tile_begin
// of host entries for the FP64 tile shape and returns the total tile count.
void gemm_f64_table(const GemmTable& t, hipStream_t s);
int tile_table_f64(GemmEntry* entries, int count, int* begins);
// a_type: DT_F64 / DT_F32 / DT_BF16 for A and aux/C per c_type; Bt is always bf16.
// split: three-product hi/lo form (A must be DT_F32 or DT_F64, Bt_lo set).
void gemm_bf16(GemmBatch& b, int a_type, int c_type, bool split, hipStream_t s);
// Split-bf16 (bf16x3, FP32 accumulation) grouped launch over a device table with 64x64
// tiles, the tiling tile_table_f64 gives.  Supported: FP64 -> FP32 store (a_vec: even
// lda and 16-byte aligned rows, else element loads), FP32 -> FP32 store / shift / lambda
// and FP32 -> FP64 store / combine (a_vec required: lda % 4 == 0 and 16-byte aligned
// rows).  Every entry carries Bt_lo.  Returns 0, or -1 for an unsupported combination.
int gemm_bf16_table(const GemmTable& t, int a_type, int c_type, bool a_vec, hipStream_t s);
// Events that the next gemm_f64 / gemm_bf16 launch records at its kernel's own start and
// end (hipExtLaunchKernelGGL), consumed by that launch; null for plain launches.
void set_gemm_events(hipEvent_t start, hipEvent_t stop);
// Tile shape (0 auto, 64, 12864, 128) and tile-row grouping of the bf16 GEMM.
void set_bf16_tiling(int tile, int group);

// ------------------------------------------------------------- stencil kernels
// Internal velocity layout: u block [ny][ldu], v block [ny-1][ldv] (ld >= row length),
// at offsets 0 and voff of each column; column stride vstride.  The reference's
// packed layout is the special case ldu = nx-1, ldv = nx, voff = nu.
struct VelLayout {
  int ldu, ldv;
  int64_t voff, stride;
};

struct GridDims {
  int nx, ny;
  double hx, hy;
};

// Per-column Schur-CG scalars, resident in device memory so the whole recurrence
// stays on the GPU; the host only reads the `active` flags once per iteration.
struct CgState {
  double bnorm, tol, rr, rho, rho_prev, alpha, beta;
  double mtol;  // > 0: the column also stops once max|r| <= mtol (posted checks only)
  int active, iters, converged, pad1;
};

constexpr int kThreads = 256;
// Blocks per column used by every reduction (fixed, so summation order is fixed).
int reduce_blocks(int64_t n);
// Partial-result blocks of defect_max per quantity (three quantities follow each other).
int defect_blocks(int64_t size, int ncols);

enum Reduce : int { RED_NONE = 0, RED_SUM = 1, RED_MAXABS = 2 };

// ---- device helpers shared by the vector kernels
// |x| with NaN mapped to +inf, so max-norm checks catch non-finite values.
__device__ inline double absn(double x) { return x != x ? (double)INFINITY : fabs(x); }

// One face of the FP64 defect mass u + visc (K u) + G p - f.  K u is the five-point row
// summed south, west, centre, east, north (a neighbour's coefficient is -1/h^2; absent
// neighbours pass nothing), with the FMAs written out, so every kernel that evaluates a
// defect rounds alike whatever the compiler would contract.
struct FaceRow {
  double kv = 0.0;
  __device__ inline void add(double coef, double value) { kv = fma(coef, value, kv); }
};
__device__ inline double face_defect(double mass, double vv, double visc, double kv, double lift, double fv) {
  return (fma(visc, kv, mass * vv) - lift) - fv;
}

// n / d for 0 <= n < 2^31 and 1 <= d < 2^15, from one FP64 product with the reciprocal
// (error far below 1, so the truncated quotient is off by at most one) and a correction
// step: a few VALU operations instead of an integer division sequence.
struct SmallDiv {
  int d;
  double inv;
  __device__ explicit SmallDiv(int d_) : d(d_), inv(1.0 / (double)d_) {}
  __device__ inline int quot(int n) const {
    int q = (int)((double)n * inv);
    const int r = n - q * d;
    if (r < 0) --q;
    else if (r >= d) ++q;
    return q;
  }
};

// 256-thread block reduction in a fixed order; result valid in thread 0.
template <int MODE>
__device__ inline double block_reduce(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_down(v, off, 64);
    v = MODE == RED_MAXABS ? fmax(v, o) : v + o;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    v = MODE == RED_MAXABS ? fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]))
                           : (sh[0] + sh[1]) + (sh[2] + sh[3]);
  }
  return v;
}


// Finalize target of a reduction: when `out` is set, the launcher also reduces the
// partials it produced into out[c] (finalize(), fixed order).
struct Fin {
  double* out = nullptr;
};

// q[c] = add[c] + sign * D vel[c] (add optional); optional per-block partial sums /
// max|.| of q.
void divergence(int in_type, const void* vel, const VelLayout& L, const GridDims& g, double* out,
                int64_t out_stride, double sign, int ncols, int reduce, double* partials,
                const CgState* st, hipStream_t s, const double* add = nullptr, Fin fin = {});
// out[c] = add[c] + D^T (p[c] - mean[c]) (add / mean optional), out in layout L.
void divergence_transpose(const double* p, int64_t p_stride, const double* mean_sums, const double* add,
                          int64_t add_stride, int out_type, void* out, const VelLayout& L,
                          const GridDims& g, int ncols, const CgState* st, hipStream_t s);
// partials of sum(x) (mode RED_SUM, y == null), sum(x*y) (RED_SUM, y != null), max|x|.
void reduce(const double* x, const double* y, int64_t n, int64_t stride, int ncols, int mode,
            double* partials, const CgState* st, hipStream_t s, Fin fin = {});
// out[c] = sum / max over the partials of column c (fixed order).
void finalize(const double* partials, int nblk, int ncols, int mode, double* out, hipStream_t s);
// x[c] -= sums[c]/n; optional partial sums of y*x afterwards.
void sub_mean_dot(double* x, int64_t n, int64_t stride, int ncols, const double* sums, const double* y,
                  double* partials, const CgState* st, hipStream_t s, Fin fin = {});
void fill(double* x, int64_t n, double v, hipStream_t s);
// x = theta g, r = b - theta q with theta = gb/gq per column (x holds g, q holds S g).
void theta_start(double* x, double* r, const double* b, const double* q, const double* gb, const double* gq,
                 int64_t n, int64_t stride, int ncols, const CgState* st, hipStream_t s);
// x = a g + beta c, r = b - a bg - beta q with (a, beta) minimising ||r||_2 per column,
// from dots = [(bg,bg) (bg,q) (q,q) (bg,b) (q,b)] x ncols (columns of n contiguous
// elements; inactive columns skipped).
void galerkin2_start(double* x, double* r, const double* g, const double* c, const double* b, const double* bg,
                     const double* q, const double* dots, int64_t n, int ncols, const CgState* st, hipStream_t s);
// out = a - b column-wise (skips inactive columns when st is given).
void sub_cols(const double* a, const double* b, double* out, int64_t n, int64_t stride, int ncols,
              const CgState* st, hipStream_t s);
// dst_k[c * s_k + i] = src[c * n + i] - sums[c] / n (sums null: the plain copy) for the
// non-null destinations, one pass over the contiguous columns of src.
void finish_cols(const double* src, int64_t n, int ncols, const double* sums, double* d0, int64_t s0, double* d1,
                 int64_t s1, double* d2, int64_t s2, hipStream_t s);
void copy_cols(const double* src, int64_t src_stride, double* dst, int64_t dst_stride, int64_t n, int ncols,
               hipStream_t s);

// Host-visible result slot in coherent pinned memory.  A post kernel writes the values,
// fences at system scope, then stores `seq`; the host spins on `seq` instead of issuing a
// device-to-host copy and a stream synchronisation (one round trip per CG iteration).
struct Mailbox {
  int seq;
  int pad[15];
  int flags[64];
  int iters[64];
  int converged[64];
  double vals[4 + 64];
};
// cg_check (ncols <= 64), then the columns' active / iteration / converged state into mb.
void cg_check_post(CgState* st, double* rr, int* flags, int ncols, Mailbox* mb, int seq, hipStream_t s,
                   const double* rr_partials = nullptr, int nblk = 0);
// mb->vals = [a[0..na), b[0..nb)] (na + nb <= 68), then mb->seq = seq.
void post_values(const double* a, int na, const double* b, int nb, Mailbox* mb, int seq, hipStream_t s);

// CG scalar steps (one thread per column).
// rr_out (optional) also receives bb, the first recurrence residual (r, r).  mtol > 0: a
// column also stops, at a posted check whose partials carry max|r| (the mixed update's),
// once max|r| <= mtol.
void cg_init(CgState* st, const double* bb, const double* bmax, double rtol, int ncols, hipStream_t s,
             double* rr_out = nullptr, double mtol = 0.0);
// A mixed sweep's inner-CG start on b (n per column, contiguous): b -= sums[c] / n,
// r = b, x = 0; (b, b) into bb[c] and max|b| into bmax[c] (partials: 2 * ncols *
// reduce_blocks(n) entries); r32 (optional) receives r rounded to FP32.
void sweep_start(double* b, double* r, double* x, int64_t n, int ncols, const double* sums, double* partials,
                 double* bb, double* bmax, hipStream_t s, float* r32 = nullptr);
void cg_check(CgState* st, const double* rr, int* flags, int ncols, hipStream_t s);
// mixed_rho's scalar step alone, from partial sums of z, z r and r that another pass
// wrote (gx per column and quantity).
void mixed_rho_partials(CgState* st, const double* partials, int gx, int64_t n, int ncols, double* zmean,
                        hipStream_t s);
// sweep_start, then in one launch per column what the finalize of its partials, cg_init
// (rtol, rr_out, mtol) and the first cg_check do: the same values in the same order.
void sweep_start_cg(double* b, double* r, double* x, int64_t n, int ncols, const double* sums, double* partials,
                    double* bb, double* bmax, CgState* st, double rtol, double* rr_out, double mtol, int* flags,
                    hipStream_t s, float* r32 = nullptr);
// out = sum_k w[k] src[k] over n elements, 1 <= count <= kMaxExtrapolation (the time
// extrapolation of recorded solutions, and differences of two fields).
constexpr int kMaxExtrapolation = 6;
struct Extrapolation {
  const double* src[kMaxExtrapolation];
  double w[kMaxExtrapolation];
  int count;
};
void extrapolate(const double* const* src, const double* w, int count, double* out, int64_t n, hipStream_t s);
void cg_rho(CgState* st, const double* rz, int ncols, hipStream_t s);
void cg_alpha(CgState* st, const double* pq, int ncols, hipStream_t s);
// The same steps from the unfinalized partials of (r, z) / (p, q) (nblk per column, as
// the reductions leave them); out receives the finalized dots.
void cg_rho_fin(CgState* st, const double* partials, int nblk, double* out, int ncols, hipStream_t s);
void cg_alpha_fin(CgState* st, const double* partials, int nblk, double* out, int ncols, hipStream_t s);
// Mixed-CG steps that fold the mean removals of z and q into their consumers: rho from
// sum(z), sum(z r), sum(r) (zmean out); p = (z - zmean) + beta p; alpha from the
// three sums of divergence_sums3 (qmean out); x += alpha p, r -= alpha (q - qmean) with
// (r, r) partials.  All over n-element columns of stride n.
void mixed_rho(const double* z, const double* r, int64_t n, int ncols, CgState* st, double* partials,
               double* zmean, hipStream_t s);
void mixed_update_p(double* p, const double* z, const double* zmean, int64_t n, int ncols, const CgState* st,
                    hipStream_t s);
void mixed_alpha(const double* partials, int64_t n, int ncols, CgState* st, double* qmean, hipStream_t s);
// r32 (optional) receives the updated r rounded to FP32.
void mixed_update_xr(double* x, double* r, const double* p, const double* q, const double* qmean, int64_t n,
                     int ncols, const CgState* st, double* partials, hipStream_t s, float* r32 = nullptr);
// out = sign * D vel (no mean removal) with partials of sum(out), sum(out y), sum(y).
void divergence_sums3(int in_type, const void* vel, const VelLayout& L, const GridDims& g, double* out,
                      int64_t out_stride, double sign, const double* y, int ncols, double* partials,
                      const CgState* st, hipStream_t s);
void cg_update_p(double* p, const double* z, int64_t n, int64_t stride, int ncols, const CgState* st,
                 hipStream_t s);
void cg_update_xr(double* x, double* r, const double* p, const double* q, int64_t n, int64_t stride,
                  int ncols, const CgState* st, double* partials, hipStream_t s, Fin fin = {});

// Final-solve checks over all columns: partial max|defect|, max|f|, max|vel| where
// defect = mass*vel + visc*K vel - D^T p - f (vel, f packed; p mean-zero).
// With fin.out set, fin.out[0..2] receive the three maxima directly.
// With div_out and/or div_fin the same launch forms div_sign * (D vel) per cell, bitwise
// as divergence(): written to div_out (np per column) when given, per-column max|.|
// into div_fin.  partials3 then needs 4 * defect_blocks entries.
// residual_out (optional) receives r_u = -defect as residual_type (DT_F64, DT_F32 or
// DT_BF16) in residual_layout (the packed layout when null).
void defect_max(const double* vel, const double* f, const double* p, const GridDims& g, int ncols,
                double mass, double visc, double* partials3, hipStream_t s, void* residual_out = nullptr,
                Fin fin = {}, double* div_out = nullptr, double div_sign = 1.0, Fin div_fin = {},
                int residual_type = DT_F64, const VelLayout* residual_layout = nullptr);
// x[c] += y[c] (FP64, contiguous columns).
// du[c] += a_c w[c] (FP32) on the columns whose CG runs: a_c = alpha of the CG state
// (gb == gq == nullptr), gb[c]/gq[c] (0 when gq[c] <= 0; the Galerkin start), or 1 (gb
// == nullptr, gq != nullptr; a plain start).
void accum_velocity(float* du, const float* w, int64_t n, int64_t stride, int ncols, const CgState* st,
                    const double* gb, const double* gq, hipStream_t s);
// x += y (FP64 += FP32), n elements.
void add_f32(double* x, const float* y, int64_t n, hipStream_t s);
void add_cols(double* x, const double* y, int64_t n, int ncols, hipStream_t s);
// x += y over nx elements and u += w (FP64 += FP32) over nu, in one launch: the
// arithmetic of add_cols and add_f32.
void add_pair(double* x, const double* y, int64_t nx, double* u, const float* w, int64_t nu, hipStream_t s);

// Learned warm start (cnn.hip): grid sizes it serves, input preparation (pool by 4,
// normalise) and the guess (network forward, bilinear up by 4, rescale).
bool warm_shape_ok(int nx, int ny);
void warm_prepare(const double* b, int64_t b_stride, const double* pprev, int64_t p_stride, const double* target,
                  int64_t t_stride, int ncols, int nx, int ny, double alpha, float* x, float* t, double* s,
                  hipStream_t st);
void warm_guess(void* cnn, const double* b, int64_t b_stride, const double* pprev, int64_t p_stride, int ncols,
                int nx, int ny, double alpha, float* x, float* y, double* s, double* guess, int64_t g_stride,
                hipStream_t st);

}  // namespace macs
