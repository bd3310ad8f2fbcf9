// Batched transform GEMMs for the DST/DCT diagonalisation, gfx950 MFMA.
//
// Every pass of a 2-D sine/cosine transform is C = epi(A . Bt^T) with the data as A
// (rows contiguous along the contraction index) and the orthonormal transform matrix
// as Bt.  One launch covers every block (u, v) of every right-hand-side column; the
// velocity shift / pressure eigenvalue divide and the preconditioner combination are
// fused into the epilogue, and the pass that must change the contiguous axis writes
// its tile transposed.
//
//   k_gemm_f64  : v_mfma_f64_16x16x4f64, FP64 in / FP64 out (the plain FP64 solve)
//   k_gemm_bf16 : v_mfma_f32_16x16x32_bf16, A converted to bf16 on the way into LDS,
//                 FP32 accumulation, output fp64 / fp32 / bf16 (the mixed solve)
#include <type_traits>

#include "internal.h"

namespace macs {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a contiguous run
// of tiles so tiles that share A row panels share one L2.
__device__ inline int xcd_tile(int bid, int total) {
  if ((total & 7) != 0) return bid;
  return (bid & 7) * (total >> 3) + (bid >> 3);
}

__device__ inline int find_entry(const GemmBatch& b, int t) {
  int e = 0;
#pragma unroll
  for (int i = 1; i < kMaxEntries; ++i)
    if (i < b.count && t >= b.e[i].tile_begin) e = i;
  return e;
}

__device__ inline double ld_as_double(const double* p, int64_t i) { return p[i]; }
__device__ inline double ld_as_double(const float* p, int64_t i) { return (double)p[i]; }
__device__ inline double ld_as_double(const __bf16* p, int64_t i) { return (double)(float)p[i]; }
__device__ inline void st_from_double(double* p, int64_t i, double v) { p[i] = v; }
__device__ inline void st_from_double(float* p, int64_t i, double v) { p[i] = (float)v; }
__device__ inline void st_from_double(__bf16* p, int64_t i, double v) { p[i] = (__bf16)(float)v; }

template <int EPI, typename TC>
__device__ inline void epilogue(const GemmBatch& b, const GemmEntry& E, int m, int n, double v) {
  if (m >= E.M || n >= E.N) return;
  const int64_t idx = b.trans_c ? (int64_t)n * E.ldc + m : (int64_t)m * E.ldc + n;
  if (EPI == EPI_SHIFT) {
    v = v / (b.s0 + b.s1 * (E.ea[m] + E.eb[n]));
  } else if (EPI == EPI_LAMBDA) {
    const double lam = E.ea[m] + E.eb[n];
    v = lam > b.thresh ? v / lam : 0.0;
  } else if (EPI == EPI_COMBINE) {
    v = b.s0 * v + b.s1 * ld_as_double(static_cast<const TC*>(E.aux), idx);
  } else if (EPI == EPI_ACCUM) {
    v = ld_as_double(static_cast<const TC*>(E.C), idx) + v;
  }
  st_from_double(static_cast<TC*>(E.C), idx, v);
}

// ------------------------------------------------------------------- FP64 GEMM
// 64x64 tile, 256 threads = 4 waves in 2x2, each wave 32x32 = 2x2 MFMA 16x16 blocks.
// K step 16, LDS double-buffered, global loads for step k+1 in flight during step k.
constexpr int F64_BM = 64, F64_BN = 64, F64_BK = 16, F64_LDS = 64 + 4;

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_f64(const GemmBatch batch) {
  __shared__ double As[2][F64_BK * F64_LDS];
  __shared__ double Bs[2][F64_BK * F64_LDS];
  const int t = xcd_tile(blockIdx.x, batch.total_tiles);
  const GemmEntry& E = batch.e[find_entry(batch, t)];
  const int lt = t - E.tile_begin;
  const int m0 = (lt / E.tiles_n) * F64_BM, n0 = (lt % E.tiles_n) * F64_BN;
  const double* __restrict__ A = static_cast<const double*>(E.A);
  const double* __restrict__ B = static_cast<const double*>(E.Bt);
  const int tid = threadIdx.x;
  const int lrow = tid >> 2, lk = (tid & 3) * 4;
  const int wave = tid >> 6, lane = tid & 63;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int li = lane & 15, lg = lane >> 4;
  const int M = E.M, N = E.N, K = E.K;
  const bool a_row = (m0 + lrow) < M, b_row = (n0 + lrow) < N;
  const double* arow = A + (int64_t)(m0 + lrow) * E.lda;
  const double* brow = B + (int64_t)(n0 + lrow) * E.ldb;

  d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  double ra[4], rb[4];
  const int nk = (K + F64_BK - 1) / F64_BK;

  auto gload = [&](int kt) {
    const int k = kt * F64_BK + lk;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      ra[i] = (a_row && k + i < K) ? arow[k + i] : 0.0;
      rb[i] = (b_row && k + i < K) ? brow[k + i] : 0.0;
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      As[buf][(lk + i) * F64_LDS + lrow] = ra[i];
      Bs[buf][(lk + i) * F64_LDS + lrow] = rb[i];
    }
  };

  gload(0);
  sstore(0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) gload(kt + 1);
#pragma unroll
    for (int kk = 0; kk < F64_BK; kk += 4) {
      const double* as = &As[cur][(kk + lg) * F64_LDS];
      const double* bs = &Bs[cur][(kk + lg) * F64_LDS];
      const double a0 = as[wm + li], a1 = as[wm + 16 + li];
      const double b0 = bs[wn + li], b1 = bs[wn + 16 + li];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    if (kt + 1 < nk) sstore(cur ^ 1);
    __syncthreads();
  }
  // v_mfma_f64_16x16x4f64 result layout: register r of lane l holds row l/16 + 4r,
  // column l%16 of the 16x16 block (the f32-result MFMAs use row 4*(l/16) + r).
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        epilogue<EPI, double>(batch, E, m0 + wm + i * 16 + lg + 4 * r, n0 + wn + j * 16 + li,
                              acc[i][j][r]);
}

// ------------------------------------------------------------------- BF16 GEMM
// 64x64 tile, 256 threads = 2x2 waves of 32x32 (2x2 MFMA 16x16x32 blocks each).
// LDS rows are k-contiguous for both operands (the bf16 MFMA takes 8 consecutive k
// per lane), padded by 8 elements to spread banks.
//
// SPLIT ("bf16x3"): every A element is split on load into hi = bf16(a) and
// lo = bf16(a - hi), the matrix is supplied as Q_hi + Q_lo, and each k-step issues
// A_hi.B_hi + A_hi.B_lo + A_lo.B_hi into the FP32 accumulator, which keeps about 16
// significant bits of both operands instead of 8.
constexpr int BF_BM = 64, BF_BN = 64;

template <bool SPLIT>
struct BfTile {
  static constexpr int BK = SPLIT ? 32 : 64;
  static constexpr int LDS = BK + 8;
  static constexpr int A_PER_THREAD = BF_BM * BK / 256;  // elements of A each thread stages
};

__device__ inline float to_f(double v) { return (float)v; }
__device__ inline float to_f(float v) { return v; }
__device__ inline float to_f(__bf16 v) { return (float)v; }

template <int EPI, typename TA, typename TC, bool SPLIT>
__global__ __launch_bounds__(256) void k_gemm_bf16(const GemmBatch batch) {
  using T = BfTile<SPLIT>;
  constexpr int BK = T::BK, LDS = T::LDS, NA = T::A_PER_THREAD;
  constexpr int NB = BF_BN * BK / 256;  // bf16 elements of B per thread (16 or 8)
  __shared__ __attribute__((aligned(16))) __bf16 As[2][(SPLIT ? 2 : 1) * BF_BM * LDS];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[2][(SPLIT ? 2 : 1) * BF_BN * LDS];
  const int t = xcd_tile(blockIdx.x, batch.total_tiles);
  const GemmEntry& E = batch.e[find_entry(batch, t)];
  const int lt = t - E.tile_begin;
  const int m0 = (lt / E.tiles_n) * BF_BM, n0 = (lt % E.tiles_n) * BF_BN;
  const TA* __restrict__ A = static_cast<const TA*>(E.A);
  const __bf16* __restrict__ B = static_cast<const __bf16*>(E.Bt);
  const __bf16* __restrict__ Blo = static_cast<const __bf16*>(E.Bt_lo);
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int li = lane & 15, lg = lane >> 4;
  const int M = E.M, N = E.N, K = E.K;
  // B staging: 4 threads per row, NB contiguous bf16 each (one or two 16-byte loads).
  const int brow_i = tid >> 2, bk = (tid & 3) * NB;
  const bool b_row = (n0 + brow_i) < N;
  const __bf16* brow = B + (int64_t)(n0 + brow_i) * E.ldb;
  const __bf16* brow_lo = SPLIT ? Blo + (int64_t)(n0 + brow_i) * E.ldb : nullptr;
  constexpr bool A_VEC = std::is_same<TA, __bf16>::value && !SPLIT;

  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  float ra[NA];
  __attribute__((aligned(16))) __bf16 rav[16];
  __attribute__((aligned(16))) __bf16 rb[NB];
  __attribute__((aligned(16))) __bf16 rbl[SPLIT ? NB : 1];
  const int nk = (K + BK - 1) / BK;

  auto load_b = [&](const __bf16* row, int k, __bf16* dst) {
    if (b_row && k + NB <= K) {
      const uint4* p = reinterpret_cast<const uint4*>(row + k);
#pragma unroll
      for (int v = 0; v < NB / 8; ++v) reinterpret_cast<uint4*>(dst)[v] = p[v];
    } else {
#pragma unroll
      for (int i = 0; i < NB; ++i) dst[i] = (b_row && k + i < K) ? row[k + i] : (__bf16)0.0f;
    }
  };
  auto gload = [&](int kt) {
    const int k0 = kt * BK;
    if constexpr (A_VEC) {
      // bf16 rows (ld % 8 == 0, 16-byte aligned): 4 threads x 16 contiguous elements.
      const int arow_i = tid >> 2, ak = (tid & 3) * 16;
      const bool ok = (m0 + arow_i) < M;
      const __bf16* row = reinterpret_cast<const __bf16*>(A) + (int64_t)(m0 + arow_i) * E.lda;
      const int k = k0 + ak;
      if (ok && k + 16 <= K) {
        const uint4* p = reinterpret_cast<const uint4*>(row + k);
        reinterpret_cast<uint4*>(rav)[0] = p[0];
        reinterpret_cast<uint4*>(rav)[1] = p[1];
      } else {
#pragma unroll
        for (int i = 0; i < 16; ++i) rav[i] = (ok && k + i < K) ? row[k + i] : (__bf16)0.0f;
      }
    } else {
      // Wide or split A: consecutive threads read consecutive k of one row (coalesced).
#pragma unroll
      for (int e = 0; e < NA; ++e) {
        const int L = e * 256 + tid, r = L / BK, c = L % BK;
        const bool ok = (m0 + r) < M && (k0 + c) < K;
        ra[e] = ok ? to_f(A[(int64_t)(m0 + r) * E.lda + k0 + c]) : 0.f;
      }
    }
    load_b(brow, k0 + bk, rb);
    if constexpr (SPLIT) load_b(brow_lo, k0 + bk, rbl);
  };
  auto sstore = [&](int buf) {
    if constexpr (A_VEC) {
      const int arow_i = tid >> 2, ak = (tid & 3) * 16;
      uint4* da = reinterpret_cast<uint4*>(&As[buf][arow_i * LDS + ak]);
      da[0] = reinterpret_cast<const uint4*>(rav)[0];
      da[1] = reinterpret_cast<const uint4*>(rav)[1];
    } else {
#pragma unroll
      for (int e = 0; e < NA; ++e) {
        const int L = e * 256 + tid, r = L / BK, c = L % BK;
        const __bf16 hi = (__bf16)ra[e];
        As[buf][r * LDS + c] = hi;
        if constexpr (SPLIT) As[buf][BF_BM * LDS + r * LDS + c] = (__bf16)(ra[e] - (float)hi);
      }
    }
    uint4* db = reinterpret_cast<uint4*>(&Bs[buf][brow_i * LDS + bk]);
#pragma unroll
    for (int v = 0; v < NB / 8; ++v) db[v] = reinterpret_cast<const uint4*>(rb)[v];
    if constexpr (SPLIT) {
      uint4* dbl = reinterpret_cast<uint4*>(&Bs[buf][BF_BN * LDS + brow_i * LDS + bk]);
#pragma unroll
      for (int v = 0; v < NB / 8; ++v) dbl[v] = reinterpret_cast<const uint4*>(rbl)[v];
    }
  };

  gload(0);
  sstore(0);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) gload(kt + 1);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 32) {
      const __bf16* as = &As[cur][0];
      const __bf16* bs = &Bs[cur][0];
      const bf8 a0 = *reinterpret_cast<const bf8*>(&as[(wm + li) * LDS + kk + lg * 8]);
      const bf8 a1 = *reinterpret_cast<const bf8*>(&as[(wm + 16 + li) * LDS + kk + lg * 8]);
      const bf8 b0 = *reinterpret_cast<const bf8*>(&bs[(wn + li) * LDS + kk + lg * 8]);
      const bf8 b1 = *reinterpret_cast<const bf8*>(&bs[(wn + 16 + li) * LDS + kk + lg * 8]);
      if constexpr (SPLIT) {
        const bf8 l0 = *reinterpret_cast<const bf8*>(&as[BF_BM * LDS + (wm + li) * LDS + kk + lg * 8]);
        const bf8 l1 = *reinterpret_cast<const bf8*>(&as[BF_BM * LDS + (wm + 16 + li) * LDS + kk + lg * 8]);
        const bf8 c0 = *reinterpret_cast<const bf8*>(&bs[BF_BN * LDS + (wn + li) * LDS + kk + lg * 8]);
        const bf8 c1 = *reinterpret_cast<const bf8*>(&bs[BF_BN * LDS + (wn + 16 + li) * LDS + kk + lg * 8]);
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(l0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(l0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(l1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(l1, b1, acc[1][1], 0, 0, 0);
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, c0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, c1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, c0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, c1, acc[1][1], 0, 0, 0);
      }
      acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc[1][1], 0, 0, 0);
    }
    if (kt + 1 < nk) sstore(cur ^ 1);
    __syncthreads();
  }
  // f32-result MFMA layout: register r of lane l holds row 4*(l/16) + r, column l%16.
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        epilogue<EPI, TC>(batch, E, m0 + wm + i * 16 + 4 * lg + r, n0 + wn + j * 16 + li,
                          (double)acc[i][j][r]);
}

int tile_batch(GemmBatch& b, int bm, int bn) {
  int total = 0;
  for (int i = 0; i < b.count; ++i) {
    GemmEntry& e = b.e[i];
    e.tiles_m = (e.M + bm - 1) / bm;
    e.tiles_n = (e.N + bn - 1) / bn;
    e.tile_begin = total;
    total += e.tiles_m * e.tiles_n;
  }
  b.total_tiles = total;
  return total;
}

template <typename TA, typename TC, bool SPLIT>
void launch_bf16(GemmBatch& b, int total, hipStream_t s) {
  switch (b.epi) {
    case EPI_STORE: k_gemm_bf16<EPI_STORE, TA, TC, SPLIT><<<total, 256, 0, s>>>(b); break;
    case EPI_SHIFT: k_gemm_bf16<EPI_SHIFT, TA, TC, SPLIT><<<total, 256, 0, s>>>(b); break;
    case EPI_LAMBDA: k_gemm_bf16<EPI_LAMBDA, TA, TC, SPLIT><<<total, 256, 0, s>>>(b); break;
    case EPI_ACCUM: k_gemm_bf16<EPI_ACCUM, TA, TC, SPLIT><<<total, 256, 0, s>>>(b); break;
    default: k_gemm_bf16<EPI_COMBINE, TA, TC, SPLIT><<<total, 256, 0, s>>>(b); break;
  }
}

template <typename TA, bool SPLIT>
void launch_bf16_c(GemmBatch& b, int c_type, int total, hipStream_t s) {
  if (c_type == DT_F64) launch_bf16<TA, double, SPLIT>(b, total, s);
  else if (c_type == DT_F32) launch_bf16<TA, float, SPLIT>(b, total, s);
  else launch_bf16<TA, __bf16, SPLIT>(b, total, s);
}

}  // namespace

void gemm_f64(GemmBatch& b, hipStream_t s) {
  if (b.count <= 0) return;
  const int total = tile_batch(b, F64_BM, F64_BN);
  switch (b.epi) {
    case EPI_STORE: k_gemm_f64<EPI_STORE><<<total, 256, 0, s>>>(b); break;
    case EPI_SHIFT: k_gemm_f64<EPI_SHIFT><<<total, 256, 0, s>>>(b); break;
    case EPI_LAMBDA: k_gemm_f64<EPI_LAMBDA><<<total, 256, 0, s>>>(b); break;
    default: k_gemm_f64<EPI_COMBINE><<<total, 256, 0, s>>>(b); break;
  }
}

void gemm_bf16(GemmBatch& b, int a_type, int c_type, bool split, hipStream_t s) {
  if (b.count <= 0) return;
  const int total = tile_batch(b, BF_BM, BF_BN);
  if (split) {
    // Split operands come from FP32/FP64 data (a bf16 input has no low part).
    if (a_type == DT_F64) launch_bf16_c<double, true>(b, c_type, total, s);
    else launch_bf16_c<float, true>(b, c_type, total, s);
  } else {
    if (a_type == DT_F64) launch_bf16_c<double, false>(b, c_type, total, s);
    else if (a_type == DT_F32) launch_bf16_c<float, false>(b, c_type, total, s);
    else launch_bf16_c<__bf16, false>(b, c_type, total, s);
  }
}

}  // namespace macs
