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
#include <hip/hip_ext.h>

#include <cstdint>
#include <stdexcept>
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
__device__ inline void epilogue(const GemmEntry& E, int trans_c, int m, int n, double v) {
  if (m >= E.M || n >= E.N) return;
  const int64_t idx = trans_c ? (int64_t)n * E.ldc + m : (int64_t)m * E.ldc + n;
  if (EPI == EPI_SHIFT) {
    v = v / (E.s0 + E.s1 * (E.ea[m] + E.eb[n]));
  } else if (EPI == EPI_LAMBDA) {
    const double lam = E.ea[m] + E.eb[n];
    v = lam > E.thresh ? v / lam : 0.0;
  } else if (EPI == EPI_COMBINE) {
    v = E.s0 * v + E.s1 * ld_as_double(static_cast<const TC*>(E.aux), idx);
  } else if (EPI == EPI_ACCUM) {
    v = ld_as_double(static_cast<const TC*>(E.C), idx) + v;
  }
  st_from_double(static_cast<TC*>(E.C), idx, v);
}

// ------------------------------------------------------------------- FP64 GEMM
// 64x64 tile, 256 threads = 4 waves in 2x2, each wave 32x32 = 2x2 MFMA 16x16 blocks.
// K step 16, LDS double-buffered, global loads for step k+1 in flight during step k.
constexpr int F64_BM = 64, F64_BN = 64, F64_BK = 16, F64_LDS = 64 + 4;

// One 64x64 output tile (index lt within entry E) of C = epi(A . Bt^T).
template <int EPI>
__device__ inline void f64_tile(const GemmEntry& E, int lt, int trans_c) {
  __shared__ double As[2][F64_BK * F64_LDS];
  __shared__ double Bs[2][F64_BK * F64_LDS];
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
        epilogue<EPI, double>(E, trans_c, m0 + wm + i * 16 + lg + 4 * r, n0 + wn + j * 16 + li,
                              acc[i][j][r]);
}

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_f64(const GemmBatch batch) {
  const int t = xcd_tile(blockIdx.x, batch.total_tiles);
  const GemmEntry& E = batch.e[find_entry(batch, t)];
  f64_tile<EPI>(E, t - E.tile_begin, batch.trans_c);
}

// Grouped launch over a device table: each block finds its entry by binary search
// over the ascending tile_begin array; entries of converged columns are skipped.
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_f64_table(const GemmTable tb) {
  if (tb.run != nullptr && *tb.run == 0) return;
  const int t = xcd_tile(blockIdx.x, tb.total_tiles);
  int lo = 0, hi = tb.count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tb.begin[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  const GemmEntry E = tb.e[lo];
  if (tb.use_state && E.st != nullptr && !E.st->active) return;
  f64_tile<EPI>(E, t - E.tile_begin, tb.trans_c);
}

// ------------------------------------------------------------------- BF16 GEMM
// BM x BN tile (128x128 for large batches, 64x64 otherwise), K step 32 (one MFMA k-step
// of v_mfma_f32_16x16x32_bf16), 256 threads = 2x2 waves, each wave (BM/2) x (BN/2), or
// 512 threads = 2x4 waves of (BM/2) x (BN/4).
// LDS rows are k-contiguous for both operands (the MFMA takes 8 consecutive k per
// lane), 64 bytes with XOR-permuted 16-byte chunks (`lx`); double-buffered, the global
// loads of step k+1 are in flight while step k computes.
//
// SPLIT ("bf16x3"): each A element is split on its way into LDS into hi = bf16(a) and
// lo = bf16(a - hi), the matrix arrives as Q_hi + Q_lo, and every k-step issues
// A_lo.B_hi + A_hi.B_lo + A_hi.B_hi into the FP32 accumulator: about 16 significant
// bits of both operands instead of 8.
//
// Tiles are visited in groups of 4 tile-rows (all tile-columns of a group before the
// next), and consecutive tiles go to the same XCD, so each XCD's L2 holds a few A
// row panels and the B panels they meet.
// Element offset of (row r, k column c) in an LDS operand part: 64-byte rows (32 bf16,
// no padding), the four 16-byte chunks of row r permuted by chunk ^ ((r >> 1) & 3).
// With 32 four-byte banks, the 8 lanes of a 16-byte access pass then hit distinct banks
// both when they read one chunk of 8 consecutive rows (MFMA fragments) and when they
// store all four chunks of 2 rows (staging).  Measured: SQ_LDS_BANK_CONFLICT from 50% of
// the LDS-active cycles (rows padded to 40 elements) to 0.
__device__ inline int lx(int r, int c) {
  return r * 32 + ((((c >> 3) ^ (r >> 1)) & 3) << 3) + (c & 7);
}

// NW waves per block: 4 as a 2x2 grid of (BM/2)x(BN/2) wave tiles, or 8 as a 2x4 grid
// of (BM/2)x(BN/4) (two waves per SIMD on one tile: the A panel is staged once for
// twice the waves).
template <int BM_, int BN_, bool SPLIT_, int NW_ = 4>
struct Cfg {
  static constexpr int BM = BM_, BN = BN_, BK = 32, LDS = 32, NW = NW_, NT = 64 * NW_;
  static constexpr int WGN = NW_ == 8 ? 4 : 2, WGM = NW_ / WGN;  // wave grid
  static constexpr int FM = BM / WGM / 16, FN = BN / WGN / 16;   // 16x16 MFMA blocks per wave
  static constexpr int AE = BM * BK / NT;              // A elements staged per thread
  static constexpr int BE = BN * BK / NT;              // B elements staged per thread
  static constexpr int PARTS = SPLIT_ ? 2 : 1;
};

// Eight bf16 from p as one 128-bit word, elements at or past `valid` zero (a k tail).
__device__ inline uint4 load8_tail(const __bf16* p, int valid) {
  unsigned h[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) h[q] = q < valid ? (unsigned)__builtin_bit_cast(unsigned short, p[q]) : 0u;
  return make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16), h[6] | (h[7] << 16));
}

__device__ inline float to_f(double v) { return (float)v; }
__device__ inline float to_f(float v) { return v; }
__device__ inline float to_f(__bf16 v) { return (float)v; }

template <typename TC>
__device__ inline void store4(TC* p, const f4& v);
template <>
__device__ inline void store4<float>(float* p, const f4& v) {
  *reinterpret_cast<f4*>(p) = v;
}
template <>
__device__ inline void store4<__bf16>(__bf16* p, const f4& v) {
  typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
  bf4 o;
  o[0] = (__bf16)v[0]; o[1] = (__bf16)v[1]; o[2] = (__bf16)v[2]; o[3] = (__bf16)v[3];
  *reinterpret_cast<bf4*>(p) = o;
}
template <>
__device__ inline void store4<double>(double* p, const f4& v) {
  p[0] = v[0]; p[1] = v[1]; p[2] = v[2]; p[3] = v[3];
}

// AMODE: how A reaches LDS.  0 = bf16 rows (16-byte loads of 8), 1 = fp32 rows with
// lda % 4 == 0 and 16-byte aligned bases (float4 loads), 2 = one element per load,
// 3 = fp64 rows with even lda and 16-byte aligned bases (two double2 loads per 4).
// PAIRED (forward passes of the even/odd split): the A element at k is staged as
// a[k] + s a[n - 1 - k] (n = E.pair_n, s = E.pair_sign = +1 or -1), combined in the input's
// precision before the bf16 split.  Every transform here has Q[k', n-1-j] = (-1)^k' Q[k', j],
// so the even (odd) rows of Q applied to the paired halves give the even (odd)
// coefficients with half the contraction length.
//
// One BM x BN output tile (index lt of entry E, tile-rows visited in groups of GROUP).
template <int EPI, typename TA, typename TC, bool SPLIT, int BM, int BN, int AMODE, int NW, bool PAIRED = false>
__device__ __forceinline__ void bf16_tile(const GemmEntry& E, const int lt, const int GROUP, const int trans_c) {
  using C = Cfg<BM, BN, SPLIT, NW>;
  constexpr int BK = C::BK, LDS = C::LDS, FM = C::FM, FN = C::FN, AE = C::AE, BE = C::BE, NT = C::NT;
  static_assert(AE >= (AMODE == 0 ? 8 : (AMODE == 1 || AMODE == 3) ? 4 : 1) && BE >= 8,
                "tile too small for the block");
  constexpr int AV = AMODE == 0 ? 8 : ((AMODE == 1 || AMODE == 3) ? 4 : 1);  // A elements per load group
  constexpr int NA = AE / AV;                                   // A loads per thread
  constexpr int NB = BE / 8;                                    // B loads per thread per part
  constexpr int APR = BK / AV;                                  // A loads per tile row
  __shared__ __attribute__((aligned(16))) __bf16 As[2][C::PARTS * BM * LDS];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[2][C::PARTS * BN * LDS];

  const int per_group = GROUP * E.tiles_n;
  const int first_m = (lt / per_group) * GROUP;
  const int gsize = min(E.tiles_m - first_m, GROUP);
  const int tm = first_m + (lt % per_group) % gsize;
  const int tn = (lt % per_group) / gsize;
  const int m0 = tm * BM, n0 = tn * BN;

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wm = (wave / C::WGN) * (BM / C::WGM), wn = (wave % C::WGN) * (BN / C::WGN);
  const int li = lane & 15, lg = lane >> 4;
  const int M = E.M, N = E.N, K = E.K;

  // Per-thread staging addresses, fixed for the whole k loop.  Rows past M (N) are
  // clamped to the last row: their products only reach output rows (columns) that the
  // epilogue discards, so no per-element guard is needed except on the k tail.
  const TA* a_ptr[NA];
  const TA* a_mir[NA];  // PAIRED: the mirrored element of column c (k = 0), read downwards
  int a_col[NA];
#pragma unroll
  for (int e = 0; e < NA; ++e) {
    const int v = e * NT + tid, r = v / APR, c = (v % APR) * AV;
    a_ptr[e] = static_cast<const TA*>(E.A) + (int64_t)min(m0 + r, M - 1) * E.lda + c;
    a_mir[e] = PAIRED ? a_ptr[e] + (E.pair_n - 1 - 2 * c) : nullptr;
    a_col[e] = c;
  }
  // PAIRED: value of A column (k0 + q) of staging group e, with its mirror combined.
  const double psign = PAIRED ? (double)E.pair_sign : 0.0;
  const float fsign = (float)psign;
  auto paired = [&](int e, int k0, int q, double a) -> float {
    return (float)(a + psign * (double)a_mir[e][-(k0 + q)]);
  };
  const __bf16* b_ptr[C::PARTS][NB];
  int b_col[NB];
#pragma unroll
  for (int e = 0; e < NB; ++e) {
    const int v = e * NT + tid, r = v >> 2, c = (v & 3) * 8;
    const int64_t off = (int64_t)min(n0 + r, N - 1) * E.ldb + c;
    b_ptr[0][e] = static_cast<const __bf16*>(E.Bt) + off;
    if constexpr (SPLIT) b_ptr[1][e] = static_cast<const __bf16*>(E.Bt_lo) + off;
    b_col[e] = c;
  }

  f4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

  // One k-tile of operands staged in registers between global memory and LDS.  bf16
  // operands stay packed in 128-bit words (8 elements): as arrays of __bf16 the
  // compiler gives every element a register of its own and unpacks each load as soon as
  // it is issued, which waits out the load and defeats the prefetch.
  struct Stage {
    float ra[AMODE == 0 ? 1 : AE];
    uint4 rab[AMODE == 0 ? NA : 1];
    uint4 rb[C::PARTS][NB];
  };
  Stage S0, S1;
  const int nk = (K + BK - 1) / BK;
  const bool ragged = (K % BK) != 0;

  // Full k-tile loads (no tail handling).
  auto gload_full = [&](int kt, Stage& st) {
    const int k0 = kt * BK;
#pragma unroll
    for (int e = 0; e < NA; ++e) {
      if constexpr (AMODE == 0) {
        st.rab[e] = *reinterpret_cast<const uint4*>(a_ptr[e] + k0);
      } else if constexpr (AMODE == 1) {
        const float4 f = *reinterpret_cast<const float4*>(a_ptr[e] + k0);
        if constexpr (PAIRED) {
          // The mirrors one element each (their groups are not 16-byte aligned for every n,
          // and a branch on the alignment would serialise the loop's loads).
          const float* mp = reinterpret_cast<const float*>(a_mir[e]) - k0;
          st.ra[e * 4 + 0] = f.x + fsign * mp[0]; st.ra[e * 4 + 1] = f.y + fsign * mp[-1];
          st.ra[e * 4 + 2] = f.z + fsign * mp[-2]; st.ra[e * 4 + 3] = f.w + fsign * mp[-3];
        } else {
          st.ra[e * 4 + 0] = f.x; st.ra[e * 4 + 1] = f.y; st.ra[e * 4 + 2] = f.z; st.ra[e * 4 + 3] = f.w;
        }
      } else if constexpr (AMODE == 3) {
        const double2 d0 = *reinterpret_cast<const double2*>(a_ptr[e] + k0);
        const double2 d1 = *reinterpret_cast<const double2*>(a_ptr[e] + k0 + 2);
        if constexpr (PAIRED) {
          st.ra[e * 4 + 0] = paired(e, k0, 0, d0.x); st.ra[e * 4 + 1] = paired(e, k0, 1, d0.y);
          st.ra[e * 4 + 2] = paired(e, k0, 2, d1.x); st.ra[e * 4 + 3] = paired(e, k0, 3, d1.y);
        } else {
          st.ra[e * 4 + 0] = (float)d0.x; st.ra[e * 4 + 1] = (float)d0.y;
          st.ra[e * 4 + 2] = (float)d1.x; st.ra[e * 4 + 3] = (float)d1.y;
        }
      } else {
        if constexpr (PAIRED) st.ra[e] = paired(e, k0, 0, (double)a_ptr[e][k0]);
        else st.ra[e] = to_f(a_ptr[e][k0]);
      }
    }
#pragma unroll
    for (int part = 0; part < C::PARTS; ++part)
#pragma unroll
      for (int e = 0; e < NB; ++e) st.rb[part][e] = *reinterpret_cast<const uint4*>(b_ptr[part][e] + k0);
  };
  auto gload = [&](int kt, Stage& st) {
    const int k0 = kt * BK;
    if (!(ragged && kt == nk - 1)) {
      gload_full(kt, st);
    } else {
      // Last, partial k-tile: zero beyond K (padding may hold anything).
#pragma unroll
      for (int e = 0; e < NA; ++e) {
        if constexpr (AMODE == 0) {
          st.rab[e] = load8_tail(a_ptr[e] + k0, K - (k0 + a_col[e]));
        } else {
#pragma unroll
          for (int q = 0; q < AV; ++q) {
            const bool in = k0 + a_col[e] + q < K;
            if constexpr (PAIRED) st.ra[e * AV + q] = in ? paired(e, k0, q, (double)a_ptr[e][k0 + q]) : 0.f;
            else st.ra[e * AV + q] = in ? to_f(a_ptr[e][k0 + q]) : 0.f;
          }
        }
      }
#pragma unroll
      for (int part = 0; part < C::PARTS; ++part)
#pragma unroll
        for (int e = 0; e < NB; ++e) st.rb[part][e] = load8_tail(b_ptr[part][e] + k0, K - (k0 + b_col[e]));
    }
  };

  auto sstore = [&](int buf, const Stage& st) {
    __bf16* as = As[buf];
#pragma unroll
    for (int e = 0; e < NA; ++e) {
      const int v = e * NT + tid, r = v / APR, c = (v % APR) * AV;
      if constexpr (AMODE == 0) {
        *reinterpret_cast<uint4*>(&as[lx(r, c)]) = st.rab[e];
      } else if constexpr (AMODE == 1 || AMODE == 3) {
        typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
        bf4 hi, lo;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          hi[q] = (__bf16)st.ra[e * 4 + q];
          lo[q] = (__bf16)(st.ra[e * 4 + q] - (float)hi[q]);
        }
        *reinterpret_cast<bf4*>(&as[lx(r, c)]) = hi;
        if constexpr (SPLIT) *reinterpret_cast<bf4*>(&as[BM * LDS + lx(r, c)]) = lo;
      } else {
        const __bf16 hi = (__bf16)st.ra[e];
        as[lx(r, c)] = hi;
        if constexpr (SPLIT) as[BM * LDS + lx(r, c)] = (__bf16)(st.ra[e] - (float)hi);
      }
    }
#pragma unroll
    for (int part = 0; part < C::PARTS; ++part) {
      __bf16* bs = Bs[buf] + part * BN * LDS;
#pragma unroll
      for (int e = 0; e < NB; ++e) {
        const int v = e * NT + tid, r = v >> 2, c = (v & 3) * 8;
        *reinterpret_cast<uint4*>(&bs[lx(r, c)]) = st.rb[part][e];
      }
    }
  };

  auto compute = [&](int cur) {
    const __bf16* as = As[cur];
    const __bf16* bs = Bs[cur];
    bf8 ah[FM], bh[FN];
#pragma unroll
    for (int i = 0; i < FM; ++i) ah[i] = *reinterpret_cast<const bf8*>(&as[lx(wm + i * 16 + li, lg * 8)]);
#pragma unroll
    for (int j = 0; j < FN; ++j) bh[j] = *reinterpret_cast<const bf8*>(&bs[lx(wn + j * 16 + li, lg * 8)]);
    if constexpr (SPLIT) {
      bf8 al[FM], bl[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i)
        al[i] = *reinterpret_cast<const bf8*>(&as[BM * LDS + lx(wm + i * 16 + li, lg * 8)]);
#pragma unroll
      for (int j = 0; j < FN; ++j)
        bl[j] = *reinterpret_cast<const bf8*>(&bs[BN * LDS + lx(wn + j * 16 + li, lg * 8)]);
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
  };

  // Pipeline: LDS buffers L0/L1 alternate; two register stages keep the global loads
  // of tiles k+1 and k+2 in flight while tile k computes.
  // The loops run over pairs of k-tiles with a single exit (an odd last tile is
  // computed after them), so the accumulators stay in place across iterations.  The
  // first loop covers the pairs whose two prefetched tiles are full: its body has no
  // branches, so each LDS store waits only for its own stage's loads (with branches in
  // the body the wait counts merge at the joins and every store drains all loads, the
  // prefetch included).  The second loop finishes the last pair or two.
  gload(0, S0);
  if (nk > 1) gload(1, S1);
  sstore(0, S0);
  __syncthreads();
  const int nfull = ragged ? nk - 1 : nk;
  int kt = 0;
  for (; kt + 3 < nfull; kt += 2) {
    gload_full(kt + 2, S0);
    compute(0);
    sstore(1, S1);
    __syncthreads();
    gload_full(kt + 3, S1);
    compute(1);
    sstore(0, S0);
    __syncthreads();
  }
  for (; kt + 2 <= nk; kt += 2) {
    // tile kt is in L0; tile kt+1 is in flight in S1
    if (kt + 2 < nk) gload(kt + 2, S0);
    compute(0);
    sstore(1, S1);
    __syncthreads();
    if (kt + 3 < nk) gload(kt + 3, S1);
    compute(1);
    if (kt + 2 < nk) sstore(0, S0);
    __syncthreads();
  }
  if (kt < nk) compute(0);

  // f32-result MFMA layout: register r of lane l holds row 4*(l/16) + r, column l%16.
  if constexpr ((EPI == EPI_SHIFT || EPI == EPI_LAMBDA) && std::is_same<TC, float>::value) {
    // The divide passes of the split form (FP32 out): eigenvalues read once per row and
    // column, the sums and the divide in FP32 (far inside the split's 16-bit operands).
    float eb[FN], ea[FM][4];
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int n = n0 + wn + j * 16 + li;
      eb[j] = n < N ? (float)E.eb[n] : 0.f;
    }
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + i * 16 + 4 * lg + r;
        ea[i][r] = m < M ? (float)E.ea[m] : 0.f;
      }
    const float s0 = (float)E.s0, s1 = (float)E.s1, th = (float)E.thresh;
    float* Cp = static_cast<float*>(E.C);
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int mb = m0 + wm + i * 16 + 4 * lg, n = n0 + wn + j * 16 + li;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = mb + r;
          if (m >= M || n >= N) continue;
          float v;
          if constexpr (EPI == EPI_SHIFT) {
            v = acc[i][j][r] / (s0 + s1 * (ea[i][r] + eb[j]));
          } else {
            const float lam = ea[i][r] + eb[j];
            v = lam > th ? acc[i][j][r] / lam : 0.f;
          }
          Cp[trans_c ? (int64_t)n * E.ldc + m : (int64_t)m * E.ldc + n] = v;
        }
      }
    return;
  }
  const bool vec_t = EPI == EPI_STORE && trans_c && (E.ldc & 3) == 0;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int mb = m0 + wm + i * 16 + 4 * lg, n = n0 + wn + j * 16 + li;
      if (vec_t && mb + 3 < M && n < N) {
        // Transposed tile: the lane's 4 rows are 4 consecutive elements of row n of C^T.
        store4<TC>(static_cast<TC*>(E.C) + (int64_t)n * E.ldc + mb, acc[i][j]);
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) epilogue<EPI, TC>(E, trans_c, mb + r, n, (double)acc[i][j][r]);
      }
    }
}


template <int EPI, typename TA, typename TC, bool SPLIT, int BM, int BN, int AMODE, int NW, bool PAIRED = false>
__global__ __launch_bounds__(64 * NW) void k_gemm_bf16(const GemmBatch batch) {
  const int t = xcd_tile(blockIdx.x, batch.total_tiles);
  const GemmEntry& E = batch.e[find_entry(batch, t)];
  bf16_tile<EPI, TA, TC, SPLIT, BM, BN, AMODE, NW, PAIRED>(E, t - E.tile_begin, batch.group, batch.trans_c);
}

// Grouped split-bf16 launch over a device table (64x64 tiles, 4 waves): each block finds
// its entry by binary search over the ascending tile_begin array; entries whose state is
// inactive (a converged segment) return at once.
template <int EPI, typename TA, typename TC, int AMODE>
__global__ __launch_bounds__(256) void k_gemm_bf16_table(const GemmTable tb) {
  if (tb.run != nullptr && *tb.run == 0) return;
  const int t = xcd_tile(blockIdx.x, tb.total_tiles);
  int lo = 0, hi = tb.count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tb.begin[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  const GemmEntry E = tb.e[lo];
  if (tb.use_state && E.st != nullptr && !E.st->active) return;
  bf16_tile<EPI, TA, TC, true, 64, 64, AMODE, 4>(E, t - E.tile_begin, 1, tb.trans_c);
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

hipEvent_t g_ev_start = nullptr, g_ev_stop = nullptr;

// One launch of `kernel`, with the pending events (if any) recorded by the dispatch
// itself: the interval is the kernel's, without the gap of separate event records.
template <typename... Args>
void launch(void (*kernel)(Args...), int blocks, int threads, hipStream_t s, Args... args) {
  if (g_ev_start != nullptr) {
    hipExtLaunchKernelGGL(kernel, dim3(blocks), dim3(threads), 0, s, g_ev_start, g_ev_stop, 0, args...);
    g_ev_start = g_ev_stop = nullptr;
  } else {
    kernel<<<blocks, threads, 0, s>>>(args...);
  }
}

template <typename TA, typename TC, bool SPLIT, int BM, int BN, int AMODE, int NW, bool PAIRED = false>
void launch_bf16_mode(GemmBatch& b, int total, hipStream_t s) {
  constexpr int nt = 64 * NW;
  if constexpr (PAIRED) {
    // Forward passes only: plain store, velocity shift, pressure eigenvalue divide.
    switch (b.epi) {
      case EPI_STORE: launch(k_gemm_bf16<EPI_STORE, TA, TC, SPLIT, BM, BN, AMODE, NW, true>, total, nt, s, (const GemmBatch)b); break;
      case EPI_SHIFT: launch(k_gemm_bf16<EPI_SHIFT, TA, TC, SPLIT, BM, BN, AMODE, NW, true>, total, nt, s, (const GemmBatch)b); break;
      case EPI_LAMBDA: launch(k_gemm_bf16<EPI_LAMBDA, TA, TC, SPLIT, BM, BN, AMODE, NW, true>, total, nt, s, (const GemmBatch)b); break;
      default: throw std::invalid_argument("a paired transform pass stores, shifts or divides");
    }
    return;
  }
  switch (b.epi) {
    case EPI_STORE: launch(k_gemm_bf16<EPI_STORE, TA, TC, SPLIT, BM, BN, AMODE, NW>, total, nt, s, (const GemmBatch)b); break;
    case EPI_SHIFT: launch(k_gemm_bf16<EPI_SHIFT, TA, TC, SPLIT, BM, BN, AMODE, NW>, total, nt, s, (const GemmBatch)b); break;
    case EPI_LAMBDA: launch(k_gemm_bf16<EPI_LAMBDA, TA, TC, SPLIT, BM, BN, AMODE, NW>, total, nt, s, (const GemmBatch)b); break;
    case EPI_ACCUM: launch(k_gemm_bf16<EPI_ACCUM, TA, TC, SPLIT, BM, BN, AMODE, NW>, total, nt, s, (const GemmBatch)b); break;
    default: launch(k_gemm_bf16<EPI_COMBINE, TA, TC, SPLIT, BM, BN, AMODE, NW>, total, nt, s, (const GemmBatch)b); break;
  }
}

template <typename TA, typename TC, bool SPLIT, int BM, int BN, int NW = 4, bool PAIRED = false>
void launch_bf16_tile(GemmBatch& b, hipStream_t s) {
  const int total = tile_batch(b, BM, BN);
  if constexpr (std::is_same<TA, __bf16>::value) {
    launch_bf16_mode<TA, TC, SPLIT, BM, BN, 0, NW>(b, total, s);
  } else if constexpr (std::is_same<TA, float>::value) {
    bool vec = true;
    for (int i = 0; i < b.count; ++i)
      vec = vec && (b.e[i].lda % 4 == 0) && (reinterpret_cast<uintptr_t>(b.e[i].A) % 16 == 0);
    if (vec) launch_bf16_mode<TA, TC, SPLIT, BM, BN, 1, NW, PAIRED>(b, total, s);
    else launch_bf16_mode<TA, TC, SPLIT, BM, BN, 2, NW, PAIRED>(b, total, s);
  } else {
    bool vec = true;
    for (int i = 0; i < b.count; ++i)
      vec = vec && (b.e[i].lda % 2 == 0) && (reinterpret_cast<uintptr_t>(b.e[i].A) % 16 == 0);
    if (vec) launch_bf16_mode<TA, TC, SPLIT, BM, BN, 3, NW, PAIRED>(b, total, s);
    else launch_bf16_mode<TA, TC, SPLIT, BM, BN, 2, NW, PAIRED>(b, total, s);
  }
}

// Tile shape: 0 = automatic, else 64 (64x64), 12864 (128x64), 128 (128x128), 1288
// (128x128 with 8 waves), 641288 (64x128 with 8 waves; FP32/FP64-staged A only).
int g_bf16_tile = 0;
int g_bf16_group = 4;

template <typename TA, typename TC, bool SPLIT, bool PAIRED = false>
void launch_bf16(GemmBatch& b, hipStream_t s) {
  int tile = g_bf16_tile, group = g_bf16_group;
  // Measured at 1024^2 (two or four 1024^3 products per launch): 128x128 tiles with 8
  // waves once that gives a tile per CU (27.9 us at four split products, against 34.0
  // with 4 waves and 35.2 for 128x64; 42.8 against 56.1 us with FP64-staged A), else
  // 64x128 tiles with 8 waves at a tile per CU (19.0 us at two split products against
  // 21.8 for 64x64; 29.3 against 35.6 with FP64-staged A; not for bf16-staged A, whose
  // 16-byte loads need 8 elements per thread), FP32-staged A at 128x64 from 1.5 tiles
  // per CU; smaller launches run 64x64 tiles (4 waves) in groups of 8 tile-rows.
  if (tile == 0) {
    if (tile_batch(b, 128, 128) >= 256) tile = 1288;
    else if (!std::is_same<TA, __bf16>::value && tile_batch(b, 64, 128) >= 256) tile = 641288;
    else if (!std::is_same<TA, double>::value && tile_batch(b, 128, 64) >= 384) tile = 12864;
    else tile = 64;
    group = tile == 64 ? 8 : 4;
  }
  if (b.group <= 0) b.group = group;
  if (tile == 128) launch_bf16_tile<TA, TC, SPLIT, 128, 128, 4, PAIRED>(b, s);
  else if (tile == 1288) launch_bf16_tile<TA, TC, SPLIT, 128, 128, 8, PAIRED>(b, s);
  else if (tile == 641288 && !std::is_same<TA, __bf16>::value) {
    if constexpr (!std::is_same<TA, __bf16>::value) launch_bf16_tile<TA, TC, SPLIT, 64, 128, 8, PAIRED>(b, s);
  }
  else if (tile == 12864) launch_bf16_tile<TA, TC, SPLIT, 128, 64, 4, PAIRED>(b, s);
  else launch_bf16_tile<TA, TC, SPLIT, 64, 64, 4, PAIRED>(b, s);
}

template <typename TA, bool SPLIT>
void launch_bf16_c(GemmBatch& b, int c_type, hipStream_t s) {
  if (c_type == DT_F64) launch_bf16<TA, double, SPLIT>(b, s);
  else if (c_type == DT_F32) launch_bf16<TA, float, SPLIT>(b, s);
  else launch_bf16<TA, __bf16, SPLIT>(b, s);
}

}  // namespace

static void set_scalars(GemmBatch& b) {
  for (int i = 0; i < b.count; ++i) {
    b.e[i].s0 = b.s0;
    b.e[i].s1 = b.s1;
    b.e[i].thresh = b.thresh;
  }
}

void gemm_f64(GemmBatch& b, hipStream_t s) {
  if (b.count <= 0) return;
  set_scalars(b);
  const int total = tile_batch(b, F64_BM, F64_BN);
  switch (b.epi) {
    case EPI_STORE: launch(k_gemm_f64<EPI_STORE>, total, 256, s, (const GemmBatch)b); break;
    case EPI_SHIFT: launch(k_gemm_f64<EPI_SHIFT>, total, 256, s, (const GemmBatch)b); break;
    case EPI_LAMBDA: launch(k_gemm_f64<EPI_LAMBDA>, total, 256, s, (const GemmBatch)b); break;
    default: launch(k_gemm_f64<EPI_COMBINE>, total, 256, s, (const GemmBatch)b); break;
  }
}

void set_gemm_events(hipEvent_t start, hipEvent_t stop) {
  g_ev_start = start;
  g_ev_stop = stop;
}

void set_bf16_tiling(int tile, int group) {
  g_bf16_tile = tile;
  g_bf16_group = group;
}

int tile_table_f64(GemmEntry* entries, int count, int* begins) {
  int total = 0;
  for (int i = 0; i < count; ++i) {
    GemmEntry& e = entries[i];
    e.tiles_m = (e.M + F64_BM - 1) / F64_BM;
    e.tiles_n = (e.N + F64_BN - 1) / F64_BN;
    e.tile_begin = total;
    begins[i] = total;
    total += e.tiles_m * e.tiles_n;
  }
  return total;
}

void gemm_f64_table(const GemmTable& t, hipStream_t s) {
  if (t.count <= 0 || t.total_tiles <= 0) return;
  switch (t.epi) {
    case EPI_STORE: k_gemm_f64_table<EPI_STORE><<<t.total_tiles, 256, 0, s>>>(t); break;
    case EPI_SHIFT: k_gemm_f64_table<EPI_SHIFT><<<t.total_tiles, 256, 0, s>>>(t); break;
    case EPI_LAMBDA: k_gemm_f64_table<EPI_LAMBDA><<<t.total_tiles, 256, 0, s>>>(t); break;
    default: k_gemm_f64_table<EPI_COMBINE><<<t.total_tiles, 256, 0, s>>>(t); break;
  }
}

namespace {
template <int EPI, typename TA, typename TC, int AMODE>
void launch_table(const GemmTable& t, hipStream_t s) {
  k_gemm_bf16_table<EPI, TA, TC, AMODE><<<t.total_tiles, 256, 0, s>>>(t);
}
}  // namespace

int gemm_bf16_table(const GemmTable& t, int a_type, int c_type, bool a_vec, hipStream_t s) {
  if (t.count <= 0 || t.total_tiles <= 0) return 0;
  if (a_type == DT_F64 && c_type == DT_F32 && t.epi == EPI_STORE) {
    if (a_vec) launch_table<EPI_STORE, double, float, 3>(t, s);
    else launch_table<EPI_STORE, double, float, 2>(t, s);
    return 0;
  }
  if (!a_vec || a_type != DT_F32) return -1;
  if (c_type == DT_F32) {
    switch (t.epi) {
      case EPI_STORE: launch_table<EPI_STORE, float, float, 1>(t, s); return 0;
      case EPI_SHIFT: launch_table<EPI_SHIFT, float, float, 1>(t, s); return 0;
      case EPI_LAMBDA: launch_table<EPI_LAMBDA, float, float, 1>(t, s); return 0;
      default: return -1;
    }
  }
  if (c_type == DT_F64) {
    switch (t.epi) {
      case EPI_STORE: launch_table<EPI_STORE, float, double, 1>(t, s); return 0;
      case EPI_COMBINE: launch_table<EPI_COMBINE, float, double, 1>(t, s); return 0;
      default: return -1;
    }
  }
  return -1;
}

void gemm_bf16(GemmBatch& b, int a_type, int c_type, bool split, hipStream_t s) {
  if (b.count <= 0) return;
  set_scalars(b);
  if (b.e[0].pair_sign != 0) {
    // Paired forward passes: split operands, FP32 intermediates.
    if (!split || c_type != DT_F32) throw std::invalid_argument("a paired pass is split and writes FP32");
    if (a_type == DT_F64) launch_bf16<double, float, true, true>(b, s);
    else launch_bf16<float, float, true, true>(b, s);
    return;
  }
  if (split) {
    // Split operands come from FP32/FP64 data (a bf16 input has no low part).
    if (a_type == DT_F64) launch_bf16_c<double, true>(b, c_type, s);
    else launch_bf16_c<float, true>(b, c_type, s);
  } else {
    if (a_type == DT_F64) launch_bf16_c<double, false>(b, c_type, s);
    else if (a_type == DT_F32) launch_bf16_c<float, false>(b, c_type, s);
    else launch_bf16_c<__bf16, false>(b, c_type, s);
  }
}

}  // namespace macs
