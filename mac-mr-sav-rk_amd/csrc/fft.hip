// Fast orthonormal DCT-II / DST-II / DST-I for the mixed solve, gfx950.
//
// The transforms of the diagonalisation (scipy.fft norm="ortho", tensor_stokes.py:58-101)
// computed through complex FFTs in FP32, for lengths whose FFT length is a power of two:
//   DCT-II (n = 2^m):      Makhoul's reordering v = [x0, x2, x4, ..., x5, x3, x1], one
//                          n-point FFT V, then C[k] = Re(e^{-i pi k/2n} V[k]);
//   DST-II (n = 2^m):      the DCT-II of (-1)^j x_j read in reverse coefficient order
//                          (sin(pi (k+1)(2j+1)/2n) = (-1)^j cos(pi (n-1-k)(2j+1)/2n));
//   DST-I  (n + 1 = 2^m):  the 2(n+1)-point FFT of the odd extension [0, x, 0, -x reversed],
//                          whose transform is -2i times the sine sums.
// Two real rows share one complex FFT (one as the real part, one as the imaginary part);
// their transforms separate by the conjugate symmetry of real data.  The inverse of the
// DCT family runs Makhoul's steps backwards (V[k] = e^{i pi k/2n}(C[k] - i C[n-k]), an
// inverse FFT, the reordering undone); DST-I is its own inverse.
//
// The FFT is a Stockham autosort over LDS (radix 4, one radix-2 stage for odd log2 M),
// 256 threads per block; every stage reads its butterflies' inputs, synchronises, and
// writes in place.
//
// A 2-D inverse Q_x^T Q_y^T diag(1/d) Q_y Q_x of one block runs as three passes over FP32
// scratch in the natural row-major layout:
//   FFT_ROWS_FWD  rows of the input (FP64 or FP32) -> x-coefficients;
//   FFT_STRIPS    strips of W columns, in place: y-forward, divide by d(kx, ky), y-inverse,
//                 the strip held in registers between its column-pair FFTs;
//   FFT_ROWS_INV  x-inverse of each row, with the output epilogue (FP32 store, FP64 add,
//                 FP64 preconditioner combination).
// Each pass reads and writes every element once: the transform costs three streams of
// the field instead of four dense products.
#include <hip/hip_ext.h>

#include <cmath>
#include <cstdint>

#include "host_common.h"
#include "internal.h"
#include "fft.h"

namespace macs {
namespace {

hipEvent_t f_ev_start = nullptr, f_ev_stop = nullptr;

constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int cmin(int a, int b) { return a < b ? a : b; }
// Row pairs per block of the row passes (LDS: pairs * lds_points(MMAX) complex floats):
// one pair from M = 1024 up, where a pair's butterflies already fill most of the block
// (measured at 1024^2: one pair per block 8 % faster than two).  The strip pass keeps
// two pairs per block: one round of forward and inverse FFTs per four-column strip.
constexpr int fft_pairs(int mmax) { return mmax <= 512 ? 4 : 1; }
constexpr int kStripPairs = 2;

__device__ inline float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// -i z (forward) or +i z (inverse)
__device__ inline float2 rot(float2 z, bool inv) { return inv ? make_float2(-z.y, z.x) : make_float2(z.y, -z.x); }

// e^{-2 pi i e / MT} (0 <= e < MT) from the half-period table th[e] = e^{-2 pi i e / MT},
// e < MT/2 (in LDS): the second half of the period is the first negated.
__device__ inline float2 twiddle(const float2* th, int e, int lg_mt) {
  const int half = 1 << (lg_mt - 1);
  const float2 w = th[e & (half - 1)];
  return e < half ? w : make_float2(-w.x, -w.y);
}

// LDS slot of FFT point i: one pad slot every 16 points, so the strided accesses of the
// early Stockham stages spread over the banks.  A buffer of M points takes lds_points(M).
__device__ inline int lp(int i) { return i + (i >> 4); }
constexpr int lds_points(int m) { return m + m / 16; }

// In-register DFTs of length 2, 4, 8 (natural order in and out).
__device__ inline void dft4(float2& x0, float2& x1, float2& x2, float2& x3, bool inv) {
  const float2 t0 = cadd(x0, x2), t1 = csub(x0, x2), t2 = cadd(x1, x3), t3 = rot(csub(x1, x3), inv);
  x0 = cadd(t0, t2);
  x1 = cadd(t1, t3);
  x2 = csub(t0, t2);
  x3 = csub(t1, t3);
}
template <int R>
__device__ inline void dft(float2 (&x)[R], bool inv) {
  if (R == 2) {
    const float2 a = x[0];
    x[0] = cadd(a, x[1]);
    x[1] = csub(a, x[1]);
  } else if (R == 4) {
    dft4(x[0], x[1], x[2], x[3], inv);
  } else {
    float2 e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6];
    float2 o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7];
    dft4(e0, e1, e2, e3, inv);
    dft4(o0, o1, o2, o3, inv);
    const float c = 0.70710678118654752f;
    // o_k *= e^{-+ i pi k / 4}
    o1 = inv ? make_float2(c * (o1.x - o1.y), c * (o1.x + o1.y)) : make_float2(c * (o1.x + o1.y), c * (o1.y - o1.x));
    o2 = rot(o2, inv);
    o3 = inv ? make_float2(-c * (o3.x + o3.y), c * (o3.x - o3.y)) : make_float2(c * (o3.y - o3.x), -c * (o3.x + o3.y));
    x[0] = cadd(e0, o0);
    x[4] = csub(e0, o0);
    x[1] = cadd(e1, o1);
    x[5] = csub(e1, o1);
    x[2] = cadd(e2, o2);
    x[6] = csub(e2, o2);
    x[3] = cadd(e3, o3);
    x[7] = csub(e3, o3);
  }
}

// One radix-R Stockham stage over P buffers of M points each, in place (the block reads
// every butterfly's inputs, synchronises, then writes).  Twiddles from the LDS table of
// length MT = 2^lg_mt (a multiple of M).
template <int R, int P, int Q>
__device__ inline void fft_stage(float2* buf, int M, int MP, int lgM, int lgNs, const float2* ct, int lg_mt, bool inv) {
  constexpr int lgR = R == 2 ? 1 : (R == 4 ? 2 : 3);
  const int nb = M >> lgR;  // butterflies per buffer
  const int tid = threadIdx.x;
  const int Ns = 1 << lgNs;
  const int tshift = lg_mt - lgNs - lgR;  // exponent k r of M/(Ns R), in table units
  float2 v[P][Q][R];
  float2 w[Q][R];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int b = tid + q * kThreads;
    if (b < nb) {
#pragma unroll
      for (int p = 0; p < P; ++p)
#pragma unroll
        for (int r = 0; r < R; ++r) v[p][q][r] = buf[p * MP + lp(b + r * nb)];
      const int k = b & (Ns - 1);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        w[q][r] = twiddle(ct, (k * r) << tshift, lg_mt);
        if (inv) w[q][r].y = -w[q][r].y;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int b = tid + q * kThreads;
    if (b >= nb) continue;
    const int k = b & (Ns - 1);
    const int base = ((b >> lgNs) << (lgNs + lgR)) + k;
#pragma unroll
    for (int p = 0; p < P; ++p) {
      float2 x[R];
#pragma unroll
      for (int r = 0; r < R; ++r) x[r] = r == 0 ? v[p][q][0] : cmul(v[p][q][r], w[q][r]);
      dft<R>(x, inv);
#pragma unroll
      for (int r = 0; r < R; ++r) buf[p * MP + lp(base + (r << lgNs))] = x[r];
    }
  }
  __syncthreads();
}

// Unnormalised complex FFT (sign -1, or +1 with inv) of P buffers of M = 2^lgM <= MMAX
// points: radix-8 stages, then one radix-2 or radix-4 stage for the remaining factor.
// The caller synchronises after filling the buffers; the result is visible to every
// thread on return.
template <int P, int MMAX>
__device__ inline void fft_run(float2* buf, int M, int MP, const float2* ct, int lg_mt, bool inv) {
  const int lgM = 31 - __clz(M);
  int lgNs = 0;
  for (; lgNs + 3 <= lgM; lgNs += 3) fft_stage<8, P, cmax(1, MMAX / 2048)>(buf, M, MP, lgM, lgNs, ct, lg_mt, inv);
  if (lgM - lgNs == 2) fft_stage<4, P, cmax(1, MMAX / 1024)>(buf, M, MP, lgM, lgNs, ct, lg_mt, inv);
  else if (lgM - lgNs == 1) fft_stage<2, P, cmax(1, MMAX / 512)>(buf, M, MP, lgM, lgNs, ct, lg_mt, inv);
}

// The launch's half-period twiddle table (MT/2 <= MMAX/2 complex floats): every thread's
// loads issued at once (fetch), then stored to LDS (commit; the caller synchronises).
template <int MMAX>
struct Table {
  static constexpr int E = cmax(1, MMAX / 2 / kThreads);
  float2 t[E];
  __device__ inline void fetch(const FftBatch& b) {
    const int q = 1 << (b.lg_mt - 1);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = threadIdx.x + e * kThreads;
      t[e] = i < q ? b.ct[i] : make_float2(0.f, 0.f);
    }
  }
  __device__ inline void commit(float2* ct, const FftBatch& b) const {
    const int q = 1 << (b.lg_mt - 1);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = threadIdx.x + e * kThreads;
      if (i < q) ct[i] = t[e];
    }
  }
};

// Makhoul's position of element j of a length-n row: even j to j/2, odd j from the end.
__device__ inline int mk_pos(int j, int n) { return (j & 1) ? n - 1 - (j >> 1) : (j >> 1); }

// Element j of two real rows (a, b) into the complex FFT input of the plan's transform.
__device__ inline void put_input(float2* buf, const FftPlan& pl, int j, float a, float b) {
  if (pl.kind == KIND_DST1) {
    buf[lp(j + 1)] = make_float2(a, b);
    buf[lp(pl.M - 1 - j)] = make_float2(-a, -b);
  } else {
    if (pl.kind == KIND_DST2 && (j & 1)) {
      a = -a;
      b = -b;
    }
    buf[lp(mk_pos(j, pl.n))] = make_float2(a, b);
  }
}

// The two zero entries of the DST-I odd extension.
__device__ inline void put_zeros(float2* buf, const FftPlan& pl) {
  if (pl.kind == KIND_DST1 && threadIdx.x == 0) {
    buf[lp(0)] = make_float2(0.f, 0.f);
    buf[lp(pl.n + 1)] = make_float2(0.f, 0.f);
  }
}

// Unnormalised coefficient kk of both rows from the forward FFT: for the DCT family the
// cosine sum C[kk] = sum_j y_j cos(pi kk (2j+1)/2n) (kk a DCT index), for DST-I the sine
// sum S[kk+1] = sum_j x_j sin(pi (kk+1)(j+1)/(n+1)).
// Unnormalised coefficient kk of both rows from the forward FFT: for the DCT family the
// cosine sum C[kk] = sum_j y_j cos(pi kk (2j+1)/2n) (kk a DCT index; cs = (cos, sin)(pi kk
// / 2n)), for DST-I the sine sum S[kk+1] = sum_j x_j sin(pi (kk+1)(j+1)/(n+1)).
__device__ inline float2 get_coef(const float2* buf, const FftPlan& pl, int kk, float2 cs) {
  if (pl.kind == KIND_DST1) {
    const float2 z = buf[lp(kk + 1)];
    return make_float2(-0.5f * z.y, 0.5f * z.x);
  }
  const float2 zk = buf[lp(kk)], zm = buf[lp((pl.M - kk) & (pl.M - 1))];
  const float pr = 0.5f * (zk.x + zm.x), pi = 0.5f * (zk.y - zm.y);
  const float dr = 0.5f * (zk.x - zm.x), di = 0.5f * (zk.y + zm.y);
  return make_float2(cs.x * pr + cs.y * pi, cs.x * di - cs.y * dr);
}

// Makhoul's inverse input V[kk] = e^{i pi kk/2n} (C[kk] - i C[n-kk]) of both rows packed as
// V_a + i V_b, from the unnormalised cosine sums ck = C[kk], cn = C[n-kk] (0 for kk = 0).
__device__ inline float2 inv_input(int kk, float2 ck, float2 cn, float2 cs) {
  const float var = cs.x * ck.x + cs.y * cn.x, vai = cs.y * ck.x - cs.x * cn.x;
  const float vbr = cs.x * ck.y + cs.y * cn.y, vbi = cs.y * ck.y - cs.x * cn.y;
  return make_float2(var - vbi, vai + vbr);
}

// The DCT half-angle rotations (cos, sin)(pi kk / 2n) of a thread's coefficients kk = tid +
// e * 256, fetched from the plan's table at the start of a kernel.
template <int E>
__device__ inline void fetch_rotations(const FftPlan& pl, float2 (&cs)[E]) {
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int kk = threadIdx.x + e * kThreads;
    cs[e] = (pl.kind != KIND_DST1 && kk < pl.n) ? pl.cs[kk] : make_float2(1.f, 0.f);
  }
}

// Coefficient index of DCT index d: DST-II coefficients run backwards.
__device__ inline int coef_index(const FftPlan& pl, int d) { return pl.kind == KIND_DST2 ? pl.n - 1 - d : d; }

__device__ inline float load_f(const void* p, int type, int64_t i) {
  return type == DT_F64 ? (float)static_cast<const double*>(p)[i] : static_cast<const float*>(p)[i];
}

// Blocks are dispatched round-robin over the 8 XCDs: give each XCD a contiguous run of
// the launch's block indices, so neighbouring strips (which share cache lines) share an L2.
__device__ inline int block_index(const FftBatch& b) {
  const int bid = blockIdx.x, total = b.total_blocks;
  if ((total & 7) != 0) return bid;
  return (bid & 7) * (total >> 3) + (bid >> 3);
}

__device__ inline int find_job(const FftBatch& b, int bid) {
  int ji = 0;
#pragma unroll
  for (int i = 1; i < kMaxFftJobs; ++i)
    if (i < b.count && bid >= b.j[i].block_begin) ji = i;
  return ji;
}

// ------------------------------------------------------------------ row passes
// One block: P row pairs (2P consecutive rows) of one job.  Every global load of a block
// is issued before its first LDS store (rows and DCT rotations into registers, unrolled
// to the largest row length), so a block waits for memory once, not once per element
// group.  LDS: P buffers of lds_points(MMAX) points, then the twiddle table.
template <int MMAX, int EPT, bool FWD>
__global__ __launch_bounds__(256) void k_fft_rows(const FftBatch b) {
  extern __shared__ float2 lds[];
  constexpr int P = fft_pairs(MMAX);
  // EPT: row elements per thread, EPT * 256 >= the launch's largest n.
  const int bid = block_index(b);
  const FftJob& J = b.j[find_job(b, bid)];
  const FftPlan pl = J.plan;
  const int n = pl.n, M = pl.M, MP = lds_points(M), rows = J.rows;
  const int r0 = (bid - J.block_begin) * 2 * P;
  const int tid = threadIdx.x;
  const bool dct = pl.kind != KIND_DST1;
  float2* ct = lds + P * lds_points(MMAX);
  // Rows a (2p) and b (2p + 1) of each pair: input rows (forward) or coefficient rows.
  float va[P][EPT], vb[P][EPT];
  const int stype = FWD ? b.src_type : DT_F32;
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int ra = r0 + 2 * p, rb = ra + 1;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int j = tid + e * kThreads;
      va[p][e] = (j < n && ra < rows) ? load_f(J.src, stype, (int64_t)ra * J.src_ld + j) : 0.f;
      vb[p][e] = (j < n && rb < rows) ? load_f(J.src, stype, (int64_t)rb * J.src_ld + j) : 0.f;
    }
  }
  float2 cs[EPT];
  fetch_rotations(pl, cs);
  Table<MMAX> tab;
  tab.fetch(b);
  tab.commit(ct, b);
  if (FWD || !dct) {
    // Forward input, or DST-I's inverse (its own forward): straight into the FFT input.
#pragma unroll
    for (int p = 0; p < P; ++p) {
      float2* buf = lds + p * MP;
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int j = tid + e * kThreads;
        if (j < n) put_input(buf, pl, j, va[p][e], vb[p][e]);
      }
      put_zeros(buf, pl);
    }
    __syncthreads();
  } else {
    // DCT family inverse: stage the coefficients, then Makhoul's inverse input from the
    // pairs (k, n - k) of unnormalised cosine sums C = X / (ortho scale).
#pragma unroll
    for (int p = 0; p < P; ++p) {
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int k = tid + e * kThreads;
        if (k < n) lds[p * MP + lp(k)] = make_float2(va[p][e], vb[p][e]);
      }
    }
    __syncthreads();
    const float sq0 = sqrtf((float)n), sqk = sqrtf(0.5f * (float)n);
    float2 vv[P][EPT];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const float2* buf = lds + p * MP;
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int kk = tid + e * kThreads;
        if (kk >= n) continue;
        float2 ck = buf[lp(coef_index(pl, kk))];
        const float s = kk == 0 ? sq0 : sqk;
        ck.x *= s;
        ck.y *= s;
        float2 cn = make_float2(0.f, 0.f);
        if (kk > 0) {
          cn = buf[lp(coef_index(pl, n - kk))];
          cn.x *= sqk;
          cn.y *= sqk;
        }
        vv[p][e] = inv_input(kk, ck, cn, cs[e]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < P; ++p) {
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int kk = tid + e * kThreads;
        if (kk < n) lds[p * MP + lp(kk)] = vv[p][e];
      }
    }
    __syncthreads();
  }
  fft_run<P, MMAX>(lds, M, MP, ct, b.lg_mt, !FWD && dct);
  if (FWD) {
    float* Y = static_cast<float*>(J.dst);
    const float s_dst1 = sqrtf(2.0f / (float)(n + 1));
    const float s0 = sqrtf(1.0f / (float)n), sk = sqrtf(2.0f / (float)n);
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int ra = r0 + 2 * p, rb = ra + 1;
      const float2* buf = lds + p * MP;
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int kk = tid + e * kThreads;
        if (kk >= n) continue;
        const float2 c = get_coef(buf, pl, kk, cs[e]);
        const float s = dct ? (kk == 0 ? s0 : sk) : s_dst1;
        const int k = coef_index(pl, kk);
        if (ra < rows) Y[(int64_t)ra * J.dst_ld + k] = c.x * s;
        if (rb < rows) Y[(int64_t)rb * J.dst_ld + k] = c.y * s;
      }
    }
  } else {
    const float s_dst1 = sqrtf(2.0f / (float)(n + 1)), inv_n = 1.0f / (float)n;
    // The FP64 epilogues read dst (accumulate) or aux (combine): those loads first too.
    const bool rd = b.dst_type == DT_F64 && (b.epi == EPI_ACCUM || b.epi == EPI_COMBINE);
    const double* rsrc = b.epi == EPI_ACCUM ? static_cast<const double*>(J.dst) : static_cast<const double*>(J.aux);
    double oa[P][EPT], ob[P][EPT];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int ra = r0 + 2 * p, rb = ra + 1;
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int j = tid + e * kThreads;
        oa[p][e] = (rd && j < n && ra < rows) ? rsrc[(int64_t)ra * J.dst_ld + j] : 0.0;
        ob[p][e] = (rd && j < n && rb < rows) ? rsrc[(int64_t)rb * J.dst_ld + j] : 0.0;
      }
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int ra = r0 + 2 * p, rb = ra + 1;
      const float2* buf = lds + p * MP;
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int j = tid + e * kThreads;
        if (j >= n) continue;
        float xa, xb;
        if (dct) {
          const float2 v = buf[lp(mk_pos(j, n))];
          const float s = (pl.kind == KIND_DST2 && (j & 1)) ? -inv_n : inv_n;
          xa = v.x * s;
          xb = v.y * s;
        } else {
          const float2 c = get_coef(buf, pl, j, make_float2(1.f, 0.f));
          xa = c.x * s_dst1;
          xb = c.y * s_dst1;
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int r = h == 0 ? ra : rb;
          if (r >= rows) continue;
          const float x = h == 0 ? xa : xb;
          const double o = h == 0 ? oa[p][e] : ob[p][e];
          const int64_t idx = (int64_t)r * J.dst_ld + j;
          if (b.dst_type == DT_F32) {
            static_cast<float*>(J.dst)[idx] = x;
          } else {
            double* D = static_cast<double*>(J.dst);
            if (b.epi == EPI_ACCUM) D[idx] = o + (double)x;
            else if (b.epi == EPI_COMBINE) D[idx] = b.s0 * (double)x + b.s1 * o;
            else D[idx] = (double)x;
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------- strip pass
// One block: a strip of W columns [kx0, kx0 + W) of an FP32 field of plan.n rows, in
// place.  Each column: forward transform along the rows, coefficient ky divided by
// s0 + s1 (lam_row[kx] + lam_col[ky]) (EPI_SHIFT) or by lam_row[kx] + lam_col[ky] where
// that exceeds thresh, else zeroed (EPI_LAMBDA), inverse transform.  Columns at or past
// J.cols read as zero and come out zero.
template <int MMAX, int NPT, int W>
__global__ __launch_bounds__(256) void k_fft_strips(const FftBatch b) {
  extern __shared__ float2 lds[];
  constexpr int P = cmin(kStripPairs, W / 2);
  const int bid = block_index(b);
  const FftJob& J = b.j[find_job(b, bid)];
  const FftPlan pl = J.plan;
  const int n = pl.n, M = pl.M, MP = lds_points(M);
  const int kx0 = (bid - J.block_begin) * W;
  const int tid = threadIdx.x;
  const bool dct = pl.kind != KIND_DST1;
  float2* ct = lds + P * lds_points(MMAX);
  float* X = static_cast<float*>(J.dst);
  const int64_t ld = J.dst_ld;
  float h[NPT][W];
  Table<MMAX> tab;
  tab.fetch(b);
  float2 cs[NPT];
  fetch_rotations(pl, cs);
  double lc[NPT];  // the coefficients' eigenvalues along the strip
#pragma unroll
  for (int i = 0; i < NPT; ++i) {
    const int kk = tid + i * kThreads;
    lc[i] = kk < n ? J.lam_col[coef_index(pl, kk)] : 0.0;
  }
#pragma unroll
  for (int i = 0; i < NPT; ++i) {
    const int y = tid + i * kThreads;
#pragma unroll
    for (int u = 0; u < W / 4; ++u) {
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (y < n) v = *reinterpret_cast<const float4*>(X + (int64_t)y * ld + kx0 + 4 * u);
      const int c = kx0 + 4 * u;
      h[i][4 * u + 0] = c + 0 < J.cols ? v.x : 0.f;
      h[i][4 * u + 1] = c + 1 < J.cols ? v.y : 0.f;
      h[i][4 * u + 2] = c + 2 < J.cols ? v.z : 0.f;
      h[i][4 * u + 3] = c + 3 < J.cols ? v.w : 0.f;
    }
  }
  tab.commit(ct, b);  // synchronised before its first use, with the first FFT input
  const float s_dst1 = sqrtf(2.0f / (float)(n + 1)), inv_n = 1.0f / (float)n;
#pragma unroll
  for (int c0 = 0; c0 < W; c0 += 2 * P) {
#pragma unroll
    for (int p = 0; p < P; ++p) {
      float2* buf = lds + p * MP;
#pragma unroll
      for (int i = 0; i < NPT; ++i) {
        const int y = tid + i * kThreads;
        if (y < n) put_input(buf, pl, y, h[i][c0 + 2 * p], h[i][c0 + 2 * p + 1]);
      }
      put_zeros(buf, pl);
    }
    __syncthreads();
    fft_run<P, MMAX>(lds, M, MP, ct, b.lg_mt, false);
    // Coefficients, divided; unnormalised cosine sums for the DCT family (the inverse
    // takes them as they are), ortho coefficients for DST-I.
    float2 cr[P][NPT];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int ca = kx0 + c0 + 2 * p, cb = ca + 1;
      const bool va = ca < J.cols, vb = cb < J.cols;
      const double la = va ? J.lam_row[ca] : 0.0, lb = vb ? J.lam_row[cb] : 0.0;
#pragma unroll
      for (int i = 0; i < NPT; ++i) {
        const int kk = tid + i * kThreads;
        if (kk >= n) continue;
        float2 c = get_coef(lds + p * MP, pl, kk, cs[i]);
        if (!dct) {
          c.x *= s_dst1;
          c.y *= s_dst1;
        }
        float da, db;
        if (b.epi == EPI_SHIFT) {
          da = va ? (float)(1.0 / (b.s0 + b.s1 * (la + lc[i]))) : 0.f;
          db = vb ? (float)(1.0 / (b.s0 + b.s1 * (lb + lc[i]))) : 0.f;
        } else {
          const double ta = la + lc[i], tb = lb + lc[i];
          da = (va && ta > b.thresh) ? (float)(1.0 / ta) : 0.f;
          db = (vb && tb > b.thresh) ? (float)(1.0 / tb) : 0.f;
        }
        cr[p][i] = make_float2(c.x * da, c.y * db);
      }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < P; ++p) {
      float2* buf = lds + p * MP;
#pragma unroll
      for (int i = 0; i < NPT; ++i) {
        const int kk = tid + i * kThreads;
        if (kk >= n) continue;
        if (dct) buf[lp(kk)] = cr[p][i];
        else put_input(buf, pl, kk, cr[p][i].x, cr[p][i].y);
      }
      put_zeros(buf, pl);
    }
    __syncthreads();
    if (dct) {
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const float2* buf = lds + p * MP;
#pragma unroll
        for (int i = 0; i < NPT; ++i) {
          const int kk = tid + i * kThreads;
          if (kk >= n) continue;
          const float2 cn = kk > 0 ? buf[lp(n - kk)] : make_float2(0.f, 0.f);
          cr[p][i] = inv_input(kk, cr[p][i], cn, cs[i]);
        }
      }
      __syncthreads();
#pragma unroll
      for (int p = 0; p < P; ++p) {
        float2* buf = lds + p * MP;
#pragma unroll
        for (int i = 0; i < NPT; ++i) {
          const int kk = tid + i * kThreads;
          if (kk < n) buf[lp(kk)] = cr[p][i];
        }
      }
      __syncthreads();
    }
    fft_run<P, MMAX>(lds, M, MP, ct, b.lg_mt, dct);
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const float2* buf = lds + p * MP;
#pragma unroll
      for (int i = 0; i < NPT; ++i) {
        const int y = tid + i * kThreads;
        if (y >= n) continue;
        float xa, xb;
        if (dct) {
          const float2 v = buf[lp(mk_pos(y, n))];
          const float s = (pl.kind == KIND_DST2 && (y & 1)) ? -inv_n : inv_n;
          xa = v.x * s;
          xb = v.y * s;
        } else {
          const float2 c = get_coef(buf, pl, y, make_float2(1.f, 0.f));
          xa = c.x * s_dst1;
          xb = c.y * s_dst1;
        }
        h[i][c0 + 2 * p] = xa;
        h[i][c0 + 2 * p + 1] = xb;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < NPT; ++i) {
    const int y = tid + i * kThreads;
    if (y >= n) continue;
#pragma unroll
    for (int u = 0; u < W / 4; ++u)
      *reinterpret_cast<float4*>(X + (int64_t)y * ld + kx0 + 4 * u) =
          make_float4(h[i][4 * u], h[i][4 * u + 1], h[i][4 * u + 2], h[i][4 * u + 3]);
  }
}

template <typename K>
void launch(K kernel, int blocks, size_t lds, hipStream_t s, const FftBatch& b) {
  if (f_ev_start != nullptr) {
    hipExtLaunchKernelGGL(kernel, dim3(blocks), dim3(kThreads), lds, s, f_ev_start, f_ev_stop, 0, b);
    f_ev_start = f_ev_stop = nullptr;
  } else {
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kThreads), lds, s, b);
  }
}

template <int MMAX, int EPT>
void launch_rows(FftBatch& b, hipStream_t s) {
  constexpr int P = fft_pairs(MMAX);
  int total = 0;
  for (int i = 0; i < b.count; ++i) {
    b.j[i].block_begin = total;
    total += (b.j[i].rows + 2 * P - 1) / (2 * P);
  }
  b.total_blocks = total;
  if (total == 0) return;
  const size_t lds = sizeof(float2) * (P * lds_points(MMAX) + MMAX / 2);
  if (b.mode == FFT_ROWS_FWD) launch(k_fft_rows<MMAX, EPT, true>, total, lds, s, b);
  else launch(k_fft_rows<MMAX, EPT, false>, total, lds, s, b);
}

template <int MMAX, int NPT>
void launch_strips(FftBatch& b, hipStream_t s) {
  constexpr int W = kFftStrip;
  constexpr int P = cmin(kStripPairs, W / 2);
  int total = 0;
  for (int i = 0; i < b.count; ++i) {
    b.j[i].block_begin = total;
    total += (b.j[i].cols + W - 1) / W;
  }
  b.total_blocks = total;
  if (total == 0) return;
  launch(k_fft_strips<MMAX, NPT, W>, total, sizeof(float2) * (P * lds_points(MMAX) + MMAX / 2), s, b);
}

int pow2_at_least(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

}  // namespace

bool fft_length_ok(int n, int kind) {
  if (n < 4 || n > kFftMaxLength) return false;
  const int m = kind == KIND_DST1 ? n + 1 : n;
  return (m & (m - 1)) == 0;
}

void set_fft_events(hipEvent_t start, hipEvent_t stop) {
  f_ev_start = start;
  f_ev_stop = stop;
}

void fft_pass(FftBatch& b, hipStream_t s) {
  if (b.count <= 0) return;
  int mmax = 256, nmax = 1, big = 0;
  for (int i = 0; i < b.count; ++i) {
    if (b.j[i].plan.M > b.j[big].plan.M) big = i;
    mmax = mmax > b.j[i].plan.M ? mmax : b.j[i].plan.M;
    nmax = nmax > b.j[i].plan.n ? nmax : b.j[i].plan.n;
  }
  // One twiddle table serves every plan of the launch: the largest M's (MT = M), of
  // which every smaller power-of-two length takes every (MT / M')-th entry.
  b.ct = b.j[big].plan.ct;
  b.lg_mt = 31 - __builtin_clz((unsigned)b.j[big].plan.M);
  mmax = pow2_at_least(mmax);
  // Elements per thread of the longest row or strip (n <= M, and n < M/2 for DST-I):
  // each branch is one (largest FFT, largest n) combination a launch can have.
  const int npt = pow2_at_least((nmax + kThreads - 1) / kThreads);
  if (b.mode == FFT_STRIPS) {
    if (mmax <= 256) launch_strips<256, 1>(b, s);
    else if (mmax == 512) launch_strips<512, 2>(b, s);
    else if (mmax == 1024) launch_strips<1024, 4>(b, s);
    else if (mmax == 2048 && npt <= 4) launch_strips<2048, 4>(b, s);
    else if (mmax == 2048) launch_strips<2048, 8>(b, s);
    else launch_strips<4096, 8>(b, s);
    return;
  }
  if (mmax <= 256) launch_rows<256, 1>(b, s);
  else if (mmax == 512) launch_rows<512, 2>(b, s);
  else if (mmax == 1024) launch_rows<1024, 4>(b, s);
  else if (mmax == 2048 && npt <= 4) launch_rows<2048, 4>(b, s);
  else if (mmax == 2048) launch_rows<2048, 8>(b, s);
  else launch_rows<4096, 8>(b, s);
}

void make_fft_tables(int n, int kind, int* M, std::vector<float2>& th, std::vector<float2>& cs) {
  const int m = kind == KIND_DST1 ? 2 * (n + 1) : n;
  *M = m;
  // e^{-2 pi i e / m}, e < m/2: half a period of the twiddles.
  th.resize(m / 2);
  for (int e = 0; e < m / 2; ++e) {
    const double a = 2.0 * M_PI * (double)e / (double)m;
    th[e] = make_float2((float)std::cos(a), (float)-std::sin(a));
  }
  // (cos, sin)(pi k / 2n), k < n: the DCT family's half-angle rotations.
  cs.assign(kind == KIND_DST1 ? 1 : n, make_float2(1.f, 0.f));
  if (kind != KIND_DST1)
    for (int k = 0; k < n; ++k) {
      const double a = M_PI * (double)k / (2.0 * n);
      cs[k] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
}

}  // namespace macs
