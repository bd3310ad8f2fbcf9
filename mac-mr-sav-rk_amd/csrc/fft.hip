// Fast orthonormal DCT-II / DST-II / DST-I for the mixed solve, gfx950.
//
// The transforms of the diagonalisation (scipy.fft norm="ortho", tensor_stokes.py:58-101)
// computed through complex FFTs in FP32, for lengths whose FFT length is a power of two:
//   DCT-II (n = 2^m):      Makhoul's reordering v = [x0, x2, x4, ..., x5, x3, x1], one
//                          n-point FFT V, then C[k] = Re(e^{-i pi k/2n} V[k]);
//   DST-II (n = 2^m):      the DCT-II of (-1)^j x_j read in reverse coefficient order
//                          (sin(pi (k+1)(2j+1)/2n) = (-1)^j cos(pi (n-1-k)(2j+1)/2n));
//   DST-I  (N = n + 1 = 2^m): with f_0 = f_N = 0 and f_j the data, one N-point FFT Y of
//                          y_j = sin(pi j/N)(f_j + f_{N-j}) + (f_j - f_{N-j})/2 gives the
//                          sine sums S_m = sum_j f_j sin(pi j m/N) as S_{2k} = -Im Y_k and
//                          S_{2k+1} = S_{2k-1} + Re Y_k (S_{-1} = -S_1), the recurrence a
//                          block prefix sum.  (The odd extension needs a 2N-point FFT; the
//                          recurrence's rounding grows as sqrt(N) eps instead of log(N) eps.)
// Two real rows share one complex FFT (one as the real part, one as the imaginary part);
// their transforms separate by the conjugate symmetry of real data.  The inverse of the
// DCT family runs Makhoul's steps backwards (V[k] = e^{i pi k/2n}(C[k] - i C[n-k]), an
// inverse FFT, the reordering undone); DST-I is its own inverse.
//
// The FFT is a Stockham autosort over LDS, 256 threads per block: radix-8 stages, then
// one radix-2 or radix-4 stage.  Every stage reads its butterflies' inputs, synchronises,
// and writes in place.  The FFT length and the transform kind are template parameters:
// a launch holds one transform group (the velocity pair DST-I and DST-II at FFT length m,
// or the pressure's DCT-II at m), each block takes the compiled body of its job's kind, and
// every stride, twiddle exponent and padded LDS offset inside is a constant or a constant
// offset from one per-thread base.
//
// A 2-D inverse Q_x^T Q_y^T diag(1/d) Q_y Q_x of one block runs as three passes over FP32
// scratch in the natural row-major layout:
//   FFT_ROWS_FWD  rows of the input (FP64 or FP32) -> x-coefficients;
//   FFT_STRIPS    strips of four columns, in place: y-forward, divide by d(kx, ky),
//                 y-inverse, the strip held in registers between its column-pair FFTs;
//   FFT_ROWS_INV  x-inverse of each row, with the output epilogue (FP32 store, FP64 add,
//                 FP64 preconditioner combination).
// Each pass reads and writes every element once: the transform costs three streams of
// the field instead of four dense products.
#include <hip/hip_ext.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>

#include "host_common.h"
#include "internal.h"
#include "fft.h"

namespace macs {
namespace {

hipEvent_t f_ev_start = nullptr, f_ev_stop = nullptr;

constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v / 2); }
// Row pairs per block of the row passes: one pair from M = 1024 up (at 1024^2 one pair
// per block is 8-12 % faster than two: twice the blocks hide more of the memory latency,
// though half the threads idle through the radix-8 stages).  The strip pass keeps two
// pairs per block: one round of forward and inverse FFTs per four-column strip.
constexpr int rows_pairs(int m) { return m <= 512 ? 4 : 1; }
constexpr int kStripPairs = 2;
static_assert(kFftStrip == 2 * kStripPairs, "one round of column-pair FFTs per strip");
// Elements of a length-n row per thread.
constexpr int per_thread(int n) { return (n + kThreads - 1) / kThreads; }

// LDS slot of FFT point i: one pad slot every 16 points, so the strided accesses of the
// early Stockham stages spread over the banks.  A buffer of M points takes lds_points(M).
__device__ inline int lp(int i) { return i + (i >> 4); }
constexpr int lds_points(int m) { return m + m / 16; }
// LDS of one block (complex floats): P buffers, the half-period twiddle table, then the
// DST-I prefix sum's wave totals (one complex float per wave and buffer).
constexpr int kWaves = kThreads / 64;
constexpr int rows_lds(int m) { return rows_pairs(m) * lds_points(m) + m / 2 + kWaves * rows_pairs(m); }
constexpr int strips_lds(int m) { return kStripPairs * lds_points(m) + m / 2 + kWaves * kStripPairs; }

__device__ inline float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// -i z (forward) or +i z (inverse)
template <bool INV>
__device__ inline float2 rot(float2 z) { return INV ? make_float2(-z.y, z.x) : make_float2(z.y, -z.x); }

// In-register DFTs of length 2, 4, 8 (natural order in and out).
template <bool INV>
__device__ inline void dft4(float2& x0, float2& x1, float2& x2, float2& x3) {
  const float2 t0 = cadd(x0, x2), t1 = csub(x0, x2), t2 = cadd(x1, x3), t3 = rot<INV>(csub(x1, x3));
  x0 = cadd(t0, t2);
  x1 = cadd(t1, t3);
  x2 = csub(t0, t2);
  x3 = csub(t1, t3);
}
template <int R, bool INV>
__device__ inline void dft(float2 (&x)[R]) {
  if constexpr (R == 2) {
    const float2 a = x[0];
    x[0] = cadd(a, x[1]);
    x[1] = csub(a, x[1]);
  } else if constexpr (R == 4) {
    dft4<INV>(x[0], x[1], x[2], x[3]);
  } else {
    float2 e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6];
    float2 o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7];
    dft4<INV>(e0, e1, e2, e3);
    dft4<INV>(o0, o1, o2, o3);
    const float c = 0.70710678118654752f;
    // o_k *= e^{-+ i pi k / 4}
    o1 = INV ? make_float2(c * (o1.x - o1.y), c * (o1.x + o1.y)) : make_float2(c * (o1.x + o1.y), c * (o1.y - o1.x));
    o2 = rot<INV>(o2);
    o3 = INV ? make_float2(-c * (o3.x + o3.y), c * (o3.x - o3.y)) : make_float2(c * (o3.y - o3.x), -c * (o3.x + o3.y));
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

// e^{-2 pi i e / M} (0 <= e < M) from the half-period table th[e] = e^{-2 pi i e / M},
// e < M/2 (in LDS): the second half of the period is the first negated.  UPPER: e may
// reach the second half (otherwise the lookup is direct).
template <int M, bool UPPER>
__device__ inline float2 twiddle(const float2* th, int e) {
  if constexpr (!UPPER) {
    return th[e];
  } else {
    const float2 w = th[e & (M / 2 - 1)];
    return e < M / 2 ? w : make_float2(-w.x, -w.y);
  }
}

// One radix-R Stockham stage (sub-transforms of length 2^LGNS before it) over P buffers of
// M points each, in place: the block reads every butterfly's inputs, synchronises, then
// writes.  Reads sit nb = M/R points apart and writes 2^LGNS apart; from 16 points on,
// their padded slots are constant offsets from one padded base.  With nb >= 256 a thread
// takes butterfly b of every buffer (its twiddles computed once); with fewer, the P nb
// butterflies of all buffers spread over the block, so no wave idles through the stage.
template <int R, int P, int M, int LGNS, bool INV>
__device__ inline void fft_stage(float2* buf, const float2* th) {
  constexpr int lgR = ilog2(R);
  constexpr int nb = M >> lgR;  // butterflies per buffer
  constexpr bool SHARED = nb >= kThreads;
  constexpr int PT = SHARED ? P : 1;                 // buffers per butterfly slot
  constexpr int G = SHARED ? nb : P * nb;            // butterfly slots
  constexpr int Q = (G + kThreads - 1) / kThreads;   // slots per thread
  constexpr bool ALL = G % kThreads == 0;            // every thread has Q slots
  constexpr int NS = 1 << LGNS;
  constexpr int MP = lds_points(M);
  constexpr int TSH = ilog2(M) - LGNS - lgR;  // twiddle exponent k r in units of M / (NS R)
  const int tid = threadIdx.x;
  float2 v[PT][Q][R];
  float2 w[Q][R];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int g = tid + q * kThreads;
    if (ALL || g < G) {
      const int b = g & (nb - 1);
      const int p0 = SHARED ? 0 : g >> ilog2(nb);
      const int lb = lp(b);
#pragma unroll
      for (int pt = 0; pt < PT; ++pt)
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int o = (p0 + pt) * MP;
          if constexpr (nb % 16 == 0) v[pt][q][r] = buf[o + lb + r * (nb + nb / 16)];
          else v[pt][q][r] = buf[o + lp(b + r * nb)];
        }
      const int k = b & (NS - 1);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        if constexpr (LGNS == 0) continue;  // k = 0: every twiddle is 1
        // e = k r M / (NS R) < r M / R: the second half of the period only for r > R/2.
        if (2 * r > R) w[q][r] = twiddle<M, true>(th, (k * r) << TSH);
        else w[q][r] = twiddle<M, false>(th, (k * r) << TSH);
        if (INV) w[q][r].y = -w[q][r].y;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int g = tid + q * kThreads;
    if (!ALL && g >= G) continue;
    const int b = g & (nb - 1);
    const int p0 = SHARED ? 0 : g >> ilog2(nb);
    const int k = b & (NS - 1);
    const int base = ((b >> LGNS) << (LGNS + lgR)) + k;
    const int lbase = lp(base);
#pragma unroll
    for (int pt = 0; pt < PT; ++pt) {
      const int o = (p0 + pt) * MP;
      float2 x[R];
#pragma unroll
      for (int r = 0; r < R; ++r) x[r] = (r == 0 || LGNS == 0) ? v[pt][q][r] : cmul(v[pt][q][r], w[q][r]);
      dft<R, INV>(x);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if constexpr (NS % 16 == 0) buf[o + lbase + r * (NS + NS / 16)] = x[r];
        else buf[o + lp(base + (r << LGNS))] = x[r];
      }
    }
  }
  __syncthreads();
}

// Unnormalised complex FFT (sign -1, or +1 with INV) of P buffers of M points: radix-8
// stages, then one radix-2 or radix-4 stage for the remaining factor.  The caller
// synchronises after filling the buffers; the result is visible to every thread on return.
template <int P, int M, bool INV, int LGNS = 0>
__device__ inline void fft_run(float2* buf, const float2* th) {
  constexpr int L = ilog2(M);
  if constexpr (LGNS + 3 <= L) {
    fft_stage<8, P, M, LGNS, INV>(buf, th);
    fft_run<P, M, INV, LGNS + 3>(buf, th);
  } else if constexpr (L - LGNS == 2) {
    fft_stage<4, P, M, LGNS, INV>(buf, th);
  } else if constexpr (L - LGNS == 1) {
    fft_stage<2, P, M, LGNS, INV>(buf, th);
  }
}

// The plan's half-period twiddle table (M/2 complex floats) into LDS: every thread's
// loads issued at once (fetch), then stored (commit; the caller synchronises).
template <int M>
struct Table {
  static constexpr int E = cmax(1, M / 2 / kThreads);
  float2 t[E];
  __device__ inline void fetch(const float2* src) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = threadIdx.x + e * kThreads;
      t[e] = i < M / 2 ? src[i] : make_float2(0.f, 0.f);
    }
  }
  __device__ inline void commit(float2* th) const {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = threadIdx.x + e * kThreads;
      if (i < M / 2) th[i] = t[e];
    }
  }
};

// Makhoul's position of element j of a length-N row: even j to j/2, odd j from the end.
template <int N>
__device__ inline int mk_pos(int j) { return (j & 1) ? N - 1 - (j >> 1) : (j >> 1); }

// Element j of two real rows (a, b) into the complex FFT input of a DCT-family transform
// (Makhoul's reordering; DST-II alternates the sign).
template <int M, int KIND>
__device__ inline void put_input(float2* buf, int j, float a, float b) {
  if (KIND == KIND_DST2 && (j & 1)) {
    a = -a;
    b = -b;
  }
  buf[lp(mk_pos<M>(j))] = make_float2(a, b);
}

// Unnormalised cosine sum C[kk] = sum_j y_j cos(pi kk (2j+1)/2n) of both rows from the
// forward FFT (kk a DCT index; cs = (cos, sin)(pi kk / 2n)).
template <int M>
__device__ inline float2 get_coef(const float2* buf, int kk, float2 cs) {
  const float2 zk = buf[lp(kk)], zm = buf[lp((M - kk) & (M - 1))];
  const float pr = 0.5f * (zk.x + zm.x), pi = 0.5f * (zk.y - zm.y);
  const float dr = 0.5f * (zk.x - zm.x), di = 0.5f * (zk.y + zm.y);
  return make_float2(cs.x * pr + cs.y * pi, cs.x * di - cs.y * dr);
}

// ------------------------------------------------------------------- DST-I
// Pairs per thread (j, N - j), j = 1 + tid + 256 e <= N/2, of the input side, and
// recurrence indices per thread k = tid C + c < N/2 of the output side.
constexpr int dst1_per_thread(int m) { return cmax(1, m / 2 / kThreads); }

// y_j and y_{N-j} of one row from f_j, f_{N-j} and w = sin(pi j / N).
__device__ inline void dst1_weigh(float fj, float fm, float w, float& yj, float& ym) {
  const float a = w * (fj + fm), d = 0.5f * (fj - fm);
  yj = a + d;
  ym = a - d;
}

// Inclusive prefix sums over the block of v (C consecutive values per thread, in thread
// order), for P pairs of sequences packed as (a, b).  tot: kWaves * P complex floats of
// LDS, free since the block's last barrier; the sums are complete on return.
template <int P, int C>
__device__ inline void block_scan(float2 (&v)[P][C], float2* tot) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float2 t[P], ex[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
#pragma unroll
    for (int c = 1; c < C; ++c) v[p][c] = cadd(v[p][c], v[p][c - 1]);
    t[p] = v[p][C - 1];
  }
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const float ux = __shfl_up(t[p].x, d, 64), uy = __shfl_up(t[p].y, d, 64);
      if (lane >= d) t[p] = cadd(t[p], make_float2(ux, uy));
    }
  }
#pragma unroll
  for (int p = 0; p < P; ++p) {
    ex[p] = make_float2(__shfl_up(t[p].x, 1, 64), __shfl_up(t[p].y, 1, 64));
    if (lane == 0) ex[p] = make_float2(0.f, 0.f);
    if (lane == 63) tot[wave * P + p] = t[p];
  }
  __syncthreads();
#pragma unroll
  for (int w = 0; w < kWaves - 1; ++w)
    if (w < wave)
#pragma unroll
      for (int p = 0; p < P; ++p) ex[p] = cadd(ex[p], tot[w * P + p]);
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int c = 0; c < C; ++c) v[p][c] = cadd(v[p][c], ex[p]);
}

// The sine sums of both rows of each of P buffers after the forward FFT of y: for the
// thread's k = tid C + c, se = S_{2k} (k >= 1) and so = S_{2k+1}, unnormalised, (a, b)
// packed.  Ends with the buffers free to overwrite.
template <int M, int P>
__device__ inline void dst1_sums(const float2* lds, float2 (&se)[P][dst1_per_thread(M)],
                                 float2 (&so)[P][dst1_per_thread(M)], float2* tot) {
  constexpr int C = dst1_per_thread(M);
  constexpr int MP = lds_points(M);
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const float2* buf = lds + p * MP;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int k = threadIdx.x * C + c;
      float2 r = make_float2(0.f, 0.f), i = make_float2(0.f, 0.f);
      if (k < M / 2) {
        const float2 zk = buf[lp(k)], zm = buf[lp((M - k) & (M - 1))];
        r = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y + zm.y));  // Re Y_k of a, b
        i = make_float2(0.5f * (zm.y - zk.y), 0.5f * (zk.x - zm.x));  // -Im Y_k of a, b
        if (k == 0) r = make_float2(0.5f * r.x, 0.5f * r.y);          // S_1 = Re Y_0 / 2
      }
      so[p][c] = r;
      se[p][c] = i;
    }
  }
  block_scan<P, C>(so, tot);  // its barrier also ends every read of the buffers
}

// Makhoul's inverse input V[kk] = e^{i pi kk/2n} (C[kk] - i C[n-kk]) of both rows packed as
// V_a + i V_b, from the unnormalised cosine sums ck = C[kk], cn = C[n-kk] (0 for kk = 0).
__device__ inline float2 inv_input(float2 ck, float2 cn, float2 cs) {
  const float var = cs.x * ck.x + cs.y * cn.x, vai = cs.y * ck.x - cs.x * cn.x;
  const float vbr = cs.x * ck.y + cs.y * cn.y, vbi = cs.y * ck.y - cs.x * cn.y;
  return make_float2(var - vbi, vai + vbr);
}

// The DCT half-angle rotations (cos, sin)(pi kk / 2n) of a thread's coefficients kk = tid +
// e * 256, fetched from the plan's table at the start of a kernel.
template <int KIND, int N, int E>
__device__ inline void fetch_rotations(const float2* table, float2 (&cs)[E]) {
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int kk = threadIdx.x + e * kThreads;
    cs[e] = (KIND != KIND_DST1 && kk < N) ? table[kk] : make_float2(1.f, 0.f);
  }
}

// Coefficient index of DCT index d: DST-II coefficients run backwards.
template <int KIND, int N>
__device__ inline int coef_index(int d) { return KIND == KIND_DST2 ? N - 1 - d : d; }

__device__ inline float load_f(const void* p, int type, int64_t i) {
  return type == DT_F64 ? (float)static_cast<const double*>(p)[i] : static_cast<const float*>(p)[i];
}

// Element j of source row r of a forward row pass (FftSource): the row's own element, or
// the FP64 pressure difference D^T makes of it, with D^T's arithmetic ((p0 - p1) / h in
// FP64, then rounded).
__device__ inline float load_row(const FftJob& J, int type, int r, int j) {
  if (J.src_mode == FFT_SRC_ROWS) return load_f(J.src, type, (int64_t)r * J.src_ld + j);
  const double* p = static_cast<const double*>(J.src) + (int64_t)r * J.src_ld + j;
  const double q = J.src_mode == FFT_SRC_DX ? p[1] : p[J.src_ld];
  return (float)((p[0] - q) / J.src_h);
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
// The epilogue of the inverse row pass: element idx of row-major output, x the FP32
// result, o the FP64 value the epilogue combines (dst for EPI_ACCUM, aux for EPI_COMBINE).
__device__ inline void store_out(const FftBatch& b, const FftJob& J, int64_t idx, float x, double o) {
  if (b.dst_type == DT_F32) {
    static_cast<float*>(J.dst)[idx] = x;
  } else {
    double* D = static_cast<double*>(J.dst);
    if (b.epi == EPI_ACCUM) D[idx] = o + (double)x;
    else if (b.epi == EPI_COMBINE) D[idx] = b.s0 * (double)x + b.s1 * o;
    else D[idx] = (double)x;
  }
}

// The FP64 value the epilogue of the inverse row pass combines with element idx (0 when
// it combines none).
__device__ inline double epi_operand(const FftBatch& b, const FftJob& J, int64_t idx, bool valid) {
  const bool rd = b.dst_type == DT_F64 && (b.epi == EPI_ACCUM || b.epi == EPI_COMBINE);
  const double* src = b.epi == EPI_ACCUM ? static_cast<const double*>(J.dst) : static_cast<const double*>(J.aux);
  return (rd && valid) ? src[idx] : 0.0;
}

// One block: P row pairs (2P consecutive rows, from row r0) of one job of FFT length M
// and DCT-family kind KIND.  Every global load of a block is issued before its first LDS
// store (rows, DCT rotations and the twiddle table into registers), so a block waits for
// memory once, not once per element group.  LDS: P buffers of lds_points(M) points, then
// the twiddle table.
template <int M, int KIND, bool FWD>
__device__ inline void rows_body(const FftBatch& b, const FftJob& J, int r0, float2* lds) {
  static_assert(KIND != KIND_DST1, "DST-I rows: rows_body_dst1");
  constexpr int P = rows_pairs(M);
  constexpr int N = M;
  constexpr int EPT = per_thread(N);
  constexpr int MP = lds_points(M);
  const int rows = J.rows;
  const int tid = threadIdx.x;
  float2* th = lds + P * MP;
  // Rows a (2p) and b (2p + 1) of each pair: input rows (forward) or coefficient rows.
  float va[P][EPT], vb[P][EPT];
  const int stype = FWD ? b.src_type : DT_F32;
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int ra = r0 + 2 * p, rb = ra + 1;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int j = tid + e * kThreads;
      va[p][e] = (j < N && ra < rows) ? load_row(J, stype, ra, j) : 0.f;
      vb[p][e] = (j < N && rb < rows) ? load_row(J, stype, rb, j) : 0.f;
    }
  }
  float2 cs[EPT];
  fetch_rotations<KIND, N>(J.plan.cs, cs);
  Table<M> tab;
  tab.fetch(J.plan.ct);
  tab.commit(th);
  if constexpr (FWD) {
#pragma unroll
    for (int p = 0; p < P; ++p) {
      float2* buf = lds + p * MP;
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int j = tid + e * kThreads;
        if (j < N) put_input<M, KIND>(buf, j, va[p][e], vb[p][e]);
      }
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
        if (k < N) lds[p * MP + lp(k)] = make_float2(va[p][e], vb[p][e]);
      }
    }
    __syncthreads();
    const float sq0 = sqrtf((float)N), sqk = sqrtf(0.5f * (float)N);
    float2 vv[P][EPT];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const float2* buf = lds + p * MP;
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int kk = tid + e * kThreads;
        if (kk >= N) continue;
        float2 ck = buf[lp(coef_index<KIND, N>(kk))];
        const float s = kk == 0 ? sq0 : sqk;
        ck.x *= s;
        ck.y *= s;
        float2 cn = make_float2(0.f, 0.f);
        if (kk > 0) {
          cn = buf[lp(coef_index<KIND, N>(N - kk))];
          cn.x *= sqk;
          cn.y *= sqk;
        }
        vv[p][e] = inv_input(ck, cn, cs[e]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < P; ++p) {
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int kk = tid + e * kThreads;
        if (kk < N) lds[p * MP + lp(kk)] = vv[p][e];
      }
    }
    __syncthreads();
  }
  fft_run<P, M, !FWD>(lds, th);
  if constexpr (FWD) {
    float* Y = static_cast<float*>(J.dst);
    const float s0 = sqrtf(1.0f / (float)N), sk = sqrtf(2.0f / (float)N);
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int ra = r0 + 2 * p, rb = ra + 1;
      const float2* buf = lds + p * MP;
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int kk = tid + e * kThreads;
        if (kk >= N) continue;
        const float2 c = get_coef<M>(buf, kk, cs[e]);
        const float s = kk == 0 ? s0 : sk;
        const int k = coef_index<KIND, N>(kk);
        if (ra < rows) Y[(int64_t)ra * J.dst_ld + k] = c.x * s;
        if (rb < rows) Y[(int64_t)rb * J.dst_ld + k] = c.y * s;
      }
    }
  } else {
    const float inv_n = 1.0f / (float)N;
    double s3[3] = {0.0, 0.0, 0.0};  // with b.sums: z, z aux, aux of the block's elements
    // The FP64 epilogues read dst (accumulate) or aux (combine): those loads first too.
    double oa[P][EPT], ob[P][EPT];
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int ra = r0 + 2 * p, rb = ra + 1;
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int j = tid + e * kThreads;
        oa[p][e] = epi_operand(b, J, (int64_t)ra * J.dst_ld + j, j < N && ra < rows);
        ob[p][e] = epi_operand(b, J, (int64_t)rb * J.dst_ld + j, j < N && rb < rows);
      }
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int ra = r0 + 2 * p, rb = ra + 1;
      const float2* buf = lds + p * MP;
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int j = tid + e * kThreads;
        if (j >= N) continue;
        const float2 v = buf[lp(mk_pos<N>(j))];
        const float s = (KIND == KIND_DST2 && (j & 1)) ? -inv_n : inv_n;
        if (ra < rows) store_out(b, J, (int64_t)ra * J.dst_ld + j, v.x * s, oa[p][e]);
        if (rb < rows) store_out(b, J, (int64_t)rb * J.dst_ld + j, v.y * s, ob[p][e]);
        if (b.sums != nullptr) {
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            if ((h == 0 ? ra : rb) >= rows) continue;
            const double o = h == 0 ? oa[p][e] : ob[p][e];
            const double z = b.s0 * (double)((h == 0 ? v.x : v.y) * s) + b.s1 * o;
            s3[0] += z;
            s3[1] += z * o;
            s3[2] += o;
          }
        }
      }
    }
    if (b.sums != nullptr) {
      // The block's partial sums: wave sums, then the four waves' in a fixed order.
      __shared__ double wsum[3][kWaves];
      const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s3[q] += __shfl_down(s3[q], off, 64);
        if (lane == 0) wsum[q][wave] = s3[q];
      }
      __syncthreads();
      if (tid < 3) {
        const double t = (wsum[tid][0] + wsum[tid][1]) + (wsum[tid][2] + wsum[tid][3]);
        const int gx = (rows + 2 * P - 1) / (2 * P);  // the job's blocks
        b.sums[((int64_t)tid * b.sums_cols + J.col) * gx + r0 / (2 * P)] = t;
      }
    }
  }
}

// One block: P row pairs (from row r0) of a DST-I job of FFT length M (N = M, n = M - 1),
// forward or inverse (the same orthonormal transform; FWD only selects the input type and
// the epilogue).  Element c of a row is f_{c+1}; each thread loads the pairs (j, N - j) it
// weighs.  The sums, held per recurrence index, return to natural order through LDS, so
// the stores (and the epilogue's FP64 loads) run along the rows.
template <int M, bool FWD>
__device__ inline void rows_body_dst1(const FftBatch& b, const FftJob& J, int r0, float2* lds) {
  constexpr int P = rows_pairs(M);
  constexpr int H = M / 2;
  constexpr int C = dst1_per_thread(M);
  constexpr int EPT = per_thread(M - 1);
  constexpr int MP = lds_points(M);
  const int rows = J.rows;
  const int tid = threadIdx.x;
  float2* th = lds + P * MP;
  float2* tot = th + M / 2;
  // f_j (lo) and f_{N-j} (hi) of rows a and b, j = 1 + tid + 256 e.
  float fl[P][C][2], fh[P][C][2];
  const int stype = FWD ? b.src_type : DT_F32;
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = r0 + 2 * p + h;
#pragma unroll
      for (int e = 0; e < C; ++e) {
        const int j = 1 + tid + e * kThreads;
        const bool ok = j <= H && r < rows;
        fl[p][e][h] = ok ? load_row(J, stype, r, j - 1) : 0.f;
        fh[p][e][h] = ok ? load_row(J, stype, r, M - j - 1) : 0.f;
      }
    }
  float w[C];
#pragma unroll
  for (int e = 0; e < C; ++e) {
    const int j = 1 + tid + e * kThreads;
    w[e] = j <= H ? J.plan.cs[j].x : 0.f;
  }
  Table<M> tab;
  tab.fetch(J.plan.ct);
  // The epilogue's FP64 operands, element j = tid + 256 e.
  double o[P][2][EPT];
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = r0 + 2 * p + h;
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int j = tid + e * kThreads;
        o[p][h][e] = epi_operand(b, J, (int64_t)r * J.dst_ld + j, !FWD && r < rows && j < M - 1);
      }
    }
  tab.commit(th);
#pragma unroll
  for (int p = 0; p < P; ++p) {
    float2* buf = lds + p * MP;
#pragma unroll
    for (int e = 0; e < C; ++e) {
      const int j = 1 + tid + e * kThreads;
      if (j > H) continue;
      float2 yj, ym;
      dst1_weigh(fl[p][e][0], fh[p][e][0], w[e], yj.x, ym.x);
      dst1_weigh(fl[p][e][1], fh[p][e][1], w[e], yj.y, ym.y);
      buf[lp(j)] = yj;
      buf[lp(M - j)] = ym;  // j = N/2: the same slot and value
    }
    if (tid == 0) buf[lp(0)] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_run<P, M, false>(lds, th);
  float2 se[P][C], so[P][C];
  dst1_sums<M, P>(lds, se, so, tot);
  // S_m to slot m (output element m - 1).
#pragma unroll
  for (int p = 0; p < P; ++p) {
    float2* buf = lds + p * MP;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int k = tid * C + c;
      if (k >= H) continue;
      if (k > 0) buf[lp(2 * k)] = se[p][c];
      buf[lp(2 * k + 1)] = so[p][c];
    }
  }
  __syncthreads();
  const float s = sqrtf(2.0f / (float)M);
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const float2* buf = lds + p * MP;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int j = tid + e * kThreads;
      if (j >= M - 1) continue;
      const float2 v = buf[lp(j + 1)];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int r = r0 + 2 * p + h;
        if (r >= rows) continue;
        const float x = (h == 0 ? v.x : v.y) * s;
        const int64_t idx = (int64_t)r * J.dst_ld + j;
        if (FWD) static_cast<float*>(J.dst)[idx] = x;
        else store_out(b, J, idx, x, o[p][h][e]);
      }
    }
  }
}

// ---------------------------------------------------------------- strip pass
// The division of the strip pass, in FP32: 1 / (s0 + s1 (lam_row[kx] + lam_col[ky]))
// (EPI_SHIFT), or 1 / (lam_row[kx] + lam_col[ky]) where that sum exceeds thresh and 0
// elsewhere (EPI_LAMBDA); 0 for a column past the job's.
__device__ inline float divisor(const FftBatch& b, bool valid, double lr, double lc) {
  if (!valid) return 0.f;
  if (b.epi == EPI_SHIFT) return __builtin_amdgcn_rcpf((float)b.s0 + (float)b.s1 * (float)(lr + lc));
  const double t = lr + lc;
  return t > b.thresh ? __builtin_amdgcn_rcpf((float)t) : 0.f;
}

// One block: a strip of four columns [kx0, kx0 + 4) of an FP32 field of N rows (one job
// of FFT length M and DCT-family kind KIND), in place.  Each column: forward transform
// along the rows, coefficient ky times divisor(kx, ky), inverse transform.  Columns at or
// past J.cols read as zero and come out zero.
template <int M, int KIND>
__device__ inline void strips_body(const FftBatch& b, const FftJob& J, int kx0, float2* lds) {
  static_assert(KIND != KIND_DST1, "DST-I strips: strips_body_dst1");
  constexpr int P = kStripPairs;
  constexpr int N = M;
  constexpr int NPT = per_thread(N);
  constexpr int MP = lds_points(M);
  const int tid = threadIdx.x;
  float2* th = lds + P * MP;
  float* X = static_cast<float*>(J.dst);
  const int64_t ld = J.dst_ld;
  float h[NPT][4];
  Table<M> tab;
  tab.fetch(J.plan.ct);
  float2 cs[NPT];
  fetch_rotations<KIND, N>(J.plan.cs, cs);
  double lc[NPT];  // the coefficients' eigenvalues along the strip
#pragma unroll
  for (int i = 0; i < NPT; ++i) {
    const int kk = tid + i * kThreads;
    lc[i] = kk < N ? J.lam_col[coef_index<KIND, N>(kk)] : 0.0;
  }
#pragma unroll
  for (int i = 0; i < NPT; ++i) {
    const int y = tid + i * kThreads;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (y < N) v = *reinterpret_cast<const float4*>(X + (int64_t)y * ld + kx0);
    h[i][0] = kx0 + 0 < J.cols ? v.x : 0.f;
    h[i][1] = kx0 + 1 < J.cols ? v.y : 0.f;
    h[i][2] = kx0 + 2 < J.cols ? v.z : 0.f;
    h[i][3] = kx0 + 3 < J.cols ? v.w : 0.f;
  }
  tab.commit(th);  // synchronised before its first use, with the first FFT input
  const float inv_n = 1.0f / (float)N;
#pragma unroll
  for (int p = 0; p < P; ++p) {
    float2* buf = lds + p * MP;
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
      const int y = tid + i * kThreads;
      if (y < N) put_input<M, KIND>(buf, y, h[i][2 * p], h[i][2 * p + 1]);
    }
  }
  __syncthreads();
  fft_run<P, M, false>(lds, th);
  // Unnormalised cosine sums, divided (the inverse takes them as they are).
  float2 cr[P][NPT];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int ca = kx0 + 2 * p, cb = ca + 1;
    const bool va = ca < J.cols, vb = cb < J.cols;
    const double la = va ? J.lam_row[ca] : 0.0, lb = vb ? J.lam_row[cb] : 0.0;
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
      const int kk = tid + i * kThreads;
      if (kk >= N) continue;
      const float2 c = get_coef<M>(lds + p * MP, kk, cs[i]);
      cr[p][i] = make_float2(c.x * divisor(b, va, la, lc[i]), c.y * divisor(b, vb, lb, lc[i]));
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < P; ++p) {
    float2* buf = lds + p * MP;
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
      const int kk = tid + i * kThreads;
      if (kk < N) buf[lp(kk)] = cr[p][i];
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const float2* buf = lds + p * MP;
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
      const int kk = tid + i * kThreads;
      if (kk >= N) continue;
      const float2 cn = kk > 0 ? buf[lp(N - kk)] : make_float2(0.f, 0.f);
      cr[p][i] = inv_input(cr[p][i], cn, cs[i]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < P; ++p) {
    float2* buf = lds + p * MP;
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
      const int kk = tid + i * kThreads;
      if (kk < N) buf[lp(kk)] = cr[p][i];
    }
  }
  __syncthreads();
  fft_run<P, M, true>(lds, th);
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const float2* buf = lds + p * MP;
#pragma unroll
    for (int i = 0; i < NPT; ++i) {
      const int y = tid + i * kThreads;
      if (y >= N) continue;
      const float2 v = buf[lp(mk_pos<N>(y))];
      const float s = (KIND == KIND_DST2 && (y & 1)) ? -inv_n : inv_n;
      h[i][2 * p] = v.x * s;
      h[i][2 * p + 1] = v.y * s;
    }
  }
#pragma unroll
  for (int i = 0; i < NPT; ++i) {
    const int y = tid + i * kThreads;
    if (y >= N) continue;
    *reinterpret_cast<float4*>(X + (int64_t)y * ld + kx0) = make_float4(h[i][0], h[i][1], h[i][2], h[i][3]);
  }
}

// The DST-I strip (FFT length M, N = M, n = M - 1 rows): each thread loads the row pairs
// (j, N - j) it weighs, and holds the coefficients and outputs 2k - 1, 2k of its
// recurrence indices k.  Both transforms are the forward one (DST-I is its own inverse).
template <int M>
__device__ inline void strips_body_dst1(const FftBatch& b, const FftJob& J, int kx0, float2* lds) {
  constexpr int P = kStripPairs;
  constexpr int H = M / 2;
  constexpr int C = dst1_per_thread(M);
  constexpr int MP = lds_points(M);
  const int tid = threadIdx.x;
  float2* th = lds + P * MP;
  float2* tot = th + M / 2;
  float* X = static_cast<float*>(J.dst);
  const int64_t ld = J.dst_ld;
  // Rows j - 1 (lo) and N - j - 1 (hi), j = 1 + tid + 256 e: element f_j and f_{N-j} of
  // the strip's four columns.
  float4 fl[C], fh[C];
#pragma unroll
  for (int e = 0; e < C; ++e) {
    const int j = 1 + tid + e * kThreads;
    fl[e] = fh[e] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (j <= H) {
      fl[e] = *reinterpret_cast<const float4*>(X + (int64_t)(j - 1) * ld + kx0);
      fh[e] = *reinterpret_cast<const float4*>(X + (int64_t)(M - j - 1) * ld + kx0);
    }
  }
  float w[C];
#pragma unroll
  for (int e = 0; e < C; ++e) {
    const int j = 1 + tid + e * kThreads;
    w[e] = j <= H ? J.plan.cs[j].x : 0.f;
  }
  Table<M> tab;
  tab.fetch(J.plan.ct);
  // Eigenvalues of the thread's coefficients 2k - 1 and 2k along the strip.
  double le[C], lo[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int k = tid * C + c;
    le[c] = (k > 0 && k < H) ? J.lam_col[2 * k - 1] : 0.0;
    lo[c] = k < H ? J.lam_col[2 * k] : 0.0;
  }
  double lr[4];
  bool vc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    vc[q] = kx0 + q < J.cols;
    lr[q] = vc[q] ? J.lam_row[kx0 + q] : 0.0;
  }
  tab.commit(th);
  // Columns 2p, 2p + 1 in buffer p: the weighted input y.
  const auto weigh = [&](float2* buf, int j, float wj, float4 a, float4 z, int p) {
    const float la = p == 0 ? a.x : a.z, lb = p == 0 ? a.y : a.w;
    const float za = p == 0 ? z.x : z.z, zb = p == 0 ? z.y : z.w;
    float2 yj, ym;
    dst1_weigh(la, za, wj, yj.x, ym.x);
    dst1_weigh(lb, zb, wj, yj.y, ym.y);
    buf[lp(j)] = yj;
    buf[lp(M - j)] = ym;
  };
#pragma unroll
  for (int p = 0; p < P; ++p) {
    float2* buf = lds + p * MP;
#pragma unroll
    for (int e = 0; e < C; ++e) {
      const int j = 1 + tid + e * kThreads;
      if (j <= H) weigh(buf, j, w[e], fl[e], fh[e], p);
    }
    if (tid == 0) buf[lp(0)] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_run<P, M, false>(lds, th);
  float2 se[P][C], so[P][C];
  dst1_sums<M, P>(lds, se, so, tot);
  // Ortho coefficients, divided, into the buffers at slot m = coefficient + 1.
  const float s = sqrtf(2.0f / (float)M);
#pragma unroll
  for (int p = 0; p < P; ++p) {
    float2* buf = lds + p * MP;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int k = tid * C + c;
      if (k >= H) continue;
      const int qa = 2 * p, qb = qa + 1;
      if (k > 0)
        buf[lp(2 * k)] = make_float2(se[p][c].x * s * divisor(b, vc[qa], lr[qa], le[c]),
                                     se[p][c].y * s * divisor(b, vc[qb], lr[qb], le[c]));
      buf[lp(2 * k + 1)] = make_float2(so[p][c].x * s * divisor(b, vc[qa], lr[qa], lo[c]),
                                       so[p][c].y * s * divisor(b, vc[qb], lr[qb], lo[c]));
    }
  }
  __syncthreads();
  // The inverse's weighted input from the pairs (j, N - j) of the divided coefficients.
  float2 yl[P][C], yh[P][C];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const float2* buf = lds + p * MP;
#pragma unroll
    for (int e = 0; e < C; ++e) {
      const int j = 1 + tid + e * kThreads;
      if (j > H) continue;
      const float2 a = buf[lp(j)], z = buf[lp(M - j)];
      dst1_weigh(a.x, z.x, w[e], yl[p][e].x, yh[p][e].x);
      dst1_weigh(a.y, z.y, w[e], yl[p][e].y, yh[p][e].y);
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < P; ++p) {
    float2* buf = lds + p * MP;
#pragma unroll
    for (int e = 0; e < C; ++e) {
      const int j = 1 + tid + e * kThreads;
      if (j > H) continue;
      buf[lp(j)] = yl[p][e];
      buf[lp(M - j)] = yh[p][e];
    }
    if (tid == 0) buf[lp(0)] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_run<P, M, false>(lds, th);
  dst1_sums<M, P>(lds, se, so, tot);
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int k = tid * C + c;
    if (k >= H) continue;
    if (k > 0)
      *reinterpret_cast<float4*>(X + (int64_t)(2 * k - 1) * ld + kx0) =
          make_float4(se[0][c].x * s, se[0][c].y * s, se[1][c].x * s, se[1][c].y * s);
    *reinterpret_cast<float4*>(X + (int64_t)(2 * k) * ld + kx0) =
        make_float4(so[0][c].x * s, so[0][c].y * s, so[1][c].x * s, so[1][c].y * s);
  }
}

// ------------------------------------------------------------------- kernels
// A launch over one transform group at FFT length MS: KS = KIND_DST2 is the velocity pair
// (jobs of DST-I or DST-II), KS = KIND_DCT2 the pressure (DCT-II).  Each block runs the
// compiled body of its job's kind.
template <int MS, int KS, bool FWD>
__global__ __launch_bounds__(256) void k_fft_rows(const FftBatch b) {
  extern __shared__ float2 lds[];
  const int bid = block_index(b);
  const FftJob& J = b.j[find_job(b, bid)];
  const int local = bid - J.block_begin;
  if constexpr (KS == KIND_DST2) {
    if (J.plan.kind == KIND_DST1) {
      rows_body_dst1<MS, FWD>(b, J, local * 2 * rows_pairs(MS), lds);
      return;
    }
  }
  rows_body<MS, KS, FWD>(b, J, local * 2 * rows_pairs(MS), lds);
}

template <int MS, int KS>
__global__ __launch_bounds__(256) void k_fft_strips(const FftBatch b) {
  extern __shared__ float2 lds[];
  const int bid = block_index(b);
  const FftJob& J = b.j[find_job(b, bid)];
  const int kx0 = (bid - J.block_begin) * kFftStrip;
  if constexpr (KS == KIND_DST2) {
    if (J.plan.kind == KIND_DST1) {
      strips_body_dst1<MS>(b, J, kx0, lds);
      return;
    }
  }
  strips_body<MS, KS>(b, J, kx0, lds);
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

template <int MS, int KS>
void launch_group(FftBatch& b, hipStream_t s) {
  int total = 0;
  for (int i = 0; i < b.count; ++i) {
    b.j[i].block_begin = total;
    const int lines = 2 * rows_pairs(b.j[i].plan.M);
    total += b.mode == FFT_STRIPS ? (b.j[i].cols + kFftStrip - 1) / kFftStrip : (b.j[i].rows + lines - 1) / lines;
  }
  b.total_blocks = total;
  if (total == 0) return;
  if (b.mode == FFT_STRIPS) {
    launch(k_fft_strips<MS, KS>, total, sizeof(float2) * strips_lds(MS), s, b);
  } else {
    constexpr size_t lds = sizeof(float2) * rows_lds(MS);
    if (b.mode == FFT_ROWS_FWD) launch(k_fft_rows<MS, KS, true>, total, lds, s, b);
    else launch(k_fft_rows<MS, KS, false>, total, lds, s, b);
  }
}

template <int KS>
void launch_length(FftBatch& b, int m, hipStream_t s) {
  switch (m) {
    case 16: launch_group<16, KS>(b, s); break;
    case 32: launch_group<32, KS>(b, s); break;
    case 64: launch_group<64, KS>(b, s); break;
    case 128: launch_group<128, KS>(b, s); break;
    case 256: launch_group<256, KS>(b, s); break;
    case 512: launch_group<512, KS>(b, s); break;
    case 1024: launch_group<1024, KS>(b, s); break;
    case 2048: launch_group<2048, KS>(b, s); break;
    default: throw std::runtime_error("fft_pass: transform length without an FFT kernel");
  }
}

}  // namespace

bool fft_length_ok(int n, int kind) {
  const int m = kind == KIND_DST1 ? n + 1 : n;
  return m >= kFftMinLength && m <= kFftMaxLength && (m & (m - 1)) == 0;
}

int fft_row_blocks(int rows, int M) { return (rows + 2 * rows_pairs(M) - 1) / (2 * rows_pairs(M)); }

void set_fft_events(hipEvent_t start, hipEvent_t stop) {
  f_ev_start = start;
  f_ev_stop = stop;
}

void fft_pass(FftBatch& b, hipStream_t s) {
  if (b.count <= 0) return;
  // The launch's group: DCT-II jobs alone, or DST-I / DST-II jobs of one FFT length m.
  const bool pressure = b.j[0].plan.kind == KIND_DCT2;
  const int m = b.j[0].plan.M;
  for (int i = 0; i < b.count; ++i) {
    const FftPlan& pl = b.j[i].plan;
    if ((pl.kind == KIND_DCT2) != pressure || pl.M != m || !fft_length_ok(pl.n, pl.kind))
      throw std::runtime_error("fft_pass: the jobs of a launch must form one transform group");
  }
  if (pressure) launch_length<KIND_DCT2>(b, m, s);
  else launch_length<KIND_DST2>(b, m, s);
}

void make_fft_tables(int n, int kind, int* M, std::vector<float2>& th, std::vector<float2>& cs) {
  const int m = kind == KIND_DST1 ? n + 1 : n;
  *M = m;
  // e^{-2 pi i e / m}, e < m/2: half a period of the twiddles.
  th.resize(m / 2);
  for (int e = 0; e < m / 2; ++e) {
    const double a = 2.0 * M_PI * (double)e / (double)m;
    th[e] = make_float2((float)std::cos(a), (float)-std::sin(a));
  }
  if (kind == KIND_DST1) {
    // (sin(pi j / m), 0), j <= m/2: DST-I's input weights.
    cs.assign(m / 2 + 1, make_float2(0.f, 0.f));
    for (int j = 0; j <= m / 2; ++j) cs[j].x = (float)std::sin(M_PI * (double)j / (double)m);
    return;
  }
  // (cos, sin)(pi k / 2n), k < n: the DCT family's half-angle rotations.
  cs.resize(n);
  for (int k = 0; k < n; ++k) {
    const double a = M_PI * (double)k / (2.0 * n);
    cs[k] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
}

}  // namespace macs
