// FFT transform passes of the mixed solve (fft.hip): declarations shared with the
// solver's host code.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "internal.h"

namespace macs {

// ------------------------------------------------------- FFT transforms (fft.hip)
// Orthonormal DCT-II / DST-II / DST-I through FP32 complex FFTs (two real rows per FFT)
// for lengths whose FFT length M = n (DCT-II, DST-II) or n + 1 (DST-I) is a power of two
// in [kFftMinLength, kFftMaxLength].  One launch holds one transform group of one M: the
// velocity pair (DST-I and DST-II) or the pressure's DCT-II.
constexpr int kFftMinLength = 16;
constexpr int kFftMaxLength = 2048;
struct FftPlan {
  int n, kind, M;       // transform length, Kind (host_common.h), complex FFT length
  const float2* ct;     // e^{-2 pi i e / M}, e < M/2: half a period of the twiddles
  const float2* cs;     // (cos, sin)(pi k / 2n), k < n: DCT-II / DST-II rotations;
                        // DST-I: (sin(pi j / M), 0), j <= M/2, its input weights
};
// Columns per block of the strip pass.
constexpr int kFftStrip = 4;
enum FftMode : int {
  FFT_ROWS_FWD = 0,  // rows of src (src_type) -> ortho coefficients, FP32 dst (natural layout)
  FFT_STRIPS = 1,    // columns of the FP32 dst, in place: forward, divide, inverse
  FFT_ROWS_INV = 2,  // rows of FP32 coefficients in src -> inverse, dst_type with epi
};
// Source of the forward row pass's rows: the rows themselves (of the batch's src_type), or
// the differences of FP64 pressure rows that D^T makes of them: element i of row r is
// (p[r][i] - p[r][i+1]) / src_h (FFT_SRC_DX, the u faces) or (p[r][i] - p[r+1][i]) /
// src_h (FFT_SRC_DY, the v faces), p[r] at src + r * src_ld, rounded to FP32 once.
enum FftSource : int { FFT_SRC_ROWS = 0, FFT_SRC_DX = 1, FFT_SRC_DY = 2 };
struct FftJob {
  const void* src;          // row r at src + r * src_ld (row modes)
  void* dst;                // row r at dst + r * dst_ld
  const void* aux;          // FFT_ROWS_INV, EPI_COMBINE: FP64 in dst's layout
  const double* lam_row;    // FFT_STRIPS: eigenvalue per column (x coefficient)
  const double* lam_col;    // FFT_STRIPS: eigenvalue per coefficient along the strip
  FftPlan plan;             // the transform along this pass's direction
  int rows;                 // row modes: rows of length plan.n
  int cols;                 // FFT_STRIPS: valid columns (the field has plan.n rows;
                            // dst_ld a multiple of 16, 16-byte aligned rows)
  int src_ld, dst_ld;
  int block_begin;          // set by the launcher
  int src_mode;             // FftSource (FFT_ROWS_FWD)
  double src_h;             // FFT_SRC_DX / FFT_SRC_DY: the grid spacing
  int col;                  // FFT_ROWS_INV with sums: the job's column
};
constexpr int kMaxFftJobs = 8;
struct FftBatch {
  FftJob j[kMaxFftJobs];
  int count, total_blocks;
  int mode;                 // FftMode
  int src_type, dst_type;   // FFT_ROWS_FWD: DT_F64 / DT_F32 src; FFT_ROWS_INV: DT_F32 / DT_F64 dst
  int epi;                  // FFT_STRIPS: EPI_SHIFT / EPI_LAMBDA; FFT_ROWS_INV: EPI_STORE /
                            // EPI_ACCUM / EPI_COMBINE (the last two with a DT_F64 dst)
  double s0, s1, thresh;    // as the GEMM epilogues use them
  // FFT_ROWS_INV, EPI_COMBINE (optional): per-block partial sums of z (the output), z aux
  // and aux at sums[(q * sums_cols + col) * gx + block], q = 0, 1, 2, gx the blocks of a
  // job (all jobs alike).
  double* sums;
  int sums_cols;
};
void fft_pass(FftBatch& b, hipStream_t s);
// Blocks per job of a row pass over `rows` rows at FFT length M (the sums' gx).
int fft_row_blocks(int rows, int M);
bool fft_length_ok(int n, int kind);
// Host tables of a plan: M, the half-period twiddle table (M/2 entries) and the DCT
// half-angle rotations (n entries) or DST-I's input weights (M/2 + 1 entries).
void make_fft_tables(int n, int kind, int* M, std::vector<float2>& th, std::vector<float2>& cs);
// Events the next fft_pass launch records at its kernel's own start and end.
void set_fft_events(hipEvent_t start, hipEvent_t stop);

}  // namespace macs
