// Host orchestration of the MAC shifted-Stokes stage solve and its C ABI.
//
// Algorithm (the reference's TensorStokes.solve, solver/mac/tensor_stokes.py:245-364):
//   W = L^-1 f                           L = mass I + viscosity K, by 2-D DST GEMMs
//   b = -D W, mean removed               Schur right-hand side
//   PCG on S = D L^-1 D^T, M^-1 = mass (D D^T)^+ + viscosity I, x0 = 0 (or a guess),
//       stop when ||r|| < krylov_tolerance ||b|| at the top of an iteration
//   p = x - mean(x);  u = L^-1 (f + D^T p)
//   residual = max|mass u + viscosity K u - D^T p - f| / (1 + max|f|); divergence = max|D u|
// All right-hand-side columns iterate together (one GEMM launch covers every column's
// u and v blocks); a converged column is frozen, so each column's iterates and
// iteration count are those of a column solved alone.
//
// MIXED precision wraps the same solve in an FP64 defect-correction loop: the inner
// solve uses bf16 transform GEMMs with FP32 accumulation and a loose Schur-CG
// tolerance; the outer residual, divergence and the velocity/pressure accumulation
// stay FP64.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "host_common.h"
#include "internal.h"
#include "fft.h"
#include "macstokes.h"

namespace macs {

// Device copies of one transform matrix: FP64 Q and Q^T (n x n), bf16 Q and Q^T
// (n x ld8, zero padded) for the mixed path.
struct DevMatrix {
  int n = 0, kind = -1, ld8 = 0;
  double* q = nullptr;
  double* qt = nullptr;
  __bf16* qb = nullptr;
  __bf16* qtb = nullptr;
  __bf16* qb_lo = nullptr;   // bf16(q - qb): the split form's second term
  __bf16* qtb_lo = nullptr;
  // Even/odd split of the mixed form's forward passes (Q[k, n-1-j] = (-1)^k Q[k, j]): the
  // even rows Q[2r, j] and the odd rows Q[2r+1, j] for j < h = ceil(n/2), rows of ldh
  // (for odd n the middle column of the even rows is halved, of the odd rows zero), hi/lo;
  // and Q^T with its columns in that coefficient order (even k first), for the inverse
  // passes.
  int h = 0, ldh = 0;
  __bf16 *qe = nullptr, *qe_lo = nullptr, *qo = nullptr, *qo_lo = nullptr;
  __bf16 *qtp = nullptr, *qtp_lo = nullptr;
  // FFT form of the same transform (fft.hip), when its FFT length is a power of two.
  bool fft_ok = false;
  FftPlan fft{};
};

// Coefficient order of the even/odd split: stored position i holds coefficient k.
inline int eo_coefficient(int i, int n) {
  const int h = (n + 1) / 2;
  return i < h ? 2 * i : 2 * (i - h) + 1;
}

// One 2-D block (u, v or pressure) of shape ny x nx with its y/x bases.
struct Block {
  int ny = 0, nx = 0;
  const DevMatrix* my = nullptr;
  const DevMatrix* mx = nullptr;
  double* lam_y = nullptr;
  double* lam_x = nullptr;
  double* lam_y_eo = nullptr;  // the same eigenvalues in the even/odd coefficient order
  double* lam_x_eo = nullptr;
  std::vector<double> host_lam_y, host_lam_x;
};

class Solver {
 public:
  Solver(int nx, int ny, double lx, double ly, const mac_stokes_opts& o) : opts_(o) {
    g_.nx = nx;
    g_.ny = ny;
    g_.hx = lx / nx;
    g_.hy = ly / ny;
    nu_ = (int64_t)(nx - 1) * ny;
    nv_ = (int64_t)nx * (ny - 1);
    size_ = nu_ + nv_;
    np_ = (int64_t)nx * ny;
    HIP_CHECK(hipSetDevice(o.device));
    HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    own_stream_ = stream_;
    packed_.ldu = nx - 1;
    packed_.ldv = nx;
    packed_.voff = nu_;
    packed_.stride = size_;
    build_bases();
    // bf16 scratch per column: u region then v region, each large enough for the
    // natural [ny'][round8(nx')] and transposed [nx'][round8(ny')] forms.
    {
      auto r8 = [](int64_t v) { return (v + 7) & ~int64_t(7); };
      const int64_t ureg = std::max<int64_t>((int64_t)ny * round8(nx - 1), (int64_t)(nx - 1) * round8(ny));
      const int64_t vreg = std::max<int64_t>((int64_t)(ny - 1) * round8(nx), (int64_t)nx * round8(ny - 1));
      const int64_t preg = std::max<int64_t>((int64_t)ny * round8(nx), (int64_t)nx * round8(ny));
      toff_v_ = r8(ureg);
      tstride_ = std::max(r8(toff_v_ + vreg), r8(preg));
      lay_b_.ldu = round8(nx - 1);
      lay_b_.ldv = round8(nx);
      lay_b_.voff = toff_v_;
      lay_b_.stride = tstride_;
      // FFT scratch per column: x-coefficients of u then v (rows of round16 floats), or of p.
      auto r16 = [](int64_t v) { return (v + 15) & ~int64_t(15); };
      foff_v_ = (int64_t)ny * r16(nx - 1);
      fstride_ = std::max<int64_t>(foff_v_ + (int64_t)(ny - 1) * r16(nx), (int64_t)ny * r16(nx));
    }
    HIP_CHECK(hipHostMalloc((void**)&host_flags_, sizeof(int) * kMaxCols, hipHostMallocDefault));
    HIP_CHECK(hipHostMalloc((void**)&mailbox_, sizeof(Mailbox), hipHostMallocCoherent | hipHostMallocMapped));
    std::memset(mailbox_, 0, sizeof(Mailbox));
    HIP_CHECK(hipHostGetDevicePointer((void**)&mailbox_dev_, mailbox_, 0));
  }

  ~Solver() {
    (void)hipSetDevice(opts_.device);
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (hipEvent_t e : all_events_) (void)hipEventDestroy(e);
    for (void* p : allocs_) (void)hipFree(p);
    if (host_flags_) (void)hipHostFree(host_flags_);
    if (hist_) (void)hipFree(hist_);
    if (hist_u_) (void)hipFree(hist_u_);
    if (hist_b_) (void)hipFree(hist_b_);
    if (mailbox_) (void)hipHostFree(mailbox_);
    if (own_stream_) (void)hipStreamDestroy(own_stream_);
  }

  const GridDims& grid() const { return g_; }
  int64_t size() const { return size_; }
  int64_t np() const { return np_; }
  hipStream_t stream() const { return stream_; }
  bool external_stream() const { return stream_ != own_stream_; }
  // The stream the handle's work runs on: the caller's (use) or its own.
  void set_stream(hipStream_t s, bool use) {
    HIP_CHECK(hipStreamSynchronize(stream_));
    stream_ = use ? s : own_stream_;
  }
  const mac_stokes_opts& opts() const { return opts_; }
  std::string error;

  static constexpr int kMaxCols = 64;

  // ---------------------------------------------------------------- transforms
  // out = L^-1 in on packed velocity columns (in/out may not alias).
  void velocity_inverse(const double* in, int64_t in_stride, double* out, int64_t out_stride,
                        const std::vector<int>& cols, double mass, double visc) {
    for (size_t base = 0; base < cols.size(); base += kMaxEntries / 2) {
      const size_t count = std::min<size_t>(kMaxEntries / 2, cols.size() - base);
      GemmBatch ga{}, gb{}, gc{}, gd{};
      for (size_t t = 0; t < count; ++t) {
        const int c = cols[base + t];
        double* s1 = s1_ + (int64_t)c * size_;
        double* s2 = s2_ + (int64_t)c * size_;
        const double* src = in + (int64_t)c * in_stride;
        double* dst = out + (int64_t)c * out_stride;
        const Block* blocks[2] = {&ublk_, &vblk_};
        const int64_t offs[2] = {0, nu_};
        for (int bi = 0; bi < 2; ++bi) {
          const Block& B = *blocks[bi];
          const int64_t off = offs[bi];
          add_pass(ga, src + off, B.nx, B.mx->q, B.mx->n, s1 + off, B.ny, B.ny, B.nx, B.nx, nullptr, nullptr);
          add_pass(gb, s1 + off, B.ny, B.my->q, B.my->n, s2 + off, B.ny, B.nx, B.ny, B.ny, B.lam_x, B.lam_y);
          add_pass(gc, s2 + off, B.ny, B.my->qt, B.my->n, s1 + off, B.nx, B.nx, B.ny, B.ny, nullptr, nullptr);
          add_pass(gd, s1 + off, B.nx, B.mx->qt, B.mx->n, dst + off, B.nx, B.ny, B.nx, B.nx, nullptr, nullptr);
        }
      }
      ga.epi = EPI_STORE; ga.trans_c = 1;
      gb.epi = EPI_SHIFT; gb.trans_c = 0; gb.s0 = mass; gb.s1 = visc;
      gc.epi = EPI_STORE; gc.trans_c = 1;
      gd.epi = EPI_STORE; gd.trans_c = 0;
      run_gemm(ga);
      run_gemm(gb);
      run_gemm(gc);
      run_gemm(gd);
    }
  }

  // out = (D D^T)^+ in on pressure columns, or with combine: out = s0*(D D^T)^+ in + s1*in.
  // The input's mean is annihilated by the masked zero mode.
  void pressure_inverse(const double* in, int64_t in_stride, double* out, int64_t out_stride,
                        const std::vector<int>& cols, bool combine, double s0, double s1) {
    for (size_t base = 0; base < cols.size(); base += kMaxEntries) {
      const size_t count = std::min<size_t>(kMaxEntries, cols.size() - base);
      GemmBatch ga{}, gb{}, gc{}, gd{};
      const Block& B = pblk_;
      for (size_t t = 0; t < count; ++t) {
        const int c = cols[base + t];
        double* s1p = s1_ + (int64_t)c * size_;
        double* s2p = s2_ + (int64_t)c * size_;
        const double* src = in + (int64_t)c * in_stride;
        double* dst = out + (int64_t)c * out_stride;
        add_pass(ga, src, B.nx, B.mx->q, B.mx->n, s1p, B.ny, B.ny, B.nx, B.nx, nullptr, nullptr);
        add_pass(gb, s1p, B.ny, B.my->q, B.my->n, s2p, B.ny, B.nx, B.ny, B.ny, B.lam_x, B.lam_y);
        add_pass(gc, s2p, B.ny, B.my->qt, B.my->n, s1p, B.nx, B.nx, B.ny, B.ny, nullptr, nullptr);
        add_pass(gd, s1p, B.nx, B.mx->qt, B.mx->n, dst, B.nx, B.ny, B.nx, B.nx, nullptr, nullptr);
        gd.e[gd.count - 1].aux = src;
      }
      ga.epi = EPI_STORE; ga.trans_c = 1;
      gb.epi = EPI_LAMBDA; gb.trans_c = 0; gb.thresh = p_thresh_;
      gc.epi = EPI_STORE; gc.trans_c = 1;
      gd.epi = combine ? EPI_COMBINE : EPI_STORE; gd.trans_c = 0; gd.s0 = s0; gd.s1 = s1;
      run_gemm(ga);
      run_gemm(gb);
      run_gemm(gc);
      run_gemm(gd);
    }
  }

  // ------------------------------------------------------ bf16 transforms
  // Mixed precision: the same four passes with bf16 operands (FP32 accumulation).
  // Intermediates are bf16 in padded scratch (leading dimensions rounded to 8 so
  // every row starts 16-byte aligned); the last pass writes FP32 (out_type DT_F32,
  // natural packed layout) or adds into an FP64 packed velocity (DT_F64, EPI_ACCUM).
  void velocity_inverse_bf16(const void* in, int in_type, const VelLayout& Lin, void* out, int out_type,
                             int64_t out_stride, const std::vector<int>& cols, double mass, double visc) {
    if (fft_velocity()) {
      velocity_inverse_fft(in, in_type, Lin, out, out_type, out_stride, cols, mass, visc);
      return;
    }
    const bool sp = split_velocity();
    const bool eo = sp && even_odd();                // paired forward passes
    const int mid = sp ? DT_F32 : DT_BF16;           // intermediate storage type
    const size_t mid_bytes = sp ? 4 : 2;
    char* t1base = sp ? reinterpret_cast<char*>(tf1_) : reinterpret_cast<char*>(tb1_);
    char* t2base = sp ? reinterpret_cast<char*>(tf2_) : reinterpret_cast<char*>(tb2_);
    const size_t per_launch = eo ? kMaxEntries / 4 : kMaxEntries / 2;  // columns per launch
    for (size_t base = 0; base < cols.size(); base += per_launch) {
      const size_t count = std::min<size_t>(per_launch, cols.size() - base);
      GemmBatch ga{}, gb{}, gc{}, gd{};
      const size_t in_bytes = in_type == DT_F64 ? 8 : (in_type == DT_F32 ? 4 : 2);
      const size_t out_bytes = out_type == DT_F64 ? 8 : 4;
      for (size_t t = 0; t < count; ++t) {
        const int c = cols[base + t];
        const Block* blocks[2] = {&ublk_, &vblk_};
        const int64_t in_off[2] = {0, Lin.voff};
        const int in_ld[2] = {Lin.ldu, Lin.ldv};
        const int64_t t_off[2] = {0, toff_v_};
        const int64_t out_off[2] = {0, nu_};
        for (int bi = 0; bi < 2; ++bi) {
          const Block& B = *blocks[bi];
          const char* src = static_cast<const char*>(in) + ((int64_t)c * Lin.stride + in_off[bi]) * in_bytes;
          char* t1 = t1base + ((int64_t)c * tstride_ + t_off[bi]) * mid_bytes;
          char* t2 = t2base + ((int64_t)c * tstride_ + t_off[bi]) * mid_bytes;
          char* dst = static_cast<char*>(out) + ((int64_t)c * out_stride + out_off[bi]) * out_bytes;
          const int ry = round8(B.ny), rx = round8(B.nx);
          if (eo) {
            // x-forward and y-forward as even/odd pairs (coefficients stored even k first;
            // the eigenvalues and the inverse matrices follow that order).
            const DevMatrix &X = *B.mx, &Y = *B.my;
            add_pass_any(ga, src, in_ld[bi], X.qe, X.ldh, t1, ry, B.ny, X.h, X.h, nullptr, nullptr, X.qe_lo);
            pair_last(ga, B.nx, 1);
            add_pass_any(ga, src, in_ld[bi], X.qo, X.ldh, t1 + (int64_t)X.h * ry * mid_bytes, ry, B.ny, B.nx / 2, X.h,
                         nullptr, nullptr, X.qo_lo);
            pair_last(ga, B.nx, -1);
            add_pass_any(gb, t1, ry, Y.qe, Y.ldh, t2, ry, B.nx, Y.h, Y.h, B.lam_x_eo, B.lam_y_eo, Y.qe_lo);
            pair_last(gb, B.ny, 1);
            add_pass_any(gb, t1, ry, Y.qo, Y.ldh, t2 + (int64_t)Y.h * mid_bytes, ry, B.nx, B.ny / 2, Y.h, B.lam_x_eo,
                         B.lam_y_eo + Y.h, Y.qo_lo);
            pair_last(gb, B.ny, -1);
            add_pass_any(gc, t2, ry, Y.qtp, Y.ld8, t1, rx, B.nx, B.ny, B.ny, nullptr, nullptr, Y.qtp_lo);
            add_pass_any(gd, t1, rx, X.qtp, X.ld8, dst, B.nx, B.ny, B.nx, B.nx, nullptr, nullptr, X.qtp_lo);
            continue;
          }
          add_pass_any(ga, src, in_ld[bi], B.mx->qb, B.mx->ld8, t1, ry, B.ny, B.nx, B.nx, nullptr, nullptr,
                       sp ? B.mx->qb_lo : nullptr);
          add_pass_any(gb, t1, ry, B.my->qb, B.my->ld8, t2, ry, B.nx, B.ny, B.ny, B.lam_x, B.lam_y,
                       sp ? B.my->qb_lo : nullptr);
          add_pass_any(gc, t2, ry, B.my->qtb, B.my->ld8, t1, rx, B.nx, B.ny, B.ny, nullptr, nullptr,
                       sp ? B.my->qtb_lo : nullptr);
          add_pass_any(gd, t1, rx, B.mx->qtb, B.mx->ld8, dst, B.nx, B.ny, B.nx, B.nx, nullptr, nullptr,
                       sp ? B.mx->qtb_lo : nullptr);
        }
      }
      ga.epi = EPI_STORE; ga.trans_c = 1;
      gb.epi = EPI_SHIFT; gb.trans_c = 0; gb.s0 = mass; gb.s1 = visc;
      gc.epi = EPI_STORE; gc.trans_c = 1;
      gd.epi = out_type == DT_F64 ? EPI_ACCUM : EPI_STORE; gd.trans_c = 0;
      run_gemm(ga, in_type, mid, true, sp);
      run_gemm(gb, mid, mid, true, sp);
      run_gemm(gc, mid, mid, true, sp);
      run_gemm(gd, mid, out_type, true, sp);
    }
  }

  // z = s0 * (D D^T)^+ r + s1 * r with bf16 transforms; r and z FP64 (np per column).
  // in32 (optional): the same input rounded to FP32, for the first pass; the last pass's
  // visc * in term reads the FP64 input either way.
  void pressure_inverse_bf16(const double* in, double* out, const std::vector<int>& cols, double s0, double s1,
                             const float* in32 = nullptr) {
    if (fft_pressure()) {
      pressure_inverse_fft(in, out, cols, s0, s1, in32);
      return;
    }
    const bool sp = split_pressure();
    const bool eo = sp && even_odd();
    const int mid = sp ? DT_F32 : DT_BF16;
    const size_t mid_bytes = sp ? 4 : 2;
    char* t1base = sp ? reinterpret_cast<char*>(tf1_) : reinterpret_cast<char*>(tb1_);
    char* t2base = sp ? reinterpret_cast<char*>(tf2_) : reinterpret_cast<char*>(tb2_);
    const size_t per_launch = eo ? kMaxEntries / 2 : kMaxEntries;
    for (size_t base = 0; base < cols.size(); base += per_launch) {
      const size_t count = std::min<size_t>(per_launch, cols.size() - base);
      GemmBatch ga{}, gb{}, gc{}, gd{};
      const Block& B = pblk_;
      const int ry = round8(B.ny), rx = round8(B.nx);
      for (size_t t = 0; t < count; ++t) {
        const int c = cols[base + t];
        const double* src = in + (int64_t)c * np_;
        const void* a0 = in32 != nullptr ? static_cast<const void*>(in32 + (int64_t)c * np_)
                                         : static_cast<const void*>(src);
        double* dst = out + (int64_t)c * np_;
        char* t1 = t1base + (int64_t)c * tstride_ * mid_bytes;
        char* t2 = t2base + (int64_t)c * tstride_ * mid_bytes;
        if (eo) {
          const DevMatrix &X = *B.mx, &Y = *B.my;
          add_pass_any(ga, a0, B.nx, X.qe, X.ldh, t1, ry, B.ny, X.h, X.h, nullptr, nullptr, X.qe_lo);
          pair_last(ga, B.nx, 1);
          add_pass_any(ga, a0, B.nx, X.qo, X.ldh, t1 + (int64_t)X.h * ry * mid_bytes, ry, B.ny, B.nx / 2, X.h,
                       nullptr, nullptr, X.qo_lo);
          pair_last(ga, B.nx, -1);
          add_pass_any(gb, t1, ry, Y.qe, Y.ldh, t2, ry, B.nx, Y.h, Y.h, B.lam_x_eo, B.lam_y_eo, Y.qe_lo);
          pair_last(gb, B.ny, 1);
          add_pass_any(gb, t1, ry, Y.qo, Y.ldh, t2 + (int64_t)Y.h * mid_bytes, ry, B.nx, B.ny / 2, Y.h, B.lam_x_eo,
                       B.lam_y_eo + Y.h, Y.qo_lo);
          pair_last(gb, B.ny, -1);
          add_pass_any(gc, t2, ry, Y.qtp, Y.ld8, t1, rx, B.nx, B.ny, B.ny, nullptr, nullptr, Y.qtp_lo);
          add_pass_any(gd, t1, rx, X.qtp, X.ld8, dst, B.nx, B.ny, B.nx, B.nx, nullptr, nullptr, X.qtp_lo);
        } else {
          add_pass_any(ga, a0, B.nx, B.mx->qb, B.mx->ld8, t1, ry, B.ny, B.nx, B.nx, nullptr, nullptr,
                       sp ? B.mx->qb_lo : nullptr);
          add_pass_any(gb, t1, ry, B.my->qb, B.my->ld8, t2, ry, B.nx, B.ny, B.ny, B.lam_x, B.lam_y,
                       sp ? B.my->qb_lo : nullptr);
          add_pass_any(gc, t2, ry, B.my->qtb, B.my->ld8, t1, rx, B.nx, B.ny, B.ny, nullptr, nullptr,
                       sp ? B.my->qtb_lo : nullptr);
          add_pass_any(gd, t1, rx, B.mx->qtb, B.mx->ld8, dst, B.nx, B.ny, B.nx, B.nx, nullptr, nullptr,
                       sp ? B.mx->qtb_lo : nullptr);
        }
        gd.e[gd.count - 1].aux = src;
      }
      ga.epi = EPI_STORE; ga.trans_c = 1;
      gb.epi = EPI_LAMBDA; gb.trans_c = 0; gb.thresh = p_thresh_;
      gc.epi = EPI_STORE; gc.trans_c = 1;
      gd.epi = EPI_COMBINE; gd.trans_c = 0; gd.s0 = s0; gd.s1 = s1;
      run_gemm(ga, in32 != nullptr ? DT_F32 : DT_F64, mid, true, sp);
      run_gemm(gb, mid, mid, true, sp);
      run_gemm(gc, mid, mid, true, sp);
      run_gemm(gd, mid, DT_F64, true, sp);
    }
  }

  // ------------------------------------------------------- FFT transforms
  // The mixed form's velocity inverse as three FP32 FFT passes per block (fft.hip): rows
  // of the input to x-coefficients (scratch fx_, rows of round16 floats), the y-transform,
  // shift division and y-inverse in place on column strips, the x-inverse of each row into
  // the output (FP32 store, or added into an FP64 output).  Input FP64 or FP32, or with
  // grad_p the FP64 pressure columns (stride np_) whose D^T the first pass forms itself.
  void velocity_inverse_fft(const void* in, int in_type, const VelLayout& Lin, void* out, int out_type,
                            int64_t out_stride, const std::vector<int>& cols, double mass, double visc,
                            const double* grad_p = nullptr) {
    const size_t in_bytes = in_type == DT_F64 ? 8 : 4;
    const size_t out_bytes = out_type == DT_F64 ? 8 : 4;
    const size_t per_launch = kMaxFftJobs / 2;
    for (size_t base = 0; base < cols.size(); base += per_launch) {
      const size_t count = std::min(per_launch, cols.size() - base);
      FftBatch fa{}, fb{}, fc{};
      fa.mode = FFT_ROWS_FWD;
      fa.src_type = in_type;
      fb.mode = FFT_STRIPS;
      fb.epi = EPI_SHIFT;
      fb.s0 = mass;
      fb.s1 = visc;
      fc.mode = FFT_ROWS_INV;
      fc.dst_type = out_type;
      fc.epi = out_type == DT_F64 ? EPI_ACCUM : EPI_STORE;
      for (size_t t = 0; t < count; ++t) {
        const int c = cols[base + t];
        const Block* blocks[2] = {&ublk_, &vblk_};
        const int64_t in_off[2] = {0, Lin.voff};
        const int in_ld[2] = {Lin.ldu, Lin.ldv};
        const int64_t x_off[2] = {0, foff_v_};
        const int64_t out_off[2] = {0, nu_};
        for (int bi = 0; bi < 2; ++bi) {
          const Block& B = *blocks[bi];
          float* X = fx_ + (int64_t)c * fstride_ + x_off[bi];
          const int ldx = (B.nx + 15) & ~15;
          FftJob& ja = fa.j[fa.count++];
          if (grad_p != nullptr) {
            ja.src = grad_p + (int64_t)c * np_;
            ja.src_ld = g_.nx;
            ja.src_mode = bi == 0 ? FFT_SRC_DX : FFT_SRC_DY;
            ja.src_h = bi == 0 ? g_.hx : g_.hy;
          } else {
            ja.src = static_cast<const char*>(in) + ((int64_t)c * Lin.stride + in_off[bi]) * in_bytes;
            ja.src_ld = in_ld[bi];
          }
          ja.dst = X;
          ja.dst_ld = ldx;
          ja.plan = B.mx->fft;
          ja.rows = B.ny;
          FftJob& jb = fb.j[fb.count++];
          jb.dst = X;
          jb.dst_ld = ldx;
          jb.plan = B.my->fft;
          jb.cols = B.nx;
          jb.lam_row = B.lam_x;
          jb.lam_col = B.lam_y;
          FftJob& jc = fc.j[fc.count++];
          jc.src = X;
          jc.src_ld = ldx;
          jc.dst = static_cast<char*>(out) + ((int64_t)c * out_stride + out_off[bi]) * out_bytes;
          jc.dst_ld = B.nx;
          jc.plan = B.mx->fft;
          jc.rows = B.ny;
        }
      }
      run_fft(fa, grad_p != nullptr ? 8.0 : (double)in_bytes, 4.0);
      run_fft(fb, 4.0, 4.0);
      run_fft(fc, 4.0, out_type == DT_F64 ? 16.0 : 4.0);
    }
  }

  // z = s0 * (D D^T)^+ r + s1 * r through the same three passes (eigenvalue division with
  // the zero mode dropped; the last pass combines with the FP64 input).  in32 (optional):
  // the input rounded to FP32, read by the first pass.
  // sums (optional): the last pass also writes the partial sums of z, z in and in that
  // rho needs (fft_row_blocks per column and quantity, sums_cols columns).
  void pressure_inverse_fft(const double* in, double* out, const std::vector<int>& cols, double s0, double s1,
                            const float* in32, double* sums = nullptr, int sums_cols = 0) {
    const Block& B = pblk_;
    const int ldx = (B.nx + 15) & ~15;
    const size_t per_launch = kMaxFftJobs;
    for (size_t base = 0; base < cols.size(); base += per_launch) {
      const size_t count = std::min(per_launch, cols.size() - base);
      FftBatch fa{}, fb{}, fc{};
      fa.mode = FFT_ROWS_FWD;
      fa.src_type = in32 != nullptr ? DT_F32 : DT_F64;
      fb.mode = FFT_STRIPS;
      fb.epi = EPI_LAMBDA;
      fb.thresh = p_thresh_;
      fc.mode = FFT_ROWS_INV;
      fc.dst_type = DT_F64;
      fc.epi = EPI_COMBINE;
      fc.s0 = s0;
      fc.s1 = s1;
      fc.sums = sums;
      fc.sums_cols = sums_cols;
      for (size_t t = 0; t < count; ++t) {
        const int c = cols[base + t];
        float* X = fx_ + (int64_t)c * fstride_;
        FftJob& ja = fa.j[fa.count++];
        ja.src = in32 != nullptr ? static_cast<const void*>(in32 + (int64_t)c * np_)
                                 : static_cast<const void*>(in + (int64_t)c * np_);
        ja.src_ld = B.nx;
        ja.dst = X;
        ja.dst_ld = ldx;
        ja.plan = B.mx->fft;
        ja.rows = B.ny;
        FftJob& jb = fb.j[fb.count++];
        jb.dst = X;
        jb.dst_ld = ldx;
        jb.plan = B.my->fft;
        jb.cols = B.nx;
        jb.lam_row = B.lam_x;
        jb.lam_col = B.lam_y;
        FftJob& jc = fc.j[fc.count++];
        jc.src = X;
        jc.src_ld = ldx;
        jc.dst = out + (int64_t)c * np_;
        jc.aux = in + (int64_t)c * np_;
        jc.col = c;
        jc.dst_ld = B.nx;
        jc.plan = B.mx->fft;
        jc.rows = B.ny;
      }
      run_fft(fa, in32 != nullptr ? 4.0 : 8.0, 4.0);
      run_fft(fb, 4.0, 4.0);
      run_fft(fc, 4.0, 16.0);
    }
  }

  // One FFT pass, event-timed when profiling: the transform work of the passes it computes
  // (2 rows n^2 per length-n transform of `rows` rows, two transforms for a strip pass) and
  // its algorithmic bytes (every element read once at in_b bytes, written once at out_b).
  void run_fft(FftBatch& b, double in_b, double out_b) {
    if (b.count <= 0) return;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (profiling_) {
      e0 = take_event();
      e1 = take_event();
      set_fft_events(e0, e1);
    }
    fft_pass(b, stream_);
    set_fft_events(nullptr, nullptr);
    HIP_CHECK(hipGetLastError());
    if (profiling_) {
      double tflops = 0, bytes = 0;
      for (int i = 0; i < b.count; ++i) {
        const FftJob& j = b.j[i];
        const double n = j.plan.n;
        const double lines = b.mode == FFT_STRIPS ? j.cols : j.rows;
        tflops += (b.mode == FFT_STRIPS ? 2.0 : 1.0) * 2.0 * lines * n * n;
        bytes += lines * n * (in_b + out_b);
      }
      pending_fft_.push_back({e0, e1, 0.0, bytes, tflops});
    }
  }

  // Bit 3: the split transforms run as FP32 FFT passes where every length allows.
  bool fft_velocity() const { return (opts_.mixed_split & 8) != 0 && split_velocity() && fft_ok_vel_; }
  bool fft_pressure() const { return (opts_.mixed_split & 8) != 0 && split_pressure() && fft_ok_p_; }
  bool split_velocity() const { return (opts_.mixed_split & 1) != 0; }
  bool split_pressure() const { return (opts_.mixed_split & 2) != 0; }
  // Bit 2: the split forward passes run as even/odd pairs of half contraction length.
  bool even_odd() const { return (opts_.mixed_split & 4) != 0; }
  static void pair_last(GemmBatch& b, int n, int sign) {
    b.e[b.count - 1].pair_n = n;
    b.e[b.count - 1].pair_sign = sign;
  }

  // ------------------------------------------------------------- profiling
  // Opt-in HIP-event timing of every transform-GEMM launch on the solver's stream,
  // with the launches' MFMA work (2MNK per issued product), the transform work of the
  // passes they compute (2MNn per pass over a length-n transform, one product: an
  // even/odd pair does the one pass with two half-length entries), and their bytes (the
  // A, Bt and C elements once each).
  void set_profiling(bool on) {
    collect_profile();
    profiling_ = on;
    prof_ms_ = prof_flops_ = prof_bytes_ = prof_tflops_ = 0.0;
    prof_launches_ = 0;
    pf_ms_ = pf_bytes_ = pf_tflops_ = 0.0;
    pf_launches_ = 0;
  }
  void profile(double* ms, double* flops, double* bytes, int64_t* launches, double* tflops) {
    collect_profile();
    *ms = prof_ms_;
    *flops = prof_flops_;
    *bytes = prof_bytes_;
    *launches = prof_launches_;
    *tflops = prof_tflops_;
  }
  void profile_fft(double* ms, double* bytes, int64_t* launches, double* tflops) {
    collect_profile();
    *ms = pf_ms_;
    *bytes = pf_bytes_;
    *launches = pf_launches_;
    *tflops = pf_tflops_;
  }

  void run_gemm(GemmBatch& b, int a_type = DT_F64, int c_type = DT_F64, bool bf16 = false, bool split = false) {
    if (b.count <= 0) return;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (profiling_) {  // recorded by the GEMM's own dispatch
      e0 = take_event();
      e1 = take_event();
      set_gemm_events(e0, e1);
    }
    if (bf16) gemm_bf16(b, a_type, c_type, split, stream_);
    else gemm_f64(b, stream_);
    set_gemm_events(nullptr, nullptr);
    HIP_CHECK(hipGetLastError());
    if (profiling_) {
      const double in_a = a_type == DT_F64 ? 8 : (a_type == DT_F32 ? 4 : 2);
      const double in_b = bf16 ? (split ? 4 : 2) : 8;
      const double out_c = c_type == DT_F64 ? 8 : (c_type == DT_F32 ? 4 : 2);
      double flops = 0, bytes = 0, tflops = 0;
      for (int i = 0; i < b.count; ++i) {
        const GemmEntry& e = b.e[i];
        flops += (split ? 6.0 : 2.0) * e.M * e.N * e.K;
        // An even/odd entry reads its rows' two halves and computes half the coefficients
        // of a length-pair_n transform.
        const double n = e.pair_sign != 0 ? e.pair_n : e.K;
        tflops += 2.0 * e.M * e.N * n;
        bytes += in_a * e.M * (e.pair_sign != 0 ? 0.5 * n : e.K) + in_b * (double)e.N * e.K + out_c * (double)e.M * e.N;
      }
      pending_.push_back({e0, e1, flops, bytes, tflops});
    }
  }

 private:
  struct Pending {
    hipEvent_t e0, e1;
    double flops, bytes, tflops;
  };
  hipEvent_t take_event() {
    if (!event_pool_.empty()) {
      hipEvent_t e = event_pool_.back();
      event_pool_.pop_back();
      return e;
    }
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    all_events_.push_back(e);
    return e;
  }
  void collect_profile() {
    if (pending_.empty() && pending_fft_.empty()) return;
    HIP_CHECK(hipStreamSynchronize(stream_));
    for (const Pending& p : pending_fft_) {
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, p.e0, p.e1));
      pf_ms_ += ms;
      pf_bytes_ += p.bytes;
      pf_tflops_ += p.tflops;
      pf_launches_ += 1;
      event_pool_.push_back(p.e0);
      event_pool_.push_back(p.e1);
    }
    pending_fft_.clear();
    for (const Pending& p : pending_) {
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, p.e0, p.e1));
      prof_ms_ += ms;
      prof_flops_ += p.flops;
      prof_bytes_ += p.bytes;
      prof_tflops_ += p.tflops;
      prof_launches_ += 1;
      event_pool_.push_back(p.e0);
      event_pool_.push_back(p.e1);
    }
    pending_.clear();
  }
  bool profiling_ = false;
  double prof_ms_ = 0, prof_flops_ = 0, prof_bytes_ = 0, prof_tflops_ = 0;
  int64_t prof_launches_ = 0;
  std::vector<Pending> pending_;
  double pf_ms_ = 0, pf_bytes_ = 0, pf_tflops_ = 0;  // FFT passes
  int64_t pf_launches_ = 0;
  std::vector<Pending> pending_fft_;
  std::vector<hipEvent_t> event_pool_, all_events_;

 public:
  // ------------------------------------------------------------------- solve
  struct Outcome {
    double residual = 0, divergence = 0, vmax = 0, fmax = 0;
    std::vector<int> iterations;
    int sweeps = 1;
    bool warm = false;
    double first_residual = 0, first_divergence = 0;  // of the first FP64 defect (mixed / refined)
    bool network = false;                             // the network's correction entered the start
  };

  // ------------------------------------------------------- learned warm start
  void set_warm_start(int mode, void* cnn, double alpha) {
    warm_mode_ = mode;
    cnn_ = cnn;
    cnn_alpha_ = alpha;
  }
  void reset_history() {
    prev_ncols_ = 0;
    hist_count_ = 0;
    hist_committed_ = 0;
  }
  // The solves recorded since the last commit belong to a trial in progress: commit keeps
  // them, rollback forgets them (a rejected or failed trial), so every later guess
  // extrapolates accepted steps only.  The ring holds one step more than the depth, so a
  // trial's records never overwrite the committed steps the extrapolation reads.
  void commit_history() { hist_committed_ = hist_count_; }
  void rollback_history() {
    hist_count_ = hist_committed_;
    prev_ncols_ = 0;
  }
  // Steps the extrapolation reaches back (1 = repeat the last step's solution, k = the
  // degree k-1 polynomial through the last k steps).  Forgets the recorded history.
  void set_history_steps(int steps) {
    hist_steps_ = steps;
    hist_slots_ = 0;
    reset_history();
  }
  int history_steps() const { return hist_steps_; }
  // Same-slot history: with a scheme of `period` stage solves per step, the solve `period`
  // back is the same stage of the previous step; the guess extrapolates polynomially in
  // time through the last `hist_steps_` steps (fewer while fewer are recorded).  The
  // step sizes come from the solves themselves: a stage's viscosity/mass is nu a_ii h,
  // so same-slot ratios are step ratios.  Consecutive same-slot solves are taken
  // (h_m + h_m-1) / 2 apart (the stage at mid-step); the weights are the Lagrange ones
  // on those nodes, (-1)^(k+1) C(K, k) for K equal steps (3, -3, 1 for K = 3).
  void set_history(int period) {
    history_period_ = period;
    reset_history();
  }
  const double* schur_rhs() const { return b_; }
  int last_ncols() const { return ncols_cur_; }

  // Fills out (io_g_ unless given) with the warm start's pressure guess (previous solve,
  // history, or the network's correction of a base guess), if there is one; nothing is
  // written when there is none.  The network reads the
  // base's residual, b - b_base (the Schur right-hand sides recorded with the base's
  // pressures, combined like them): with one operator S p_m = b_m, so the correction
  // it predicts, p - base, solves S (p - base) = b - b_base.
  bool learned_guess(const double* b, int ncols, double* out = nullptr) {
    if (out == nullptr) out = io_g_;
    if (warm_mode_ == MAC_WARM_PREVIOUS) {
      // Previous solve's pressures as the guess (same column count only).
      if (prev_ncols_ != ncols) return false;
      copy_cols(pprev_, np_, out, np_, np_, ncols, stream_);
      return true;
    }
    net_base_ = nullptr;
    if (warm_mode_ == MAC_WARM_HISTORY) return history_guess(ncols, out);
    if (warm_mode_ != MAC_WARM_NETWORK || cnn_ == nullptr || !warm_shape_ok(g_.nx, g_.ny)) return false;
    // The network corrects a base guess: the history extrapolation when a period is set,
    // else the previous solve's pressures.  Without a base there is no guess (zero start),
    // as in training.  Once the history reaches back net_depth_limit_ steps its
    // extrapolation alone is the guess: there the network's correction (at a quarter of
    // the resolution, bf16 weights) is coarser than the extrapolation's error.
    if (history_period_ > 0 && recorded_steps(ncols) >= net_depth_limit_) return history_guess(ncols, out);
    const double* base = nullptr;
    const int64_t n = np_ * ncols;
    if (history_period_ > 0) {
      if (history_b_ok(ncols) && history_guess(ncols, warm_base_) && extrapolate_history(hist_b_, n, warm_db_))
        base = warm_base_;
    } else if (prev_ncols_ == ncols && prev_b_ok_) {
      base = pprev_;
      copy_cols(bprev_, np_, warm_db_, np_, np_, ncols, stream_);
    }
    if (base == nullptr) return false;
    const double w[2] = {1.0, -1.0};
    const double* src[2] = {b, warm_db_};
    extrapolate(src, w, 2, warm_db_, n, stream_);  // b - b_base
    warm_guess(cnn_, warm_db_, np_, base, np_, ncols, g_.nx, g_.ny, cnn_alpha_, warm_x_, warm_y_, warm_s_, out,
               np_, stream_);
    net_base_ = base;
    return true;
  }
  // Same-slot steps recorded for `ncols`-column solves (0 without a history).
  int64_t recorded_steps(int ncols) const {
    if (history_period_ <= 0 || hist_ncols_ != ncols) return 0;
    return std::min<int64_t>(hist_count_ / history_period_, hist_steps_);
  }
  void set_net_depth_limit(int k) { net_depth_limit_ = k; }
  // Every slot the history extrapolation reads has its Schur right-hand side recorded.
  bool history_b_ok(int ncols) const {
    if (history_period_ <= 0 || hist_ncols_ != ncols || hist_b_ == nullptr) return false;
    const int P = history_period_, slots = hist_ring_steps() * P;
    const int64_t steps = std::min<int64_t>(hist_count_ / P, hist_steps_);
    for (int64_t k = 1; k <= steps; ++k)
      if (!hist_b_ok_[(int)((hist_count_ - k * P) % slots)]) return false;
    return steps > 0;
  }
  bool history_guess(int ncols, double* out) {
    if (history_period_ <= 0 || hist_ncols_ != ncols) return false;
    return extrapolate_history(hist_, np_ * ncols, out);
  }
  // out = the same-slot polynomial extrapolation from ring `ring` (elements per slot n).
  bool extrapolate_history(const double* ring, int64_t n, double* out) {
    const int P = history_period_;
    const int64_t steps = std::min<int64_t>(hist_count_ / P, hist_steps_);
    if (steps == 0) return false;
    const int slots = hist_ring_steps() * P;
    auto index = [&](int back) { return (int)((hist_count_ - (int64_t)back * P) % slots); };
    if (steps == 1) {
      copy_cols(ring + index(1) * n, n, out, n, n, 1, stream_);
      return true;
    }
    double w[kMaxHistSteps];
    const double* src[kMaxHistSteps];
    history_weights((int)steps, index, w);
    for (int k = 0; k < steps; ++k) src[k] = ring + index(k + 1) * n;
    extrapolate(src, w, (int)steps, out, n, stream_);
    return true;
  }
  // Lagrange weights at the new solve's time for the last `steps` same-slot solves.
  template <class Index>
  void history_weights(int steps, Index index, double* w) const {
    double h[kMaxHistSteps + 1] = {shift_cur_};
    for (int k = 1; k <= steps; ++k) h[k] = hist_shift_[index(k)];
    macs::history_weights(h, steps, w);
  }
  // Records the solve's pressures (and, for the mixed solve, velocities: its outer
  // iterate starts from both) for the warm starts, with the zero start's Schur
  // right-hand side b when the solve formed it (the network's input).
  void remember_pressure(const double* p, int ncols, const double* u = nullptr, const double* b = nullptr) {
    const Record r = record(ncols, u != nullptr, b);
    copy_cols(p, np_, pprev_, np_, np_, ncols, stream_);
    if (r.p) copy_cols(p, np_, r.p, np_, np_, ncols, stream_);
    if (r.u) copy_cols(u, size_, r.u, size_, size_, ncols, stream_);
  }
  // Where a solve's pressures and velocities go in the history ring (null: not recorded).
  struct Record {
    double* p;
    double* u;
  };
  // The bookkeeping of recording a solve (b is copied here), without the copies of p and
  // u: those go to pprev_ and the returned slots.
  Record record(int ncols, bool with_u, const double* b) {
    Record rec{nullptr, nullptr};
    prev_ncols_ = ncols;
    prev_b_ok_ = b != nullptr;
    if (b) copy_cols(b, np_, bprev_, np_, np_, ncols, stream_);
    if (history_period_ > 0) {
      const int slots = hist_ring_steps() * history_period_;
      if (hist_ncols_ != ncols || hist_slots_ < slots || hist_has_u_ != with_u) {
        HIP_CHECK(hipStreamSynchronize(stream_));
        if (hist_) (void)hipFree(hist_);
        if (hist_u_) (void)hipFree(hist_u_);
        if (hist_b_) (void)hipFree(hist_b_);
        hist_ = hist_u_ = hist_b_ = nullptr;
        HIP_CHECK(hipMalloc(&hist_, sizeof(double) * np_ * ncols * slots));
        HIP_CHECK(hipMalloc(&hist_b_, sizeof(double) * np_ * ncols * slots));
        if (with_u) HIP_CHECK(hipMalloc(&hist_u_, sizeof(double) * size_ * ncols * slots));
        hist_ncols_ = ncols;
        hist_slots_ = slots;
        hist_has_u_ = with_u;
        hist_count_ = 0;
        hist_committed_ = 0;
      }
      const int slot = (int)(hist_count_ % slots);
      hist_shift_[slot] = shift_cur_;
      hist_b_ok_[slot] = b != nullptr;
      rec.p = hist_ + (int64_t)slot * np_ * ncols;
      if (b) copy_cols(b, np_, hist_b_ + (int64_t)slot * np_ * ncols, np_, np_, ncols, stream_);
      if (with_u) rec.u = hist_u_ + (int64_t)slot * size_ * ncols;
      ++hist_count_;
    }
    return rec;
  }
  // The history extrapolation of the velocities (mixed solve), into out; false when
  // there is none.
  bool history_velocity(int ncols, double* out) {
    if (history_period_ <= 0 || !hist_has_u_ || hist_ncols_ != ncols) return false;
    return extrapolate_history(hist_u_, size_ * ncols, out);
  }

  Outcome solve(const double* f, int64_t f_stride, int ncols, double mass, double visc, double* vel,
                int64_t vel_stride, double* pres, int64_t pres_stride, const double* guess, int64_t guess_stride) {
    if (opts_.precision == MAC_PRECISION_MIXED)
      return solve_mixed(f, f_stride, ncols, mass, visc, vel, vel_stride, pres, pres_stride, guess, guess_stride);
    mixed_ = false;
    ensure_capacity(ncols);
    ncols_cur_ = ncols;
    shift_cur_ = mass > 0.0 ? visc / mass : 0.0;
    std::vector<int> all(ncols);
    for (int c = 0; c < ncols; ++c) all[c] = c;
    Outcome out;
    out.iterations.assign(ncols, 0);
    // W = L^-1 f ; b = -D W ; b -= mean(b)
    velocity_inverse(f, f_stride, w_, size_, all, mass, visc);
    divergence(DT_F64, w_, packed_, g_, b_, np_, -1.0, ncols, RED_SUM, partials_, nullptr, stream_, nullptr,
               fin(sums_));
    sub_mean_dot(b_, np_, np_, ncols, sums_, b_, partials_, nullptr, stream_, fin(sums2_));  // ||b||^2
    reduce(b_, nullptr, np_, np_, ncols, RED_MAXABS, partials_, nullptr, stream_, fin(sums3_));
    cg_init(cg_, sums2_, sums3_, opts_.krylov_tolerance, ncols, stream_);
    // x = 0, r = b  (or x = guess, r = b - S guess on columns whose CG runs)
    HIP_CHECK(hipMemsetAsync(x_, 0, sizeof(double) * np_ * ncols, stream_));
    copy_cols(b_, np_, r_, np_, np_, ncols, stream_);
    HIP_CHECK(hipMemcpyAsync(rr_, sums2_, sizeof(double) * ncols, hipMemcpyDeviceToDevice, stream_));
    out.warm = false;
    if (guess != nullptr) {
      start_from(guess, guess_stride, ncols, mass, visc);
    } else if (learned_guess(b_, ncols)) {
      out.warm = true;
      if (net_base_ != nullptr) {
        out.network = true;
        start_two_term(net_base_, ncols, mass, visc);
      } else {
        start_from(io_g_, np_, ncols, mass, visc, true);
      }
    }
    pcg(ncols, mass, visc, out.iterations, true);
    // p = x - mean(x)
    reduce(x_, nullptr, np_, np_, ncols, RED_SUM, partials_, nullptr, stream_, fin(sums_));
    sub_mean_dot(x_, np_, np_, ncols, sums_, nullptr, nullptr, nullptr, stream_);
    if (opts_.fp64_refine > 0) {
      refine_fp64(f, f_stride, ncols, mass, visc, vel, vel_stride, pres, pres_stride, out);
      return out;
    }
    copy_cols(x_, np_, pres, pres_stride, np_, ncols, stream_);
    remember_pressure(x_, ncols, nullptr, b_);
    // u = L^-1 (f + D^T p)
    divergence_transpose(x_, np_, nullptr, f, f_stride, DT_F64, w_, packed_, g_, ncols, nullptr, stream_);
    velocity_inverse(w_, size_, vel, vel_stride, all, mass, visc);
    checks(f, f_stride, vel, vel_stride, x_, ncols, mass, visc, out);
    return out;
  }

  // FP64 defect correction after the reference solve (p in x_, mean-zero): with the
  // FP64 residuals r_u = f - (mass u + visc K u) + D^T p and r_c = -D u of the current
  // iterate, each sweep solves the Stokes system for the correction exactly as the
  // reference solves for u and p, (W = L^-1 r_u, S dp = r_c - D W by the Schur CG to
  // inner_tolerance, du = L^-1 (r_u + D^T dp)), all in FP64.  Exact operators make each
  // sweep contract the defect by about the inner tolerance; the loop stops once the
  // reference's residual and max|D u| are both <= refine_tolerance.
  void refine_fp64(const double* f, int64_t f_stride, int ncols, double mass, double visc, double* vel,
                   int64_t vel_stride, double* pres, int64_t pres_stride, Outcome& out) {
    std::vector<int> all(ncols), inner(ncols);
    for (int c = 0; c < ncols; ++c) all[c] = c;
    const double* fc = f;
    if (f_stride != size_) {
      copy_cols(f, f_stride, fcomp_, size_, size_, ncols, stream_);
      fc = fcomp_;
    }
    const double* b0 = b_;
    copy_cols(b0, np_, bzero_, np_, np_, ncols, stream_);  // the zero start's b, for the history
    copy_cols(x_, np_, pacc_, np_, np_, ncols, stream_);
    divergence_transpose(pacc_, np_, nullptr, fc, size_, DT_F64, w_, packed_, g_, ncols, nullptr, stream_);
    velocity_inverse(w_, size_, uacc_, size_, all, mass, visc);
    out.sweeps = 1;
    for (int sweep = 0;; ++sweep) {
      defect_max(uacc_, fc, pacc_, g_, ncols, mass, visc, partials_, stream_, ru64_, fin(sums_), rc_, -1.0,
                 fin(sums2_), DT_F64, nullptr);
      post_values(sums_, 3, sums2_, ncols, mailbox_dev_, ++seq_, stream_);
      const volatile Mailbox* mb = wait_post(seq_);
      out.residual = mb->vals[0] / (1.0 + mb->vals[1]);
      out.fmax = mb->vals[1];
      out.vmax = mb->vals[2];
      out.divergence = 0.0;
      for (int c = 0; c < ncols; ++c) out.divergence = std::max(out.divergence, (double)mb->vals[3 + c]);
      if (sweep == 0) {
        out.first_residual = out.residual;
        out.first_divergence = out.divergence;
      }
      const bool done = out.residual <= opts_.refine_tolerance && out.divergence <= opts_.refine_tolerance;
      if (done || sweep >= opts_.fp64_refine || !std::isfinite(out.residual) || !std::isfinite(out.divergence))
        break;
      velocity_inverse(ru64_, size_, w_, size_, all, mass, visc);  // W = L^-1 r_u
      divergence(DT_F64, w_, packed_, g_, b_, np_, -1.0, ncols, RED_SUM, partials_, nullptr, stream_, rc_,
                 fin(sums_));                                      // b = r_c - D W
      sub_mean_dot(b_, np_, np_, ncols, sums_, b_, partials_, nullptr, stream_, fin(sums2_));
      reduce(b_, nullptr, np_, np_, ncols, RED_MAXABS, partials_, nullptr, stream_, fin(sums3_));
      cg_init(cg_, sums2_, sums3_, opts_.inner_tolerance, ncols, stream_);
      HIP_CHECK(hipMemsetAsync(x_, 0, sizeof(double) * np_ * ncols, stream_));
      copy_cols(b_, np_, r_, np_, np_, ncols, stream_);
      HIP_CHECK(hipMemcpyAsync(rr_, sums2_, sizeof(double) * ncols, hipMemcpyDeviceToDevice, stream_));
      pcg(ncols, mass, visc, inner, false);
      for (int c = 0; c < ncols; ++c) out.iterations[c] += inner[c];
      add_cols(pacc_, x_, np_, ncols, stream_);
      divergence_transpose(x_, np_, nullptr, ru64_, size_, DT_F64, w2_, packed_, g_, ncols, nullptr, stream_);
      velocity_inverse(w2_, size_, w_, size_, all, mass, visc);    // du = L^-1 (r_u + D^T dp)
      add_cols(uacc_, w_, size_, ncols, stream_);
      out.sweeps = sweep + 2;
    }
    reduce(pacc_, nullptr, np_, np_, ncols, RED_SUM, partials_, nullptr, stream_, fin(sums_));
    sub_mean_dot(pacc_, np_, np_, ncols, sums_, nullptr, nullptr, nullptr, stream_);
    copy_cols(pacc_, np_, pres, pres_stride, np_, ncols, stream_);
    copy_cols(uacc_, size_, vel, vel_stride, size_, ncols, stream_);
    remember_pressure(pacc_, ncols, nullptr, bzero_);
    copy_cols(bzero_, np_, b_, np_, np_, ncols, stream_);  // b_ is the solve's zero-start b again
  }

  // MIXED: FP64 defect correction around bf16-transform inner solves.
  //   u = 0, p = guess (or 0)
  //   repeat: r_u = f - (mass u + visc K u) + D^T p, r_c = -D u   (FP64; r_u kept in the
  //           precision the transform reads, FP32 for the split form)
  //           stop when max|r_u|/(1+max|f|) <= refine_tol and max|D u| <= refine_tol
  //           W = L~^-1 r_u; b = r_c - D W; dp = PCG(S~, b, inner_tol); p += dp
  //           u += L~^-1 (r_u + D^T dp) = W + sum_k alpha_k w_k, where w_k = L~^-1 D^T p_k
  //           is the intermediate of the k-th Schur product (no extra transforms)
  // (L~, S~: the velocity inverse and Schur operator with bf16 transforms.)
  Outcome solve_mixed(const double* f, int64_t f_stride, int ncols, double mass, double visc, double* vel,
                      int64_t vel_stride, double* pres, int64_t pres_stride, const double* guess,
                      int64_t guess_stride) {
    ensure_capacity(ncols);
    ncols_cur_ = ncols;
    shift_cur_ = mass > 0.0 ? visc / mass : 0.0;
    std::vector<int> all(ncols);
    for (int c = 0; c < ncols; ++c) all[c] = c;
    Outcome out;
    out.iterations.assign(ncols, 0);
    out.sweeps = 0;
    const double* fc = f;
    if (f_stride != size_) {
      copy_cols(f, f_stride, fcomp_, size_, size_, ncols, stream_);
      fc = fcomp_;
    }
    bool u_set = false;  // uacc_ holds a velocity guess (else it starts from zero)
    // Starting pressure: the caller's guess, else the warm start's (history, previous
    // solve or network), else zero.  It is the outer iterate, so the first FP64 defect
    // measures its true error and every sweep improves on it; started inside a bf16 inner
    // solve instead, its accuracy below the bf16 operator's would be lost.
    out.warm = false;
    bool bzero_ok = false;  // bzero_ holds this solve's zero-start Schur right-hand side
    if (guess) {
      copy_cols(guess, guess_stride, pacc_, np_, np_, ncols, stream_);
    } else if ((warm_mode_ == MAC_WARM_HISTORY || warm_mode_ == MAC_WARM_PREVIOUS ||
                (warm_mode_ == MAC_WARM_NETWORK && recorded_steps(ncols) >= net_depth_limit_)) &&
               learned_guess(nullptr, ncols, pacc_)) {
      // With a history, the velocities start from their own extrapolation too: a good
      // pressure with u = 0 leaves an O(1) velocity defect, whose bf16 correction (after
      // D) would swamp the pressure's accuracy.  (A network handle with a deep enough
      // history takes the extrapolation alone, without forming b for the network.)
      if (history_period_ > 0) u_set = history_velocity(ncols, uacc_);
      out.warm = true;
    } else {
      HIP_CHECK(hipMemsetAsync(pacc_, 0, sizeof(double) * np_ * ncols, stream_));
    }
    if (guess == nullptr && !out.warm && warm_mode_ == MAC_WARM_NETWORK) {
      // The network reads the Schur right-hand side of the zero start, b = -D L~^-1 f
      // (mean removed): one bf16 velocity inverse before the sweeps.
      mixed_ = true;
      velocity_inverse_bf16(fc, DT_F64, packed_, du_, DT_F32, size_, all, mass, visc);
      divergence(DT_F32, du_, packed_, g_, b_, np_, -1.0, ncols, RED_SUM, partials_, nullptr, stream_, nullptr,
                 fin(sums_));
      sub_mean_dot(b_, np_, np_, ncols, sums_, nullptr, nullptr, nullptr, stream_);
      mixed_ = false;
      copy_cols(b_, np_, bzero_, np_, np_, ncols, stream_);  // the sweeps reuse b_
      bzero_ok = true;
      if (learned_guess(b_, ncols, pacc_)) {
        if (history_period_ > 0) u_set = history_velocity(ncols, uacc_);
        out.warm = true;
        out.network = net_base_ != nullptr;
      }
    }
    if (!u_set) HIP_CHECK(hipMemsetAsync(uacc_, 0, sizeof(double) * size_ * ncols, stream_));
    std::vector<int> inner(ncols);
    // r_u goes straight into the transform's input buffer, in the precision the split
    // (or plain bf16) transform keeps of it: FP32 holds every bit of its bf16 hi + lo.
    void* ru_buf = split_velocity() ? static_cast<void*>(tfb_) : static_cast<void*>(tb_);
    const int ru_type = split_velocity() ? DT_F32 : DT_BF16;
    for (int sweep = 0;; ++sweep) {
      // FP64 residuals of the full system at the current iterate.
      defect_max(uacc_, fc, pacc_, g_, ncols, mass, visc, partials_, stream_, ru_buf, fin(sums_), rc_, -1.0,
                 fin(sums2_), ru_type, &lay_b_);
      post_values(sums_, 3, sums2_, ncols, mailbox_dev_, ++seq_, stream_);
      // From a history start a sweep is all but certain (the extrapolation's defect lies
      // above the bar), so the inner solve's first transforms are queued before the host
      // reads the defect and run while it does; should the start already meet the bar,
      // they are dropped (they write scratch only).
      const bool ahead = sweep == 0 && u_set;
      if (ahead) {
        mixed_ = true;
        velocity_inverse_bf16(ru_buf, ru_type, lay_b_, du_, DT_F32, size_, all, mass, visc);
        mixed_ = false;
      }
      const volatile Mailbox* mb = wait_post(seq_);
      out.residual = mb->vals[0] / (1.0 + mb->vals[1]);
      out.fmax = mb->vals[1];
      out.vmax = mb->vals[2];
      out.divergence = 0.0;
      for (int c = 0; c < ncols; ++c) out.divergence = std::max(out.divergence, (double)mb->vals[3 + c]);
      if (sweep == 0) {
        // The first defect also measures max|f|: a non-finite right-hand side is an
        // invalid argument, as the reference rejects it (tensor_stokes.py:246-257).
        if (!std::isfinite(out.fmax)) throw ArgFailure("Invalid Stokes RHS");
        out.first_residual = out.residual;
        out.first_divergence = out.divergence;
      }
      const bool done = out.residual <= opts_.refine_tolerance && out.divergence <= opts_.refine_tolerance;
      if (done || sweep >= opts_.max_sweeps || !std::isfinite(out.residual) || !std::isfinite(out.divergence))
        break;
      // Inner solve for the correction (dp, du).
      mixed_ = true;
      if (!ahead) velocity_inverse_bf16(ru_buf, ru_type, lay_b_, du_, DT_F32, size_, all, mass, visc);
      divergence(DT_F32, du_, packed_, g_, b_, np_, -1.0, ncols, RED_SUM, partials_, nullptr, stream_, rc_,
                 fin(sums_));
      // The inner CG's recurrence residual is r = b - S~ x = -D (u + du): the sweep's new
      // divergence.  Besides the relative test, a column stops once max|r| is half the bar,
      // where the sweep is certain to meet it (the momentum defect it leaves is the split
      // operator's error on an already small correction).
      // This sweep's inner tolerance: a sweep cuts the defect by about the inner tolerance,
      // so 0.25 x bar / defect would reach the bar in one sweep; inner_tolerance caps it
      // (a history start's small defect), 1e-4 floors it (a cold start's large one).
      const double cap = opts_.inner_tolerance;
      const double sweep_tol = std::min(cap, std::max(std::min(1e-4, cap), 0.25 * opts_.refine_tolerance /
                                                                              std::max(out.residual, out.divergence)));
      // b -= mean, r = b, x = 0, with (b, b) and max|b|; the CG's state and its first
      // check in the same launch sequence.
      sweep_start_cg(b_, r_, x_, np_, ncols, sums_, partials_, sums2_, sums3_, cg_, sweep_tol, rr_,
                     0.5 * opts_.refine_tolerance, flags_, stream_, tfb_);
      r32_ = tfb_;
      first_checked_ = true;
      accumulate_ = true;
      pcg(ncols, mass, visc, inner, false);
      accumulate_ = false;
      r32_ = nullptr;
      for (int c = 0; c < ncols; ++c) out.iterations[c] += inner[c];
      // p += dp, and u += du with du = L~^-1 (r_u + D^T dp) = W + sum_k alpha_k L~^-1 D^T p_k,
      // already summed in du_.
      add_pair(pacc_, x_, np_ * ncols, uacc_, du_, size_ * ncols, stream_);
      mixed_ = false;
      out.sweeps = sweep + 1;
    }
    // p mean-zero; hand out u and p, and record them for the warm starts, in one pass each.
    reduce(pacc_, nullptr, np_, np_, ncols, RED_SUM, partials_, nullptr, stream_, fin(sums_));
    const Record rec = record(ncols, true, bzero_ok ? bzero_ : nullptr);
    finish_cols(pacc_, np_, ncols, sums_, pres, pres_stride, pprev_, np_, rec.p, np_, stream_);
    finish_cols(uacc_, size_, ncols, nullptr, vel, vel_stride, rec.u, size_, nullptr, 0, stream_);
    return out;
  }

 private:
  // One GEMM pass of a 2-D transform: C = A . Q^T with A [M][K] (lda), Q (or Q^T) as
  // Bt [N][K] (ldb = n), C written transposed [N][M] (ldc) when the batch says so.
  static void add_pass(GemmBatch& b, const double* A, int lda, const double* Q, int n, double* C, int ldc,
                       int M, int N, int K, const double* ea, const double* eb) {
    add_pass_any(b, A, lda, Q, n, C, ldc, M, N, K, ea, eb);
  }
  static void add_pass_any(GemmBatch& b, const void* A, int lda, const void* Q, int n, void* C, int ldc,
                           int M, int N, int K, const double* ea, const double* eb, const void* Qlo = nullptr) {
    GemmEntry& e = b.e[b.count++];
    e.A = A;
    e.Bt = Q;
    e.Bt_lo = Qlo;
    e.C = C;
    e.aux = nullptr;
    e.ea = ea;
    e.eb = eb;
    e.M = M;
    e.N = N;
    e.K = K;
    e.lda = lda;
    e.ldb = n;
    e.ldc = ldc;
  }

  // q = S p = D L^-1 D^T p, mean removed; optional partial (y, q) dots after the projection.
  void schur(const double* pin, double* qout, const std::vector<int>& cols, double mass, double visc,
             const double* dot_with, double* dot_out = nullptr) {
    const int ncols = ncols_cur_;
    const CgState* st = use_state_ ? cg_ : nullptr;
    if (mixed_) {
      const int dt_type = split_velocity() ? DT_F32 : DT_BF16;
      void* dt_buf = split_velocity() ? static_cast<void*>(tfb_) : static_cast<void*>(tb_);
      divergence_transpose(pin, np_, nullptr, nullptr, 0, dt_type, dt_buf, lay_b_, g_, ncols, st, stream_);
      velocity_inverse_bf16(dt_buf, dt_type, lay_b_, wf_, DT_F32, size_, cols, mass, visc);
      divergence(DT_F32, wf_, packed_, g_, qout, np_, 1.0, ncols, RED_SUM, partials_, st, stream_, nullptr,
                 fin(sums_));
    } else {
      divergence_transpose(pin, np_, nullptr, nullptr, 0, DT_F64, w2_, packed_, g_, ncols, st, stream_);
      velocity_inverse(w2_, size_, w_, size_, cols, mass, visc);
      divergence(DT_F64, w_, packed_, g_, qout, np_, 1.0, ncols, RED_SUM, partials_, st, stream_, nullptr,
                 fin(sums_));
    }
    sub_mean_dot(qout, np_, np_, ncols, sums_, dot_with, partials_, st, stream_,
                 dot_with != nullptr ? fin(dot_out) : Fin{});
  }

  // x = guess, r = b - S x, rr = (r, r) on the columns whose CG runs (a column with a
  // zero right-hand side keeps x = 0, as the reference skips its CG).
  void start_from(const double* guess, int64_t guess_stride, int ncols, double mass, double visc,
                  bool scale_guess = false) {
    std::vector<CgState> st(ncols);
    HIP_CHECK(hipMemcpyAsync(st.data(), cg_, sizeof(CgState) * ncols, hipMemcpyDeviceToHost, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
    std::vector<int> live;
    for (int c = 0; c < ncols; ++c) {
      if (!st[c].active) continue;
      live.push_back(c);
      HIP_CHECK(hipMemcpyAsync(x_ + (int64_t)c * np_, guess + (int64_t)c * guess_stride, sizeof(double) * np_,
                               hipMemcpyDeviceToDevice, stream_));
    }
    if (live.empty()) return;
    use_state_ = true;
    schur(x_, q_, live, mass, visc, nullptr);
    use_state_ = false;
    if (scale_guess) {
      // Galerkin step along the guess: never a worse start, in the S-norm, than zero.
      reduce(x_, b_, np_, np_, ncols, RED_SUM, partials_, cg_, stream_, fin(sums_));
      reduce(x_, q_, np_, np_, ncols, RED_SUM, partials_, cg_, stream_, fin(sums2_));
      theta_start(x_, r_, b_, q_, sums_, sums2_, np_, np_, ncols, cg_, stream_);
      if (accumulate_) accum_velocity(du_, wf_, size_, size_, ncols, cg_, sums_, sums2_, stream_);
    } else {
      sub_cols(b_, q_, r_, np_, np_, ncols, cg_, stream_);
      if (accumulate_) accum_velocity(du_, wf_, size_, size_, ncols, cg_, nullptr, sums_, stream_);
    }
    reduce(r_, r_, np_, np_, ncols, RED_SUM, partials_, cg_, stream_, fin(rr_));
  }

  // The network's start in the FP64 solve: the combination a g + beta c of the base g and
  // the network's correction c = guess - g with the smallest starting residual
  // ||b - S(a g + beta c)||_2, never larger than the best multiple of the base alone.
  // S c costs one Schur product (as the one-term start); S g is the base's recorded
  // right-hand side b - (b - b_base), exact to the recorded solves' tolerance.  On the
  // columns whose CG runs.
  void start_two_term(const double* g, int ncols, double mass, double visc) {
    std::vector<CgState> st(ncols);
    HIP_CHECK(hipMemcpyAsync(st.data(), cg_, sizeof(CgState) * ncols, hipMemcpyDeviceToHost, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
    std::vector<int> live;
    for (int c = 0; c < ncols; ++c)
      if (st[c].active) live.push_back(c);
    if (live.empty()) return;
    const int64_t n = np_ * ncols;
    {
      const double w[2] = {1.0, -1.0};
      const double* cdiff[2] = {io_g_, g};
      extrapolate(cdiff, w, 2, z_, n, stream_);  // c = guess - base
      const double* bdiff[2] = {b_, warm_db_};
      extrapolate(bdiff, w, 2, p_, n, stream_);  // b_base = b - (b - b_base)
    }
    use_state_ = true;
    schur(z_, q_, live, mass, visc, nullptr);    // q = S c
    use_state_ = false;
    const double* pairs[5][2] = {{p_, p_}, {p_, q_}, {q_, q_}, {p_, b_}, {q_, b_}};
    for (int k = 0; k < 5; ++k)
      reduce(pairs[k][0], pairs[k][1], np_, np_, ncols, RED_SUM, partials_, cg_, stream_, fin(gal_ + k * ncols));
    galerkin2_start(x_, r_, g, z_, b_, p_, q_, gal_, np_, ncols, cg_, stream_);
    reduce(r_, r_, np_, np_, ncols, RED_SUM, partials_, cg_, stream_, fin(rr_));
  }

  // Schur PCG on all columns at once (x, r, rr, cg_ prepared by the caller).  With
  // `fail`, a column not converged after max_iterations raises MAC_NOT_CONVERGED.
  void pcg(int ncols, double mass, double visc, std::vector<int>& iters, bool fail) {
    ncols_cur_ = ncols;
    use_state_ = true;
    std::vector<int> active;
    bool settled = false;  // every column left the loop through the convergence test
    for (int it = 0; it < opts_.max_iterations; ++it) {
      if (it == 0 && mixed_) {
        // A mixed sweep's inner CG starts from x = 0 on b != 0 (a sweep runs only while the
        // defect is above the bar): every column with b != 0 takes at least one iteration,
        // so the first check runs on the device alone and the iteration is queued behind
        // it without a host round trip.  A column with b = 0 is inactive on the device and
        // every vector kernel skips it; its transform columns compute into scratch only.
        if (!first_checked_) cg_check(cg_, rr_, flags_, ncols, stream_);
        first_checked_ = false;
        active.clear();
        for (int c = 0; c < ncols; ++c) active.push_back(c);
      } else {
        // From the second iteration on, (r, r) arrives as the update's partials.
        cg_check_post(cg_, rr_, flags_, ncols, mailbox_dev_, ++seq_, stream_, it == 0 ? nullptr : partials_,
                      reduce_blocks(np_));
        const volatile Mailbox* mb = wait_post(seq_);
        active.clear();
        for (int c = 0; c < ncols; ++c)
          if (mb->flags[c]) active.push_back(c);
        if (active.empty()) {
          use_state_ = false;
          for (int c = 0; c < ncols; ++c) {
            iters[c] = mb->iters[c];
            if (fail && !mb->converged[c])
              throw SolveFailure(MAC_NOT_CONVERGED,
                                 "Pressure Schur CG failed with info=" + std::to_string(opts_.max_iterations));
          }
          settled = true;
          break;
        }
      }
      if (mixed_) {
        // The mixed inner CG folds the mean removals of z and q into their consumers:
        // rho = (z - mean z, r) from one pass of sums, p = (z - mean z) + beta p, and
        // alpha = rho / (q - mean q, p), r -= alpha (q - mean q) from sums the
        // divergence accumulates (the same iteration in exact arithmetic, two fewer
        // passes and four fewer launches).
        double* zmean = means_;
        double* qmean = means_ + kMaxCols;
        const int gxs = fft_pressure() ? fft_row_blocks(pblk_.ny, pblk_.mx->fft.M) : 0;
        if (fft_pressure() && (size_t)3 * ncols * gxs <= partials_cap_) {
          // z, and the sums rho needs, from the pressure inverse's last pass.
          pressure_inverse_fft(r_, z_, active, mass, visc, r32_, partials_, ncols);
          mixed_rho_partials(cg_, partials_, gxs, np_, ncols, zmean, stream_);
        } else {
          pressure_inverse_bf16(r_, z_, active, mass, visc, r32_);
          mixed_rho(z_, r_, np_, ncols, cg_, partials_, zmean, stream_);
        }
        mixed_update_p(p_, z_, zmean, np_, ncols, cg_, stream_);
        if (fft_velocity()) {
          // D^T p formed by the velocity inverse's first pass from the pressure rows.
          velocity_inverse_fft(nullptr, DT_F64, lay_b_, wf_, DT_F32, size_, active, mass, visc, p_);
        } else {
          const int dt_type = split_velocity() ? DT_F32 : DT_BF16;
          void* dt_buf = split_velocity() ? static_cast<void*>(tfb_) : static_cast<void*>(tb_);
          divergence_transpose(p_, np_, nullptr, nullptr, 0, dt_type, dt_buf, lay_b_, g_, ncols, cg_, stream_);
          velocity_inverse_bf16(dt_buf, dt_type, lay_b_, wf_, DT_F32, size_, active, mass, visc);
        }
        divergence_sums3(DT_F32, wf_, packed_, g_, q_, np_, 1.0, p_, ncols, partials_, cg_, stream_);
        mixed_alpha(partials_, np_, ncols, cg_, qmean, stream_);
        mixed_update_xr(x_, r_, p_, q_, qmean, np_, ncols, cg_, partials_, stream_, r32_);
        if (accumulate_) accum_velocity(du_, wf_, size_, size_, ncols, cg_, nullptr, nullptr, stream_);
        continue;
      }
      // z = mass (D D^T)^+ r + visc r, mean removed; rho = (r, z)
      pressure_inverse(r_, np_, z_, np_, active, true, mass, visc);
      reduce(z_, nullptr, np_, np_, ncols, RED_SUM, partials_, cg_, stream_, fin(sums_));
      sub_mean_dot(z_, np_, np_, ncols, sums_, r_, partials_, cg_, stream_);
      cg_rho_fin(cg_, partials_, reduce_blocks(np_), sums2_, ncols, stream_);
      cg_update_p(p_, z_, np_, np_, ncols, cg_, stream_);
      // q = S p; alpha = rho / (p, q)
      schur(p_, q_, active, mass, visc, p_);
      cg_alpha_fin(cg_, partials_, reduce_blocks(np_), sums2_, ncols, stream_);
      cg_update_xr(x_, r_, p_, q_, np_, np_, ncols, cg_, partials_, stream_);
      if (accumulate_) accum_velocity(du_, wf_, size_, size_, ncols, cg_, nullptr, nullptr, stream_);
    }
    use_state_ = false;
    if (settled) return;
    // A column still active after max_iterations completed iterations failed (the
    // convergence test is only made at the top of an iteration, as in scipy's cg).
    std::vector<CgState> st(ncols);
    HIP_CHECK(hipMemcpyAsync(st.data(), cg_, sizeof(CgState) * ncols, hipMemcpyDeviceToHost, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
    for (int c = 0; c < ncols; ++c) {
      iters[c] = st[c].iters;
      if (fail && !st[c].converged)
        throw SolveFailure(MAC_NOT_CONVERGED,
                           "Pressure Schur CG failed with info=" + std::to_string(opts_.max_iterations));
    }
  }

  void checks(const double* f, int64_t f_stride, const double* vel, int64_t vel_stride, const double* p, int ncols,
              double mass, double visc, Outcome& out) {
    // The defect kernel reads contiguous columns; the caller's strides are compacted.
    const double* fc = f;
    const double* vc = vel;
    if (f_stride != size_) {
      copy_cols(f, f_stride, w_, size_, size_, ncols, stream_);
      fc = w_;
    }
    if (vel_stride != size_) {
      copy_cols(vel, vel_stride, w2_, size_, size_, ncols, stream_);
      vc = w2_;
    }
    defect_max(vc, fc, p, g_, ncols, mass, visc, partials_, stream_, nullptr, fin(sums_), nullptr, 1.0,
               fin(sums2_));
    post_values(sums_, 3, sums2_, ncols, mailbox_dev_, ++seq_, stream_);
    const volatile Mailbox* mb = wait_post(seq_);
    out.residual = mb->vals[0] / (1.0 + mb->vals[1]);
    out.fmax = mb->vals[1];
    out.vmax = mb->vals[2];
    out.divergence = 0.0;
    for (int c = 0; c < ncols; ++c) out.divergence = std::max(out.divergence, (double)mb->vals[3 + c]);
  }

  void build_bases() {
    const int nx = g_.nx, ny = g_.ny;
    ublk_.ny = ny; ublk_.nx = nx - 1;
    ublk_.my = matrix(ny, KIND_DST2); ublk_.mx = matrix(nx - 1, KIND_DST1);
    vblk_.ny = ny - 1; vblk_.nx = nx;
    vblk_.my = matrix(ny - 1, KIND_DST1); vblk_.mx = matrix(nx, KIND_DST2);
    pblk_.ny = ny; pblk_.nx = nx;
    pblk_.my = matrix(ny, KIND_DCT2); pblk_.mx = matrix(nx, KIND_DCT2);
    set_lam(ublk_, KIND_DST2, KIND_DST1);
    set_lam(vblk_, KIND_DST1, KIND_DST2);
    set_lam(pblk_, KIND_DCT2, KIND_DCT2);
    fft_ok_vel_ = ublk_.mx->fft_ok && ublk_.my->fft_ok && vblk_.mx->fft_ok && vblk_.my->fft_ok;
    fft_ok_p_ = pblk_.mx->fft_ok && pblk_.my->fft_ok;
    double lmax = 0.0;
    for (double a : pblk_.host_lam_y)
      for (double b : pblk_.host_lam_x) lmax = std::max(lmax, a + b);
    p_thresh_ = 100.0 * 2.220446049250313e-16 * lmax;
  }

  void set_lam(Block& B, int ky, int kx) {
    B.host_lam_y = eigenvalues_1d(B.ny, g_.hy, ky);
    B.host_lam_x = eigenvalues_1d(B.nx, g_.hx, kx);
    B.lam_y = upload(B.host_lam_y);
    B.lam_x = upload(B.host_lam_x);
    auto eo = [](const std::vector<double>& lam) {
      const int n = (int)lam.size();
      std::vector<double> out(n);
      for (int i = 0; i < n; ++i) out[i] = lam[eo_coefficient(i, n)];
      return out;
    };
    B.lam_y_eo = upload(eo(B.host_lam_y));
    B.lam_x_eo = upload(eo(B.host_lam_x));
  }

  const DevMatrix* matrix(int n, int kind) {
    for (auto& m : mats_)
      if (m->n == n && m->kind == kind) return m.get();
    auto m = std::make_unique<DevMatrix>();
    m->n = n;
    m->kind = kind;
    m->ld8 = round8(n);
    std::vector<double> q = transform_matrix(n, kind);
    std::vector<double> qt((size_t)n * n);
    for (int a = 0; a < n; ++a)
      for (int b = 0; b < n; ++b) qt[(size_t)b * n + a] = q[(size_t)a * n + b];
    m->q = upload(q);
    m->qt = upload(qt);
    std::vector<__bf16> qb((size_t)n * m->ld8, (__bf16)0.0f), qtb((size_t)n * m->ld8, (__bf16)0.0f);
    std::vector<__bf16> qbl((size_t)n * m->ld8, (__bf16)0.0f), qtbl((size_t)n * m->ld8, (__bf16)0.0f);
    for (int a = 0; a < n; ++a)
      for (int b = 0; b < n; ++b) {
        const double v = q[(size_t)a * n + b], vt = qt[(size_t)a * n + b];
        const __bf16 h = (__bf16)(float)v, ht = (__bf16)(float)vt;
        qb[(size_t)a * m->ld8 + b] = h;
        qtb[(size_t)a * m->ld8 + b] = ht;
        qbl[(size_t)a * m->ld8 + b] = (__bf16)(float)(v - (double)(float)h);
        qtbl[(size_t)a * m->ld8 + b] = (__bf16)(float)(vt - (double)(float)ht);
      }
    m->qb = static_cast<__bf16*>(upload_raw(qb.data(), qb.size() * sizeof(__bf16)));
    m->qtb = static_cast<__bf16*>(upload_raw(qtb.data(), qtb.size() * sizeof(__bf16)));
    m->qb_lo = static_cast<__bf16*>(upload_raw(qbl.data(), qbl.size() * sizeof(__bf16)));
    m->qtb_lo = static_cast<__bf16*>(upload_raw(qtbl.data(), qtbl.size() * sizeof(__bf16)));
    // Even/odd halves and the coefficient-ordered transpose.
    m->h = (n + 1) / 2;
    m->ldh = round8(m->h);
    const int h = m->h, hf = n / 2;
    std::vector<__bf16> qe((size_t)h * m->ldh, (__bf16)0.0f), qel(qe.size(), (__bf16)0.0f);
    std::vector<__bf16> qo((size_t)h * m->ldh, (__bf16)0.0f), qol(qo.size(), (__bf16)0.0f);
    std::vector<__bf16> qtp((size_t)n * m->ld8, (__bf16)0.0f), qtpl(qtp.size(), (__bf16)0.0f);
    auto split = [](double v, __bf16& hi, __bf16& lo) {
      hi = (__bf16)(float)v;
      lo = (__bf16)(float)(v - (double)(float)hi);
    };
    for (int r = 0; r < h; ++r)
      for (int j = 0; j < h; ++j) {
        const bool middle = j >= hf;  // odd n: the self-mirrored column
        split(q[(size_t)(2 * r) * n + j] * (middle ? 0.5 : 1.0), qe[(size_t)r * m->ldh + j],
              qel[(size_t)r * m->ldh + j]);
        if (r < hf && !middle) split(q[(size_t)(2 * r + 1) * n + j], qo[(size_t)r * m->ldh + j], qol[(size_t)r * m->ldh + j]);
      }
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i)
        split(q[(size_t)eo_coefficient(i, n) * n + j], qtp[(size_t)j * m->ld8 + i], qtpl[(size_t)j * m->ld8 + i]);
    m->qe = static_cast<__bf16*>(upload_raw(qe.data(), qe.size() * sizeof(__bf16)));
    m->qe_lo = static_cast<__bf16*>(upload_raw(qel.data(), qel.size() * sizeof(__bf16)));
    m->qo = static_cast<__bf16*>(upload_raw(qo.data(), qo.size() * sizeof(__bf16)));
    m->qo_lo = static_cast<__bf16*>(upload_raw(qol.data(), qol.size() * sizeof(__bf16)));
    m->qtp = static_cast<__bf16*>(upload_raw(qtp.data(), qtp.size() * sizeof(__bf16)));
    m->qtp_lo = static_cast<__bf16*>(upload_raw(qtpl.data(), qtpl.size() * sizeof(__bf16)));
    if (fft_length_ok(n, kind)) {
      std::vector<float2> ct, cs;
      int M = 0;
      make_fft_tables(n, kind, &M, ct, cs);
      m->fft.n = n;
      m->fft.kind = kind;
      m->fft.M = M;
      m->fft.ct = static_cast<const float2*>(upload_raw(ct.data(), ct.size() * sizeof(float2)));
      m->fft.cs = static_cast<const float2*>(upload_raw(cs.data(), cs.size() * sizeof(float2)));
      m->fft_ok = true;
    }
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

 public:
  void ensure_capacity(int ncols) {
    if (ncols <= cap_) return;
    // Work buffers are replaced, never freed mid-solve; old ones are released here.
    HIP_CHECK(hipStreamSynchronize(stream_));
    for (void* p : work_) {
      (void)hipFree(p);
      allocs_.erase(std::remove(allocs_.begin(), allocs_.end(), p), allocs_.end());
    }
    work_.clear();
    auto take = [&](size_t bytes) {
      void* p = dev_alloc(bytes);
      work_.push_back(p);
      return p;
    };
    const size_t vb = sizeof(double) * size_ * ncols, pb = sizeof(double) * np_ * ncols;
    w_ = (double*)take(vb);
    w2_ = (double*)take(vb);
    s1_ = (double*)take(vb);
    s2_ = (double*)take(vb);
    io_f_ = (double*)take(vb);
    io_v_ = (double*)take(vb);
    io_p_ = (double*)take(pb);
    io_g_ = (double*)take(pb);
    b_ = (double*)take(pb);
    x_ = (double*)take(pb);
    r_ = (double*)take(pb);
    z_ = (double*)take(pb);
    p_ = (double*)take(pb);
    q_ = (double*)take(pb);
    fcomp_ = (double*)take(vb);
    pprev_ = (double*)take(pb);
    bprev_ = (double*)take(pb);
    bzero_ = (double*)take(pb);
    warm_base_ = (double*)take(pb);
    warm_db_ = (double*)take(pb);
    warm_x_ = (float*)take(sizeof(float) * 2 * (np_ / 16 + 64) * ncols);
    warm_y_ = (float*)take(sizeof(float) * (np_ / 16 + 64) * ncols);
    warm_s_ = (double*)take(sizeof(double) * (2 * ncols + 4));
    prev_ncols_ = 0;
    prev_b_ok_ = false;
    uacc_ = (double*)take(vb);
    ru64_ = (double*)take(vb);
    pacc_ = (double*)take(pb);
    rc_ = (double*)take(pb);
    wf_ = (float*)take(sizeof(float) * size_ * ncols);
    du_ = (float*)take(sizeof(float) * size_ * ncols);
    tb_ = (__bf16*)take(sizeof(__bf16) * tstride_ * ncols);
    tb1_ = (__bf16*)take(sizeof(__bf16) * tstride_ * ncols);
    tb2_ = (__bf16*)take(sizeof(__bf16) * tstride_ * ncols);
    tf1_ = (float*)take(sizeof(float) * tstride_ * ncols);
    tf2_ = (float*)take(sizeof(float) * tstride_ * ncols);
    tfb_ = (float*)take(sizeof(float) * tstride_ * ncols);
    fx_ = (fft_ok_vel_ || fft_ok_p_) ? (float*)take(sizeof(float) * fstride_ * ncols) : nullptr;
    const int nb = std::max(defect_blocks(size_, ncols), reduce_blocks(np_) * ncols);
    partials_cap_ = 4 * (size_t)nb + 16;
    partials_ = (double*)take(sizeof(double) * partials_cap_);
    sums_ = (double*)take(sizeof(double) * (ncols + 4));
    means_ = (double*)take(sizeof(double) * 2 * kMaxCols);
    sums2_ = (double*)take(sizeof(double) * (ncols + 4));
    sums3_ = (double*)take(sizeof(double) * (ncols + 4));
    rr_ = (double*)take(sizeof(double) * (ncols + 4));
    cg_ = (CgState*)take(sizeof(CgState) * ncols);
    gal_ = (double*)take(sizeof(double) * 8 * ncols);
    flags_ = (int*)take(sizeof(int) * ncols);
    cap_ = ncols;
  }

  double* io_f_ = nullptr;
  double* io_v_ = nullptr;
  double* io_p_ = nullptr;
  double* io_g_ = nullptr;
  double* scratch_sums() { return sums_; }
  double* scratch_partials() { return partials_; }
  const VelLayout& packed() const { return packed_; }

 private:
  mac_stokes_opts opts_;
  GridDims g_{};
  int64_t nu_ = 0, nv_ = 0, size_ = 0, np_ = 0;
  hipStream_t stream_ = nullptr;
  hipStream_t own_stream_ = nullptr;
  VelLayout packed_{};
  std::vector<std::unique_ptr<DevMatrix>> mats_;
  Block ublk_, vblk_, pblk_;
  double p_thresh_ = 0.0;
  std::vector<void*> allocs_, work_;
  int cap_ = 0;
  int ncols_cur_ = 0;
  bool use_state_ = false;
  double *w_ = nullptr, *w2_ = nullptr, *s1_ = nullptr, *s2_ = nullptr;
  double *b_ = nullptr, *x_ = nullptr, *r_ = nullptr, *z_ = nullptr, *p_ = nullptr, *q_ = nullptr;
  double *fcomp_ = nullptr, *uacc_ = nullptr, *pacc_ = nullptr, *rc_ = nullptr;
  double* ru64_ = nullptr;   // refined FP64 solve: the FP64 velocity residual r_u
  float* wf_ = nullptr;
  static Fin fin(double* out) { return Fin{out}; }
  float* du_ = nullptr;       // mixed sweeps: velocity correction, accumulated in the CG
  bool accumulate_ = false;   // pcg / start_from add alpha * (L~^-1 D^T p) into du_
  bool first_checked_ = false;  // the next pcg's first check is queued already (sweep_start_cg)
  __bf16 *tb_ = nullptr, *tb1_ = nullptr, *tb2_ = nullptr;
  float *tf1_ = nullptr, *tf2_ = nullptr, *tfb_ = nullptr;
  int64_t tstride_ = 0, toff_v_ = 0;
  float* fx_ = nullptr;              // FFT transform scratch (fstride_ floats per column)
  int64_t fstride_ = 0, foff_v_ = 0;
  bool fft_ok_vel_ = false, fft_ok_p_ = false;  // every length of the blocks has an FFT plan
  VelLayout lay_b_{};
  // During a mixed sweep's inner CG: r rounded to FP32 (np per column), kept in tfb_
  // between the kernels that update r and the next pressure inverse (tfb_ takes D^T p
  // only after that inverse has read it).
  float* r32_ = nullptr;
  bool mixed_ = false;
  void* cnn_ = nullptr;
  double cnn_alpha_ = 1.0;
  int warm_mode_ = 0;
  int history_period_ = 0, hist_ncols_ = 0, hist_slots_ = 0;
  int64_t hist_count_ = 0, hist_committed_ = 0;
  static constexpr int kMaxHistSteps = kMaxExtrapolation;  // deepest extrapolation (degree 5)
  int hist_steps_ = 6;                         // steps the extrapolation uses
  int hist_ring_steps() const { return hist_steps_ + 1; }  // + one step for a trial in progress
  double shift_cur_ = 0.0;                             // viscosity / mass of the solve in progress
  double hist_shift_[(kMaxHistSteps + 1) * 16] = {};   // and of each recorded slot (period <= 16)
  double* hist_ = nullptr;      // ring of (steps+1)*period recorded pressures (np x ncols each)
  double* hist_u_ = nullptr;    // and velocities (size x ncols each; mixed solves)
  double* hist_b_ = nullptr;    // and zero-start Schur right-hand sides (np x ncols each)
  bool hist_b_ok_[(kMaxHistSteps + 1) * 16] = {};      // slot's right-hand side recorded
  double* bprev_ = nullptr;     // the previous solve's zero-start Schur right-hand side
  bool prev_b_ok_ = false;
  double* bzero_ = nullptr;     // mixed network solves: this solve's, kept from the sweeps
  double* warm_db_ = nullptr;   // b_base, then b - b_base
  bool hist_has_u_ = false;
  double* warm_base_ = nullptr;
  double* pprev_ = nullptr;
  int prev_ncols_ = 0;
  float *warm_x_ = nullptr, *warm_y_ = nullptr;
  double* warm_s_ = nullptr;
  double *partials_ = nullptr, *sums_ = nullptr, *sums2_ = nullptr, *sums3_ = nullptr, *rr_ = nullptr;
  size_t partials_cap_ = 0;  // doubles at partials_
  double* means_ = nullptr;  // mixed CG: column means of z and q (kMaxCols each)
  CgState* cg_ = nullptr;
  double* gal_ = nullptr;          // two-term Galerkin start: five dot products per column
  const double* net_base_ = nullptr;  // base guess the network corrected in this solve, if it ran
  int net_depth_limit_ = 3;        // the network serves solves with fewer recorded steps than this
  int* flags_ = nullptr;
  int* host_flags_ = nullptr;
  Mailbox* mailbox_ = nullptr;      // host view (coherent pinned memory)
  Mailbox* mailbox_dev_ = nullptr;  // device view of the same memory
  int seq_ = 0;

  // Spin until the mailbox holds sequence number `seq`.  The stream is polled now and
  // then, so a failed or finished stream ends the wait with an error instead of a hang.
  const volatile Mailbox* wait_post(int seq) {
    const volatile Mailbox* mb = mailbox_;
    for (long spin = 1;; ++spin) {
      if (__atomic_load_n(&mailbox_->seq, __ATOMIC_ACQUIRE) == seq) return mb;
      if ((spin & 255) == 0) {
        const hipError_t e = hipStreamQuery(stream_);
        if (e == hipSuccess) {
          if (__atomic_load_n(&mailbox_->seq, __ATOMIC_ACQUIRE) == seq) return mb;
          throw HipFailure("result slot was not written by the finished stream");
        }
        if (e != hipErrorNotReady) throw HipFailure(std::string("stream failed: ") + hipGetErrorString(e));
      }
    }
  }
};

}  // namespace macs

// ======================================================================= C ABI
using macs::Solver;

namespace {
thread_local std::string g_create_error;

int fail(Solver* s, int code, const std::string& msg) {
  if (s) s->error = msg;
  return code;
}

template <typename F>
int guarded(Solver* s, F&& body) {
  try {
    s->error.clear();
    body();
    return MAC_OK;
  } catch (const macs::ArgFailure& e) {
    return fail(s, MAC_INVALID_ARGUMENT, e.what());
  } catch (const macs::SolveFailure& e) {
    return fail(s, e.code, e.what());
  } catch (const std::exception& e) {
    return fail(s, MAC_HIP_ERROR, e.what());
  }
}

bool is_finite(double v) { return std::isfinite(v); }
}  // namespace

extern "C" {

void mac_stokes_default_opts(mac_stokes_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->tolerance = 1e-9;
  o->krylov_tolerance = 1e-12;
  o->max_iterations = 100;
  o->precision = MAC_PRECISION_FP64;
  o->refine_tolerance = 1e-10;
  o->inner_tolerance = 1e-3;
  o->max_sweeps = 30;
  o->mixed_split = 15;
  o->device = 0;
}

void* mac_stokes_create(int32_t nx, int32_t ny, double lx, double ly, const mac_stokes_opts* opts) {
  g_create_error.clear();
  mac_stokes_opts o;
  mac_stokes_default_opts(&o);
  if (opts) o = *opts;
  if (nx < 3 || ny < 3 || nx > 16384 || ny > 16384 || !is_finite(lx) || !is_finite(ly) || lx <= 0 || ly <= 0) {
    g_create_error = "Require 3 <= nx,ny <= 16384 and positive finite domain lengths";
    return nullptr;
  }
  if (!(o.tolerance > 0) || !(o.krylov_tolerance > 0) || !is_finite(o.tolerance) || !is_finite(o.krylov_tolerance) ||
      o.max_iterations < 1 || (o.precision != MAC_PRECISION_FP64 && o.precision != MAC_PRECISION_MIXED) ||
      !(o.refine_tolerance > 0) || !(o.inner_tolerance > 0) || o.max_sweeps < 1 || o.mixed_split < 0 ||
      o.mixed_split > 15 || o.fp64_refine < 0 || o.fp64_refine > 30) {
    g_create_error = "Invalid solver controls";
    return nullptr;
  }
  try {
    return new Solver(nx, ny, lx, ly, o);
  } catch (const std::exception& e) {
    g_create_error = e.what();
    return nullptr;
  }
}

const char* mac_stokes_create_error(void) { return g_create_error.c_str(); }

void mac_stokes_destroy(void* h) { delete static_cast<Solver*>(h); }

const char* mac_stokes_last_error(void* h) { return h ? static_cast<Solver*>(h)->error.c_str() : ""; }

int32_t mac_stokes_solve(void* h, const double* rhs, int64_t ld_rhs, int32_t ncols, double mass, double viscosity,
                         double* velocity, int64_t ld_velocity, double* pressure, int64_t ld_pressure,
                         const double* guess, int64_t ld_guess, int32_t* iterations, mac_solve_report* report,
                         int32_t device_pointers) {
  Solver* s = static_cast<Solver*>(h);
  if (!s) return MAC_INVALID_ARGUMENT;
  return guarded(s, [&] {
    if (!is_finite(mass) || !is_finite(viscosity) || std::min(mass, viscosity) < 0 || mass + viscosity <= 0)
      throw macs::ArgFailure("Require mass,viscosity >= 0 and nonzero total");
    const int64_t size = s->size(), np = s->np();
    if (ncols < 1 || ncols > Solver::kMaxCols || !rhs || !velocity || !pressure || ld_rhs < size ||
        ld_velocity < size || ld_pressure < np || (guess && ld_guess < np))
      throw macs::ArgFailure("Invalid Stokes RHS");
    HIP_CHECK(hipSetDevice(s->opts().device));
    const auto t0 = std::chrono::steady_clock::now();
    s->ensure_capacity(ncols);
    hipStream_t st = s->stream();
    const double* f = rhs;
    const double* gs = guess;
    double* vel = velocity;
    double* pr = pressure;
    int64_t lf = ld_rhs, lv = ld_velocity, lp = ld_pressure, lg = ld_guess;
    if (!device_pointers) {
      // Host arrays are pageable: blocking copies, with the solver's stream idle (an
      // asynchronous copy into or out of pageable memory followed by a stream
      // synchronisation has been seen to leave the host array unwritten).
      HIP_CHECK(hipStreamSynchronize(st));
      HIP_CHECK(hipMemcpy2D(s->io_f_, size * sizeof(double), rhs, ld_rhs * sizeof(double), size * sizeof(double),
                            ncols, hipMemcpyHostToDevice));
      f = s->io_f_;
      lf = size;
      if (guess) {
        HIP_CHECK(hipMemcpy2D(s->io_g_, np * sizeof(double), guess, ld_guess * sizeof(double), np * sizeof(double),
                              ncols, hipMemcpyHostToDevice));
        gs = s->io_g_;
        lg = np;
      }
      vel = s->io_v_;
      lv = size;
      pr = s->io_p_;
      lp = np;
    }
    // Non-finite RHS values are rejected like the reference (ValueError); the mixed solve
    // checks max|f| from its first defect instead of a pass of its own.
    if (s->opts().precision != MAC_PRECISION_MIXED) {
      macs::reduce(f, nullptr, size, lf, ncols, macs::RED_MAXABS, s->scratch_partials(), nullptr, st);
      macs::finalize(s->scratch_partials(), macs::reduce_blocks(size), ncols, macs::RED_MAXABS, s->scratch_sums(),
                     st);
      std::vector<double> fm(ncols);
      HIP_CHECK(hipMemcpyAsync(fm.data(), s->scratch_sums(), sizeof(double) * ncols, hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
      for (double v : fm)
        if (!is_finite(v)) throw macs::ArgFailure("Invalid Stokes RHS");
    }
    Solver::Outcome out = s->solve(f, lf, ncols, mass, viscosity, vel, lv, pr, lp, gs, lg);
    // On the caller's stream, device outputs are stream-ordered: no wait for the hand-out.
    if (!device_pointers || !s->external_stream()) HIP_CHECK(hipStreamSynchronize(st));
    if (!device_pointers) {
      HIP_CHECK(hipMemcpy2D(velocity, ld_velocity * sizeof(double), vel, size * sizeof(double), size * sizeof(double),
                            ncols, hipMemcpyDeviceToHost));
      HIP_CHECK(hipMemcpy2D(pressure, ld_pressure * sizeof(double), pr, np * sizeof(double), np * sizeof(double),
                            ncols, hipMemcpyDeviceToHost));
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    int total = 0;
    for (int c = 0; c < ncols; ++c) {
      total += out.iterations[c];
      if (iterations) iterations[c] = out.iterations[c];
    }
    if (report) {
      report->residual = out.residual;
      report->divergence_inf = out.divergence;
      report->velocity_max = out.vmax;
      report->seconds = secs;
      report->sweeps = out.sweeps;
      report->total_iterations = total;
      report->warm_started = out.warm ? (out.network ? 2 : 1) : 0;
      report->first_residual = out.first_residual;
      report->first_divergence = out.first_divergence;
    }
    const double tol = s->opts().tolerance;
    if (!is_finite(out.residual) || !is_finite(out.divergence) || !is_finite(out.vmax) || out.residual > tol ||
        out.divergence > tol * (1.0 + out.vmax)) {
      char buf[128];
      std::snprintf(buf, sizeof buf, "Stokes residual=%.3e, divergence=%.3e", out.residual, out.divergence);
      throw macs::SolveFailure(MAC_GATE_FAILED, buf);
    }
  });
}

int32_t mac_stokes_set_stream(void* h, void* stream, int32_t use) {
  Solver* s = static_cast<Solver*>(h);
  if (!s || (use != 0 && use != 1)) return MAC_INVALID_ARGUMENT;
  return guarded(s, [&] {
    HIP_CHECK(hipSetDevice(s->opts().device));
    s->set_stream(static_cast<hipStream_t>(stream), use == 1);
  });
}

void mac_stokes_set_profiling(void* h, int32_t on) {
  Solver* s = static_cast<Solver*>(h);
  if (!s) return;
  (void)guarded(s, [&] { s->set_profiling(on != 0); });
}

int32_t mac_stokes_get_profile(void* h, mac_kernel_profile* out) {
  Solver* s = static_cast<Solver*>(h);
  if (!s || !out) return MAC_INVALID_ARGUMENT;
  return guarded(s, [&] {
    s->profile(&out->gemm_ms, &out->gemm_flops, &out->gemm_bytes, &out->gemm_launches, &out->gemm_transform_flops);
    s->profile_fft(&out->fft_ms, &out->fft_bytes, &out->fft_launches, &out->fft_transform_flops);
  });
}

int32_t mac_stokes_set_warm_history(void* h, int32_t period) {
  Solver* s = static_cast<Solver*>(h);
  if (!s || period < 0 || period > 16) return MAC_INVALID_ARGUMENT;
  return guarded(s, [&] { s->set_history(period); });
}

int32_t mac_stokes_set_history_steps(void* h, int32_t steps) {
  Solver* s = static_cast<Solver*>(h);
  if (!s || steps < 1 || steps > macs::kMaxExtrapolation) return MAC_INVALID_ARGUMENT;
  return guarded(s, [&] { s->set_history_steps(steps); });
}

int32_t mac_stokes_history_mark(void* h, int32_t rollback) {
  Solver* s = static_cast<Solver*>(h);
  if (!s) return MAC_INVALID_ARGUMENT;
  return guarded(s, [&] {
    if (rollback) s->rollback_history();
    else s->commit_history();
  });
}

int32_t mac_stokes_set_network_depth(void* h, int32_t steps) {
  Solver* s = static_cast<Solver*>(h);
  if (!s || steps < 1 || steps > macs::kMaxExtrapolation + 1) return MAC_INVALID_ARGUMENT;
  return guarded(s, [&] { s->set_net_depth_limit(steps); });
}

int32_t mac_stokes_set_warm_start(void* h, int32_t mode, void* cnn, double alpha) {
  Solver* s = static_cast<Solver*>(h);
  if (!s || mode < MAC_WARM_OFF || mode > MAC_WARM_HISTORY || (mode == MAC_WARM_NETWORK && !cnn) ||
      !(alpha > 0) || !is_finite(alpha))
    return MAC_INVALID_ARGUMENT;
  return guarded(s, [&] { s->set_warm_start(mode, cnn, alpha); });
}

int32_t mac_stokes_reset_history(void* h) {
  Solver* s = static_cast<Solver*>(h);
  if (!s) return MAC_INVALID_ARGUMENT;
  return guarded(s, [&] { s->reset_history(); });
}

int32_t mac_stokes_last_schur_rhs(void* h, double* out, int64_t ld, int32_t ncols) {
  Solver* s = static_cast<Solver*>(h);
  if (!s || !out || ld < s->np() || ncols < 1 || ncols > s->last_ncols()) return MAC_INVALID_ARGUMENT;
  return guarded(s, [&] {
    HIP_CHECK(hipSetDevice(s->opts().device));
    macs::copy_cols(s->schur_rhs(), s->np(), out, ld, s->np(), ncols, s->stream());
    HIP_CHECK(hipStreamSynchronize(s->stream()));
  });
}

int32_t mac_warm_prepare(int32_t nx, int32_t ny, const double* b, int64_t ldb, const double* pprev, int64_t ldp,
                         const double* target, int64_t ldt, int32_t ncols, double alpha, float* x, float* t,
                         double* scale) {
  if (!macs::warm_shape_ok(nx, ny) || !b || !x || !scale || ncols < 1 || !(alpha > 0) || (target && !t))
    return MAC_INVALID_ARGUMENT;
  macs::warm_prepare(b, ldb, pprev, ldp, target, ldt, ncols, nx, ny, alpha, x, t, scale, nullptr);
  return hipDeviceSynchronize() == hipSuccess ? MAC_OK : MAC_HIP_ERROR;
}

int32_t mac_apply_divergence(void* h, const double* vel, int64_t ldv, int32_t ncols, double* out, int64_t ldo) {
  Solver* s = static_cast<Solver*>(h);
  if (!s) return MAC_INVALID_ARGUMENT;
  return guarded(s, [&] {
    if (ncols < 1 || !vel || !out || ldv < s->size() || ldo < s->np()) throw macs::ArgFailure("Invalid arguments");
    HIP_CHECK(hipSetDevice(s->opts().device));
    macs::VelLayout L = s->packed();
    L.stride = ldv;
    macs::divergence(macs::DT_F64, vel, L, s->grid(), out, ldo, 1.0, ncols, macs::RED_NONE, nullptr, nullptr,
                     s->stream());
    HIP_CHECK(hipStreamSynchronize(s->stream()));
  });
}

int32_t mac_apply_divergence_transpose(void* h, const double* p, int64_t ldp, int32_t ncols, double* out,
                                       int64_t ldo) {
  Solver* s = static_cast<Solver*>(h);
  if (!s) return MAC_INVALID_ARGUMENT;
  return guarded(s, [&] {
    if (ncols < 1 || !p || !out || ldp < s->np() || ldo < s->size()) throw macs::ArgFailure("Invalid arguments");
    HIP_CHECK(hipSetDevice(s->opts().device));
    macs::VelLayout L = s->packed();
    L.stride = ldo;
    macs::divergence_transpose(p, ldp, nullptr, nullptr, 0, macs::DT_F64, out, L, s->grid(), ncols, nullptr,
                               s->stream());
    HIP_CHECK(hipStreamSynchronize(s->stream()));
  });
}

int32_t mac_velocity_inverse(void* h, const double* in, int64_t ldi, int32_t ncols, double mass, double visc,
                             double* out, int64_t ldo) {
  Solver* s = static_cast<Solver*>(h);
  if (!s) return MAC_INVALID_ARGUMENT;
  return guarded(s, [&] {
    if (ncols < 1 || ncols > Solver::kMaxCols || !in || !out || ldi < s->size() || ldo < s->size() ||
        !is_finite(mass) || !is_finite(visc))
      throw macs::ArgFailure("Invalid arguments");
    HIP_CHECK(hipSetDevice(s->opts().device));
    s->ensure_capacity(ncols);
    std::vector<int> cols(ncols);
    for (int c = 0; c < ncols; ++c) cols[c] = c;
    if (s->opts().precision == MAC_PRECISION_MIXED) {
      // bf16 transforms, FP32 accumulation, added into a zeroed FP64 output.
      HIP_CHECK(hipMemset2DAsync(out, ldo * sizeof(double), 0, s->size() * sizeof(double), ncols, s->stream()));
      macs::VelLayout L = s->packed();
      L.stride = ldi;
      s->velocity_inverse_bf16(in, macs::DT_F64, L, out, macs::DT_F64, ldo, cols, mass, visc);
    } else {
      s->velocity_inverse(in, ldi, out, ldo, cols, mass, visc);
    }
    HIP_CHECK(hipStreamSynchronize(s->stream()));
  });
}

int32_t mac_pressure_poisson_inverse(void* h, const double* in, int64_t ldi, int32_t ncols, double* out,
                                     int64_t ldo) {
  Solver* s = static_cast<Solver*>(h);
  if (!s) return MAC_INVALID_ARGUMENT;
  return guarded(s, [&] {
    if (ncols < 1 || ncols > Solver::kMaxCols || !in || !out || ldi < s->np() || ldo < s->np())
      throw macs::ArgFailure("Invalid arguments");
    HIP_CHECK(hipSetDevice(s->opts().device));
    s->ensure_capacity(ncols);
    std::vector<int> cols(ncols);
    for (int c = 0; c < ncols; ++c) cols[c] = c;
    const int64_t np = s->np();
    if (s->opts().precision == MAC_PRECISION_MIXED) {
      double* tin = s->io_g_;
      double* tout = s->io_p_;
      macs::copy_cols(in, ldi, tin, np, np, ncols, s->stream());
      s->pressure_inverse_bf16(tin, tout, cols, 1.0, 0.0);
      macs::copy_cols(tout, np, out, ldo, np, ncols, s->stream());
    } else {
      s->pressure_inverse(in, ldi, out, ldo, cols, false, 1.0, 0.0);
    }
    macs::reduce(out, nullptr, np, ldo, ncols, macs::RED_SUM, s->scratch_partials(), nullptr, s->stream());
    macs::finalize(s->scratch_partials(), macs::reduce_blocks(np), ncols, macs::RED_SUM, s->scratch_sums(),
                   s->stream());
    macs::sub_mean_dot(out, np, ldo, ncols, s->scratch_sums(), nullptr, nullptr, nullptr, s->stream());
    HIP_CHECK(hipStreamSynchronize(s->stream()));
  });
}

}  // extern "C"
