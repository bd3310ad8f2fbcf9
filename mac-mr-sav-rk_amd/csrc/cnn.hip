// Pressure warm-start network: forward, backward and Adam on gfx950, plus the
// resampling glue that turns a Schur right-hand side and the previous pressure into a
// CG starting guess.
//
// Network (width w, 4 levels, 3x3 convs, zero padding, ReLU after all convs but the
// last; pool = 2x2 average, up = 2x nearest):
//   e0 = relu(c0(x)); e1 = relu(c1(pool e0)); e2 = relu(c2(pool e1)); e3 = relu(c3(pool e2))
//   g2 = relu(c4(up e3 + e2)); g1 = relu(c5(up g2 + e1)); y = c6(up g1 + e0)
// Weights live in FP32 (master, Adam state) and in a bf16 copy that every forward and
// backward-data convolution reads; accumulation is FP32.  Activations are NCHW FP32.
// Every reduction (weight gradients, loss) is two-level with a fixed block count, so a
// training step is bitwise reproducible.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "internal.h"
#include "macstokes.h"

#define HIP_CHECK_CNN(expr)                                                                          \
  do {                                                                                               \
    hipError_t e_ = (expr);                                                                          \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace macs {
namespace {

constexpr int TS = 16;  // spatial tile edge of the conv kernels

// y[n][co][p] = act(b[co] + sum_{ci,tap} w[co][ci][tap] x[n][ci][p + tap - 1]) (+ add)
// One thread per output pixel, CT output channels held in registers, input tiles and
// the weight slice staged in LDS in chunks of 8 input channels.
template <int CT, bool RELU>
__global__ __launch_bounds__(256) void k_conv_fwd(const float* __restrict__ x, int ci, const __bf16* __restrict__ w,
                                                  const float* __restrict__ bias, float* __restrict__ y, int co, int H,
                                                  int W, int co_groups) {
  constexpr int CH = 8;
  __shared__ float xs[CH][TS + 2][TS + 2];
  __shared__ float ws[CT][CH][9];
  const int n = blockIdx.z / co_groups, cg = blockIdx.z % co_groups;
  const int co0 = cg * CT;
  const int tx = threadIdx.x % TS, ty = threadIdx.x / TS;
  const int ox = blockIdx.x * TS + tx, oy = blockIdx.y * TS + ty;
  float acc[CT];
#pragma unroll
  for (int c = 0; c < CT; ++c) acc[c] = 0.f;
  const float* xn = x + (int64_t)n * ci * H * W;
  for (int c0 = 0; c0 < ci; c0 += CH) {
    // stage input tile with halo
    for (int idx = threadIdx.x; idx < CH * (TS + 2) * (TS + 2); idx += 256) {
      const int c = idx / ((TS + 2) * (TS + 2)), rem = idx % ((TS + 2) * (TS + 2));
      const int yy = blockIdx.y * TS + rem / (TS + 2) - 1, xx = blockIdx.x * TS + rem % (TS + 2) - 1;
      float v = 0.f;
      if (c0 + c < ci && yy >= 0 && yy < H && xx >= 0 && xx < W) v = xn[((int64_t)(c0 + c) * H + yy) * W + xx];
      xs[c][rem / (TS + 2)][rem % (TS + 2)] = v;
    }
    for (int idx = threadIdx.x; idx < CT * CH * 9; idx += 256) {
      const int o = idx / (CH * 9), rem = idx % (CH * 9), c = rem / 9, tap = rem % 9;
      float v = 0.f;
      if (co0 + o < co && c0 + c < ci) v = (float)w[((int64_t)(co0 + o) * ci + c0 + c) * 9 + tap];
      ws[o][c][tap] = v;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      float nb[9];
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) nb[tap] = xs[c][ty + tap / 3][tx + tap % 3];
#pragma unroll
      for (int o = 0; o < CT; ++o)
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) acc[o] = fmaf(ws[o][c][tap], nb[tap], acc[o]);
    }
    __syncthreads();
  }
  if (ox < W && oy < H) {
#pragma unroll
    for (int o = 0; o < CT; ++o) {
      if (co0 + o >= co) break;
      float v = acc[o] + (bias ? bias[co0 + o] : 0.f);
      if (RELU) v = fmaxf(v, 0.f);
      y[(((int64_t)n * co + co0 + o) * H + oy) * W + ox] = v;
    }
  }
}

// Per spatial tile, sample and chunk of OC output channels: partial
// dW[co][ci][tap] = sum_p dy[co][p] x[ci][p+tap-1] and db[co] = sum_p dy[co][p], written
// to part[tile][...] (each chunk fills its own entries; summed later in fixed order).
constexpr int OC = 16;
__global__ __launch_bounds__(256) void k_conv_wgrad(const float* __restrict__ x, int ci, const float* __restrict__ dy,
                                                    int co, int H, int W, int chunks, float* __restrict__ part) {
  extern __shared__ float sm[];
  float* ds = sm;                              // [OC][TS*TS]
  float* xs = sm + OC * TS * TS;               // [ci][TS+2][TS+2]
  const int n = blockIdx.z / chunks, oc0 = (blockIdx.z % chunks) * OC;
  const int noc = min(OC, co - oc0);
  const int blk = (n * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  const int nw = co * ci * 9 + co;
  for (int idx = threadIdx.x; idx < noc * TS * TS; idx += 256) {
    const int o = idx / (TS * TS), rem = idx % (TS * TS);
    const int yy = blockIdx.y * TS + rem / TS, xx = blockIdx.x * TS + rem % TS;
    ds[idx] = (yy < H && xx < W) ? dy[(((int64_t)n * co + oc0 + o) * H + yy) * W + xx] : 0.f;
  }
  for (int idx = threadIdx.x; idx < ci * (TS + 2) * (TS + 2); idx += 256) {
    const int c = idx / ((TS + 2) * (TS + 2)), rem = idx % ((TS + 2) * (TS + 2));
    const int yy = blockIdx.y * TS + rem / (TS + 2) - 1, xx = blockIdx.x * TS + rem % (TS + 2) - 1;
    xs[idx] = (yy >= 0 && yy < H && xx >= 0 && xx < W) ? x[(((int64_t)n * ci + c) * H + yy) * W + xx] : 0.f;
  }
  __syncthreads();
  const int local = noc * ci * 9 + noc;
  for (int li = threadIdx.x; li < local; li += 256) {
    float s = 0.f;
    int widx;
    if (li < noc * ci * 9) {
      const int o = li / (ci * 9), rem = li % (ci * 9), c = rem / 9, tap = rem % 9;
      const int dyy = tap / 3, dxx = tap % 3;
      const float* d = ds + o * TS * TS;
      const float* xc = xs + c * (TS + 2) * (TS + 2);
      for (int p = 0; p < TS * TS; ++p) s = fmaf(d[p], xc[(p / TS + dyy) * (TS + 2) + p % TS + dxx], s);
      widx = (oc0 + o) * ci * 9 + rem;
    } else {
      const int o = li - noc * ci * 9;
      const float* d = ds + o * TS * TS;
      for (int p = 0; p < TS * TS; ++p) s += d[p];
      widx = co * ci * 9 + oc0 + o;
    }
    part[(int64_t)blk * nw + widx] = s;
  }
}

// g[k] = sum over blocks b of part[b][k] (fixed order).
__global__ void k_sum_parts(const float* __restrict__ part, int nblk, int nw, float* __restrict__ g) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nw) return;
  float s = 0.f;
  for (int b = 0; b < nblk; ++b) s += part[(int64_t)b * nw + k];
  g[k] = s;
}

// wt[ci][co][8 - tap] = w[co][ci][tap]: backward-data as a forward conv.
__global__ void k_flip(const __bf16* __restrict__ w, int co, int ci, __bf16* __restrict__ wt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= co * ci * 9) return;
  const int o = i / (ci * 9), rem = i % (ci * 9), c = rem / 9, tap = rem % 9;
  wt[((int64_t)c * co + o) * 9 + (8 - tap)] = w[i];
}

__global__ void k_pool(const float* __restrict__ x, float* __restrict__ y, int64_t planes, int H, int W) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int h = H / 2, w = W / 2;
  if (i >= planes * h * w) return;
  const int64_t pl = i / ((int64_t)h * w);
  const int r = (int)(i % ((int64_t)h * w)), yy = r / w, xx = r % w;
  const float* s = x + pl * H * W + (int64_t)(2 * yy) * W + 2 * xx;
  y[i] = 0.25f * (s[0] + s[W] + s[1] + s[W + 1]);
}

// y = up(a) + b   (a at half resolution)
__global__ void k_up_add(const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ y,
                         int64_t planes, int H, int W) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= planes * H * W) return;
  const int64_t pl = i / ((int64_t)H * W);
  const int r = (int)(i % ((int64_t)H * W)), yy = r / W, xx = r % W;
  y[i] = a[pl * (H / 2) * (W / 2) + (int64_t)(yy / 2) * (W / 2) + xx / 2] + b[i];
}

// out = sum of the 2x2 block of d (adjoint of nearest up), times relu mask of act.
__global__ void k_up_bwd_mask(const float* __restrict__ d, const float* __restrict__ act, float* __restrict__ out,
                              int64_t planes, int h, int w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= planes * h * w) return;
  const int64_t pl = i / ((int64_t)h * w);
  const int r = (int)(i % ((int64_t)h * w)), yy = r / w, xx = r % w;
  const int W = 2 * w;
  const float* s = d + pl * 4 * h * w + (int64_t)(2 * yy) * W + 2 * xx;
  const float v = s[0] + s[W] + s[1] + s[W + 1];
  out[i] = act[i] > 0.f ? v : 0.f;
}

// out = (a + 0.25 up(b)) * (act > 0)   (a, act, out full resolution; b half)
__global__ void k_pool_bwd_add_mask(const float* __restrict__ a, const float* __restrict__ b,
                                    const float* __restrict__ act, float* __restrict__ out, int64_t planes, int H,
                                    int W) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= planes * H * W) return;
  const int64_t pl = i / ((int64_t)H * W);
  const int r = (int)(i % ((int64_t)H * W)), yy = r / W, xx = r % W;
  const float v = a[i] + 0.25f * b[pl * (H / 2) * (W / 2) + (int64_t)(yy / 2) * (W / 2) + xx / 2];
  out[i] = act[i] > 0.f ? v : 0.f;
}

// dy = 2 (y - t) / count; per-block partial sums of (y - t)^2.
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ y, const float* __restrict__ t,
                                             float* __restrict__ dy, int64_t n, float* __restrict__ part) {
  __shared__ float sh[256];
  float s = 0.f;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const float d = y[i] - t[i];
    dy[i] = 2.f * d / (float)n;
    s += d * d;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_adam(float* __restrict__ w, __bf16* __restrict__ wb, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  wb[i] = (__bf16)wi;
}

__global__ void k_to_bf16(const float* __restrict__ w, __bf16* __restrict__ wb, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) wb[i] = (__bf16)w[i];
}


// ------------------------------------------------------ warm-start resampling
// Network inputs for column c (grid nx x ny, network grid w = nx/4, h = ny/4):
//   x[c][0] = pool4(b_c) / s_c,  x[c][1] = pool4(pprev_c) / (alpha s_c),  s_c = rms(pool4(b_c)),
// target t[c] = pool4(p_c - pprev_c) / (alpha s_c) (the caller passes p - pprev): the network
// predicts the change from the previous pressure, and the guess is
// pprev_c + alpha s_c * bilinear_up4(y_c) (pprev_c = 0 when there is none).
__global__ void k_pool4(const double* __restrict__ src, int64_t stride, int nx, int ny, float* __restrict__ dst,
                        int64_t dst_stride) {
  const int w = nx / 4, h = ny / 4;
  const int c = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w * h) return;
  const int J = i / w, I = i % w;
  const double* s = src + (int64_t)c * stride + (int64_t)(4 * J) * nx + 4 * I;
  double acc = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc += s[(int64_t)a * nx + b];
  dst[(int64_t)c * dst_stride + i] = (float)(acc / 16.0);
}

__global__ void k_zero_f(float* __restrict__ p, int64_t stride, int64_t n) {
  const int c = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[(int64_t)c * stride + i] = 0.f;
}

// s[c] = rms of the plane at x + c*stride (one block per column, fixed order; 1 for a
// zero plane).
__global__ __launch_bounds__(256) void k_rms(const float* __restrict__ x, int64_t stride, int64_t n,
                                             double* __restrict__ s) {
  __shared__ double sh[256];
  const int c = blockIdx.x;
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) {
    const double v = x[(int64_t)c * stride + i];
    acc += v * v;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double r = sqrt(sh[0] / (double)n);
    s[c] = r > 0.0 ? r : 1.0;
  }
}

// Scale the network planes of column c: plane 0 (b) by 1/s, plane 1 (the base guess) by
// its own rms 1/s2, the target by 1/(alpha s).
__global__ void k_normalize(float* __restrict__ x, float* __restrict__ t, int64_t n, const double* __restrict__ s,
                            const double* __restrict__ s2, double alpha) {
  const int c = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float* xc = x + (int64_t)c * 2 * n;
  xc[i] = (float)(xc[i] / s[c]);
  xc[n + i] = (float)(xc[n + i] / s2[c]);
  if (t) t[(int64_t)c * n + i] = (float)(t[(int64_t)c * n + i] / (alpha * s[c]));
}

// guess[c] (nx x ny) = alpha s_c * bilinear interpolation of y[c] (nx/4 x ny/4), cell centres.
__global__ void k_up4(const float* __restrict__ y, int nx, int ny, const double* __restrict__ s, double alpha,
                      const double* __restrict__ base, int64_t base_stride, double* __restrict__ out,
                      int64_t out_stride) {
  const int w = nx / 4, h = ny / 4;
  const int c = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)nx * ny) return;
  const int j = (int)(i / nx), k = (int)(i % nx);
  const double cy = (j + 0.5) / 4.0 - 0.5, cx = (k + 0.5) / 4.0 - 0.5;
  int j0 = (int)floor(cy), k0 = (int)floor(cx);
  double fy = cy - j0, fx = cx - k0;
  if (j0 < 0) { j0 = 0; fy = 0.0; }
  if (k0 < 0) { k0 = 0; fx = 0.0; }
  if (j0 > h - 1) { j0 = h - 1; fy = 0.0; }
  if (k0 > w - 1) { k0 = w - 1; fx = 0.0; }
  const int j1 = min(j0 + 1, h - 1), k1 = min(k0 + 1, w - 1);
  const float* yc = y + (int64_t)c * w * h;
  const double v = (1 - fy) * ((1 - fx) * yc[j0 * w + k0] + fx * yc[j0 * w + k1]) +
                   fy * ((1 - fx) * yc[j1 * w + k0] + fx * yc[j1 * w + k1]);
  const double b0 = base ? base[(int64_t)c * base_stride + i] : 0.0;
  out[(int64_t)c * out_stride + i] = b0 + alpha * s[c] * v;
}

inline int nblocks(int64_t n) { return (int)((n + 255) / 256); }

}  // namespace

// ============================================================== buffer plan
// Element counts of the activation buffers of one forward + backward pass of n inputs
// of H x W, in the order of the Acts members (e0 ... de0, part, lossp, wt), from the
// width and the shapes alone (no device calls), so the plan can be checked on a host.
// Every count is the largest extent any kernel touches in that buffer.
enum { kPlanBuffers = 29 };
bool cnn_plan(int width, int n, int H, int W, int64_t* count) {
  if (H % 8 || W % 8 || H < 8 || W < 8 || n < 1 || width < 1) return false;
  const int64_t w = width, f = (int64_t)n * w * H * W, h2 = f / 4, h4 = f / 16, h8 = f / 64;
  const int64_t one = (int64_t)n * H * W;
  // e0 p1 e1 p2 e2 p3 e3 u2 g2 u1 g1 u0 y dy du0 dg1 du1 dg2 du2 de3 dp3 de2 dp2 de1 dp1 de0
  const int64_t acts[26] = {f, h2, h2, h4, h4, h8, h8, h4, h4, h2, h2, f, one, one,
                            f, h2, h2, h4, h4, h8, h8, h4, h4, h2, h2, f};
  for (int k = 0; k < 26; ++k) count[k] = acts[k];
  // The weight-gradient partials: one row of co*ci*9 + co per (spatial block, input) of
  // the widest layer at full resolution, the largest of the seven layers' needs.
  const int64_t nblk = (int64_t)((W + TS - 1) / TS) * ((H + TS - 1) / TS) * n;
  count[26] = nblk * (w * w * 9 + w);
  count[27] = 256;                       // loss partials (at most 256 blocks)
  count[28] = (w * w * 9 + 1) / 2;       // flipped bf16 weights, in float units
  return true;
}

// ============================================================== network object
class Cnn {
 public:
  Cnn(int width, int device) : width_(width), device_(device) {
    HIP_CHECK_CNN(hipSetDevice(device));
    HIP_CHECK_CNN(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    const int shapes[7][2] = {{width, 2}, {width, width}, {width, width}, {width, width},
                              {width, width}, {width, width}, {1, width}};
    int off = 0;
    for (int l = 0; l < 7; ++l) {
      co_[l] = shapes[l][0];
      ci_[l] = shapes[l][1];
      woff_[l] = off;
      off += co_[l] * ci_[l] * 9;
      boff_[l] = off;
      off += co_[l];
    }
    nparams_ = off;
    w_ = alloc<float>(nparams_);
    wb_ = alloc<__bf16>(nparams_);
    g_ = alloc<float>(nparams_);
    m_ = alloc<float>(nparams_);
    v_ = alloc<float>(nparams_);
    HIP_CHECK_CNN(hipMemset(w_, 0, sizeof(float) * nparams_));
    HIP_CHECK_CNN(hipMemset(m_, 0, sizeof(float) * nparams_));
    HIP_CHECK_CNN(hipMemset(v_, 0, sizeof(float) * nparams_));
    sync_bf16();
  }
  ~Cnn() {
    (void)hipSetDevice(device_);
    (void)hipStreamSynchronize(stream_);
    for (void* p : allocs_) (void)hipFree(p);
    (void)hipStreamDestroy(stream_);
  }

  int nparams() const { return nparams_; }
  hipStream_t stream() const { return stream_; }
  std::string error;

  void set_params(const float* host) {
    HIP_CHECK_CNN(hipMemcpy(w_, host, sizeof(float) * nparams_, hipMemcpyHostToDevice));
    HIP_CHECK_CNN(hipMemset(m_, 0, sizeof(float) * nparams_));
    HIP_CHECK_CNN(hipMemset(v_, 0, sizeof(float) * nparams_));
    step_ = 0;
    sync_bf16();
  }
  void get_params(float* host) {
    HIP_CHECK_CNN(hipMemcpy(host, w_, sizeof(float) * nparams_, hipMemcpyDeviceToHost));
  }

  // x: [n][2][H][W] device, y: [n][1][H][W] device.
  void forward(const float* x, float* y, int n, int H, int W) {
    ensure(n, H, W);
    forward_planned(x, y, n, H, W);
  }

  // Loss = mean (y - t)^2 over all n*H*W outputs; gradient into g_ (and host copy).
  float loss_grad(const float* x, const float* t, int n, int H, int W) {
    ensure(n, H, W);
    return loss_grad_planned(x, t, n, H, W);
  }

 private:
  // The passes proper.  They use the activation buffers as planned by the caller's
  // ensure() and never re-plan, so no buffer address they hold (a_.y passed to the
  // forward pass, for one) can go stale while they run.
  void forward_planned(const float* x, float* y, int n, int H, int W) {
    const int w = width_;
    conv(x, 0, a_.e0, n, H, W, true);
    pool(a_.e0, a_.p1, (int64_t)n * w, H, W);
    conv(a_.p1, 1, a_.e1, n, H / 2, W / 2, true);
    pool(a_.e1, a_.p2, (int64_t)n * w, H / 2, W / 2);
    conv(a_.p2, 2, a_.e2, n, H / 4, W / 4, true);
    pool(a_.e2, a_.p3, (int64_t)n * w, H / 4, W / 4);
    conv(a_.p3, 3, a_.e3, n, H / 8, W / 8, true);
    up_add(a_.e3, a_.e2, a_.u2, (int64_t)n * w, H / 4, W / 4);
    conv(a_.u2, 4, a_.g2, n, H / 4, W / 4, true);
    up_add(a_.g2, a_.e1, a_.u1, (int64_t)n * w, H / 2, W / 2);
    conv(a_.u1, 5, a_.g1, n, H / 2, W / 2, true);
    up_add(a_.g1, a_.e0, a_.u0, (int64_t)n * w, H, W);
    conv(a_.u0, 6, y, n, H, W, false);
  }

  float loss_grad_planned(const float* x, const float* t, int n, int H, int W) {
    forward_planned(x, a_.y, n, H, W);
    const int w = width_;
    const int64_t cnt = (int64_t)n * H * W;
    const int mb = std::min<int64_t>(nblocks(cnt), 256);
    k_mse<<<mb, 256, 0, stream_>>>(a_.y, t, a_.dy, cnt, a_.lossp);
    // conv6
    wgrad(a_.u0, 6, a_.dy, n, H, W);
    dgrad(a_.dy, 6, a_.du0, n, H, W);                       // du0 (w, H, W)
    upbwd_mask(a_.du0, a_.g1, a_.dg1, (int64_t)n * w, H / 2, W / 2);
    wgrad(a_.u1, 5, a_.dg1, n, H / 2, W / 2);
    dgrad(a_.dg1, 5, a_.du1, n, H / 2, W / 2);
    upbwd_mask(a_.du1, a_.g2, a_.dg2, (int64_t)n * w, H / 4, W / 4);
    wgrad(a_.u2, 4, a_.dg2, n, H / 4, W / 4);
    dgrad(a_.dg2, 4, a_.du2, n, H / 4, W / 4);
    upbwd_mask(a_.du2, a_.e3, a_.de3, (int64_t)n * w, H / 8, W / 8);
    wgrad(a_.p3, 3, a_.de3, n, H / 8, W / 8);
    dgrad(a_.de3, 3, a_.dp3, n, H / 8, W / 8);
    poolbwd_add_mask(a_.du2, a_.dp3, a_.e2, a_.de2, (int64_t)n * w, H / 4, W / 4);
    wgrad(a_.p2, 2, a_.de2, n, H / 4, W / 4);
    dgrad(a_.de2, 2, a_.dp2, n, H / 4, W / 4);
    poolbwd_add_mask(a_.du1, a_.dp2, a_.e1, a_.de1, (int64_t)n * w, H / 2, W / 2);
    wgrad(a_.p1, 1, a_.de1, n, H / 2, W / 2);
    dgrad(a_.de1, 1, a_.dp1, n, H / 2, W / 2);
    poolbwd_add_mask(a_.du0, a_.dp1, a_.e0, a_.de0, (int64_t)n * w, H, W);
    wgrad(x, 0, a_.de0, n, H, W);
    std::vector<float> lp(mb);
    HIP_CHECK_CNN(hipMemcpyAsync(lp.data(), a_.lossp, sizeof(float) * mb, hipMemcpyDeviceToHost, stream_));
    HIP_CHECK_CNN(hipStreamSynchronize(stream_));
    double s = 0;
    for (float v : lp) s += v;
    return (float)(s / (double)cnt);
  }

 public:
  float train_step(const float* x, const float* t, int n, int H, int W, float lr) {
    const float loss = loss_grad(x, t, n, H, W);
    ++step_;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)step_), c2 = 1.f - std::pow(b2, (float)step_);
    k_adam<<<nblocks(nparams_), 256, 0, stream_>>>(w_, wb_, g_, m_, v_, nparams_, lr, b1, b2, 1e-8f, c1, c2);
    HIP_CHECK_CNN(hipStreamSynchronize(stream_));
    return loss;
  }

  const float* grad() const { return g_; }

 private:
  template <typename T>
  T* alloc(int64_t count) {
    void* p = nullptr;
    HIP_CHECK_CNN(hipMalloc(&p, std::max<int64_t>(count, 64) * sizeof(T)));
    allocs_.push_back(p);
    return static_cast<T*>(p);
  }

  void sync_bf16() {
    k_to_bf16<<<nblocks(nparams_), 256, 0, stream_>>>(w_, wb_, nparams_);
    HIP_CHECK_CNN(hipStreamSynchronize(stream_));
  }

  void conv(const float* x, int l, float* y, int n, int H, int W, bool relu) {
    const int co = co_[l], ci = ci_[l];
    const __bf16* w = wb_ + woff_[l];
    const float* b = w_ + boff_[l];
    dim3 grid((W + TS - 1) / TS, (H + TS - 1) / TS, 1);
    if (co == 1) {
      grid.z = n;
      if (relu) k_conv_fwd<1, true><<<grid, 256, 0, stream_>>>(x, ci, w, b, y, co, H, W, 1);
      else k_conv_fwd<1, false><<<grid, 256, 0, stream_>>>(x, ci, w, b, y, co, H, W, 1);
    } else {
      const int groups = (co + 15) / 16;
      grid.z = n * groups;
      if (relu) k_conv_fwd<16, true><<<grid, 256, 0, stream_>>>(x, ci, w, b, y, co, H, W, groups);
      else k_conv_fwd<16, false><<<grid, 256, 0, stream_>>>(x, ci, w, b, y, co, H, W, groups);
    }
    HIP_CHECK_CNN(hipGetLastError());
  }

  // dx = conv^T(dy): forward conv with flipped, transposed weights, no bias.
  void dgrad(const float* dy, int l, float* dx, int n, int H, int W) {
    const int co = co_[l], ci = ci_[l];
    k_flip<<<nblocks(co * ci * 9), 256, 0, stream_>>>(wb_ + woff_[l], co, ci, a_.wt);
    dim3 grid((W + TS - 1) / TS, (H + TS - 1) / TS, 1);
    const int groups = (ci + 15) / 16;
    grid.z = n * groups;
    k_conv_fwd<16, false><<<grid, 256, 0, stream_>>>(dy, co, a_.wt, nullptr, dx, ci, H, W, groups);
    HIP_CHECK_CNN(hipGetLastError());
  }

  void wgrad(const float* x, int l, const float* dy, int n, int H, int W) {
    const int co = co_[l], ci = ci_[l];
    const int nw = co * ci * 9 + co;
    const int chunks = (co + OC - 1) / OC;
    dim3 grid((W + TS - 1) / TS, (H + TS - 1) / TS, n * chunks);
    const int nblk = grid.x * grid.y * n;
    const size_t sm = sizeof(float) * (OC * TS * TS + ci * (TS + 2) * (TS + 2));
    k_conv_wgrad<<<grid, 256, sm, stream_>>>(x, ci, dy, co, H, W, chunks, a_.part);
    k_sum_parts<<<nblocks(nw), 256, 0, stream_>>>(a_.part, nblk, nw, g_ + woff_[l]);
    HIP_CHECK_CNN(hipGetLastError());
  }

  void pool(const float* x, float* y, int64_t planes, int H, int W) {
    k_pool<<<nblocks(planes * (H / 2) * (W / 2)), 256, 0, stream_>>>(x, y, planes, H, W);
  }
  void up_add(const float* a, const float* b, float* y, int64_t planes, int H, int W) {
    k_up_add<<<nblocks(planes * H * W), 256, 0, stream_>>>(a, b, y, planes, H, W);
  }
  void upbwd_mask(const float* d, const float* act, float* out, int64_t planes, int h, int w) {
    k_up_bwd_mask<<<nblocks(planes * h * w), 256, 0, stream_>>>(d, act, out, planes, h, w);
  }
  void poolbwd_add_mask(const float* a, const float* b, const float* act, float* out, int64_t planes, int H, int W) {
    k_pool_bwd_add_mask<<<nblocks(planes * H * W), 256, 0, stream_>>>(a, b, act, out, planes, H, W);
  }

  struct Acts {
    float *e0, *p1, *e1, *p2, *e2, *p3, *e3, *u2, *g2, *u1, *g1, *u0, *y, *dy;
    float *du0, *dg1, *du1, *dg2, *du2, *de3, *dp3, *de2, *dp2, *de1, *dp1, *de0, *part, *lossp;
    __bf16* wt;
  };

  // Sizes the activation buffers for (n, H, W) by the plan; called once at the top of
  // each entry point, before any buffer address is taken.  (Re-planning inside a pass
  // once freed buffers whose addresses the pass already held.)
  void ensure(int n, int H, int W) {
    int64_t plan[kPlanBuffers];
    if (!cnn_plan(width_, n, H, W, plan)) throw std::invalid_argument("network input H, W must be multiples of 8");
    if (n == cap_n_ && H == cap_h_ && W == cap_w_) return;
    HIP_CHECK_CNN(hipStreamSynchronize(stream_));
    for (void* p : act_allocs_) {
      (void)hipFree(p);
      allocs_.erase(std::remove(allocs_.begin(), allocs_.end(), p), allocs_.end());
    }
    act_allocs_.clear();
    int k = 0;
    auto take = [&]() {
      float* p = alloc<float>(plan[k++]);
      act_allocs_.push_back(p);
      return p;
    };
    float** order[28] = {&a_.e0, &a_.p1, &a_.e1, &a_.p2, &a_.e2, &a_.p3, &a_.e3, &a_.u2, &a_.g2, &a_.u1,
                         &a_.g1, &a_.u0, &a_.y, &a_.dy, &a_.du0, &a_.dg1, &a_.du1, &a_.dg2, &a_.du2, &a_.de3,
                         &a_.dp3, &a_.de2, &a_.dp2, &a_.de1, &a_.dp1, &a_.de0, &a_.part, &a_.lossp};
    for (float** slot : order) *slot = take();
    a_.wt = reinterpret_cast<__bf16*>(take());
    cap_n_ = n;
    cap_h_ = H;
    cap_w_ = W;
  }

  int width_, device_;
  hipStream_t stream_ = nullptr;
  int co_[7], ci_[7], woff_[7], boff_[7];
  int nparams_ = 0, step_ = 0;
  float *w_ = nullptr, *g_ = nullptr, *m_ = nullptr, *v_ = nullptr;
  __bf16* wb_ = nullptr;
  Acts a_{};
  int cap_n_ = 0, cap_h_ = 0, cap_w_ = 0;
  std::vector<void*> allocs_, act_allocs_;
};

// ---------------------------------------------------------- solver-side glue
bool warm_shape_ok(int nx, int ny) { return nx % 32 == 0 && ny % 32 == 0 && nx >= 32 && ny >= 32; }

// x: [ncols][2][ny/4][nx/4], t (nullable): [ncols][ny/4][nx/4], s: [ncols] scales.
void warm_prepare(const double* b, int64_t b_stride, const double* pprev, int64_t p_stride, const double* target,
                  int64_t t_stride, int ncols, int nx, int ny, double alpha, float* x, float* t, double* s,
                  hipStream_t st) {
  const int64_t n = (int64_t)(nx / 4) * (ny / 4);
  const dim3 grid((unsigned)((n + 255) / 256), (unsigned)ncols);
  k_pool4<<<grid, 256, 0, st>>>(b, b_stride, nx, ny, x, 2 * n);
  if (pprev) k_pool4<<<grid, 256, 0, st>>>(pprev, p_stride, nx, ny, x + n, 2 * n);
  else k_zero_f<<<grid, 256, 0, st>>>(x + n, 2 * n, n);
  if (target) k_pool4<<<grid, 256, 0, st>>>(target, t_stride, nx, ny, t, n);
  k_rms<<<ncols, 256, 0, st>>>(x, 2 * n, n, s);
  k_rms<<<ncols, 256, 0, st>>>(x + n, 2 * n, n, s + ncols);
  k_normalize<<<grid, 256, 0, st>>>(x, target ? t : nullptr, n, s, s + ncols, alpha);
}

void warm_guess(void* cnn, const double* b, int64_t b_stride, const double* pprev, int64_t p_stride, int ncols,
                int nx, int ny, double alpha, float* x, float* y, double* s, double* guess, int64_t g_stride,
                hipStream_t st) {
  Cnn* net = static_cast<Cnn*>(cnn);
  warm_prepare(b, b_stride, pprev, p_stride, nullptr, 0, ncols, nx, ny, alpha, x, nullptr, s, st);
  if (hipStreamSynchronize(st) != hipSuccess) throw std::runtime_error("warm-start input preparation failed");
  net->forward(x, y, ncols, ny / 4, nx / 4);
  if (hipStreamSynchronize(net->stream()) != hipSuccess) throw std::runtime_error("warm-start forward failed");
  const dim3 grid((unsigned)(((int64_t)nx * ny + 255) / 256), (unsigned)ncols);
  k_up4<<<grid, 256, 0, st>>>(y, nx, ny, s, alpha, pprev, p_stride, guess, g_stride);
}

}  // namespace macs

// ======================================================================= C ABI
using macs::Cnn;

namespace {
thread_local std::string g_cnn_error;

template <typename F>
int cnn_guard(Cnn* c, F&& body) {
  try {
    c->error.clear();
    body();
    return MAC_OK;
  } catch (const std::invalid_argument& e) {
    c->error = e.what();
    return MAC_INVALID_ARGUMENT;
  } catch (const std::exception& e) {
    c->error = e.what();
    return MAC_HIP_ERROR;
  }
}
}  // namespace

extern "C" {

void* mac_cnn_create(int32_t width, int32_t device) {
  g_cnn_error.clear();
  if (width < 1 || width > 64) {
    g_cnn_error = "width must be in [1, 64]";
    return nullptr;
  }
  try {
    return new Cnn(width, device);
  } catch (const std::exception& e) {
    g_cnn_error = e.what();
    return nullptr;
  }
}

const char* mac_cnn_create_error(void) { return g_cnn_error.c_str(); }
void mac_cnn_destroy(void* h) { delete static_cast<Cnn*>(h); }
const char* mac_cnn_last_error(void* h) { return h ? static_cast<Cnn*>(h)->error.c_str() : ""; }
int64_t mac_cnn_num_params(void* h) { return h ? static_cast<Cnn*>(h)->nparams() : 0; }

int32_t mac_cnn_set_params(void* h, const float* host_params) {
  Cnn* c = static_cast<Cnn*>(h);
  if (!c || !host_params) return MAC_INVALID_ARGUMENT;
  return cnn_guard(c, [&] { c->set_params(host_params); });
}

int32_t mac_cnn_get_params(void* h, float* host_params) {
  Cnn* c = static_cast<Cnn*>(h);
  if (!c || !host_params) return MAC_INVALID_ARGUMENT;
  return cnn_guard(c, [&] { c->get_params(host_params); });
}

int32_t mac_cnn_plan(int32_t width, int32_t n, int32_t H, int32_t W, int64_t* counts) {
  if (!counts) return MAC_INVALID_ARGUMENT;
  return macs::cnn_plan(width, n, H, W, counts) ? MAC_OK : MAC_INVALID_ARGUMENT;
}

int32_t mac_cnn_forward(void* h, const float* x, float* y, int32_t n, int32_t H, int32_t W) {
  Cnn* c = static_cast<Cnn*>(h);
  if (!c || !x || !y || n < 1) return MAC_INVALID_ARGUMENT;
  return cnn_guard(c, [&] {
    c->forward(x, y, n, H, W);
    if (hipStreamSynchronize(c->stream()) != hipSuccess) throw std::runtime_error("forward failed");
  });
}

int32_t mac_cnn_loss_grad(void* h, const float* x, const float* target, int32_t n, int32_t H, int32_t W,
                          float* grad_out, float* loss_out) {
  Cnn* c = static_cast<Cnn*>(h);
  if (!c || !x || !target || n < 1) return MAC_INVALID_ARGUMENT;
  return cnn_guard(c, [&] {
    const float loss = c->loss_grad(x, target, n, H, W);
    if (loss_out) *loss_out = loss;
    if (grad_out &&
        hipMemcpy(grad_out, c->grad(), sizeof(float) * c->nparams(), hipMemcpyDeviceToDevice) != hipSuccess)
      throw std::runtime_error("gradient copy failed");
  });
}

int32_t mac_cnn_train_step(void* h, const float* x, const float* target, int32_t n, int32_t H, int32_t W, float lr,
                           float* loss_out) {
  Cnn* c = static_cast<Cnn*>(h);
  if (!c || !x || !target || n < 1 || !(lr > 0)) return MAC_INVALID_ARGUMENT;
  return cnn_guard(c, [&] {
    const float loss = c->train_step(x, target, n, H, W, lr);
    if (loss_out) *loss_out = loss;
  });
}

}  // extern "C"
