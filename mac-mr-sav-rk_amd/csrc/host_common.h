// Host-side helpers shared by the solver and the batched solver: error types, the
// HIP status check, and the orthonormal transform matrices with their eigenvalues.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace macs {

#define HIP_CHECK(expr)                                                                              \
  do {                                                                                               \
    hipError_t e_ = (expr);                                                                          \
    if (e_ != hipSuccess)                                                                            \
      throw macs::HipFailure(std::string(#expr) + ": " + hipGetErrorString(e_));                           \
  } while (0)

struct HipFailure : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct ArgFailure : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct SolveFailure : std::runtime_error {
  int code;
  SolveFailure(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum Kind { KIND_DST1 = 0, KIND_DST2 = 1, KIND_DCT2 = 2 };

// Orthonormal analysis matrix (forward = Q x, inverse = Q^T y), as scipy.fft's
// norm="ortho" dst/dct (tensor_stokes.py:58-82).  The angle's integer numerator is
// reduced modulo the period first, so every entry is a correctly-argued sin/cos.
inline std::vector<double> transform_matrix(int n, int kind) {
  std::vector<double> q((size_t)n * n);
  for (int k = 0; k < n; ++k) {
    for (int j = 0; j < n; ++j) {
      double v;
      if (kind == KIND_DST1) {
        const int64_t period = 2 * (int64_t)(n + 1);
        const int64_t a = ((int64_t)(k + 1) * (j + 1)) % period;
        v = std::sqrt(2.0 / (n + 1)) * std::sin(M_PI * (double)a / (double)(n + 1));
      } else if (kind == KIND_DST2) {
        const int64_t period = 4 * (int64_t)n;
        const int64_t a = ((int64_t)(k + 1) * (2 * j + 1)) % period;
        v = std::sqrt(2.0 / n) * std::sin(M_PI * (double)a / (2.0 * n));
        if (k == n - 1) v /= std::sqrt(2.0);
      } else {
        const int64_t period = 4 * (int64_t)n;
        const int64_t a = ((int64_t)k * (2 * j + 1)) % period;
        v = std::sqrt(2.0 / n) * std::cos(M_PI * (double)a / (2.0 * n));
        if (k == 0) v /= std::sqrt(2.0);
      }
      q[(size_t)k * n + j] = v;
    }
  }
  return q;
}

// (2 - 2 cos theta_k)/h^2, theta = pi k/n (dct2, k=0..n-1), pi k/(n+1) (dst1, k=1..n),
// pi k/n (dst2, k=1..n): tensor_stokes.py:31-38.
inline std::vector<double> eigenvalues_1d(int n, double h, int kind) {
  std::vector<double> lam(n);
  for (int k = 0; k < n; ++k) {
    double theta;
    if (kind == KIND_DCT2) theta = M_PI * k / n;
    else if (kind == KIND_DST1) theta = M_PI * (k + 1) / (n + 1);
    else theta = M_PI * (k + 1) / n;
    lam[k] = (2.0 - 2.0 * std::cos(theta)) / (h * h);
  }
  return lam;
}

inline int round8(int n) { return (n + 7) & ~7; }

// Weights w[0..steps-1] of the last `steps` same-stage solves (newest first) in the
// Lagrange extrapolation to the new solve's time.  h[0] is the new solve's step size and
// h[k] that of the solve k steps back (each known up to a common factor: a stage's
// viscosity / mass is nu a_ii dt); consecutive same-stage solves are taken (h[k-1] +
// h[k]) / 2 apart.  Equal or unknown steps give (-1)^(k+1) C(steps, k).
inline void history_weights(const double* h, int steps, double* w) {
  bool equal = true, known = std::isfinite(h[0]) && h[0] > 0.0;
  for (int k = 1; k <= steps; ++k) {
    equal = equal && h[k] == h[0];
    known = known && std::isfinite(h[k]) && h[k] > 0.0;
  }
  if (equal || !known) {
    double c = 1.0;  // C(steps, k), built up k by k
    for (int k = 1; k <= steps; ++k) {
      c = c * (steps - k + 1) / k;
      w[k - 1] = (k % 2 == 1) ? c : -c;
    }
    return;
  }
  double x[16] = {0.0};  // node times relative to the new solve (steps < 16)
  for (int k = 1; k <= steps; ++k) x[k] = x[k - 1] - 0.5 * (h[k - 1] + h[k]);
  for (int k = 1; k <= steps; ++k) {
    double num = 1.0, den = 1.0;
    for (int m = 1; m <= steps; ++m) {
      if (m == k) continue;
      num *= -x[m];
      den *= x[k] - x[m];
    }
    w[k - 1] = num / den;
  }
}

}  // namespace macs
