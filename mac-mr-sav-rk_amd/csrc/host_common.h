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

}  // namespace macs
