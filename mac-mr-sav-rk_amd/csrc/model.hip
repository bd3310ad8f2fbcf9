// Device-resident MAC model operations the schemes call between stage solves: the
// skew-symmetric convection (solver/mac/kernels.py:6-35), linear combinations of fields
// (solver/mac_ns.py:40-45), the viscous operator K (solver/mac/operators.py:6-17, CSR row
// order), D and D^T, and the FP64 inner product / max-norm (solver/mac/grid.py:67-68).
//
// Convection and combination use FP64 operations in the order of the reference's
// expressions, and this file is built with -ffp-contract=off (no fused multiply-adds),
// so their results are bitwise those of the reference; K and G apply their rows in CSR
// column order.
// Kernels run on the null stream (ordered with torch's default stream); entry points
// that return a host scalar synchronise, the others return once enqueued.
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "host_common.h"
#include "internal.h"
#include "macstokes.h"

namespace macs {
namespace {

struct Dims {
  int nx, ny;
  double hx, hy;
  int nu, size;
};

Dims dims(int nx, int ny, double lx, double ly) {
  Dims d;
  d.nx = nx;
  d.ny = ny;
  d.hx = lx / nx;
  d.hy = ly / ny;
  d.nu = (nx - 1) * ny;
  d.size = d.nu + nx * (ny - 1);
  return d;
}

// Packed face accessors with the wall values of the full (ny, nx+1) / (ny+1, nx) arrays:
// normal velocity zero on the walls.
__device__ inline double U(const double* u, const Dims& d, int j, int i) {  // i in 0..nx
  return (i <= 0 || i >= d.nx) ? 0.0 : u[j * (d.nx - 1) + (i - 1)];
}
__device__ inline double V(const double* v, const Dims& d, int j, int i) {  // j in 0..ny
  return (j <= 0 || j >= d.ny) ? 0.0 : v[(j - 1) * d.nx + i];
}

__device__ inline double mul(double a, double b) { return __dmul_rn(a, b); }
__device__ inline double add(double a, double b) { return __dadd_rn(a, b); }
__device__ inline double sub(double a, double b) { return __dsub_rn(a, b); }
__device__ inline double dvd(double a, double b) { return __ddiv_rn(a, b); }

// N(u) on every interior face (kernels.py:6-35): centred dual-cell transport velocities
// e, w, n, s; flux difference; minus half the dual divergence times the face velocity;
// odd ghosts (-u, -v) across the tangential walls.
__global__ __launch_bounds__(256) void k_convection(const double* __restrict__ vel, int64_t ld, Dims d,
                                                    double* __restrict__ out, int64_t ld_out) {
  const int c = blockIdx.y;
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= d.size) return;
  const double* u = vel + (int64_t)c * ld;
  const double* v = u + d.nu;
  double r;
  if (f < d.nu) {
    const SmallDiv dw(d.nx - 1);
    const int j = dw.quot(f), i = f - j * (d.nx - 1) + 1;  // full index of the u face
    const double uc = U(u, d, j, i), ue = U(u, d, j, i + 1), uw = U(u, d, j, i - 1);
    const double un = j + 1 < d.ny ? U(u, d, j + 1, i) : -uc;
    const double us = j > 0 ? U(u, d, j - 1, i) : -uc;
    const double e = mul(0.5, add(uc, ue)), w = mul(0.5, add(uc, uw));
    const double n = mul(0.5, add(V(v, d, j + 1, i - 1), V(v, d, j + 1, i)));
    const double s = mul(0.5, add(V(v, d, j, i - 1), V(v, d, j, i)));
    double flux = dvd(sub(mul(mul(e, 0.5), add(uc, ue)), mul(mul(w, 0.5), add(uc, uw))), d.hx);
    flux = add(flux, dvd(sub(mul(mul(n, 0.5), add(uc, un)), mul(mul(s, 0.5), add(uc, us))), d.hy));
    r = sub(flux, mul(mul(0.5, uc), add(dvd(sub(e, w), d.hx), dvd(sub(n, s), d.hy))));
  } else {
    const SmallDiv dn(d.nx);
    const int g = f - d.nu;
    const int j = dn.quot(g) + 1, i = g - (j - 1) * d.nx;  // full index of the v face
    const double vc = V(v, d, j, i), vn = V(v, d, j + 1, i), vs = V(v, d, j - 1, i);
    const double ve = i + 1 < d.nx ? V(v, d, j, i + 1) : -vc;
    const double vw = i > 0 ? V(v, d, j, i - 1) : -vc;
    const double e = mul(0.5, add(U(u, d, j - 1, i + 1), U(u, d, j, i + 1)));
    const double w = mul(0.5, add(U(u, d, j - 1, i), U(u, d, j, i)));
    const double n = mul(0.5, add(vc, vn)), s = mul(0.5, add(vc, vs));
    double flux = dvd(sub(mul(mul(e, 0.5), add(vc, ve)), mul(mul(w, 0.5), add(vc, vw))), d.hx);
    flux = add(flux, dvd(sub(mul(mul(n, 0.5), add(vc, vn)), mul(mul(s, 0.5), add(vc, vs))), d.hy));
    r = sub(flux, mul(mul(0.5, vc), add(dvd(sub(e, w), d.hx), dvd(sub(n, s), d.hy))));
  }
  out[(int64_t)c * ld_out + f] = r;
}

constexpr int kMaxTerms = 8;
struct Terms {
  const double* f[kMaxTerms];
  double c[kMaxTerms];
  int count;
};

// out = ((0 + c0 f0) + c1 f1) + ...: the reference's left-to-right accumulation.
__global__ __launch_bounds__(256) void k_combine(Terms t, int64_t n, double* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int k = 0; k < t.count; ++k) acc = add(acc, mul(t.c[k], t.f[k][i]));
    out[i] = acc;
  }
}

// (K u) row by row, nonzeros in ascending column order (the CSR matvec's order):
// x-neighbours -1/hx^2, y-neighbours -1/hy^2, diagonal 2/h^2 in the normal direction and
// 3/h^2 (tangential wall rows) or 2/h^2 in the tangential one.
__global__ __launch_bounds__(256) void k_apply_K(const double* __restrict__ vel, Dims d, double* __restrict__ out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= d.size) return;
  const double hx2 = mul(d.hx, d.hx), hy2 = mul(d.hy, d.hy);
  const double ox = dvd(-1.0, hx2), oy = dvd(-1.0, hy2);
  double acc = 0.0;
  if (f < d.nu) {
    const int w = d.nx - 1;
    const SmallDiv dw(w);
    const int j = dw.quot(f), i = f - j * w;
    const double diag = add(dvd(2.0, hx2), dvd((j == 0 || j == d.ny - 1) ? 3.0 : 2.0, hy2));
    if (j > 0) acc = add(acc, mul(oy, vel[f - w]));
    if (i > 0) acc = add(acc, mul(ox, vel[f - 1]));
    acc = add(acc, mul(diag, vel[f]));
    if (i < w - 1) acc = add(acc, mul(ox, vel[f + 1]));
    if (j < d.ny - 1) acc = add(acc, mul(oy, vel[f + w]));
  } else {
    const int e = f - d.nu;
    const SmallDiv dn(d.nx);
    const int j = dn.quot(e), i = e - j * d.nx;
    const double* v = vel + d.nu;
    const double diag = add(dvd((i == 0 || i == d.nx - 1) ? 3.0 : 2.0, hx2), dvd(2.0, hy2));
    if (j > 0) acc = add(acc, mul(oy, v[e - d.nx]));
    if (i > 0) acc = add(acc, mul(ox, v[e - 1]));
    acc = add(acc, mul(diag, v[e]));
    if (i < d.nx - 1) acc = add(acc, mul(ox, v[e + 1]));
    if (j < d.ny - 2) acc = add(acc, mul(oy, v[e + d.nx]));
  }
  out[f] = acc;
}

// G p = -D^T p as the CSR rows of -D^T (operators.py:25-29): each face row holds
// -1/h at its lower cell and +1/h at its upper cell, summed in column order.
__global__ __launch_bounds__(256) void k_gradient(const double* __restrict__ p, Dims d, double* __restrict__ out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= d.size) return;
  double lo, hi, h;
  if (f < d.nu) {
    const SmallDiv dw(d.nx - 1);
    const int j = dw.quot(f), i = f - j * (d.nx - 1);
    lo = p[j * d.nx + i];
    hi = p[j * d.nx + i + 1];
    h = d.hx;
  } else {
    const int e = f - d.nu;
    lo = p[e];
    hi = p[e + d.nx];
    h = d.hy;
  }
  const double inv = dvd(1.0, h);
  out[f] = add(add(0.0, mul(-inv, lo)), mul(inv, hi));
}

// Scratch for the scalar reductions (one per device, guarded; calls are synchronous).
std::mutex g_scratch_mutex;
std::vector<double*> g_scratch;
std::vector<size_t> g_scratch_len;

double* scratch(size_t n) {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  if ((int)g_scratch.size() <= dev) {
    g_scratch.resize(dev + 1, nullptr);
    g_scratch_len.resize(dev + 1, 0);
  }
  if (g_scratch_len[dev] < n) {
    if (g_scratch[dev]) HIP_CHECK(hipFree(g_scratch[dev]));
    HIP_CHECK(hipMalloc(&g_scratch[dev], sizeof(double) * n));
    g_scratch_len[dev] = n;
  }
  return g_scratch[dev];
}

thread_local std::string g_model_error;

template <typename F>
int32_t model_guard(F&& body) {
  try {
    g_model_error.clear();
    body();
    return MAC_OK;
  } catch (const ArgFailure& e) {
    g_model_error = e.what();
    return MAC_INVALID_ARGUMENT;
  } catch (const std::exception& e) {
    g_model_error = e.what();
    return MAC_HIP_ERROR;
  }
}

void check_grid(int nx, int ny, double lx, double ly) {
  if (nx < 3 || ny < 3 || nx > 16384 || ny > 16384 || !(lx > 0) || !(ly > 0) || !std::isfinite(lx) ||
      !std::isfinite(ly))
    throw ArgFailure("Require 3 <= nx,ny <= 16384 and positive finite domain lengths");
}

}  // namespace
}  // namespace macs

using namespace macs;

extern "C" {

const char* mac_model_last_error(void) { return g_model_error.c_str(); }

int32_t mac_convection(int32_t nx, int32_t ny, double lx, double ly, const double* velocity, int64_t ld,
                       int32_t ncols, double* out, int64_t ld_out) {
  return model_guard([&] {
    check_grid(nx, ny, lx, ly);
    const Dims d = dims(nx, ny, lx, ly);
    if (!velocity || !out || ncols < 1 || ncols > 64 || ld < d.size || ld_out < d.size)
      throw ArgFailure("Invalid convection arguments");
    const dim3 grid((unsigned)((d.size + 255) / 256), (unsigned)ncols);
    k_convection<<<grid, 256, 0, 0>>>(velocity, ld, d, out, ld_out);
    HIP_CHECK(hipGetLastError());
  });
}

int32_t mac_combine(int64_t n, int32_t nterms, const double* coefficients, const double* const* fields,
                    double* out) {
  return model_guard([&] {
    if (n < 1 || nterms < 1 || nterms > kMaxTerms || !coefficients || !fields || !out)
      throw ArgFailure("Invalid combination arguments");
    Terms t{};
    t.count = nterms;
    for (int k = 0; k < nterms; ++k) {
      if (!fields[k] || !std::isfinite(coefficients[k])) throw ArgFailure("Invalid combination term");
      t.f[k] = fields[k];
      t.c[k] = coefficients[k];
    }
    k_combine<<<reduce_blocks(n), 256, 0, 0>>>(t, n, out);
    HIP_CHECK(hipGetLastError());
  });
}

int32_t mac_apply_K(int32_t nx, int32_t ny, double lx, double ly, const double* velocity, double* out) {
  return model_guard([&] {
    check_grid(nx, ny, lx, ly);
    const Dims d = dims(nx, ny, lx, ly);
    if (!velocity || !out) throw ArgFailure("Invalid K arguments");
    k_apply_K<<<(unsigned)((d.size + 255) / 256), 256, 0, 0>>>(velocity, d, out);
    HIP_CHECK(hipGetLastError());
  });
}

int32_t mac_grid_divergence(int32_t nx, int32_t ny, double lx, double ly, const double* velocity, double* out) {
  return model_guard([&] {
    check_grid(nx, ny, lx, ly);
    if (!velocity || !out) throw ArgFailure("Invalid divergence arguments");
    const Dims d = dims(nx, ny, lx, ly);
    GridDims g{nx, ny, d.hx, d.hy};
    VelLayout L{nx - 1, nx, d.nu, d.size};
    divergence(DT_F64, velocity, L, g, out, (int64_t)nx * ny, 1.0, 1, RED_NONE, nullptr, nullptr, 0);
    HIP_CHECK(hipGetLastError());
  });
}

int32_t mac_grid_gradient(int32_t nx, int32_t ny, double lx, double ly, const double* pressure, double* out) {
  return model_guard([&] {
    check_grid(nx, ny, lx, ly);
    if (!pressure || !out) throw ArgFailure("Invalid gradient arguments");
    const Dims d = dims(nx, ny, lx, ly);
    k_gradient<<<(unsigned)((d.size + 255) / 256), 256, 0, 0>>>(pressure, d, out);
    HIP_CHECK(hipGetLastError());
  });
}

int32_t mac_dot(int64_t n, const double* a, const double* b, double* result) {
  return model_guard([&] {
    if (n < 1 || !a || !b || !result) throw ArgFailure("Invalid dot arguments");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    const int nb = reduce_blocks(n);
    double* part = scratch(nb + 1);
    reduce(a, b, n, n, 1, RED_SUM, part, nullptr, 0);
    finalize(part, nb, 1, RED_SUM, part + nb, 0);
    HIP_CHECK(hipMemcpy(result, part + nb, sizeof(double), hipMemcpyDeviceToHost));
  });
}

int32_t mac_max_abs(int64_t n, const double* a, double* result) {
  return model_guard([&] {
    if (n < 1 || !a || !result) throw ArgFailure("Invalid max-norm arguments");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    const int nb = reduce_blocks(n);
    double* part = scratch(nb + 1);
    reduce(a, nullptr, n, n, 1, RED_MAXABS, part, nullptr, 0);
    finalize(part, nb, 1, RED_MAXABS, part + nb, 0);
    HIP_CHECK(hipMemcpy(result, part + nb, sizeof(double), hipMemcpyDeviceToHost));
  });
}

int32_t mac_reduce_into(int64_t n, const double* a, const double* b, int32_t mode, double* out) {
  return model_guard([&] {
    if (n < 1 || !a || !out || mode < 0 || mode > 1 || (mode == 0 && !b))
      throw ArgFailure("Invalid reduction arguments");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    const int nb = reduce_blocks(n);
    const int red = mode == 0 ? RED_SUM : RED_MAXABS;
    double* part = scratch(nb);
    reduce(a, mode == 0 ? b : nullptr, n, n, 1, red, part, nullptr, 0);
    finalize(part, nb, 1, red, out, 0);
    HIP_CHECK(hipGetLastError());
  });
}

}  // extern "C"
