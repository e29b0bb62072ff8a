// Test infrastructure (the oracle): helpers shared by the ODT and BPM restatements.
//
// Restated from the reference's src/common.hpp (index maps, fista_t) and src/cpu/tv.cpp with
// src/cpu/tv.hpp (the FGP TV prox, used by both the ODT solver and the BPM inversion).
#pragma once

#include <cmath>
#include <cstddef>
#include <vector>

namespace oqh {

template <class T>
constexpr T kPi = static_cast<T>(3.14159265358979323846);

template <class T>
T parity(int i) { return (i & 1) ? T(-1) : T(1); }                       // common.hpp:19
inline int signed_bin(int k, int n) { return k < n / 2 ? k : k - n; }      // common.hpp:22
inline int dct_position(int i, int n) { return (i & 1) ? n - 1 - (i >> 1) : (i >> 1); }  // :25
inline int negative_bin(int k, int n) { return k == 0 ? 0 : n - k; }       // common.hpp:70

template <class T>
T fista_t(T t) { return T(0.5) * (T(1) + std::sqrt(T(1) + T(4) * t * t)); }  // common.hpp:153

// Dual of the TV prox, one array per axis (z, y, x): tv.hpp:9-18.
template <class T>
struct Dual {
  std::vector<T> p[3];
  std::vector<T> r[3];
  explicit Dual(std::size_t size) {
    for (int a = 0; a < 3; ++a) {
      p[a].assign(size, T(0));
      r[a].assign(size, T(0));
    }
  }
};

// common.hpp:107-118: L(p) at (z, y, x), the negative divergence.
template <class T>
T tv_adjoint(const T* pz, const T* py, const T* px, int nz, int ny, int nx, int z, int y, int x) {
  const long i = (static_cast<long>(z) * ny + y) * nx + x;
  const long plane = static_cast<long>(ny) * nx;
  T s = 0;
  if (z < nz - 1) s += pz[i];
  if (z > 0) s -= pz[i - plane];
  if (y < ny - 1) s += py[i];
  if (y > 0) s -= py[i - nx];
  if (x < nx - 1) s += px[i];
  if (x > 0) s -= px[i - 1];
  return s;
}

template <class T>
T tv_primal(T b, T lambda, T adjoint) {  // common.hpp:121-124
  const T x = b - lambda * adjoint;
  return x > T(0) ? x : T(0);
}

// tv.cpp:7-44, with tv_gradient (common.hpp:144-150) and tv_dual_step (common.hpp:128-141).
template <class T>
void tv_prox(const T* b, T* x, Dual<T>& d, int nz, int ny, int nx, T lambda, int inner) {
  const std::ptrdiff_t total = static_cast<std::ptrdiff_t>(nz) * ny * nx;
  const long plane = static_cast<long>(ny) * nx;
  const T step = T(1) / (T(12) * lambda);
  for (int a = 0; a < 3; ++a) d.r[a] = d.p[a];
  const auto primal = [&](const std::vector<T>* q) {
#pragma omp parallel for
    for (std::ptrdiff_t i = 0; i < total; ++i) {
      const int xx = static_cast<int>(i % nx);
      const int yy = static_cast<int>((i / nx) % ny);
      const int zz = static_cast<int>(i / plane);
      x[i] = tv_primal<T>(b[i], lambda,
                          tv_adjoint<T>(q[0].data(), q[1].data(), q[2].data(), nz, ny, nx, zz, yy, xx));
    }
  };
  T t = 1;
  for (int k = 0; k < inner; ++k) {
    primal(d.r);
    const T t_next = fista_t<T>(t);
    const T momentum = (t - T(1)) / t_next;
#pragma omp parallel for
    for (std::ptrdiff_t i = 0; i < total; ++i) {
      const int xx = static_cast<int>(i % nx);
      const int yy = static_cast<int>((i / nx) % ny);
      const int zz = static_cast<int>(i / plane);
      const T c = x[i];
      T g[3];
      g[0] = zz < nz - 1 ? c - x[i + plane] : T(0);
      g[1] = yy < ny - 1 ? c - x[i + nx] : T(0);
      g[2] = xx < nx - 1 ? c - x[i + 1] : T(0);
      T q[3];
      T norm2 = 0;
      for (int a = 0; a < 3; ++a) {
        q[a] = d.r[a][i] + step * g[a];
        norm2 += q[a] * q[a];
      }
      const T scale = norm2 > T(1) ? T(1) / std::sqrt(norm2) : T(1);
      for (int a = 0; a < 3; ++a) {
        const T next = q[a] * scale;
        d.r[a][i] = next + momentum * (next - d.p[a][i]);
        d.p[a][i] = next;
      }
    }
    t = t_next;
  }
  primal(d.p);
}

}  // namespace oqh
