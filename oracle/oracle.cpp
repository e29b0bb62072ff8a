// Test infrastructure: the CPU oracle. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load this library, as the checker; the product never links it.
//
// A restatement of the reference's CPU backend (NaNIsaac/QoraHolo, read as text), templated
// on the working precision so the same algorithm runs in float (the reference's precision)
// and in double (the FP64 statement). Each function cites the reference lines it follows.
// Expressions that decide integer results (Ewald bins, crop masks, DCT reorder) and the order
// of float operations follow the reference; the translation unit is compiled with
// -ffp-contract=off so no multiply-add is fused, as in the reference's host build.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <numbers>
#include <stdexcept>
#include <string>
#include <vector>

#include <omp.h>

#include "fft.hpp"
#include "shared.hpp"

namespace oqh {

// ---------------------------------------------------------------------------------------------
// Shared helpers: src/common.hpp:19-183.


template <class T>
T neumann_eigenvalue(int ky, int kx, int ny, int nx) {  // common.hpp:28-31
  return T(2) * std::cos(kPi<T> * static_cast<T>(ky) / static_cast<T>(ny)) +
         T(2) * std::cos(kPi<T> * static_cast<T>(kx) / static_cast<T>(nx)) - T(4);
}

template <class T>
T wrap_phase(T p) {  // common.hpp:33
  return p - T(2) * kPi<T> * std::floor((p + kPi<T>) / (T(2) * kPi<T>));
}

template <class T>
struct Grid {  // common.hpp:36-44
  int n, nz;
  T df, dfz, fm2, fna2, gain;
};

// api.cpp:32-42 (make_grid): float arithmetic on the geometry values.
template <class T>
Grid<T> make_grid(int n, int nz, T wavelength, T n_medium, T na, T pixel) {
  const T fm = n_medium / wavelength;
  const T fna = na / wavelength;
  return {n, nz, T(1) / (static_cast<T>(n) * pixel), T(1) / (static_cast<T>(nz) * pixel),
          fm * fm, fna * fna, T(4) * kPi<T> / pixel};
}

// common.hpp:56-67
template <class T>
bool ewald_target(const Grid<T>& g, int ky, int kx, T fiy, T fix, T fiz, int& kz, T& nuz) {
  const T fy = static_cast<T>(signed_bin(ky, g.n)) * g.df + fiy;
  const T fx = static_cast<T>(signed_bin(kx, g.n)) * g.df + fix;
  const T rho2 = fy * fy + fx * fx;
  if (rho2 > g.fna2 || rho2 >= g.fm2) return false;
  nuz = std::sqrt(g.fm2 - rho2);
  const int k = static_cast<int>(std::floor((nuz - fiz) / g.dfz + T(0.5)));
  if (k < -g.nz / 2 || k >= g.nz / 2) return false;
  kz = k < 0 ? k + g.nz : k;
  return true;
}

template <class T>
T index_from_potential(T re, T im, T nm2, T inv_k02) {  // common.hpp:179-183
  const T a = nm2 + re * inv_k02;
  const T b = im * inv_k02;
  return std::sqrt(T(0.5) * (std::hypot(a, b) + a));
}


// ---------------------------------------------------------------------------------------------
// FFT over chosen axes of a (d0, d1, d2) array: src/cpu/fft.cpp:18-66.

template <class T>
void fft3(cx<T>* data, int d0, int d1, int d2, unsigned axes, bool forward, T scale) {
  const int dims[3] = {d0, d1, d2};
  const long strides[3] = {static_cast<long>(d1) * d2, d2, 1};
  bool first = true;
  for (int axis = 2; axis >= 0; --axis) {
    if (!(axes & (1u << axis))) continue;
    const int len = dims[axis];
    const long st = strides[axis];
    const int o1 = axis == 0 ? d1 : d0;
    const int o2 = axis == 2 ? d1 : d2;
    const long s1 = axis == 0 ? strides[1] : strides[0];
    const long s2 = axis == 2 ? strides[1] : strides[2];
    const Fft1<T>& p = plan<T>(len);
    const T f = first ? scale : T(1);
    first = false;
#pragma omp parallel
    {
      std::vector<cx<T>> line(len), work(len);
#pragma omp for schedule(static)
      for (long l = 0; l < static_cast<long>(o1) * o2; ++l) {
        const long base = (l / o2) * s1 + (l % o2) * s2;
        for (int i = 0; i < len; ++i) line[i] = data[base + i * st];
        p.run(line.data(), work.data(), forward);
        for (int i = 0; i < len; ++i) data[base + i * st] = f == T(1) ? line[i] : line[i] * f;
      }
    }
  }
}

// r2c over all three axes into (nz, ny, nx/2+1): fft.cpp:49-54 (last axis first, then c2c).
template <class T>
void rfft3(const T* in, cx<T>* out, int nz, int ny, int nx) {
  const int h = nx / 2 + 1;
#pragma omp parallel
  {
    std::vector<cx<T>> work(nx);
#pragma omp for schedule(static)
    for (long l = 0; l < static_cast<long>(nz) * ny; ++l) {
      rfft_line<T>(in + l * nx, 1, out + l * h, 1, nx, work.data());
    }
  }
  fft3<T>(out, nz, ny, h, 0b011u, true, T(1));
}

// fft.cpp:56-66: c2c over axes 0 and 1 in place, then c2r along axis 2 times scale.
template <class T>
void irfft3(cx<T>* in, T* out, int nz, int ny, int nx, T scale) {
  const int h = nx / 2 + 1;
  fft3<T>(in, nz, ny, h, 0b011u, false, T(1));
#pragma omp parallel
  {
    std::vector<cx<T>> work(nx);
#pragma omp for schedule(static)
    for (long l = 0; l < static_cast<long>(nz) * ny; ++l) {
      irfft_line<T>(in + l * h, 1, out + l * nx, 1, nx, scale, work.data());
    }
  }
}

// DCT-II along axis 2 (x) or 1 (y) via a complex FFT (Makhoul), fft.cpp:70-133.
template <class T>
std::vector<cx<T>> dct_twiddles(int len, double sign) {  // fft.cpp:72-79
  std::vector<cx<T>> w(len);
  for (int k = 0; k < len; ++k) {
    const double a = sign * std::numbers::pi * k / (2.0 * len);
    w[k] = {static_cast<T>(std::cos(a)), static_cast<T>(std::sin(a))};
  }
  return w;
}

template <class T, class F>
void for_each_line(int nz, int ny, int nx, int axis, F&& f) {  // fft.cpp:81-99
  const int len = axis == 2 ? nx : ny;
  const long step = axis == 2 ? 1 : nx;
  const int lines = nz * (axis == 2 ? ny : nx);
#pragma omp parallel for
  for (int line = 0; line < lines; ++line) {
    long base = 0;
    if (axis == 2) {
      base = static_cast<long>(line) * nx;
    } else {
      const int z = line / nx;
      const int x = line % nx;
      base = static_cast<long>(z) * ny * nx + x;
    }
    f(base, step, len);
  }
}

template <class T>
void dct2(T* a, int nz, int ny, int nx, int axis) {  // fft.cpp:101-115
  const int len = axis == 2 ? nx : ny;
  std::vector<cx<T>> buf(static_cast<std::size_t>(nz) * ny * nx);
  for_each_line<T>(nz, ny, nx, axis, [&](long base, long step, int n) {
    for (int i = 0; i < n; ++i) buf[base + dct_position(i, n) * step] = a[base + i * step];
  });
  fft3<T>(buf.data(), nz, ny, nx, 1u << axis, true, T(1));
  const auto w = dct_twiddles<T>(len, -1.0);
  for_each_line<T>(nz, ny, nx, axis, [&](long base, long step, int n) {
    for (int k = 0; k < n; ++k) a[base + k * step] = (w[k] * buf[base + k * step]).real();
  });
}

template <class T>
void idct2(T* a, int nz, int ny, int nx, int axis) {  // fft.cpp:117-133
  const int len = axis == 2 ? nx : ny;
  std::vector<cx<T>> buf(static_cast<std::size_t>(nz) * ny * nx);
  const auto w = dct_twiddles<T>(len, 1.0);
  for_each_line<T>(nz, ny, nx, axis, [&](long base, long step, int n) {
    for (int k = 0; k < n; ++k) {
      const T mirror = k == 0 ? T(0) : a[base + (n - k) * step];
      buf[base + k * step] = w[k] * cx<T>(a[base + k * step], -mirror);
    }
  });
  fft3<T>(buf.data(), nz, ny, nx, 1u << axis, false, T(1) / static_cast<T>(len));
  for_each_line<T>(nz, ny, nx, axis, [&](long base, long step, int n) {
    for (int i = 0; i < n; ++i) a[base + i * step] = buf[base + dct_position(i, n) * step].real();
  });
}

// ---------------------------------------------------------------------------------------------
// Field extraction: src/cpu/field.cpp:78-121 (crop plan resolved by the caller, api.cpp:52-60).

template <class T>
void extract_fields(const float* holo, int count, int n, int cy, int cx_, float radius, int m,
                    cx<T>* out) {
  const std::size_t nn = static_cast<std::size_t>(n) * n;
  std::vector<cx<T>> spec(nn * count);
#pragma omp parallel for
  for (int z = 0; z < count; ++z) {
    for (int y = 0; y < n; ++y) {
      for (int x = 0; x < n; ++x) {
        spec[z * nn + y * n + x] = static_cast<T>(holo[z * nn + y * n + x]) * parity<T>(x + y);
      }
    }
  }
  fft3<T>(spec.data(), count, n, n, 0b110u, true, T(1));
  const T sign = parity<T>(cy + cx_);
  const T r2 = static_cast<T>(radius) * static_cast<T>(radius);
  const std::size_t mm = static_cast<std::size_t>(m) * m;
  for (int z = 0; z < count; ++z) {
    const cx<T>* s = spec.data() + z * nn;
    cx<T>* o = out + z * mm;
    for (int j = 0; j < m; ++j) {
      const int dy = j - m / 2;
      const int sy = n / 2 + cy + dy;
      for (int i = 0; i < m; ++i) {
        const int dx = i - m / 2;
        const int sx = n / 2 + cx_ + dx;
        const bool inside =
            static_cast<T>(dy * dy + dx * dx) <= r2 && sy >= 0 && sy < n && sx >= 0 && sx < n;
        o[j * m + i] = inside ? s[sy * n + sx] * sign : cx<T>{};
      }
    }
  }
  const T scale = T(1) / static_cast<T>(nn);
  fft3<T>(out, count, m, m, 0b110u, false, scale);
  for (int z = 0; z < count; ++z) {
    cx<T>* o = out + z * mm;
    for (int y = 0; y < m; ++y) {
      for (int x = 0; x < m; ++x) o[y * m + x] *= parity<T>(x + y);
    }
  }
}

// field.cpp:62-74: sample / background, one background or one per sample.
template <class T>
void field_ratio(const cx<T>* s, int count, const cx<T>* b, int bcount, int plane, cx<T>* out) {
  for (int z = 0; z < count; ++z) {
    const cx<T>* bb = b + static_cast<std::size_t>(bcount == 1 ? 0 : z) * plane;
    for (int i = 0; i < plane; ++i) {
      out[static_cast<std::size_t>(z) * plane + i] = s[static_cast<std::size_t>(z) * plane + i] / bb[i];
    }
  }
}

// field.cpp:14-60: sideband search on the modulated spectrum of the first hologram.
inline float refine(std::complex<float> lo, std::complex<float> mid, std::complex<float> hi,
                    int n) {
  const std::complex<float> den = 2.0f * mid - lo - hi;
  if (std::abs(den) == 0.0f) return 0.0f;
  const float c = std::numbers::pi_v<float> / static_cast<float>(n);
  const float d = ((lo - hi) / den).real() * std::tan(c) / c;
  return std::clamp(d, -0.5f, 0.5f);
}

// ---------------------------------------------------------------------------------------------
// Unwrap and Rytov: src/cpu/unwrap.cpp:11-82.

template <class T>
void unwrap(const T* psi, int nz, int ny, int nx, T* phi) {
  const std::size_t plane = static_cast<std::size_t>(ny) * nx;
#pragma omp parallel for
  for (int z = 0; z < nz; ++z) {
    for (int y = 0; y < ny; ++y) {
      for (int x = 0; x < nx; ++x) {
        const std::size_t i = z * plane + static_cast<std::size_t>(y) * nx + x;
        const T p = psi[i];
        T r = 0;
        if (x + 1 < nx) r += wrap_phase<T>(psi[i + 1] - p);
        if (x > 0) r -= wrap_phase<T>(p - psi[i - 1]);
        if (y + 1 < ny) r += wrap_phase<T>(psi[i + nx] - p);
        if (y > 0) r -= wrap_phase<T>(p - psi[i - nx]);
        phi[i] = r;
      }
    }
  }
  dct2<T>(phi, nz, ny, nx, 2);
  dct2<T>(phi, nz, ny, nx, 1);
#pragma omp parallel for
  for (int z = 0; z < nz; ++z) {
    for (int ky = 0; ky < ny; ++ky) {
      for (int kx = 0; kx < nx; ++kx) {
        T& v = phi[z * plane + static_cast<std::size_t>(ky) * nx + kx];
        v = (ky == 0 && kx == 0) ? T(0) : v / neumann_eigenvalue<T>(ky, kx, ny, nx);
      }
    }
  }
  idct2<T>(phi, nz, ny, nx, 1);
  idct2<T>(phi, nz, ny, nx, 2);
  // unwrap.cpp:48-57: congruence constant, summed in double.
#pragma omp parallel for
  for (int z = 0; z < nz; ++z) {
    std::complex<double> acc = 0;
    const T* w = psi + z * plane;
    T* u = phi + z * plane;
    for (std::size_t i = 0; i < plane; ++i) acc += std::polar(1.0, double(w[i] - u[i]));
    const auto c = static_cast<T>(std::arg(acc));
    for (std::size_t i = 0; i < plane; ++i) u[i] += c;
  }
}

template <class T>
void rytov(const cx<T>* ratio, int nz, int ny, int nx, cx<T>* out) {  // unwrap.cpp:61-82
  const std::size_t plane = static_cast<std::size_t>(ny) * nx;
  const std::size_t total = plane * nz;
  std::vector<T> wrapped(total), phase(total);
  for (std::size_t i = 0; i < total; ++i) wrapped[i] = std::arg(ratio[i]);
  unwrap<T>(wrapped.data(), nz, ny, nx, phase.data());
  for (int z = 0; z < nz; ++z) {
    const T* ph = phase.data() + z * plane;
    double sum = 0;
    for (int x = 0; x < nx; ++x) sum += ph[x] + ph[static_cast<std::size_t>(ny - 1) * nx + x];
    for (int y = 1; y < ny - 1; ++y) sum += ph[static_cast<std::size_t>(y) * nx] + ph[static_cast<std::size_t>(y) * nx + nx - 1];
    const double mean = sum / (2.0 * nx + 2.0 * (ny - 2));
    const auto shift =
        static_cast<T>(2.0 * std::numbers::pi * std::round(mean / (2.0 * std::numbers::pi)));
    for (std::size_t i = z * plane; i < (z + 1) * plane; ++i) {
      out[i] = {std::log(std::abs(ratio[i])), phase[i] - shift};
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Reconstruction: src/cpu/odt.cpp:31-242 and the TV prox src/cpu/tv.cpp:7-44.

template <class T>
struct HalfData {  // odt.cpp:31-35
  std::vector<cx<T>> a;
  std::vector<std::uint8_t> hits;
};

template <class T>
HalfData<T> hermitian_data(const std::vector<cx<T>>& p, const std::vector<T>& weight, int nz,
                           int n) {  // odt.cpp:37-62, common.hpp:79-82
  const int h = n / 2 + 1;
  HalfData<T> d{std::vector<cx<T>>(static_cast<std::size_t>(nz) * n * h),
                std::vector<std::uint8_t>(static_cast<std::size_t>(nz) * n * h)};
#pragma omp parallel for
  for (int kz = 0; kz < nz; ++kz) {
    const int mz = negative_bin(kz, nz);
    for (int ky = 0; ky < n; ++ky) {
      const int my = negative_bin(ky, n);
      for (int kx = 0; kx < h; ++kx) {
        const std::size_t i = (static_cast<std::size_t>(kz) * n + ky) * n + kx;
        const std::size_t j = (static_cast<std::size_t>(mz) * n + my) * n + negative_bin(kx, n);
        const std::size_t o = (static_cast<std::size_t>(kz) * n + ky) * h + kx;
        const T re = T(0.5) * (p[i].real() + p[j].real());
        const T im = T(0.5) * (p[i].imag() - p[j].imag());
        d.a[o] = {re, im};
        d.hits[o] = static_cast<std::uint8_t>((weight[i] > 0) + (weight[j] > 0));
      }
    }
  }
  return d;
}

template <class T>
void restore(std::vector<cx<T>>& e, const HalfData<T>& d) {  // odt.cpp:64-73, common.hpp:83-87
  const auto total = static_cast<std::ptrdiff_t>(e.size());
#pragma omp parallel for
  for (std::ptrdiff_t i = 0; i < total; ++i) {
    const T b = T(1) - T(0.5) * static_cast<T>(d.hits[i]);
    e[i] = {d.a[i].real() + b * e[i].real(), d.a[i].imag() + b * e[i].imag()};
  }
}

template <class T>
std::vector<T> gp_real(const HalfData<T>& d, int nz, int n, int iterations) {  // odt.cpp:76-90
  const T inv = T(1) / (static_cast<T>(nz) * static_cast<T>(n) * static_cast<T>(n));
  std::vector<T> x(static_cast<std::size_t>(nz) * n * n);
  std::vector<cx<T>> e = d.a;
  irfft3<T>(e.data(), x.data(), nz, n, n, inv);
  const auto total = static_cast<std::ptrdiff_t>(x.size());
  for (int it = 0; it < iterations; ++it) {
#pragma omp parallel for
    for (std::ptrdiff_t i = 0; i < total; ++i) x[i] = std::max(x[i], T(0));
    rfft3<T>(x.data(), e.data(), nz, n, n);
    restore<T>(e, d);
    irfft3<T>(e.data(), x.data(), nz, n, n, inv);
  }
  return x;
}


// odt.cpp:94-120
template <class T>
std::vector<T> tv_real(const HalfData<T>& d, int nz, int n, int iterations, T lambda, int inner) {
  const T inv = T(1) / (static_cast<T>(nz) * static_cast<T>(n) * static_cast<T>(n));
  const std::size_t size = static_cast<std::size_t>(nz) * n * n;
  std::vector<T> y(size);
  std::vector<cx<T>> e = d.a;
  irfft3<T>(e.data(), y.data(), nz, n, n, inv);
  std::vector<T> x_prev = y;
  std::vector<T> x(size);
  Dual<T> dual(size);
  const auto total = static_cast<std::ptrdiff_t>(size);
  T t = 1;
  for (int it = 0; it < iterations; ++it) {
    rfft3<T>(y.data(), e.data(), nz, n, n);
    restore<T>(e, d);
    irfft3<T>(e.data(), y.data(), nz, n, n, inv);
    tv_prox<T>(y.data(), x.data(), dual, nz, n, n, lambda, inner);
    const T t_next = fista_t<T>(t);
    const T momentum = (t - T(1)) / t_next;
#pragma omp parallel for
    for (std::ptrdiff_t i = 0; i < total; ++i) y[i] = x[i] + momentum * (x[i] - x_prev[i]);
    std::swap(x, x_prev);
    t = t_next;
  }
  return x_prev;
}

struct Options {
  int nz;
  int iterations;
  int solver;        // 0: Gerchberg-Papoulis, 1: FISTA-TV
  int gp_transform;  // 0: r2c, 1: c2c
  float tv_weight;
  int tv_inner;
};

// odt.cpp:124-242. Optional outputs: the normalised k-space potential (nz, n, n) and the hit
// weights, the intermediates the mapping produces before any solver runs.
template <class T>
double reconstruct(const cx<T>* fields, int count, int n, const float* illum, const float* geom,
                   const Options& o, T* ri, cx<T>* kspace, float* weights) {
  const int nz = o.nz > 0 ? o.nz : n;
  const Grid<T> g = make_grid<T>(n, nz, static_cast<T>(geom[0]), static_cast<T>(geom[1]),
                                 static_cast<T>(geom[2]), static_cast<T>(geom[3]));
  const std::size_t plane = static_cast<std::size_t>(n) * n;
  std::vector<cx<T>> spectra(fields, fields + plane * count);
  fft3<T>(spectra.data(), count, n, n, 0b110u, true, T(1));

  std::vector<T> fiz(count);
  for (int a = 0; a < count; ++a) {
    const T fy = static_cast<T>(illum[2 * a]);
    const T fx = static_cast<T>(illum[2 * a + 1]);
    fiz[a] = std::sqrt(g.fm2 - fy * fy - fx * fx);
  }

  const std::size_t total = plane * nz;
  std::vector<cx<T>> potential(total);
  std::vector<T> weight(total, T(0));
#pragma omp parallel for schedule(dynamic, 64)
  for (int col = 0; col < static_cast<int>(plane); ++col) {
    const int ky = col / n;
    const int kx = col % n;
    for (int a = 0; a < count; ++a) {
      int kz = 0;
      T nuz = 0;
      if (!ewald_target<T>(g, ky, kx, static_cast<T>(illum[2 * a]), static_cast<T>(illum[2 * a + 1]),
                           fiz[a], kz, nuz)) {
        continue;
      }
      const cx<T> v = cx<T>(T(0), -g.gain * nuz) * spectra[a * plane + col];
      const std::size_t i = kz * plane + col;
      potential[i] += parity<T>(kz) * v;
      weight[i] += T(1);
    }
  }
  std::ptrdiff_t filled = 0;
#pragma omp parallel for reduction(+ : filled)
  for (std::ptrdiff_t i = 0; i < static_cast<std::ptrdiff_t>(total); ++i) {
    if (weight[i] > 0) {
      potential[i] /= weight[i];
      ++filled;
    }
  }
  if (kspace) std::copy(potential.begin(), potential.end(), kspace);
  if (weights) {
    for (std::size_t i = 0; i < total; ++i) weights[i] = static_cast<float>(weight[i]);
  }

  const T k0 = T(2) * kPi<T> / static_cast<T>(geom[0]);
  const T nm2 = static_cast<T>(geom[1]) * static_cast<T>(geom[1]);
  const T inv_k02 = T(1) / (k0 * k0);
  const double result_filled = static_cast<double>(filled) / static_cast<double>(total);

  const bool tv = o.solver == 1;
  if (o.iterations > 0 && (tv || o.gp_transform == 0)) {
    const HalfData<T> d = hermitian_data<T>(potential, weight, nz, n);
    potential = {};
    weight = {};
    std::vector<T> x = tv ? tv_real<T>(d, nz, n, o.iterations,
                                       static_cast<T>(o.tv_weight) * k0 * k0, o.tv_inner)
                          : gp_real<T>(d, nz, n, o.iterations);
#pragma omp parallel for
    for (std::ptrdiff_t i = 0; i < static_cast<std::ptrdiff_t>(total); ++i) {
      ri[i] = index_from_potential<T>(x[i], T(0), nm2, inv_k02);
    }
    return result_filled;
  }

  const T inv = T(1) / static_cast<T>(total);
  if (o.iterations > 0) {
    const std::vector<cx<T>> measured = potential;
    fft3<T>(potential.data(), nz, n, n, 0b111u, false, inv);
    for (int it = 0; it < o.iterations; ++it) {
#pragma omp parallel for
      for (std::ptrdiff_t i = 0; i < static_cast<std::ptrdiff_t>(total); ++i) {
        potential[i] = {std::max(potential[i].real(), T(0)), T(0)};
      }
      fft3<T>(potential.data(), nz, n, n, 0b111u, true, T(1));
#pragma omp parallel for
      for (std::ptrdiff_t i = 0; i < static_cast<std::ptrdiff_t>(total); ++i) {
        if (weight[i] > 0) potential[i] = measured[i];
      }
      fft3<T>(potential.data(), nz, n, n, 0b111u, false, inv);
    }
  } else {
    fft3<T>(potential.data(), nz, n, n, 0b111u, false, inv);
  }
#pragma omp parallel for
  for (std::ptrdiff_t i = 0; i < static_cast<std::ptrdiff_t>(total); ++i) {
    ri[i] = index_from_potential<T>(potential[i].real(), potential[i].imag(), nm2, inv_k02);
  }
  return result_filled;
}

}  // namespace oqh

// =============================================================================================
// C interface for the Python test harness. Complex arrays are interleaved (re, im) pairs.
// Every entry returns 0 on success and 1 on a thrown exception (message via oqh_last_error).

namespace {
thread_local std::string g_error;

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return 1;
  }
}

template <class T>
oqh::cx<T>* as_cx(T* p) { return reinterpret_cast<oqh::cx<T>*>(p); }
template <class T>
const oqh::cx<T>* as_cx(const T* p) { return reinterpret_cast<const oqh::cx<T>*>(p); }
}  // namespace

extern "C" {

const char* oqh_last_error() { return g_error.c_str(); }

// OpenMP threads of later calls (the CPU baseline times one thread and all of them).
void oqh_set_threads(int threads) { omp_set_num_threads(threads > 0 ? threads : 1); }
int oqh_max_threads() { return omp_get_max_threads(); }

int oqh_fft_c2c_f(float* d, int d0, int d1, int d2, unsigned axes, int forward, float scale) {
  return guarded([&] { oqh::fft3<float>(as_cx(d), d0, d1, d2, axes, forward != 0, scale); });
}
int oqh_fft_c2c_d(double* d, int d0, int d1, int d2, unsigned axes, int forward, double scale) {
  return guarded([&] { oqh::fft3<double>(as_cx(d), d0, d1, d2, axes, forward != 0, scale); });
}
int oqh_rfft3_f(const float* in, float* out, int nz, int ny, int nx) {
  return guarded([&] { oqh::rfft3<float>(in, as_cx(out), nz, ny, nx); });
}
int oqh_rfft3_d(const double* in, double* out, int nz, int ny, int nx) {
  return guarded([&] { oqh::rfft3<double>(in, as_cx(out), nz, ny, nx); });
}
int oqh_irfft3_f(float* in, float* out, int nz, int ny, int nx, float scale) {
  return guarded([&] { oqh::irfft3<float>(as_cx(in), out, nz, ny, nx, scale); });
}
int oqh_irfft3_d(double* in, double* out, int nz, int ny, int nx, double scale) {
  return guarded([&] { oqh::irfft3<double>(as_cx(in), out, nz, ny, nx, scale); });
}
int oqh_dct2_f(float* a, int nz, int ny, int nx, int axis, int inverse) {
  return guarded([&] {
    inverse ? oqh::idct2<float>(a, nz, ny, nx, axis) : oqh::dct2<float>(a, nz, ny, nx, axis);
  });
}
int oqh_dct2_d(double* a, int nz, int ny, int nx, int axis, int inverse) {
  return guarded([&] {
    inverse ? oqh::idct2<double>(a, nz, ny, nx, axis) : oqh::dct2<double>(a, nz, ny, nx, axis);
  });
}

// Spectrum helper for the sideband search: modulated 2D FFT of one hologram (field.cpp:15-22).
int oqh_shifted_spectrum(const float* h, int n, float* out) {
  return guarded([&] {
    auto* s = as_cx(out);
    for (int y = 0; y < n; ++y) {
      for (int x = 0; x < n; ++x) s[y * n + x] = h[y * n + x] * oqh::parity<float>(x + y);
    }
    oqh::fft3<float>(s, 1, n, n, 0b110u, true, 1.0f);
  });
}

// field.cpp:35-60. out: iy, ix (as float), y, x, magnitude.
int oqh_find_sideband(const float* h, int n, float dc_exclusion, float* out) {
  return guarded([&] {
    if (n % 2 != 0 || n < 4) throw std::invalid_argument("find_sideband: even side >= 4");
    if (dc_exclusion <= 0) dc_exclusion = static_cast<float>(n) / 8.0f;
    std::vector<std::complex<float>> s(static_cast<std::size_t>(n) * n);
    for (int y = 0; y < n; ++y) {
      for (int x = 0; x < n; ++x) s[y * n + x] = h[y * n + x] * oqh::parity<float>(x + y);
    }
    oqh::fft3<float>(s.data(), 1, n, n, 0b110u, true, 1.0f);
    const int c = n / 2;
    int biy = 0, bix = 0;
    float mag = 0;
    for (int y = 1; y < n - 1; ++y) {
      for (int x = c; x < n - 1; ++x) {
        const int dy = y - c;
        const int dx = x - c;
        if (dx == 0 && dy <= 0) continue;
        if (static_cast<float>(dy * dy + dx * dx) <= dc_exclusion * dc_exclusion) continue;
        const float m = std::abs(s[y * n + x]);
        if (m > mag) {
          mag = m;
          biy = dy;
          bix = dx;
        }
      }
    }
    if (mag == 0) throw std::runtime_error("find_sideband: no peak outside the DC region");
    const int y = biy + c;
    const int x = bix + c;
    out[0] = static_cast<float>(biy);
    out[1] = static_cast<float>(bix);
    out[2] = static_cast<float>(biy) + oqh::refine(s[(y - 1) * n + x], s[y * n + x], s[(y + 1) * n + x], n);
    out[3] = static_cast<float>(bix) + oqh::refine(s[y * n + x - 1], s[y * n + x], s[y * n + x + 1], n);
    out[4] = mag;
  });
}

int oqh_extract_fields_f(const float* holo, int count, int n, int cy, int cx, float radius, int m,
                         float* out) {
  return guarded([&] { oqh::extract_fields<float>(holo, count, n, cy, cx, radius, m, as_cx(out)); });
}
int oqh_extract_fields_d(const float* holo, int count, int n, int cy, int cx, float radius, int m,
                         double* out) {
  return guarded([&] { oqh::extract_fields<double>(holo, count, n, cy, cx, radius, m, as_cx(out)); });
}
int oqh_field_ratio_f(const float* s, int count, const float* b, int bcount, int plane, float* out) {
  return guarded([&] { oqh::field_ratio<float>(as_cx(s), count, as_cx(b), bcount, plane, as_cx(out)); });
}
int oqh_field_ratio_d(const double* s, int count, const double* b, int bcount, int plane,
                      double* out) {
  return guarded([&] { oqh::field_ratio<double>(as_cx(s), count, as_cx(b), bcount, plane, as_cx(out)); });
}
int oqh_unwrap_f(const float* psi, int nz, int ny, int nx, float* out) {
  return guarded([&] { oqh::unwrap<float>(psi, nz, ny, nx, out); });
}
int oqh_unwrap_d(const double* psi, int nz, int ny, int nx, double* out) {
  return guarded([&] { oqh::unwrap<double>(psi, nz, ny, nx, out); });
}
int oqh_rytov_f(const float* ratio, int nz, int ny, int nx, float* out) {
  return guarded([&] { oqh::rytov<float>(as_cx(ratio), nz, ny, nx, as_cx(out)); });
}
int oqh_rytov_d(const double* ratio, int nz, int ny, int nx, double* out) {
  return guarded([&] { oqh::rytov<double>(as_cx(ratio), nz, ny, nx, as_cx(out)); });
}

// geom: wavelength, n_medium, na, pixel. opts: nz, iterations, solver, gp_transform, tv_inner.
int oqh_reconstruct_f(const float* fields, int count, int n, const float* illum, const float* geom,
                      const int* opts, float tv_weight, float* ri, float* kspace, float* weights,
                      double* filled) {
  return guarded([&] {
    const oqh::Options o{opts[0], opts[1], opts[2], opts[3], tv_weight, opts[4]};
    *filled = oqh::reconstruct<float>(as_cx(fields), count, n, illum, geom, o, ri,
                                      kspace ? as_cx(kspace) : nullptr, weights);
  });
}
int oqh_reconstruct_d(const double* fields, int count, int n, const float* illum, const float* geom,
                      const int* opts, float tv_weight, double* ri, double* kspace, float* weights,
                      double* filled) {
  return guarded([&] {
    const oqh::Options o{opts[0], opts[1], opts[2], opts[3], tv_weight, opts[4]};
    *filled = oqh::reconstruct<double>(as_cx(fields), count, n, illum, geom, o, ri,
                                       kspace ? as_cx(kspace) : nullptr, weights);
  });
}

// One TV prox call from a zero dual (tv.cpp:7-44), for unit tests of the FGP kernel.
int oqh_tv_prox_f(const float* b, float* x, int nz, int ny, int nx, float lambda, int inner) {
  return guarded([&] {
    oqh::Dual<float> d(static_cast<std::size_t>(nz) * ny * nx);
    oqh::tv_prox<float>(b, x, d, nz, ny, nx, lambda, inner);
  });
}
int oqh_tv_prox_d(const double* b, double* x, int nz, int ny, int nx, double lambda, int inner) {
  return guarded([&] {
    oqh::Dual<double> d(static_cast<std::size_t>(nz) * ny * nx);
    oqh::tv_prox<double>(b, x, d, nz, ny, nx, lambda, inner);
  });
}

// api.cpp:203-216 (circular_scan): float arithmetic with std::sin / std::cos on float.
int oqh_circular_scan(int count, float na_illumination, float wavelength, float* out) {
  return guarded([&] {
    out[0] = 0.0f;
    out[1] = 0.0f;
    const float f = na_illumination / wavelength;
    for (int i = 1; i < count; ++i) {
      const float phi = 2.0f * std::numbers::pi_v<float> * static_cast<float>(i - 1) /
                        static_cast<float>(count - 1);
      out[2 * i] = f * std::sin(phi);
      out[2 * i + 1] = f * std::cos(phi);
    }
  });
}

// The Ewald bin of one (column, illumination) pair in float (common.hpp:56-67), for the
// bit-exactness tests of the device mapping. Returns kz or -1 when rejected.
int oqh_ewald_targets(int n, int nz, const float* geom, const float* illum, int count, int* kz_out,
                      float* nuz_out) {
  return guarded([&] {
    const auto g = oqh::make_grid<float>(n, nz, geom[0], geom[1], geom[2], geom[3]);
    const std::size_t plane = static_cast<std::size_t>(n) * n;
    for (int a = 0; a < count; ++a) {
      const float fy = illum[2 * a];
      const float fx = illum[2 * a + 1];
      const float fiz = std::sqrt(g.fm2 - fy * fy - fx * fx);
      for (std::size_t col = 0; col < plane; ++col) {
        int kz = 0;
        float nuz = 0;
        const bool ok = oqh::ewald_target<float>(g, static_cast<int>(col / n),
                                                 static_cast<int>(col % n), fy, fx, fiz, kz, nuz);
        kz_out[a * plane + col] = ok ? kz : -1;
        nuz_out[a * plane + col] = ok ? nuz : 0.0f;
      }
    }
  });
}

}  // extern "C"
