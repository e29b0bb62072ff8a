// Test infrastructure: the CPU oracle's beam propagation model. Only tests/ and bench.py's
// cpu_baseline leg load it, as the checker.
//
// Restates the reference's src/bpm.cpp (bpm_tables, in double on the host) and
// src/cpu/bpm.cpp (split-step forward, hand-written adjoint, tangent-linear model, gradient and
// the FISTA / batched proximal-gradient inversion), templated on the working precision: float
// is the reference's arithmetic, double the FP64 statement used for finite differences.
//
// One deliberate difference: the reference merges per-thread gradients under `omp critical`
// in whatever order the threads arrive (src/cpu/bpm.cpp:108-124), so its float gradient is not
// reproducible bit for bit. Here every illumination's contribution is kept apart and the
// contributions are added to the gradient in illumination order a = 0, 1, ..., the order the
// device path uses too; the per-angle loss sums are added in the same order.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include <omp.h>

#include "fft.hpp"
#include "shared.hpp"

namespace oqh {

namespace {

// Per-illumination tables (src/backends.hpp:61-70), lateral bins in FFT order.
template <class T>
struct BpmTables {
  int n = 0;
  int nz = 0;
  int count = 0;
  T phase = 0;
  std::vector<cx<T>> step;
  std::vector<cx<T>> last;
  std::vector<T> q;
  T lipschitz = 0;
};

// src/bpm.cpp:10-60. Every value is computed in double and stored in T.
template <class T>
BpmTables<T> bpm_tables(int n, int nz, const float* illum, int count, const float* geom,
                        bool obliquity) {
  constexpr double pi = kPi<double>;
  const double lambda = geom[0];
  const double fm = static_cast<double>(geom[1]) / lambda;
  const double fna = static_cast<double>(geom[2]) / lambda;
  const double d = geom[3];
  const double df = 1.0 / (n * d);
  const double nuz_min = std::sqrt(std::max(fm * fm - fna * fna, 0.01 * fm * fm));
  const double back = -(nz - 1 - nz / 2) * d;
  BpmTables<T> t;
  t.n = n;
  t.nz = nz;
  t.count = count;
  t.phase = static_cast<T>(2.0 * pi / lambda * d);
  const std::size_t plane = static_cast<std::size_t>(n) * n;
  t.step.assign(plane * count, cx<T>(0));
  t.last.assign(plane * count, cx<T>(0));
  t.q.assign(plane * count, T(0));
  double q0 = 0;
  for (int a = 0; a < count; ++a) {
    const double fiy = illum[2 * a];
    const double fix = illum[2 * a + 1];
    const double fiz = std::sqrt(fm * fm - fiy * fiy - fix * fix);
    for (int ky = 0; ky < n; ++ky) {
      for (int kx = 0; kx < n; ++kx) {
        const std::size_t i = a * plane + static_cast<std::size_t>(ky) * n + kx;
        const double fy = signed_bin(ky, n) * df + fiy;
        const double fx = signed_bin(kx, n) * df + fix;
        const double rho2 = fy * fy + fx * fx;
        double q = 1;
        if (rho2 < fm * fm) {
          const double nuz = std::sqrt(fm * fm - rho2);
          t.step[i] = cx<T>(std::polar(1.0, 2.0 * pi * (nuz - fiz) * d));
          if (rho2 <= fna * fna) t.last[i] = cx<T>(std::polar(1.0, 2.0 * pi * (nuz - fiz) * back));
          if (obliquity) q = fm / std::max(nuz, nuz_min);
        }
        t.q[i] = static_cast<T>(q);
      }
    }
    q0 += static_cast<double>(t.q[a * plane]) * t.q[a * plane];
  }
  t.lipschitz = static_cast<T>(nz * static_cast<double>(t.phase) * t.phase * q0 / count);
  return t;
}

// 2D transform of one n x n plane, unnormalised, times `scale` (src/cpu/fft.cpp:30-34).
template <class T>
void fft2(cx<T>* data, int n, bool forward, T scale, std::vector<cx<T>>& line,
          std::vector<cx<T>>& work) {
  const Fft1<T>& p = plan<T>(n);
  line.resize(n);
  work.resize(n);
  for (int y = 0; y < n; ++y) p.run(data + static_cast<std::size_t>(y) * n, work.data(), forward);
  for (int x = 0; x < n; ++x) {
    for (int y = 0; y < n; ++y) line[y] = data[static_cast<std::size_t>(y) * n + x];
    p.run(line.data(), work.data(), forward);
    for (int y = 0; y < n; ++y) data[static_cast<std::size_t>(y) * n + x] = line[y];
  }
  if (scale != T(1)) {
    const std::size_t plane = static_cast<std::size_t>(n) * n;
    for (std::size_t i = 0; i < plane; ++i) data[i] *= scale;
  }
}

// src/common.hpp:158-163
template <class T>
void bpm_step(T cre, T cim, T q, T are, T aim, T& sre, T& sim) {
  const T tre = q * are + (T(1) - q) * sre;
  const T tim = q * aim + (T(1) - q) * sim;
  sre = cre * tre - cim * tim;
  sim = cre * tim + cim * tre;
}

// src/common.hpp:168-176
template <class T>
void bpm_step_adjoint(T cre, T cim, T q, T& gre, T& gim, T& are, T& aim) {
  const T ure = cre * gre + cim * gim;
  const T uim = cre * gim - cim * gre;
  are = q * ure;
  aim = q * uim;
  gre = (T(1) - q) * ure;
  gim = (T(1) - q) * uim;
}

// src/cpu/bpm.cpp:18-24: exp(i phase x) per voxel.
template <class T>
std::vector<cx<T>> screens(const T* x, std::size_t size, T phase) {
  std::vector<cx<T>> t(size);
#pragma omp parallel for
  for (std::ptrdiff_t i = 0; i < static_cast<std::ptrdiff_t>(size); ++i) {
    t[i] = std::polar(T(1), phase * x[i]);
  }
  return t;
}

template <class T>
struct Scratch {
  std::vector<cx<T>> line, work;
};

// src/cpu/bpm.cpp:26-37
template <class T>
void step(const BpmTables<T>& t, int a, int k, const cx<T>* spectrum, cx<T>* s) {
  const std::size_t plane = static_cast<std::size_t>(t.n) * t.n;
  const cx<T>* c = (k == t.nz - 1 ? t.last.data() : t.step.data()) + a * plane;
  const T* q = t.q.data() + a * plane;
  for (std::size_t i = 0; i < plane; ++i) {
    T re = s[i].real();
    T im = s[i].imag();
    bpm_step<T>(c[i].real(), c[i].imag(), q[i], spectrum[i].real(), spectrum[i].imag(), re, im);
    s[i] = {re, im};
  }
}

// src/cpu/bpm.cpp:39-62: envelope of illumination a at z = 0 into y; store (nz planes)
// receives the envelope in front of each screen when given.
template <class T>
void forward_angle(const BpmTables<T>& t, const std::vector<cx<T>>& screen, int a, cx<T>* store,
                   cx<T>* y, Scratch<T>& sc) {
  const int n = t.n;
  const std::size_t plane = static_cast<std::size_t>(n) * n;
  const T inv = T(1) / static_cast<T>(plane);
  std::vector<cx<T>> s(plane, cx<T>(0));
  std::vector<cx<T>> w(plane, cx<T>(1));
  std::vector<cx<T>> v(plane);
  s[0] = static_cast<T>(plane);
  for (int k = 0; k < t.nz; ++k) {
    if (k > 0) {
      std::copy(s.begin(), s.end(), w.begin());
      fft2<T>(w.data(), n, false, inv, sc.line, sc.work);
    }
    if (store) std::copy(w.begin(), w.end(), store + k * plane);
    const cx<T>* tk = screen.data() + k * plane;
    for (std::size_t i = 0; i < plane; ++i) v[i] = tk[i] * w[i];
    fft2<T>(v.data(), n, true, T(1), sc.line, sc.work);
    step<T>(t, a, k, v.data(), s.data());
  }
  std::copy(s.begin(), s.end(), y);
  fft2<T>(y, n, false, inv, sc.line, sc.work);
}

// src/cpu/bpm.cpp:64-98: grad += J_a^T r from the envelopes stored by forward_angle.
template <class T>
void adjoint_angle(const BpmTables<T>& t, const std::vector<cx<T>>& screen, int a,
                   const cx<T>* store, const cx<T>* r, T* grad, Scratch<T>& sc) {
  const int n = t.n;
  const std::size_t plane = static_cast<std::size_t>(n) * n;
  const T inv = T(1) / static_cast<T>(plane);
  std::vector<cx<T>> g(r, r + plane);
  std::vector<cx<T>> b(plane);
  std::vector<cx<T>> back(plane);
  fft2<T>(g.data(), n, true, T(1), sc.line, sc.work);
  for (int k = t.nz - 1; k >= 0; --k) {
    const cx<T>* c = (k == t.nz - 1 ? t.last.data() : t.step.data()) + a * plane;
    const T* q = t.q.data() + a * plane;
    for (std::size_t i = 0; i < plane; ++i) {
      const cx<T> gi = g[i] + (k < t.nz - 1 ? back[i] : cx<T>{});
      T gre = gi.real();
      T gim = gi.imag();
      T are = 0;
      T aim = 0;
      bpm_step_adjoint<T>(c[i].real(), c[i].imag(), q[i], gre, gim, are, aim);
      g[i] = {gre, gim};
      b[i] = {are, aim};
    }
    fft2<T>(b.data(), n, false, inv, sc.line, sc.work);
    const cx<T>* tk = screen.data() + k * plane;
    const cx<T>* wk = store + k * plane;
    T* gk = grad + k * plane;
    for (std::size_t i = 0; i < plane; ++i) {
      const cx<T> v = tk[i] * wk[i];
      gk[i] += t.phase * (v.real() * b[i].imag() - v.imag() * b[i].real());
      back[i] = std::conj(tk[i]) * b[i];
    }
    fft2<T>(back.data(), n, true, T(1), sc.line, sc.work);
  }
}

// Runs per_angle(a, contribution) for a in [first, last) on the OpenMP threads, then adds the
// contributions to grad in illumination order.
template <class T, class F>
void ordered_gradient(int first, int last, std::size_t size, T* grad, F&& per_angle) {
  const int threads = std::max(1, omp_get_max_threads());
  std::vector<std::vector<T>> part(std::min(threads, std::max(last - first, 1)));
  for (auto& p : part) p.assign(size, T(0));
  for (int a0 = first; a0 < last; a0 += static_cast<int>(part.size())) {
    const int group = std::min(static_cast<int>(part.size()), last - a0);
#pragma omp parallel for schedule(dynamic)
    for (int j = 0; j < group; ++j) {
      std::fill(part[j].begin(), part[j].end(), T(0));
      per_angle(a0 + j, part[j].data());
    }
#pragma omp parallel for
    for (std::ptrdiff_t i = 0; i < static_cast<std::ptrdiff_t>(size); ++i) {
      for (int j = 0; j < group; ++j) grad[i] += part[j][i];
    }
  }
}

// src/cpu/bpm.cpp:100-129: gradient of 0.5 / (last - first) sum_a |bpm_a(x) - fields_a|^2,
// returning the data term with the factor 0.5 / count.
template <class T>
double gradient(const BpmTables<T>& t, const T* x, const cx<T>* fields, int first, int last,
                T* grad) {
  const std::size_t plane = static_cast<std::size_t>(t.n) * t.n;
  const std::size_t size = plane * t.nz;
  const std::vector<cx<T>> screen = screens<T>(x, size, t.phase);
  std::fill(grad, grad + size, T(0));
  std::vector<double> loss_a(t.count, 0.0);
  ordered_gradient<T>(first, last, size, grad, [&](int a, T* contribution) {
    Scratch<T> sc;
    std::vector<cx<T>> store(size);
    std::vector<cx<T>> y(plane);
    forward_angle<T>(t, screen, a, store.data(), y.data(), sc);
    const cx<T>* d = fields + a * plane;
    double loss = 0;
    for (std::size_t i = 0; i < plane; ++i) {
      y[i] -= d[i];
      loss += std::norm(std::complex<double>(y[i]));
    }
    loss_a[a] = loss;
    adjoint_angle<T>(t, screen, a, store.data(), y.data(), contribution, sc);
  });
  const T scale = T(1) / static_cast<T>(last - first);
  for (std::size_t i = 0; i < size; ++i) grad[i] *= scale;
  double loss = 0;
  for (int a = first; a < last; ++a) loss += loss_a[a];
  return 0.5 * loss / t.count;
}

// src/cpu/bpm.cpp:133-139
template <class T>
void bpm_forward(const BpmTables<T>& t, const T* x, cx<T>* out) {
  const std::size_t plane = static_cast<std::size_t>(t.n) * t.n;
  const std::vector<cx<T>> screen = screens<T>(x, plane * t.nz, t.phase);
#pragma omp parallel for schedule(dynamic)
  for (int a = 0; a < t.count; ++a) {
    Scratch<T> sc;
    forward_angle<T>(t, screen, a, nullptr, out + a * plane, sc);
  }
}

// src/cpu/bpm.cpp:141-176: tangent-linear model J h.
template <class T>
void bpm_jvp(const BpmTables<T>& t, const T* x, const T* h, cx<T>* out) {
  const int n = t.n;
  const std::size_t plane = static_cast<std::size_t>(n) * n;
  const T inv = T(1) / static_cast<T>(plane);
  const std::vector<cx<T>> screen = screens<T>(x, plane * t.nz, t.phase);
#pragma omp parallel for schedule(dynamic)
  for (int a = 0; a < t.count; ++a) {
    Scratch<T> sc;
    std::vector<cx<T>> s(plane, cx<T>(0)), ds(plane, cx<T>(0)), w(plane, cx<T>(1)),
        dw(plane, cx<T>(0)), v(plane), dv(plane);
    s[0] = static_cast<T>(plane);
    for (int k = 0; k < t.nz; ++k) {
      if (k > 0) {
        w = s;
        dw = ds;
        fft2<T>(w.data(), n, false, inv, sc.line, sc.work);
        fft2<T>(dw.data(), n, false, inv, sc.line, sc.work);
      }
      for (std::size_t i = 0; i < plane; ++i) {
        const cx<T> tk = screen[k * plane + i];
        v[i] = tk * w[i];
        dv[i] = tk * (dw[i] + cx<T>(T(0), t.phase * h[k * plane + i]) * w[i]);
      }
      fft2<T>(v.data(), n, true, T(1), sc.line, sc.work);
      fft2<T>(dv.data(), n, true, T(1), sc.line, sc.work);
      step<T>(t, a, k, v.data(), s.data());
      step<T>(t, a, k, dv.data(), ds.data());
    }
    fft2<T>(ds.data(), n, false, inv, sc.line, sc.work);
    std::copy(ds.begin(), ds.end(), out + a * plane);
  }
}

// src/cpu/bpm.cpp:178-189: Re(J^H r), summed over illuminations in order.
template <class T>
void bpm_vjp(const BpmTables<T>& t, const T* x, const cx<T>* r, T* grad) {
  const std::size_t plane = static_cast<std::size_t>(t.n) * t.n;
  const std::size_t size = plane * t.nz;
  const std::vector<cx<T>> screen = screens<T>(x, size, t.phase);
  std::fill(grad, grad + size, T(0));
  ordered_gradient<T>(0, t.count, size, grad, [&](int a, T* contribution) {
    Scratch<T> sc;
    std::vector<cx<T>> store(size);
    std::vector<cx<T>> y(plane);
    forward_angle<T>(t, screen, a, store.data(), y.data(), sc);
    adjoint_angle<T>(t, screen, a, store.data(), r + a * plane, contribution, sc);
  });
}

struct SolveOptions {
  int iterations;
  int batch;
  float step;
  float tv_weight;
  int tv_inner;
};

// src/cpu/bpm.cpp:191-243. x enters as the contrast ri - n_medium and leaves as the result's
// contrast; loss receives one value per iteration.
template <class T>
void bpm_reconstruct(const BpmTables<T>& t, const cx<T>* fields, T* x_io, const SolveOptions& o,
                     double* loss_out) {
  const int count = t.count;
  const int batch = o.batch <= 0 || o.batch >= count ? count : o.batch;
  const T alpha = static_cast<T>(o.step) / t.lipschitz;
  const bool tv = o.tv_weight > 0;
  const std::size_t size = static_cast<std::size_t>(t.nz) * t.n * t.n;
  std::vector<T> x(x_io, x_io + size);
  for (T& v : x) v = std::max(v, T(0));
  std::vector<T> grad(size), z(size);
  Dual<T> dual(tv ? size : 0);
  const auto prox_step = [&](const std::vector<T>& point, std::vector<T>& out) {
    for (std::size_t i = 0; i < size; ++i) z[i] = point[i] - alpha * grad[i];
    if (tv) {
      tv_prox<T>(z.data(), out.data(), dual, t.nz, t.n, t.n, alpha * static_cast<T>(o.tv_weight),
                 o.tv_inner);
    } else {
      for (std::size_t i = 0; i < size; ++i) out[i] = std::max(z[i], T(0));
    }
  };
  if (batch == count) {
    std::vector<T> y = x;
    std::vector<T> next(size);
    T tk = 1;
    for (int it = 0; it < o.iterations; ++it) {
      loss_out[it] = gradient<T>(t, y.data(), fields, 0, count, grad.data());
      prox_step(y, next);
      const T tn = fista_t<T>(tk);
      const T m = (tk - T(1)) / tn;
      for (std::size_t i = 0; i < size; ++i) {
        y[i] = next[i] + m * (next[i] - x[i]);
        x[i] = next[i];
      }
      tk = tn;
    }
  } else {
    for (int it = 0; it < o.iterations; ++it) {
      double loss = 0;
      for (int first = 0; first < count; first += batch) {
        loss += gradient<T>(t, x.data(), fields, first, std::min(first + batch, count), grad.data());
        prox_step(x, x);
      }
      loss_out[it] = loss;
    }
  }
  std::copy(x.begin(), x.end(), x_io);
}

thread_local std::string g_bpm_error;

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_bpm_error = e.what();
    return 1;
  }
}

template <class T>
cx<T>* as_cx(T* p) { return reinterpret_cast<cx<T>*>(p); }
template <class T>
const cx<T>* as_cx(const T* p) { return reinterpret_cast<const cx<T>*>(p); }

}  // namespace

}  // namespace oqh

// C interface. Volumes are contrasts x = ri - n_medium (nz, n, n); fields and residuals are
// interleaved complex (count, n, n). geom: wavelength, n_medium, na, pixel.
extern "C" {

const char* oqh_bpm_last_error() { return oqh::g_bpm_error.c_str(); }

// Tables as the reference stores them: step, last (interleaved complex) and q, each
// (count, n, n); scalars[0] = phase, scalars[1] = Lipschitz constant.
int oqh_bpm_tables_f(int n, int nz, const float* illum, int count, const float* geom, int obliquity,
                     float* step, float* last, float* q, float* scalars) {
  return oqh::guarded([&] {
    const auto t = oqh::bpm_tables<float>(n, nz, illum, count, geom, obliquity != 0);
    std::memcpy(step, t.step.data(), t.step.size() * sizeof(t.step[0]));
    std::memcpy(last, t.last.data(), t.last.size() * sizeof(t.last[0]));
    std::memcpy(q, t.q.data(), t.q.size() * sizeof(float));
    scalars[0] = t.phase;
    scalars[1] = t.lipschitz;
  });
}

#define OQH_BPM_ENTRIES(T, S)                                                                    \
  int oqh_bpm_forward_##S(const T* x, int nz, int n, const float* illum, int count,              \
                          const float* geom, int obliquity, T* out) {                            \
    return oqh::guarded([&] {                                                                    \
      const auto t = oqh::bpm_tables<T>(n, nz, illum, count, geom, obliquity != 0);              \
      oqh::bpm_forward<T>(t, x, oqh::as_cx(out));                                                \
    });                                                                                          \
  }                                                                                              \
  int oqh_bpm_jvp_##S(const T* x, const T* h, int nz, int n, const float* illum, int count,      \
                      const float* geom, int obliquity, T* out) {                                \
    return oqh::guarded([&] {                                                                    \
      const auto t = oqh::bpm_tables<T>(n, nz, illum, count, geom, obliquity != 0);              \
      oqh::bpm_jvp<T>(t, x, h, oqh::as_cx(out));                                                 \
    });                                                                                          \
  }                                                                                              \
  int oqh_bpm_vjp_##S(const T* x, const T* r, int nz, int n, const float* illum, int count,      \
                      const float* geom, int obliquity, T* grad) {                               \
    return oqh::guarded([&] {                                                                    \
      const auto t = oqh::bpm_tables<T>(n, nz, illum, count, geom, obliquity != 0);              \
      oqh::bpm_vjp<T>(t, x, oqh::as_cx(r), grad);                                                \
    });                                                                                          \
  }                                                                                              \
  int oqh_bpm_gradient_##S(const T* x, const T* fields, int nz, int n, const float* illum,       \
                           int count, const float* geom, int obliquity, int first, int last,     \
                           T* grad, double* loss) {                                              \
    return oqh::guarded([&] {                                                                    \
      if (first < 0 || last > count || first >= last) throw std::invalid_argument("bad range"); \
      const auto t = oqh::bpm_tables<T>(n, nz, illum, count, geom, obliquity != 0);              \
      *loss = oqh::gradient<T>(t, x, oqh::as_cx(fields), first, last, grad);                     \
    });                                                                                          \
  }                                                                                              \
  int oqh_bpm_reconstruct_##S(const T* fields, int count, int n, const float* illum,             \
                              const float* geom, T* x, int nz, const int* opts, float step,      \
                              float tv_weight, double* loss) {                                   \
    return oqh::guarded([&] {                                                                    \
      const auto t = oqh::bpm_tables<T>(n, nz, illum, count, geom, opts[3] != 0);                \
      const oqh::SolveOptions o{opts[0], opts[1], step, tv_weight, opts[2]};                     \
      oqh::bpm_reconstruct<T>(t, oqh::as_cx(fields), x, o, loss);                                \
    });                                                                                          \
  }

OQH_BPM_ENTRIES(float, f)
OQH_BPM_ENTRIES(double, d)

#undef OQH_BPM_ENTRIES

}  // extern "C"
