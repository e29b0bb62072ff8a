// Test infrastructure: CPU oracle only. Nothing in the product links this file.
//
// Mixed-radix complex FFT for the CPU oracle. The reference's CPU backend calls pocketfft
// (src/cpu/fft.cpp:18-66, pocketfft commit 4d865cba, fetched at configure time and absent here),
// so the oracle carries its own transform: Stockham autosort passes of radix 4, 2, 3 and 5 with
// twiddles evaluated in double and rounded once to T, a direct DFT for any other prime factor,
// and real transforms of even length through a half-length complex transform. Conventions are
// pocketfft's: unnormalised, forward sign -1, `scale` multiplies the result.
#pragma once

#include <cmath>
#include <complex>
#include <cstddef>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <unordered_map>
#include <vector>

namespace oqh {

template <class T>
using cx = std::complex<T>;

template <class T>
class Fft1 {
 public:
  explicit Fft1(int n) : n_(n) {
    if (n < 1) throw std::invalid_argument("fft length must be positive");
    int m = n;
    while (m % 4 == 0) { fac_.push_back(4); m /= 4; }
    while (m % 2 == 0) { fac_.push_back(2); m /= 2; }
    for (int p = 3; p * p <= m; p += 2) {
      while (m % p == 0) { fac_.push_back(p); m /= p; }
    }
    if (m > 1) fac_.push_back(m);
    tw_.resize(n);
    for (int t = 0; t < n; ++t) {
      // exp(-2 pi i t / n) in double; octant symmetry keeps the argument small.
      const double a = 2.0 * M_PI * static_cast<double>(t) / static_cast<double>(n);
      tw_[t] = {static_cast<T>(std::cos(a)), static_cast<T>(-std::sin(a))};
    }
  }

  int size() const { return n_; }

  // In place over `data` (n contiguous values); `work` holds n values.
  void run(cx<T>* data, cx<T>* work, bool forward) const {
    cx<T>* in = data;
    cx<T>* out = work;
    int ns = 1;
    for (int r : fac_) {
      pass(in, out, r, ns, forward);
      ns *= r;
      std::swap(in, out);
    }
    if (in != data) {
      for (int i = 0; i < n_; ++i) data[i] = in[i];
    }
  }

 private:
  // Twiddle exp(sign 2 pi i t / n).
  cx<T> w(long t, bool forward) const {
    const cx<T> v = tw_[static_cast<std::size_t>(t % n_)];
    return forward ? v : std::conj(v);
  }

  // One Stockham pass of radix r over sub-transforms of length ns: reads stride n / r,
  // writes expanded, output in natural order after the last pass.
  void pass(const cx<T>* in, cx<T>* out, int r, int ns, bool fwd) const {
    const int q = n_ / r;
    const long step = n_ / (static_cast<long>(ns) * r);
    cx<T> v[64];
    std::vector<cx<T>> big;
    cx<T>* buf = v;
    if (r > 64) {
      big.resize(2 * static_cast<std::size_t>(r));
      buf = big.data();
    }
    for (int j = 0; j < q; ++j) {
      const int k = j % ns;
      for (int s = 0; s < r; ++s) {
        cx<T> x = in[j + s * q];
        if (k != 0 && s != 0) x = x * w(static_cast<long>(k) * s * step, fwd);
        buf[s] = x;
      }
      butterfly(buf, r, fwd);
      const int base = (j / ns) * ns * r + k;
      for (int s = 0; s < r; ++s) out[base + s * ns] = buf[s];
    }
  }

  static cx<T> mul_i(cx<T> a, bool forward) {
    // forward: -i a; inverse: +i a
    return forward ? cx<T>(a.imag(), -a.real()) : cx<T>(-a.imag(), a.real());
  }

  void butterfly(cx<T>* v, int r, bool fwd) const {
    switch (r) {
      case 2: {
        const cx<T> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
        return;
      }
      case 4: {
        const cx<T> s02 = v[0] + v[2], d02 = v[0] - v[2];
        const cx<T> s13 = v[1] + v[3], d13 = mul_i(v[1] - v[3], fwd);
        v[0] = s02 + s13;
        v[1] = d02 + d13;
        v[2] = s02 - s13;
        v[3] = d02 - d13;
        return;
      }
      case 3: {
        const T c = static_cast<T>(-0.5);
        const T s = static_cast<T>(0.86602540378443864676);  // sqrt(3) / 2
        const cx<T> t1 = v[1] + v[2];
        const cx<T> t2 = v[0] + c * t1;
        const cx<T> t3 = mul_i(s * (v[1] - v[2]), fwd);
        v[0] = v[0] + t1;
        v[1] = t2 + t3;
        v[2] = t2 - t3;
        return;
      }
      case 5: {
        const T c1 = static_cast<T>(0.30901699437494742410);   // cos(2 pi / 5)
        const T c2 = static_cast<T>(-0.80901699437494742410);  // cos(4 pi / 5)
        const T s1 = static_cast<T>(0.95105651629515357212);   // sin(2 pi / 5)
        const T s2 = static_cast<T>(0.58778525229247312917);   // sin(4 pi / 5)
        const cx<T> a14 = v[1] + v[4], d14 = v[1] - v[4];
        const cx<T> a23 = v[2] + v[3], d23 = v[2] - v[3];
        const cx<T> r1 = v[0] + c1 * a14 + c2 * a23;
        const cx<T> r2 = v[0] + c2 * a14 + c1 * a23;
        const cx<T> i1 = mul_i(s1 * d14 + s2 * d23, fwd);
        const cx<T> i2 = mul_i(s2 * d14 - s1 * d23, fwd);
        v[0] = v[0] + a14 + a23;
        v[1] = r1 + i1;
        v[4] = r1 - i1;
        v[2] = r2 + i2;
        v[3] = r2 - i2;
        return;
      }
      default: {
        // Direct DFT for a prime factor without a dedicated butterfly.
        cx<T>* o = v + r;
        const long stride = n_ / r;
        for (int k = 0; k < r; ++k) {
          cx<T> acc = 0;
          for (int s = 0; s < r; ++s) acc += v[s] * w(static_cast<long>(k) * s % r * stride, fwd);
          o[k] = acc;
        }
        for (int k = 0; k < r; ++k) v[k] = o[k];
        return;
      }
    }
  }

  int n_;
  std::vector<int> fac_;
  std::vector<cx<T>> tw_;
};

// Plans are cached per length; the cache is shared by all threads.
template <class T>
const Fft1<T>& plan(int n) {
  static std::mutex mu;
  static std::unordered_map<int, std::unique_ptr<Fft1<T>>> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto& p = cache[n];
  if (!p) p = std::make_unique<Fft1<T>>(n);
  return *p;
}

// Real-to-half-complex transform of one line of even length n: out[0..n/2].
// z[m] = x[2m] + i x[2m+1], Z = FFT_{n/2}(z),
// X[k] = (Z[k] + conj Z[-k]) / 2 - i/2 exp(-2 pi i k / n) (Z[k] - conj Z[-k]).
template <class T>
void rfft_line(const T* x, long xstride, cx<T>* out, long ostride, int n, cx<T>* work) {
  const int m = n / 2;
  const Fft1<T>& p = plan<T>(m);
  const Fft1<T>& pn = plan<T>(n);
  cx<T>* z = work;
  cx<T>* tmp = work + m;
  for (int j = 0; j < m; ++j) z[j] = {x[(2 * j) * xstride], x[(2 * j + 1) * xstride]};
  p.run(z, tmp, true);
  (void)pn;
  for (int k = 0; k <= m; ++k) {
    const cx<T> zk = z[k % m];
    const cx<T> zm = std::conj(z[(m - k) % m]);
    const cx<T> e = static_cast<T>(0.5) * (zk + zm);
    const cx<T> d = static_cast<T>(0.5) * (zk - zm);
    const double a = 2.0 * M_PI * static_cast<double>(k) / static_cast<double>(n);
    const cx<T> wk(static_cast<T>(std::cos(a)), static_cast<T>(-std::sin(a)));
    // -i * wk * d
    const cx<T> od = wk * d;
    const cx<T> o(od.imag(), -od.real());
    if (k == 0 || k == m) {
      out[k * ostride] = {e.real() + o.real(), T(0)};
    } else {
      out[k * ostride] = e + o;
    }
  }
}

// Half-complex-to-real inverse of one line of even length n (unnormalised, times scale). The
// imaginary parts of bins 0 and n/2 are ignored, as a Hermitian input implies.
// Z[k] = (X[k] + conj X[m-k]) + i exp(2 pi i k / n) (X[k] - conj X[m-k]), z = IFFT_m(Z),
// x[2j] = Re z[j], x[2j+1] = Im z[j].
template <class T>
void irfft_line(const cx<T>* in, long istride, T* x, long xstride, int n, T scale, cx<T>* work) {
  const int m = n / 2;
  const Fft1<T>& p = plan<T>(m);
  cx<T>* z = work;
  cx<T>* tmp = work + m;
  for (int k = 0; k < m; ++k) {
    cx<T> xk = in[k * istride];
    cx<T> xm = in[(m - k) * istride];
    if (k == 0) {
      xk = {xk.real(), T(0)};
      xm = {xm.real(), T(0)};
    }
    const cx<T> s = xk + std::conj(xm);
    const cx<T> d = xk - std::conj(xm);
    const double a = 2.0 * M_PI * static_cast<double>(k) / static_cast<double>(n);
    const cx<T> wk(static_cast<T>(std::cos(a)), static_cast<T>(std::sin(a)));
    const cx<T> wd = wk * d;
    z[k] = s + cx<T>(-wd.imag(), wd.real());
  }
  p.run(z, tmp, false);
  for (int j = 0; j < m; ++j) {
    x[(2 * j) * xstride] = z[j].real() * scale;
    x[(2 * j + 1) * xstride] = z[j].imag() * scale;
  }
}

}  // namespace oqh
