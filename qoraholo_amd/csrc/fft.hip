// Batched FFT kernels on LDS-resident lines and their host launchers.
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "fft.hpp"
#include "fft4s.hpp"
#include "runtime.hpp"

namespace qh {

namespace {

constexpr int kThreads = 256;

std::mutex g_tables_mu;
// (device, kind, n) -> table; kind 0: FFT twiddles, 1: DCT twiddles.
std::map<std::tuple<int, int, int>, DeviceBuffer<float2>> g_tables;

const float2* table(int kind, int n) {
  int dev = 0;
  check(hipGetDevice(&dev), "hipGetDevice");
  std::lock_guard<std::mutex> lock(g_tables_mu);
  auto& buf = g_tables[{dev, kind, n}];
  if (buf.size() == 0) {
    std::vector<float2> h(n);
    for (int t = 0; t < n; ++t) {
      // kind 0: exp(-2 pi i t / n); kind 1: exp(-i pi t / (2 n)). Evaluated in double.
      const double a = kind == 0 ? 2.0 * M_PI * t / static_cast<double>(n)
                                 : M_PI * t / (2.0 * static_cast<double>(n));
      h[t] = {static_cast<float>(std::cos(a)), static_cast<float>(-std::sin(a))};
    }
    DeviceBuffer<float2> d(n);
    check(hipMemcpy(d.get(), h.data(), n * sizeof(float2), hipMemcpyHostToDevice), "upload twiddles");
    buf = std::move(d);
  }
  return buf.get();
}

__device__ __forceinline__ long line_base(long l, const Lines& g) {
  return (l / g.inner) * g.outer + (l % g.inner) * g.inner_stride;
}

struct LinesArgs {
  const float2* in;
  float2* out;
  Lines g;
  int n;
  int lines_per_block;
  FftFactors f;
  const float2* tw;
  int inverse;
  float scale;
  int alt_mod;
};

// CONTIG: lines contiguous in memory (stride 1), LDS line-major, threads walk along lines.
// Otherwise lines are adjacent (inner_stride 1): each row of the block is one coalesced
// segment, LDS is interleaved (line fastest) and threads walk across lines.
template <bool CONTIG>
__global__ void __launch_bounds__(kThreads) k_fft_lines(LinesArgs a) {
  extern __shared__ float2 smem[];
  __shared__ long base[64];
  const int n = a.n;
  float2* tw = smem;
  float2* b0 = smem + n;
  float2* b1 = b0 + a.lines_per_block * n;
  for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = a.tw[i];
  const long l0 = xcd_block() * a.lines_per_block;
  const int nl = static_cast<int>(min(static_cast<long>(a.lines_per_block), a.g.count - l0));
  for (int i = threadIdx.x; i < nl; i += blockDim.x) base[i] = line_base(l0 + i, a.g);
  __syncthreads();
  const int ls = CONTIG ? n : 1;
  const int es = CONTIG ? 1 : nl;
  const int total = nl * n;
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    int line, i;
    if (CONTIG) {
      line = e / n;
      i = e - line * n;
    } else {
      i = e / nl;
      line = e - i * nl;
    }
    b0[line * ls + i * es] = a.in[base[line] + i * a.g.stride];
  }
  __syncthreads();
  const float2* r = lds_fft(b0, b1, ls, es, nl, a.f, tw, a.inverse != 0, !CONTIG);
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    int line, i;
    if (CONTIG) {
      line = e / n;
      i = e - line * n;
    } else {
      i = e / nl;
      line = e - i * nl;
    }
    float2 v = r[line * ls + i * es];
    if (a.scale != 1.0f) v = cscale(v, a.scale);
    if (a.alt_mod > 0 && ((i + (l0 + line) % a.alt_mod) & 1)) v = {-v.x, -v.y};
    a.out[base[line] + i * a.g.stride] = v;
  }
}

struct RowsArgs {
  const float* x;
  const float* x2;
  float* xo;
  long x_pitch;
  float2* c;
  long c_pitch;
  long rows;
  int n;
  int rows_per_block;
  FftFactors f;      // of n / 2
  const float2* tw;  // exp(-2 pi i t / n), t in [0, n)
  RealOp op;
};

// Real rows of even length n through a complex transform of n/2 packed samples
// z[j] = x[2j] + i x[2j+1]; X[k] = (Z[k] + conj Z[-k]) / 2 - i/2 w^k (Z[k] - conj Z[-k]).
__global__ void __launch_bounds__(kThreads) k_rfft_rows(RowsArgs a) {
  extern __shared__ float2 smem[];
  const int n = a.n;
  const int m = n / 2;
  float2* tw = smem;        // n entries
  float2* twm = tw + n;     // m entries, exp(-2 pi i t / m)
  float2* b0 = twm + m;
  float2* b1 = b0 + a.rows_per_block * (m + 1);
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const float2 w = a.tw[i];
    tw[i] = w;
    if ((i & 1) == 0) twm[i >> 1] = w;
  }
  const long r0 = xcd_block() * a.rows_per_block;
  const int nl = static_cast<int>(min(static_cast<long>(a.rows_per_block), a.rows - r0));
  for (int e = threadIdx.x; e < nl * m; e += blockDim.x) {
    const int line = e / m;
    const int j = e - line * m;
    const long at = (r0 + line) * a.x_pitch + 2 * j;
    float u = a.x[at];
    float v = a.x[at + 1];
    if (a.op.in == RealIn::nonneg) {
      u = u < 0.0f ? 0.0f : u;
      v = v < 0.0f ? 0.0f : v;
    } else if (a.op.in == RealIn::momentum) {
      const float pu = a.x2[at];
      const float pv = a.x2[at + 1];
      u = u + a.op.a * (u - pu);
      v = v + a.op.a * (v - pv);
    }
    b0[line * m + j] = {u, v};
  }
  __syncthreads();
  const float2* z = lds_fft(b0, b1, m, 1, nl, a.f, twm, false, false);
  for (int e = threadIdx.x; e < nl * (m + 1); e += blockDim.x) {
    const int line = e / (m + 1);
    const int k = e - line * (m + 1);
    const float2 zk = z[line * m + (k == m ? 0 : k)];
    const float2 zm = cconj(z[line * m + (k == 0 ? 0 : m - k)]);
    const float2 ev = {0.5f * (zk.x + zm.x), 0.5f * (zk.y + zm.y)};
    const float2 d = {0.5f * (zk.x - zm.x), 0.5f * (zk.y - zm.y)};
    const float2 od = cmul(tw[k], d);
    const float2 o = {od.y, -od.x};
    float2 out;
    if (k == 0 || k == m) {
      out = {ev.x + o.x, 0.0f};
    } else {
      out = cadd(ev, o);
    }
    a.c[(r0 + line) * a.c_pitch + k] = out;
  }
}

// Inverse: Z[k] = (X[k] + conj X[m-k]) + i w^-k (X[k] - conj X[m-k]), z = IFFT_m(Z),
// x[2j] = Re z[j], x[2j+1] = Im z[j]; imaginary parts of X[0] and X[m] are dropped.
__global__ void __launch_bounds__(kThreads) k_irfft_rows(RowsArgs a) {
  extern __shared__ float2 smem[];
  const int n = a.n;
  const int m = n / 2;
  float2* tw = smem;
  float2* twm = tw + n;
  float2* b0 = twm + m;
  float2* b1 = b0 + a.rows_per_block * (m + 1);
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const float2 w = a.tw[i];
    tw[i] = w;
    if ((i & 1) == 0) twm[i >> 1] = w;
  }
  const long r0 = xcd_block() * a.rows_per_block;
  const int nl = static_cast<int>(min(static_cast<long>(a.rows_per_block), a.rows - r0));
  for (int e = threadIdx.x; e < nl * (m + 1); e += blockDim.x) {
    const int line = e / (m + 1);
    const int k = e - line * (m + 1);
    b0[e] = a.c[(r0 + line) * a.c_pitch + k];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < nl * m; e += blockDim.x) {
    const int line = e / m;
    const int k = e - line * m;
    float2 xk = b0[line * (m + 1) + k];
    float2 xm = b0[line * (m + 1) + m - k];
    if (k == 0) {
      xk.y = 0.0f;
      xm.y = 0.0f;
    }
    const float2 s = cadd(xk, cconj(xm));
    const float2 d = csub(xk, cconj(xm));
    const float2 wd = cmul(cconj(tw[k]), d);
    b1[line * m + k] = cadd(s, float2{-wd.y, wd.x});
  }
  __syncthreads();
  const float2* z = lds_fft(b1, b0, m, 1, nl, a.f, twm, true, false);
  for (int e = threadIdx.x; e < nl * m; e += blockDim.x) {
    const int line = e / m;
    const int j = e - line * m;
    const float2 v = z[line * m + j];
    float u0 = v.x * a.op.a;
    float u1 = v.y * a.op.a;
    if (a.op.out == RealOut::index) {
      u0 = index_from_potential(u0, 0.0f, a.op.nm2, a.op.inv_k02);
      u1 = index_from_potential(u1, 0.0f, a.op.nm2, a.op.inv_k02);
    }
    float* px = a.xo + (r0 + line) * a.x_pitch + 2 * j;
    px[0] = u0;
    px[1] = u1;
  }
}

struct ZArgs {
  float2* e;
  const float2* a;
  const std::uint8_t* hits;
  long lines;
  long plane;
  int nz;
  int lines_per_block;
  FftFactors f;
  const float2* tw;
  long vol;  // elements between the volumes of a batch (0: one volume, lines == plane)
};

// First element of z line l: lines past a plane's belong to the next volumes of the batch.
__device__ __forceinline__ long z_line_base(const ZArgs& a, long l) {
  return a.vol == 0 ? l : (l / a.plane) * a.vol + l % a.plane;
}

__global__ void __launch_bounds__(kThreads) k_z_roundtrip_restore(ZArgs a) {
  extern __shared__ float2 smem[];
  const int n = a.nz;
  float2* tw = smem;
  float2* b0 = smem + n;
  float2* b1 = b0 + a.lines_per_block * n;
  for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = a.tw[i];
  const long l0 = xcd_block() * a.lines_per_block;
  const int nl = static_cast<int>(min(static_cast<long>(a.lines_per_block), a.lines - l0));
  const int total = nl * n;
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int i = e / nl;
    const int line = e - i * nl;
    b0[e] = a.e[i * a.plane + z_line_base(a, l0 + line)];
  }
  __syncthreads();
  float2* r = lds_fft(b0, b1, 1, nl, nl, a.f, tw, false, true);
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int i = e / nl;
    const int line = e - i * nl;
    const long g = i * a.plane + z_line_base(a, l0 + line);
    const float2 d = a.a[g];
    float2 v = r[e];
    hermitian_restore(d.x, d.y, a.hits[g], v.x, v.y);
    r[e] = v;
  }
  __syncthreads();
  const float2* r2 = lds_fft(r, r == b0 ? b1 : b0, 1, nl, nl, a.f, tw, true, true);
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int i = e / nl;
    const int line = e - i * nl;
    a.e[i * a.plane + z_line_base(a, l0 + line)] = r2[e];
  }
}

// ---- four-step kernels for lengths 2^a 3^b from 32 to 768 ----

constexpr int floor_pow2(int v) {
  int p = 1;
  while (2 * p <= v) p *= 2;
  return p;
}

// T = max(N1, N2) threads per line (a power of two, or 24). Blocks whose threads walk across
// lines (LINE_FAST) hold at least 16 lines, so that each row of the block is a whole 128-byte
// segment: with 8 lines, two blocks share every cache line and, scheduled on different XCDs,
// fetch it twice. Other blocks hold a power of two of lines, about 256 threads.
template <int N1, int N2, bool LINE_FAST = false>
struct Shape4 {
  static constexpr int N = N1 * N2;
  static constexpr int T = N1 > N2 ? N1 : N2;
  static constexpr int LINES = (LINE_FAST && 256 / T < 16) ? 16 : floor_pow2(256 / T);
  static constexpr int BLOCK = T * LINES;
  static_assert(BLOCK % 64 == 0, "whole waves");
};

template <int N1, int N2, bool LINE_FAST>
__global__ void __launch_bounds__((Shape4<N1, N2, LINE_FAST>::BLOCK)) k_fft4s_lines(LinesArgs a) {
  using S = Shape4<N1, N2, LINE_FAST>;
  __shared__ float2 tw[S::N];
  __shared__ float2 t[t_size<N1, N2, S::LINES, LINE_FAST>()];
  const int tid = static_cast<int>(threadIdx.x);
  if constexpr (LINE_FAST) {
    for (int i = tid; i < S::N; i += S::BLOCK) tw[i] = a.tw[i];
  } else {
    fill_twiddle_matrix<N2, N1, N1>(tw, a.tw, 1, tid, S::BLOCK);
  }
  const int line = LINE_FAST ? tid % S::LINES : tid / S::T;
  const int j = LINE_FAST ? tid / S::LINES : tid % S::T;
  const long l = xcd_block() * S::LINES + line;
  const bool valid = l < a.g.count;
  const long base = valid ? line_base(l, a.g) : 0;
  float2 v[N2];
  float2 w[N1];
#pragma unroll
  for (int n2 = 0; n2 < N2; ++n2) {
    v[n2] = (valid && j < N1) ? a.in[base + static_cast<long>(j + N1 * n2) * a.g.stride]
                              : float2{0.0f, 0.0f};
  }
  __syncthreads();
  fft_4step<N1, N2, S::LINES, LINE_FAST, LINE_FAST ? kTwPlain : kTwByK, N1>(v, w, j, line, t, tw,
                                                                          a.inverse != 0);
  if (valid && j < N2) {
    const int lm = a.alt_mod > 0 ? static_cast<int>(l % a.alt_mod) : 0;
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
      const int k = N2 * k1 + j;
      float2 val = w[k1];
      if (a.scale != 1.0f) val = cscale(val, a.scale);
      if (a.alt_mod > 0 && ((k + lm) & 1)) val = {-val.x, -val.y};
      a.out[base + static_cast<long>(k) * a.g.stride] = val;
    }
  }
}

// z lines of the half spectrum: forward transform, restore, inverse transform. BATCH: the
// lines of several volumes (the line offsets' divisions are left out of single volumes).
template <int N1, int N2, bool BATCH>
__global__ void __launch_bounds__((Shape4<N1, N2, true>::BLOCK)) k_fft4s_zrt(ZArgs a) {
  using S = Shape4<N1, N2, true>;
  __shared__ float2 tw[S::N];
  __shared__ float2 t[t_size<N1, N2, S::LINES, true>()];
  const int tid = static_cast<int>(threadIdx.x);
  for (int i = tid; i < S::N; i += S::BLOCK) tw[i] = a.tw[i];
  const int line = tid % S::LINES;
  const int j = tid / S::LINES;
  const long l = xcd_block() * S::LINES + line;
  const bool valid = l < a.lines;
  const long lb = BATCH ? z_line_base(a, l) : l;
  float2 v[N2];
  float2 w[N1];
#pragma unroll
  for (int n2 = 0; n2 < N2; ++n2) {
    v[n2] = (valid && j < N1) ? a.e[static_cast<long>(j + N1 * n2) * a.plane + lb] : float2{0.0f, 0.0f};
  }
  __syncthreads();
  fft_4step<N1, N2, S::LINES, true>(v, w, j, line, t, tw, false);
  if (valid && j < N2) {
    // Where neither k nor -k was measured (hits 0) the data term is +0 by construction (the
    // cleared potential folded: 0.5 (+0 + +0), 0.5 (+0 - +0)), so it need not be read there:
    // most of the half spectrum lies outside the measured caps. The compiler then holds more of
    // the loads in registers; the lengths whose occupancy that leaves unchanged (256, 384) skip.
    if constexpr (N1 == 16 && (N2 == 16 || N2 == 24)) {
      std::uint8_t h[N1];
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) h[k1] = a.hits[static_cast<long>(N2 * k1 + j) * a.plane + lb];
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) {
        const long g = static_cast<long>(N2 * k1 + j) * a.plane + lb;
        const float2 d = h[k1] != 0 ? a.a[g] : float2{0.0f, 0.0f};
        hermitian_restore(d.x, d.y, h[k1], w[k1].x, w[k1].y);
      }
    } else {
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) {
        const long g = static_cast<long>(N2 * k1 + j) * a.plane + lb;
        const float2 d = a.a[g];
        hermitian_restore(d.x, d.y, a.hits[g], w[k1].x, w[k1].y);
      }
    }
  }
  // Thread j now holds X[j + N2 k1]: the input layout of the transform with the factors swapped.
  float2 u[N2];
  fft_4step<N2, N1, S::LINES, true>(w, u, j, line, t, tw, true);
  if (valid && j < N1) {
#pragma unroll
    for (int k1 = 0; k1 < N2; ++k1) a.e[static_cast<long>(N1 * k1 + j) * a.plane + lb] = u[k1];
  }
}

// Real rows of n = 2 N1 N2 through the packed complex transform, then the split into bins
// 0 .. n/2 (see k_rfft_rows).
template <int N1, int N2>
__global__ void __launch_bounds__((Shape4<N1, N2>::BLOCK)) k_rfft4s_rows(RowsArgs a) {
  using S = Shape4<N1, N2>;
  constexpr int M = S::N;
  __shared__ float2 twm[M];
  __shared__ float2 twn[M + 1];
  __shared__ float2 t[t_size<N1, N2, S::LINES, false>()];
  const int tid = static_cast<int>(threadIdx.x);
  fill_twiddle_matrix<N2, N1, N1>(twm, a.tw, 2, tid, S::BLOCK);  // W_M^t = W_n^(2t)
  for (int i = tid; i <= M; i += S::BLOCK) twn[i] = a.tw[i];
  const int line = tid / S::T;
  const int j = tid % S::T;
  const long row = xcd_block() * S::LINES + line;
  const bool valid = row < a.rows;
  float2 v[N2];
  float2 w[N1];
  // All loads of a thread are issued before the first use (the input operation is uniform over
  // the grid), so their latencies overlap.
  const bool live = valid && j < N1;
  const bool momentum = a.op.in == RealIn::momentum;
  const long at0 = row * a.x_pitch + 2 * j;
  float2 q[N2];
#pragma unroll
  for (int n2 = 0; n2 < N2; ++n2) {
    v[n2] = live ? *reinterpret_cast<const float2*>(a.x + at0 + 2 * N1 * n2) : float2{0.0f, 0.0f};
  }
#pragma unroll
  for (int n2 = 0; n2 < N2; ++n2) {
    q[n2] = (live && momentum) ? *reinterpret_cast<const float2*>(a.x2 + at0 + 2 * N1 * n2)
                               : float2{0.0f, 0.0f};
  }
  if (a.op.in == RealIn::nonneg) {
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) {
      v[n2].x = v[n2].x < 0.0f ? 0.0f : v[n2].x;
      v[n2].y = v[n2].y < 0.0f ? 0.0f : v[n2].y;
    }
  } else if (momentum) {
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) {
      v[n2].x = v[n2].x + a.op.a * (v[n2].x - q[n2].x);
      v[n2].y = v[n2].y + a.op.a * (v[n2].y - q[n2].y);
    }
  }
  __syncthreads();
  fft_4step<N1, N2, S::LINES, false, kTwByK, N1>(v, w, j, line, t, twm, false);
  float2* zrow = t + line * (M + 1);
  if (j < N2) {
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) zrow[N2 * k1 + j] = w[k1];
  }
  __syncthreads();
  if (valid) {
    for (int k = j; k <= M; k += S::T) {
      const float2 zk = zrow[k == M ? 0 : k];
      const float2 zm = cconj(zrow[k == 0 ? 0 : M - k]);
      const float2 ev = {0.5f * (zk.x + zm.x), 0.5f * (zk.y + zm.y)};
      const float2 d = {0.5f * (zk.x - zm.x), 0.5f * (zk.y - zm.y)};
      const float2 od = cmul(twn[k], d);
      const float2 o = {od.y, -od.x};
      a.c[row * a.c_pitch + k] = (k == 0 || k == M) ? float2{ev.x + o.x, 0.0f} : cadd(ev, o);
    }
  }
}

template <int N1, int N2>
__global__ void __launch_bounds__((Shape4<N1, N2>::BLOCK)) k_irfft4s_rows(RowsArgs a) {
  using S = Shape4<N1, N2>;
  constexpr int M = S::N;
  __shared__ float2 twm[M];
  __shared__ float2 twn[M + 1];
  __shared__ float2 t[t_size<N1, N2, S::LINES, false>()];
  const int tid = static_cast<int>(threadIdx.x);
  fill_twiddle_matrix<N2, N1, N1>(twm, a.tw, 2, tid, S::BLOCK);  // W_M^t = W_n^(2t)
  for (int i = tid; i <= M; i += S::BLOCK) twn[i] = a.tw[i];
  const int line = tid / S::T;
  const int j = tid % S::T;
  const long row = xcd_block() * S::LINES + line;
  const bool valid = row < a.rows;
  float2* xrow = t + line * (M + 1);
  {
    // Bins 0 .. M of the row, loaded together before they are staged in LDS.
    constexpr int kPer = (M + S::T) / S::T;
    float2 c[kPer];
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int k = j + i * S::T;
      c[i] = (valid && k <= M) ? a.c[row * a.c_pitch + k] : float2{0.0f, 0.0f};
    }
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int k = j + i * S::T;
      if (k <= M) xrow[k] = c[i];
    }
  }
  __syncthreads();
  float2 v[N2];
  float2 w[N1];
#pragma unroll
  for (int n2 = 0; n2 < N2; ++n2) {
    const int k = j + N1 * n2;
    float2 z{0.0f, 0.0f};
    if (j < N1) {
      float2 xk = xrow[k];
      float2 xm = xrow[M - k];
      if (k == 0) {
        xk.y = 0.0f;
        xm.y = 0.0f;
      }
      const float2 s = cadd(xk, cconj(xm));
      const float2 d = csub(xk, cconj(xm));
      const float2 wd = cmul(cconj(twn[k]), d);
      z = cadd(s, float2{-wd.y, wd.x});
    }
    v[n2] = z;
  }
  __syncthreads();
  fft_4step<N1, N2, S::LINES, false, kTwByK, N1>(v, w, j, line, t, twm, true);
  if (valid && j < N2) {
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
      const int m = N2 * k1 + j;
      float u0 = w[k1].x * a.op.a;
      float u1 = w[k1].y * a.op.a;
      if (a.op.out == RealOut::index) {
        u0 = index_from_potential(u0, 0.0f, a.op.nm2, a.op.inv_k02);
        u1 = index_from_potential(u1, 0.0f, a.op.nm2, a.op.inv_k02);
      }
      *reinterpret_cast<float2*>(a.xo + row * a.x_pitch + 2 * m) = float2{u0, u1};
    }
  }
}

// Runs F<N1, N2> for the lengths with a four-step kernel; false for the others.
template <template <int, int> class F, class... Args>
bool with_four_step(int n, Args&&... args) {
  switch (n) {
    case 32: F<8, 4>::run(args...); return true;
    case 48: F<8, 6>::run(args...); return true;
    case 64: F<8, 8>::run(args...); return true;
    case 96: F<16, 6>::run(args...); return true;
    case 128: F<16, 8>::run(args...); return true;
    case 192: F<16, 12>::run(args...); return true;
    case 256: F<16, 16>::run(args...); return true;
    case 384: F<16, 24>::run(args...); return true;
    case 512: F<32, 16>::run(args...); return true;
    case 768: F<32, 24>::run(args...); return true;
    default: return false;
  }
}

template <int N1, int N2>
struct LaunchLines {
  static void run(const LinesArgs& a, bool contig, hipStream_t s) {
    if (contig) {
      using S = Shape4<N1, N2, false>;
      const unsigned grid = static_cast<unsigned>((a.g.count + S::LINES - 1) / S::LINES);
      hipLaunchKernelGGL((k_fft4s_lines<N1, N2, false>), dim3(grid), dim3(S::BLOCK), 0, s, a);
    } else {
      using S = Shape4<N1, N2, true>;
      const unsigned grid = static_cast<unsigned>((a.g.count + S::LINES - 1) / S::LINES);
      hipLaunchKernelGGL((k_fft4s_lines<N1, N2, true>), dim3(grid), dim3(S::BLOCK), 0, s, a);
    }
  }
};

template <int N1, int N2>
struct LaunchZrt {
  static void run(const ZArgs& z, hipStream_t s) {
    using S = Shape4<N1, N2, true>;
    const unsigned grid = static_cast<unsigned>((z.lines + S::LINES - 1) / S::LINES);
    if (z.vol != 0) {
      hipLaunchKernelGGL((k_fft4s_zrt<N1, N2, true>), dim3(grid), dim3(S::BLOCK), 0, s, z);
    } else {
      hipLaunchKernelGGL((k_fft4s_zrt<N1, N2, false>), dim3(grid), dim3(S::BLOCK), 0, s, z);
    }
  }
};

template <int N1, int N2>
struct LaunchRows {
  static void run(const RowsArgs& a, bool inverse, hipStream_t s) {
    const unsigned grid = static_cast<unsigned>((a.rows + Shape4<N1, N2>::LINES - 1) / Shape4<N1, N2>::LINES);
    if (inverse) {
      hipLaunchKernelGGL((k_irfft4s_rows<N1, N2>), dim3(grid), dim3(Shape4<N1, N2>::BLOCK), 0, s, a);
    } else {
      hipLaunchKernelGGL((k_rfft4s_rows<N1, N2>), dim3(grid), dim3(Shape4<N1, N2>::BLOCK), 0, s, a);
    }
  }
};

int contig_lines_per_block(int n) { return std::max(1, std::min(64, 2048 / n)); }
// Lines per block keep the two LDS buffers within 32 KiB (dynamic LDS stays below 64 KiB).
int strided_lines_per_block(int n) { return std::max(1, std::min(16, 2048 / n)); }

}  // namespace

FftFactors fft_factors(int n) {
  require(n >= 1, "fft: length must be positive");
  FftFactors f;
  f.n = n;
  int m = n;
  const auto push = [&](int r) {
    require(f.count < kMaxFactors, "fft: too many factors");
    f.r[f.count++] = r;
    m /= r;
  };
  while (m % 8 == 0) push(8);
  while (m % 4 == 0) push(4);
  while (m % 2 == 0) push(2);
  while (m % 3 == 0) push(3);
  while (m % 5 == 0) push(5);
  require(m == 1, "fft: lengths must factor into 2, 3 and 5, got " + std::to_string(n));
  return f;
}

bool fft_supported(int n) {
  // One line of up to 2048 complex values fits the LDS budget of the transform kernels.
  if (n < 1 || n > 2048) return false;
  for (int r : {2, 3, 5}) {
    while (n % r == 0) n /= r;
  }
  return n == 1;
}

const float2* fft_twiddles(int n) { return table(0, n); }
const float2* dct_twiddles(int n) { return table(1, n); }

Lines lines_along(int axis, int d0, int d1, int d2, long pitch) {
  switch (axis) {
    case 2: return {static_cast<long>(d0) * d1, d1, d1 * pitch, pitch, 1};
    case 1: return {static_cast<long>(d0) * d2, d2, d1 * pitch, 1, pitch};
    default: return {static_cast<long>(d1) * d2, d2, pitch, 1, d1 * pitch};
  }
}

void fft_lines(const float2* in, float2* out, int n, const Lines& l, bool inverse, float scale,
               int alt_mod, hipStream_t s) {
  if (l.count == 0) return;
  LinesArgs a{in, out, l, n, 0, fft_factors(n), fft_twiddles(n), inverse ? 1 : 0, scale, alt_mod};
  if (l.stride != 1) require(l.inner_stride == 1, "fft: strided lines must be adjacent");
  if (with_four_step<LaunchLines>(n, a, l.stride == 1, s)) {
    check_launch("fft_lines");
    return;
  }
  if (l.stride == 1) {
    a.lines_per_block = contig_lines_per_block(n);
    const std::size_t lds = (n + 2 * static_cast<std::size_t>(a.lines_per_block) * n) * sizeof(float2);
    const unsigned grid = static_cast<unsigned>((l.count + a.lines_per_block - 1) / a.lines_per_block);
    hipLaunchKernelGGL(k_fft_lines<true>, dim3(grid), dim3(kThreads), lds, s, a);
  } else {
    require(l.inner_stride == 1, "fft: strided lines must be adjacent");
    a.lines_per_block = strided_lines_per_block(n);
    const std::size_t lds = (n + 2 * static_cast<std::size_t>(a.lines_per_block) * n) * sizeof(float2);
    const unsigned grid = static_cast<unsigned>((l.count + a.lines_per_block - 1) / a.lines_per_block);
    hipLaunchKernelGGL(k_fft_lines<false>, dim3(grid), dim3(kThreads), lds, s, a);
  }
  check_launch("fft_lines");
}

void rfft_rows(const float* x, const float* x2, long x_pitch, float2* out, long out_pitch,
               long rows, int n, const RealOp& op, hipStream_t s) {
  require(n % 2 == 0, "rfft: even length required");
  if (rows == 0) return;
  const int m = n / 2;
  RowsArgs a{x, x2, nullptr, x_pitch, out, out_pitch, rows, n, contig_lines_per_block(m),
             fft_factors(m), fft_twiddles(n), op};
  if (x_pitch % 2 == 0 && with_four_step<LaunchRows>(m, a, false, s)) {
    check_launch("rfft_rows");
    return;
  }
  const std::size_t lds =
      (n + m + static_cast<std::size_t>(a.rows_per_block) * (2 * m + 1)) * sizeof(float2);
  const unsigned grid = static_cast<unsigned>((rows + a.rows_per_block - 1) / a.rows_per_block);
  hipLaunchKernelGGL(k_rfft_rows, dim3(grid), dim3(kThreads), lds, s, a);
  check_launch("rfft_rows");
}

void irfft_rows(const float2* in, long in_pitch, float* x, long x_pitch, long rows, int n,
                const RealOp& op, hipStream_t s) {
  require(n % 2 == 0, "irfft: even length required");
  if (rows == 0) return;
  const int m = n / 2;
  RowsArgs a{nullptr, nullptr, x, x_pitch, const_cast<float2*>(in), in_pitch, rows, n,
             contig_lines_per_block(m), fft_factors(m), fft_twiddles(n), op};
  if (x_pitch % 2 == 0 && with_four_step<LaunchRows>(m, a, true, s)) {
    check_launch("irfft_rows");
    return;
  }
  const std::size_t lds =
      (n + m + static_cast<std::size_t>(a.rows_per_block) * (2 * m + 1)) * sizeof(float2);
  const unsigned grid = static_cast<unsigned>((rows + a.rows_per_block - 1) / a.rows_per_block);
  hipLaunchKernelGGL(k_irfft_rows, dim3(grid), dim3(kThreads), lds, s, a);
  check_launch("irfft_rows");
}

void z_roundtrip_restore(float2* e, const float2* a, const std::uint8_t* hits, int nz, int ny,
                         int h, hipStream_t s, int batch) {
  require(batch >= 1, "z_roundtrip_restore: batch must be positive");
  const long plane = static_cast<long>(ny) * h;
  ZArgs z{e, a, hits, plane * batch, plane, nz, strided_lines_per_block(nz), fft_factors(nz),
          fft_twiddles(nz), batch > 1 ? plane * nz : 0};
  if (with_four_step<LaunchZrt>(nz, z, s)) {
    check_launch("z_roundtrip_restore");
    return;
  }
  const std::size_t lds = (nz + 2 * static_cast<std::size_t>(z.lines_per_block) * nz) * sizeof(float2);
  const unsigned grid = static_cast<unsigned>((z.lines + z.lines_per_block - 1) / z.lines_per_block);
  hipLaunchKernelGGL(k_z_roundtrip_restore, dim3(grid), dim3(kThreads), lds, s, z);
  check_launch("z_roundtrip_restore");
}

void rfft3(const float* x, float2* out, int nz, int ny, int nx, hipStream_t s) {
  const int h = nx / 2 + 1;
  rfft_rows(x, nullptr, nx, out, h, static_cast<long>(nz) * ny, nx, RealOp{}, s);
  fft_lines(out, out, ny, lines_along(1, nz, ny, h, h), false, 1.0f, 0, s);
  fft_lines(out, out, nz, lines_along(0, nz, ny, h, h), false, 1.0f, 0, s);
}

void irfft3(float2* in, float* x, int nz, int ny, int nx, float scale, hipStream_t s) {
  const int h = nx / 2 + 1;
  fft_lines(in, in, nz, lines_along(0, nz, ny, h, h), true, 1.0f, 0, s);
  fft_lines(in, in, ny, lines_along(1, nz, ny, h, h), true, 1.0f, 0, s);
  RealOp op;
  op.a = scale;
  irfft_rows(in, h, x, nx, static_cast<long>(nz) * ny, nx, op, s);
}

void fft3(float2* data, int nz, int ny, int nx, unsigned axes, bool inverse, float scale,
          hipStream_t s) {
  const int dims[3] = {nz, ny, nx};
  int first = 3;
  for (int ax = 0; ax < 3; ++ax) {
    if (axes & (1u << ax)) first = std::min(first, ax);
  }
  // Passes run x, y, z; the scale goes with the last one.
  for (int ax = 2; ax >= 0; --ax) {
    if (!(axes & (1u << ax))) continue;
    const bool final_pass = ax == first;
    fft_lines(data, data, dims[ax], lines_along(ax, nz, ny, nx, nx), inverse,
              final_pass ? scale : 1.0f, 0, s);
  }
}

}  // namespace qh
