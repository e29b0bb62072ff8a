// Beam propagation model on the device (reference src/cpu/bpm.cpp restated for gfx950).
//
// State of one illumination between slices is a spectrum S (n x n, k-space) and, for the next
// slice, T = inverse transform of S along y. Every slice is two fused kernels over the lines
// of all illuminations of a chunk:
//   rows:    T row -> inverse x transform, / n^2 = envelope w -> (store w) -> times the screen
//            exp(i phase x_k) -> forward x transform = U row
//   columns: U column -> forward y transform = A -> S = c (q A + (1 - q) S) -> S stored ->
//            inverse y transform = T column
// so each slice reads and writes every plane twice, with no separate scale, screen or
// propagation pass. The adjoint runs the same way backwards (columns: forward y transform,
// adjoint propagation step, inverse y transform; rows: inverse x transform, gradient term,
// conjugate screen, forward x transform), and its per-illumination gradient terms are added to
// the gradient in illumination order by one small kernel per slice, so the gradient does not
// depend on the chunking and is reproducible bit for bit.
//
// Checkpointing: the forward pass stores (S, T) in front of every `checkpoint`-th slice and the
// envelopes of the last segment; the adjoint recomputes each earlier segment's envelopes from
// its checkpoint just before it needs them, with the same kernels, so they are the same bits.
#include <algorithm>
#include <cmath>
#include <memory>
#include <numbers>

#include "bpm.hpp"
#include "common.hpp"
#include "fft.hpp"
#include "fft_engine.hpp"
#include "tv.hpp"

namespace qh {

// src/bpm.cpp:10-60: every value in double, stored in float.
BpmTables bpm_tables(int n, int nz, const qh_illumination* il, int count, const qh_geometry& g,
                     bool obliquity) {
  constexpr double pi = std::numbers::pi;
  const double lambda = g.wavelength;
  const double fm = g.n_medium / lambda;
  const double fna = g.na / lambda;
  const double d = g.pixel;
  const double df = 1.0 / (n * d);
  const double nuz_min = std::sqrt(std::max(fm * fm - fna * fna, 0.01 * fm * fm));
  const double back = -(nz - 1 - nz / 2) * d;
  BpmTables t;
  t.n = n;
  t.nz = nz;
  t.count = count;
  t.phase = static_cast<float>(2.0 * pi / lambda * d);
  const std::size_t plane = static_cast<std::size_t>(n) * n;
  t.step.assign(plane * count, float2{0.0f, 0.0f});
  t.last.assign(plane * count, float2{0.0f, 0.0f});
  t.q.assign(plane * count, 0.0f);
  double q0 = 0;
  for (int a = 0; a < count; ++a) {
    const double fiy = il[a].fy;
    const double fix = il[a].fx;
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
          const double ps = 2.0 * pi * (nuz - fiz) * d;
          t.step[i] = {static_cast<float>(std::cos(ps)), static_cast<float>(std::sin(ps))};
          if (rho2 <= fna * fna) {
            const double pl = 2.0 * pi * (nuz - fiz) * back;
            t.last[i] = {static_cast<float>(std::cos(pl)), static_cast<float>(std::sin(pl))};
          }
          if (obliquity) q = fm / std::max(nuz, nuz_min);
        }
        t.q[i] = static_cast<float>(q);
      }
    }
    q0 += static_cast<double>(t.q[a * plane]) * t.q[a * plane];
  }
  t.lipschitz = static_cast<float>(nz * static_cast<double>(t.phase) * t.phase * q0 / count);
  return t;
}

struct BpmModel::Args {
  LineGeom g;
  const float2* in;
  float2* out;
  const float2* screen;  // exp(i phase x) of the slice, (n, n)
  float2* w;             // envelopes of the chunk for this slice, or null
  const float2* data;    // measured fields of the chunk (residual), or null
  double* loss;          // per-line data term of the chunk
  const float2* s_in;    // spectrum in front of the slice; null: the unit envelope
  float2* s_out;
  float2* g_state;       // adjoint state, (chunk, n, n)
  float* p;              // per-illumination gradient terms, (chunk, n, n)
  const float2* c;       // propagator of the slice from the chunk's first illumination
  const float* q;
  float inv;             // 1 / n^2
  float phase;
  int first;             // adjoint: the first slice processed (no state yet)
  int s_zero;            // columns with s_in == null: the state in front of slice 0 is zero
                         // (the tangent), not the unit envelope
  const float* h;        // tangent direction of the slice, (n, n)
  // Adjoint columns also sum the gradient terms of the slice processed before (red_p, left by
  // the previous rows launch) into red_grad, in blocks after the line blocks; red_init: the
  // first chunk, whose sum starts from zero rather than from red_grad.
  const float* red_p;
  float* red_grad;
  int red_chunk;
  int red_blocks;
  int red_init;
  int line_blocks;  // with red_blocks: the grid's first line_blocks blocks run the lines
};

namespace {

using Args = BpmModel::Args;

// Side of the lines of engine E: a compile-time constant for the register engines, so that the
// line arithmetic below compiles to multiplications; the LDS engine's lines have any length.
template <class E>
__device__ __forceinline__ int line_side(int n) {
  if constexpr (requires { E::N; }) {
    return E::N;
  } else {
    return n;
  }
}

// Row (or column) r of line l = b n + r within its illumination's plane, in 32-bit arithmetic
// (a chunk holds far fewer than 2^31 lines; 64-bit remainders by a runtime n had been expanded
// once per element).
template <class E>
__device__ __forceinline__ int line_in_plane(long l, int n) {
  return static_cast<int>(static_cast<unsigned>(l) % static_cast<unsigned>(line_side<E>(n)));
}

// Envelopes are stored for the adjoint, which reads them back only after the rest of their
// segment has been recomputed and every other plane of it has passed through the caches:
// written non-temporally they displace nothing that the next kernels read (384/81 gradient
// 189 -> 184 ms, 256/49 30.8 -> 30.4 ms; the same hint on the propagation state, which the
// next slice reads, was slower at 256, where that state stays in the MALL).
__device__ __forceinline__ void store_envelope(float2* p, float2 v) {
  typedef float vec2 __attribute__((ext_vector_type(2)));
  __builtin_nontemporal_store(vec2{v.x, v.y}, reinterpret_cast<vec2*>(p));
}

// Bins whose operands the slice kernels load together before using any of them (mid_loaded):
// 8 for lines of 384 and more (384/81 gradient 210 -> 191 ms; its column blocks hold one block
// per CU either way), 4 for shorter lines, where 8 would take the 256-point column kernels past
// 128 registers and one of their two blocks per CU (256/49 gradient 32.4 -> 30.7 ms with 4, 32.8
// with 8). Same results bit for bit.
template <class E>
constexpr int mid_group() {
  if constexpr (requires { E::N; }) {
    return E::N >= 384 ? 8 : 4;
  } else {
    return 1;
  }
}

// Rows (line l = b n + y, element i at l n + i): inverse x transform of T, envelope, screen,
// forward x transform. in == null: the unit envelope of the first slice. out == null: store the
// envelope only.
template <class E>
__global__ void __launch_bounds__(E::kBlock) k_bpm_rows_fwd(Args a) {
  E e(a.g);
  const int n = a.g.n;
  if (a.in == nullptr) {
    e.load([&](long l, int i) {
      if (a.w) a.w[l * n + i] = float2{1.0f, 0.0f};
      return a.screen[line_in_plane<E>(l, n) * n + i];
    });
    if (a.out == nullptr) return;
    e.first(false);
    e.store1([&](long l, int k, float2 v) { a.out[l * n + k] = v; });
    return;
  }
  e.load([&](long l, int i) { return a.in[l * n + i]; });
  e.first(true);
  e.template mid_loaded<mid_group<E>()>(
      [&](long l, int k) { return a.screen[line_in_plane<E>(l, n) * n + k]; },
      [&](long l, int k, float2& v, float2 t) {
        const float2 w = cscale(v, a.inv);
        if (a.w) store_envelope(a.w + l * n + k, w);
        v = cmul(t, w);
      });
  if (a.out == nullptr) return;
  e.second(false);
  e.store2([&](long l, int m, float2 v) { a.out[l * n + m] = v; });
}

// Tangent rows (bpm_jvp, src/cpu/bpm.cpp:141-176): inverse x transform of the tangent dT,
// dw = dT / n^2, dv = t (dw + i phase h w) with the envelope w the primal rows stored for this
// slice, forward x transform. in == null: dw = 0 (the first slice).
template <class E>
__global__ void __launch_bounds__(E::kBlock) k_bpm_rows_jvp(Args a) {
  E e(a.g);
  const int n = a.g.n;
  // The complex operations of the reference, in its order: (0, phase h) w, + dw, t times that.
  const auto tangent = [&](long l, int k, float2 dw) {
    const float2 w = a.w[l * n + k];
    const float p = a.phase * a.h[line_in_plane<E>(l, n) * n + k];
    const float2 d{dw.x + (0.0f * w.x - p * w.y), dw.y + (0.0f * w.y + p * w.x)};
    return cmul(a.screen[line_in_plane<E>(l, n) * n + k], d);
  };
  if (a.in == nullptr) {
    e.load([&](long l, int i) { return tangent(l, i, float2{0.0f, 0.0f}); });
    e.first(false);
    e.store1([&](long l, int k, float2 v) { a.out[l * n + k] = v; });
    return;
  }
  e.load([&](long l, int i) { return a.in[l * n + i]; });
  e.first(true);
  e.mid([&](long l, int k, float2& v) { v = tangent(l, k, cscale(v, a.inv)); });
  e.second(false);
  e.store2([&](long l, int m, float2 v) { a.out[l * n + m] = v; });
}

// Column lines: element i of line l = b n + c at b n^2 + i n + c. The division takes the runtime
// n: with the compile-time side the 384-point column kernel ran 10% slower (measured).
template <class E>
__device__ __forceinline__ long col_at(long l, int i, int n, long plane) {
  const unsigned m = static_cast<unsigned>(n);
  const unsigned lu = static_cast<unsigned>(l);
  const unsigned b = lu / m;
  return static_cast<long>(b) * plane + static_cast<long>(i) * m + (lu - b * m);
}

// Columns (line l = b n + kx, element i at b n^2 + i n + kx): forward y transform, propagation
// step S = c (q A + (1 - q) S) (src/common.hpp:158-163), inverse y transform.
template <class E>
__global__ void __launch_bounds__(E::kBlock) k_bpm_cols_fwd(Args a) {
  E e(a.g);
  const int n = a.g.n;
  const long plane = static_cast<long>(n) * n;
  e.load([&](long l, int i) { return a.in[col_at<E>(l, i, n, plane)]; });
  e.first(false);
  // Operands of a bin: the state S in front of the slice, the propagator c and the weight q.
  struct Ops {
    float2 s, c;
    float q;
  };
  e.template mid_loaded<mid_group<E>()>(
      [&](long l, int ky) {
        const long x = col_at<E>(l, ky, n, plane);
        Ops o;
        if (a.s_in) {
          o.s = a.s_in[x];
        } else if (a.s_zero) {
          o.s = float2{0.0f, 0.0f};
        } else {
          o.s = (ky == 0 && line_in_plane<E>(l, n) == 0) ? float2{static_cast<float>(plane), 0.0f}
                                                          : float2{0.0f, 0.0f};
        }
        o.c = a.c[x];
        o.q = a.q[x];
        return o;
      },
      [&](long l, int ky, float2& v, Ops o) {
        bpm_step(o.c.x, o.c.y, o.q, v.x, v.y, o.s.x, o.s.y);
        a.s_out[col_at<E>(l, ky, n, plane)] = o.s;
        v = o.s;
      });
  if (a.out == nullptr) return;
  e.second(true);
  e.store2([&](long l, int m, float2 v) { a.out[col_at<E>(l, m, n, plane)] = v; });
}

// Rows after the last slice: y = inverse x transform / n^2. Without data, y is the output;
// with data, r = y - d, the per-row data term, and the forward x transform of r, the first
// half of the transform that starts the adjoint.
template <class E>
__global__ void __launch_bounds__(E::kBlock) k_bpm_rows_out(Args a) {
  E e(a.g);
  const int n = a.g.n;
  e.load([&](long l, int i) { return a.in[l * n + i]; });
  e.first(true);
  if (a.data == nullptr) {
    e.store1([&](long l, int k, float2 v) { a.out[l * n + k] = cscale(v, a.inv); });
    return;
  }
  e.mid([&](long l, int k, float2& v) {
    const float2 y = cscale(v, a.inv);
    const float2 d = a.data[l * n + k];
    v = float2{y.x - d.x, y.y - d.y};
  });
  e.reduce_lines(
      [&](long, int, const float2& v) {
        const double re = v.x;
        const double im = v.y;
        return re * re + im * im;
      },
      a.loss);
  e.second(false);
  e.store2([&](long l, int m, float2 v) { a.out[l * n + m] = v; });
}

// grad[i] += p[b][i] for b = 0 .. chunk-1 in order (from 0 instead of grad[i] with init, the
// same bits as adding to a zeroed grad); the loads of eight illuminations are issued together
// so that each thread keeps several in flight.
__device__ __forceinline__ void reduce_terms(const float* p, int chunk, long plane, float* grad,
                                             bool init, long i) {
  float g = init ? 0.0f : grad[i];
  int b = 0;
  for (; b + 8 <= chunk; b += 8) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = p[(b + u) * plane + i];
#pragma unroll
    for (int u = 0; u < 8; ++u) g += v[u];
  }
  for (; b < chunk; ++b) g += p[b * plane + i];
  grad[i] = g;
}

// Adjoint columns: forward y transform completes back = FFT2(conj(t) b) of the slice above (or
// FFT2(r) on the first slice processed); with the carried g, U = conj(c) (g + back), g <- (1-q) U,
// b = q U (src/common.hpp:168-176); inverse y transform of b.
template <class E>
__global__ void __launch_bounds__(E::kBlock) k_bpm_cols_adj(Args a) {
  const int nb = a.red_blocks > 0 ? a.line_blocks : static_cast<int>(gridDim.x);
  if (static_cast<int>(blockIdx.x) >= nb) {
    // A reduction block (the last red_blocks of the grid, dealt to the XCDs in turn after the
    // line blocks): the previous slice's gradient terms, one plane element per thread.
    const long plane = static_cast<long>(a.g.n) * a.g.n;
    const long i = static_cast<long>(static_cast<int>(blockIdx.x) - nb) * E::kBlock + threadIdx.x;
    if (i < plane) reduce_terms(a.red_p, a.red_chunk, plane, a.red_grad, a.red_init != 0, i);
    return;
  }
  E e(a.g, nb);
  const int n = a.g.n;
  const long plane = static_cast<long>(n) * n;
  e.load([&](long l, int i) { return a.in[col_at<E>(l, i, n, plane)]; });
  e.first(false);
  // Operands of a bin: the carried adjoint state g, the propagator c and the weight q.
  struct Ops {
    float2 g, c;
    float q;
  };
  e.template mid_loaded<mid_group<E>()>(
      [&](long l, int ky) {
        const long x = col_at<E>(l, ky, n, plane);
        Ops o;
        o.g = a.first ? float2{0.0f, 0.0f} : a.g_state[x];
        o.c = a.c[x];
        o.q = a.q[x];
        return o;
      },
      [&](long l, int ky, float2& v, Ops o) {
        float gre = v.x;
        float gim = v.y;
        if (!a.first) {
          gre = o.g.x + v.x;
          gim = o.g.y + v.y;
        }
        float are = 0.0f;
        float aim = 0.0f;
        bpm_step_adjoint(o.c.x, o.c.y, o.q, gre, gim, are, aim);
        a.g_state[col_at<E>(l, ky, n, plane)] = float2{gre, gim};
        v = float2{are, aim};
      });
  e.second(true);
  e.store2([&](long l, int m, float2 v) { a.out[col_at<E>(l, m, n, plane)] = v; });
}

// Adjoint rows: b = inverse x transform / n^2; gradient term phase Im(conj(t w) b) per
// illumination; back = conj(t) b and its forward x transform (skipped on slice 0). Four waves
// per SIMD: the 384-point kernel otherwise takes 129 registers, one over the limit for four.
template <class E>
__global__ void __launch_bounds__(E::kBlock) __attribute__((amdgpu_waves_per_eu(4)))
k_bpm_rows_adj(Args a) {
  E e(a.g);
  const int n = a.g.n;
  e.load([&](long l, int i) { return a.in[l * n + i]; });
  e.first(true);
  // Operands of a bin: the screen t and the stored envelope w.
  struct Ops {
    float2 t, w;
  };
  e.template mid_loaded<mid_group<E>()>(
      [&](long l, int k) {
        return Ops{a.screen[line_in_plane<E>(l, n) * n + k], a.w[l * n + k]};
      },
      [&](long l, int k, float2& v, Ops o) {
        const float2 b = cscale(v, a.inv);
        const float2 tw = cmul(o.t, o.w);
        a.p[l * n + k] = a.phase * (tw.x * b.y - tw.y * b.x);
        v = cmul(cconj(o.t), b);
      });
  if (a.out == nullptr) return;
  e.second(false);
  e.store2([&](long l, int m, float2 v) { a.out[l * n + m] = v; });
}

__global__ void __launch_bounds__(256) k_bpm_reduce(const float* p, int chunk, long plane,
                                                    float* grad, int init) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= plane) return;
  reduce_terms(p, chunk, plane, grad, init != 0, i);
}

// exp(i phase x) per voxel (src/cpu/bpm.cpp:18-24).
__global__ void __launch_bounds__(256) k_bpm_screen(const float* x, long total, float phase,
                                                    float2* out) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  float s, c;
  sincosf(phase * x[i], &s, &c);
  out[i] = float2{c, s};
}

// z = point - alpha (grad scale); with clamp, max(z, 0) (src/cpu/bpm.cpp:204-212).
__global__ void __launch_bounds__(256) k_bpm_descent(const float* point, const float* grad,
                                                     float scale, float alpha, long total,
                                                     float* z, int clamp) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const float g = grad[i] * scale;
  const float v = point[i] - alpha * g;
  z[i] = (clamp && v < 0.0f) ? 0.0f : v;
}

// y = next + m (next - x); x = next (src/cpu/bpm.cpp:225-228).
__global__ void __launch_bounds__(256) k_bpm_momentum(const float* next, float* x, float* y,
                                                      float m, long total) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const float v = next[i];
  y[i] = v + m * (v - x[i]);
  x[i] = v;
}

__global__ void __launch_bounds__(256) k_bpm_scale(float* x, long total, float scale) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  x[i] *= scale;
}

__global__ void __launch_bounds__(256) k_bpm_offset(const float* in, float* out, long total,
                                                    float v) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  out[i] = in[i] + v;
}

__global__ void __launch_bounds__(256) k_bpm_clamp(float* x, long total) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const float v = x[i];
  x[i] = v < 0.0f ? 0.0f : v;
}

enum class Kind { rows_fwd, rows_out, rows_adj, rows_jvp, cols_fwd, cols_adj };

template <bool ROWS, class E>
void launch_with(Kind k, Args a, unsigned grid, std::size_t lds, hipStream_t s) {
  if constexpr (ROWS) {
    switch (k) {
      case Kind::rows_fwd: hipLaunchKernelGGL(k_bpm_rows_fwd<E>, dim3(grid), dim3(E::kBlock), lds, s, a); break;
      case Kind::rows_out: hipLaunchKernelGGL(k_bpm_rows_out<E>, dim3(grid), dim3(E::kBlock), lds, s, a); break;
      case Kind::rows_jvp: hipLaunchKernelGGL(k_bpm_rows_jvp<E>, dim3(grid), dim3(E::kBlock), lds, s, a); break;
      default: hipLaunchKernelGGL(k_bpm_rows_adj<E>, dim3(grid), dim3(E::kBlock), lds, s, a); break;
    }
  } else {
    if (k == Kind::cols_adj && a.red_grad != nullptr) {
      const long plane = static_cast<long>(a.g.n) * a.g.n;
      a.line_blocks = static_cast<int>(grid);
      a.red_blocks = static_cast<int>((plane + E::kBlock - 1) / E::kBlock);
      grid += static_cast<unsigned>(a.red_blocks);
    }
    switch (k) {
      case Kind::cols_fwd: hipLaunchKernelGGL(k_bpm_cols_fwd<E>, dim3(grid), dim3(E::kBlock), lds, s, a); break;
      default: hipLaunchKernelGGL(k_bpm_cols_adj<E>, dim3(grid), dim3(E::kBlock), lds, s, a); break;
    }
  }
}

// Rows: blocks of 8 lines (2 waves at 256; 4 or 16 lines were no faster). Columns: at least
// 16 lines per block, so that each block row is a whole 128-byte segment (blocks of 8 lines
// would share every cache line with a block that may run on another XCD); 32 at 256 and 384,
// where that was faster (384: forward 62.7 -> 59.5 ms; 192 was slower with 32).
template <bool ROWS, int N1, int N2>
void launch_reg(Kind k, Args a, hipStream_t s) {
  constexpr int kT = N1 > N2 ? N1 : N2;
  constexpr int kRows = 8 * kT;
  constexpr int kCols = (N1 * N2 == 256 || N1 * N2 == 384) ? 32 * kT
                                                             : (16 * kT > 256 ? 16 * kT : 256);
  using E = RegLines<N1, N2, !ROWS, ROWS ? kRows : kCols>;
  a.g.lines_per_block = E::LINES;
  launch_with<ROWS, E>(k, a, E::grid(a.g.lines), 0, s);
}

template <bool ROWS>
void launch(Kind k, Args a, hipStream_t s) {
  switch (a.g.n) {
    case 32: launch_reg<ROWS, 8, 4>(k, a, s); break;
    case 48: launch_reg<ROWS, 8, 6>(k, a, s); break;
    case 64: launch_reg<ROWS, 8, 8>(k, a, s); break;
    case 96: launch_reg<ROWS, 16, 6>(k, a, s); break;
    case 128: launch_reg<ROWS, 16, 8>(k, a, s); break;
    case 192: launch_reg<ROWS, 16, 12>(k, a, s); break;
    case 256: launch_reg<ROWS, 16, 16>(k, a, s); break;
    case 384: launch_reg<ROWS, 16, 24>(k, a, s); break;
    case 512: launch_reg<ROWS, 32, 16>(k, a, s); break;
    default: {
      using E = LdsLines<ROWS>;
      a.g.lines_per_block = E::lines_per_block(a.g.n);
      launch_with<ROWS, E>(k, a, E::grid(a.g.lines, a.g.lines_per_block), E::dynamic_lds(a.g.n), s);
      break;
    }
  }
  check_launch("bpm");
}

}  // namespace

BpmModel::BpmModel(const BpmTables& t, int chunk, int checkpoint)
    : n_(t.n), nz_(t.nz), count_(t.count), plane_(static_cast<long>(t.n) * t.n),
      phase_(t.phase), lipschitz_(t.lipschitz) {
  require(fft_supported(n_), "bpm: n must factor into 2, 3 and 5 and be at most 2048");
  require(chunk >= 0 && checkpoint >= -1, "bpm: chunk must be non-negative, checkpoint at least -1");
  require(static_cast<long>(t.count) * n_ < (1L << 31), "bpm: too many lines for one plan");
  if (checkpoint < 0) {
    // Every envelope kept when the illuminations of one pass (the chunk asked for, or all of
    // them) then fit within half the free device memory: the adjoint recomputes nothing, at
    // 8 n^2 nz bytes per illumination (384^2 x 384 with 81 illuminations: 37 GB, gradient
    // 184 -> about 130 ms against checkpoints every 16 slices). Otherwise checkpoints keep the
    // passes wide.
    std::size_t free_bytes = 0, total_bytes = 0;
    check(hipMemGetInfo(&free_bytes, &total_bytes), "hipMemGetInfo");
    const std::size_t keep_all = static_cast<std::size_t>(plane_) *
                                 (5 * sizeof(float2) + sizeof(float) + sizeof(float2) * nz_);
    const std::size_t angles = static_cast<std::size_t>(chunk > 0 ? std::min(chunk, count_) : count_);
    checkpoint = keep_all * angles <= free_bytes / 2 ? 0 : kBpmCheckpoint;
  }
  seg_len_ = checkpoint == 0 || checkpoint > nz_ ? nz_ : checkpoint;
  nseg_ = (nz_ + seg_len_ - 1) / seg_len_;
  if (chunk == 0) {
    // As many illuminations per pass as keep the adjoint's buffers within half the free
    // device memory (src/cuda/bpm.cu:225-232 sizes its chunks the same way).
    std::size_t free_bytes = 0, total_bytes = 0;
    check(hipMemGetInfo(&free_bytes, &total_bytes), "hipMemGetInfo");
    const std::size_t per_angle =
        static_cast<std::size_t>(plane_) * (5 * sizeof(float2) + sizeof(float) +
                                            2 * sizeof(float2) * (nseg_ - 1) +
                                            sizeof(float2) * seg_len_);
    const std::size_t fit = std::max<std::size_t>(1, free_bytes / 2 / per_angle);
    chunk_ = static_cast<int>(std::min<std::size_t>(fit, count_));
  } else {
    chunk_ = chunk > count_ ? count_ : chunk;
  }
  const std::size_t table = t.step.size();
  step_.resize(table);
  last_.resize(table);
  q_.resize(table);
  check(hipMemcpy(step_.get(), t.step.data(), table * sizeof(float2), hipMemcpyHostToDevice), "bpm tables");
  check(hipMemcpy(last_.get(), t.last.data(), table * sizeof(float2), hipMemcpyHostToDevice), "bpm tables");
  check(hipMemcpy(q_.get(), t.q.data(), table * sizeof(float), hipMemcpyHostToDevice), "bpm tables");
  const std::size_t cp = static_cast<std::size_t>(chunk_) * plane_;
  screen_.resize(static_cast<std::size_t>(nz_) * plane_);
  f_.resize(cp);
  s_.resize(cp);
}

// The adjoint's buffers, allocated by the first gradient or vjp.
void BpmModel::ensure_adjoint() {
  if (w_.size() != 0) return;
  const std::size_t cp = static_cast<std::size_t>(chunk_) * plane_;
  a_.resize(cp);
  g_.resize(cp);
  p_.resize(cp);
  ck_s_.resize(cp * (nseg_ - 1));
  ck_t_.resize(cp * (nseg_ - 1));
  w_.resize(cp * seg_len_);
}

std::size_t BpmModel::bytes() const {
  return step_.bytes() + last_.bytes() + q_.bytes() + screen_.bytes() + f_.bytes() + s_.bytes() +
         a_.bytes() + g_.bytes() + p_.bytes() + ck_s_.bytes() + ck_t_.bytes() + w_.bytes();
}

float2* BpmModel::ck_s(int j) const { return ck_s_.get() + (j - 1) * static_cast<long>(chunk_) * plane_; }
float2* BpmModel::ck_t(int j) const { return ck_t_.get() + (j - 1) * static_cast<long>(chunk_) * plane_; }

BpmModel::Args BpmModel::base(int chunk_count) const {
  Args a{};
  a.g.tw = fft_twiddles(n_);
  a.g.f = fft_factors(n_);
  a.g.n = n_;
  a.g.lines = static_cast<long>(chunk_count) * n_;
  a.inv = 1.0f / static_cast<float>(plane_);
  a.phase = phase_;
  return a;
}

void BpmModel::set_screen(const float* x, hipStream_t s) {
  const long total = static_cast<long>(nz_) * plane_;
  hipLaunchKernelGGL(k_bpm_screen, dim3(blocks_for(total)), dim3(256), 0, s, x, total, phase_,
                     screen_.get());
  check_launch("bpm_screen");
}

// Slices 0 .. nz-1 of illuminations [a0, a0 + chunk_count); the final T is left in f_. With
// keep, (S, T) in front of every segment after the first go to the checkpoints and the last
// segment's envelopes to w_.
void BpmModel::forward_pass(int a0, int chunk_count, bool keep, hipStream_t s) {
  const long cp = static_cast<long>(chunk_count) * plane_;
  for (int k = 0; k < nz_; ++k) {
    const int seg = k / seg_len_;
    const bool restart = keep && k > 0 && k % seg_len_ == 0;
    Args r = base(chunk_count);
    r.in = k == 0 ? nullptr : (restart ? ck_t(seg) : f_.get());
    r.out = f_.get();
    r.screen = screen_.get() + k * plane_;
    r.w = keep && seg == nseg_ - 1 ? w_.get() + (k - seg * seg_len_) * cp : nullptr;
    launch<true>(Kind::rows_fwd, r, s);
    const bool boundary = keep && (k + 1) % seg_len_ == 0 && k + 1 < nz_;
    Args c = base(chunk_count);
    c.in = f_.get();
    c.out = boundary ? ck_t((k + 1) / seg_len_) : f_.get();
    c.s_in = k == 0 ? nullptr : (restart ? ck_s(seg) : s_.get());
    c.s_out = boundary ? ck_s((k + 1) / seg_len_) : s_.get();
    c.c = (k == nz_ - 1 ? last_.get() : step_.get()) + a0 * plane_;
    c.q = q_.get() + a0 * plane_;
    launch<false>(Kind::cols_fwd, c, s);
  }
}

// Envelopes of segment `seg` into w_ from its checkpoint, with the kernels of forward_pass.
void BpmModel::recompute(int seg, int a0, int chunk_count, hipStream_t s) {
  const long cp = static_cast<long>(chunk_count) * plane_;
  const int k0 = seg * seg_len_;
  const int k1 = std::min(k0 + seg_len_, nz_) - 1;
  for (int k = k0; k <= k1; ++k) {
    Args r = base(chunk_count);
    r.in = k == 0 ? nullptr : (k == k0 ? ck_t(seg) : f_.get());
    r.out = k == k1 ? nullptr : f_.get();
    r.screen = screen_.get() + k * plane_;
    r.w = w_.get() + (k - k0) * cp;
    launch<true>(Kind::rows_fwd, r, s);
    if (k == k1) break;
    Args c = base(chunk_count);
    c.in = f_.get();
    c.out = f_.get();
    c.s_in = k == 0 ? nullptr : (k == k0 ? ck_s(seg) : s_.get());
    c.s_out = s_.get();
    c.c = step_.get() + a0 * plane_;
    c.q = q_.get() + a0 * plane_;
    launch<false>(Kind::cols_fwd, c, s);
  }
}

// Slices nz-1 .. 0 from the forward x transform of the residual in a_. The gradient terms a
// rows launch leaves in p_ are summed into grad by the next slice's column launch (its
// reduction blocks), those of slice 0 by a launch of their own; with init (the first chunk)
// the sums start from zero, and grad needs no clearing.
void BpmModel::adjoint(int a0, int chunk_count, float* grad, bool init, hipStream_t s) {
  const long cp = static_cast<long>(chunk_count) * plane_;
  int pending = -1;  // slice whose terms are in p_, not yet summed
  for (int seg = nseg_ - 1; seg >= 0; --seg) {
    const int k0 = seg * seg_len_;
    const int k1 = std::min(k0 + seg_len_, nz_) - 1;
    if (seg < nseg_ - 1) recompute(seg, a0, chunk_count, s);
    for (int k = k1; k >= k0; --k) {
      Args c = base(chunk_count);
      c.in = a_.get();
      c.out = a_.get();
      c.g_state = g_.get();
      c.first = k == nz_ - 1 ? 1 : 0;
      c.c = (k == nz_ - 1 ? last_.get() : step_.get()) + a0 * plane_;
      c.q = q_.get() + a0 * plane_;
      if (pending >= 0) {
        c.red_p = p_.get();
        c.red_grad = grad + pending * plane_;
        c.red_chunk = chunk_count;
        c.red_init = init ? 1 : 0;
      }
      launch<false>(Kind::cols_adj, c, s);
      Args r = base(chunk_count);
      r.in = a_.get();
      r.out = k > 0 ? a_.get() : nullptr;
      r.screen = screen_.get() + k * plane_;
      r.w = w_.get() + (k - k0) * cp;
      r.p = p_.get();
      launch<true>(Kind::rows_adj, r, s);
      pending = k;
    }
  }
  hipLaunchKernelGGL(k_bpm_reduce, dim3(blocks_for(plane_)), dim3(256), 0, s, p_.get(),
                     chunk_count, plane_, grad + pending * plane_, init ? 1 : 0);
  check_launch("bpm_reduce");
}

void BpmModel::forward(const float* x, float2* y, hipStream_t s) {
  set_screen(x, s);
  for (int a0 = 0; a0 < count_; a0 += chunk_) {
    const int b = std::min(chunk_, count_ - a0);
    forward_pass(a0, b, false, s);
    Args r = base(b);
    r.in = f_.get();
    r.out = y + a0 * plane_;
    launch<true>(Kind::rows_out, r, s);
  }
}

void BpmModel::vjp(const float* x, const float2* r, float* grad, hipStream_t s) {
  ensure_adjoint();
  set_screen(x, s);
  for (int a0 = 0; a0 < count_; a0 += chunk_) {
    const int b = std::min(chunk_, count_ - a0);
    forward_pass(a0, b, true, s);
    fft_lines(r + a0 * plane_, a_.get(), n_, lines_along(2, 1, b * n_, n_, n_), false, 1.0f, 0, s);
    adjoint(a0, b, grad, a0 == 0, s);
  }
}

// Primal and tangent march slice by slice: the primal rows store the slice's envelope in w_,
// the tangent rows read it; T, S in f_, s_ and dT, dS in a_, g_ (the adjoint's buffers).
void BpmModel::jvp(const float* x, const float* h, float2* y, hipStream_t s) {
  ensure_adjoint();
  set_screen(x, s);
  for (int a0 = 0; a0 < count_; a0 += chunk_) {
    const int b = std::min(chunk_, count_ - a0);
    for (int k = 0; k < nz_; ++k) {
      Args r = base(b);
      r.in = k == 0 ? nullptr : f_.get();
      r.out = f_.get();
      r.screen = screen_.get() + k * plane_;
      r.w = w_.get();
      launch<true>(Kind::rows_fwd, r, s);
      Args t = base(b);
      t.in = k == 0 ? nullptr : a_.get();
      t.out = a_.get();
      t.screen = screen_.get() + k * plane_;
      t.w = w_.get();
      t.h = h + k * plane_;
      launch<true>(Kind::rows_jvp, t, s);
      Args c = base(b);
      c.in = f_.get();
      c.out = f_.get();
      c.s_in = k == 0 ? nullptr : s_.get();
      c.s_out = s_.get();
      c.c = (k == nz_ - 1 ? last_.get() : step_.get()) + a0 * plane_;
      c.q = q_.get() + a0 * plane_;
      launch<false>(Kind::cols_fwd, c, s);
      Args d = c;
      d.in = a_.get();
      d.out = a_.get();
      d.s_in = k == 0 ? nullptr : g_.get();
      d.s_out = g_.get();
      d.s_zero = 1;
      launch<false>(Kind::cols_fwd, d, s);
    }
    Args r = base(b);
    r.in = a_.get();
    r.out = y + a0 * plane_;
    launch<true>(Kind::rows_out, r, s);
  }
}

void BpmModel::gradient(const float* x, const float2* fields, int first, int last, float* grad,
                        double* loss_lines, hipStream_t s) {
  require(first >= 0 && last <= count_ && first < last, "bpm: illumination range out of bounds");
  ensure_adjoint();
  set_screen(x, s);
  for (int a0 = first; a0 < last; a0 += chunk_) {
    const int b = std::min(chunk_, last - a0);
    forward_pass(a0, b, true, s);
    Args r = base(b);
    r.in = f_.get();
    r.out = a_.get();
    r.data = fields + a0 * plane_;
    r.loss = loss_lines + static_cast<long>(a0) * n_;
    launch<true>(Kind::rows_out, r, s);
    adjoint(a0, b, grad, a0 == first, s);
  }
}

void bpm_scale(float* x, long total, float scale, hipStream_t s) {
  hipLaunchKernelGGL(k_bpm_scale, dim3(blocks_for(total)), dim3(256), 0, s, x, total, scale);
  check_launch("bpm_scale");
}

void bpm_offset(const float* in, float* out, long total, float v, hipStream_t s) {
  hipLaunchKernelGGL(k_bpm_offset, dim3(blocks_for(total)), dim3(256), 0, s, in, out, total, v);
  check_launch("bpm_offset");
}

BpmSolver::BpmSolver(BpmModel& m, const BpmSolveConfig& c)
    : m_(m), c_(c), total_(static_cast<long>(m.nz()) * m.n() * m.n()),
      batch_(c.batch <= 0 || c.batch >= m.count() ? m.count() : c.batch),
      grad_(total_), z_(total_),
      lines_(static_cast<std::size_t>(std::max(c.iterations, 1)) * m.count() * m.n()),
      lines_host_(lines_.size()) {
  require(c.iterations >= 0 && c.step > 0 && c.tv_weight >= 0 && c.tv_inner >= 1,
          "bpm: invalid solver options");
  if (batch_ == m.count()) {
    y_.resize(total_);
    next_.resize(total_);
  }
  if (c.tv_weight > 0.0f) dual_ = std::make_unique<TvState>(m.nz(), m.n(), m.n());
  m.reserve();
}

std::size_t BpmSolver::bytes() const {
  return grad_.bytes() + z_.bytes() + y_.bytes() + next_.bytes() + lines_.bytes() +
         (dual_ ? dual_->bytes() : 0);
}

void BpmSolver::run(const float2* fields, float* x, double* loss, hipStream_t s) {
  BpmModel& m = m_;
  const BpmSolveConfig& o = c_;
  const int count = m.count();
  const int n = m.n();
  const long total = total_;
  const int batch = batch_;
  const float alpha = o.step / m.lipschitz();
  const bool tv = o.tv_weight > 0.0f;
  const unsigned grid = blocks_for(total);
  if (tv) dual_->reset(s);
  hipLaunchKernelGGL(k_bpm_clamp, dim3(grid), dim3(256), 0, s, x, total);
  check_launch("bpm_clamp");
  // z = point - alpha grad / (last - first), then out = prox(z).
  const auto prox_step = [&](const float* point, float* out, int first, int last) {
    const float scale = 1.0f / static_cast<float>(last - first);
    hipLaunchKernelGGL(k_bpm_descent, dim3(grid), dim3(256), 0, s, point, grad_.get(), scale,
                       alpha, total, tv ? z_.get() : out, tv ? 0 : 1);
    check_launch("bpm_descent");
    if (tv) dual_->prox(z_.get(), out, alpha * o.tv_weight, o.tv_inner, s);
  };
  std::vector<int> firsts;
  if (batch == count) {
    check(hipMemcpyAsync(y_.get(), x, total * sizeof(float), hipMemcpyDeviceToDevice, s), "bpm copy");
    float tk = 1.0f;
    for (int it = 0; it < o.iterations; ++it) {
      m.gradient(y_.get(), fields, 0, count, grad_.get(), lines_.get() + static_cast<long>(it) * count * n, s);
      prox_step(y_.get(), next_.get(), 0, count);
      const float tn = fista_t(tk);
      const float mm = (tk - 1.0f) / tn;
      hipLaunchKernelGGL(k_bpm_momentum, dim3(grid), dim3(256), 0, s, next_.get(), x, y_.get(), mm,
                         total);
      check_launch("bpm_momentum");
      tk = tn;
    }
    firsts.push_back(0);
  } else {
    for (int it = 0; it < o.iterations; ++it) {
      for (int first = 0; first < count; first += batch) {
        const int last = std::min(first + batch, count);
        m.gradient(x, fields, first, last, grad_.get(), lines_.get() + static_cast<long>(it) * count * n, s);
        prox_step(x, x, first, last);
      }
    }
    for (int first = 0; first < count; first += batch) firsts.push_back(first);
  }
  // Data term per iteration: per batch 0.5 / count times its sum over illuminations, in order.
  const std::size_t used = static_cast<std::size_t>(o.iterations) * count * n;
  if (used > 0) {
    check(hipMemcpyAsync(lines_host_.get(), lines_.get(), used * sizeof(double),
                         hipMemcpyDeviceToHost, s),
          "bpm loss");
  }
  check(hipStreamSynchronize(s), "bpm solve");
  const double* h = lines_host_.get();
  for (int it = 0; it < o.iterations; ++it) {
    double value = 0;
    for (std::size_t bi = 0; bi < firsts.size(); ++bi) {
      const int first = firsts[bi];
      const int last = bi + 1 < firsts.size() ? firsts[bi + 1] : count;
      double sum = 0;
      for (int a = first; a < last; ++a) {
        double la = 0;
        for (int y = 0; y < n; ++y) la += h[(static_cast<std::size_t>(it) * count + a) * n + y];
        sum += la;
      }
      value += 0.5 * sum / count;
    }
    if (loss) loss[it] = value;
  }
}

void bpm_solve(BpmModel& m, const float2* fields, float* x, const BpmSolveConfig& o, double* loss,
               hipStream_t s) {
  BpmSolver(m, o).run(fields, x, loss, s);
}

}  // namespace qh
