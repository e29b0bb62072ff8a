// Line engines for fused kernels that transform a batch of lines twice with an elementwise
// step in between (the split-step beam propagation and its adjoint).
//
// A kernel body is written once against this interface and instantiated with either engine:
//   load(f)          input sample i of global line l is f(l, i)
//   first(inverse)   transform every line
//   mid(f)           f(l, k, v&) may read and replace output bin k of line l
//   mid_loaded<G>(ld, f)  the same with operands: ld(l, k) loads bin k's operands, f(l, k, v&,
//                    ops) uses them; the operands of G bins are loaded before any f runs.
//                    f's stores may alias ld's loads as far as the compiler can tell, so in a
//                    plain per-bin loop each bin waits for its own loads (measured: 16 to 32
//                    memory round trips in a row per thread in the BPM slice kernels).
//   reduce_lines(f, out)  out[l] = sum over k of f(l, k, v), summed in a fixed order
//   second(inverse)  transform every line again
//   store1(f) / store2(f)  f(l, m, v) receives element m after the first / second transform
// Lines of a block that lie beyond `lines` are neither read nor written.
//
// RegLines: lengths N = N1 N2 of the form 2^a 3^b (32 .. 512) as register-resident four-step
// transforms, max(N1, N2) threads per line; the first
// transform leaves thread j with bins N2 k1 + j, which is the input layout of the transform with
// the factors swapped, so the second transform needs no reordering. BLOCK threads hold
// BLOCK / max(N1, N2) lines: one-wave blocks suit contiguous rows (many small blocks in flight,
// cheap barriers); column blocks hold at least 16 lines so that each load spans whole 128-byte
// lines.
// LdsLines: any length that factors into 2, 3 and 5 (up to 2048), lines held in LDS between
// radix passes, 256 threads.
#pragma once

#include <hip/hip_runtime.h>

#include "fft4s.hpp"
#include "fft_core.hpp"

namespace qh {

// Shape of a batch of lines and the twiddles exp(-2 pi i t / n), t < n.
struct LineGeom {
  const float2* tw;
  FftFactors f;
  int n;
  long lines;
  int lines_per_block;  // LdsLines only
};

constexpr int ce_max(int a, int b) { return a > b ? a : b; }

template <int N1, int N2, bool LINE_FAST, int BLOCK = 256>
struct RegLines {
  static constexpr int kBlock = BLOCK;
  static constexpr int N = N1 * N2;
  static constexpr int T = N1 > N2 ? N1 : N2;
  static constexpr int LINES = BLOCK / T;
  static_assert(LINES >= 1 && BLOCK % T == 0, "a block holds whole lines");
  static constexpr int kLds =
      ce_max(t_size<N1, N2, LINES, LINE_FAST>(), t_size<N2, N1, LINES, LINE_FAST>());
  // Threads that walk along lines read the twiddles from one matrix W^(x y), x < N2, y < N1,
  // rows padded to N1 + 1 entries: by rows in the first transform, by columns in the second
  // (fft_4step's kTwByK and kTwByJ). Threads that walk across lines read the plain table.
  static constexpr bool kMatrix = !LINE_FAST;
  static constexpr int kTwPitch = N1 + 1;
  static constexpr int kTw = kMatrix ? N2 * kTwPitch : N;

  float2* t;
  float2* tw;
  int line;
  int j;
  long l;
  bool valid;
  float2 v[N2];
  float2 w[N1];

  __device__ static float2* lds_t() {
    __shared__ float2 s[kLds];
    return s;
  }
  __device__ static float2* lds_tw() {
    __shared__ float2 s[kTw];
    return s;
  }
  __device__ static double* lds_red() {
    __shared__ double s[BLOCK];
    return s;
  }

  static unsigned grid(long lines) { return static_cast<unsigned>((lines + LINES - 1) / LINES); }
  static std::size_t dynamic_lds(int) { return 0; }

  __device__ explicit RegLines(const LineGeom& g) { init<false>(g, 0); }
  // The grid's first nb blocks run the lines (its later blocks run other work).
  __device__ RegLines(const LineGeom& g, int nb) { init<true>(g, nb); }
  // The block's line index is formed after the twiddle fill has been issued, so that the fill's
  // loads do not wait for the grid size (a short kernel's start is on its critical path).
  template <bool PART>
  __device__ void init(const LineGeom& g, int nb) {
    t = lds_t();
    tw = lds_tw();
    const int tid = static_cast<int>(threadIdx.x);
    if constexpr (kMatrix) {
      fill_twiddle_matrix<N2, N1, kTwPitch>(tw, g.tw, 1, tid, BLOCK);
    } else {
      for (int i = tid; i < N; i += BLOCK) tw[i] = g.tw[i];
    }
    line = LINE_FAST ? tid % LINES : tid / T;
    j = LINE_FAST ? tid / LINES : tid % T;
    l = (PART ? xcd_block(nb) : xcd_block()) * LINES + line;
    valid = l < g.lines;
  }

  template <class F>
  __device__ void load(F f) {
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) {
      v[n2] = (valid && j < N1) ? f(l, j + N1 * n2) : float2{0.0f, 0.0f};
    }
    __syncthreads();
  }
  __device__ void first(bool inverse) {
    fft_4step<N1, N2, LINES, LINE_FAST, kMatrix ? kTwByK : kTwPlain, kTwPitch>(v, w, j, line, t, tw,
                                                                             inverse);
  }
  template <class F>
  __device__ void mid(F f) {
    if (valid && j < N2) {
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) f(l, N2 * k1 + j, w[k1]);
    }
  }
  template <int G, class L, class F>
  __device__ void mid_loaded(L ld, F f) {
    if (valid && j < N2) {
#pragma unroll
      for (int k0 = 0; k0 < N1; k0 += G) {
        decltype(ld(l, 0)) ops[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
          if (k0 + g < N1) ops[g] = ld(l, N2 * (k0 + g) + j);
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
          if (k0 + g < N1) f(l, N2 * (k0 + g) + j, w[k0 + g], ops[g]);
        }
      }
    }
  }
  template <class F>
  __device__ void reduce_lines(F f, double* out) {
    double* red = lds_red();
    double s = 0.0;
    if (valid && j < N2) {
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) s += f(l, N2 * k1 + j, w[k1]);
    }
    red[line * T + j] = s;
    __syncthreads();
    if (valid && j == 0) {
      double total = 0.0;
      for (int jj = 0; jj < N2; ++jj) total += red[line * T + jj];
      out[l] = total;
    }
    __syncthreads();
  }
  __device__ void second(bool inverse) {
    fft_4step<N2, N1, LINES, LINE_FAST, kMatrix ? kTwByJ : kTwPlain, kTwPitch>(w, v, j, line, t, tw,
                                                                             inverse);
  }
  template <class F>
  __device__ void store1(F f) {
    if (valid && j < N2) {
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) f(l, N2 * k1 + j, w[k1]);
    }
  }
  template <class F>
  __device__ void store2(F f) {
    if (valid && j < N1) {
#pragma unroll
      for (int k1 = 0; k1 < N2; ++k1) f(l, N1 * k1 + j, v[k1]);
    }
  }
};

// CONTIG: lines contiguous in memory, threads walk along lines. Otherwise lines are adjacent
// columns: threads walk across lines so that each row of the block is one coalesced segment.
template <bool CONTIG>
struct LdsLines {
  static constexpr int kBlock = 256;
  float2* tw;
  float2* b0;
  float2* b1;
  float2* cur;
  const FftFactors* f;
  int n;
  int nl;
  int ls;
  int es;
  long l0;

  static int lines_per_block(int n) {
    return CONTIG ? (2048 / n > 64 ? 64 : (2048 / n < 1 ? 1 : 2048 / n))
                  : (2048 / n > 16 ? 16 : (2048 / n < 1 ? 1 : 2048 / n));
  }
  static unsigned grid(long lines, int lpb) { return static_cast<unsigned>((lines + lpb - 1) / lpb); }
  static std::size_t dynamic_lds(int n) {
    return (static_cast<std::size_t>(n) + 2 * static_cast<std::size_t>(lines_per_block(n)) * n) *
           sizeof(float2);
  }

  __device__ explicit LdsLines(const LineGeom& g) { init<false>(g, 0); }
  // The grid's first nb blocks run the lines (its later blocks run other work).
  __device__ LdsLines(const LineGeom& g, int nb) { init<true>(g, nb); }
  template <bool PART>
  __device__ void init(const LineGeom& g, int nb) {
    extern __shared__ float2 smem[];
    n = g.n;
    f = &g.f;
    tw = smem;
    b0 = smem + n;
    b1 = b0 + g.lines_per_block * n;
    cur = b0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = g.tw[i];
    l0 = (PART ? xcd_block(nb) : xcd_block()) * g.lines_per_block;
    const long left = g.lines - l0;
    nl = static_cast<int>(left < g.lines_per_block ? left : g.lines_per_block);
    ls = CONTIG ? n : 1;
    es = CONTIG ? 1 : nl;
  }

  __device__ void split(int e, int& line, int& i) const {
    if (CONTIG) {
      line = e / n;
      i = e - line * n;
    } else {
      i = e / nl;
      line = e - i * nl;
    }
  }

  template <class F>
  __device__ void load(F fn) {
    for (int e = threadIdx.x; e < nl * n; e += blockDim.x) {
      int line, i;
      split(e, line, i);
      b0[line * ls + i * es] = fn(l0 + line, i);
    }
    cur = b0;
    __syncthreads();
  }
  __device__ void first(bool inverse) { cur = lds_fft(b0, b1, ls, es, nl, *f, tw, inverse, !CONTIG); }
  template <class F>
  __device__ void mid(F fn) {
    for (int e = threadIdx.x; e < nl * n; e += blockDim.x) {
      int line, k;
      split(e, line, k);
      fn(l0 + line, k, cur[line * ls + k * es]);
    }
    __syncthreads();
  }
  template <int G, class L, class F>
  __device__ void mid_loaded(L ld, F fn) {
    mid([&](long l, int k, float2& v) { fn(l, k, v, ld(l, k)); });
  }
  template <class F>
  __device__ void reduce_lines(F fn, double* out) {
    if (static_cast<int>(threadIdx.x) < nl) {
      const int line = static_cast<int>(threadIdx.x);
      double s = 0.0;
      for (int k = 0; k < n; ++k) s += fn(l0 + line, k, cur[line * ls + k * es]);
      out[l0 + line] = s;
    }
    __syncthreads();
  }
  __device__ void second(bool inverse) {
    float2* other = cur == b0 ? b1 : b0;
    cur = lds_fft(cur, other, ls, es, nl, *f, tw, inverse, !CONTIG);
  }
  template <class F>
  __device__ void store1(F fn) {
    for (int e = threadIdx.x; e < nl * n; e += blockDim.x) {
      int line, m;
      split(e, line, m);
      fn(l0 + line, m, cur[line * ls + m * es]);
    }
  }
  template <class F>
  __device__ void store2(F fn) { store1(fn); }
};

}  // namespace qh
