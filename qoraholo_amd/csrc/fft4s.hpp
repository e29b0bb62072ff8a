// Register-resident four-step FFT of lines of length N = N1 * N2 for gfx950.
//
// With n = n1 + N1 n2 and k = N2 k1 + k2:
//   X[N2 k1 + k2] = sum_n1 W_N^(n1 k2) W_N1^(n1 k1) sum_n2 x[n1 + N1 n2] W_N2^(n2 k2).
// Thread n1 of a line holds its N2 samples x[n1 + N1 n2] in registers and transforms them
// (step 1), multiplies by W_N^(n1 k2), and the block transposes through LDS once so that
// thread k2 holds Y[., k2] and does the N1-point transform (step 2), ending with
// X[N2 k1 + k2] for all k1 in its registers. The small DFTs (2, 3, 4, 6, 8, 12, 16, 24, 32) run
// entirely in registers with constant twiddles. Compared with radix passes between LDS buffers this moves
// each value through LDS once per transform.
#pragma once

#include <hip/hip_runtime.h>

#include "fft_core.hpp"

namespace qh {

// exp(-2 pi i j / 16), exp(-2 pi i j / 32) and exp(-2 pi i j / 24), rounded from double.
__device__ constexpr float kW16[16][2] = {
    {1.0f, 0.0f}, {0.92387953251128673848f, -0.38268343236508978178f},
    {0.70710678118654757274f, -0.70710678118654746172f}, {0.38268343236508983729f, -0.92387953251128673848f},
    {0.0f, -1.0f}, {-0.38268343236508972627f, -0.92387953251128673848f},
    {-0.70710678118654746172f, -0.70710678118654757274f}, {-0.92387953251128673848f, -0.38268343236508989280f},
    {-1.0f, 0.0f}, {-0.92387953251128684951f, 0.38268343236508967076f},
    {-0.70710678118654768376f, 0.70710678118654746172f}, {-0.38268343236509033689f, 0.92387953251128651644f},
    {0.0f, 1.0f}, {0.38268343236509000382f, 0.92387953251128662746f},
    {0.70710678118654735069f, 0.70710678118654768376f}, {0.92387953251128651644f, 0.38268343236509039240f}};

__device__ constexpr float kW32[32][2] = {
    {1.0f, 0.0f}, {0.98078528040323043058f, -0.19509032201612824808f},
    {0.92387953251128673848f, -0.38268343236508978178f}, {0.83146961230254523567f, -0.55557023301960217765f},
    {0.70710678118654757274f, -0.70710678118654746172f}, {0.55557023301960228867f, -0.83146961230254523567f},
    {0.38268343236508983729f, -0.92387953251128673848f}, {0.19509032201612833135f, -0.98078528040323043058f},
    {0.0f, -1.0f}, {-0.19509032201612819257f, -0.98078528040323043058f},
    {-0.38268343236508972627f, -0.92387953251128673848f}, {-0.55557023301960195560f, -0.83146961230254545772f},
    {-0.70710678118654746172f, -0.70710678118654757274f}, {-0.83146961230254534669f, -0.55557023301960217765f},
    {-0.92387953251128673848f, -0.38268343236508989280f}, {-0.98078528040323043058f, -0.19509032201612860891f},
    {-1.0f, 0.0f}, {-0.98078528040323043058f, 0.19509032201612835911f},
    {-0.92387953251128684951f, 0.38268343236508967076f}, {-0.83146961230254545772f, 0.55557023301960195560f},
    {-0.70710678118654768376f, 0.70710678118654746172f}, {-0.55557023301960217765f, 0.83146961230254523567f},
    {-0.38268343236509033689f, 0.92387953251128651644f}, {-0.19509032201612866442f, 0.98078528040323031956f},
    {0.0f, 1.0f}, {0.19509032201612830359f, 0.98078528040323043058f},
    {0.38268343236509000382f, 0.92387953251128662746f}, {0.55557023301960184458f, 0.83146961230254545772f},
    {0.70710678118654735069f, 0.70710678118654768376f}, {0.83146961230254523567f, 0.55557023301960217765f},
    {0.92387953251128651644f, 0.38268343236509039240f}, {0.98078528040323031956f, 0.19509032201612871993f}};

__device__ constexpr float kW24[24][2] = {
    {1.0f, 0.0f}, {0.96592582628906831221f, -0.25881904510252073948f},
    {0.86602540378443870761f, -0.49999999999999994449f}, {0.70710678118654757274f, -0.70710678118654746172f},
    {0.50000000000000011102f, -0.86602540378443859659f}, {0.25881904510252073948f, -0.96592582628906831221f},
    {0.0f, -1.0f}, {-0.25881904510252062845f, -0.96592582628906831221f},
    {-0.49999999999999977796f, -0.86602540378443870761f}, {-0.70710678118654746172f, -0.70710678118654757274f},
    {-0.86602540378443870761f, -0.49999999999999994449f}, {-0.96592582628906820119f, -0.25881904510252101703f},
    {-1.0f, 0.0f}, {-0.96592582628906831221f, 0.25881904510252079499f},
    {-0.86602540378443881863f, 0.49999999999999972244f}, {-0.7071067811865479058f, 0.70710678118654712865f},
    {-0.50000000000000044409f, 0.86602540378443837454f}, {-0.25881904510252062845f, 0.96592582628906831221f},
    {0.0f, 1.0f}, {0.25881904510252029539f, 0.96592582628906842324f},
    {0.50000000000000011102f, 0.86602540378443859659f}, {0.70710678118654735069f, 0.70710678118654768376f},
    {0.86602540378443837454f, 0.50000000000000044409f}, {0.96592582628906809017f, 0.25881904510252157214f}};


// exp(-+2 pi i j / R) for R dividing 32 or 24.
template <int R>
__device__ __forceinline__ float2 wconst(int j, bool inv) {
  static_assert(32 % R == 0 || 24 % R == 0, "constant twiddles for R dividing 32 or 24");
  float2 w;
  if constexpr (R == 16) {
    w = {kW16[j][0], kW16[j][1]};
  } else if constexpr (32 % R == 0) {
    w = {kW32[j * (32 / R)][0], kW32[j * (32 / R)][1]};
  } else {
    w = {kW24[j * (24 / R)][0], kW24[j * (24 / R)][1]};
  }
  if (inv) w.y = -w.y;
  return w;
}

template <int R>
__device__ __forceinline__ void dft(float2* v, bool inv);

template <>
__device__ __forceinline__ void dft<1>(float2*, bool) {}
template <>
__device__ __forceinline__ void dft<2>(float2* v, bool inv) { bfly<2>(v, inv); }
template <>
__device__ __forceinline__ void dft<4>(float2* v, bool inv) { bfly<4>(v, inv); }
template <>
__device__ __forceinline__ void dft<8>(float2* v, bool inv) { bfly<8>(v, inv); }
template <>
__device__ __forceinline__ void dft<3>(float2* v, bool inv) { bfly<3>(v, inv); }

// Four-step DFT of A * B values in registers, natural order in and out.
template <int A, int B>
__device__ __forceinline__ void dft_regs(float2* v, bool inv) {
  float2 t[A][B];
#pragma unroll
  for (int a = 0; a < A; ++a) {
#pragma unroll
    for (int b = 0; b < B; ++b) t[a][b] = v[a + A * b];
    dft<B>(t[a], inv);
#pragma unroll
    for (int k2 = 1; k2 < B; ++k2) {
      if (a > 0) t[a][k2] = cmul(t[a][k2], wconst<A * B>(a * k2, inv));
    }
  }
#pragma unroll
  for (int k2 = 0; k2 < B; ++k2) {
    float2 u[A];
#pragma unroll
    for (int a = 0; a < A; ++a) u[a] = t[a][k2];
    dft<A>(u, inv);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) v[B * k1 + k2] = u[k1];
  }
}

template <>
__device__ __forceinline__ void dft<16>(float2* v, bool inv) { dft_regs<4, 4>(v, inv); }
template <>
__device__ __forceinline__ void dft<32>(float2* v, bool inv) { dft_regs<4, 8>(v, inv); }
template <>
__device__ __forceinline__ void dft<6>(float2* v, bool inv) { dft_regs<2, 3>(v, inv); }
template <>
__device__ __forceinline__ void dft<12>(float2* v, bool inv) { dft_regs<4, 3>(v, inv); }
template <>
__device__ __forceinline__ void dft<24>(float2* v, bool inv) { dft_regs<8, 3>(v, inv); }

// LDS index of Y[line][n1][k2] during the transpose. LINE_FAST blocks (threads that walk
// across lines) keep lines contiguous; the others pad each n1 row so that threads that walk
// along a line touch different banks.
template <int N1, int N2, int LINES, bool LINE_FAST>
__device__ __forceinline__ int t_index(int line, int n1, int k2) {
  if constexpr (LINE_FAST) {
    return (n1 * N2 + k2) * LINES + line;
  } else {
    return line * (N1 * (N2 + 1)) + n1 * (N2 + 1) + k2;
  }
}

template <int N1, int N2, int LINES, bool LINE_FAST>
constexpr int t_size() {
  return LINE_FAST ? N1 * N2 * LINES : LINES * N1 * (N2 + 1);
}

// Twiddle layouts of fft_4step, where thread j needs W_N^(j k2) for k2 = 1 .. N2-1:
//   kTwPlain: tw[t] = W_N^t (t < N), read at tw[j k2];
//   kTwByK:   a matrix A[x][y] = W_N^(x y) at tw[x P + y] (fill_twiddle_matrix), read at
//             A[k2][j] = tw[k2 P + j];
//   kTwByJ:   the same matrix read by the transform with the factors swapped, A[j][k2].
// Threads that walk along lines (j differs across the lanes of a wave) read tw[j k2] with a
// stride of k2 entries, up to 16 lanes on one LDS bank (k2 = 16: 41% of the LDS cycles of the
// 384-point BPM row kernels were bank conflicts); A[k2][j] is read at consecutive addresses,
// and A[j][k2] with a stride of P = N1 + 1 entries (2 LDS banks per step) meets at most two
// lanes per bank. One matrix with rows k2 < N2 and columns j < N1 serves both transforms of a
// line engine. Threads that walk across lines share j within a line group and read kTwPlain.
enum TwLayout { kTwPlain, kTwByK, kTwByJ };

// A[x][y] = W_N^(x y) for x < R, y < C at dst[x P + y], gathered from src[stride t] = W_N^t
// (the table's own values, so results do not change with the layout).
template <int R, int C, int P>
__device__ __forceinline__ void fill_twiddle_matrix(float2* dst, const float2* src, int stride,
                                                    int tid, int threads) {
  for (int i = tid; i < R * C; i += threads) {
    const int x = i / C;
    const int y = i - x * C;
    dst[x * P + y] = src[stride * (x * y)];
  }
}

// One line-batch transform. Thread j of `line` enters with v[n2] = x[j + N1 n2] (j < N1) and
// leaves with w[k1] = X[N2 k1 + j] (j < N2); twiddles in LDS in layout TWL (pitch P). Every
// thread of the block calls it; it ends with a barrier so `t` can be reused.
template <int N1, int N2, int LINES, bool LINE_FAST, int TWL = kTwPlain, int P = 0>
__device__ __forceinline__ void fft_4step(float2 (&v)[N2], float2 (&w)[N1], int j, int line,
                                          float2* t, const float2* tw, bool inv) {
  if (j < N1) {
    dft<N2>(v, inv);
#pragma unroll
    for (int k2 = 1; k2 < N2; ++k2) {
      if (j > 0) {
        float2 c = TWL == kTwByK ? tw[k2 * P + j] : (TWL == kTwByJ ? tw[j * P + k2] : tw[j * k2]);
        if (inv) c.y = -c.y;
        v[k2] = cmul(v[k2], c);
      }
    }
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) t[t_index<N1, N2, LINES, LINE_FAST>(line, j, k2)] = v[k2];
  }
  __syncthreads();
  if (j < N2) {
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) w[n1] = t[t_index<N1, N2, LINES, LINE_FAST>(line, n1, j)];
    dft<N1>(w, inv);
  }
  __syncthreads();
}

}  // namespace qh
