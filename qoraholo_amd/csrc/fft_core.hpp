// LDS-resident mixed-radix FFT used by every transform kernel of the library.
//
// A workgroup holds a batch of lines in LDS, element (line, i) at buf[line * ls + i * es], and
// runs Stockham autosort passes of radix 8, 4, 2, 3 and 5 between two LDS buffers: each pass
// reads R values at stride n/R, applies the twiddles exp(-+2 pi i k s / (ns R)) from an LDS
// copy of a table computed in double on the host, does an R-point DFT in registers and writes
// the values to their expanded positions. After the last pass the lines are in natural order.
// Transforms are unnormalised; `inverse` flips the sign of the exponent.
#pragma once

#include <hip/hip_runtime.h>

namespace qh {

inline constexpr int kMaxFactors = 16;

// Block index with consecutive values on one XCD. Workgroups are dealt to the 8 XCDs in turn,
// so block b runs on XCD b mod 8; renumbering gives each XCD a contiguous range of blocks, and
// neighbouring blocks that share 128-byte lines (line batches whose rows are not line aligned,
// byte arrays read 16 bytes per row) share them through one L2 instead of fetching them once
// per XCD. The last nb mod 8 blocks keep their numbers.
__device__ __forceinline__ long xcd_block() {
  const long nb = gridDim.x;
  const long b = blockIdx.x;
  const long full = nb - nb % 8;
  return b < full ? (b % 8) * (full / 8) + b / 8 : b;
}
// The same renumbering over the grid's first nb blocks only (the later blocks run other work
// and must not call it).
__device__ __forceinline__ long xcd_block(int nb) {
  const int b = static_cast<int>(blockIdx.x);
  const int full = nb - nb % 8;
  return b < full ? (b % 8) * (full / 8) + b / 8 : b;
}

struct FftFactors {
  int n = 0;
  int count = 0;
  int r[kMaxFactors] = {};
};

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return {a.x * s, a.y * s}; }
__device__ __forceinline__ float2 cconj(float2 a) { return {a.x, -a.y}; }
// -i a for the forward transform, +i a for the inverse.
__device__ __forceinline__ float2 rot(float2 a, bool inv) {
  return inv ? float2{-a.y, a.x} : float2{a.y, -a.x};
}

template <int R>
__device__ __forceinline__ void bfly(float2* v, bool inv);

template <>
__device__ __forceinline__ void bfly<2>(float2* v, bool) {
  const float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <>
__device__ __forceinline__ void bfly<4>(float2* v, bool inv) {
  const float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
  const float2 s13 = cadd(v[1], v[3]), d13 = rot(csub(v[1], v[3]), inv);
  v[0] = cadd(s02, s13);
  v[1] = cadd(d02, d13);
  v[2] = csub(s02, s13);
  v[3] = csub(d02, d13);
}

template <>
__device__ __forceinline__ void bfly<8>(float2* v, bool inv) {
  // Two radix-4 DFTs over even and odd samples, combined with W8^k = exp(-+i pi k / 4).
  float2 e[4] = {v[0], v[2], v[4], v[6]};
  float2 o[4] = {v[1], v[3], v[5], v[7]};
  bfly<4>(e, inv);
  bfly<4>(o, inv);
  const float h = 0.70710678118654752440f;
  // W8^1 o1, W8^2 o2 = -+i o2, W8^3 o3.
  const float2 o1 = inv ? float2{h * (o[1].x - o[1].y), h * (o[1].x + o[1].y)}
                        : float2{h * (o[1].x + o[1].y), h * (o[1].y - o[1].x)};
  const float2 o2 = rot(o[2], inv);
  const float2 o3 = inv ? float2{-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y)}
                        : float2{h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y)};
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

template <>
__device__ __forceinline__ void bfly<3>(float2* v, bool inv) {
  const float c = -0.5f;
  const float s = 0.86602540378443864676f;
  const float2 t1 = cadd(v[1], v[2]);
  const float2 t2 = {v[0].x + c * t1.x, v[0].y + c * t1.y};
  const float2 t3 = rot(cscale(csub(v[1], v[2]), s), inv);
  v[0] = cadd(v[0], t1);
  v[1] = cadd(t2, t3);
  v[2] = csub(t2, t3);
}

template <>
__device__ __forceinline__ void bfly<5>(float2* v, bool inv) {
  const float c1 = 0.30901699437494742410f;
  const float c2 = -0.80901699437494742410f;
  const float s1 = 0.95105651629515357212f;
  const float s2 = 0.58778525229247312917f;
  const float2 a14 = cadd(v[1], v[4]), d14 = csub(v[1], v[4]);
  const float2 a23 = cadd(v[2], v[3]), d23 = csub(v[2], v[3]);
  const float2 r1 = {v[0].x + c1 * a14.x + c2 * a23.x, v[0].y + c1 * a14.y + c2 * a23.y};
  const float2 r2 = {v[0].x + c2 * a14.x + c1 * a23.x, v[0].y + c2 * a14.y + c1 * a23.y};
  const float2 i1 = rot({s1 * d14.x + s2 * d23.x, s1 * d14.y + s2 * d23.y}, inv);
  const float2 i2 = rot({s2 * d14.x - s1 * d23.x, s2 * d14.y - s1 * d23.y}, inv);
  v[0] = cadd(cadd(v[0], a14), a23);
  v[1] = cadd(r1, i1);
  v[4] = csub(r1, i1);
  v[2] = cadd(r2, i2);
  v[3] = csub(r2, i2);
}

// One Stockham pass of radix R over `lines` lines of length n; ns is the length of the
// sub-transforms already combined. line_fast maps consecutive threads to consecutive lines.
template <int R>
__device__ __forceinline__ void radix_pass(const float2* src, float2* dst, int ls, int es,
                                           int lines, int n, int ns, const float2* tw, bool inv,
                                           bool line_fast) {
  const int q = n / R;
  const int step = n / (ns * R);
  const int nb = lines * q;
  for (int t = threadIdx.x; t < nb; t += blockDim.x) {
    int line, j;
    if (line_fast) {
      line = t % lines;
      j = t / lines;
    } else {
      line = t / q;
      j = t - line * q;
    }
    const int k = j % ns;
    float2 v[R];
#pragma unroll
    for (int s = 0; s < R; ++s) v[s] = src[line * ls + (j + s * q) * es];
    if (k != 0) {
#pragma unroll
      for (int s = 1; s < R; ++s) {
        float2 w = tw[k * s * step];
        if (inv) w.y = -w.y;
        v[s] = cmul(v[s], w);
      }
    }
    bfly<R>(v, inv);
    const int o = (j - k) * R + k;
#pragma unroll
    for (int s = 0; s < R; ++s) dst[line * ls + (o + s * ns) * es] = v[s];
  }
}

// Runs all passes; the data starts in `a` and the returned pointer (a or b) holds the result.
// Every thread of the block must call it. The caller synchronises before the first pass.
__device__ __forceinline__ float2* lds_fft(float2* a, float2* b, int ls, int es, int lines,
                                           const FftFactors& f, const float2* tw, bool inv,
                                           bool line_fast) {
  float2* src = a;
  float2* dst = b;
  int ns = 1;
  for (int p = 0; p < f.count; ++p) {
    const int r = f.r[p];
    switch (r) {
      case 8: radix_pass<8>(src, dst, ls, es, lines, f.n, ns, tw, inv, line_fast); break;
      case 4: radix_pass<4>(src, dst, ls, es, lines, f.n, ns, tw, inv, line_fast); break;
      case 2: radix_pass<2>(src, dst, ls, es, lines, f.n, ns, tw, inv, line_fast); break;
      case 3: radix_pass<3>(src, dst, ls, es, lines, f.n, ns, tw, inv, line_fast); break;
      default: radix_pass<5>(src, dst, ls, es, lines, f.n, ns, tw, inv, line_fast); break;
    }
    __syncthreads();
    float2* t = src;
    src = dst;
    dst = t;
    ns *= r;
  }
  return src;
}

}  // namespace qh
