// Host interface of the library's own FFT: factor plans, device twiddle tables and launchers
// for batched transforms along one axis of a 3D array. 2D and 3D transforms are sequences of
// these passes (one HBM read and write each); the solver fuses its elementwise steps into them.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fft_core.hpp"

namespace qh {

// Radix plan for length n: factors 8, 4, 2, 3, 5. Throws InvalidArgument for other primes.
FftFactors fft_factors(int n);
bool fft_supported(int n);

// Device table exp(-2 pi i t / n), t in [0, n), evaluated in double; cached per device.
const float2* fft_twiddles(int n);
// Device table exp(-i pi k / (2 n)), k in [0, n): the DCT-II twiddles (Makhoul).
const float2* dct_twiddles(int n);

// Lines of a 3D array: line l starts at (l / inner) * outer + (l % inner) * inner_stride and
// its elements are `stride` apart. Units are elements (float2 for complex arrays).
struct Lines {
  long count;
  long inner;
  long outer;
  long inner_stride;
  long stride;
};

// Lines along axis 2 (x), 1 (y) or 0 (z) of an (d0, d1, d2) array with rows `pitch` elements
// apart and planes d1 * pitch apart.
Lines lines_along(int axis, int d0, int d1, int d2, long pitch);

// Complex transform of every line (in place when in == out). Output element (line, i) is
// multiplied by `scale` and, with alt_mod > 0, by (-1)^(i + line % alt_mod).
void fft_lines(const float2* in, float2* out, int n, const Lines& l, bool inverse, float scale,
               int alt_mod, hipStream_t s);

// Elementwise operation fused into the real input of a row transform.
enum class RealIn : int {
  plain = 0,     // x
  nonneg = 1,    // max(x, 0)
  momentum = 2,  // x + m (x - x2)
};
// Elementwise operation fused into the real output of an inverse row transform.
enum class RealOut : int {
  scale = 0,  // v = y * scale
  index = 1,  // refractive index of v = y * scale (real potential)
};

struct RealOp {
  RealIn in = RealIn::plain;
  RealOut out = RealOut::scale;
  float a = 1.0f;  // momentum m, or output scale
  float nm2 = 0.0f;
  float inv_k02 = 0.0f;
};

// Real rows (rows x n, pitch in floats) -> half-spectrum rows (n/2 + 1 bins, pitch in float2).
// x2 is the second operand of RealIn::momentum.
void rfft_rows(const float* x, const float* x2, long x_pitch, float2* out, long out_pitch,
               long rows, int n, const RealOp& op, hipStream_t s);
// Half-spectrum rows -> real rows; the imaginary parts of bins 0 and n/2 are ignored.
void irfft_rows(const float2* in, long in_pitch, float* x, long x_pitch, long rows, int n,
                const RealOp& op, hipStream_t s);

// Half-spectrum solver step along z: forward transform of each z line of e, restore
// e <- a + (1 - hits / 2) e, inverse transform. e, a, hits are (nz, ny, h) arrays, or `batch`
// of them one after another. a must be +0 (both parts) wherever hits is 0, as the folded
// Hermitian data is: lengths 256 and 384 do not read a there.
void z_roundtrip_restore(float2* e, const float2* a, const std::uint8_t* hits, int nz, int ny,
                         int h, hipStream_t s, int batch = 1);

// 3D real <-> half-complex transforms over dense (nz, ny, nx) real arrays and (nz, ny, nx/2+1)
// complex arrays. The inverse scales by `scale` and may overwrite its input.
void rfft3(const float* x, float2* out, int nz, int ny, int nx, hipStream_t s);
void irfft3(float2* in, float* x, int nz, int ny, int nx, float scale, hipStream_t s);
// 3D complex transform over the axes in `axes` (bit 0: z, bit 1: y, bit 2: x).
void fft3(float2* data, int nz, int ny, int nx, unsigned axes, bool inverse, float scale,
          hipStream_t s);

}  // namespace qh
