// Field extraction from off-axis holograms and the sample / background ratio.
//
// Holograms are transformed without modulation (row R2C, then columns), so the spectrum has
// its origin at bin 0 and the crop reads signed frequency f at bin f mod n, with the other
// half of the plane read as conj(H(-f)). For centre-origin holograms and fields the origin
// bookkeeping of the reference's CPU path (src/cpu/field.cpp:84-119) reduces to one sign per
// image, (-1)^(cy + cx), and, in the per-call path, a demodulation of the output by (-1)^(x+y).
#include "common.hpp"
#include "fft.hpp"
#include "runtime.hpp"
#include "stages.hpp"

namespace qh {

namespace {

__global__ void k_crop(const float2* half, int count, int n, int m, int cy, int cx, float r2,
                       float sign, float2* out) {
  const long total = static_cast<long>(count) * m * m;
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int x = static_cast<int>(i % m);
  const int y = static_cast<int>((i / m) % m);
  const int z = static_cast<int>(i / (static_cast<long>(m) * m));
  const int dy = y - m / 2;
  const int dx = x - m / 2;
  const int fy = cy + dy;
  const int fx = cx + dx;
  float2 v{0.0f, 0.0f};
  if (static_cast<float>(dy * dy + dx * dx) <= r2 && fy >= -n / 2 && fy < n / 2 && fx >= -n / 2 &&
      fx < n / 2) {
    const int h = n / 2 + 1;
    const float2* spec = half + static_cast<long>(z) * n * h;
    if (fx >= 0) {
      v = spec[static_cast<long>((fy + n) % n) * h + fx];
    } else {
      const float2 c = spec[static_cast<long>((n - fy) % n) * h - fx];
      v = {c.x, -c.y};
    }
    v = {v.x * sign, v.y * sign};
  }
  out[i] = v;
}

// b holds `period` values, repeated over the samples (one background stack for several
// requests); 32-bit indices (a batch of stacks holds fewer than 2^31 pixels).
__global__ void k_ratio_phase(const float2* s, const float2* b, long total, unsigned period,
                              float* psi, float* logamp) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const float2 u = s[i];
  const float2 v = b[static_cast<unsigned>(i) % period];
  const float d = 1.0f / (v.x * v.x + v.y * v.y);
  const float re = (u.x * v.x + u.y * v.y) * d;
  const float im = (u.y * v.x - u.x * v.y) * d;
  psi[i] = atan2f(im, re);
  logamp[i] = logf(hypotf(re, im));
}

__global__ void k_split_ratio(const float2* r, long total, float* psi, float* logamp) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const float2 v = r[i];
  psi[i] = atan2f(v.y, v.x);
  logamp[i] = logf(hypotf(v.x, v.y));
}

// Element (y, x) of the centred spectrum S = FFT((-1)^(x+y) h) read from the unmodulated half
// spectrum: S[y][x] = H((y - c) mod n, x - c), columns left of the centre as conj(H(-f)).
__device__ __forceinline__ float2 shifted(const float2* spec, int n, int y, int x) {
  const int h = n / 2 + 1;
  const int c = n / 2;
  const int fy = y - c;
  const int fx = x - c;
  if (fx >= 0) return spec[static_cast<long>((fy + n) % n) * h + fx];
  const float2 v = spec[static_cast<long>((n - fy) % n) * h - fx];
  return {v.x, -v.y};
}

// Largest |S| over y in [1, n-1), x in [n/2, n-1), outside the DC disc and the half axis
// dx = 0, dy <= 0 (src/cpu/field.cpp:35-49), one block per hologram. Each thread scans its
// candidates in row-major order keeping the first maximum; the block keeps the larger
// magnitude and, on a tie, the earlier position, so the result is the reference loop's.
__global__ void __launch_bounds__(1024) k_sideband_peak(const float2* half, int n, float dc2,
                                                        SidebandPeak* out) {
  __shared__ float sm[1024];
  __shared__ long si[1024];
  const int img = static_cast<int>(blockIdx.x);
  const int tid = static_cast<int>(threadIdx.x);
  const float2* spec = half + static_cast<long>(img) * n * (n / 2 + 1);
  const int c = n / 2;
  const int w = n - 1 - c;
  const long total = static_cast<long>(n - 2) * w;
  float best = 0.0f;
  long bi = -1;
  for (long e = tid; e < total; e += blockDim.x) {
    const int y = 1 + static_cast<int>(e / w);
    const int x = c + static_cast<int>(e % w);
    const int dy = y - c;
    const int dx = x - c;
    if (dx == 0 && dy <= 0) continue;
    if (static_cast<float>(dy * dy + dx * dx) <= dc2) continue;
    const float2 v = shifted(spec, n, y, x);
    const float m = hypotf(v.x, v.y);
    if (m > best) {
      best = m;
      bi = e;
    }
  }
  sm[tid] = best;
  si[tid] = bi;
  __syncthreads();
  for (int stride = static_cast<int>(blockDim.x) / 2; stride > 0; stride >>= 1) {
    if (tid < stride) {
      const float mo = sm[tid + stride];
      const long io = si[tid + stride];
      if (mo > sm[tid] || (mo == sm[tid] && io >= 0 && (si[tid] < 0 || io < si[tid]))) {
        sm[tid] = mo;
        si[tid] = io;
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    SidebandPeak r{};
    r.mag = sm[0];
    if (si[0] >= 0) {
      const int y = 1 + static_cast<int>(si[0] / w);
      const int x = c + static_cast<int>(si[0] % w);
      r.iy = y - c;
      r.ix = x - c;
      r.s[0] = shifted(spec, n, y, x);
      r.s[1] = shifted(spec, n, y - 1, x);
      r.s[2] = shifted(spec, n, y + 1, x);
      r.s[3] = shifted(spec, n, y, x - 1);
      r.s[4] = shifted(spec, n, y, x + 1);
    }
    out[img] = r;
  }
}

}  // namespace

void sideband_peaks(const float* holo, int count, int n, float dc_exclusion, float2* half,
                    SidebandPeak* peaks, hipStream_t s) {
  hologram_spectra(holo, count, n, half, s);
  hipLaunchKernelGGL(k_sideband_peak, dim3(count), dim3(1024), 0, s, half, n,
                     dc_exclusion * dc_exclusion, peaks);
  check_launch("sideband_peak");
}

void hologram_spectra(const float* holo, int count, int n, float2* half, hipStream_t s) {
  const int h = n / 2 + 1;
  rfft_rows(holo, nullptr, n, half, h, static_cast<long>(count) * n, n, RealOp{}, s);
  fft_lines(half, half, n, lines_along(1, count, n, h, h), false, 1.0f, 0, s);
}

void crop_fields(const float2* half, int count, int n, const CropPlan& p, float2* fields,
                 hipStream_t s) {
  const long total = static_cast<long>(count) * p.out * p.out;
  hipLaunchKernelGGL(k_crop, dim3(blocks_for(total)), dim3(256), 0, s, half, count, n, p.out,
                     p.cy, p.cx, p.radius * p.radius, parity(p.cy + p.cx), fields);
  check_launch("crop_fields");
}

void fields_inverse(float2* fields, int count, int m, float scale, bool demodulate, hipStream_t s) {
  fft_lines(fields, fields, m, lines_along(2, count, m, m, m), true, 1.0f, 0, s);
  fft_lines(fields, fields, m, lines_along(1, count, m, m, m), true, scale, demodulate ? m : 0, s);
}

void ratio_phase(const float2* sample, const float2* background, long total, float* psi,
                 float* logamp, hipStream_t s, long period) {
  if (period == 0) period = total;
  require(total < (1L << 31) && period >= 1 && total % period == 0,
          "ratio_phase: the samples must be whole background stacks");
  hipLaunchKernelGGL(k_ratio_phase, dim3(blocks_for(total)), dim3(256), 0, s, sample, background,
                     total, static_cast<unsigned>(period), psi, logamp);
  check_launch("ratio_phase");
}

void split_ratio(const float2* ratio, long total, float* psi, float* logamp, hipStream_t s) {
  hipLaunchKernelGGL(k_split_ratio, dim3(blocks_for(total)), dim3(256), 0, s, ratio, total, psi,
                     logamp);
  check_launch("split_ratio");
}

void fft2_images(float2* data, int count, int n, bool inverse, hipStream_t s) {
  fft_lines(data, data, n, lines_along(2, count, n, n, n), inverse, 1.0f, 0, s);
  fft_lines(data, data, n, lines_along(1, count, n, n, n), inverse, 1.0f, 0, s);
}

}  // namespace qh
