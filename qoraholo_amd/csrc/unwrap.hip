// Least-squares phase unwrapping (Ghiglia & Romero) and the Rytov field on the device.
//
// rho = div of the wrapped phase differences; DCT-II along x (rows), then one column kernel
// that does the DCT-II along y, the division by the Neumann eigenvalues and the exact inverse
// DCT along y in LDS; the inverse DCT along x; then the congruence constant from per-image sums
// of exp(i (psi - phi)) in double, reduced in a fixed order. The DCTs follow Makhoul: reorder,
// complex FFT, twiddle exp(-i pi k / 2n); the inverse is V_k = exp(i pi k / 2n)(X_k - i X_{n-k}),
// inverse FFT, reorder back, / n (reference src/cpu/fft.cpp:101-133, src/cpu/unwrap.cpp:11-82).
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "fft.hpp"
#include "runtime.hpp"
#include "stages.hpp"

namespace qh {

namespace {

constexpr int kThreads = 256;

__global__ void k_divergence(const float* psi, float* rho, long total, int ny, int nx) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int x = static_cast<int>(i % nx);
  const int y = static_cast<int>((i / nx) % ny);
  const float p = psi[i];
  float r = 0.0f;
  if (x + 1 < nx) r += wrap_phase(psi[i + 1] - p);
  if (x > 0) r -= wrap_phase(p - psi[i - 1]);
  if (y + 1 < ny) r += wrap_phase(psi[i + nx] - p);
  if (y > 0) r -= wrap_phase(p - psi[i - nx]);
  rho[i] = r;
}

struct DctArgs {
  float* a;
  long rows;
  int n;
  int rows_per_block;
  FftFactors f;
  const float2* tw;
  const float2* dtw;  // exp(-i pi k / (2 n))
  float inv_n;
};

// DCT-II (FORWARD) or its exact inverse along contiguous rows of length n.
template <bool FORWARD>
__global__ void __launch_bounds__(kThreads) k_dct_rows(DctArgs a) {
  extern __shared__ float2 smem[];
  const int n = a.n;
  float2* tw = smem;
  float2* dtw = tw + n;
  float2* b0 = dtw + n;
  float2* b1 = b0 + a.rows_per_block * n;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    tw[i] = a.tw[i];
    dtw[i] = a.dtw[i];
  }
  const long r0 = static_cast<long>(blockIdx.x) * a.rows_per_block;
  const int nl = static_cast<int>(min(static_cast<long>(a.rows_per_block), a.rows - r0));
  __syncthreads();
  for (int e = threadIdx.x; e < nl * n; e += blockDim.x) {
    const int line = e / n;
    const int i = e - line * n;
    const float* row = a.a + (r0 + line) * n;
    if (FORWARD) {
      b0[line * n + dct_position(i, n)] = {row[i], 0.0f};
    } else {
      const float mirror = i == 0 ? 0.0f : row[n - i];
      b0[line * n + i] = cmul(cconj(dtw[i]), float2{row[i], -mirror});
    }
  }
  __syncthreads();
  const float2* r = lds_fft(b0, b1, n, 1, nl, a.f, tw, !FORWARD, false);
  for (int e = threadIdx.x; e < nl * n; e += blockDim.x) {
    const int line = e / n;
    const int i = e - line * n;
    float* row = a.a + (r0 + line) * n;
    if (FORWARD) {
      const float2 v = r[line * n + i];
      row[i] = dtw[i].x * v.x - dtw[i].y * v.y;
    } else {
      row[i] = r[line * n + dct_position(i, n)].x * a.inv_n;
    }
  }
}

struct ColArgs {
  float* a;
  long lines;  // count * nx
  int ny;
  int nx;
  int lines_per_block;
  FftFactors f;
  const float2* tw;
  const float2* dtw;
  const float* eig_y;
  const float* eig_x;
  float inv_n;
};

// Along y: DCT-II, division by the Neumann eigenvalue (DC set to 0), inverse DCT-II.
__global__ void __launch_bounds__(kThreads) k_dct_cols_poisson(ColArgs a) {
  extern __shared__ float2 smem[];
  const int n = a.ny;
  float2* tw = smem;
  float2* dtw = tw + n;
  float2* b0 = dtw + n;
  float2* b1 = b0 + a.lines_per_block * n;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    tw[i] = a.tw[i];
    dtw[i] = a.dtw[i];
  }
  const long l0 = static_cast<long>(blockIdx.x) * a.lines_per_block;
  const int nl = static_cast<int>(min(static_cast<long>(a.lines_per_block), a.lines - l0));
  const int total = nl * n;
  const auto global_at = [&](int line, int i) {
    const long l = l0 + line;
    return (l / a.nx) * static_cast<long>(n) * a.nx + (l % a.nx) + static_cast<long>(i) * a.nx;
  };
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int i = e / nl;
    const int line = e - i * nl;
    b0[dct_position(i, n) * nl + line] = {a.a[global_at(line, i)], 0.0f};
  }
  __syncthreads();
  float2* r = lds_fft(b0, b1, 1, nl, nl, a.f, tw, false, true);
  float2* other = r == b0 ? b1 : b0;
  float* sv = reinterpret_cast<float*>(other);
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int k = e / nl;
    const int line = e - k * nl;
    const int kx = static_cast<int>((l0 + line) % a.nx);
    const float2 v = r[e];
    const float t = dtw[k].x * v.x - dtw[k].y * v.y;
    sv[e] = (k == 0 && kx == 0) ? 0.0f : t / (a.eig_y[k] + a.eig_x[kx] - 4.0f);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int k = e / nl;
    const int line = e - k * nl;
    const float mirror = k == 0 ? 0.0f : sv[(n - k) * nl + line];
    r[e] = cmul(cconj(dtw[k]), float2{sv[e], -mirror});
  }
  __syncthreads();
  const float2* r2 = lds_fft(r, other, 1, nl, nl, a.f, tw, true, true);
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int i = e / nl;
    const int line = e - i * nl;
    a.a[global_at(line, i)] = r2[dct_position(i, n) * nl + line].x * a.inv_n;
  }
}

__device__ double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if (static_cast<int>(threadIdx.x) < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  const double s = red[0];
  __syncthreads();
  return s;
}

// Per (block, image) partial sums of exp(i (psi - phi)) in double.
__global__ void __launch_bounds__(kThreads) k_congruence_partial(const float* psi, const float* phi,
                                                                 long plane, double2* partial) {
  __shared__ double red[kThreads];
  const int z = static_cast<int>(blockIdx.y);
  double c = 0.0;
  double sn = 0.0;
  for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < plane;
       i += static_cast<long>(gridDim.x) * blockDim.x) {
    const long j = z * plane + i;
    double s1, c1;
    sincos(static_cast<double>(psi[j] - phi[j]), &s1, &c1);
    c += c1;
    sn += s1;
  }
  c = block_sum(c, red);
  sn = block_sum(sn, red);
  if (threadIdx.x == 0) partial[static_cast<long>(z) * gridDim.x + blockIdx.x] = {c, sn};
}

// One block per image: congruence constant c = arg(sum) and, for Rytov, the 2 pi multiple
// nearest to the mean of (phi + c) on the border rows and columns.
__global__ void __launch_bounds__(kThreads) k_congruence_finish(const double2* partial, int nblk,
                                                                const float* phi, int ny, int nx,
                                                                int border, float2* shift) {
  __shared__ double red[kThreads];
  __shared__ float sc;
  const int z = static_cast<int>(blockIdx.x);
  if (threadIdx.x == 0) {
    double c = 0.0, s = 0.0;
    for (int b = 0; b < nblk; ++b) {
      c += partial[static_cast<long>(z) * nblk + b].x;
      s += partial[static_cast<long>(z) * nblk + b].y;
    }
    sc = static_cast<float>(atan2(s, c));
  }
  __syncthreads();
  const float c = sc;
  float sh = 0.0f;
  if (border) {
    const float* im = phi + static_cast<long>(z) * ny * nx;
    double acc = 0.0;
    const int pairs = nx + (ny - 2);
    for (int p = threadIdx.x; p < pairs; p += blockDim.x) {
      float pair;
      if (p < nx) {
        pair = (im[p] + c) + (im[static_cast<long>(ny - 1) * nx + p] + c);
      } else {
        const int y = 1 + (p - nx);
        pair = (im[static_cast<long>(y) * nx] + c) + (im[static_cast<long>(y) * nx + nx - 1] + c);
      }
      acc += pair;
    }
    const double sum = block_sum(acc, red);
    const double mean = sum / (2.0 * nx + 2.0 * (ny - 2));
    const double two_pi = 6.283185307179586;
    sh = static_cast<float>(two_pi * round(mean / two_pi));
  }
  if (threadIdx.x == 0) shift[z] = {c, sh};
}

__global__ void k_add_congruence(const float* phi, const float2* shift, long plane, long total,
                                 float* out) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  out[i] = phi[i] + shift[i / plane].x;
}

__global__ void k_assemble_rytov(const float* logamp, const float* phi, const float2* shift,
                                 long plane, long total, float2* out) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const float2 cs = shift[i / plane];
  out[i] = {logamp[i], (phi[i] + cs.x) - cs.y};
}

int rows_per_block(int n) { return std::max(1, std::min(64, 2048 / n)); }

}  // namespace

Unwrapper::Unwrapper(int count, int ny, int nx)
    : count_(count), ny_(ny), nx_(nx),
      nblk_(std::max(1, std::min(64, static_cast<int>((static_cast<long>(ny) * nx + 255) / 256)))),
      phi_(static_cast<std::size_t>(count) * ny * nx),
      partial_(static_cast<std::size_t>(count) * nblk_),
      shift_(count),
      eig_y_(ny),
      eig_x_(nx) {
  require(fft_supported(ny) && fft_supported(nx), "unwrap: image sides must factor into 2, 3, 5");
  const auto upload = [](DeviceBuffer<float>& d, int n) {
    // 2 cos(pi k / n) in float on the host, the terms of the Neumann eigenvalue.
    std::vector<float> h(n);
    for (int k = 0; k < n; ++k) h[k] = 2.0f * cosf(kPi * static_cast<float>(k) / static_cast<float>(n));
    check(hipMemcpy(d.get(), h.data(), n * sizeof(float), hipMemcpyHostToDevice), "upload eigenvalues");
  };
  upload(eig_y_, ny);
  upload(eig_x_, nx);
}

std::size_t Unwrapper::bytes() const {
  return phi_.bytes() + partial_.bytes() + shift_.bytes() + eig_y_.bytes() + eig_x_.bytes();
}

void Unwrapper::solve(const float* psi, hipStream_t s, bool border, int images) {
  require(images >= 1 && images <= count_, "unwrap: image count beyond the capacity");
  const long plane = static_cast<long>(ny_) * nx_;
  const long total = plane * images;
  float* phi = phi_.get();
  hipLaunchKernelGGL(k_divergence, dim3(blocks_for(total)), dim3(256), 0, s, psi, phi, total, ny_, nx_);
  check_launch("divergence");
  for (int pass = 0; pass < 2; ++pass) {
    const bool forward = pass == 0;
    if (!forward) {
      ColArgs c{phi, static_cast<long>(images) * nx_, ny_, nx_,
                std::max(1, std::min(16, 2048 / ny_)), fft_factors(ny_), fft_twiddles(ny_),
                dct_twiddles(ny_), eig_y_.get(), eig_x_.get(), 1.0f / static_cast<float>(ny_)};
      const std::size_t lds = (2 * ny_ + 2 * static_cast<std::size_t>(c.lines_per_block) * ny_) * sizeof(float2);
      const unsigned grid = static_cast<unsigned>((c.lines + c.lines_per_block - 1) / c.lines_per_block);
      hipLaunchKernelGGL(k_dct_cols_poisson, dim3(grid), dim3(kThreads), lds, s, c);
      check_launch("dct_cols_poisson");
    }
    DctArgs d{phi, static_cast<long>(images) * ny_, nx_, rows_per_block(nx_), fft_factors(nx_),
              fft_twiddles(nx_), dct_twiddles(nx_), 1.0f / static_cast<float>(nx_)};
    const std::size_t lds = (2 * nx_ + 2 * static_cast<std::size_t>(d.rows_per_block) * nx_) * sizeof(float2);
    const unsigned grid = static_cast<unsigned>((d.rows + d.rows_per_block - 1) / d.rows_per_block);
    if (forward) {
      hipLaunchKernelGGL(k_dct_rows<true>, dim3(grid), dim3(kThreads), lds, s, d);
    } else {
      hipLaunchKernelGGL(k_dct_rows<false>, dim3(grid), dim3(kThreads), lds, s, d);
    }
    check_launch("dct_rows");
  }
  hipLaunchKernelGGL(k_congruence_partial, dim3(nblk_, images), dim3(kThreads), 0, s, psi, phi,
                     plane, partial_.get());
  check_launch("congruence_partial");
  hipLaunchKernelGGL(k_congruence_finish, dim3(images), dim3(kThreads), 0, s, partial_.get(),
                     nblk_, phi, ny_, nx_, border ? 1 : 0, shift_.get());
  check_launch("congruence_finish");
}

void Unwrapper::unwrap(const float* psi, float* out, hipStream_t s, int images) {
  if (images == 0) images = count_;
  solve(psi, s, false, images);
  const long plane = static_cast<long>(ny_) * nx_;
  const long total = plane * images;
  hipLaunchKernelGGL(k_add_congruence, dim3(blocks_for(total)), dim3(256), 0, s, phi_.get(),
                     shift_.get(), plane, total, out);
  check_launch("add_congruence");
}

void Unwrapper::rytov(const float* psi, const float* logamp, float2* out, hipStream_t s,
                      int images) {
  if (images == 0) images = count_;
  solve(psi, s, true, images);
  const long plane = static_cast<long>(ny_) * nx_;
  const long total = plane * images;
  hipLaunchKernelGGL(k_assemble_rytov, dim3(blocks_for(total)), dim3(256), 0, s, logamp,
                     phi_.get(), shift_.get(), plane, total, out);
  check_launch("assemble_rytov");
}

}  // namespace qh
