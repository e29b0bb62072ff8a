// Ewald-sphere mapping and the missing-cone solvers (Gerchberg-Papoulis, FISTA-TV).
//
// Mapping is a deterministic gather: one thread owns a lateral column (ky, kx) and walks the
// illuminations in order a = 0..A-1, adding into its own column of the zeroed volume, so every
// voxel receives its values in the reference's order (src/cpu/odt.cpp:147-161) and the counts
// are exact integers. No float atomics anywhere. The real solvers fold the normalised data
// into its Hermitian part on the half spectrum in one pass (normalise + hermitian_data + hits
// + filled count), then iterate on half spectra with the restore fused into the z transforms.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "fft.hpp"
#include "runtime.hpp"
#include "stages.hpp"

namespace qh {

namespace {

__global__ void k_clear(float2* potential, std::uint16_t* counts, long total) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  potential[i] = {0.0f, 0.0f};
  counts[i] = 0;
}

// A thread's column is zero before its first contribution (k_clear), so the first value a
// voxel receives is written without reading the voxel back; only a voxel hit again (another
// illumination landing on the same axial bin) is read, added to and written. A bit per axial
// bin, in LDS, records which voxels of the column already hold a value. The spectra loads do
// not depend on earlier iterations, so several are in flight at once (with the stores to the
// volume, which they do not alias).
__global__ void __launch_bounds__(64) k_ewald_gather(const float2* __restrict__ spectra,
                                                     const float4* __restrict__ illum, int count,
                                                     OdtGrid g, float2* __restrict__ potential,
                                                     std::uint16_t* __restrict__ counts) {
  extern __shared__ unsigned seen[];  // (nz + 31) / 32 words per thread, word-major
  const int t = static_cast<int>(threadIdx.x);
  const int words = (g.nz + 31) >> 5;
  for (int w = 0; w < words; ++w) seen[w * 64 + t] = 0u;
  const long plane = static_cast<long>(g.n) * g.n;
  const long col = static_cast<long>(blockIdx.x) * blockDim.x + t;
  if (col >= plane) return;
  const int ky = static_cast<int>(col / g.n);
  const int kx = static_cast<int>(col % g.n);
#pragma unroll 4
  for (int a = 0; a < count; ++a) {
    const float4 il = illum[a];
    int kz = 0;
    float nuz = 0.0f;
    if (!ewald_target(g, ky, kx, il.x, il.y, il.z, kz, nuz)) continue;
    const float2 s = spectra[a * plane + col];
    // (0, -gain nu_z) * s, then times (-1)^kz; the same float operations as the reference.
    const float b = -g.gain * nuz;
    const float sign = parity(kz);
    const long i = kz * plane + col;
    unsigned& word = seen[(kz >> 5) * 64 + t];
    const unsigned bit = 1u << (kz & 31);
    float2 p{0.0f, 0.0f};
    unsigned c = 0;
    if (word & bit) {
      p = potential[i];
      c = counts[i];
    }
    word |= bit;
    p.x += sign * (0.0f * s.x - b * s.y);
    p.y += sign * (0.0f * s.y + b * s.x);
    potential[i] = p;
    counts[i] = static_cast<std::uint16_t>(c + 1);
  }
}

// Blocks of the counting kernels: 8 per CU of 256 threads, a few thousand global atomics.
constexpr unsigned kCountBlocks = 2048;

// Adds c (0, 1 or 2) of every lane of the wave to *count with one LDS atomic per wave (one per
// lane serialised the whole block on a single address). All lanes of the wave call it.
__device__ __forceinline__ void add_wave_count(unsigned int* count, unsigned c) {
  const unsigned long long one = __ballot(c >= 1u);
  const unsigned long long two = __ballot(c >= 2u);
  if ((threadIdx.x & 63u) == 0) {
    const unsigned w = static_cast<unsigned>(__popcll(one) + __popcll(two));
    if (w) atomicAdd(count, w);
  }
}

// Divides every voxel with hits by its count and counts the filled voxels.
__global__ void k_normalize(float2* potential, const std::uint16_t* counts, long total,
                            unsigned long long* filled) {
  __shared__ unsigned int block_count;
  if (threadIdx.x == 0) block_count = 0;
  __syncthreads();
  // Grid-stride: a bounded grid keeps the global atomics (one per block, on a single address)
  // few; one per block of 256 voxels had serialised tens of thousands of them.
  const long stride = static_cast<long>(gridDim.x) * blockDim.x;
  for (long base = static_cast<long>(blockIdx.x) * blockDim.x; base < total; base += stride) {
    const long i = base + threadIdx.x;
    const bool hit = i < total && counts[i] > 0;
    if (hit) {
      const float w = static_cast<float>(counts[i]);
      float2 p = potential[i];
      potential[i] = {p.x / w, p.y / w};
    }
    add_wave_count(&block_count, hit ? 1u : 0u);
  }
  __syncthreads();
  if (threadIdx.x == 0 && block_count > 0) atomicAdd(filled, static_cast<unsigned long long>(block_count));
}

// Half spectrum (nz, n, n/2+1): a(k) = (P(k) + conj P(-k)) / 2 with P the normalised data,
// hits = [w(k) > 0] + [w(-k) > 0]. Every voxel of the full volume is visited once as k or as
// -k, which gives the filled count (integer atomics, exact and order-free).
__global__ void k_fold_half(const float2* potential, const std::uint16_t* counts, int nz, int n,
                            float2* a, std::uint8_t* hits, unsigned long long* filled) {
  __shared__ unsigned int block_count;
  if (threadIdx.x == 0) block_count = 0;
  __syncthreads();
  const int h = n / 2 + 1;
  const long half = static_cast<long>(nz) * n * h;
  // Grid-stride over the half spectrum (see k_normalize).
  const long stride = static_cast<long>(gridDim.x) * blockDim.x;
  for (long base = static_cast<long>(blockIdx.x) * blockDim.x; base < half; base += stride) {
  const long o = base + threadIdx.x;
  unsigned c = 0;
  if (o < half) {
    // 32-bit index arithmetic (a half spectrum has fewer than 2^31 bins): a 64-bit division
    // per bin had cost more than the bin's loads.
    const unsigned row = static_cast<unsigned>(o) / static_cast<unsigned>(h);
    const int kx = static_cast<int>(static_cast<unsigned>(o) - row * static_cast<unsigned>(h));
    const unsigned zq = row / static_cast<unsigned>(n);
    const int ky = static_cast<int>(row - zq * static_cast<unsigned>(n));
    const int kz = static_cast<int>(zq);
    const long i = (static_cast<long>(kz) * n + ky) * n + kx;
    const long j = (static_cast<long>(negative_bin(kz, nz)) * n + negative_bin(ky, n)) * n +
                   negative_bin(kx, n);
    const unsigned wi = counts[i];
    const unsigned wj = counts[j];
    float2 pi = potential[i];
    float2 pj = potential[j];
    if (wi > 0) pi = {pi.x / static_cast<float>(wi), pi.y / static_cast<float>(wi)};
    if (wj > 0) pj = {pj.x / static_cast<float>(wj), pj.y / static_cast<float>(wj)};
    float2 v;
    hermitian_data(pi.x, pi.y, pj.x, pj.y, v.x, v.y);
    a[o] = v;
    hits[o] = static_cast<std::uint8_t>((wi > 0) + (wj > 0));
    c = wi > 0 ? 1u : 0u;
    if (kx >= 1 && kx <= n / 2 - 1 && wj > 0) ++c;
  }
  add_wave_count(&block_count, c);
  }
  __syncthreads();
  if (threadIdx.x == 0 && block_count > 0) atomicAdd(filled, static_cast<unsigned long long>(block_count));
}

__global__ void k_gp_nonneg(float2* p, long total) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const float re = p[i].x;
  p[i] = {re < 0.0f ? 0.0f : re, 0.0f};
}

__global__ void k_gp_restore_measured(float2* p, const float2* measured,
                                      const std::uint16_t* counts, long total) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  if (counts[i] > 0) p[i] = measured[i];
}

__global__ void k_index_complex(const float2* p, float* ri, long total, float nm2, float inv_k02) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  ri[i] = index_from_potential(p[i].x, p[i].y, nm2, inv_k02);
}

}  // namespace

OdtSolver::OdtSolver(const OdtGrid& g, int count, const OdtConfig& c, int batch)
    : g_(g), count_(count), batch_(batch), c_(c),
      real_(c.iterations > 0 && (c.solver == 1 || c.gp_transform == 0)),
      tv_(c.iterations > 0 && c.solver == 1),
      vol_(static_cast<long>(g.nz) * g.n * g.n),
      half_(static_cast<long>(g.nz) * g.n * (g.n / 2 + 1)) {
  require(fft_supported(g.n) && fft_supported(g.nz), "reconstruct: sizes must factor into 2, 3, 5");
  require(count <= 65535, "reconstruct: at most 65535 illuminations");
  require(half_ < (1L << 31), "reconstruct: at most 2^31 half-spectrum bins");
  require(batch >= 1, "reconstruct: the batch must be positive");
  const float k0 = 2.0f * kPi / c.wavelength;
  nm2_ = c.n_medium * c.n_medium;
  inv_k02_ = 1.0f / (k0 * k0);
  lambda_ = c.tv_weight * k0 * k0;
  illum_.resize(count);
  potential_.resize(vol_);
  counts_.resize(vol_);
  ri_.resize(vol_);
  filled_.resize(1);
  if (real_) {
    // A batched solve keeps the data, hits, spectrum, iterates and TV duals of every request of
    // the batch.
    const long nb = batch;
    a_.resize(nb * half_);
    hits_.resize(nb * half_);
    e_.resize(nb * half_);
    x_[0].resize(nb * vol_);
    if (tv_) {
      // FISTA's previous and prox-input iterates (Gerchberg-Papoulis updates x_[0] in place).
      x_[1].resize(nb * vol_);
      x_[2].resize(nb * vol_);
      tv_state_ = std::make_unique<TvState>(g.nz, g.n, g.n, batch);
      fgp_events_.resize(2 * c.iterations);
    }
  } else if (c.iterations > 0) {
    measured_.resize(vol_);
  }
}

void OdtSolver::set_illuminations(const float* illum) {
  std::vector<float4> il(count_);
  for (int a = 0; a < count_; ++a) {
    const float fy = illum[2 * a];
    const float fx = illum[2 * a + 1];
    // fiz on the host in float, as the reference evaluates it (src/cpu/odt.cpp:141-144).
    il[a] = {fy, fx, std::sqrt(g_.fm2 - fy * fy - fx * fx), 0.0f};
  }
  check(hipMemcpy(illum_.get(), il.data(), count_ * sizeof(float4), hipMemcpyHostToDevice),
        "upload illuminations");
}

std::size_t OdtSolver::bytes() const {
  std::size_t b = illum_.bytes() + potential_.bytes() + counts_.bytes() + measured_.bytes() +
                  a_.bytes() + hits_.bytes() + e_.bytes() + ri_.bytes() + filled_.bytes();
  for (const auto& x : x_) b += x.bytes();
  if (tv_state_) b += tv_state_->bytes();
  return b;
}

void OdtSolver::map(float2* fields, hipStream_t s, StageClock* clock, int first, bool fold,
                    int slot) {
  const auto mark = [&](int m) {
    if (clock) clock->mark(first + m, s);
  };
  const long plane = static_cast<long>(g_.n) * g_.n;
  fft2_images(fields, count_, g_.n, false, s);
  mark(kFft2);
  hipLaunchKernelGGL(k_clear, dim3(blocks_for(vol_)), dim3(256), 0, s, potential_.get(),
                     counts_.get(), vol_);
  check_launch("clear");
  fill(reinterpret_cast<float*>(filled_.get()), 2, 0.0f, s);
  mark(kClear);
  const std::size_t seen_bytes = static_cast<std::size_t>((g_.nz + 31) / 32) * 64 * sizeof(unsigned);
  hipLaunchKernelGGL(k_ewald_gather, dim3(blocks_for(plane, 64)), dim3(64), seen_bytes, s, fields,
                     illum_.get(), count_, g_, potential_.get(), counts_.get());
  check_launch("ewald_gather");
  mark(kMapping);
  if (fold) {
    hipLaunchKernelGGL(k_fold_half, dim3(std::min(blocks_for(half_), kCountBlocks)), dim3(256), 0, s,
                       potential_.get(),
                       counts_.get(), g_.nz, g_.n, a_.get() + slot * half_,
                       hits_.get() + slot * half_, filled_.get());
    check_launch("fold_half");
  } else {
    hipLaunchKernelGGL(k_normalize, dim3(std::min(blocks_for(vol_), kCountBlocks)), dim3(256), 0, s,
                       potential_.get(),
                       counts_.get(), vol_, filled_.get());
    check_launch("normalize");
  }
  mark(kNormalize);
}

void OdtSolver::run_batch(float2* fields, long field_stride, int nb, float* out, hipStream_t s) {
  require(batchable(), "reconstruct: only the solves on half spectra batch");
  require(nb >= 1 && nb <= batch_, "reconstruct: batch beyond the solver's capacity");
  for (int b = 0; b < nb; ++b) map(fields + b * field_stride, s, nullptr, 0, true, b);
  const int n = g_.n;
  const int nz = g_.nz;
  const int h = n / 2 + 1;
  const float inv = 1.0f / static_cast<float>(vol_);
  // Volumes lie one after another, so the row and y passes see one array of nb * nz planes;
  // the z passes take the batch.
  const long rows = static_cast<long>(nb) * nz * n;
  for (int b = 0; b < nb; ++b) {
    fft_lines(a_.get() + b * half_, e_.get() + b * half_, nz, lines_along(0, nz, n, h, h), true,
              1.0f, 0, s);
  }
  fft_lines(e_.get(), e_.get(), n, lines_along(1, nb * nz, n, h, h), true, 1.0f, 0, s);
  RealOp first_op;
  first_op.a = inv;
  irfft_rows(e_.get(), h, x_[0].get(), n, rows, n, first_op, s);
  // One gradient step of every volume, as run()'s gradient_step.
  const auto gradient_step = [&](const float* x, const float* x2, RealIn in, float m, float* v,
                                 RealOut kind) {
    RealOp op;
    op.in = in;
    op.a = m;
    rfft_rows(x, x2, n, e_.get(), h, rows, n, op, s);
    fft_lines(e_.get(), e_.get(), n, lines_along(1, nb * nz, n, h, h), false, 1.0f, 0, s);
    z_roundtrip_restore(e_.get(), a_.get(), hits_.get(), nz, n, h, s, nb);
    fft_lines(e_.get(), e_.get(), n, lines_along(1, nb * nz, n, h, h), true, 1.0f, 0, s);
    RealOp oo;
    oo.out = kind;
    oo.a = inv;
    oo.nm2 = nm2_;
    oo.inv_k02 = inv_k02_;
    irfft_rows(e_.get(), h, v, n, rows, n, oo, s);
  };
  if (!tv_) {
    for (int it = 0; it < c_.iterations; ++it) {
      const bool last = it == c_.iterations - 1;
      gradient_step(x_[0].get(), nullptr, RealIn::nonneg, 0.0f, last ? out : x_[0].get(),
                    last ? RealOut::index : RealOut::scale);
    }
    return;
  }
  // FISTA-TV as run() does it, every step over the batch; the TV prox takes the batch too.
  tv_state_->reset(s, nb);
  int cur = 0;
  int prev = 0;
  float t = 1.0f;
  float m = 0.0f;
  float* v = x_[2].get();
  for (int it = 0; it < c_.iterations; ++it) {
    const bool last = it == c_.iterations - 1;
    gradient_step(x_[cur].get(), x_[prev].get(), RealIn::momentum, m, v, RealOut::scale);
    const int next = prev != cur ? prev : 1 - cur;
    tv_state_->prox(v, last ? out : x_[next].get(), lambda_, c_.tv_inner, s, last, nm2_,
                    inv_k02_);
    prev = cur;
    cur = next;
    const float tn = fista_t(t);
    m = (t - 1.0f) / tn;
    t = tn;
  }
}

double OdtSolver::collect_fgp_ms() {
  fgp_ms_ = 0;
  if (fgp_steps_ == 0) return 0;
  for (int it = 0; it < c_.iterations; ++it) fgp_ms_ += elapsed_ms(fgp_events_[2 * it], fgp_events_[2 * it + 1]);
  return fgp_ms_;
}

void OdtSolver::map_only(float2* fields, hipStream_t s) { map(fields, s, nullptr, 0, false); }

void OdtSolver::run(float2* fields, hipStream_t s, StageClock* clock, int first) {
  const auto mark = [&](int m) {
    if (clock) clock->mark(first + m, s);
  };
  map(fields, s, clock, first, real_);
  float* out = out_ ? out_ : ri_.get();
  const int n = g_.n;
  const int nz = g_.nz;
  const int h = n / 2 + 1;
  const float inv = 1.0f / static_cast<float>(vol_);
  fgp_steps_ = 0;
  fgp_launches_ = 0;
  fgp_ms_ = 0;

  if (real_) {
    // x = C2R(a) / N: the first z pass reads a and writes e, so a is kept.
    fft_lines(a_.get(), e_.get(), nz, lines_along(0, nz, n, h, h), true, 1.0f, 0, s);
    fft_lines(e_.get(), e_.get(), n, lines_along(1, nz, n, h, h), true, 1.0f, 0, s);
    RealOp out_op;
    out_op.a = inv;
    irfft_rows(e_.get(), h, x_[0].get(), n, static_cast<long>(nz) * n, n, out_op, s);
    const auto gradient_step = [&](const float* x, const float* x2, RealIn in, float m, float* v,
                                   RealOut out) {
      RealOp op;
      op.in = in;
      op.a = m;
      rfft_rows(x, x2, n, e_.get(), h, static_cast<long>(nz) * n, n, op, s);
      fft_lines(e_.get(), e_.get(), n, lines_along(1, nz, n, h, h), false, 1.0f, 0, s);
      z_roundtrip_restore(e_.get(), a_.get(), hits_.get(), nz, n, h, s);
      fft_lines(e_.get(), e_.get(), n, lines_along(1, nz, n, h, h), true, 1.0f, 0, s);
      RealOp oo;
      oo.out = out;
      oo.a = inv;
      oo.nm2 = nm2_;
      oo.inv_k02 = inv_k02_;
      irfft_rows(e_.get(), h, v, n, static_cast<long>(nz) * n, n, oo, s);
    };
    if (!tv_) {
      // Gerchberg-Papoulis (src/cpu/odt.cpp:76-90): x = max(x, 0), R2C, restore, C2R / N.
      for (int it = 0; it < c_.iterations; ++it) {
        const bool last = it == c_.iterations - 1;
        gradient_step(x_[0].get(), nullptr, RealIn::nonneg, 0.0f, last ? out : x_[0].get(),
                      last ? RealOut::index : RealOut::scale);
      }
      mark(kSolve);
      mark(kFinish);
      return;
    }
    // FISTA-TV (src/cpu/odt.cpp:94-120): y = x + m (x - x_prev) is formed inside the R2C load.
    tv_state_->reset(s);
    int cur = 0;
    int prev = 0;
    float t = 1.0f;
    float m = 0.0f;
    float* v = x_[2].get();
    for (int it = 0; it < c_.iterations; ++it) {
      const bool last = it == c_.iterations - 1;
      gradient_step(x_[cur].get(), x_[prev].get(), RealIn::momentum, m, v, RealOut::scale);
      const int next = prev != cur ? prev : 1 - cur;
      tv_state_->prox(v, last ? out : x_[next].get(), lambda_, c_.tv_inner, s, last, nm2_,
                      inv_k02_, time_fgp_ ? &fgp_events_[2 * it] : nullptr,
                      time_fgp_ ? &fgp_events_[2 * it + 1] : nullptr);
      prev = cur;
      cur = next;
      const float tn = fista_t(t);
      m = (t - 1.0f) / tn;
      t = tn;
    }
    mark(kSolve);
    mark(kFinish);
    if (time_fgp_) {
      fgp_steps_ = c_.iterations * c_.tv_inner;
      fgp_launches_ = c_.iterations * tv_state_->launches();
    }
    return;
  }

  // Complex path (src/cpu/odt.cpp:208-241): plain inverse, or GP on full spectra.
  float2* p = potential_.get();
  if (c_.iterations > 0) {
    check(hipMemcpyAsync(measured_.get(), p, vol_ * sizeof(float2), hipMemcpyDeviceToDevice, s),
          "copy measured");
    fft3(p, nz, n, n, 0b111u, true, inv, s);
    for (int it = 0; it < c_.iterations; ++it) {
      hipLaunchKernelGGL(k_gp_nonneg, dim3(blocks_for(vol_)), dim3(256), 0, s, p, vol_);
      check_launch("gp_nonneg");
      fft3(p, nz, n, n, 0b111u, false, 1.0f, s);
      hipLaunchKernelGGL(k_gp_restore_measured, dim3(blocks_for(vol_)), dim3(256), 0, s, p,
                         measured_.get(), counts_.get(), vol_);
      check_launch("gp_restore");
      fft3(p, nz, n, n, 0b111u, true, inv, s);
    }
  } else {
    fft3(p, nz, n, n, 0b111u, true, inv, s);
  }
  mark(kSolve);
  hipLaunchKernelGGL(k_index_complex, dim3(blocks_for(vol_)), dim3(256), 0, s, p, out, vol_,
                     nm2_, inv_k02_);
  check_launch("index");
  mark(kFinish);
}

}  // namespace qh
