// Index and physics expressions evaluated on the host and on gfx950. Every float expression
// keeps the operand order of the reference's shared helpers (src/common.hpp:19-183), and the
// library is compiled with -ffp-contract=off, so no multiply-add is fused and the integer
// results that follow from float arithmetic (Ewald bins) match the host bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#define QH_HD __host__ __device__ __forceinline__

namespace qh {

inline constexpr float kPi = 3.14159265358979323846f;

// (-1)^i, src/common.hpp:19.
QH_HD float parity(int i) { return (i & 1) ? -1.0f : 1.0f; }

// Signed frequency of DFT bin k of an n-point transform; bin n/2 is negative. common.hpp:22.
QH_HD int signed_bin(int k, int n) { return k < n / 2 ? k : k - n; }

// Makhoul position of sample i in a DCT-II of length n. common.hpp:25.
QH_HD int dct_position(int i, int n) { return (i & 1) ? n - 1 - (i >> 1) : (i >> 1); }

// Bin of -k. common.hpp:70.
QH_HD int negative_bin(int k, int n) { return k == 0 ? 0 : n - k; }

// Wrapped phase in [-pi, pi). common.hpp:33.
QH_HD float wrap_phase(float p) {
  return p - 2.0f * kPi * floorf((p + kPi) / (2.0f * kPi));
}

// k-space grid of the Ewald mapping; frequencies in cycles per micrometre. common.hpp:36-44.
struct OdtGrid {
  int n;
  int nz;
  float df;
  float dfz;
  float fm2;
  float fna2;
  float gain;
};

// Nearest axial bin of lateral bin (ky, kx) for an illumination (fiy, fix, fiz); false outside
// the detection pupil or the volume. common.hpp:56-67.
QH_HD bool ewald_target(const OdtGrid& g, int ky, int kx, float fiy, float fix, float fiz,
                        int& kz, float& nuz) {
  const float fy = static_cast<float>(signed_bin(ky, g.n)) * g.df + fiy;
  const float fx = static_cast<float>(signed_bin(kx, g.n)) * g.df + fix;
  const float rho2 = fy * fy + fx * fx;
  if (rho2 > g.fna2 || rho2 >= g.fm2) return false;
  nuz = sqrtf(g.fm2 - rho2);
  const int k = static_cast<int>(floorf((nuz - fiz) / g.dfz + 0.5f));
  if (k < -g.nz / 2 || k >= g.nz / 2) return false;
  kz = k < 0 ? k + g.nz : k;
  return true;
}

// Hermitian part of the mapped data and the half-spectrum restore of the real GP/FISTA loop.
// common.hpp:79-87.
QH_HD void hermitian_data(float pre, float pim, float nre, float nim, float& are, float& aim) {
  are = 0.5f * (pre + nre);
  aim = 0.5f * (pim - nim);
}
QH_HD void hermitian_restore(float are, float aim, int hits, float& re, float& im) {
  const float b = 1.0f - 0.5f * static_cast<float>(hits);
  re = are + b * re;
  im = aim + b * im;
}

// Primal point of the TV dual: max(b - lambda L(p), 0). common.hpp:121-124.
QH_HD float tv_primal(float b, float lambda, float adjoint) {
  const float x = b - lambda * adjoint;
  return x > 0.0f ? x : 0.0f;
}

// One FGP dual step at a voxel: q = r + step g, p' = q / max(1, |q|), r' = p' + m (p' - p).
// common.hpp:128-141.
QH_HD void tv_dual_step(const float g[3], float step, float momentum, float p[3], float r[3]) {
  float q[3];
  float norm2 = 0.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    q[a] = r[a] + step * g[a];
    norm2 += q[a] * q[a];
  }
  const float scale = norm2 > 1.0f ? 1.0f / sqrtf(norm2) : 1.0f;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float next = q[a] * scale;
    r[a] = next + momentum * (next - p[a]);
    p[a] = next;
  }
}

// t_{k+1} of FISTA and FGP. common.hpp:153.
QH_HD float fista_t(float t) { return 0.5f * (1.0f + sqrtf(1.0f + 4.0f * t * t)); }

// Real part of the principal sqrt(n_m^2 + O / k0^2). common.hpp:179-183.
QH_HD float index_from_potential(float re, float im, float nm2, float inv_k02) {
  const float a = nm2 + re * inv_k02;
  const float b = im * inv_k02;
  return sqrtf(0.5f * (hypotf(a, b) + a));
}

// One slice of the multi-slice BPM on a spectrum bin and its adjoint. common.hpp:158-176.
QH_HD void bpm_step(float cre, float cim, float q, float are, float aim, float& sre, float& sim) {
  const float tre = q * are + (1.0f - q) * sre;
  const float tim = q * aim + (1.0f - q) * sim;
  sre = cre * tre - cim * tim;
  sim = cre * tim + cim * tre;
}
QH_HD void bpm_step_adjoint(float cre, float cim, float q, float& gre, float& gim, float& are,
                            float& aim) {
  const float ure = cre * gre + cim * gim;
  const float uim = cre * gim - cim * gre;
  are = q * ure;
  aim = q * uim;
  gre = (1.0f - q) * ure;
  gim = (1.0f - q) * uim;
}

}  // namespace qh
