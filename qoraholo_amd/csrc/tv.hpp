// FGP total-variation proximal step on the device.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "runtime.hpp"

namespace qh {

// d[0, n) = v, by a kernel (safe inside captured graphs, unlike a memset node).
void fill(float* d, long n, float v, hipStream_t s);

// One fused FGP dual step over a dense (nz, ny, nx) volume: r = p_k + m (p_k - p_prev) (r = p_k
// on the first step), z = max(b - lambda L(r), 0) at each voxel and its upper neighbours,
// g = L^T z, and the projected dual update p_out = P(r + step g). p_out is a third buffer.
// With batch > 1, every array holds that many volumes one after another (b, out: nz ny nx
// floats apart; duals: three components, 3 nz ny nx floats apart), all in one launch.
void fgp_step(const float* b, const float* p_in, const float* p_prev, float* p_out, int nz,
              int ny, int nx, float lambda, float step, float momentum, bool first,
              hipStream_t s, int batch = 1);

// Two FGP steps in one pass (rows up to 256 wide): from p_k, p_{k-1} to p_{k+1} and p_{k+2},
// bit for bit what two fgp_step calls with momenta m1 then m2 give. The four buffers are
// distinct; p_k1 may be null (not written).
bool fgp_pair_supported(int nx);
void fgp_pair(const float* b, const float* p_k, const float* p_km1, float* p_k1, float* p_k2,
              int nz, int ny, int nx, float lambda, float step, float m1, bool first, float m2,
              hipStream_t s, int batch = 1);

// out = max(b - lambda L(p), 0); with to_index the refractive index of that potential.
void tv_primal(const float* b, const float* p, int nz, int ny, int nx, float lambda, float* out,
               bool to_index, float nm2, float inv_k02, hipStream_t s, int batch = 1);

// Dual state of the TV prox, kept across calls (FISTA warm start): four rotating dual buffers
// of three components each. Steps run two per pass where the row width allows (set_fused(false)
// runs them one per pass, for comparisons).
class TvState {
 public:
  // The dual buffers in device memory of its own, for up to `batch` volumes proxed together.
  TvState(int nz, int ny, int nx, int batch = 1);
  // The dual buffers in `workspace` (workspace_floats(nz, ny, nx) floats of device memory,
  // 256-byte aligned, not owned: it must outlive the state).
  TvState(int nz, int ny, int nx, float* workspace);
  static std::size_t workspace_floats(int nz, int ny, int nx);
  // Zero duals for the next `batch` volumes (at most the capacity), which every prox until the
  // next reset takes together: b and out then hold that many volumes one after another.
  void reset(hipStream_t s, int batch = 1);
  // out = prox_{lambda TV, x >= 0}(b) by `inner` FGP steps from the current dual.
  // Events t0 and t1, when given, bracket the FGP steps (not the final primal kernel).
  void prox(const float* b, float* out, float lambda, int inner, hipStream_t s,
            bool to_index = false, float nm2 = 0.0f, float inv_k02 = 0.0f, Event* t0 = nullptr,
            Event* t1 = nullptr);
  std::size_t bytes() const;
  void set_fused(bool on) { fused_ = on; }
  // Kernel passes of the last prox call.
  int launches() const { return launches_; }

 private:
  int nz_, ny_, nx_;
  long vol_;
  int cap_ = 1;    // volumes the buffers hold
  int batch_ = 1;  // volumes of the current solve
  DeviceBuffer<float> owned_;
  float* p_[4];  // dual buffers of three components each, inside owned_ or the workspace
  int cur_ = 0;
  bool fused_ = true;
  int launches_ = 0;
};

}  // namespace qh
