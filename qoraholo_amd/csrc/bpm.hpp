// Multi-slice beam propagation (split step) on the device: forward model, hand-written adjoint
// with slice checkpointing, data-term gradient and the proximal-gradient inversion.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <memory>
#include <vector>

#include "../../include/qoraholo_hip.h"
#include "runtime.hpp"
#include "tv.hpp"

namespace qh {

// Per-illumination tables (reference src/backends.hpp:61-70), computed in double on the host
// (src/bpm.cpp:10-60) and stored in float; lateral bins in FFT order, each (count, n, n).
struct BpmTables {
  int n = 0;
  int nz = 0;
  int count = 0;
  float phase = 0;            // k0 * pixel
  std::vector<float2> step;   // one-voxel propagator, zero outside the medium band
  std::vector<float2> last;   // from the last slice back to z = 0, times the detection pupil
  std::vector<float> q;       // weight of the scattered part of a slice
  float lipschitz = 0;        // of the data-term gradient for a weak object
};

BpmTables bpm_tables(int n, int nz, const qh_illumination* illuminations, int count,
                     const qh_geometry& geometry, bool obliquity);

// Device-resident tables and work buffers for one volume shape and illumination set.
// Illuminations are processed in chunks of `chunk` (0: as many as keep the adjoint's buffers
// within half the free device memory); the adjoint recomputes envelopes from a checkpoint of
// the propagation state every `checkpoint` slices (0: keep the envelope of every slice, no
// recomputation; negative: keep every envelope when the chunk of illuminations asked for, or
// all of them, then fits within half the free device memory, as src/cuda/bpm.cu:139-141 stores
// them all, else checkpoint every kBpmCheckpoint slices). The adjoint's buffers are allocated
// by the first gradient or vjp.
constexpr int kBpmCheckpoint = 16;
class BpmModel {
 public:
  BpmModel(const BpmTables& tables, int chunk, int checkpoint);
  BpmModel(const BpmModel&) = delete;
  BpmModel& operator=(const BpmModel&) = delete;

  int n() const { return n_; }
  int nz() const { return nz_; }
  int count() const { return count_; }
  int chunk() const { return chunk_; }
  // Slices per recomputed segment as resolved (0: every envelope kept).
  int checkpoint() const { return seg_len_ == nz_ ? 0 : seg_len_; }
  float phase() const { return phase_; }
  float lipschitz() const { return lipschitz_; }
  std::size_t bytes() const;
  // Allocates the adjoint's buffers now rather than on the first gradient or vjp.
  void reserve() { ensure_adjoint(); }

  // Kernel arguments, defined with the kernels.
  struct Args;

  // Contrast x (nz, n, n) -> field ratios y (count, n, n) at z = 0.
  void forward(const float* x, float2* y, hipStream_t s);
  // grad (nz, n, n) = Re(J^H r) at x, summed over all illuminations in order.
  void vjp(const float* x, const float2* r, float* grad, hipStream_t s);
  // Tangent fields y (count, n, n) = J h at x: the linearised model of bpm_jvp
  // (src/cpu/bpm.cpp:141-176), marched beside the forward model slice by slice.
  void jvp(const float* x, const float* h, float2* y, hipStream_t s);
  // grad = sum over a in [first, last) of Re(J_a^H (bpm_a(x) - fields_a)), unscaled;
  // loss_lines[a n + y] = sum over x of |bpm_a(x) - fields_a|^2 in double, per image row.
  void gradient(const float* x, const float2* fields, int first, int last, float* grad,
                double* loss_lines, hipStream_t s);

 private:
  Args base(int chunk_count) const;
  void ensure_adjoint();
  void set_screen(const float* x, hipStream_t s);
  void forward_pass(int a0, int chunk_count, bool keep, hipStream_t s);
  void recompute(int seg, int a0, int chunk_count, hipStream_t s);
  void adjoint(int a0, int chunk_count, float* grad, bool init, hipStream_t s);
  float2* ck_s(int j) const;
  float2* ck_t(int j) const;

  int n_, nz_, count_, chunk_, seg_len_, nseg_;
  long plane_;
  float phase_, lipschitz_;
  DeviceBuffer<float2> step_, last_;
  DeviceBuffer<float> q_;
  DeviceBuffer<float2> screen_;
  DeviceBuffer<float2> f_, s_, a_, g_;
  DeviceBuffer<float> p_;
  DeviceBuffer<float2> ck_s_, ck_t_, w_;
};

// x[i] *= scale (the 1 / batch factor of the gradient, src/cpu/bpm.cpp:126-127).
void bpm_scale(float* x, long total, float scale, hipStream_t s);

struct BpmSolveConfig {
  int iterations = 20;
  int batch = 0;
  float step = 1.0f;
  float tv_weight = 0.0f;
  int tv_inner = 10;
};

// out[i] = in[i] + v (in place allowed): the contrast x = ri - n_medium and back, with the
// single float rounding of the reference's host conversion (src/api.cpp:159-163).
void bpm_offset(const float* in, float* out, long total, float v, hipStream_t s);

// Reference src/cpu/bpm.cpp:191-243 on the device, with its work buffers (gradient, prox input,
// FISTA point, TV dual, per-row loss terms) allocated once, so that repeated solves allocate and
// free nothing (a free would wait for all work on the device).
class BpmSolver {
 public:
  BpmSolver(BpmModel& model, const BpmSolveConfig& c);
  BpmSolver(const BpmSolver&) = delete;
  BpmSolver& operator=(const BpmSolver&) = delete;
  // x (contrast, device) is the starting volume and receives the result; fields (count, n, n)
  // are device field ratios; loss (host, may be null) receives one data-term value per
  // iteration. Returns when the stream has finished the solve.
  void run(const float2* fields, float* x, double* loss, hipStream_t s);
  std::size_t bytes() const;

 private:
  BpmModel& m_;
  BpmSolveConfig c_;
  long total_;
  int batch_;
  DeviceBuffer<float> grad_, z_, y_, next_;
  DeviceBuffer<double> lines_;
  PinnedBuffer<double> lines_host_;
  std::unique_ptr<TvState> dual_;
};

// One solve with buffers of its own (BpmSolver constructed for the call).
void bpm_solve(BpmModel& model, const float2* fields, float* x, const BpmSolveConfig& c,
               double* loss, hipStream_t s);

}  // namespace qh
