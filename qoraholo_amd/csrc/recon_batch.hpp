// A reconstruction slot of the served loop that runs several requests of one geometry together:
// the front end (hologram transforms, crop, field inverse, ratio, unwrap, Rytov) takes every
// hologram stack of the batch in one set of launches, then each request's Ewald mapping runs in
// turn, and the Gerchberg-Papoulis or FISTA-TV solves of all of them run as one (every
// transform, elementwise and TV launch over all the batch's volumes). The device work of a batch of b requests is one
// hipGraph, captured for every b up to the capacity when the slot is built. Every result equals
// what the request gets alone on a Pipeline, bit for bit: each image and each volume goes
// through the same kernels with the same operations, only more of them per launch.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "../../include/qoraholo_hip.h"
#include "runtime.hpp"
#include "stages.hpp"

namespace qh {

class ReconBatch {
 public:
  // Background holograms (count, N, N) in host memory; their fields are kept on the device.
  ReconBatch(const qh_pipeline_options& o, const float* host_backgrounds, int capacity);
  int capacity() const { return cap_; }
  // inputs[b]: holograms (count, N, N), outputs[b]: index volume (nz, n, n), host memory, for
  // b < nb <= capacity. Returns with every output complete.
  void run(const float* const* inputs, float* const* outputs, int nb);
  double filled() const;
  // Batch sizes whose device work replays a captured graph (all of them unless a capture
  // failed, which leaves kernel-by-kernel launches with the same results).
  bool graphs_ready() const { return graphs_ok_; }

 private:
  void device_work(int nb, hipStream_t s);

  int count_, n_, m_, cap_;
  CropPlan crop_;
  Stream stream_;
  Unwrapper unwrapper_;
  OdtSolver solver_;
  DeviceBuffer<float> holo_;
  DeviceBuffer<float2> half_;
  DeviceBuffer<float2> background_;
  DeviceBuffer<float2> fields_;
  DeviceBuffer<float> psi_;
  DeviceBuffer<float> logamp_;
  DeviceBuffer<float> ri_;
  PinnedBuffer<unsigned long long> filled_;
  std::vector<std::unique_ptr<GraphExec>> graphs_;  // graphs_[b - 1]: a batch of b
  bool graphs_ok_ = false;
};

}  // namespace qh
