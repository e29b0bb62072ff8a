// One BPM inversion request on device-resident state, as the reference's recipe runs it
// (examples/bpm_figure.py:74-88): Rytov fields of the measured field ratios, a Gerchberg-Papoulis
// start from them, then the proximal-gradient BPM solve (src/cpu/bpm.cpp:191-243) from that
// start. Tables, plans and every work buffer are allocated by the constructor; run() allocates
// nothing, so several inversions run side by side on their own streams.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "bpm.hpp"
#include "runtime.hpp"
#include "stages.hpp"

namespace qh {

struct BpmInversionSpec {
  BpmTables tables;
  int chunk = 0;        // illuminations per device pass (0: chosen from the free memory)
  int checkpoint = -1;  // slices per recomputed adjoint segment (-1: chosen from the free memory)
  OdtGrid grid;         // the start's reconstruction grid (n, nz of the BPM volume)
  OdtConfig start;      // Gerchberg-Papoulis start
  std::vector<float> illuminations;  // (count, 2): fy, fx
  BpmSolveConfig solve;
  float n_medium = 1.337f;
};

class BpmInversion {
 public:
  explicit BpmInversion(const BpmInversionSpec& spec);
  BpmInversion(const BpmInversion&) = delete;
  BpmInversion& operator=(const BpmInversion&) = delete;

  // Field ratios (count, n, n) complex in host memory -> refractive index (nz, n, n) in host
  // memory; loss (host, may be null) receives one data-term value per iteration. Blocks until
  // the result is in ri.
  void run(const float2* fields, float* ri, double* loss);
  int n() const { return model_.n(); }
  int nz() const { return model_.nz(); }
  int count() const { return model_.count(); }
  int chunk() const { return model_.chunk(); }
  int checkpoint() const { return model_.checkpoint(); }

 private:
  Stream stream_;
  float n_medium_;
  long plane_;
  BpmModel model_;
  BpmSolver solver_;
  OdtSolver start_;
  Unwrapper unwrap_;
  DeviceBuffer<float2> fields_, work_;
  DeviceBuffer<float> psi_, logamp_, x_;
};

}  // namespace qh
