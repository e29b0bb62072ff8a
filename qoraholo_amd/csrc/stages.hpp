// Device stages of the reconstruction chain: field extraction, unwrapping and Rytov, and the
// ODT solver. Each stage owns its device buffers; callers pass device pointers and a stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <memory>
#include <vector>

#include "common.hpp"
#include "runtime.hpp"
#include "tv.hpp"

namespace qh {

// Integer crop centre, radius and output side of field extraction (reference backends.hpp:20-25).
struct CropPlan {
  int cy;
  int cx;
  float radius;
  int out;
};

// ---- field extraction (src/cpu/field.cpp:78-121 on the device) ----

// Half spectra (count, n, n/2+1) of real holograms (count, n, n): row R2C then y transform.
void hologram_spectra(const float* holo, int count, int n, float2* half, hipStream_t s);
// Disc of radius r around the carrier into (count, m, m) with the carrier at m/2, times the
// per-image sign (-1)^(cy+cx); bins with kx < 0 are read as conj(H(-ky, -kx)).
void crop_fields(const float2* half, int count, int n, const CropPlan& p, float2* fields,
                 hipStream_t s);
// Inverse 2D transform of (count, m, m) fields times scale, optionally demodulated by
// (-1)^(x+y) (the centre-origin output of the per-call path).
void fields_inverse(float2* fields, int count, int m, float scale, bool demodulate, hipStream_t s);
// psi = arg(s / b), logamp = log|s / b| with r = s conj(b) / |b|^2. The background holds
// `period` values (0: total), repeated over the samples: one stack for several requests.
void ratio_phase(const float2* sample, const float2* background, long total, float* psi,
                 float* logamp, hipStream_t s, long period = 0);
// psi = arg r, logamp = log|r| of a given ratio.
void split_ratio(const float2* ratio, long total, float* psi, float* logamp, hipStream_t s);
// Sideband peak of each hologram of a stack (src/cpu/field.cpp:35-49 on the device): integer
// offset from DC, magnitude, and the centred spectrum at the peak and its four neighbours
// (centre, y-1, y+1, x-1, x+1) for the sub-bin refinement on the host. half: (count, n, n/2+1).
struct SidebandPeak {
  int iy;
  int ix;
  float mag;
  int pad;
  float2 s[5];
};
void sideband_peaks(const float* holo, int count, int n, float dc_exclusion, float2* half,
                    SidebandPeak* peaks, hipStream_t s);
// Plain 2D transform of (count, n, n) complex images in place.
void fft2_images(float2* data, int count, int n, bool inverse, hipStream_t s);

// ---- unwrapping and Rytov (src/cpu/unwrap.cpp:11-82 on the device) ----

class Unwrapper {
 public:
  Unwrapper(int count, int ny, int nx);
  // Least-squares unwrap of psi into phi, congruent to psi modulo 2 pi. `images` (0: count)
  // takes the first images of the capacity; each image is solved on its own.
  void unwrap(const float* psi, float* phi, hipStream_t s, int images = 0);
  // out = logamp + i (unwrap(psi) - 2 pi round(mean over the border / 2 pi)).
  void rytov(const float* psi, const float* logamp, float2* out, hipStream_t s, int images = 0);
  std::size_t bytes() const;

 private:
  void solve(const float* psi, hipStream_t s, bool border, int images);
  int count_, ny_, nx_;
  int nblk_;
  DeviceBuffer<float> phi_;
  DeviceBuffer<double2> partial_;
  DeviceBuffer<float2> shift_;  // per image: congruence constant c, border shift
  DeviceBuffer<float> eig_y_;   // 2 cos(pi k / ny), host-evaluated in float
  DeviceBuffer<float> eig_x_;
};

// ---- ODT solver (src/cpu/odt.cpp:124-242 on the device) ----

struct OdtConfig {
  int iterations = 0;
  int solver = 0;        // 0: Gerchberg-Papoulis, 1: FISTA-TV
  int gp_transform = 0;  // 0: r2c, 1: c2c
  float tv_weight = 1e-2f;
  int tv_inner = 10;
  float wavelength = 0.532f;
  float n_medium = 1.337f;
};

// Event marks the solver records when a clock is attached.
enum OdtMark { kFft2 = 0, kClear, kMapping, kNormalize, kSolve, kFinish, kOdtMarks };

class StageClock {
 public:
  explicit StageClock(int marks) : events_(marks) {}
  void mark(int i, hipStream_t s) { events_[i].record(s); }
  double ms(int a, int b) const { return elapsed_ms(events_[a], events_[b]); }
  Event& event(int i) { return events_[i]; }

 private:
  std::vector<Event> events_;
};

class OdtSolver {
 public:
  // `batch`: the requests run_batch takes at once (the solves on half spectra).
  OdtSolver(const OdtGrid& g, int count, const OdtConfig& c, int batch = 1);
  void set_illuminations(const float* illum_host);  // (count, 2) fy, fx
  // fields: (count, n, n) complex on the device, overwritten by their spectra. The index volume
  // (nz, n, n) is left in index(). Marks go to clock[first + OdtMark] when given.
  void run(float2* fields, hipStream_t s, StageClock* clock = nullptr, int first = 0);
  // Whether run_batch applies: the Gerchberg-Papoulis and FISTA-TV solves on half spectra.
  bool batchable() const { return real_; }
  // nb <= batch requests whose fields are `field_stride` elements apart: each is mapped on its
  // own, then their solves run together, every launch over all nb volumes (the same operations
  // per line and per row as run(), so each volume has the bits run() gives it). The index
  // volumes go to out, (nb, nz, n, n).
  void run_batch(float2* fields, long field_stride, int nb, float* out, hipStream_t s);
  // Normalised k-space potential and hit counts after the mapping (for intermediate checks).
  void map_only(float2* fields, hipStream_t s);
  const float* index() const { return ri_.get(); }
  // Where the next run() writes the index volume: a caller's device buffer of (nz, n, n)
  // floats, or null for index().
  void set_output(float* out) { out_ = out; }
  const float2* potential() const { return potential_.get(); }
  const std::uint16_t* counts() const { return counts_.get(); }
  const unsigned long long* filled_device() const { return filled_.get(); }
  std::size_t bytes() const;
  int nz() const { return g_.nz; }
  int n() const { return g_.n; }
  // Sum of the FGP step durations of the last run and the number of steps, when timed.
  double fgp_ms() const { return fgp_ms_; }
  // Reads the FGP events of the last run; call after the stream has finished.
  double collect_fgp_ms();
  int fgp_steps() const { return fgp_steps_; }
  int fgp_launches() const { return fgp_launches_; }
  void time_fgp(bool on) { time_fgp_ = on; }
  bool timing_fgp() const { return time_fgp_; }

 private:
  void map(float2* fields, hipStream_t s, StageClock* clock, int first, bool fold, int slot = 0);
  OdtGrid g_;
  int count_;
  int batch_;
  OdtConfig c_;
  bool real_;
  bool tv_;
  long vol_;
  long half_;
  float nm2_, inv_k02_, lambda_;
  DeviceBuffer<float4> illum_;
  DeviceBuffer<float2> potential_;
  DeviceBuffer<std::uint16_t> counts_;
  DeviceBuffer<float2> measured_;
  DeviceBuffer<float2> a_;
  DeviceBuffer<std::uint8_t> hits_;
  DeviceBuffer<float2> e_;
  DeviceBuffer<float> x_[3];
  DeviceBuffer<float> ri_;
  float* out_ = nullptr;
  DeviceBuffer<unsigned long long> filled_;
  std::unique_ptr<TvState> tv_state_;
  std::vector<Event> fgp_events_;
  bool time_fgp_ = false;
  double fgp_ms_ = 0;
  int fgp_steps_ = 0;
  int fgp_launches_ = 0;
};

}  // namespace qh
