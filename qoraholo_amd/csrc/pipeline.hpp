// Device-resident reconstruction pipeline: holograms in, refractive index out, with every
// buffer, twiddle table and background field allocated once (include/qoraholo/pipeline.hpp of
// the reference, src/cuda/pipeline.cu:12-161).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "../../include/qoraholo_hip.h"
#include "runtime.hpp"
#include "stages.hpp"

namespace qh {

class Pipeline {
 public:
  explicit Pipeline(const qh_pipeline_options& o);
  void set_background(const float* host_backgrounds);
  float* input() { return input_.get(); }
  const float* output() const { return output_.get(); }
  void process();
  // process() from and to pageable host memory: each chunk of holograms is copied into the
  // pinned input (on the copy threads) just before its upload is queued, so the copies overlap
  // the uploads and transforms of the chunks before it; the volume leaves in chunks whose host
  // copies overlap the DMA of the next chunk.
  void run(const float* holograms, float* ri);
  void process_device(const float* holo_dev, float* ri_dev, hipStream_t s);
  double filled() const;
  // Stage timings of the last call; after process_device they are read once its work is done.
  const qh_pipeline_timings& timings();
  double setup_ms() const { return setup_ms_; }
  std::size_t device_bytes() const { return device_bytes_; }
  int nz() const { return solver_.nz(); }
  int n() const { return solver_.n(); }
  void time_fgp(bool on) { solver_.time_fgp(on); }
  // Replay the device work of a request (front end and solver) as one captured graph (default)
  // or launch it kernel by kernel. Graph replays time the whole device work only (reported as
  // the solve stage); stage splits and FGP intervals come from kernel-by-kernel runs, which
  // time_fgp(true) selects.
  void use_graph(bool on) {
    use_graph_ = on;
    if (!on) graph_.reset();
  }
  // Captures the graph of a request on the pipeline's own device input now, so that the first
  // process() or run() only replays it (a server builds its slots this way, under the lock that
  // also serialises the allocations of other slots). No device work runs.
  void capture();
  bool graph_ready() const { return graph_.ready(); }
  bool has_background() const { return has_background_; }
  void set_setup(double ms, std::size_t bytes) {
    setup_ms_ = ms;
    device_bytes_ = bytes;
  }

 private:
  enum { kStart = 0, kFields, kRytov, kSolver, kDownloaded = kSolver + kOdtMarks, kMarks };
  void front_end(const float* holo_dev, hipStream_t s, bool from_host, const float* user = nullptr);
  void process_host(const float* user_in, float* user_out);
  // Queues the front end and the solver on s reading holograms at holo_dev: through the graph
  // captured for holo_dev and ri_dev (captured now if there is none), or kernel by kernel.
  // ri_dev: where the index volume goes (null: the solver's own buffer, which run() and
  // process() download from).
  void compute(const float* holo_dev, float* ri_dev, hipStream_t s);
  // The graph captured for holo_dev and ri_dev (captured now if there is none); false when the
  // capture failed and the pipeline fell back to kernel-by-kernel launches.
  bool ensure_graph(const float* holo_dev, float* ri_dev);
  void collect(bool downloaded);
  // Orders stream s after the work the last process_device() queued on its caller's stream:
  // every request shares the pipeline's buffers, whatever stream it runs on.
  void after_device_work(hipStream_t s);

  int count_, n_, m_, chunk_, chunks_;
  CropPlan crop_;
  Stream compute_;
  Stream copy_;
  Unwrapper unwrapper_;
  OdtSolver solver_;
  DeviceBuffer<float> holo_;
  DeviceBuffer<float2> half_;
  DeviceBuffer<float2> background_;
  DeviceBuffer<float2> fields_;
  DeviceBuffer<float> psi_;
  DeviceBuffer<float> logamp_;
  PinnedBuffer<float> input_;
  PinnedBuffer<float> output_;
  PinnedBuffer<unsigned long long> filled_;
  std::vector<Event> chunk_done_;
  std::vector<Event> out_done_;
  Event upload_start_;
  Event device_done_;  // recorded after the last process_device() on its caller's stream
  bool device_pending_ = false;
  GraphExec graph_;
  const float* graph_input_ = nullptr;
  float* graph_output_ = nullptr;
  bool use_graph_ = true;
  bool graph_ran_ = false;  // the last request ran through the graph
  bool capturing_ = false;  // stage marks are skipped while the graph is captured
  StageClock clock_;
  qh_pipeline_timings timings_{};
  double setup_ms_ = 0;
  std::size_t device_bytes_ = 0;
  bool has_background_ = false;
  bool timings_pending_ = false;
};

}  // namespace qh
