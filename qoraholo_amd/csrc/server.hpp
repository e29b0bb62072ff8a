// Served loop: requests of different kinds and geometries (reconstructions of holograms of a
// given count and size into a volume of a given size and solver; BPM inversions of field ratios
// of a given volume and illumination set) arrive on host threads; each geometry keeps its
// background fields, plans and solver state resident in device-resident slots; a dispatcher
// admits queued requests into free slots, creating slots while the HBM budget allows, so
// requests that arrive while others are on the device start as soon as a slot frees up and run
// concurrently on their own streams.
#pragma once

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/qoraholo_hip.h"
#include "bpm_inversion.hpp"
#include "pipeline.hpp"

namespace qh {

// What a slot runs: one request at a time, host input to host output, on device state of its
// own that it allocated when it was built.
class SlotWorker {
 public:
  virtual ~SlotWorker() = default;
  // input: the request's host input; out: its host result; extra: an optional host output of
  // the request kind (the loss of each BPM iteration).
  virtual void run(const void* input, float* out, double* extra) = 0;
  // Filled fraction of the last reconstruction's k-space (0 for other kinds).
  virtual double filled() const { return 0.0; }
  // The slot's requests replay a captured graph of their device work.
  virtual bool replays_graph() const { return false; }
  // Requests of its geometry the slot runs together; run_batch runs nb <= capacity() of them.
  virtual int capacity() const { return 1; }
  virtual void run_batch(const void* const* inputs, float* const* outs, double* const* extras,
                         int nb) {
    for (int b = 0; b < nb; ++b) run(inputs[b], outs[b], extras[b]);
  }
};

class Server {
 public:
  Server(std::size_t budget_bytes, int max_slots);
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  // Registers a reconstruction geometry with its background holograms (count, N, N); returns
  // its id. max_batch > 1: its slots take up to that many queued requests of the geometry at
  // once (one front end for all their stacks, one graph replay for the batch).
  int add_geometry(const qh_pipeline_options& o, const float* backgrounds, int max_batch = 1);
  // Registers a BPM inversion geometry; returns its id.
  int add_bpm(const BpmInversionSpec& spec);
  // Queues a request; its input (holograms (count, N, N) of a reconstruction geometry, field
  // ratios (count, n, n) complex of a BPM one), ri_out (nz, n, n) and extra (the per-iteration
  // loss of a BPM request, may be null) are host memory that must stay valid until wait()
  // returns for this id.
  // kind: the request kind the caller's input is for (kReconstruction or kBpm); a geometry of
  // the other kind is refused.
  enum Kind { kReconstruction = 0, kBpm = 1 };
  std::int64_t submit(int geometry, Kind kind, const void* input, float* ri_out,
                      double* extra = nullptr);
  // Blocks until the request is done; returns its status and message.
  int wait(std::int64_t id, double* filled, double* latency_ms, std::string* error);
  bool done(std::int64_t id);
  qh_server_stats stats();
  void reset_stats();

 private:
  struct Geometry {
    std::function<std::unique_ptr<SlotWorker>()> make;
    Kind kind = kReconstruction;
    std::size_t slot_bytes = 0;  // device bytes of one slot, measured on the first
    int slots = 0;
  };
  struct Request {
    std::int64_t id;
    int geometry;
    const void* input;
    float* out;
    double* extra;
    std::chrono::steady_clock::time_point arrival;
    bool finished = false;
    int status = QH_OK;
    std::string error;
    double filled = 0;
    double latency_ms = 0;
  };
  struct Slot {
    int geometry;
    std::unique_ptr<SlotWorker> pipeline;
    std::size_t bytes = 0;
    std::vector<std::shared_ptr<Request>> work;  // the requests it runs now (a batch)
    bool busy = false;
    bool stop = false;
    std::condition_variable cv;
    std::thread thread;
  };

  int add_kind(Kind kind, std::function<std::unique_ptr<SlotWorker>()> make);
  // Called with mu_ held: slots running a request, and the eviction of one idle slot of a
  // geometry other than `keep` (false when there is none).
  int busy_slots() const;
  bool evict_idle(int keep);
  void dispatch_loop();
  void slot_loop(Slot* slot);
  // Called with mu_ held: an idle slot of the request's geometry, or a new one when the budget
  // and the slot limit allow (evicting idle slots of other geometries if needed); null when the
  // request must wait.
  Slot* slot_for(const std::shared_ptr<Request>& r);

  std::size_t budget_;
  int max_slots_;
  std::mutex mu_;
  std::condition_variable queue_cv_;
  std::condition_variable done_cv_;
  std::deque<std::shared_ptr<Request>> queue_;
  std::map<std::int64_t, std::shared_ptr<Request>> requests_;
  std::vector<std::unique_ptr<Geometry>> geometries_;
  std::vector<std::unique_ptr<Slot>> slots_;
  std::vector<std::unique_ptr<Slot>> retired_;
  std::size_t used_bytes_ = 0;
  int peak_busy_ = 0;  // most slots running requests at once since the last reset_stats
  std::int64_t batched_ = 0;  // requests that ran in a batch of two or more
  int peak_batch_ = 0;
  std::int64_t next_id_ = 1;
  bool stop_ = false;
  std::vector<double> latencies_;
  std::int64_t submitted_ = 0, completed_ = 0, failed_ = 0;
  std::chrono::steady_clock::time_point first_arrival_, last_done_;
  std::thread dispatcher_;
};

}  // namespace qh
