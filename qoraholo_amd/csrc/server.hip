// Served reconstruction loop (see server.hpp).
#include <algorithm>

#include "recon_batch.hpp"
#include "server.hpp"

namespace qh {

void check_pipeline_options(const qh_pipeline_options& o);

namespace {

double ms_between(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

std::size_t free_device_bytes() {
  std::size_t free_b = 0, total_b = 0;
  check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
  return free_b;
}

// A reconstruction slot: a Pipeline of the geometry with its background fields set.
class ReconWorker final : public SlotWorker {
 public:
  struct Spec {
    qh_pipeline_options options;
    std::vector<qh_illumination> illuminations;
    std::vector<float> backgrounds;
  };
  explicit ReconWorker(const Spec& spec) : pipeline_(spec.options) {
    pipeline_.set_background(spec.backgrounds.data());
    // Slots are built under the server lock, which also covers every allocation and free of
    // other slots, so no other slot's setup can invalidate this capture; every request of the
    // slot then replays the graph (the slot falls back to kernel-by-kernel launches, with the
    // same results, if the capture fails).
    pipeline_.capture();
  }
  void run(const void* input, float* out, double*) override {
    pipeline_.run(static_cast<const float*>(input), out);
  }
  double filled() const override { return pipeline_.filled(); }
  bool replays_graph() const override { return pipeline_.graph_ready(); }

 private:
  Pipeline pipeline_;
};

// A reconstruction slot that takes up to `capacity` requests of its geometry at once.
class ReconBatchWorker final : public SlotWorker {
 public:
  ReconBatchWorker(const ReconWorker::Spec& spec, int capacity)
      : batch_(spec.options, spec.backgrounds.data(), capacity) {}
  void run(const void* input, float* out, double*) override {
    const float* in = static_cast<const float*>(input);
    batch_.run(&in, &out, 1);
  }
  void run_batch(const void* const* inputs, float* const* outs, double* const*, int nb) override {
    std::vector<const float*> in(nb);
    for (int b = 0; b < nb; ++b) in[b] = static_cast<const float*>(inputs[b]);
    batch_.run(in.data(), outs, nb);
  }
  double filled() const override { return batch_.filled(); }
  bool replays_graph() const override { return batch_.graphs_ready(); }
  int capacity() const override { return batch_.capacity(); }

 private:
  ReconBatch batch_;
};

// A BPM inversion slot.
class BpmWorker final : public SlotWorker {
 public:
  explicit BpmWorker(const BpmInversionSpec& spec) : inversion_(spec) {}
  void run(const void* input, float* out, double* extra) override {
    inversion_.run(static_cast<const float2*>(input), out, extra);
  }
  int chunk() const { return inversion_.chunk(); }
  int checkpoint() const { return inversion_.checkpoint(); }

 private:
  BpmInversion inversion_;
};

}  // namespace

Server::Server(std::size_t budget_bytes, int max_slots)
    : budget_(budget_bytes), max_slots_(std::max(1, max_slots)) {
  if (budget_ == 0) budget_ = free_device_bytes() / 10 * 9;
  dispatcher_ = std::thread([this] { dispatch_loop(); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lock(mu_);
    stop_ = true;
    for (auto& s : slots_) s->stop = true;
  }
  queue_cv_.notify_all();
  for (auto& s : slots_) s->cv.notify_all();
  if (dispatcher_.joinable()) dispatcher_.join();
  for (auto* list : {&slots_, &retired_}) {
    for (auto& s : *list) {
      s->cv.notify_all();
      if (s->thread.joinable()) s->thread.join();
    }
  }
}

void Server::reset_stats() {
  std::lock_guard<std::mutex> lock(mu_);
  latencies_.clear();
  submitted_ = completed_ = failed_ = 0;
  batched_ = 0;
  peak_batch_ = 0;
  peak_busy_ = busy_slots();
}

int Server::add_geometry(const qh_pipeline_options& o, const float* backgrounds, int max_batch) {
  check_pipeline_options(o);
  require(backgrounds != nullptr, "Server: background holograms are required");
  require(max_batch >= 1 && max_batch <= 64, "Server: max_batch must be between 1 and 64");
  auto spec = std::make_shared<ReconWorker::Spec>();
  spec->options = o;
  spec->illuminations.assign(o.illuminations, o.illuminations + o.count);
  spec->options.illuminations = spec->illuminations.data();
  const std::size_t image = static_cast<std::size_t>(o.hologram_size) * o.hologram_size;
  spec->backgrounds.assign(backgrounds, backgrounds + image * o.count);
  if (max_batch > 1) {
    return add_kind(kReconstruction, [spec, max_batch]() -> std::unique_ptr<SlotWorker> {
      return std::make_unique<ReconBatchWorker>(*spec, max_batch);
    });
  }
  return add_kind(kReconstruction, [spec] { return std::make_unique<ReconWorker>(*spec); });
}

int Server::add_bpm(const BpmInversionSpec& spec) {
  auto shared = std::make_shared<BpmInversionSpec>(spec);
  // An automatic illumination chunk and checkpoint interval are chosen from the free memory
  // when the first slot is built (inside add_kind, before the geometry is visible to the
  // dispatcher) and kept for every later slot, so that all slots have the footprint admission
  // measured on the first.
  return add_kind(kBpm, [shared] {
    auto w = std::make_unique<BpmWorker>(*shared);
    if (shared->chunk == 0) shared->chunk = w->chunk();
    if (shared->checkpoint < 0) shared->checkpoint = w->checkpoint();
    return w;
  });
}

int Server::add_kind(Kind kind, std::function<std::unique_ptr<SlotWorker>()> make) {
  auto g = std::make_unique<Geometry>();
  g->make = std::move(make);
  g->kind = kind;
  // The first slot is built now: it measures the footprint and fails early when it does not fit.
  // The lock keeps the dispatcher from building slots meanwhile (it builds them under the same
  // lock), so the change in free memory is this slot's alone; running slots allocate nothing.
  std::lock_guard<std::mutex> lock(mu_);
  check(hipDeviceSynchronize(), "Server::add_geometry");
  const std::size_t before = free_device_bytes();
  auto slot = std::make_unique<Slot>();
  slot->pipeline = g->make();
  check(hipDeviceSynchronize(), "Server::add_geometry");
  const std::size_t after = free_device_bytes();
  slot->bytes = before > after ? before - after : 0;
  g->slot_bytes = std::max<std::size_t>(slot->bytes, 1);
  if (slot->bytes > budget_) {
    throw OutOfMemory("Server: one slot of this geometry needs " + std::to_string(slot->bytes) +
                      " bytes, over the HBM budget");
  }
  const int id = static_cast<int>(geometries_.size());
  // Idle slots of other geometries make room as they do for a request; when the slot still does
  // not fit (the others are busy), the geometry is registered without a resident slot and its
  // first request builds one.
  while (used_bytes_ + slot->bytes > budget_) {
    if (!evict_idle(id)) break;
  }
  geometries_.push_back(std::move(g));
  if (used_bytes_ + slot->bytes > budget_) return id;
  slot->geometry = id;
  geometries_[id]->slots = 1;
  used_bytes_ += slot->bytes;
  Slot* raw = slot.get();
  slots_.push_back(std::move(slot));
  raw->thread = std::thread([this, raw] { slot_loop(raw); });
  return id;
}

int Server::busy_slots() const {
  int n = 0;
  for (const auto& s : slots_) n += s->busy ? 1 : 0;
  return n;
}

bool Server::evict_idle(int keep) {
  auto it = std::find_if(slots_.begin(), slots_.end(), [&](const std::unique_ptr<Slot>& s) {
    return s->pipeline && !s->busy && s->geometry != keep;
  });
  if (it == slots_.end()) return false;
  // The victim is idle: free its device memory now; its thread leaves its wait once the lock is
  // released and is joined with the retired slots.
  Slot* victim = it->get();
  victim->pipeline.reset();
  victim->stop = true;
  victim->cv.notify_all();
  used_bytes_ -= std::min(used_bytes_, victim->bytes);
  --geometries_[victim->geometry]->slots;
  retired_.push_back(std::move(*it));
  slots_.erase(it);
  return true;
}

std::int64_t Server::submit(int geometry, Kind kind, const void* input, float* ri_out,
                            double* extra) {
  require(input != nullptr && ri_out != nullptr, "Server::submit: null buffer");
  auto r = std::make_shared<Request>();
  r->input = input;
  r->out = ri_out;
  r->extra = extra;
  r->arrival = std::chrono::steady_clock::now();
  {
    std::lock_guard<std::mutex> lock(mu_);
    require(geometry >= 0 && geometry < static_cast<int>(geometries_.size()),
            "Server::submit: unknown geometry");
    require(geometries_[geometry]->kind == kind,
            kind == kBpm ? "Server::submit_bpm: not a BPM geometry"
                         : "Server::submit: a BPM geometry takes submit_bpm");
    r->geometry = geometry;
    r->id = next_id_++;
    if (submitted_ == 0) first_arrival_ = r->arrival;
    ++submitted_;
    requests_[r->id] = r;
    queue_.push_back(r);
  }
  queue_cv_.notify_all();
  return r->id;
}

Server::Slot* Server::slot_for(const std::shared_ptr<Request>& r) {
  // max_slots bounds the requests in flight; idle slots stay resident (background fields, plans
  // and solver state of their geometry) until the HBM budget needs their memory.
  if (busy_slots() >= max_slots_) return nullptr;
  for (auto& s : slots_) {
    if (!s->busy && !s->stop && s->geometry == r->geometry && s->pipeline) return s.get();
  }
  Geometry& g = *geometries_[r->geometry];
  while (used_bytes_ + g.slot_bytes > budget_) {
    // Evict one idle slot of another geometry; stop when none is idle.
    if (!evict_idle(r->geometry)) return nullptr;
  }
  auto slot = std::make_unique<Slot>();
  slot->geometry = r->geometry;
  slot->pipeline = g.make();
  slot->bytes = g.slot_bytes;
  used_bytes_ += slot->bytes;
  ++g.slots;
  Slot* raw = slot.get();
  slots_.push_back(std::move(slot));
  raw->thread = std::thread([this, raw] { slot_loop(raw); });
  return raw;
}

void Server::dispatch_loop() {
  std::unique_lock<std::mutex> lock(mu_);
  while (!stop_) {
    bool progress = false;
    // Indexes, not iterators: erasing inside a deque invalidates its iterators.
    for (std::size_t i = 0; i < queue_.size();) {
      Slot* slot = nullptr;
      try {
        slot = slot_for(queue_[i]);
      } catch (const std::exception& e) {
        // A slot that cannot be built fails this request only.
        auto& r = *queue_[i];
        r.finished = true;
        r.status = dynamic_cast<const OutOfMemory*>(&e) ? QH_ERR_OUT_OF_MEMORY : QH_ERR_RUNTIME;
        r.error = e.what();
        ++failed_;
        done_cv_.notify_all();
        queue_.erase(queue_.begin() + static_cast<std::ptrdiff_t>(i));
        progress = true;
        continue;
      }
      if (!slot) {
        ++i;
        continue;
      }
      // The slot takes this request and, up to its capacity, the later queued requests of the
      // same geometry, in arrival order.
      const int geometry = queue_[i]->geometry;
      const std::size_t cap = static_cast<std::size_t>(slot->pipeline->capacity());
      std::vector<std::shared_ptr<Request>> batch{queue_[i]};
      queue_.erase(queue_.begin() + static_cast<std::ptrdiff_t>(i));
      for (std::size_t j = i; j < queue_.size() && batch.size() < cap;) {
        if (queue_[j]->geometry == geometry) {
          batch.push_back(queue_[j]);
          queue_.erase(queue_.begin() + static_cast<std::ptrdiff_t>(j));
        } else {
          ++j;
        }
      }
      if (batch.size() > 1) batched_ += static_cast<std::int64_t>(batch.size());
      peak_batch_ = std::max(peak_batch_, static_cast<int>(batch.size()));
      slot->busy = true;
      slot->work = std::move(batch);
      peak_busy_ = std::max(peak_busy_, busy_slots());
      slot->cv.notify_all();
      progress = true;
    }
    if (!progress) queue_cv_.wait(lock);
  }
}

void Server::slot_loop(Slot* slot) {
  std::unique_lock<std::mutex> lock(mu_);
  while (true) {
    slot->cv.wait(lock, [&] { return slot->stop || !slot->work.empty(); });
    if (slot->stop && slot->work.empty()) return;
    std::vector<std::shared_ptr<Request>> batch = slot->work;
    lock.unlock();
    int status = QH_OK;
    std::string error;
    double filled = 0;
    try {
      if (batch.size() == 1) {
        slot->pipeline->run(batch[0]->input, batch[0]->out, batch[0]->extra);
      } else {
        std::vector<const void*> in;
        std::vector<float*> out;
        std::vector<double*> extra;
        for (const auto& r : batch) {
          in.push_back(r->input);
          out.push_back(r->out);
          extra.push_back(r->extra);
        }
        slot->pipeline->run_batch(in.data(), out.data(), extra.data(),
                                  static_cast<int>(batch.size()));
      }
      filled = slot->pipeline->filled();
    } catch (const InvalidArgument& e) {
      status = QH_ERR_INVALID_ARGUMENT;
      error = e.what();
    } catch (const OutOfMemory& e) {
      status = QH_ERR_OUT_OF_MEMORY;
      error = e.what();
    } catch (const std::exception& e) {
      status = QH_ERR_RUNTIME;
      error = e.what();
    }
    const auto now = std::chrono::steady_clock::now();
    lock.lock();
    for (const auto& r : batch) {
      r->finished = true;
      r->status = status;
      r->error = error;
      r->filled = filled;
      r->latency_ms = ms_between(r->arrival, now);
      if (status == QH_OK) {
        ++completed_;
        latencies_.push_back(r->latency_ms);
        last_done_ = now;
      } else {
        ++failed_;
      }
    }
    slot->work.clear();
    slot->busy = false;
    done_cv_.notify_all();
    queue_cv_.notify_all();
  }
}

int Server::wait(std::int64_t id, double* filled, double* latency_ms, std::string* error) {
  std::unique_lock<std::mutex> lock(mu_);
  auto it = requests_.find(id);
  require(it != requests_.end(), "Server::wait: unknown request id");
  std::shared_ptr<Request> r = it->second;
  done_cv_.wait(lock, [&] { return r->finished; });
  requests_.erase(id);
  if (filled) *filled = r->filled;
  if (latency_ms) *latency_ms = r->latency_ms;
  if (error) *error = r->error;
  return r->status;
}

bool Server::done(std::int64_t id) {
  std::lock_guard<std::mutex> lock(mu_);
  auto it = requests_.find(id);
  require(it != requests_.end(), "Server::poll: unknown request id");
  return it->second->finished;
}

qh_server_stats Server::stats() {
  std::lock_guard<std::mutex> lock(mu_);
  qh_server_stats s{};
  s.submitted = submitted_;
  s.completed = completed_;
  s.failed = failed_;
  s.queued = static_cast<std::int64_t>(queue_.size());
  std::vector<double> l = latencies_;
  std::sort(l.begin(), l.end());
  const auto pct = [&](double q) {
    if (l.empty()) return 0.0;
    const double pos = q * static_cast<double>(l.size() - 1);
    const std::size_t i = static_cast<std::size_t>(pos);
    const double f = pos - static_cast<double>(i);
    return i + 1 < l.size() ? l[i] * (1.0 - f) + l[i + 1] * f : l[i];
  };
  s.p50_ms = pct(0.50);
  s.p99_ms = pct(0.99);
  double sum = 0;
  for (double v : l) sum += v;
  s.mean_ms = l.empty() ? 0.0 : sum / static_cast<double>(l.size());
  const double span = completed_ > 0 ? ms_between(first_arrival_, last_done_) : 0.0;
  s.recon_per_s = span > 0 ? 1000.0 * static_cast<double>(completed_) / span : 0.0;
  s.device_bytes = used_bytes_;
  s.budget_bytes = budget_;
  int live = 0, busy = 0, graphs = 0;
  for (auto& sl : slots_) {
    live += sl->pipeline ? 1 : 0;
    busy += sl->busy ? 1 : 0;
    graphs += sl->pipeline && sl->pipeline->replays_graph() ? 1 : 0;
  }
  s.slots = live;
  s.graph_slots = graphs;
  s.batched_requests = batched_;
  s.peak_batch = peak_batch_;
  s.busy_slots = busy;
  s.peak_busy_slots = peak_busy_;
  return s;
}

}  // namespace qh
