// Host ingest and egress (see hostio.hpp).
#include "hostio.hpp"

#include <sched.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "runtime.hpp"

namespace qh {

namespace {

// Copies below this size go through one hipMemcpy (or one memcpy) on the calling thread.
constexpr std::size_t kDirect = std::size_t(4) << 20;
// Staging chunk: large enough that each copy thread moves a few MB per chunk.
constexpr std::size_t kChunk = std::size_t(16) << 20;
// Host copies are split into parts of at least this size.
constexpr std::size_t kMinPart = std::size_t(1) << 20;

int usable_cpus() {
  cpu_set_t set;
  CPU_ZERO(&set);
  if (sched_getaffinity(0, sizeof set, &set) == 0) return std::max(1, CPU_COUNT(&set));
  return std::max(1u, std::thread::hardware_concurrency());
}

// A fixed set of copy threads. A job is a number of parts; the submitting thread works on its
// own job too, so a job always finishes even when every pool thread is busy with other jobs.
class CopyPool {
 public:
  explicit CopyPool(int threads) {
    for (int i = 0; i < threads; ++i) workers_.emplace_back([this] { loop(); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lock(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
  }
  CopyPool(const CopyPool&) = delete;
  CopyPool& operator=(const CopyPool&) = delete;

  void run(int parts, const std::function<void(int)>& part) {
    auto job = std::make_shared<Job>();
    job->parts = parts;
    job->part = &part;
    {
      std::lock_guard<std::mutex> lock(mu_);
      jobs_.push_back(job);
    }
    cv_.notify_all();
    work(*job);
    retire(job);
    std::unique_lock<std::mutex> lock(job->mu);
    job->done_cv.wait(lock, [&] { return job->done == job->parts; });
  }

 private:
  struct Job {
    int parts = 0;
    const std::function<void(int)>* part = nullptr;
    std::atomic<int> next{0};
    int done = 0;  // guarded by mu
    std::mutex mu;
    std::condition_variable done_cv;
  };

  static void work(Job& job) {
    for (;;) {
      const int i = job.next.fetch_add(1);
      if (i >= job.parts) return;
      (*job.part)(i);
      std::lock_guard<std::mutex> lock(job.mu);
      if (++job.done == job.parts) job.done_cv.notify_all();
    }
  }

  // Takes a job off the queue once all its parts are claimed (work() has returned for it).
  void retire(const std::shared_ptr<Job>& job) {
    std::lock_guard<std::mutex> lock(mu_);
    const auto it = std::find(jobs_.begin(), jobs_.end(), job);
    if (it != jobs_.end()) jobs_.erase(it);
  }

  void loop() {
    for (;;) {
      std::shared_ptr<Job> job;
      {
        std::unique_lock<std::mutex> lock(mu_);
        for (;;) {
          cv_.wait(lock, [&] { return stop_ || !jobs_.empty(); });
          if (stop_) return;
          job = jobs_.front();
          if (job->next.load() < job->parts) break;
          jobs_.pop_front();  // every part is claimed already
        }
      }
      work(*job);
      retire(job);
    }
  }

  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<std::shared_ptr<Job>> jobs_;
  bool stop_ = false;
};

CopyPool& pool() {
  static CopyPool p(host_copy_threads() - 1);
  return p;
}

// Two pinned chunks of the calling thread and the events that mark their last DMA.
struct Staging {
  PinnedBuffer<char> buf[2]{PinnedBuffer<char>(kChunk), PinnedBuffer<char>(kChunk)};
  Event done[2];
  bool pending[2] = {false, false};

  void wait(int b) {
    if (pending[b]) check(hipEventSynchronize(done[b].get()), "hipEventSynchronize");
    pending[b] = false;
  }
};

Staging& staging() {
  thread_local std::unique_ptr<Staging> s;
  if (!s) s = std::make_unique<Staging>();
  return *s;
}

}  // namespace

int host_copy_threads() {
  static const int n = std::clamp(usable_cpus() / 2, 1, 8);
  return n;
}

void host_copy(void* dst, const void* src, std::size_t bytes) {
  const int threads = host_copy_threads();
  if (bytes < 2 * kMinPart || threads == 1) {
    std::memcpy(dst, src, bytes);
    return;
  }
  const int parts = static_cast<int>(std::min<std::size_t>(threads, bytes / kMinPart));
  // Part boundaries on 4 KiB pages.
  const std::size_t step = ((bytes / parts) + 4095) & ~std::size_t(4095);
  auto* d = static_cast<char*>(dst);
  const auto* s = static_cast<const char*>(src);
  pool().run(parts, [&](int i) {
    const std::size_t a = std::min(bytes, i * step);
    const std::size_t b = std::min(bytes, a + step);
    if (b > a) std::memcpy(d + a, s + a, b - a);
  });
}

void parallel_for(int parts, const std::function<void(int)>& part) {
  if (parts <= 1) {
    if (parts == 1) part(0);
    return;
  }
  pool().run(parts, part);
}

void upload(void* dst, const void* src, std::size_t bytes, hipStream_t s) {
  if (bytes < kDirect) {
    check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s), "upload");
    return;
  }
  Staging& st = staging();
  auto* d = static_cast<char*>(dst);
  const auto* h = static_cast<const char*>(src);
  int b = 0;
  for (std::size_t off = 0; off < bytes; off += kChunk, b ^= 1) {
    const std::size_t len = std::min(kChunk, bytes - off);
    st.wait(b);
    host_copy(st.buf[b].get(), h + off, len);
    check(hipMemcpyAsync(d + off, st.buf[b].get(), len, hipMemcpyHostToDevice, s), "upload");
    st.done[b].record(s);
    st.pending[b] = true;
  }
}

void download(void* dst, const void* src, std::size_t bytes, hipStream_t s) {
  if (bytes < kDirect) {
    check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s), "download");
    check(hipStreamSynchronize(s), "download");
    return;
  }
  Staging& st = staging();
  auto* h = static_cast<char*>(dst);
  const auto* d = static_cast<const char*>(src);
  const std::size_t chunks = (bytes + kChunk - 1) / kChunk;
  auto enqueue = [&](std::size_t i) {
    const int b = static_cast<int>(i & 1);
    const std::size_t off = i * kChunk;
    st.wait(b);
    check(hipMemcpyAsync(st.buf[b].get(), d + off, std::min(kChunk, bytes - off),
                         hipMemcpyDeviceToHost, s),
          "download");
    st.done[b].record(s);
    st.pending[b] = true;
  };
  enqueue(0);
  if (chunks > 1) enqueue(1);
  for (std::size_t i = 0; i < chunks; ++i) {
    const int b = static_cast<int>(i & 1);
    const std::size_t off = i * kChunk;
    st.wait(b);
    host_copy(h + off, st.buf[b].get(), std::min(kChunk, bytes - off));
    if (i + 2 < chunks) enqueue(i + 2);
  }
}

}  // namespace qh
