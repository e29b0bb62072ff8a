// HIP runtime plumbing: status errors, device and pinned buffers, streams, events.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace qh {

// Errors the C-ABI maps to status codes: invalid arguments, misuse, out-of-memory and device
// failures keep their own types so the host layers can rethrow what the reference throws.
struct InvalidArgument : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};
struct LogicError : std::logic_error {
  using std::logic_error::logic_error;
};
struct OutOfMemory : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void require(bool ok, const std::string& what) {
  if (!ok) throw InvalidArgument(what);
}

inline void check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  const std::string msg = std::string(what) + ": " + hipGetErrorString(e);
  if (e == hipErrorOutOfMemory) throw OutOfMemory(msg);
  throw DeviceError(msg);
}

inline void check_launch(const char* what) { check(hipGetLastError(), what); }

inline unsigned blocks_for(std::size_t n, unsigned block = 256) {
  return static_cast<unsigned>((n + block - 1) / block);
}

template <class T>
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(std::size_t n) { resize(n); }
  ~DeviceBuffer() { release(); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : p_(std::exchange(o.p_, nullptr)), n_(std::exchange(o.n_, 0)) {}
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
      release();
      p_ = std::exchange(o.p_, nullptr);
      n_ = std::exchange(o.n_, 0);
    }
    return *this;
  }
  void resize(std::size_t n) {
    release();
    if (n == 0) return;
    check(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)), "hipMalloc");
    n_ = n;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* get() const { return p_; }
  std::size_t size() const { return n_; }
  std::size_t bytes() const { return n_ * sizeof(T); }

 private:
  T* p_ = nullptr;
  std::size_t n_ = 0;
};

template <class T>
class PinnedBuffer {
 public:
  PinnedBuffer() = default;
  explicit PinnedBuffer(std::size_t n) {
    if (n == 0) return;
    check(hipHostMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T), hipHostMallocDefault),
          "hipHostMalloc");
    n_ = n;
  }
  ~PinnedBuffer() {
    if (p_) (void)hipHostFree(p_);
  }
  PinnedBuffer(const PinnedBuffer&) = delete;
  PinnedBuffer& operator=(const PinnedBuffer&) = delete;
  PinnedBuffer(PinnedBuffer&& o) noexcept : p_(std::exchange(o.p_, nullptr)), n_(std::exchange(o.n_, 0)) {}
  T* get() const { return p_; }
  std::size_t size() const { return n_; }

 private:
  T* p_ = nullptr;
  std::size_t n_ = 0;
};

class Stream {
 public:
  Stream() { check(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking), "hipStreamCreate"); }
  ~Stream() {
    if (s_) (void)hipStreamDestroy(s_);
  }
  Stream(const Stream&) = delete;
  Stream& operator=(const Stream&) = delete;
  hipStream_t get() const { return s_; }

 private:
  hipStream_t s_ = nullptr;
};

class Event {
 public:
  Event() { check(hipEventCreate(&e_), "hipEventCreate"); }
  ~Event() {
    if (e_) (void)hipEventDestroy(e_);
  }
  Event(const Event&) = delete;
  Event& operator=(const Event&) = delete;
  Event(Event&& o) noexcept : e_(std::exchange(o.e_, nullptr)) {}
  void record(hipStream_t s) { check(hipEventRecord(e_, s), "hipEventRecord"); }
  hipEvent_t get() const { return e_; }

 private:
  hipEvent_t e_ = nullptr;
};

// An instantiated graph: the kernels, copies and event records captured from a stream, replayed
// with one launch.
class GraphExec {
 public:
  GraphExec() = default;
  ~GraphExec() { reset(); }
  GraphExec(const GraphExec&) = delete;
  GraphExec& operator=(const GraphExec&) = delete;
  void reset() {
    if (g_) (void)hipGraphExecDestroy(g_);
    g_ = nullptr;
  }
  // Ends the capture begun on s and instantiates what it recorded.
  void end_capture(hipStream_t s) {
    reset();
    hipGraph_t graph = nullptr;
    check(hipStreamEndCapture(s, &graph), "hipStreamEndCapture");
    const hipError_t e = hipGraphInstantiate(&g_, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    check(e, "hipGraphInstantiate");
  }
  // Ends a capture that failed part way and drops what it recorded.
  static void abandon_capture(hipStream_t s) {
    hipGraph_t graph = nullptr;
    (void)hipStreamEndCapture(s, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
  }
  void launch(hipStream_t s) const { check(hipGraphLaunch(g_, s), "hipGraphLaunch"); }
  bool ready() const { return g_ != nullptr; }

 private:
  hipGraphExec_t g_ = nullptr;
};

inline double elapsed_ms(const Event& a, const Event& b) {
  float ms = 0;
  check(hipEventElapsedTime(&ms, a.get(), b.get()), "hipEventElapsedTime");
  return ms;
}

}  // namespace qh
