// Device-resident pipeline. The sample holograms are copied in `upload_chunks` parts on a copy
// stream while the compute stream transforms and crops the parts already present; fields stay
// in HBM from the crop to the index volume, which leaves with one copy.
#include <algorithm>
#include <chrono>

#include "fft.hpp"
#include "hostio.hpp"
#include "pipeline.hpp"

namespace qh {

OdtGrid grid_from(const qh_pipeline_options& o);
OdtConfig config_from(const qh_odt_options& o, const qh_geometry& g);

namespace {

double since_ms(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// The volume leaves the device in this many parts; holograms from a pageable buffer enter in
// pieces of this size.
constexpr int kOutChunks = 4;
constexpr std::size_t kInPiece = std::size_t(8) << 20;

}  // namespace

Pipeline::Pipeline(const qh_pipeline_options& o)
    : count_(o.count),
      n_(o.hologram_size),
      m_(o.crop.out),
      chunk_((o.count + o.upload_chunks - 1) / o.upload_chunks),
      chunks_((o.count + chunk_ - 1) / chunk_),
      crop_{o.crop.cy, o.crop.cx, o.crop.radius, o.crop.out},
      unwrapper_(o.count, o.crop.out, o.crop.out),
      solver_(grid_from(o), o.count, config_from(o.odt, o.geometry)),
      holo_(static_cast<std::size_t>(o.count) * o.hologram_size * o.hologram_size),
      half_(static_cast<std::size_t>(o.count) * o.hologram_size * (o.hologram_size / 2 + 1)),
      background_(static_cast<std::size_t>(o.count) * o.crop.out * o.crop.out),
      fields_(background_.size()),
      psi_(background_.size()),
      logamp_(background_.size()),
      input_(holo_.size()),
      output_(static_cast<std::size_t>(solver_.nz()) * solver_.n() * solver_.n()),
      filled_(1),
      chunk_done_(chunks_),
      out_done_(kOutChunks),
      clock_(kMarks) {
  std::vector<float> il(2 * static_cast<std::size_t>(count_));
  for (int a = 0; a < count_; ++a) {
    il[2 * a] = o.illuminations[a].fy;
    il[2 * a + 1] = o.illuminations[a].fx;
  }
  solver_.set_illuminations(il.data());
  // Twiddle tables of every transform length, created now so process() allocates nothing.
  for (int len : {n_, n_ / 2, m_, solver_.n(), solver_.n() / 2, solver_.nz()}) {
    if (len >= 1) {
      fft_twiddles(len);
      dct_twiddles(len);
    }
  }
  *filled_.get() = 0;
}

void Pipeline::after_device_work(hipStream_t s) {
  if (device_pending_) check(hipStreamWaitEvent(s, device_done_.get(), 0), "hipStreamWaitEvent");
}

void Pipeline::set_background(const float* host) {
  const hipStream_t s = compute_.get();
  after_device_work(s);
  const std::size_t image = static_cast<std::size_t>(n_) * n_;
  check(hipMemcpyAsync(holo_.get(), host, count_ * image * sizeof(float), hipMemcpyHostToDevice, s),
        "upload backgrounds");
  hologram_spectra(holo_.get(), count_, n_, half_.get(), s);
  crop_fields(half_.get(), count_, n_, crop_, background_.get(), s);
  fields_inverse(background_.get(), count_, m_, 1.0f, false, s);
  check(hipStreamSynchronize(s), "set_background");
  device_pending_ = false;
  has_background_ = true;
}

// Fields of the sample holograms, their ratio to the background fields and the Rytov field.
// The 1/N^2 scale and the (-1)^(x+y) output sign cancel in the ratio and are skipped.
void Pipeline::front_end(const float* holo_dev, hipStream_t s, bool from_host, const float* user) {
  const std::size_t image = static_cast<std::size_t>(n_) * n_;
  const std::size_t field = static_cast<std::size_t>(m_) * m_;
  const int h = n_ / 2 + 1;
  if (!capturing_) clock_.mark(kStart, s);
  const hipStream_t c = copy_.get();
  if (from_host) check(hipStreamWaitEvent(c, clock_.event(kStart).get(), 0), "hipStreamWaitEvent");
  for (int k = 0; k < chunks_; ++k) {
    const int first = k * chunk_;
    const int valid = std::min(chunk_, count_ - first);
    if (from_host) {
      // From the caller's pageable buffer: pieces of the chunk are copied into the pinned input
      // and queued for upload one after the other, so host copies and DMA overlap.
      const std::size_t bytes = valid * image * sizeof(float);
      const std::size_t piece = user ? kInPiece : bytes;
      for (std::size_t a = 0; a < bytes; a += piece) {
        const std::size_t len = std::min(piece, bytes - a);
        char* pinned = reinterpret_cast<char*>(input_.get() + first * image) + a;
        if (user) host_copy(pinned, reinterpret_cast<const char*>(user + first * image) + a, len);
        check(hipMemcpyAsync(reinterpret_cast<char*>(holo_.get() + first * image) + a, pinned, len,
                             hipMemcpyHostToDevice, c),
              "upload holograms");
      }
      chunk_done_[k].record(c);
      check(hipStreamWaitEvent(s, chunk_done_[k].get(), 0), "hipStreamWaitEvent");
    }
    const float* src = (from_host ? holo_.get() : holo_dev) + first * image;
    float2* half = half_.get() + static_cast<std::size_t>(first) * n_ * h;
    hologram_spectra(src, valid, n_, half, s);
    crop_fields(half, valid, n_, crop_, fields_.get() + first * field, s);
  }
  fields_inverse(fields_.get(), count_, m_, 1.0f, false, s);
  if (!capturing_) clock_.mark(kFields, s);
  ratio_phase(fields_.get(), background_.get(), static_cast<long>(count_) * field, psi_.get(),
              logamp_.get(), s);
  unwrapper_.rytov(psi_.get(), logamp_.get(), fields_.get(), s);
  if (!capturing_) clock_.mark(kRytov, s);
}

void Pipeline::process() { process_host(nullptr, nullptr); }

void Pipeline::run(const float* holograms, float* ri) { process_host(holograms, ri); }

void Pipeline::process_host(const float* user_in, float* user_out) {
  if (!has_background_) throw LogicError("Pipeline: set_background() first");
  const auto t0 = std::chrono::steady_clock::now();
  const hipStream_t s = compute_.get();
  // Both streams start after any request still queued on a caller's stream (the copy stream
  // below waits on events recorded on s after this point).
  after_device_work(s);
  after_device_work(copy_.get());
  if (use_graph_ && !solver_.timing_fgp()) {
    // Uploads into holo_ on the copy stream once the previous request's work on s is done,
    // then the captured front end and solver, which read holo_.
    const hipStream_t c = copy_.get();
    const std::size_t image = static_cast<std::size_t>(n_) * n_;
    upload_start_.record(s);
    check(hipStreamWaitEvent(c, upload_start_.get(), 0), "hipStreamWaitEvent");
    const std::size_t bytes = count_ * image * sizeof(float);
    const std::size_t piece = user_in ? kInPiece : bytes;
    for (std::size_t a = 0; a < bytes; a += piece) {
      const std::size_t len = std::min(piece, bytes - a);
      char* pinned = reinterpret_cast<char*>(input_.get()) + a;
      if (user_in) host_copy(pinned, reinterpret_cast<const char*>(user_in) + a, len);
      check(hipMemcpyAsync(reinterpret_cast<char*>(holo_.get()) + a, pinned, len,
                           hipMemcpyHostToDevice, c),
            "upload holograms");
    }
    chunk_done_.back().record(c);
    check(hipStreamWaitEvent(s, chunk_done_.back().get(), 0), "hipStreamWaitEvent");
    compute(holo_.get(), nullptr, s);
  } else {
    graph_ran_ = false;
    front_end(nullptr, s, true, user_in);
    solver_.run(fields_.get(), s, &clock_, kSolver);
  }
  const std::size_t vol = static_cast<std::size_t>(solver_.nz()) * solver_.n() * solver_.n();
  const std::size_t part = (vol + kOutChunks - 1) / kOutChunks;
  for (int k = 0; k < kOutChunks; ++k) {
    const std::size_t a = std::min(vol, k * part);
    const std::size_t len = std::min(vol, a + part) - a;
    if (len) {
      check(hipMemcpyAsync(output_.get() + a, solver_.index() + a, len * sizeof(float),
                           hipMemcpyDeviceToHost, s),
            "download index");
    }
    out_done_[k].record(s);
  }
  check(hipMemcpyAsync(filled_.get(), solver_.filled_device(), sizeof(unsigned long long),
                       hipMemcpyDeviceToHost, s),
        "download filled");
  clock_.mark(kDownloaded, s);
  if (user_out) {
    for (int k = 0; k < kOutChunks; ++k) {
      const std::size_t a = std::min(vol, k * part);
      const std::size_t len = std::min(vol, a + part) - a;
      check(hipEventSynchronize(out_done_[k].get()), "hipEventSynchronize");
      if (len) host_copy(user_out + a, output_.get() + a, len * sizeof(float));
    }
  }
  check(hipStreamSynchronize(s), "process");
  device_pending_ = false;
  collect(true);
  timings_pending_ = false;
  timings_.upload = graph_ran_ ? elapsed_ms(upload_start_, chunk_done_.back())
                               : elapsed_ms(clock_.event(kStart), chunk_done_.back());
  timings_.total = since_ms(t0);
}

bool Pipeline::ensure_graph(const float* holo_dev, float* ri_dev) {
  if (graph_.ready() && graph_input_ == holo_dev && graph_output_ == ri_dev) return true;
  // Captured on the pipeline's own stream (a caller's stream may be the null stream, which
  // cannot be captured), without the stage marks: events recorded inside a capture only
  // order work within the graph and are not time-stamped by its replays.
  // A capture can be invalidated by work other threads start meanwhile (allocations and
  // synchronous copies on the null stream); the pipeline then stays with kernel-by-kernel
  // launches, whose results are the same.
  const hipStream_t cs = compute_.get();
  check(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  capturing_ = true;
  try {
    front_end(holo_dev, cs, false);
    solver_.set_output(ri_dev);
    solver_.run(fields_.get(), cs, nullptr, 0);
    solver_.set_output(nullptr);
    capturing_ = false;
    graph_.end_capture(cs);
    graph_input_ = holo_dev;
    graph_output_ = ri_dev;
    return true;
  } catch (const DeviceError&) {
    capturing_ = false;
    solver_.set_output(nullptr);
    GraphExec::abandon_capture(cs);
    graph_.reset();
    use_graph_ = false;
    return false;
  } catch (...) {
    capturing_ = false;
    solver_.set_output(nullptr);
    GraphExec::abandon_capture(cs);
    graph_.reset();
    throw;
  }
}

void Pipeline::capture() {
  if (!has_background_) throw LogicError("Pipeline: set_background() first");
  if (use_graph_ && !solver_.timing_fgp()) ensure_graph(holo_.get(), nullptr);
}

void Pipeline::compute(const float* holo_dev, float* ri_dev, hipStream_t s) {
  graph_ran_ = use_graph_ && !solver_.timing_fgp() && ensure_graph(holo_dev, ri_dev);
  if (!graph_ran_) {
    front_end(holo_dev, s, false);
    solver_.set_output(ri_dev);
    solver_.run(fields_.get(), s, &clock_, kSolver);
    solver_.set_output(nullptr);
    return;
  }
  clock_.mark(kStart, s);
  graph_.launch(s);
  clock_.mark(kSolver + kFinish, s);
}

void Pipeline::process_device(const float* holo_dev, float* ri_dev, hipStream_t s) {
  if (!has_background_) throw LogicError("Pipeline: set_background() first");
  after_device_work(s);
  // The index volume is written straight into ri_dev (the graph is captured for that address).
  compute(holo_dev, ri_dev, s);
  check(hipMemcpyAsync(filled_.get(), solver_.filled_device(), sizeof(unsigned long long),
                       hipMemcpyDeviceToHost, s),
        "download filled");
  clock_.mark(kDownloaded, s);
  device_done_.record(s);
  device_pending_ = true;
  timings_pending_ = true;
}

const qh_pipeline_timings& Pipeline::timings() {
  if (timings_pending_) {
    check(hipEventSynchronize(clock_.event(kDownloaded).get()), "timings");
    collect(false);
    timings_.upload = 0;
    timings_.total = clock_.ms(kStart, kDownloaded);
    timings_pending_ = false;
  }
  return timings_;
}

void Pipeline::collect(bool downloaded) {
  const int o = kSolver;
  if (graph_ran_) {
    timings_.fields = timings_.rytov = timings_.fft2 = timings_.clear = timings_.mapping = 0;
    timings_.normalize = timings_.finish = 0;
    timings_.solve = clock_.ms(kStart, o + kFinish);
    timings_.download = downloaded ? clock_.ms(o + kFinish, kDownloaded) : 0.0;
    timings_.fgp_ms = 0;
    timings_.fgp_steps = 0;
    timings_.fgp_launches = 0;
    return;
  }
  timings_.fields = clock_.ms(kStart, kFields);
  timings_.rytov = clock_.ms(kFields, kRytov);
  timings_.fft2 = clock_.ms(kRytov, o + kFft2);
  timings_.clear = clock_.ms(o + kFft2, o + kClear);
  timings_.mapping = clock_.ms(o + kClear, o + kMapping);
  timings_.normalize = clock_.ms(o + kMapping, o + kNormalize);
  timings_.solve = clock_.ms(o + kNormalize, o + kSolve);
  timings_.finish = clock_.ms(o + kSolve, o + kFinish);
  timings_.download = downloaded ? clock_.ms(o + kFinish, kDownloaded) : 0.0;
  timings_.fgp_ms = solver_.collect_fgp_ms();
  timings_.fgp_steps = solver_.fgp_steps();
  timings_.fgp_launches = solver_.fgp_launches();
}

double Pipeline::filled() const {
  const double vol = static_cast<double>(solver_.nz()) * solver_.n() * solver_.n();
  return static_cast<double>(*filled_.get()) / vol;
}

}  // namespace qh
