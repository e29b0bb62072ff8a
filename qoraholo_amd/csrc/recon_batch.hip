// Batched reconstruction slot (see recon_batch.hpp).
#include "recon_batch.hpp"

#include <algorithm>

#include "fft.hpp"
#include "hostio.hpp"

namespace qh {

OdtGrid grid_from(const qh_pipeline_options& o);
OdtConfig config_from(const qh_odt_options& o, const qh_geometry& g);
void check_pipeline_options(const qh_pipeline_options& o);

namespace {

// Images the unwrapper holds: every stack of a full batch (the options checked first).
int batch_images(const qh_pipeline_options& o, int capacity) {
  check_pipeline_options(o);
  require(capacity >= 1, "Server: the batch size must be positive");
  return capacity * o.count;
}

}  // namespace

ReconBatch::ReconBatch(const qh_pipeline_options& o, const float* host_backgrounds, int capacity)
    : count_(o.count),
      n_(o.hologram_size),
      m_(o.crop.out),
      cap_(capacity),
      crop_{o.crop.cy, o.crop.cx, o.crop.radius, o.crop.out},
      unwrapper_(batch_images(o, capacity), o.crop.out, o.crop.out),
      solver_(grid_from(o), o.count, config_from(o.odt, o.geometry), capacity),
      holo_(static_cast<std::size_t>(capacity) * o.count * o.hologram_size * o.hologram_size),
      half_(static_cast<std::size_t>(capacity) * o.count * o.hologram_size *
            (o.hologram_size / 2 + 1)),
      background_(static_cast<std::size_t>(o.count) * o.crop.out * o.crop.out),
      fields_(static_cast<std::size_t>(capacity) * background_.size()),
      psi_(fields_.size()),
      logamp_(fields_.size()),
      ri_(static_cast<std::size_t>(capacity) * solver_.nz() * solver_.n() * solver_.n()),
      filled_(1) {
  require(host_backgrounds != nullptr, "Server: background holograms are required");
  std::vector<float> il(2 * static_cast<std::size_t>(count_));
  for (int a = 0; a < count_; ++a) {
    il[2 * a] = o.illuminations[a].fy;
    il[2 * a + 1] = o.illuminations[a].fx;
  }
  solver_.set_illuminations(il.data());
  for (int len : {n_, n_ / 2, m_, solver_.n(), solver_.n() / 2, solver_.nz()}) {
    if (len >= 1) {
      fft_twiddles(len);
      dct_twiddles(len);
    }
  }
  *filled_.get() = 0;
  // Background fields, as Pipeline::set_background forms them.
  const hipStream_t s = stream_.get();
  const std::size_t image = static_cast<std::size_t>(n_) * n_;
  upload(holo_.get(), host_backgrounds, count_ * image * sizeof(float), s);
  hologram_spectra(holo_.get(), count_, n_, half_.get(), s);
  crop_fields(half_.get(), count_, n_, crop_, background_.get(), s);
  fields_inverse(background_.get(), count_, m_, 1.0f, false, s);
  check(hipStreamSynchronize(s), "ReconBatch background");
  // One graph per batch size, captured now (the server builds slots under the lock that also
  // covers the allocations of other slots). A capture that fails leaves kernel-by-kernel
  // launches.
  graphs_.resize(cap_);
  graphs_ok_ = true;
  for (int b = 1; b <= cap_ && graphs_ok_; ++b) {
    graphs_[b - 1] = std::make_unique<GraphExec>();
    check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
      device_work(b, s);
      graphs_[b - 1]->end_capture(s);
    } catch (const DeviceError&) {
      GraphExec::abandon_capture(s);
      graphs_ok_ = false;
    } catch (...) {
      GraphExec::abandon_capture(s);
      throw;
    }
  }
  if (!graphs_ok_) graphs_.clear();
}

// The device work of nb requests whose holograms are in holo_: the front end of every stack in
// one set of launches (the background fields repeat over the stacks), then each request's
// mapping and, on half spectra, one solve over all nb volumes (complex-spectrum solves: each
// request's in turn), each index volume into its own part of ri_.
void ReconBatch::device_work(int nb, hipStream_t s) {
  const int images = nb * count_;
  const std::size_t field = static_cast<std::size_t>(m_) * m_;
  const std::size_t vol = static_cast<std::size_t>(solver_.nz()) * solver_.n() * solver_.n();
  hologram_spectra(holo_.get(), images, n_, half_.get(), s);
  crop_fields(half_.get(), images, n_, crop_, fields_.get(), s);
  fields_inverse(fields_.get(), images, m_, 1.0f, false, s);
  ratio_phase(fields_.get(), background_.get(), static_cast<long>(images) * field, psi_.get(),
              logamp_.get(), s, static_cast<long>(count_) * field);
  unwrapper_.rytov(psi_.get(), logamp_.get(), fields_.get(), s, images);
  if (solver_.batchable()) {
    solver_.run_batch(fields_.get(), static_cast<long>(count_) * field, nb, ri_.get(), s);
    return;
  }
  for (int b = 0; b < nb; ++b) {
    solver_.set_output(ri_.get() + b * vol);
    solver_.run(fields_.get() + static_cast<std::size_t>(b) * count_ * field, s);
  }
  solver_.set_output(nullptr);
}

void ReconBatch::run(const float* const* inputs, float* const* outputs, int nb) {
  require(nb >= 1 && nb <= cap_, "ReconBatch: batch size beyond the capacity");
  const hipStream_t s = stream_.get();
  const std::size_t stack = static_cast<std::size_t>(count_) * n_ * n_;
  const std::size_t vol = static_cast<std::size_t>(solver_.nz()) * solver_.n() * solver_.n();
  for (int b = 0; b < nb; ++b) upload(holo_.get() + b * stack, inputs[b], stack * sizeof(float), s);
  if (graphs_ok_) {
    graphs_[nb - 1]->launch(s);
  } else {
    device_work(nb, s);
  }
  // The filled fraction depends on the geometry only (which bins the illuminations reach), so
  // every request of the batch has the last solve's.
  check(hipMemcpyAsync(filled_.get(), solver_.filled_device(), sizeof(unsigned long long),
                       hipMemcpyDeviceToHost, s),
        "download filled");
  for (int b = 0; b < nb; ++b) download(outputs[b], ri_.get() + b * vol, vol * sizeof(float), s);
  check(hipStreamSynchronize(s), "ReconBatch::run");
}

double ReconBatch::filled() const {
  const double vol = static_cast<double>(solver_.nz()) * solver_.n() * solver_.n();
  return static_cast<double>(*filled_.get()) / vol;
}

}  // namespace qh
