// Backend::hip for the reference's C++ library: each function turns Array3 and the option
// structs into pointers and POD structs of include/qoraholo_hip.h, calls the HIP library and
// turns its status codes back into the exceptions the reference throws (std::invalid_argument,
// std::logic_error, std::runtime_error: src/api.cpp:13-23, src/cuda/util.cuh:16-26,
// src/pipeline.cpp:108). Compiled with the reference's headers (include/ and src/) and
// qoraholo_hip.h on the include path; links libqoraholo_hip.so.
#include "hip_backend.hpp"

#include <cstddef>
#include <stdexcept>
#include <string>

#include "qoraholo_hip.h"

namespace qoraholo::hip {

namespace {

void check(int rc) {
  if (rc == QH_OK) return;
  const std::string msg = qh_hip_last_error();
  if (rc == QH_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
  if (rc == QH_ERR_LOGIC) throw std::logic_error(msg);
  throw std::runtime_error(msg);
}

const qh_c64* c64(const cfloat* p) { return reinterpret_cast<const qh_c64*>(p); }
qh_c64* c64(cfloat* p) { return reinterpret_cast<qh_c64*>(p); }
const qh_illumination* illum(std::span<const Illumination> s) {
  return reinterpret_cast<const qh_illumination*>(s.data());
}

qh_geometry geometry_of(const Geometry& g) { return {g.wavelength, g.n_medium, g.na, g.pixel}; }

qh_odt_options odt_of(const OdtOptions& o) {
  return {o.nz, o.iterations, static_cast<int>(o.solver), static_cast<int>(o.gp_transform),
          o.tv_weight, o.tv_inner, static_cast<int>(o.mapping)};
}

qh_bpm_table_view view_of(const BpmTables& t) {
  return {t.n, t.nz, t.count, t.phase, c64(t.step.data()), c64(t.last.data()), t.q.data(),
          t.lipschitz};
}

// Illuminations per pass chosen by the library (0) and the automatic checkpoint interval (-1):
// every envelope kept, as src/cuda/bpm.cu stores them, when a pass's illuminations fit within
// half the free device memory, else a checkpoint every 16 slices; the tables carry the
// obliquity.
qh_bpm_options bpm_layout(int chunk) { return {1, chunk, -1}; }

class HipPipeline final : public Pipeline::Impl {
 public:
  HipPipeline(const PipelineOptions& o, const CropPlan& c) {
    const qh_pipeline_options po{o.count,
                                 o.hologram_size,
                                 {c.cy, c.cx, c.radius, c.out},
                                 illum(o.illuminations),
                                 geometry_of(o.geometry),
                                 odt_of(o.odt),
                                 o.upload_chunks};
    check(qh_hip_pipeline_create(&po, &p_));
    int nz = 0, ny = 0, nx = 0;
    check(qh_hip_pipeline_output_shape(p_, &nz, &ny, &nx));
    shape = {nz, ny, nx};
    input_size_ = static_cast<std::size_t>(o.count) * o.hologram_size * o.hologram_size;
    device_bytes = qh_hip_pipeline_device_bytes(p_);
  }
  ~HipPipeline() override { qh_hip_pipeline_destroy(p_); }
  HipPipeline(const HipPipeline&) = delete;
  HipPipeline& operator=(const HipPipeline&) = delete;

  void set_background(const Array3<float>& b) override {
    check(qh_hip_pipeline_set_background(p_, b.data.data()));
  }
  std::span<float> input() override { return {qh_hip_pipeline_input(p_), input_size_}; }
  void process() override {
    check(qh_hip_pipeline_process(p_));
    qh_pipeline_timings t{};
    check(qh_hip_pipeline_timings(p_, &t));
    timings = {t.upload, t.fields, t.rytov,  t.fft2,     t.clear, t.mapping,
               t.normalize, t.solve, t.finish, t.download, t.total};
    filled = qh_hip_pipeline_filled(p_);
  }
  [[nodiscard]] std::span<const float> output() const override {
    return {qh_hip_pipeline_output(p_),
            static_cast<std::size_t>(shape[0]) * shape[1] * shape[2]};
  }

 private:
  qh_pipeline* p_ = nullptr;
  std::size_t input_size_ = 0;
};

}  // namespace

bool device_present() { return qh_hip_device_present() != 0; }

std::unique_ptr<Pipeline::Impl> make_pipeline(const PipelineOptions& options, const CropPlan& crop,
                                              const detail::OdtGrid&) {
  return std::make_unique<HipPipeline>(options, crop);
}

Array3<cfloat> extract_fields(const Array3<float>& holograms, const CropPlan& plan) {
  const qh_crop_plan p{plan.cy, plan.cx, plan.radius, plan.out};
  Array3<cfloat> out(holograms.nz, plan.out, plan.out);
  check(qh_hip_extract_fields(holograms.data.data(), holograms.nz, holograms.nx, &p,
                              c64(out.data.data())));
  return out;
}

Array3<float> unwrap(const Array3<float>& wrapped) {
  Array3<float> out(wrapped.nz, wrapped.ny, wrapped.nx);
  check(qh_hip_unwrap(wrapped.data.data(), wrapped.nz, wrapped.ny, wrapped.nx, out.data.data()));
  return out;
}

Array3<cfloat> rytov(const Array3<cfloat>& ratio) {
  Array3<cfloat> out(ratio.nz, ratio.ny, ratio.nx);
  check(qh_hip_rytov(c64(ratio.data.data()), ratio.nz, ratio.ny, ratio.nx,
                     c64(out.data.data())));
  return out;
}

OdtResult reconstruct(const Array3<cfloat>& rytov_fields, std::span<const Illumination> illums,
                      const detail::OdtGrid& grid, const Geometry& geometry,
                      const OdtOptions& options, OdtTimings* timings, TvKernel) {
  const qh_odt_grid g{grid.n, grid.nz, grid.df, grid.dfz, grid.fm2, grid.fna2, grid.gain};
  const qh_geometry ge = geometry_of(geometry);
  const qh_odt_options op = odt_of(options);
  OdtResult r;
  r.ri = Array3<float>(grid.nz, grid.n, grid.n);
  qh_odt_timings t{};
  check(qh_hip_reconstruct(c64(rytov_fields.data.data()), rytov_fields.nz, rytov_fields.nx,
                           illum(illums), &g, &ge, &op, r.ri.data.data(), &r.filled,
                           timings ? &t : nullptr));
  if (timings) {
    *timings = {t.bytes, t.setup,     t.upload, t.fft2,   t.clear,
                t.mapping, t.normalize, t.solve, t.finish, t.download};
  }
  return r;
}

Array3<cfloat> bpm_forward(const Array3<float>& x, const BpmTables& t) {
  const qh_bpm_table_view v = view_of(t);
  const qh_bpm_options o = bpm_layout(0);
  Array3<cfloat> out(t.count, t.n, t.n);
  check(qh_hip_bpm_forward_tables(x.data.data(), &v, &o, c64(out.data.data())));
  return out;
}

BpmResult bpm_reconstruct(const Array3<cfloat>& fields, Array3<float> x, const BpmTables& t,
                          const BpmSolveOptions& options, int chunk) {
  const qh_bpm_table_view v = view_of(t);
  const qh_bpm_options o = bpm_layout(chunk);
  const qh_bpm_solve_options s{options.iterations, options.batch, options.step,
                               options.tv_weight, options.tv_inner};
  BpmResult r;
  r.loss.resize(static_cast<std::size_t>(options.iterations));
  check(qh_hip_bpm_reconstruct_tables(c64(fields.data.data()), x.data.data(), &v, &o, &s,
                                      r.loss.empty() ? nullptr : r.loss.data()));
  r.ri = std::move(x);
  return r;
}

Array3<float> bpm_vjp(const Array3<float>& x, const Array3<cfloat>& r, const BpmTables& t) {
  const qh_bpm_table_view v = view_of(t);
  const qh_bpm_options o = bpm_layout(0);
  Array3<float> grad(t.nz, t.n, t.n);
  check(qh_hip_bpm_vjp_tables(x.data.data(), c64(r.data.data()), &v, &o, grad.data.data()));
  return grad;
}

Array3<cfloat> bpm_jvp(const Array3<float>& x, const Array3<float>& h, const BpmTables& t) {
  const qh_bpm_table_view v = view_of(t);
  const qh_bpm_options o = bpm_layout(0);
  Array3<cfloat> out(t.count, t.n, t.n);
  check(qh_hip_bpm_jvp_tables(x.data.data(), h.data.data(), &v, &o, c64(out.data.data())));
  return out;
}

}  // namespace qoraholo::hip
