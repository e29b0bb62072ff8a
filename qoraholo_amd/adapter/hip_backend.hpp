// Backend::hip entry points for the reference's dispatch layer: the declarations a maintainer
// adds to src/backends.hpp beside `namespace cuda` (src/backends.hpp:107-126), with the same
// signatures, so api.cpp and pipeline.cpp call hip:: exactly where they call cuda:: today.
// Built into the reference's library with hip_backend.cpp, linked against libqoraholo_hip.so.
#pragma once

#include <memory>
#include <span>

#include "backends.hpp"

namespace qoraholo::hip {

bool device_present();
std::unique_ptr<Pipeline::Impl> make_pipeline(const PipelineOptions& options, const CropPlan& crop,
                                              const detail::OdtGrid& grid);
Array3<cfloat> extract_fields(const Array3<float>& holograms, const CropPlan& plan);
Array3<float> unwrap(const Array3<float>& wrapped);
Array3<cfloat> rytov(const Array3<cfloat>& ratio);
// tv_kernel selects among the CUDA backend's FGP kernels for its benchmark (all do the same
// arithmetic); the HIP backend has one FGP implementation, two steps per pass and equal to
// single steps bit for bit, which every value runs.
OdtResult reconstruct(const Array3<cfloat>& rytov_fields, std::span<const Illumination> illum,
                      const detail::OdtGrid& grid, const Geometry& geometry,
                      const OdtOptions& options, OdtTimings* timings = nullptr,
                      TvKernel tv_kernel = TvKernel::fused);
Array3<cfloat> bpm_forward(const Array3<float>& x, const BpmTables& t);
BpmResult bpm_reconstruct(const Array3<cfloat>& fields, Array3<float> x, const BpmTables& t,
                          const BpmSolveOptions& options, int chunk = 0);
Array3<float> bpm_vjp(const Array3<float>& x, const Array3<cfloat>& r, const BpmTables& t);
// The device form of cpu::bpm_jvp (src/backends.hpp:100), which has no CUDA counterpart.
Array3<cfloat> bpm_jvp(const Array3<float>& x, const Array3<float>& h, const BpmTables& t);

}  // namespace qoraholo::hip
