// One BPM inversion request on device-resident state (see bpm_inversion.hpp).
#include "bpm_inversion.hpp"

#include "fft.hpp"
#include "hostio.hpp"

namespace qh {

BpmInversion::BpmInversion(const BpmInversionSpec& spec)
    : n_medium_(spec.n_medium),
      plane_(static_cast<long>(spec.tables.n) * spec.tables.n),
      model_(spec.tables, spec.chunk, spec.checkpoint),
      solver_(model_, spec.solve),
      start_(spec.grid, spec.tables.count, spec.start),
      unwrap_(spec.tables.count, spec.tables.n, spec.tables.n),
      fields_(static_cast<std::size_t>(spec.tables.count) * plane_),
      work_(fields_.size()),
      psi_(fields_.size()),
      logamp_(fields_.size()),
      x_(static_cast<std::size_t>(spec.tables.nz) * plane_) {
  require(spec.grid.n == spec.tables.n && spec.grid.nz == spec.tables.nz,
          "BpmInversion: the start's grid must be the BPM volume");
  require(static_cast<int>(spec.illuminations.size()) == 2 * spec.tables.count,
          "BpmInversion: need one illumination per field");
  start_.set_illuminations(spec.illuminations.data());
  // Transform tables are built on first use (synchronous copies); build them now, outside any
  // request.
  for (int len : {spec.tables.n, spec.tables.n / 2, spec.tables.nz}) fft_twiddles(len);
}

void BpmInversion::run(const float2* fields, float* ri, double* loss) {
  const hipStream_t s = stream_.get();
  const long total = static_cast<long>(fields_.size());
  const long vol = static_cast<long>(x_.size());
  upload(fields_.get(), fields, fields_.bytes(), s);
  // Rytov fields of the ratios (reference rytov(), src/cpu/unwrap.cpp:61-82), then the start.
  split_ratio(fields_.get(), total, psi_.get(), logamp_.get(), s);
  unwrap_.rytov(psi_.get(), logamp_.get(), work_.get(), s);
  start_.run(work_.get(), s);
  bpm_offset(start_.index(), x_.get(), vol, -n_medium_, s);  // contrast x = ri - n_medium
  solver_.run(fields_.get(), x_.get(), loss, s);
  bpm_offset(x_.get(), x_.get(), vol, n_medium_, s);
  download(ri, x_.get(), x_.bytes(), s);
}

}  // namespace qh
