// The C-ABI: argument checks with the reference's rules, status codes and per-thread error
// messages, and the per-call entry points that copy host arrays in and out.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <complex>
#include <numbers>
#include <cstring>
#include <limits>
#include <memory>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "../../include/qoraholo_hip.h"
#include "bpm.hpp"
#include "fft.hpp"
#include "pipeline.hpp"
#include "hostio.hpp"
#include "runtime.hpp"
#include "server.hpp"
#include "stages.hpp"
#include "status.hpp"

struct qh_pipeline {
  std::unique_ptr<qh::Pipeline> impl;
};

struct qh_server {
  std::unique_ptr<qh::Server> impl;
};

struct qh_bpm_plan {
  std::unique_ptr<qh::BpmModel> impl;
};

namespace qh {

std::string& last_error() {
  thread_local std::string message;
  return message;
}

namespace {

// Per-thread stream of the per-call entry points, so calls from different host threads run
// independently.
hipStream_t call_stream() {
  thread_local std::unique_ptr<Stream> s;
  if (!s) s = std::make_unique<Stream>();
  return s->get();
}

void require_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) throw DeviceError("no HIP device available");
}

void require_even_square(int ny, int nx, const char* who) {
  require(ny == nx && nx >= 4 && nx % 2 == 0,
          std::string(who) + ": images must be square with an even side");
}

void require_fft(int n, const char* who) {
  require(fft_supported(n), std::string(who) + ": the HIP backend transforms sides of the form "
                                               "2^a 3^b 5^c, got " + std::to_string(n));
}

OdtGrid make_grid(int n, int nz, const qh_geometry& g) {
  const float fm = g.n_medium / g.wavelength;
  const float fna = g.na / g.wavelength;
  return {n, nz, 1.0f / (static_cast<float>(n) * g.pixel), 1.0f / (static_cast<float>(nz) * g.pixel),
          fm * fm, fna * fna, 4.0f * kPi / g.pixel};
}

// src/api.cpp:107-123 (odt_grid): checks of reconstruct() and of the Pipeline.
OdtGrid checked_grid(int n, int count, const qh_illumination* illum, const qh_geometry& g,
                     const qh_odt_options& o, const char* who) {
  require_even_square(n, n, who);
  require(count >= 1 && illum != nullptr, std::string(who) + ": need one illumination per field");
  require(g.wavelength > 0 && g.pixel > 0 && g.n_medium > 0 && g.na > 0,
          std::string(who) + ": geometry values must be positive");
  const int nz = o.nz > 0 ? o.nz : n;
  require(nz % 2 == 0, std::string(who) + ": nz must be even");
  require(o.iterations >= 0, std::string(who) + ": iterations must not be negative");
  require(o.solver == 0 || o.solver == 1, std::string(who) + ": unknown solver");
  require(o.gp_transform == 0 || o.gp_transform == 1, std::string(who) + ": unknown gp_transform");
  require(o.solver != 1 || o.iterations == 0 || (o.tv_weight > 0 && o.tv_inner >= 0),
          std::string(who) + ": TV needs tv_weight > 0");
  const float fm = g.n_medium / g.wavelength;
  for (int a = 0; a < count; ++a) {
    require(illum[a].fy * illum[a].fy + illum[a].fx * illum[a].fx < fm * fm,
            std::string(who) + ": illumination frequency beyond n_medium / wavelength");
  }
  require_fft(n, who);
  require_fft(nz, who);
  return make_grid(n, nz, g);
}

void check_plan(const qh_crop_plan& p, int n) {
  require(p.out % 2 == 0 && p.out <= n && p.out >= 4,
          "extract_fields: output_size must be even and <= N");
  require(p.radius > 0 && p.radius <= static_cast<float>(p.out) / 2.0f,
          "extract_fields: crop radius must fit in the output window");
}

// src/cpu/field.cpp:24-33 (refine): Jacobsen's three-bin estimate with Candan's tan(pi/N)/(pi/N)
// correction, clamped to half a bin; float arithmetic as the reference's host code.
float refine_peak(std::complex<float> lo, std::complex<float> mid, std::complex<float> hi, int n) {
  const std::complex<float> den = 2.0f * mid - lo - hi;
  if (std::abs(den) == 0.0f) return 0.0f;
  const float c = std::numbers::pi_v<float> / static_cast<float>(n);
  const float d = ((lo - hi) / den).real() * std::tan(c) / c;
  return std::clamp(d, -0.5f, 0.5f);
}

// Sideband (iy, ix, y, x, magnitude) of every hologram of a host stack.
std::vector<std::array<float, 5>> device_sidebands(const float* holo, int count, int n,
                                                   float dc_exclusion) {
  require(holo != nullptr && count >= 1, "find_sideband: need at least one hologram");
  require(n >= 4 && n % 2 == 0, "find_sideband: holograms must be square with an even side");
  require_fft(n, "find_sideband");
  if (dc_exclusion <= 0) dc_exclusion = static_cast<float>(n) / 8.0f;
  require_device();
  const hipStream_t s = call_stream();
  const std::size_t plane = static_cast<std::size_t>(n) * n;
  DeviceBuffer<float> dh(plane * count);
  DeviceBuffer<float2> half(static_cast<std::size_t>(count) * n * (n / 2 + 1));
  DeviceBuffer<SidebandPeak> peaks(count);
  upload(dh.get(), holo, dh.bytes(), s);
  sideband_peaks(dh.get(), count, n, dc_exclusion, half.get(), peaks.get(), s);
  std::vector<SidebandPeak> h(count);
  download(h.data(), peaks.get(), peaks.bytes(), s);
  check(hipStreamSynchronize(s), "find_sideband");
  std::vector<std::array<float, 5>> out(count);
  for (int a = 0; a < count; ++a) {
    const SidebandPeak& p = h[a];
    if (p.mag == 0.0f) throw DeviceError("find_sideband: no peak outside the DC region");
    const auto cx = [](float2 v) { return std::complex<float>(v.x, v.y); };
    out[a] = {static_cast<float>(p.iy), static_cast<float>(p.ix),
              static_cast<float>(p.iy) + refine_peak(cx(p.s[1]), cx(p.s[0]), cx(p.s[2]), n),
              static_cast<float>(p.ix) + refine_peak(cx(p.s[3]), cx(p.s[0]), cx(p.s[4]), n),
              p.mag};
  }
  return out;
}

// src/api.cpp:141-157 (checked_bpm_tables): checks of bpm_forward and bpm_reconstruct.
BpmTables checked_bpm_tables(int nz, int n, int count, const qh_illumination* illum,
                             const qh_geometry* g, const qh_bpm_options* o, const char* who) {
  const std::string w(who);
  require(g != nullptr && o != nullptr, w + ": missing geometry or options");
  require(n >= 4 && n % 2 == 0 && nz >= 1, w + ": volume must be (nz, n, n) with an even n");
  require(count >= 1 && illum != nullptr, w + ": need one illumination per field");
  require(g->wavelength > 0 && g->pixel > 0 && g->n_medium > 0 && g->na > 0,
          w + ": geometry values must be positive");
  const float fm = g->n_medium / g->wavelength;
  for (int a = 0; a < count; ++a) {
    require(illum[a].fy * illum[a].fy + illum[a].fx * illum[a].fx < fm * fm,
            w + ": illumination frequency beyond n_medium / wavelength");
  }
  require(o->chunk >= 0 && o->checkpoint >= -1,
          w + ": chunk must not be negative, checkpoint at least -1");
  require_fft(n, who);
  return bpm_tables(n, nz, illum, count, *g, o->obliquity != 0);
}

// BpmTables from the caller's arrays (the layout of src/backends.hpp:61-70), with the shape
// checks of checked_bpm_tables.
BpmTables tables_from_view(const qh_bpm_table_view* v, const qh_bpm_options* o, const char* who) {
  const std::string w(who);
  require(v != nullptr && o != nullptr, w + ": missing tables or options");
  require(v->n >= 4 && v->n % 2 == 0 && v->nz >= 1, w + ": volume must be (nz, n, n) with an even n");
  require(v->count >= 1, w + ": need one illumination per field");
  require(v->step != nullptr && v->last != nullptr && v->q != nullptr, w + ": missing table arrays");
  require(o->chunk >= 0 && o->checkpoint >= -1,
          w + ": chunk must not be negative, checkpoint at least -1");
  require_fft(v->n, who);
  BpmTables t;
  t.n = v->n;
  t.nz = v->nz;
  t.count = v->count;
  t.phase = v->phase;
  t.lipschitz = v->lipschitz;
  const std::size_t total = static_cast<std::size_t>(v->count) * v->n * v->n;
  const float2* step = reinterpret_cast<const float2*>(v->step);
  const float2* last = reinterpret_cast<const float2*>(v->last);
  t.step.assign(step, step + total);
  t.last.assign(last, last + total);
  t.q.assign(v->q, v->q + total);
  return t;
}

// Complex float division as the reference's host build evaluates std::complex<float> / (GCC's
// runtime on x86-64, which the reference's CPU loop calls): numerator and denominator in double,
// the simple formula (double carries the range and precision Smith's scaling would otherwise
// buy), each part rounded once to float; NaN results recovered as in C99 Annex G.3.
std::complex<float> cdiv(std::complex<float> num, std::complex<float> den) {
  float a = num.real(), b = num.imag(), c = den.real(), d = den.imag();
  const double aa = a, bb = b, cc = c, dd = d;
  const double denom = cc * cc + dd * dd;
  float x = static_cast<float>((aa * cc + bb * dd) / denom);
  float y = static_cast<float>((bb * cc - aa * dd) / denom);
  if (std::isnan(x) && std::isnan(y)) {
    const float inf = std::numeric_limits<float>::infinity();
    if (c == 0.0f && d == 0.0f && (!std::isnan(a) || !std::isnan(b))) {
      x = std::copysign(inf, c) * a;
      y = std::copysign(inf, c) * b;
    } else if ((std::isinf(a) || std::isinf(b)) && std::isfinite(c) && std::isfinite(d)) {
      a = std::copysign(std::isinf(a) ? 1.0f : 0.0f, a);
      b = std::copysign(std::isinf(b) ? 1.0f : 0.0f, b);
      x = inf * (a * c + b * d);
      y = inf * (b * c - a * d);
    } else if ((std::isinf(c) || std::isinf(d)) && std::isfinite(a) && std::isfinite(b)) {
      c = std::copysign(std::isinf(c) ? 1.0f : 0.0f, c);
      d = std::copysign(std::isinf(d) ? 1.0f : 0.0f, d);
      x = 0.0f * (a * c + b * d);
      y = 0.0f * (b * c - a * d);
    }
  }
  return {x, y};
}

}  // namespace

// Checks of the Pipeline constructor (src/pipeline.cpp:72-91) and of odt_grid.
void check_pipeline_options(const qh_pipeline_options& o) {
  const int n = o.hologram_size;
  require(o.count >= 1, "Pipeline: count must be positive");
  require(n >= 4 && n % 2 == 0, "Pipeline: hologram_size must be even");
  require(o.upload_chunks >= 1 && o.upload_chunks <= o.count,
          "Pipeline: upload_chunks must be between 1 and count");
  check_plan(o.crop, n);
  require_fft(n, "Pipeline");
  require_fft(o.crop.out, "Pipeline");
  checked_grid(o.crop.out, o.count, o.illuminations, o.geometry, o.odt, "Pipeline");
}

OdtGrid grid_from(const qh_pipeline_options& o) {
  return checked_grid(o.crop.out, o.count, o.illuminations, o.geometry, o.odt, "Pipeline");
}

OdtConfig config_from(const qh_odt_options& o, const qh_geometry& g) {
  OdtConfig c;
  c.iterations = o.iterations;
  c.solver = o.solver;
  c.gp_transform = o.gp_transform;
  c.tv_weight = o.tv_weight;
  c.tv_inner = o.tv_inner;
  c.wavelength = g.wavelength;
  c.n_medium = g.n_medium;
  return c;
}

namespace {

void bpm_forward_call(const BpmTables& t, const float* x, const qh_bpm_options* o, qh_c64* out) {
  require_device();
  const hipStream_t s = call_stream();
  BpmModel m(t, o->chunk, o->checkpoint);
  const std::size_t vol = static_cast<std::size_t>(t.nz) * t.n * t.n;
  DeviceBuffer<float> dx(vol);
  DeviceBuffer<float2> dy(static_cast<std::size_t>(t.count) * t.n * t.n);
  upload(dx.get(), x, dx.bytes(), s);
  m.forward(dx.get(), dy.get(), s);
  download(out, dy.get(), dy.bytes(), s);
  check(hipStreamSynchronize(s), "bpm_forward");
}

void bpm_vjp_call(const BpmTables& t, const float* x, const qh_c64* r, const qh_bpm_options* o,
                  float* grad) {
  require_device();
  const hipStream_t s = call_stream();
  BpmModel m(t, o->chunk, o->checkpoint);
  const std::size_t vol = static_cast<std::size_t>(t.nz) * t.n * t.n;
  DeviceBuffer<float> dx(vol), dg(vol);
  DeviceBuffer<float2> dr(static_cast<std::size_t>(t.count) * t.n * t.n);
  upload(dx.get(), x, dx.bytes(), s);
  upload(dr.get(), r, dr.bytes(), s);
  m.vjp(dx.get(), dr.get(), dg.get(), s);
  download(grad, dg.get(), dg.bytes(), s);
  check(hipStreamSynchronize(s), "bpm_vjp");
}

void bpm_jvp_call(const BpmTables& t, const float* x, const float* h, const qh_bpm_options* o,
                  qh_c64* out) {
  require_device();
  const hipStream_t s = call_stream();
  BpmModel m(t, o->chunk, o->checkpoint);
  const std::size_t vol = static_cast<std::size_t>(t.nz) * t.n * t.n;
  DeviceBuffer<float> dx(vol), dh(vol);
  DeviceBuffer<float2> dy(static_cast<std::size_t>(t.count) * t.n * t.n);
  upload(dx.get(), x, dx.bytes(), s);
  upload(dh.get(), h, dh.bytes(), s);
  m.jvp(dx.get(), dh.get(), dy.get(), s);
  download(out, dy.get(), dy.bytes(), s);
  check(hipStreamSynchronize(s), "bpm_jvp");
}

void check_solve_options(const qh_bpm_solve_options* solve, int n, const char* who) {
  const std::string w(who);
  require(solve != nullptr, w + ": missing solver options");
  require(solve->iterations >= 0 && solve->step > 0 && solve->tv_weight >= 0 &&
              solve->tv_inner >= 1 && solve->batch >= 0,
          w + ": invalid solver options");
  require(solve->tv_weight == 0 || n <= 512, w + ": the TV prox handles n <= 512");
}

void bpm_reconstruct_call(const BpmTables& t, const qh_c64* fields, float* x,
                          const qh_bpm_options* o, const qh_bpm_solve_options* solve,
                          double* loss) {
  require_device();
  const hipStream_t s = call_stream();
  BpmModel m(t, o->chunk, o->checkpoint);
  const std::size_t vol = static_cast<std::size_t>(t.nz) * t.n * t.n;
  DeviceBuffer<float> dx(vol);
  DeviceBuffer<float2> df(static_cast<std::size_t>(t.count) * t.n * t.n);
  upload(dx.get(), x, dx.bytes(), s);
  upload(df.get(), fields, df.bytes(), s);
  BpmSolveConfig c;
  c.iterations = solve->iterations;
  c.batch = solve->batch;
  c.step = solve->step;
  c.tv_weight = solve->tv_weight;
  c.tv_inner = solve->tv_inner;
  std::vector<double> l(std::max(c.iterations, 1));
  bpm_solve(m, df.get(), dx.get(), c, l.data(), s);
  download(x, dx.get(), dx.bytes(), s);
  check(hipStreamSynchronize(s), "bpm_reconstruct");
  if (c.iterations > 0) std::memcpy(loss, l.data(), c.iterations * sizeof(double));
}

}  // namespace

}  // namespace qh

using namespace qh;

extern "C" {

const char* qh_hip_last_error(void) { return last_error().c_str(); }

int qh_hip_device_present(void) {
  int n = 0;
  return hipGetDeviceCount(&n) == hipSuccess && n > 0 ? 1 : 0;
}

int qh_make_grid(int n, int nz, const qh_geometry* geometry, qh_odt_grid* out) {
  return guarded([&] {
    require(geometry && out, "make_grid: null argument");
    const OdtGrid g = make_grid(n, nz, *geometry);
    *out = {g.n, g.nz, g.df, g.dfz, g.fm2, g.fna2, g.gain};
  });
}

int qh_resolve_crop(float center_y, float center_x, float radius, int output_size, int n,
                 qh_crop_plan* out) {
  return guarded([&] {
    require(out != nullptr, "crop_plan: null argument");
    qh_crop_plan p{static_cast<int>(std::lround(center_y)), static_cast<int>(std::lround(center_x)),
                   radius, output_size > 0 ? output_size : n};
    if (p.radius <= 0) p.radius = std::hypot(center_y, center_x) / 3.0f;
    check_plan(p, n);
    *out = p;
  });
}

int qh_circular_scan(int count, float na_illumination, const qh_geometry* geometry,
                     qh_illumination* out) {
  return guarded([&] {
    require(geometry && out, "circular_scan: null argument");
    require(count >= 1, "circular_scan: count must be positive");
    require(na_illumination >= 0 && na_illumination < geometry->n_medium,
            "circular_scan: na_illumination must be below n_medium");
    out[0] = {0.0f, 0.0f};
    const float f = na_illumination / geometry->wavelength;
    for (int i = 1; i < count; ++i) {
      const float phi = 2.0f * kPi * static_cast<float>(i - 1) / static_cast<float>(count - 1);
      out[i] = {f * std::sin(phi), f * std::cos(phi)};
    }
  });
}

int qh_field_ratio(const qh_c64* sample, int count, const qh_c64* background, int bcount,
                   int plane, qh_c64* out) {
  return guarded([&] {
    require(sample && background && out, "field_ratio: null argument");
    require(count >= 1 && plane >= 1 && (bcount == 1 || bcount == count),
            "field_ratio: background shape does not match");
    using C = std::complex<float>;
    const auto* s = reinterpret_cast<const C*>(sample);
    const auto* b = reinterpret_cast<const C*>(background);
    auto* o = reinterpret_cast<C*>(out);
    const std::size_t total = static_cast<std::size_t>(count) * plane;
    const int parts = static_cast<int>(
        std::clamp<std::size_t>(total >> 16, 1, static_cast<std::size_t>(host_copy_threads())));
    const std::size_t step = (total + parts - 1) / parts;
    parallel_for(parts, [&](int p) {
      const std::size_t end = std::min(total, (p + 1) * step);
      for (std::size_t i = p * step; i < end;) {
        const std::size_t z = i / plane;
        const std::size_t stop = std::min(end, (z + 1) * plane);
        const C* bz = b + (bcount == 1 ? 0 : z * plane);
        const std::size_t z0 = z * plane;
        for (; i < stop; ++i) o[i] = cdiv(s[i], bz[i - z0]);
      }
    });
  });
}

int qh_host_copy(void* dst, const void* src, size_t bytes) {
  return guarded([&] {
    require(bytes == 0 || (dst && src), "host_copy: null buffer");
    host_copy(dst, src, bytes);
  });
}

int qh_gaussian_noise(uint64_t seed, float sigma, size_t count, float* out) {
  return guarded([&] {
    require(out != nullptr || count == 0, "gaussian_noise: null output");
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> noise(0.0f, sigma);
    for (size_t i = 0; i < count; ++i) out[i] = noise(rng);
  });
}

int qh_hip_extract_fields(const float* holograms, int count, int n, const qh_crop_plan* plan,
                          qh_c64* fields) {
  return guarded([&] {
    require(holograms && plan && fields, "extract_fields: null argument");
    require_even_square(n, n, "extract_fields");
    require(count >= 1, "extract_fields: empty stack");
    check_plan(*plan, n);
    require_fft(n, "extract_fields");
    require_fft(plan->out, "extract_fields");
    require_device();
    const hipStream_t s = call_stream();
    const int m = plan->out;
    const std::size_t image = static_cast<std::size_t>(n) * n;
    DeviceBuffer<float> holo(count * image);
    DeviceBuffer<float2> half(static_cast<std::size_t>(count) * n * (n / 2 + 1));
    DeviceBuffer<float2> out(static_cast<std::size_t>(count) * m * m);
    upload(holo.get(), holograms, holo.bytes(), s);
    hologram_spectra(holo.get(), count, n, half.get(), s);
    const CropPlan cp{plan->cy, plan->cx, plan->radius, m};
    crop_fields(half.get(), count, n, cp, out.get(), s);
    fields_inverse(out.get(), count, m, 1.0f / static_cast<float>(image), true, s);
    download(fields, out.get(), out.bytes(), s);
    check(hipStreamSynchronize(s), "extract_fields");
  });
}

int qh_hip_unwrap(const float* wrapped, int nz, int ny, int nx, float* out) {
  return guarded([&] {
    require(wrapped && out, "unwrap: null argument");
    require(nz >= 1 && ny % 2 == 0 && nx % 2 == 0 && ny >= 2 && nx >= 2,
            "unwrap: image sides must be even");
    require_fft(ny, "unwrap");
    require_fft(nx, "unwrap");
    require_device();
    const hipStream_t s = call_stream();
    const std::size_t total = static_cast<std::size_t>(nz) * ny * nx;
    DeviceBuffer<float> psi(total);
    DeviceBuffer<float> phi(total);
    Unwrapper u(nz, ny, nx);
    upload(psi.get(), wrapped, psi.bytes(), s);
    u.unwrap(psi.get(), phi.get(), s);
    download(out, phi.get(), phi.bytes(), s);
    check(hipStreamSynchronize(s), "unwrap");
  });
}

int qh_hip_rytov(const qh_c64* ratio, int nz, int ny, int nx, qh_c64* out) {
  return guarded([&] {
    require(ratio && out, "rytov: null argument");
    require(nz >= 1 && ny % 2 == 0 && nx % 2 == 0 && ny >= 4 && nx >= 4,
            "rytov: image sides must be even");
    require_fft(ny, "rytov");
    require_fft(nx, "rytov");
    require_device();
    const hipStream_t s = call_stream();
    const std::size_t total = static_cast<std::size_t>(nz) * ny * nx;
    DeviceBuffer<float2> r(total);
    DeviceBuffer<float> psi(total);
    DeviceBuffer<float> logamp(total);
    Unwrapper u(nz, ny, nx);
    upload(r.get(), ratio, r.bytes(), s);
    split_ratio(r.get(), static_cast<long>(total), psi.get(), logamp.get(), s);
    u.rytov(psi.get(), logamp.get(), r.get(), s);
    download(out, r.get(), r.bytes(), s);
    check(hipStreamSynchronize(s), "rytov");
  });
}

int qh_hip_reconstruct(const qh_c64* fields, int count, int n, const qh_illumination* illum,
                       const qh_odt_grid* grid, const qh_geometry* geometry,
                       const qh_odt_options* options, float* ri, double* filled,
                       qh_odt_timings* timings) {
  return guarded([&] {
    require(fields && illum && geometry && options && ri, "reconstruct: null argument");
    const OdtGrid checked = checked_grid(n, count, illum, *geometry, *options, "reconstruct");
    OdtGrid g = checked;
    if (grid) g = {grid->n, grid->nz, grid->df, grid->dfz, grid->fm2, grid->fna2, grid->gain};
    require(g.n == n && g.nz == checked.nz, "reconstruct: grid does not match the fields");
    require_device();
    const auto t0 = std::chrono::steady_clock::now();
    std::size_t free_before = 0, free_after = 0, total_mem = 0;
    check(hipMemGetInfo(&free_before, &total_mem), "hipMemGetInfo");
    const hipStream_t s = call_stream();
    OdtSolver solver(g, count, config_from(*options, *geometry));
    std::vector<float> il(2 * static_cast<std::size_t>(count));
    for (int a = 0; a < count; ++a) {
      il[2 * a] = illum[a].fy;
      il[2 * a + 1] = illum[a].fx;
    }
    solver.set_illuminations(il.data());
    solver.time_fgp(timings != nullptr);
    const std::size_t plane = static_cast<std::size_t>(n) * n;
    DeviceBuffer<float2> f(plane * count);
    for (int len : {n, n / 2, g.nz}) fft_twiddles(len);
    check(hipMemGetInfo(&free_after, &total_mem), "hipMemGetInfo");
    const double setup =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    StageClock clock(kOdtMarks + 3);
    const int start = kOdtMarks, uploaded = kOdtMarks + 1, downloaded = kOdtMarks + 2;
    clock.mark(start, s);
    upload(f.get(), fields, f.bytes(), s);
    clock.mark(uploaded, s);
    solver.run(f.get(), s, &clock, 0);
    const std::size_t vol = plane * g.nz;
    unsigned long long nfilled = 0;
    download(ri, solver.index(), vol * sizeof(float), s);
    download(&nfilled, solver.filled_device(), sizeof nfilled, s);
    clock.mark(downloaded, s);
    check(hipStreamSynchronize(s), "reconstruct");
    if (filled) *filled = static_cast<double>(nfilled) / static_cast<double>(vol);
    if (timings) {
      timings->bytes = free_before > free_after ? free_before - free_after : 0;
      timings->setup = setup;
      timings->upload = clock.ms(start, uploaded);
      timings->fft2 = clock.ms(uploaded, kFft2);
      timings->clear = clock.ms(kFft2, kClear);
      timings->mapping = clock.ms(kClear, kMapping);
      timings->normalize = clock.ms(kMapping, kNormalize);
      timings->solve = clock.ms(kNormalize, kSolve);
      timings->finish = clock.ms(kSolve, kFinish);
      timings->download = clock.ms(kFinish, downloaded);
      timings->fgp_ms = solver.collect_fgp_ms();
      timings->fgp_steps = solver.fgp_steps();
      timings->fgp_launches = solver.fgp_launches();
    }
  });
}

int qh_hip_map_kspace(const qh_c64* fields, int count, int n, const qh_illumination* illum,
                      const qh_odt_grid* grid, qh_c64* potential, uint16_t* counts,
                      double* filled) {
  return guarded([&] {
    require(fields && illum && grid && potential && counts, "map_kspace: null argument");
    require_even_square(n, n, "map_kspace");
    require(grid->n == n && grid->nz % 2 == 0 && grid->nz >= 2, "map_kspace: bad grid");
    require(count >= 1, "map_kspace: need one illumination per field");
    require_fft(n, "map_kspace");
    require_device();
    const hipStream_t s = call_stream();
    const OdtGrid g{grid->n, grid->nz, grid->df, grid->dfz, grid->fm2, grid->fna2, grid->gain};
    OdtSolver solver(g, count, OdtConfig{});
    std::vector<float> il(2 * static_cast<std::size_t>(count));
    for (int a = 0; a < count; ++a) {
      il[2 * a] = illum[a].fy;
      il[2 * a + 1] = illum[a].fx;
    }
    solver.set_illuminations(il.data());
    const std::size_t plane = static_cast<std::size_t>(n) * n;
    DeviceBuffer<float2> f(plane * count);
    upload(f.get(), fields, f.bytes(), s);
    solver.map_only(f.get(), s);
    const std::size_t vol = plane * g.nz;
    unsigned long long nfilled = 0;
    download(potential, solver.potential(), vol * sizeof(float2), s);
    download(counts, solver.counts(), vol * sizeof(uint16_t), s);
    download(&nfilled, solver.filled_device(), sizeof nfilled, s);
    check(hipStreamSynchronize(s), "map_kspace");
    if (filled) *filled = static_cast<double>(nfilled) / static_cast<double>(vol);
  });
}

int qh_hip_pipeline_create(const qh_pipeline_options* o, qh_pipeline** out) {
  return guarded([&] {
    require(o && out, "Pipeline: null argument");
    *out = nullptr;
    check_pipeline_options(*o);
    require_device();
    const auto t0 = std::chrono::steady_clock::now();
    std::size_t free_before = 0, free_after = 0, total_mem = 0;
    check(hipMemGetInfo(&free_before, &total_mem), "hipMemGetInfo");
    auto p = std::make_unique<qh_pipeline>();
    p->impl = std::make_unique<Pipeline>(*o);
    check(hipDeviceSynchronize(), "Pipeline");
    check(hipMemGetInfo(&free_after, &total_mem), "hipMemGetInfo");
    p->impl->set_setup(
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(),
        free_before > free_after ? free_before - free_after : 0);
    *out = p.release();
  });
}

void qh_hip_pipeline_destroy(qh_pipeline* p) { delete p; }

int qh_hip_pipeline_set_background(qh_pipeline* p, const float* backgrounds) {
  return guarded([&] {
    require(p && backgrounds, "Pipeline::set_background: null argument");
    p->impl->set_background(backgrounds);
  });
}

float* qh_hip_pipeline_input(qh_pipeline* p) { return p ? p->impl->input() : nullptr; }

const float* qh_hip_pipeline_output(const qh_pipeline* p) { return p ? p->impl->output() : nullptr; }

int qh_hip_pipeline_process(qh_pipeline* p) {
  return guarded([&] {
    require(p != nullptr, "Pipeline: null handle");
    p->impl->process();
  });
}

int qh_hip_pipeline_run(qh_pipeline* p, const float* holograms, float* ri) {
  return guarded([&] {
    require(p != nullptr, "Pipeline: null handle");
    require(holograms != nullptr && ri != nullptr, "Pipeline: null buffer");
    p->impl->run(holograms, ri);
  });
}

int qh_hip_pipeline_process_device(qh_pipeline* p, const float* holograms_device, float* ri_device,
                                   void* stream) {
  return guarded([&] {
    require(p && holograms_device && ri_device, "Pipeline: null argument");
    p->impl->process_device(holograms_device, ri_device, static_cast<hipStream_t>(stream));
  });
}

double qh_hip_pipeline_filled(const qh_pipeline* p) { return p ? p->impl->filled() : 0.0; }

int qh_hip_pipeline_timings(const qh_pipeline* p, qh_pipeline_timings* out) {
  return guarded([&] {
    require(p && out, "Pipeline: null argument");
    *out = const_cast<qh_pipeline*>(p)->impl->timings();
  });
}

double qh_hip_pipeline_setup_ms(const qh_pipeline* p) { return p ? p->impl->setup_ms() : 0.0; }

size_t qh_hip_pipeline_device_bytes(const qh_pipeline* p) { return p ? p->impl->device_bytes() : 0; }

int qh_hip_pipeline_output_shape(const qh_pipeline* p, int* nz, int* ny, int* nx) {
  return guarded([&] {
    require(p && nz && ny && nx, "Pipeline: null argument");
    *nz = p->impl->nz();
    *ny = p->impl->n();
    *nx = p->impl->n();
  });
}

int qh_hip_pipeline_time_fgp(qh_pipeline* p, int on) {
  return guarded([&] {
    require(p != nullptr, "Pipeline: null handle");
    p->impl->time_fgp(on != 0);
  });
}

int qh_hip_pipeline_use_graph(qh_pipeline* p, int on) {
  return guarded([&] {
    require(p != nullptr, "Pipeline: null handle");
    p->impl->use_graph(on != 0);
  });
}

int qh_hip_server_create(size_t budget_bytes, int max_slots, qh_server** out) {
  return guarded([&] {
    require(out != nullptr, "Server: null argument");
    *out = nullptr;
    require(max_slots >= 1, "Server: max_slots must be positive");
    require_device();
    auto s = std::make_unique<qh_server>();
    s->impl = std::make_unique<Server>(budget_bytes, max_slots);
    *out = s.release();
  });
}

void qh_hip_server_destroy(qh_server* s) { delete s; }

int qh_hip_server_add_geometry(qh_server* s, const qh_pipeline_options* options,
                               const float* backgrounds, int* geometry) {
  return guarded([&] {
    require(s && options && backgrounds && geometry, "Server: null argument");
    *geometry = s->impl->add_geometry(*options, backgrounds);
  });
}

int qh_hip_server_add_geometry_batched(qh_server* s, const qh_pipeline_options* options,
                                       const float* backgrounds, int max_batch, int* geometry) {
  return guarded([&] {
    require(s && options && backgrounds && geometry, "Server: null argument");
    *geometry = s->impl->add_geometry(*options, backgrounds, max_batch);
  });
}

int qh_hip_server_submit(qh_server* s, int geometry, const float* holograms, float* ri_out,
                         int64_t* id) {
  return guarded([&] {
    require(s && holograms && ri_out && id, "Server: null argument");
    *id = s->impl->submit(geometry, Server::kReconstruction, holograms, ri_out);
  });
}

int qh_hip_server_add_bpm(qh_server* s, const qh_bpm_request_options* o, int* geometry) {
  return guarded([&] {
    require(s && o && geometry, "Server: null argument");
    const qh_bpm_solve_options& v = o->solve;
    require(v.iterations >= 0 && v.step > 0 && v.tv_weight >= 0 && v.tv_inner >= 1 && v.batch >= 0,
            "Server: invalid BPM solver options");
    require(v.tv_weight == 0 || o->n <= 512, "Server: the TV prox handles n <= 512");
    require(o->start_iterations >= 0, "Server: start iterations must not be negative");
    BpmInversionSpec spec;
    spec.tables = checked_bpm_tables(o->nz, o->n, o->count, o->illuminations, &o->geometry,
                                     &o->bpm, "Server");
    spec.chunk = o->bpm.chunk;
    spec.checkpoint = o->bpm.checkpoint;
    qh_odt_options start{};
    start.nz = o->nz;
    start.iterations = o->start_iterations;
    start.solver = 0;
    start.gp_transform = 0;
    start.tv_weight = 1e-2f;
    start.tv_inner = 10;
    spec.grid = checked_grid(o->n, o->count, o->illuminations, o->geometry, start, "Server");
    spec.start = config_from(start, o->geometry);
    spec.illuminations.resize(2 * static_cast<std::size_t>(o->count));
    for (int a = 0; a < o->count; ++a) {
      spec.illuminations[2 * a] = o->illuminations[a].fy;
      spec.illuminations[2 * a + 1] = o->illuminations[a].fx;
    }
    spec.solve.iterations = v.iterations;
    spec.solve.batch = v.batch;
    spec.solve.step = v.step;
    spec.solve.tv_weight = v.tv_weight;
    spec.solve.tv_inner = v.tv_inner;
    spec.n_medium = o->geometry.n_medium;
    require_device();
    *geometry = s->impl->add_bpm(spec);
  });
}

int qh_hip_server_submit_bpm(qh_server* s, int geometry, const qh_c64* fields, float* ri_out,
                             double* loss, int64_t* id) {
  return guarded([&] {
    require(s && fields && ri_out && id, "Server: null argument");
    *id = s->impl->submit(geometry, Server::kBpm, fields, ri_out, loss);
  });
}

int qh_hip_server_wait(qh_server* s, int64_t id, double* filled, double* latency_ms) {
  int status = QH_OK;
  const int rc = guarded([&] {
    require(s != nullptr, "Server: null handle");
    std::string error;
    status = s->impl->wait(id, filled, latency_ms, &error);
    if (status != QH_OK) throw std::runtime_error(error);
  });
  if (rc != QH_OK && status != QH_OK) return status;
  return rc;
}

int qh_hip_server_poll(qh_server* s, int64_t id, int* done) {
  return guarded([&] {
    require(s && done, "Server: null argument");
    *done = s->impl->done(id) ? 1 : 0;
  });
}

int qh_hip_server_stats(qh_server* s, qh_server_stats* out) {
  return guarded([&] {
    require(s && out, "Server: null argument");
    *out = s->impl->stats();
  });
}

int qh_hip_server_reset_stats(qh_server* s) {
  return guarded([&] {
    require(s != nullptr, "Server: null handle");
    s->impl->reset_stats();
  });
}

int qh_hip_fft_c2c(qh_c64* data, int d0, int d1, int d2, unsigned axes, int forward, float scale) {
  return guarded([&] {
    require(data && d0 >= 1 && d1 >= 1 && d2 >= 1, "fft: bad arguments");
    if (axes & 1u) require_fft(d0, "fft");
    if (axes & 2u) require_fft(d1, "fft");
    if (axes & 4u) require_fft(d2, "fft");
    require_device();
    const hipStream_t s = call_stream();
    DeviceBuffer<float2> d(static_cast<std::size_t>(d0) * d1 * d2);
    upload(d.get(), data, d.bytes(), s);
    fft3(d.get(), d0, d1, d2, axes, forward == 0, scale, s);
    download(data, d.get(), d.bytes(), s);
    check(hipStreamSynchronize(s), "fft");
  });
}

int qh_hip_rfft3(const float* x, int nz, int ny, int nx, qh_c64* out) {
  return guarded([&] {
    require(x && out && nz >= 1 && ny >= 1 && nx >= 2 && nx % 2 == 0, "rfft3: bad arguments");
    require_fft(nz, "rfft3");
    require_fft(ny, "rfft3");
    require_fft(nx / 2, "rfft3");
    require_device();
    const hipStream_t s = call_stream();
    const std::size_t vol = static_cast<std::size_t>(nz) * ny * nx;
    const std::size_t half = static_cast<std::size_t>(nz) * ny * (nx / 2 + 1);
    DeviceBuffer<float> dx(vol);
    DeviceBuffer<float2> dh(half);
    upload(dx.get(), x, dx.bytes(), s);
    rfft3(dx.get(), dh.get(), nz, ny, nx, s);
    download(out, dh.get(), dh.bytes(), s);
    check(hipStreamSynchronize(s), "rfft3");
  });
}

int qh_hip_irfft3(const qh_c64* in, int nz, int ny, int nx, float scale, float* x) {
  return guarded([&] {
    require(x && in && nz >= 1 && ny >= 1 && nx >= 2 && nx % 2 == 0, "irfft3: bad arguments");
    require_fft(nz, "irfft3");
    require_fft(ny, "irfft3");
    require_fft(nx / 2, "irfft3");
    require_device();
    const hipStream_t s = call_stream();
    const std::size_t vol = static_cast<std::size_t>(nz) * ny * nx;
    const std::size_t half = static_cast<std::size_t>(nz) * ny * (nx / 2 + 1);
    DeviceBuffer<float> dx(vol);
    DeviceBuffer<float2> dh(half);
    upload(dh.get(), in, dh.bytes(), s);
    irfft3(dh.get(), dx.get(), nz, ny, nx, scale, s);
    download(x, dx.get(), dx.bytes(), s);
    check(hipStreamSynchronize(s), "irfft3");
  });
}

int qh_hip_find_sidebands(const float* holograms, int count, int n, float dc_exclusion,
                          float* out) {
  return guarded([&] {
    require(out != nullptr, "find_sideband: null output");
    const auto r = device_sidebands(holograms, count, n, dc_exclusion);
    for (int a = 0; a < count; ++a) std::memcpy(out + 5 * a, r[a].data(), 5 * sizeof(float));
  });
}

int qh_hip_estimate_illuminations(const float* backgrounds, int count, int n,
                                  float hologram_pixel, float dc_exclusion, qh_illumination* out,
                                  float* carrier) {
  return guarded([&] {
    require(out != nullptr && carrier != nullptr, "estimate_illuminations: null output");
    require(hologram_pixel > 0, "estimate_illuminations: pixel must be positive");
    const auto r = device_sidebands(backgrounds, count, n, dc_exclusion);
    // examples/readme_figure.py:39-44: df = 1 / (N pixel), f_a = (s_a - c) df, in double.
    const double df = 1.0 / (static_cast<double>(n) * static_cast<double>(hologram_pixel));
    for (int a = 0; a < count; ++a) {
      out[a].fy = static_cast<float>((static_cast<double>(r[a][2]) - r[0][2]) * df);
      out[a].fx = static_cast<float>((static_cast<double>(r[a][3]) - r[0][3]) * df);
    }
    std::memcpy(carrier, r[0].data(), 5 * sizeof(float));
  });
}

int qh_bpm_tables(int n, int nz, const qh_illumination* illum, int count,
                  const qh_geometry* geometry, int obliquity, qh_c64* step, qh_c64* last,
                  float* q, float* phase, float* lipschitz) {
  return guarded([&] {
    require(step && last && q && phase && lipschitz, "bpm_tables: null output");
    const qh_bpm_options o{obliquity, 0, 0};
    const BpmTables t = checked_bpm_tables(nz, n, count, illum, geometry, &o, "bpm_tables");
    std::memcpy(step, t.step.data(), t.step.size() * sizeof(float2));
    std::memcpy(last, t.last.data(), t.last.size() * sizeof(float2));
    std::memcpy(q, t.q.data(), t.q.size() * sizeof(float));
    *phase = t.phase;
    *lipschitz = t.lipschitz;
  });
}

int qh_hip_bpm_forward(const float* x, int nz, int n, const qh_illumination* illum, int count,
                       const qh_geometry* geometry, const qh_bpm_options* options, qh_c64* out) {
  return guarded([&] {
    require(x && out, "bpm_forward: null buffer");
    const BpmTables t = checked_bpm_tables(nz, n, count, illum, geometry, options, "bpm_forward");
    bpm_forward_call(t, x, options, out);
  });
}

int qh_hip_bpm_vjp(const float* x, const qh_c64* r, int nz, int n, const qh_illumination* illum,
                   int count, const qh_geometry* geometry, const qh_bpm_options* options,
                   float* grad) {
  return guarded([&] {
    require(x && r && grad, "bpm_vjp: null buffer");
    const BpmTables t = checked_bpm_tables(nz, n, count, illum, geometry, options, "bpm_vjp");
    bpm_vjp_call(t, x, r, options, grad);
  });
}

int qh_hip_bpm_reconstruct(const qh_c64* fields, int count, int n, const qh_illumination* illum,
                           const qh_geometry* geometry, float* x, int nz,
                           const qh_bpm_options* options, const qh_bpm_solve_options* solve,
                           double* loss) {
  return guarded([&] {
    require(fields && x && solve && (loss || solve->iterations == 0), "bpm_reconstruct: null buffer");
    check_solve_options(solve, n, "bpm_reconstruct");
    const BpmTables t = checked_bpm_tables(nz, n, count, illum, geometry, options, "bpm_reconstruct");
    bpm_reconstruct_call(t, fields, x, options, solve, loss);
  });
}

int qh_hip_bpm_jvp(const float* x, const float* h, int nz, int n, const qh_illumination* illum,
                   int count, const qh_geometry* geometry, const qh_bpm_options* options,
                   qh_c64* out) {
  return guarded([&] {
    require(x && h && out, "bpm_jvp: null buffer");
    const BpmTables t = checked_bpm_tables(nz, n, count, illum, geometry, options, "bpm_jvp");
    bpm_jvp_call(t, x, h, options, out);
  });
}

int qh_hip_bpm_forward_tables(const float* x, const qh_bpm_table_view* tables,
                              const qh_bpm_options* options, qh_c64* out) {
  return guarded([&] {
    require(x && out, "bpm_forward: null buffer");
    bpm_forward_call(tables_from_view(tables, options, "bpm_forward"), x, options, out);
  });
}

int qh_hip_bpm_vjp_tables(const float* x, const qh_c64* r, const qh_bpm_table_view* tables,
                          const qh_bpm_options* options, float* grad) {
  return guarded([&] {
    require(x && r && grad, "bpm_vjp: null buffer");
    bpm_vjp_call(tables_from_view(tables, options, "bpm_vjp"), x, r, options, grad);
  });
}

int qh_hip_bpm_jvp_tables(const float* x, const float* h, const qh_bpm_table_view* tables,
                          const qh_bpm_options* options, qh_c64* out) {
  return guarded([&] {
    require(x && h && out, "bpm_jvp: null buffer");
    bpm_jvp_call(tables_from_view(tables, options, "bpm_jvp"), x, h, options, out);
  });
}

int qh_hip_bpm_reconstruct_tables(const qh_c64* fields, float* x,
                                  const qh_bpm_table_view* tables, const qh_bpm_options* options,
                                  const qh_bpm_solve_options* solve, double* loss) {
  return guarded([&] {
    require(fields && x && solve && (loss || solve->iterations == 0), "bpm_reconstruct: null buffer");
    const BpmTables t = tables_from_view(tables, options, "bpm_reconstruct");
    check_solve_options(solve, t.n, "bpm_reconstruct");
    require(t.lipschitz > 0, "bpm_reconstruct: the tables' Lipschitz constant must be positive");
    bpm_reconstruct_call(t, fields, x, options, solve, loss);
  });
}

int qh_hip_bpm_plan_create(int nz, int n, const qh_illumination* illum, int count,
                           const qh_geometry* geometry, const qh_bpm_options* options,
                           qh_bpm_plan** out) {
  return guarded([&] {
    require(out != nullptr, "bpm_plan: null output");
    *out = nullptr;
    const BpmTables t = checked_bpm_tables(nz, n, count, illum, geometry, options, "bpm_plan");
    require_device();
    auto p = std::make_unique<qh_bpm_plan>();
    p->impl = std::make_unique<BpmModel>(t, options->chunk, options->checkpoint);
    *out = p.release();
  });
}

void qh_hip_bpm_plan_destroy(qh_bpm_plan* plan) { delete plan; }

int qh_hip_bpm_plan_info(const qh_bpm_plan* plan, float* phase, float* lipschitz, int* chunk,
                         size_t* device_bytes) {
  return guarded([&] {
    require(plan != nullptr, "bpm_plan: null plan");
    if (phase) *phase = plan->impl->phase();
    if (lipschitz) *lipschitz = plan->impl->lipschitz();
    if (chunk) *chunk = plan->impl->chunk();
    if (device_bytes) *device_bytes = plan->impl->bytes();
  });
}

int qh_hip_bpm_plan_checkpoint(const qh_bpm_plan* plan, int* checkpoint) {
  return guarded([&] {
    require(plan != nullptr && checkpoint != nullptr, "bpm_plan: null plan or output");
    *checkpoint = plan->impl->checkpoint();
  });
}

namespace {
hipStream_t plan_stream(void* stream) { return static_cast<hipStream_t>(stream); }
}  // namespace

int qh_hip_bpm_forward_dev(qh_bpm_plan* plan, const float* x, qh_c64* out, void* stream) {
  return guarded([&] {
    require(plan && x && out, "bpm_forward_dev: null argument");
    plan->impl->forward(x, reinterpret_cast<float2*>(out), plan_stream(stream));
  });
}

int qh_hip_bpm_vjp_dev(qh_bpm_plan* plan, const float* x, const qh_c64* r, float* grad,
                       void* stream) {
  return guarded([&] {
    require(plan && x && r && grad, "bpm_vjp_dev: null argument");
    plan->impl->vjp(x, reinterpret_cast<const float2*>(r), grad, plan_stream(stream));
  });
}

int qh_hip_bpm_jvp_dev(qh_bpm_plan* plan, const float* x, const float* h, qh_c64* out,
                       void* stream) {
  return guarded([&] {
    require(plan && x && h && out, "bpm_jvp_dev: null argument");
    plan->impl->jvp(x, h, reinterpret_cast<float2*>(out), plan_stream(stream));
  });
}

int qh_hip_bpm_gradient_dev(qh_bpm_plan* plan, const float* x, const qh_c64* fields, int first,
                            int last, float* grad, double* loss, void* stream) {
  return guarded([&] {
    require(plan && x && fields && grad && loss, "bpm_gradient_dev: null argument");
    BpmModel& m = *plan->impl;
    require(first >= 0 && last <= m.count() && first < last,
            "bpm_gradient_dev: illumination range out of bounds");
    const hipStream_t s = plan_stream(stream);
    DeviceBuffer<double> lines(static_cast<std::size_t>(m.count()) * m.n());
    m.gradient(x, reinterpret_cast<const float2*>(fields), first, last, grad, lines.get(), s);
    bpm_scale(grad, static_cast<long>(m.nz()) * m.n() * m.n(),
              1.0f / static_cast<float>(last - first), s);
    std::vector<double> h(lines.size());
    download(h.data(), lines.get(), lines.bytes(), s);
    check(hipStreamSynchronize(s), "bpm_gradient_dev");
    double sum = 0;
    for (int a = first; a < last; ++a) {
      double la = 0;
      for (int y = 0; y < m.n(); ++y) la += h[static_cast<std::size_t>(a) * m.n() + y];
      sum += la;
    }
    *loss = 0.5 * sum / m.count();
  });
}

int qh_hip_tv_prox(const float* b, int nz, int ny, int nx, float lambda, int inner, int fused,
                   float* x) {
  return guarded([&] {
    require(b && x && nz >= 1 && ny >= 1 && nx >= 1 && lambda > 0 && inner >= 0,
            "tv_prox: bad arguments");
    require_device();
    const hipStream_t s = call_stream();
    const std::size_t vol = static_cast<std::size_t>(nz) * ny * nx;
    DeviceBuffer<float> db(vol);
    DeviceBuffer<float> dx(vol);
    TvState state(nz, ny, nx);
    state.set_fused(fused != 0);
    state.reset(s);
    upload(db.get(), b, db.bytes(), s);
    state.prox(db.get(), dx.get(), lambda, inner, s);
    download(x, dx.get(), dx.bytes(), s);
    check(hipStreamSynchronize(s), "tv_prox");
  });
}

int qh_hip_tv_prox_batch(const float* b, int batch, int nz, int ny, int nx, float lambda,
                         int inner, int fused, float* x) {
  return guarded([&] {
    require(b && x && batch >= 1 && batch <= 65535 && nz >= 1 && ny >= 1 && nx >= 1 &&
                lambda > 0 && inner >= 0,
            "tv_prox_batch: bad arguments (batch 1 to 65535)");
    require_device();
    const hipStream_t s = call_stream();
    const std::size_t vol = static_cast<std::size_t>(nz) * ny * nx;
    DeviceBuffer<float> db(vol * batch);
    DeviceBuffer<float> dx(vol * batch);
    TvState state(nz, ny, nx, batch);
    state.set_fused(fused != 0);
    state.reset(s, batch);
    upload(db.get(), b, db.bytes(), s);
    state.prox(db.get(), dx.get(), lambda, inner, s);
    download(x, dx.get(), dx.bytes(), s);
    check(hipStreamSynchronize(s), "tv_prox_batch");
  });
}

}  // extern "C"
