// Total-variation proximal step by fast gradient projection on the dual (FGP), the inner loop
// of FISTA-TV. One kernel per dual step: a 64 x 8 tile of (x, y) columns marches along z, the
// dual r of two planes (with a one-voxel halo) and the primal point z of two planes live in
// LDS, so every voxel's b, p and r are read from HBM once per step and p, r written once.
#include "common.hpp"
#include "runtime.hpp"
#include "tv.hpp"

namespace qh {

namespace {

constexpr int kTx = 64;
constexpr int kTy = 8;
constexpr int kZChunk = 32;
constexpr int kThreadsTv = kTx * kTy;

struct TvArgs {
  const float* b;     // primal data, dense (nz, ny, nx)
  const float* r_in;  // dual extrapolation r, three components `vol` floats apart
  const float* p_in;  // dual p at the start of the step
  float* p_out;       // dual p after the step (p_in or another buffer)
  float* r_out;       // r after the step (never r_in)
  long vol;
  int nz, ny, nx;
  float lambda;
  float step;
  float momentum;
};

__device__ __forceinline__ long at3(const TvArgs& a, int z, int y, int x) {
  return (static_cast<long>(z) * a.ny + y) * a.nx + x;
}

__global__ void __launch_bounds__(kThreadsTv) k_fgp_step(TvArgs a) {
  constexpr int hx = kTx + 2;
  constexpr int hy = kTy + 2;
  // r on two planes, local y, x in [-1, kT]; z on two planes, local y, x in [0, kT].
  __shared__ float sr[2][3][hy][hx];
  __shared__ float sz[2][kTy + 1][kTx + 1];
  const int x0 = static_cast<int>(blockIdx.x) * kTx;
  const int y0 = static_cast<int>(blockIdx.y) * kTy;
  const int z0 = static_cast<int>(blockIdx.z) * kZChunk;
  const int z1 = min(z0 + kZChunk, a.nz);
  const int tid = static_cast<int>(threadIdx.x);
  const int tx = tid % kTx;
  const int ty = tid / kTx;

  const auto load = [&](int slot, int z) {
    for (int i = tid; i < hx * hy; i += kThreadsTv) {
      const int lx = i % hx;
      const int ly = i / hx;
      const int x = x0 - 1 + lx;
      const int y = y0 - 1 + ly;
      const bool in = x >= 0 && x < a.nx && y >= 0 && y < a.ny && z >= 0 && z < a.nz;
      const long g = in ? at3(a, z, y, x) : 0;
#pragma unroll
      for (int c = 0; c < 3; ++c) sr[slot][c][ly][lx] = in ? a.r_in[c * a.vol + g] : 0.0f;
    }
  };
  // z on plane z from r on that plane (slot) and on the plane below (below); the additions run
  // in the order of the negative divergence L(r) = sum_a r_a(i) - r_a(i - e_a), axes z, y, x.
  const auto primal = [&](int out, int slot, int below, int z) {
    for (int i = tid; i < (kTx + 1) * (kTy + 1); i += kThreadsTv) {
      const int lx = i % (kTx + 1);
      const int ly = i / (kTx + 1);
      const int x = x0 + lx;
      const int y = y0 + ly;
      float zv = 0.0f;
      if (x < a.nx && y < a.ny && z < a.nz) {
        const int cx = lx + 1;
        const int cy = ly + 1;
        float s = 0.0f;
        if (z < a.nz - 1) s += sr[slot][0][cy][cx];
        if (z > 0) s -= sr[below][0][cy][cx];
        if (y < a.ny - 1) s += sr[slot][1][cy][cx];
        if (y > 0) s -= sr[slot][1][cy - 1][cx];
        if (x < a.nx - 1) s += sr[slot][2][cy][cx];
        if (x > 0) s -= sr[slot][2][cy][cx - 1];
        zv = tv_primal(a.b[at3(a, z, y, x)], a.lambda, s);
      }
      sz[out][ly][lx] = zv;
    }
  };

  int cur = 0;
  int zc = 0;
  load(1, z0 - 1);
  load(cur, z0);
  __syncthreads();
  primal(zc, cur, 1, z0);
  const int x = x0 + tx;
  const int y = y0 + ty;
  const bool inside = x < a.nx && y < a.ny;
  for (int z = z0; z < z1; ++z) {
    const int next = 1 - cur;
    const long g = inside ? at3(a, z, y, x) : 0;
    float pv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) pv[c] = inside ? a.p_in[c * a.vol + g] : 0.0f;
    __syncthreads();  // plane z - 1 in `next` and z-slot 1 - zc are no longer read
    load(next, z + 1);
    __syncthreads();
    primal(1 - zc, next, cur, z + 1);
    __syncthreads();
    if (inside) {
      const float c = sz[zc][ty][tx];
      float gr[3];
      gr[0] = z < a.nz - 1 ? c - sz[1 - zc][ty][tx] : 0.0f;
      gr[1] = y < a.ny - 1 ? c - sz[zc][ty + 1][tx] : 0.0f;
      gr[2] = x < a.nx - 1 ? c - sz[zc][ty][tx + 1] : 0.0f;
      float rv[3];
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) rv[cc] = sr[cur][cc][ty + 1][tx + 1];
      tv_dual_step(gr, a.step, a.momentum, pv, rv);
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        a.p_out[cc * a.vol + g] = pv[cc];
        a.r_out[cc * a.vol + g] = rv[cc];
      }
    }
    cur = next;
    zc = 1 - zc;
  }
}

// out = max(b - lambda L(p), 0), optionally followed by the refractive index of out.
__global__ void k_tv_primal(const float* b, const float* p, long vol, int nz, int ny, int nx,
                            float lambda, float* out, int to_index, float nm2, float inv_k02) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= vol) return;
  const int x = static_cast<int>(i % nx);
  const int y = static_cast<int>((i / nx) % ny);
  const int z = static_cast<int>(i / (static_cast<long>(nx) * ny));
  const long plane = static_cast<long>(nx) * ny;
  const float* pz = p;
  const float* py = p + vol;
  const float* px = p + 2 * vol;
  float s = 0.0f;
  if (z < nz - 1) s += pz[i];
  if (z > 0) s -= pz[i - plane];
  if (y < ny - 1) s += py[i];
  if (y > 0) s -= py[i - nx];
  if (x < nx - 1) s += px[i];
  if (x > 0) s -= px[i - 1];
  const float v = tv_primal(b[i], lambda, s);
  out[i] = to_index ? index_from_potential(v, 0.0f, nm2, inv_k02) : v;
}

__global__ void k_fill(float* d, long n, float v) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) d[i] = v;
}

}  // namespace

void fill(float* d, long n, float v, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(n)), dim3(256), 0, s, d, n, v);
  check_launch("fill");
}

void fgp_step(const float* b, const float* r_in, const float* p_in, float* p_out, float* r_out,
              int nz, int ny, int nx, float lambda, float step, float momentum, hipStream_t s) {
  TvArgs a{b, r_in, p_in, p_out, r_out, static_cast<long>(nz) * ny * nx, nz, ny, nx,
           lambda, step, momentum};
  const dim3 grid((nx + kTx - 1) / kTx, (ny + kTy - 1) / kTy, (nz + kZChunk - 1) / kZChunk);
  hipLaunchKernelGGL(k_fgp_step, grid, dim3(kThreadsTv), 0, s, a);
  check_launch("fgp_step");
}

void tv_primal(const float* b, const float* p, int nz, int ny, int nx, float lambda, float* out,
               bool to_index, float nm2, float inv_k02, hipStream_t s) {
  const long vol = static_cast<long>(nz) * ny * nx;
  hipLaunchKernelGGL(k_tv_primal, dim3(blocks_for(vol)), dim3(256), 0, s, b, p, vol, nz, ny, nx,
                     lambda, out, to_index ? 1 : 0, nm2, inv_k02);
  check_launch("tv_primal");
}

TvState::TvState(int nz, int ny, int nx)
    : nz_(nz), ny_(ny), nx_(nx), vol_(static_cast<long>(nz) * ny * nx),
      p_{DeviceBuffer<float>(3 * vol_), DeviceBuffer<float>(3 * vol_)},
      r_{DeviceBuffer<float>(3 * vol_), DeviceBuffer<float>(3 * vol_)} {}

void TvState::reset(hipStream_t s) {
  fill(p_[0].get(), 3 * vol_, 0.0f, s);
  cur_ = 0;
}

std::size_t TvState::bytes() const {
  return p_[0].bytes() + p_[1].bytes() + r_[0].bytes() + r_[1].bytes();
}

// tv_prox of the reference (src/cpu/tv.cpp:7-44): r <- p, `inner` FGP steps, out = primal(p).
// The copy r <- p is folded into the first step, which reads p as r and writes the new p to
// the second p buffer; later steps update p in place and ping-pong r.
void TvState::prox(const float* b, float* out, float lambda, int inner, hipStream_t s,
                   bool to_index, float nm2, float inv_k02, Event* t0, Event* t1) {
  const float step = 1.0f / (12.0f * lambda);
  float t = 1.0f;
  float* p_old = p_[cur_].get();
  float* p_new = p_[1 - cur_].get();
  const float* r_in = p_old;
  int rs = 0;
  if (t0) t0->record(s);
  for (int k = 0; k < inner; ++k) {
    const float tn = fista_t(t);
    const float m = (t - 1.0f) / tn;
    float* r_out = r_[rs].get();
    fgp_step(b, r_in, k == 0 ? p_old : p_new, p_new, r_out, nz_, ny_, nx_, lambda, step, m, s);
    r_in = r_out;
    rs = 1 - rs;
    t = tn;
  }
  if (t1) t1->record(s);
  if (inner > 0) cur_ = 1 - cur_;
  tv_primal(b, p_[cur_].get(), nz_, ny_, nx_, lambda, out, to_index, nm2, inv_k02, s);
}

}  // namespace qh
