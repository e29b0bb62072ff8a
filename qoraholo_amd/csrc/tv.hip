// Total-variation proximal step by fast gradient projection on the dual (FGP), the inner loop
// of FISTA-TV, on gfx950.
//
// One wavefront per block owns a strip of TY rows of the volume, each lane VX consecutive x
// positions of every row (a whole 256-wide row is 64 lanes x 4 floats, loaded as float4), and
// marches along z through a chunk of planes. y neighbours are in the lane's own registers, x
// neighbours one lane away (shuffles), z neighbours in the registers of the previous plane; only
// one halo row above and below the strip (and one plane at the chunk start) is read twice.
//
// The extrapolated dual r_k = p_k + m_{k-1} (p_k - p_{k-1}) is not stored: each step reads the
// two latest dual iterates and forms r with the same float operations the reference's dual step
// uses to produce it (src/common.hpp:128-141), so a step reads b, p_k, p_{k-1} and writes p_{k+1}
// (10 volumes instead of 13 with r stored). The first step of a prox uses r = p_k
// (src/cpu/tv.cpp:10).
#include <algorithm>
#include <atomic>

#include "common.hpp"
#include "runtime.hpp"
#include "tv.hpp"

namespace qh {

namespace {

constexpr int kZChunk = 16;

struct TvArgs {
  const float* b;       // primal data, dense (nz, ny, nx)
  const float* p_in;    // dual p_k, three components `vol` floats apart
  const float* p_prev;  // dual p_{k-1} (unused on the first step)
  float* p_out;         // dual p_{k+1}, a third buffer
  long vol;
  int nz, ny, nx;
  float lambda;
  float step;
  float momentum;  // m_{k-1}, the momentum of the step that produced p_k
  int first;       // r = p_k exactly
  int zlen = kZChunk;  // planes per chunk of the strip kernels (strip_grid)
};

// Volume blockIdx.y of a batch of volumes that lie one after another: its primal data and
// output start blockIdx.y vol floats in, its duals (three components each) 3 blockIdx.y vol in.
// Kernels apply the offsets only in their BATCH instantiations, which single volumes never run:
// the offsets cost the 256^3 two-step pass 1.3% (154.7 -> 156.8 us) when always applied.
__device__ __forceinline__ long batch_volume() { return static_cast<long>(blockIdx.y); }
template <class T>
__device__ __forceinline__ T* shifted(T* p, long by) {
  return p == nullptr ? p : p + by;
}

// Block b runs tile t: with a grid that is a multiple of 8, blocks b, b+8, b+16, ... (which the
// dispatcher places on the same XCD) get consecutive tiles, so the halo rows and the planes
// shared by neighbouring tiles are read through the same L2.
__device__ __forceinline__ int xcd_tile() {
  const int nb = static_cast<int>(gridDim.x);
  const int b = static_cast<int>(blockIdx.x);
  return (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
}

// Row access. VEC: nx is a multiple of VX and VX of 2, so a lane's VX values are all inside the
// row or all outside it and every access is whole float4s (float2s when VX is not a multiple of
// 4) under one predicate per lane; otherwise element by element.
// XNEG: xb may be negative (the left halo of an x tile); other callers start rows at column 0.
template <int VX, bool VEC, bool XNEG = false, bool NT = false>
__device__ __forceinline__ void load_row(const float* base, const TvArgs& a, int z, int y, int xb,
                                         float (&out)[VX]) {
  const bool inside = z >= 0 && z < a.nz && y >= 0 && y < a.ny;
  if constexpr (VEC) {
    static_assert(VX % 2 == 0, "whole float2s");
    if (inside && (!XNEG || xb >= 0) && xb < a.nx) {
      const float* p = base + (static_cast<long>(z) * a.ny + y) * a.nx + xb;
      if constexpr (VX % 4 == 0) {
#pragma unroll
        for (int v = 0; v < VX; v += 4) {
          typedef float vec4 __attribute__((ext_vector_type(4)));
          const vec4 q = NT ? __builtin_nontemporal_load(reinterpret_cast<const vec4*>(p + v))
                            : *reinterpret_cast<const vec4*>(p + v);
          out[v] = q.x;
          out[v + 1] = q.y;
          out[v + 2] = q.z;
          out[v + 3] = q.w;
        }
      } else {
#pragma unroll
        for (int v = 0; v < VX; v += 2) {
          typedef float vec2 __attribute__((ext_vector_type(2)));
          const vec2 q = NT ? __builtin_nontemporal_load(reinterpret_cast<const vec2*>(p + v))
                            : *reinterpret_cast<const vec2*>(p + v);
          out[v] = q.x;
          out[v + 1] = q.y;
        }
      }
    } else {
#pragma unroll
      for (int v = 0; v < VX; ++v) out[v] = 0.0f;
    }
  } else {
#pragma unroll
    for (int v = 0; v < VX; ++v) out[v] = 0.0f;
    if (!inside) return;
    const long row = (static_cast<long>(z) * a.ny + y) * a.nx;
#pragma unroll
    for (int v = 0; v < VX; ++v) {
      if ((!XNEG || xb + v >= 0) && xb + v < a.nx) out[v] = base[row + xb + v];
    }
  }
}

// NT: non-temporal stores, for the dual iterates. A pass writes them (400 MB at 256^3) long
// before the next pass reads them, so kept in the caches they only displace the halo rows and
// columns that neighbouring tiles of the same pass read back (measured: 169 -> 165 us per pass
// at 256^3, 657 -> 620 us at 384^3 together with 64-plane z chunks).
template <int VX, bool VEC, bool NT = false>
__device__ __forceinline__ void store_row(float* base, const TvArgs& a, int z, int y, int xb,
                                          const float (&in)[VX]) {
  if (z >= a.nz || y >= a.ny) return;
  float* p = base + (static_cast<long>(z) * a.ny + y) * a.nx + xb;
  if constexpr (VEC) {
    static_assert(VX % 2 == 0, "whole float2s");
    typedef float vec4 __attribute__((ext_vector_type(4)));
    typedef float vec2 __attribute__((ext_vector_type(2)));
    if (xb < a.nx) {
      if constexpr (VX % 4 == 0) {
#pragma unroll
        for (int v = 0; v < VX; v += 4) {
          const vec4 q = {in[v], in[v + 1], in[v + 2], in[v + 3]};
          if constexpr (NT) {
            __builtin_nontemporal_store(q, reinterpret_cast<vec4*>(p + v));
          } else {
            *reinterpret_cast<vec4*>(p + v) = q;
          }
        }
      } else {
#pragma unroll
        for (int v = 0; v < VX; v += 2) {
          const vec2 q = {in[v], in[v + 1]};
          if constexpr (NT) {
            __builtin_nontemporal_store(q, reinterpret_cast<vec2*>(p + v));
          } else {
            *reinterpret_cast<vec2*>(p + v) = q;
          }
        }
      }
    }
  } else {
#pragma unroll
    for (int v = 0; v < VX; ++v) {
      if (xb + v < a.nx) p[v] = in[v];
    }
  }
}

// Primal point z = max(b - lambda L(r), 0) of plane zp for the strip's rows y0..y0+TY (the
// last one is the halo row below the strip). R holds r at plane zp, ry_top the y component of
// r on row y0 - 1, rz_below the z component on plane zp - 1. The additions run in the order of
// the negative divergence L(r)(i) = sum_a r_a(i) - r_a(i - e_a), axes z, y, x.
template <int VX, int TY>
__device__ __forceinline__ void primal_plane(const TvArgs& a, int zp, int y0, int xb,
                                             const float (&R)[3][TY + 1][VX],
                                             const float (&ry_top)[VX],
                                             const float (&rz_below)[TY + 1][VX],
                                             const float (&B)[TY + 1][VX],
                                             float (&Z)[TY + 1][VX]) {
#pragma unroll
  for (int j = 0; j <= TY; ++j) {
    const int y = y0 + j;
    const float left = __shfl_up(R[2][j][VX - 1], 1);
#pragma unroll
    for (int v = 0; v < VX; ++v) {
      const int x = xb + v;
      float s = 0.0f;
      if (zp < a.nz - 1) s += R[0][j][v];
      if (zp > 0) s -= rz_below[j][v];
      if (y < a.ny - 1) s += R[1][j][v];
      if (y > 0) s -= (j == 0 ? ry_top[v] : R[1][j - 1][v]);
      if (x < a.nx - 1) s += R[2][j][v];
      if (x > 0) s -= (v == 0 ? left : R[2][j][v - 1]);
      Z[j][v] = (y < a.ny && x < a.nx && zp < a.nz) ? tv_primal(B[j][v], a.lambda, s) : 0.0f;
    }
  }
}

// r = p_k + m (p_k - p_{k-1}) on one row (or p_k on the first step of a prox).
template <int VX>
__device__ __forceinline__ void extrapolate(const TvArgs& a, const float (&pk)[VX],
                                            const float (&pm)[VX], float (&r)[VX]) {
#pragma unroll
  for (int v = 0; v < VX; ++v) r[v] = a.first ? pk[v] : pk[v] + a.momentum * (pk[v] - pm[v]);
}

// r of plane z on the strip rows and the halo row below (R), its y component on the row above
// (ry_top), and b on the strip rows and the halo row. P keeps p_k of the strip rows for the
// dual step of this plane.
template <int VX, int TY, bool VEC>
__device__ __forceinline__ void load_plane(const TvArgs& a, int z, int y0, int xb,
                                           float (&R)[3][TY + 1][VX], float (&ry_top)[VX],
                                           float (&B)[TY + 1][VX], float (&P)[3][TY][VX]) {
#pragma unroll
  for (int j = 0; j <= TY; ++j) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float pk[VX], pm[VX];
      load_row<VX, VEC>(a.p_in + c * a.vol, a, z, y0 + j, xb, pk);
      if (!a.first) {
        load_row<VX, VEC>(a.p_prev + c * a.vol, a, z, y0 + j, xb, pm);
      } else {
#pragma unroll
        for (int v = 0; v < VX; ++v) pm[v] = 0.0f;
      }
      extrapolate<VX>(a, pk, pm, R[c][j]);
      if (j < TY) {
#pragma unroll
        for (int v = 0; v < VX; ++v) P[c][j][v] = pk[v];
      }
    }
    load_row<VX, VEC>(a.b, a, z, y0 + j, xb, B[j]);
  }
  float pk[VX], pm[VX];
  load_row<VX, VEC>(a.p_in + a.vol, a, z, y0 - 1, xb, pk);
  if (!a.first) {
    load_row<VX, VEC>(a.p_prev + a.vol, a, z, y0 - 1, xb, pm);
  } else {
#pragma unroll
    for (int v = 0; v < VX; ++v) pm[v] = 0.0f;
  }
  extrapolate<VX>(a, pk, pm, ry_top);
}

template <int VX, int TY, bool VEC, bool BATCH>
__global__ void __launch_bounds__(64) k_fgp_step(TvArgs a, int chunks) {
  if constexpr (BATCH) {
    a.b = shifted(a.b, batch_volume() * a.vol);
    a.p_in = shifted(a.p_in, 3 * batch_volume() * a.vol);
    a.p_prev = shifted(a.p_prev, 3 * batch_volume() * a.vol);
    a.p_out = shifted(a.p_out, 3 * batch_volume() * a.vol);
  }
  const int t = xcd_tile();
  const int y0 = (t / chunks) * TY;
  const int z0 = (t % chunks) * a.zlen;
  const int z1 = min(z0 + a.zlen, a.nz);
  const int xb = static_cast<int>(threadIdx.x) * VX;

  float r[3][TY + 1][VX];  // r at plane z, rows y0..y0+TY
  float pc[3][TY][VX];     // p_k at plane z, strip rows
  float zc[TY + 1][VX];    // primal at plane z
  {
    // Warm-up: the primal on plane z0 needs the z component of r on plane z0 - 1.
    float rz_below[TY + 1][VX];
#pragma unroll
    for (int j = 0; j <= TY; ++j) {
      float pk[VX], pm[VX];
      load_row<VX, VEC>(a.p_in, a, z0 - 1, y0 + j, xb, pk);
      if (!a.first) {
        load_row<VX, VEC>(a.p_prev, a, z0 - 1, y0 + j, xb, pm);
      } else {
#pragma unroll
        for (int v = 0; v < VX; ++v) pm[v] = 0.0f;
      }
      extrapolate<VX>(a, pk, pm, rz_below[j]);
    }
    float ry_top[VX], bb[TY + 1][VX];
    load_plane<VX, TY, VEC>(a, z0, y0, xb, r, ry_top, bb, pc);
    primal_plane<VX, TY>(a, z0, y0, xb, r, ry_top, rz_below, bb, zc);
  }
  for (int z = z0; z < z1; ++z) {
    float rn[3][TY + 1][VX], ryn[VX], bn[TY + 1][VX], pn[3][TY][VX], zn[TY + 1][VX];
    load_plane<VX, TY, VEC>(a, z + 1, y0, xb, rn, ryn, bn, pn);
    primal_plane<VX, TY>(a, z + 1, y0, xb, rn, ryn, r[0], bn, zn);
#pragma unroll
    for (int j = 0; j < TY; ++j) {
      const int y = y0 + j;
      const float right = __shfl_down(zc[j][0], 1);
      float po[3][VX];
#pragma unroll
      for (int v = 0; v < VX; ++v) {
        const int x = xb + v;
        const float c = zc[j][v];
        float g[3];
        g[0] = z < a.nz - 1 ? c - zn[j][v] : 0.0f;
        g[1] = y < a.ny - 1 ? c - zc[j + 1][v] : 0.0f;
        g[2] = x < a.nx - 1 ? c - (v == VX - 1 ? right : zc[j][v + 1]) : 0.0f;
        float p3[3] = {pc[0][j][v], pc[1][j][v], pc[2][j][v]};
        float r3[3] = {r[0][j][v], r[1][j][v], r[2][j][v]};
        tv_dual_step(g, a.step, 0.0f, p3, r3);
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) po[cc][v] = p3[cc];
      }
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) store_row<VX, VEC, true>(a.p_out + cc * a.vol, a, z, y, xb, po[cc]);
    }
#pragma unroll
    for (int j = 0; j <= TY; ++j) {
#pragma unroll
      for (int v = 0; v < VX; ++v) {
        zc[j][v] = zn[j][v];
#pragma unroll
        for (int c = 0; c < 3; ++c) r[c][j][v] = rn[c][j][v];
      }
    }
#pragma unroll
    for (int j = 0; j < TY; ++j) {
#pragma unroll
      for (int v = 0; v < VX; ++v) {
#pragma unroll
        for (int c = 0; c < 3; ++c) pc[c][j][v] = pn[c][j][v];
      }
    }
  }
}

// ---- two FGP steps per pass ----
//
// A block of NW waves owns NW - 4 output rows; wave w holds row y0 - 2 + w across its 64 lanes
// (VX floats each) and the block marches z through a chunk of planes. Per plane t the waves
// load b, p_k, p_{k-1} of plane t + 2, form r_k and z_k there, the dual p_{k+1} of plane t + 1
// and its r_{k+1}, z_{k+1} of plane t + 1, and p_{k+2} of plane t. y neighbours come through
// LDS (one row buffer per quantity, read back before the owner overwrites it), x neighbours by
// shuffles, z neighbours from the registers of the previous plane. The two halo rows on each
// side make step one available where step two needs it; their results are not stored. Every
// value is formed with the float operations of k_fgp_step, so a pass equals two single steps
// bit for bit, while HBM sees 7 volumes read and 6 written per two steps instead of 20.
// Planes per block: at most 64 and at least 4, chosen per volume by pair_chunk; each chunk
// starts two planes early, so longer chunks read fewer warm-up planes (384^3 pass with x tiles
// of 2 columns per lane 675 -> 637 us from 32 to 64 planes).
constexpr int kPairChunk = 64;
constexpr int kPairChunkMin = 4;

struct PairArgs {
  const float* b;
  const float* p_k;    // p_k (r_k is formed from it)
  const float* p_km1;  // p_{k-1} (unused on the first step of a prox)
  float* p_k1;         // out: p_{k+1}
  float* p_k2;         // out: p_{k+2}
  long vol;
  int nz, ny, nx;
  float lambda;
  float step;
  float m1;  // momentum of r_k = p_k + m1 (p_k - p_{k-1})
  int first;
  float m2;  // momentum of r_{k+1} = p_{k+1} + m2 (p_{k+1} - p_k)
  int zlen;  // planes per chunk
  int xt;       // output columns per x tile (nx when one tile spans the row)
  int tiles_x;  // x tiles per row
};

// Whether column xb + v of a lane can lie on the volume's x faces. Rows that start at column 0
// (not XNEG) with whole-vector access (VEC: nx a multiple of VX, xb a multiple of VX) put the
// faces only in a lane's first (x = 0) and last (x = nx - 1) element, so the middle elements of
// an active lane (xb < nx) never need the x tests.
template <int VX, bool XNEG, bool VEC>
__device__ __forceinline__ constexpr bool x_left_face(int v) {
  return XNEG || !VEC || v == 0;
}
template <int VX, bool XNEG, bool VEC>
__device__ __forceinline__ constexpr bool x_right_face(int v) {
  return XNEG || !VEC || v == VX - 1;
}

// z = max(b - lambda L(r), 0) on one row of plane zp (the expression order of primal_plane).
// zp and y are the same for every lane of the wave: on planes and rows off the volume's z and y
// faces every z and y term is taken, so that branch runs the same float operations without the
// per-element selects (the kernel's time is set by its VALU issue rate, not only by its bytes).
template <int VX, bool XNEG, bool VEC>
__device__ __forceinline__ void pair_primal(const PairArgs& a, int zp, int y, int xb,
                                            const float (&R)[3][VX], const float (&rz_below)[VX],
                                            const float (&ry_up)[VX], const float (&B)[VX],
                                            float (&Z)[VX]) {
  const float left = __shfl_up(R[2][VX - 1], 1);
  if (zp > 0 && zp < a.nz - 1 && y > 0 && y < a.ny - 1) {
#pragma unroll
    for (int v = 0; v < VX; ++v) {
      const int x = xb + v;
      float s = 0.0f;
      s += R[0][v];
      s -= rz_below[v];
      s += R[1][v];
      s -= ry_up[v];
      if (!x_right_face<VX, XNEG, VEC>(v) || x < a.nx - 1) s += R[2][v];
      if (!x_left_face<VX, XNEG, VEC>(v) || x > 0) s -= (v == 0 ? left : R[2][v - 1]);
      // Without XNEG and with VEC every element of an active lane is inside the row; inactive
      // lanes (xb >= nx) are never stored and feed no active lane's result.
      Z[v] = (XNEG || !VEC) ? (((!XNEG || x >= 0) && x < a.nx) ? tv_primal(B[v], a.lambda, s) : 0.0f)
                            : tv_primal(B[v], a.lambda, s);
    }
    return;
  }
#pragma unroll
  for (int v = 0; v < VX; ++v) {
    const int x = xb + v;
    float s = 0.0f;
    if (zp < a.nz - 1) s += R[0][v];
    if (zp > 0) s -= rz_below[v];
    if (y < a.ny - 1) s += R[1][v];
    if (y > 0) s -= ry_up[v];
    if (x < a.nx - 1) s += R[2][v];
    if (x > 0) s -= (v == 0 ? left : R[2][v - 1]);
    Z[v] = (y >= 0 && y < a.ny && (!XNEG || x >= 0) && x < a.nx && zp >= 0 && zp < a.nz)
               ? tv_primal(B[v], a.lambda, s)
               : 0.0f;
  }
}

// p' = P(r + step g) on one row of plane zp, g from z of this plane (zc), the plane above (zn)
// and the row below (zdn): the dual step of k_fgp_step. Off the last plane and the last row
// (wave-uniform) the z and y differences are taken without selects.
template <int VX, bool XNEG, bool VEC>
__device__ __forceinline__ void pair_dual(const PairArgs& a, int zp, int y, int xb,
                                          const float (&zc)[VX], const float (&zn)[VX],
                                          const float (&zdn)[VX], const float (&R)[3][VX],
                                          float (&P)[3][VX]) {
  const float right = __shfl_down(zc[0], 1);
  const bool inner = zp < a.nz - 1 && y < a.ny - 1;
#pragma unroll
  for (int v = 0; v < VX; ++v) {
    const int x = xb + v;
    const float c = zc[v];
    float g[3];
    if (inner) {
      g[0] = c - zn[v];
      g[1] = c - zdn[v];
    } else {
      g[0] = zp < a.nz - 1 ? c - zn[v] : 0.0f;
      g[1] = y < a.ny - 1 ? c - zdn[v] : 0.0f;
    }
    g[2] = (!x_right_face<VX, XNEG, VEC>(v) || x < a.nx - 1)
               ? c - (v == VX - 1 ? right : zc[v + 1])
               : 0.0f;
    float p3[3] = {0.0f, 0.0f, 0.0f};
    float r3[3] = {R[0][v], R[1][v], R[2][v]};
    tv_dual_step(g, a.step, 0.0f, p3, r3);
#pragma unroll
    for (int c3 = 0; c3 < 3; ++c3) P[c3][v] = p3[c3];
  }
}

// Rows in LDS (row r of a buffer at r * nx), with the access rules of load_row.
template <int VX, bool VEC>
__device__ __forceinline__ void lds_put(float* buf, int row, int nx, int xb, const float (&v)[VX]) {
  if constexpr (VEC && VX % 4 == 0) {
    if (xb < nx) {
#pragma unroll
      for (int i = 0; i < VX; i += 4) {
        *reinterpret_cast<float4*>(buf + row * nx + xb + i) = float4{v[i], v[i + 1], v[i + 2], v[i + 3]};
      }
    }
  } else if constexpr (VEC && VX % 2 == 0) {
    if (xb < nx) {
#pragma unroll
      for (int i = 0; i < VX; i += 2) {
        *reinterpret_cast<float2*>(buf + row * nx + xb + i) = float2{v[i], v[i + 1]};
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < VX; ++i) {
      if (xb + i < nx) buf[row * nx + xb + i] = v[i];
    }
  }
}

template <int VX, bool VEC>
__device__ __forceinline__ void lds_get(const float* buf, int row, int nx, int xb, float (&v)[VX]) {
  if constexpr (VEC && VX % 4 == 0) {
    if (xb < nx) {
#pragma unroll
      for (int i = 0; i < VX; i += 4) {
        const float4 q = *reinterpret_cast<const float4*>(buf + row * nx + xb + i);
        v[i] = q.x;
        v[i + 1] = q.y;
        v[i + 2] = q.z;
        v[i + 3] = q.w;
      }
    } else {
#pragma unroll
      for (int i = 0; i < VX; ++i) v[i] = 0.0f;
    }
  } else if constexpr (VEC && VX % 2 == 0) {
    if (xb < nx) {
#pragma unroll
      for (int i = 0; i < VX; i += 2) {
        const float2 q = *reinterpret_cast<const float2*>(buf + row * nx + xb + i);
        v[i] = q.x;
        v[i + 1] = q.y;
      }
    } else {
#pragma unroll
      for (int i = 0; i < VX; ++i) v[i] = 0.0f;
    }
  } else {
#pragma unroll
    for (int i = 0; i < VX; ++i) v[i] = (xb + i < nx) ? buf[row * nx + xb + i] : 0.0f;
  }
}

// VEC: see load_row. XT: rows are cut into x tiles (more than one per row).
template <int VX, int NW, bool VEC, bool XT, bool BATCH>
__global__ void __launch_bounds__(NW * 64) k_fgp_pair(PairArgs a, int chunks) {
  if constexpr (BATCH) {
    a.b = shifted(a.b, batch_volume() * a.vol);
    a.p_k = shifted(a.p_k, 3 * batch_volume() * a.vol);
    a.p_km1 = shifted(a.p_km1, 3 * batch_volume() * a.vol);
    a.p_k1 = shifted(a.p_k1, 3 * batch_volume() * a.vol);
    a.p_k2 = shifted(a.p_k2, 3 * batch_volume() * a.vol);
  }
  constexpr int TOP = 2;  // halo rows above the output rows (and two below)
  constexpr int T = NW - 4;
  // LDS row width: the row when a wave holds it whole, else the columns one wave holds.
  const int LW = XT ? 64 * VX : a.nx;
  extern __shared__ float4 lds4[];
  float* lds = reinterpret_cast<float*>(lds4);
  float* ry = lds;              // y component of r_k of each row
  float* rqy = ry + NW * LW;    // y component of r_{k+1}
  float* zk = rqy + NW * LW;    // z_k
  float* zq = zk + NW * LW;     // z_{k+1}
  TvArgs bd{};
  bd.nz = a.nz;
  bd.ny = a.ny;
  bd.nx = a.nx;
  // The wave index is uniform across the wave: as a scalar it keeps the row arithmetic, the
  // row addresses and the y tests off the vector ALU.
  const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
  const int lane = static_cast<int>(threadIdx.x) & 63;
  // Tiles of one z chunk are consecutive (x fastest, then y), so the neighbours that share halo
  // rows and columns run on the same XCD. Rows wider than a wave are cut into x tiles of xt
  // output columns; each tile's wave also holds one lane of columns on either side, which two
  // steps of the 7-point stencil need, and stores only its own columns.
  const int tile = xcd_tile();
  const int tx = XT ? tile % a.tiles_x : 0;
  const int rest = XT ? tile / a.tiles_x : tile;
  const int tiles_y = static_cast<int>(gridDim.x) / (chunks * (XT ? a.tiles_x : 1));
  const int y0 = (rest % tiles_y) * T;
  const int z0 = (rest / tiles_y) * a.zlen;
  const int z1 = min(z0 + a.zlen, a.nz);
  const int y = y0 - TOP + w;
  const int xo = tx * a.xt;            // first output column of the tile
  const int x0 = XT ? xo - VX : 0;     // column of lane 0
  const int xl = lane * VX;            // column within the LDS row
  const int xb = x0 + xl;
  const bool out_col = !XT || (xb >= xo && xb < min(xo + a.xt, a.nx));
  const int wu = w > 0 ? w - 1 : 0;       // the first and last waves read their own row: their
  const int wd = w < NW - 1 ? w + 1 : w;  // results only feed other halo values, never a store
  const bool out_row = w >= TOP && w < TOP + T && y < a.ny;
  const int pre = 2;  // planes of warm-up before z0
  // p_k, p_{k-1} and b of this row on plane f, as loaded.
  auto load_raw_as = [&]<bool NT>(int f, float (&Pk)[3][VX], float (&Pm)[3][VX], float (&B)[VX]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      load_row<VX, VEC, XT, NT>(a.p_k + c * a.vol, bd, f, y, xb, Pk[c]);
      if (!a.first) {
        load_row<VX, VEC, XT, NT>(a.p_km1 + c * a.vol, bd, f, y, xb, Pm[c]);
      } else {
#pragma unroll
        for (int v = 0; v < VX; ++v) Pm[c][v] = 0.0f;
      }
    }
    load_row<VX, VEC, XT, NT>(a.b, bd, f, y, xb, B);
  };
  // Rows that no other tile reads (waves 4 .. NW-5 of an untiled row: the halo rows of the
  // tiles above and below are this tile's first and last two output rows) are loaded
  // non-temporally, which keeps them from displacing the rows that neighbouring tiles read back
  // (256^3 pass 153.6 -> 147.8 us). In x tiles every row's edge columns are a neighbouring
  // tile's halo, and the same hint made the 384^3 pass slower (587 -> 608 us); so did loading
  // every row non-temporally (256^3: 165 us).
  const bool stream_row = !XT && w >= 4 && w <= NW - 5;
  auto load_raw = [&](int f, float (&Pk)[3][VX], float (&Pm)[3][VX], float (&B)[VX]) {
    if (stream_row) {
      load_raw_as.template operator()<true>(f, Pk, Pm, B);
    } else {
      load_raw_as.template operator()<false>(f, Pk, Pm, B);
    }
  };
  // r_k = p_k + m1 (p_k - p_{k-1}), or p_k on the first step.
  auto extrapolate_k = [&](const float (&Pk)[3][VX], const float (&Pm)[3][VX], float (&R)[3][VX]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
      for (int v = 0; v < VX; ++v) R[c][v] = a.first ? Pk[c][v] : Pk[c][v] + a.m1 * (Pk[c][v] - Pm[c][v]);
    }
  };
  // r_k, p_k and b of this row on plane f.
  auto load_step1 = [&](int f, float (&R)[3][VX], float (&Pk)[3][VX], float (&B)[VX]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float pm[VX];
      load_row<VX, VEC, XT>(a.p_k + c * a.vol, bd, f, y, xb, Pk[c]);
      if (!a.first) {
        load_row<VX, VEC, XT>(a.p_km1 + c * a.vol, bd, f, y, xb, pm);
      } else {
#pragma unroll
        for (int v = 0; v < VX; ++v) pm[v] = 0.0f;
      }
#pragma unroll
      for (int v = 0; v < VX; ++v) R[c][v] = a.first ? Pk[c][v] : Pk[c][v] + a.m1 * (Pk[c][v] - pm[v]);
    }
    load_row<VX, VEC, XT>(a.b, bd, f, y, xb, B);
  };

  // Prologue: r_k on plane z0 - pre (z component) and z_k on plane z0 - pre + 1.
  const int zs = z0 - pre + 1;
  float rz_m2[VX];
  {
    float pk[VX], pm[VX];
    load_row<VX, VEC, XT>(a.p_k, bd, zs - 1, y, xb, pk);
    if (!a.first) {
      load_row<VX, VEC, XT>(a.p_km1, bd, zs - 1, y, xb, pm);
    } else {
#pragma unroll
      for (int v = 0; v < VX; ++v) pm[v] = 0.0f;
    }
#pragma unroll
    for (int v = 0; v < VX; ++v) rz_m2[v] = a.first ? pk[v] : pk[v] + a.m1 * (pk[v] - pm[v]);
  }
  float rk1[3][VX], pk1[3][VX], b1[VX], zk1[VX];
  load_step1(zs, rk1, pk1, b1);
  lds_put<VX, VEC>(ry, w, LW, xl, rk1[1]);
  __syncthreads();
  {
    float up[VX];
    lds_get<VX, VEC>(ry, wu, LW, xl, up);
    pair_primal<VX, XT, VEC>(a, zs, y, xb, rk1, rz_m2, up, b1, zk1);
  }
  lds_put<VX, VEC>(zk, w, LW, xl, zk1);
  float rq0[3][VX], zq0[VX], zdn_k[VX], zdn_q[VX];
#pragma unroll
  for (int v = 0; v < VX; ++v) {
    zq0[v] = 0.0f;
    zdn_q[v] = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) rq0[c][v] = 0.0f;
  }
  __syncthreads();
  lds_get<VX, VEC>(zk, wd, LW, xl, zdn_k);  // z_k of the row below on plane zs

  // The loads of the next plane are issued once the first dual step is done, before the second
  // barrier, so that their latency overlaps the second primal and the last dual step instead
  // of stalling the block at the next barrier (issued after the third barrier the pass took 6%
  // longer at 256 and 7% at 384; any earlier, the registers spill); r_k is formed from them at
  // the top of the next iteration.
  // What each wave's row feeds: the output rows (waves 2 .. NW-3) need p_{k+2}, which needs
  // z_{k+1} one row further down (waves 2 .. NW-2), which needs r_{k+1} one row up (waves
  // 1 .. NW-2), which needs z_k one row down (waves 1 .. NW-1); wave 0 only supplies r_k. A
  // wave skips the quantities no output depends on (wave-uniform branches; the variant with
  // element-wise row access keeps computing them, which holds it below the register limit).
  const bool need_zk = !VEC || w >= 1;
  const bool need_pk1 = !VEC || (w >= 1 && w <= NW - 2);
  const bool need_zq = !VEC || (w >= 2 && w <= NW - 2);
  const bool need_pk2 = !VEC || (w >= TOP && w < TOP + T);
  float pk2[3][VX], pm2[3][VX], b2[VX];
  load_raw(z0 - pre + 2, pk2, pm2, b2);
  for (int t = z0 - pre; t < z1; ++t) {
    const int f = t + 2;
    float rk2[3][VX];
    extrapolate_k(pk2, pm2, rk2);
    lds_put<VX, VEC>(ry, w, LW, xl, rk2[1]);
    __syncthreads();  // (1) r_k of plane f visible
    float zk2[VX] = {};
    if (need_zk) {
      float up[VX];
      lds_get<VX, VEC>(ry, wu, LW, xl, up);
      pair_primal<VX, XT, VEC>(a, f, y, xb, rk2, rk1[0], up, b2, zk2);
    }
    // p_{k+1} and r_{k+1} on plane t + 1.
    float rq1[3][VX] = {};
    if (need_pk1) {
      float pq1[3][VX];
      pair_dual<VX, XT, VEC>(a, t + 1, y, xb, zk1, zk2, zdn_k, rk1, pq1);
      if (a.p_k1 && out_row && out_col && t + 1 >= z0 && t + 1 < z1) {
#pragma unroll
        for (int c = 0; c < 3; ++c) store_row<VX, VEC, true>(a.p_k1 + c * a.vol, bd, t + 1, y, xb, pq1[c]);
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int v = 0; v < VX; ++v) rq1[c][v] = pq1[c][v] + a.m2 * (pq1[c][v] - pk1[c][v]);
      }
    }
    float pk3[3][VX], pm3[3][VX], b3[VX];
    load_raw(f + 1, pk3, pm3, b3);
    lds_put<VX, VEC>(zk, w, LW, xl, zk2);
    lds_put<VX, VEC>(rqy, w, LW, xl, rq1[1]);
    __syncthreads();  // (2) z_k of plane f and r_{k+1} of plane t + 1 visible
    lds_get<VX, VEC>(zk, wd, LW, xl, zdn_k);  // for the next plane's dual step
    float zq1[VX] = {};
    if (need_zq) {
      float up[VX];
      lds_get<VX, VEC>(rqy, wu, LW, xl, up);
      pair_primal<VX, XT, VEC>(a, t + 1, y, xb, rq1, rq0[0], up, b1, zq1);
    }
    lds_put<VX, VEC>(zq, w, LW, xl, zq1);
    __syncthreads();  // (3) z_{k+1} of plane t + 1 visible
    // p_{k+2} on plane t.
    if (t >= z0 && need_pk2) {
      float p2[3][VX];
      pair_dual<VX, XT, VEC>(a, t, y, xb, zq0, zq1, zdn_q, rq0, p2);
      if (out_row && out_col) {
#pragma unroll
        for (int c = 0; c < 3; ++c) store_row<VX, VEC, true>(a.p_k2 + c * a.vol, bd, t, y, xb, p2[c]);
      }
    }
    lds_get<VX, VEC>(zq, wd, LW, xl, zdn_q);  // z_{k+1} of the row below on plane t + 1
#pragma unroll
    for (int v = 0; v < VX; ++v) {
      b1[v] = b2[v];
      b2[v] = b3[v];
      zk1[v] = zk2[v];
      zq0[v] = zq1[v];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        rk1[c][v] = rk2[c][v];
        pk1[c][v] = pk2[c][v];
        pk2[c][v] = pk3[c][v];
        pm2[c][v] = pm3[c][v];
        rq0[c][v] = rq1[c][v];
      }
    }
  }
}

// out = max(b - lambda L(p), 0) (optionally its refractive index), same strip layout.
template <int VX, int TY, bool VEC, bool BATCH>
__global__ void __launch_bounds__(64) k_tv_primal(TvArgs a, int chunks, float* out, int to_index,
                                                  float nm2, float inv_k02) {
  if constexpr (BATCH) {
    a.b = shifted(a.b, batch_volume() * a.vol);
    a.p_in = shifted(a.p_in, 3 * batch_volume() * a.vol);
    out += batch_volume() * a.vol;
  }
  const int t = xcd_tile();
  const int y0 = (t / chunks) * TY;
  const int z0 = (t % chunks) * a.zlen;
  const int z1 = min(z0 + a.zlen, a.nz);
  const int xb = static_cast<int>(threadIdx.x) * VX;
  float p[3][TY + 1][VX];
  float pz_below[TY + 1][VX];
  float py_top[VX];
  float bb[TY + 1][VX];
  float x3[TY + 1][VX];
#pragma unroll
  for (int j = 0; j < TY; ++j) load_row<VX, VEC>(a.p_in, a, z0 - 1, y0 + j, xb, pz_below[j]);
#pragma unroll
  for (int v = 0; v < VX; ++v) pz_below[TY][v] = 0.0f;
  for (int z = z0; z < z1; ++z) {
#pragma unroll
    for (int j = 0; j < TY; ++j) {
      // Read once (non-temporal), except the y component of the strip's last row, which the
      // strip below reads as its row above (256^3 FISTA-TV request 20.33 -> 20.17 ms).
      load_row<VX, VEC, false, true>(a.p_in, a, z, y0 + j, xb, p[0][j]);
      if (j == TY - 1) {
        load_row<VX, VEC>(a.p_in + a.vol, a, z, y0 + j, xb, p[1][j]);
      } else {
        load_row<VX, VEC, false, true>(a.p_in + a.vol, a, z, y0 + j, xb, p[1][j]);
      }
      load_row<VX, VEC, false, true>(a.p_in + 2 * a.vol, a, z, y0 + j, xb, p[2][j]);
      load_row<VX, VEC, false, true>(a.b, a, z, y0 + j, xb, bb[j]);
    }
#pragma unroll
    for (int v = 0; v < VX; ++v) {
      p[0][TY][v] = p[1][TY][v] = p[2][TY][v] = 0.0f;
      bb[TY][v] = 0.0f;
    }
    load_row<VX, VEC>(a.p_in + a.vol, a, z, y0 - 1, xb, py_top);
    primal_plane<VX, TY>(a, z, y0, xb, p, py_top, pz_below, bb, x3);
#pragma unroll
    for (int j = 0; j < TY; ++j) {
      float o[VX];
#pragma unroll
      for (int v = 0; v < VX; ++v) {
        o[v] = to_index ? index_from_potential(x3[j][v], 0.0f, nm2, inv_k02) : x3[j][v];
        pz_below[j][v] = p[0][j][v];
      }
      store_row<VX, VEC>(out, a, z, y0 + j, xb, o);
    }
  }
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

namespace {

constexpr int kRows = 2;

int lanes_x(int nx) {
  const int vx = (nx + 63) / 64;
  if (vx <= 1) return 1;
  if (vx <= 2) return 2;
  if (vx <= 4) return 4;
  if (vx <= 6) return 6;
  if (vx <= 8) return 8;
  throw InvalidArgument("tv: volume side above 512");
}

// CUs of the current device, queried once per device.
int device_cus() {
  static std::atomic<int> cache[64];
  int dev = 0;
  check(hipGetDevice(&dev), "hipGetDevice");
  require(dev >= 0 && dev < 64, "tv: device index above 63");
  int cus = cache[dev].load(std::memory_order_relaxed);
  if (cus == 0) {
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
          "hipDeviceGetAttribute");
    cus = std::max(cus, 1);
    cache[dev].store(cus, std::memory_order_relaxed);
  }
  return cus;
}

// One wave per strip of kRows rows and z chunk: chunks of kZChunk planes, halved (down to 2)
// while the grid holds fewer than two waves per SIMD (128^3: 4 planes and 2048 waves, where
// 16-plane chunks gave 512 waves for 1024 SIMDs).
dim3 strip_grid(int ny, int nz, TvArgs& a, int* chunks, int batch) {
  const long strips = (ny + kRows - 1) / kRows;
  const long want = 8L * device_cus();
  int zlen = kZChunk;
  while (zlen > 2 && strips * ((nz + zlen - 1) / zlen) * batch < want) zlen /= 2;
  a.zlen = zlen;
  *chunks = (nz + zlen - 1) / zlen;
  return dim3(static_cast<unsigned>(strips * *chunks), static_cast<unsigned>(batch));
}

// Whole-vector row access (see load_row) when the row splits into whole lane chunks.
template <int VX>
bool rows_vec(int nx) {
  return VX % 2 == 0 && nx % VX == 0;
}

template <int VX, bool BATCH>
void launch_step_as(const TvArgs& a, int chunks, dim3 grid, hipStream_t s) {
  if constexpr (VX % 2 == 0) {
    if (rows_vec<VX>(a.nx)) {
      hipLaunchKernelGGL((k_fgp_step<VX, kRows, true, BATCH>), grid, dim3(64), 0, s, a, chunks);
      return;
    }
  }
  hipLaunchKernelGGL((k_fgp_step<VX, kRows, false, BATCH>), grid, dim3(64), 0, s, a, chunks);
}

// A grid with a second dimension runs a batch of volumes (strip_grid).
template <int VX>
void launch_step(const TvArgs& a, int chunks, dim3 grid, hipStream_t s) {
  if (grid.y > 1) {
    launch_step_as<VX, true>(a, chunks, grid, s);
  } else {
    launch_step_as<VX, false>(a, chunks, grid, s);
  }
}

template <int VX, bool BATCH>
void launch_primal_as(const TvArgs& a, int chunks, dim3 grid, float* out, int ti, float nm2,
                      float inv_k02, hipStream_t s) {
  if constexpr (VX % 2 == 0) {
    if (rows_vec<VX>(a.nx)) {
      hipLaunchKernelGGL((k_tv_primal<VX, kRows, true, BATCH>), grid, dim3(64), 0, s, a, chunks,
                         out, ti, nm2, inv_k02);
      return;
    }
  }
  hipLaunchKernelGGL((k_tv_primal<VX, kRows, false, BATCH>), grid, dim3(64), 0, s, a, chunks, out,
                     ti, nm2, inv_k02);
}

template <int VX>
void launch_primal(const TvArgs& a, int chunks, dim3 grid, float* out, int ti, float nm2,
                   float inv_k02, hipStream_t s) {
  if (grid.y > 1) {
    launch_primal_as<VX, true>(a, chunks, grid, out, ti, nm2, inv_k02, s);
  } else {
    launch_primal_as<VX, false>(a, chunks, grid, out, ti, nm2, inv_k02, s);
  }
}

}  // namespace

void fgp_step(const float* b, const float* p_in, const float* p_prev, float* p_out, int nz,
              int ny, int nx, float lambda, float step, float momentum, bool first,
              hipStream_t s, int batch) {
  require(batch >= 1 && batch <= 65535, "fgp_step: batch must be 1 to 65535");
  TvArgs a{b, p_in, p_prev, p_out, static_cast<long>(nz) * ny * nx, nz, ny, nx,
           lambda, step, momentum, first ? 1 : 0};
  int chunks = 0;
  const dim3 grid = strip_grid(ny, nz, a, &chunks, batch);
  switch (lanes_x(nx)) {
    case 1: launch_step<1>(a, chunks, grid, s); break;
    case 2: launch_step<2>(a, chunks, grid, s); break;
    case 4: launch_step<4>(a, chunks, grid, s); break;
    case 6: launch_step<6>(a, chunks, grid, s); break;
    default: launch_step<8>(a, chunks, grid, s); break;
  }
  check_launch("fgp_step");
}

void tv_primal(const float* b, const float* p, int nz, int ny, int nx, float lambda, float* out,
               bool to_index, float nm2, float inv_k02, hipStream_t s, int batch) {
  require(batch >= 1 && batch <= 65535, "tv_primal: batch must be 1 to 65535");
  TvArgs a{b, p, nullptr, nullptr, static_cast<long>(nz) * ny * nx, nz, ny, nx,
           lambda, 0.0f, 0.0f, 1};
  int chunks = 0;
  const dim3 grid = strip_grid(ny, nz, a, &chunks, batch);
  const int ti = to_index ? 1 : 0;
  switch (lanes_x(nx)) {
    case 1: launch_primal<1>(a, chunks, grid, out, ti, nm2, inv_k02, s); break;
    case 2: launch_primal<2>(a, chunks, grid, out, ti, nm2, inv_k02, s); break;
    case 4: launch_primal<4>(a, chunks, grid, out, ti, nm2, inv_k02, s); break;
    case 6: launch_primal<6>(a, chunks, grid, out, ti, nm2, inv_k02, s); break;
    default: launch_primal<8>(a, chunks, grid, out, ti, nm2, inv_k02, s); break;
  }
  check_launch("tv_primal");
}

namespace {

// Rows per block of the two-step kernel: NW waves, NW - 4 output rows, as many waves as the
// registers of VX floats per lane leave room for in one block per CU. Rows up to 256 wide are
// held whole by a wave; wider rows are cut into x tiles of at most 62 VX output columns (one
// lane of VX columns of halo on either side), balanced. Tiles of 4 columns per lane (12 waves,
// 8 output rows) when the row width is a multiple of 4, so that every access is a float4: at
// 384^3, 2 tiles of 192 columns ran a pass in 593 us against 618 us for 4 tiles of 96 at 2
// columns per lane (16 waves, 12 output rows), which leave fewer lanes and rows idle; at 512^3
// and 320^3 the two were equal (1667 against 1658 us, 417 against 420 us). Other widths take
// tiles of 2 columns per lane, whole float2s for even widths.
int pair_waves(int vx) { return vx == 1 ? 16 : (vx == 2 ? 16 : 12); }
int pair_tile_vx(int nx) { return nx % 4 == 0 ? 4 : 2; }

// Planes per z chunk for a grid of `tiles` row tiles: the length from kPairChunkMin to
// kPairChunk that minimises the rounds of blocks each CU runs times the planes a block marches
// (its chunk and two warm-up planes), the longest on ties. 256^3: 32 planes, 256 blocks, one
// round; 384^3 in x tiles of 4 columns per lane: 48 planes, 768 blocks, three full rounds (64
// planes would leave the third round a quarter full; tiles of 2 columns per lane took 64);
// 128^3: 6 planes and 242 blocks (8-plane chunks left 80 CUs idle and spent a fifth of each
// block on warm-up).
int pair_chunk(const void* kernel, std::atomic<int>& per_cu_cache, int threads, std::size_t lds,
               int tiles, int nz) {
  int per_cu = per_cu_cache.load(std::memory_order_relaxed);
  if (per_cu == 0) {
    check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds),
          "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    per_cu = std::max(per_cu, 1);
    per_cu_cache.store(per_cu, std::memory_order_relaxed);
  }
  const long slots = static_cast<long>(device_cus()) * per_cu;
  int best = kPairChunk;
  long best_cost = -1;
  for (int zl = kPairChunk; zl >= kPairChunkMin; --zl) {
    const long blocks = static_cast<long>(tiles) * ((nz + zl - 1) / zl);
    const long cost = (blocks + slots - 1) / slots * (zl + 2);
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = zl;
    }
  }
  return best;
}

// Volumes of a batch are the grid's second dimension; their blocks count in the chunk choice.
template <int VX, int NW, bool VEC, bool XT, bool BATCH>
void launch_pair_kernel_as(PairArgs a, int tiles, std::size_t lds, hipStream_t s, int batch) {
  // Blocks per CU of this kernel (its registers and LDS; the same on every gfx950 device).
  static std::atomic<int> per_cu{0};
  a.zlen = pair_chunk(reinterpret_cast<const void*>(&k_fgp_pair<VX, NW, VEC, XT, BATCH>), per_cu,
                      NW * 64, lds, tiles * batch, a.nz);
  const int chunks = (a.nz + a.zlen - 1) / a.zlen;
  hipLaunchKernelGGL((k_fgp_pair<VX, NW, VEC, XT, BATCH>),
                     dim3(static_cast<unsigned>(tiles * chunks), static_cast<unsigned>(batch)),
                     dim3(NW * 64), lds, s, a, chunks);
}

template <int VX, int NW, bool VEC, bool XT>
void launch_pair_kernel(PairArgs a, int tiles, std::size_t lds, hipStream_t s, int batch) {
  if (batch > 1) {
    launch_pair_kernel_as<VX, NW, VEC, XT, true>(a, tiles, lds, s, batch);
  } else {
    launch_pair_kernel_as<VX, NW, VEC, XT, false>(a, tiles, lds, s, batch);
  }
}

template <int VX, int NW, bool XT>
void launch_pair(const PairArgs& a, int tiles, hipStream_t s, int batch) {
  // Four row buffers of NW rows of nx floats (64 VX per x tile): at most 48 KiB.
  const std::size_t lds = 4 * static_cast<std::size_t>(NW) * (XT ? 64 * VX : a.nx) * sizeof(float);
  if constexpr (VX % 2 == 0) {
    if (rows_vec<VX>(a.nx)) {
      launch_pair_kernel<VX, NW, true, XT>(a, tiles, lds, s, batch);
      return;
    }
  }
  launch_pair_kernel<VX, NW, false, XT>(a, tiles, lds, s, batch);
}

}  // namespace

bool fgp_pair_supported(int nx) { return nx >= 1 && nx <= 512; }

void fgp_pair(const float* b, const float* p_k, const float* p_km1, float* p_k1, float* p_k2,
              int nz, int ny, int nx, float lambda, float step, float m1, bool first, float m2,
              hipStream_t s, int batch) {
  require(fgp_pair_supported(nx), "fgp_pair: rows up to 512 wide");
  require(batch >= 1 && batch <= 65535, "fgp_pair: batch must be 1 to 65535");
  int vx = std::min(lanes_x(nx), 4);
  int tiles_x = 1;
  int xt = nx;
  if (nx > 64 * 4) {
    vx = pair_tile_vx(nx);
    const int widest = 62 * vx;
    tiles_x = (nx + widest - 1) / widest;
    xt = ((nx + tiles_x - 1) / tiles_x + vx - 1) / vx * vx;  // balanced tiles, whole lanes
  }
  const int nw = pair_waves(vx);
  const int rows = nw - 4;
  const int tiles_y = (ny + rows - 1) / rows;
  const int tiles = tiles_x * tiles_y;
  PairArgs a{b,  p_k,    p_km1, p_k1, p_k2,          static_cast<long>(nz) * ny * nx,
             nz, ny,     nx,    lambda, step,        m1, first ? 1 : 0, m2, 0, xt, tiles_x};
  if (tiles_x > 1) {
    if (vx == 4) {
      launch_pair<4, 12, true>(a, tiles, s, batch);
    } else {
      launch_pair<2, 16, true>(a, tiles, s, batch);
    }
  } else {
    switch (vx) {
      case 1: launch_pair<1, 16, false>(a, tiles, s, batch); break;
      case 2: launch_pair<2, 16, false>(a, tiles, s, batch); break;
      default: launch_pair<4, 12, false>(a, tiles, s, batch); break;
    }
  }
  check_launch("fgp_pair");
}

namespace {

// Floats of one dual buffer (three components per volume of the batch, one volume's after
// another's), rounded up to 256 bytes so that every buffer of a workspace starts 256-byte
// aligned.
std::size_t dual_floats(int nz, int ny, int nx, int batch = 1) {
  const std::size_t three = 3 * static_cast<std::size_t>(nz) * ny * nx * batch;
  return (three + 63) / 64 * 64;
}

}  // namespace

std::size_t TvState::workspace_floats(int nz, int ny, int nx) { return 4 * dual_floats(nz, ny, nx); }

TvState::TvState(int nz, int ny, int nx, int batch)
    : nz_(nz), ny_(ny), nx_(nx), vol_(static_cast<long>(nz) * ny * nx), cap_(batch),
      owned_(4 * dual_floats(nz, ny, nx, std::max(batch, 1))) {
  require(batch >= 1 && batch <= 65535, "tv: the batch must be 1 to 65535");
  const std::size_t stride = dual_floats(nz, ny, nx, batch);
  for (int i = 0; i < 4; ++i) p_[i] = owned_.get() + i * stride;
}

TvState::TvState(int nz, int ny, int nx, float* workspace)
    : nz_(nz), ny_(ny), nx_(nx), vol_(static_cast<long>(nz) * ny * nx) {
  const std::size_t stride = dual_floats(nz, ny, nx);
  for (int i = 0; i < 4; ++i) p_[i] = workspace + i * stride;
}

void TvState::reset(hipStream_t s, int batch) {
  require(batch >= 1 && batch <= cap_, "tv: batch beyond the state's capacity");
  batch_ = batch;
  fill(p_[0], 3 * vol_ * batch, 0.0f, s);
  cur_ = 0;
}

std::size_t TvState::bytes() const { return owned_.bytes(); }

// tv_prox of the reference (src/cpu/tv.cpp:7-44): r <- p, `inner` FGP steps, out = primal(p).
// The dual iterates rotate through four buffers: a pass reads p_k and p_{k-1} (r is formed
// inside it) and writes p_{k+1} (and p_{k+2} for a two-step pass) to buffers nobody reads in
// that pass. Steps go in pairs where the row width allows, with an odd last step on its own; the
// momentum sequence is that of single steps. A last pass that is a pair skips p_{k+1}, which the
// next call does not read (its first step uses r = p). The prox leaves p in buffer cur_ for the
// next call.
void TvState::prox(const float* b, float* out, float lambda, int inner, hipStream_t s,
                   bool to_index, float nm2, float inv_k02, Event* t0, Event* t1) {
  const float step = 1.0f / (12.0f * lambda);
  const bool pairs = fused_ && fgp_pair_supported(nx_);
  float t = 1.0f;
  float m_prev = 0.0f;
  int cur = cur_;
  int prev = (cur_ + 1) % 4;
  launches_ = 0;
  const auto free_buffers = [&](int& f1, int& f2) {
    f1 = -1;
    f2 = -1;
    for (int i = 0; i < 4; ++i) {
      if (i == cur || i == prev) continue;
      if (f1 < 0) {
        f1 = i;
      } else {
        f2 = i;
      }
    }
  };
  if (t0) t0->record(s);
  for (int k = 0; k < inner;) {
    int f1 = 0, f2 = 0;
    free_buffers(f1, f2);
    if (pairs && k + 1 < inner) {
      const float m1 = m_prev;
      float tn = fista_t(t);
      const float m2 = (t - 1.0f) / tn;
      t = tn;
      // The last pass skips p_{k+1}: the next prox starts from p alone.
      const bool last = k + 2 == inner;
      fgp_pair(b, p_[cur], p_[prev], last ? nullptr : p_[f1], p_[f2],
               nz_, ny_, nx_, lambda, step, m1, k == 0, m2, s, batch_);
      tn = fista_t(t);
      m_prev = (t - 1.0f) / tn;
      t = tn;
      prev = f1;
      cur = f2;
      k += 2;
    } else {
      fgp_step(b, p_[cur], p_[prev], p_[f1], nz_, ny_, nx_, lambda, step,
               m_prev, k == 0, s, batch_);
      const float tn = fista_t(t);
      m_prev = (t - 1.0f) / tn;
      t = tn;
      prev = cur;
      cur = f1;
      k += 1;
    }
    ++launches_;
  }
  if (t1) t1->record(s);
  cur_ = cur;
  tv_primal(b, p_[cur_], nz_, ny_, nx_, lambda, out, to_index, nm2, inv_k02, s, batch_);
}

}  // namespace qh
