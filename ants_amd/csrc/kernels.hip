// SyN + CC kernels for MI355X (gfx950, wave64).
//
// Every kernel here is a stencil, gather or elementwise pass over fields of
// a few bytes per voxel, so the path is HBM-bound and uses no MFMA.  Global
// reductions are deterministic: each block writes one partial, the last
// block to finish sums the partials in block order.
#include "kernels.hpp"

#include <hip/hip_runtime.h>
#include <atomic>
#include <cfloat>
#include <cstdint>
#include <type_traits>
#include <utility>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace ants {
namespace {

constexpr int kBlock = 256;

struct PtrPack4 {
  const float* p[4];
};
struct MutPack4 {
  float* p[4];
};
struct SlotPack4 {
  RedSlot* p[4];
};

__device__ __forceinline__ int64_t lin(const GridDev& g, int x, int y, int z) {
  return static_cast<int64_t>(x) + static_cast<int64_t>(g.n[0]) * (y + static_cast<int64_t>(g.n[1]) * z);
}

__device__ __forceinline__ uint32_t fast_div(uint32_t n, uint32_t m, uint32_t s1, uint32_t s2) {
  const uint32_t t = __umulhi(m, n);
  return (t + ((n - t) >> s1)) >> s2;
}

// Voxel index -> (x, y, z); grids hold fewer than 2^31 voxels.
__device__ __forceinline__ void unlin(const GridDev& g, int64_t i, int& x, int& y, int& z) {
  const uint32_t n = static_cast<uint32_t>(i);
  const uint32_t r = fast_div(n, g.div_m[0], g.div_s1[0], g.div_s2[0]);
  x = static_cast<int>(n - r * static_cast<uint32_t>(g.n[0]));
  const uint32_t q = fast_div(r, g.div_m[1], g.div_s1[1], g.div_s2[1]);
  y = static_cast<int>(r - q * static_cast<uint32_t>(g.n[1]));
  z = static_cast<int>(q);
}

__device__ __forceinline__ bool on_boundary(const GridDev& g, int x, int y, int z) {
  bool b = x == 0 || x == g.n[0] - 1 || y == 0 || y == g.n[1] - 1;
  if (g.D == 3) b = b || z == 0 || z == g.n[2] - 1;
  return b;
}

// Continuous index in grid `gi` of the point (level-grid voxel of `go`) + u.
template <int D>
__device__ __forceinline__ void cidx_of(const GridDev& go, int x, int y, int z, const float* u, const GridDev& gi,
                                        float c[3]) {
  const int ix[3] = {x, y, z};
  if (go.oblique | gi.oblique) {
    // point = org_o + dir_o (ix .* sp_o) + u ; index = (dir_i^T (point - org_i)) .* isp_i
    float q[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int r = 0; r < D; ++r) {
      float p = go.org[r];
#pragma unroll
      for (int k = 0; k < D; ++k) p += go.dir[r * 3 + k] * (static_cast<float>(ix[k]) * go.sp[k]);
      if (u) p += u[r];
      q[r] = p - gi.org[r];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float t = 0.f;
#pragma unroll
      for (int k = 0; k < D; ++k) t += gi.dir[k * 3 + d] * q[k];
      c[d] = d < D ? t * gi.isp[d] : 0.f;
    }
    return;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (d < D) {
      float p = go.org[d] + static_cast<float>(ix[d]) * go.sp[d];
      if (u) p += u[d];
      c[d] = (p - gi.org[d]) * gi.isp[d];
    } else {
      c[d] = 0.f;
    }
  }
}

// c = G ix + Q u + h (WarpMap): the level-grid point displaced by u, mapped
// through an affine initial transform, as a continuous index.
template <int D>
__device__ __forceinline__ void map_cidx(const WarpMap& m, int x, int y, int z, const float* u, float c[3]) {
  const float ix[3] = {static_cast<float>(x), static_cast<float>(y), static_cast<float>(z)};
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    if (r >= D) {
      c[r] = 0.f;
      continue;
    }
    float t = m.h[r];
#pragma unroll
    for (int k = 0; k < D; ++k) t += m.G[r * 3 + k] * ix[k];
#pragma unroll
    for (int k = 0; k < D; ++k) t += m.Q[r * 3 + k] * u[k];
    c[r] = t;
  }
}

// ITK ImageFunction::IsInsideBuffer on a continuous index: -0.5 <= c < n - 0.5.
template <int D>
__device__ __forceinline__ bool inside(const GridDev& g, const float c[3]) {
#pragma unroll
  for (int d = 0; d < D; ++d)
    if (!(c[d] >= -0.5f && c[d] < static_cast<float>(g.n[d]) - 0.5f)) return false;
  return true;
}

// BF16 storage of smoothed pyramid images: round to nearest even on store,
// exact widening on load.
__device__ __forceinline__ uint16_t f2bf(float f) {
  const uint32_t u = __float_as_uint(f);
  const uint32_t r = ((u >> 16) & 1u) + 0x7fffu;
  return static_cast<uint16_t>((u + r) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(static_cast<uint32_t>(b) << 16); }
__device__ __forceinline__ float ld(const float* p, int64_t i) { return p[i]; }
__device__ __forceinline__ float ld(const uint16_t* p, int64_t i) { return bf2f(p[i]); }

#ifndef ANTS_IMG_PAIR_LOADS
#define ANTS_IMG_PAIR_LOADS 1
#endif
// two floats at 4-byte alignment: one dwordx2 global load
typedef float float2u __attribute__((ext_vector_type(2), aligned(4)));
// Trilinear sample of a scalar image of element type T (float or bf16) at an
// inside point, neighbours clamped to the buffer.
template <int D, class T>
__device__ __forceinline__ float interp_img(const GridDev& g, const T* __restrict__ img, const float c[3]) {
  int lo[3], hi[3];
  float w[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (d < D) {
      float fl = floorf(c[d]);
      int b = static_cast<int>(fl);
      float t = c[d] - fl;
      if (b < 0) {
        b = 0;
        t = 0.f;
      }
      lo[d] = b;
      hi[d] = b + 1 < g.n[d] ? b + 1 : b;
      w[d] = t;
    } else {
      lo[d] = hi[d] = 0;
      w[d] = 0.f;
    }
  }
  const int64_t sy = g.n[0], sz = static_cast<int64_t>(g.n[0]) * g.n[1];
  if (D == 2) {
    const float v00 = ld(img, lo[0] + sy * lo[1]), v10 = ld(img, hi[0] + sy * lo[1]);
    const float v01 = ld(img, lo[0] + sy * hi[1]), v11 = ld(img, hi[0] + sy * hi[1]);
    const float a = v00 + (v10 - v00) * w[0];
    const float b = v01 + (v11 - v01) * w[0];
    return a + (b - a) * w[1];
  }
  const int64_t z0 = sz * lo[2], z1 = sz * hi[2], y0 = sy * lo[1], y1 = sy * hi[1];
#if ANTS_IMG_PAIR_LOADS
  if (std::is_same<T, float>::value && g.n[0] >= 2) {
    // the x neighbours are adjacent: one 8-byte load per row (dword
    // alignment suffices for global loads).  At the last column hi == lo:
    // the pair ending at lo is loaded instead (n[0] >= 2 keeps it in the row)
    const bool pair = hi[0] == lo[0] + 1;
    const int64_t x0 = pair ? lo[0] : lo[0] - 1;
    auto row = [&](int64_t r, float& v0, float& v1) {
      const float2u q = *reinterpret_cast<const float2u*>(reinterpret_cast<const float*>(img) + x0 + r);
      v0 = pair ? q.x : q.y;
      v1 = q.y;
    };
    float p00, q00, p10, q10, p01, q01, p11, q11;
    row(y0 + z0, p00, q00);
    row(y1 + z0, p10, q10);
    row(y0 + z1, p01, q01);
    row(y1 + z1, p11, q11);
    const float a00 = p00 + (q00 - p00) * w[0];
    const float a10 = p10 + (q10 - p10) * w[0];
    const float a01 = p01 + (q01 - p01) * w[0];
    const float a11 = p11 + (q11 - p11) * w[0];
    const float b0 = a00 + (a10 - a00) * w[1];
    const float b1 = a01 + (a11 - a01) * w[1];
    return b0 + (b1 - b0) * w[2];
  }
#endif
  const float a00 = ld(img, lo[0] + y0 + z0) + (ld(img, hi[0] + y0 + z0) - ld(img, lo[0] + y0 + z0)) * w[0];
  const float a10 = ld(img, lo[0] + y1 + z0) + (ld(img, hi[0] + y1 + z0) - ld(img, lo[0] + y1 + z0)) * w[0];
  const float a01 = ld(img, lo[0] + y0 + z1) + (ld(img, hi[0] + y0 + z1) - ld(img, lo[0] + y0 + z1)) * w[0];
  const float a11 = ld(img, lo[0] + y1 + z1) + (ld(img, hi[0] + y1 + z1) - ld(img, lo[0] + y1 + z1)) * w[0];
  const float b0 = a00 + (a10 - a00) * w[1];
  const float b1 = a01 + (a11 - a01) * w[1];
  return b0 + (b1 - b0) * w[2];
}

// Linear interpolation of NC interleaved components at an inside point:
// neighbours clamped to the buffer (ITK Linear/VectorLinear interpolators).
// kClampAll also clamps neighbours above the buffer, so that the loads are
// in bounds for any point and may be issued before the inside test; for an
// inside point the result is the same.
// IT is the element-index type: int when the buffer has fewer than 2^31
// elements (cheaper address arithmetic), int64_t otherwise.
template <int D, int NC, bool kClampAll = false, class IT = int64_t>
__device__ __forceinline__ void interp(const GridDev& g, const float* __restrict__ img, const float c[3], float* v) {
  int lo[3], hi[3];
  float w[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (d < D) {
      float f = floorf(c[d]);
      int b = static_cast<int>(f);
      float t = c[d] - f;
      if (b < 0) {
        b = 0;
        t = 0.f;
      }
      if (kClampAll && b > g.n[d] - 1) {
        b = g.n[d] - 1;
        t = 0.f;
      }
      lo[d] = b;
      hi[d] = b + 1 < g.n[d] ? b + 1 : b;
      w[d] = t;
    } else {
      lo[d] = hi[d] = 0;
      w[d] = 0.f;
    }
  }
#pragma unroll
  for (int k = 0; k < NC; ++k) v[k] = 0.f;
  const IT sy = g.n[0], sz = static_cast<IT>(g.n[0]) * g.n[1];
  if (D == 2) {
    const IT i00 = lo[0] + sy * lo[1], i10 = hi[0] + sy * lo[1];
    const IT i01 = lo[0] + sy * hi[1], i11 = hi[0] + sy * hi[1];
    const float wx = w[0], wy = w[1];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const float a = img[i00 * NC + k] + (img[i10 * NC + k] - img[i00 * NC + k]) * wx;
      const float b = img[i01 * NC + k] + (img[i11 * NC + k] - img[i01 * NC + k]) * wx;
      v[k] = a + (b - a) * wy;
    }
  } else {
    const IT z0 = sz * lo[2], z1 = sz * hi[2], y0 = sy * lo[1], y1 = sy * hi[1];
    const IT i000 = lo[0] + y0 + z0, i100 = hi[0] + y0 + z0, i010 = lo[0] + y1 + z0, i110 = hi[0] + y1 + z0;
    const IT i001 = lo[0] + y0 + z1, i101 = hi[0] + y0 + z1, i011 = lo[0] + y1 + z1, i111 = hi[0] + y1 + z1;
    const float wx = w[0], wy = w[1], wz = w[2];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const float a00 = img[i000 * NC + k] + (img[i100 * NC + k] - img[i000 * NC + k]) * wx;
      const float a10 = img[i010 * NC + k] + (img[i110 * NC + k] - img[i010 * NC + k]) * wx;
      const float a01 = img[i001 * NC + k] + (img[i101 * NC + k] - img[i001 * NC + k]) * wx;
      const float a11 = img[i011 * NC + k] + (img[i111 * NC + k] - img[i011 * NC + k]) * wx;
      const float b0 = a00 + (a10 - a00) * wy;
      const float b1 = a01 + (a11 - a01) * wy;
      v[k] = b0 + (b1 - b0) * wz;
    }
  }
}

// ---------------------------------------------------------------------------
// Deterministic block + grid reduction.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// Block-level: returns the block total in thread 0 (fixed order).
template <int NT>
__device__ __forceinline__ void block_sum_max(double& s, float& m) {
  __shared__ double ss[NT / 64];
  __shared__ float sm[NT / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  s = wave_sum(s);
  m = wave_max(m);
  if (lane == 0) {
    ss[wid] = s;
    sm[wid] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    float mm = 0.f;
    for (int w = 0; w < NT / 64; ++w) {
      t += ss[w];
      mm = fmaxf(mm, sm[w]);
    }
    s = t;
    m = mm;
  }
  __syncthreads();
}

// Every block calls this once with its per-thread values and writes one
// (sum, max) partial with a plain store; the kernel boundary publishes it.
template <int NT>
__device__ __forceinline__ void block_partial(double s, float m, Partial* out) {
  block_sum_max<NT>(s, m);
  if (threadIdx.x == 0) {
    out->s = s;
    out->m = m;
  }
}

// One block sums n partials: thread t takes t, t+NT, ... in order, then the
// block tree of block_sum_max; every thread receives the result.  The order
// is fixed, so every block that reduces the same partials gets the same bits.
// Workgroups are dealt to the 8 XCDs round-robin by id.  slab_block renumbers
// them so that XCD k takes the k-th contiguous slab of block indices: the
// neighbouring rows and planes a trilinear gather touches are then fetched
// into that XCD's L2 by its own blocks.  A bijection of [0, nb) onto itself;
// the nb % 8 leftover blocks keep their place.
__device__ __forceinline__ int64_t slab_block(int64_t b, int64_t nb) {
  const int64_t full = nb / 8;
  return b < 8 * full ? (b % 8) * full + b / 8 : b;
}

// block index of k_compose (measured: 4.74 -> 4.64 ms per headline pair;
// the image warp measured slower with it and keeps the plain order)
#ifndef ANTS_GATHER_XCD_MAP
#define ANTS_GATHER_XCD_MAP 1
#endif
__device__ __forceinline__ int64_t gather_block() {
#if ANTS_GATHER_XCD_MAP
  return slab_block(blockIdx.x, gridDim.x);
#else
  return blockIdx.x;
#endif
}

// Tile (x, y, z) of a block of a 3-D tiled launch.  With `xcd` the tiles
// are renumbered by slab_block so that each XCD holds a contiguous run of
// tiles in x-fastest order (neighbouring tiles, whose halos overlap, share
// an L2); otherwise the hardware order.
__device__ __forceinline__ void tile_of(bool xcd, int& tx, int& ty, int& tz) {
  if (!xcd) {
    tx = blockIdx.x;
    ty = blockIdx.y;
    tz = blockIdx.z;
    return;
  }
  const int64_t gx = gridDim.x, gy = gridDim.y;
  const int64_t nb = gx * gy * gridDim.z;
  const int64_t t = slab_block(blockIdx.x + gx * (blockIdx.y + gy * static_cast<int64_t>(blockIdx.z)), nb);
  tx = static_cast<int>(t % gx);
  const int64_t r = t / gx;
  ty = static_cast<int>(r % gy);
  tz = static_cast<int>(r / gy);
}
// measured per 192x224x192 pair: 41.28 -> 41.12 ms with both (CC 5.84 ->
// 5.79 ms; the smoothing unchanged)
#ifndef ANTS_SMOOTH_XCD_MAP
#define ANTS_SMOOTH_XCD_MAP 1
#endif
#ifndef ANTS_CC_XCD_MAP
#define ANTS_CC_XCD_MAP 1
#endif

template <int NT>
__device__ __forceinline__ void reduce_partials(const Partial* p, int n, double& s, float& m) {
  __shared__ double r_s;
  __shared__ float r_m;
  double a = 0.0;
  float b = 0.f;
  for (int i = threadIdx.x; i < n; i += NT) {
    a += p[i].s;
    b = fmaxf(b, p[i].m);
  }
  block_sum_max<NT>(a, b);
  if (threadIdx.x == 0) {
    r_s = a;
    r_m = b;
  }
  __syncthreads();
  s = r_s;
  m = r_m;
}

template <int NT>
__global__ void __launch_bounds__(NT) k_reduce_slot(const Partial* p, int n, RedSlot* slot) {
  double s;
  float m;
  reduce_partials<NT>(p, n, s, m);
  if (threadIdx.x == 0) {
    slot->sum = s;
    slot->maxbits = __float_as_uint(m);
    slot->computed = 1;
  }
}

// Closes an iteration (one block): reduces the CC partials into `energy`
// when given, copies the nwords statistics words at `src` to pinned host
// memory `dst`, and then sets *flag = 1.  The system-scope release fence
// makes the copy visible to the host before the flag.
template <int NT>
__global__ void __launch_bounds__(NT) k_publish_stats(const Partial* cc_parts, int cc_n, RedSlot* energy,
                                                      const uint32_t* src, uint32_t* dst, int nwords,
                                                      uint32_t* flag) {
  if (cc_parts) {
    double sum;
    float m;
    reduce_partials<NT>(cc_parts, cc_n, sum, m);
    if (threadIdx.x == 0) {
      energy->sum = sum;
      energy->maxbits = __float_as_uint(m);
      energy->computed = 1;
    }
    __syncthreads();  // the energy words are copied below by other threads
  }
  for (int w = threadIdx.x; w < nwords; w += NT) dst[w] = src[w];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) {
    *reinterpret_cast<volatile uint32_t*>(flag) = 1u;
    __threadfence_system();
  }
}

// ---------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}

__device__ __forceinline__ uint32_t float_order(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ void __launch_bounds__(kBlock) k_minmax(const float* __restrict__ in, int64_t n, uint32_t* out2) {
  uint32_t mn = 0xffffffffu, mx = 0u;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const uint32_t o = float_order(in[i]);
    mn = min(mn, o);
    mx = max(mx, o);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    mn = min(mn, static_cast<uint32_t>(__shfl_xor(static_cast<int>(mn), off, 64)));
    mx = max(mx, static_cast<uint32_t>(__shfl_xor(static_cast<int>(mx), off, 64)));
  }
  __shared__ uint32_t smn[kBlock / 64], smx[kBlock / 64];
  if ((threadIdx.x & 63) == 0) {
    smn[threadIdx.x >> 6] = mn;
    smx[threadIdx.x >> 6] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / 64; ++w) {
      mn = min(mn, smn[w]);
      mx = max(mx, smx[w]);
    }
    atomicMin(&out2[0], mn);
    atomicMax(&out2[1], mx);
  }
}

// IntensityWindowingImageFilter: clamp to [lo,hi], x*scale + shift in double.
__global__ void k_window(const float* __restrict__ in, float* __restrict__ out, int64_t n, float lo, float hi,
                         double scale, double shift) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const float x = in[i];
    float v;
    if (x < lo)
      v = 0.f;
    else if (x > hi)
      v = 1.f;
    else
      v = static_cast<float>(static_cast<double>(x) * scale + shift);
    out[i] = v;
  }
}

// ---------------------------------------------------------------------------
// Separable Gaussian along one axis, edge replicated (ZeroFluxNeumann).
template <int NC>
__global__ void __launch_bounds__(kBlock) k_gauss_axis(GridDev g, const float* __restrict__ in, float* __restrict__ out,
                                                       Taps t, int axis) {
  const int64_t N = g.count();
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= N) return;
  int x, y, z;
  unlin(g, i, x, y, z);
  const int pos = axis == 0 ? x : (axis == 1 ? y : z);
  const int len = g.n[axis];
  const int64_t stride = axis == 0 ? 1 : (axis == 1 ? g.n[0] : static_cast<int64_t>(g.n[0]) * g.n[1]);
  const int64_t base = i - static_cast<int64_t>(pos) * stride;
  float acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0.f;
  for (int k = -t.r; k <= t.r; ++k) {
    int p = pos + k;
    p = p < 0 ? 0 : (p >= len ? len - 1 : p);
    const float w = t.c[k < 0 ? -k : k];
    const float* src = in + (base + p * stride) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] += w * src[c];
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) out[i * NC + c] = acc[c];
}

template <int D>
__global__ void __launch_bounds__(kBlock) k_gauss_field_final(GridDev g, PtrPack4 in, PtrPack4 orig, MutPack4 out,
                                                              Taps t, int axis, float w1, float w2,
                                                              int boundary, unsigned int* maxbits) {
  const int k = blockIdx.y;
  const float* __restrict__ src = in.p[k];
  const float* __restrict__ org = orig.p[k];
  float* __restrict__ dst = out.p[k];
  const int64_t N = g.count();
  const int len = g.n[axis];
  const int64_t stride = axis == 0 ? 1 : (axis == 1 ? g.n[0] : static_cast<int64_t>(g.n[0]) * g.n[1]);
  float nrm = 0.f;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < N;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    int x, y, z;
    unlin(g, i, x, y, z);
    const int pos = axis == 0 ? x : (axis == 1 ? y : z);
    const int64_t base = i - static_cast<int64_t>(pos) * stride;
    float acc[D];
#pragma unroll
    for (int c = 0; c < D; ++c) acc[c] = 0.f;
    for (int kk = -t.r; kk <= t.r; ++kk) {
      int p = pos + kk;
      p = p < 0 ? 0 : (p >= len ? len - 1 : p);
      const float w = t.c[kk < 0 ? -kk : kk];
      const float* s = src + (base + p * stride) * D;
#pragma unroll
      for (int c = 0; c < D; ++c) acc[c] += w * s[c];
    }
    const bool bnd = boundary && on_boundary(g, x, y, z);
    float s2 = 0.f;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      const float v = bnd ? 0.f : (w2 != 0.f ? acc[c] * w1 + org[i * D + c] * w2 : acc[c] * w1);
      dst[i * D + c] = v;
      const float q = v / g.sp[c];
      s2 += q * q;
    }
    nrm = fmaxf(nrm, sqrtf(s2));
  }
  double dummy = 0.0;
  block_sum_max<kBlock>(dummy, nrm);
  if (threadIdx.x == 0) atomicMax(&maxbits[k], __float_as_uint(nrm));
}

// Fused 3-D separable Gaussian of 3-vector fields (SyN field smoothing).
// Block = TX x TY output columns over a z-chunk of one field (blockIdx.z =
// field * nzc + chunk).  Each plane (edges replicated, ZeroFluxNeumann) is
// staged once in LDS, convolved along x (four outputs per task from aligned
// 16-byte LDS reads, so each input is read once per four outputs) and along
// y from LDS, and pushed into a (2R+1)-deep register ring that forms the z
// pass.  The epilogue zeroes the boundary, blends with the input when
// w2 != 0, and takes the max norm.  The next planes are prefetched into
// registers while the current one is used.
// f(integral_constant<I>, zz++) for I = I0 .. W-1 while zz < zend
template <int I, int W, class F>
__device__ __forceinline__ void unrolled_steps(F& f, int& zz, int zend) {
  if constexpr (I < W) {
    if (zz >= zend) return;
    f(std::integral_constant<int, I>{}, zz);
    ++zz;
    unrolled_steps<I + 1, W>(f, zz, zend);
  }
}

#ifndef ANTS_SMOOTH_PREFETCH
#define ANTS_SMOOTH_PREFETCH 1
#endif
#ifndef ANTS_SMOOTH_ROTATE
#define ANTS_SMOOTH_ROTATE 1
#endif
constexpr bool kSmoothRotate = ANTS_SMOOTH_ROTATE != 0;
// max norm of u / spacing as a product with 1 / spacing (at most an ulp
// from the quotient) instead of an IEEE division per component
#ifndef ANTS_SMOOTH_NORM_MUL
#define ANTS_SMOOTH_NORM_MUL 1
#endif

constexpr int kSmoothPrefetch = ANTS_SMOOTH_PREFETCH;
// Tile of the fused smoothing: 16 x 32 outputs per plane measured fastest
// over the three pyramid levels of a 192x224x192 registration.
#ifndef ANTS_SMOOTH_TX
#define ANTS_SMOOTH_TX 16
#endif
#ifndef ANTS_SMOOTH_TY
#define ANTS_SMOOTH_TY 32
#endif
constexpr int kSmoothTX = ANTS_SMOOTH_TX, kSmoothTY = ANTS_SMOOTH_TY;
// Occupancy floor of the fused smoothing (waves per SIMD): 4 keeps two
// 512-thread blocks per CU, which caps the kernel at 128 VGPRs.
#ifndef ANTS_SMOOTH_WAVES
#define ANTS_SMOOTH_WAVES 4
#endif
constexpr int kSmoothWavesPerEU = ANTS_SMOOTH_WAVES;
// y pass: 0 = each thread convolves its own column (2R+1 LDS reads per
// output and component); S > 0 = a task convolves S rows of one column and
// component from S + 2R reads into a staging buffer (same arithmetic, same
// bits).
#ifndef ANTS_SMOOTH_YSEG
#define ANTS_SMOOTH_YSEG 0
#endif
constexpr int kSmoothYSeg = ANTS_SMOOTH_YSEG;
// NC interleaved components (3: a displacement field, 1: an image); with
// maxbits null no norm is taken.
template <int R, int TX, int TY, int NC = 3>
__global__ void __launch_bounds__(TX* TY, kSmoothWavesPerEU) k_smooth3d(GridDev g, PtrPack4 in, PtrPack4 orig, MutPack4 out, Taps t,
                                                     int zchunk, int nzc, float w1, float w2, int boundary,
                                                     unsigned int* maxbits) {
  constexpr int W = 2 * R + 1;
  constexpr int LX = TX + 2 * R, LY = TY + 2 * R;
  constexpr int LXP = (LX + 3) / 4 * 4;  // row pitch: whole float4 chunks
  constexpr int NT = TX * TY;
  constexpr int NE = NC * LX * LY;
  constexpr int KP = (NE + NT - 1) / NT;
  __shared__ __align__(16) float tile[NC][LY][LXP];
  __shared__ __align__(16) float xs[NC][LY][TX];
  __shared__ float ysg[kSmoothYSeg > 0 ? NC : 1][kSmoothYSeg > 0 ? TY : 1][TX];
  int btx, bty, btz;
  tile_of(ANTS_SMOOTH_XCD_MAP != 0, btx, bty, btz);
  const int f = btz / nzc, zb = btz - f * nzc;
  const float* __restrict__ src = in.p[f];
  const float* __restrict__ org = orig.p[f];
  float* __restrict__ dst = out.p[f];
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const int x0 = btx * TX, y0 = bty * TY, z0 = zb * zchunk;
  const int z1 = min(z0 + zchunk, nz);
  const int x = x0 + tx, y = y0 + ty;
  const bool active = x < nx && y < ny;
  const int64_t plane = static_cast<int64_t>(nx) * ny;
  float c[R + 1];
#pragma unroll
  for (int b = 0; b <= R; ++b) c[b] = b <= t.r ? t.c[b] : 0.f;
  float ring[NC][W];
#pragma unroll
  for (int a = 0; a < NC; ++a)
#pragma unroll
    for (int b = 0; b < W; ++b) ring[a][b] = 0.f;
  // planes are fetched kSmoothPrefetch ahead into registers: slot
  // (zz - (z0 - R)) % P holds plane zz until it is stored to LDS, and is
  // then refilled with plane zz + P
  constexpr int P = kSmoothPrefetch;
  float pv[P][KP];
  // Loop-invariant part of each staged element's addresses: the in-plane
  // source offset (edges replicated) and the LDS slot; -1 marks the unused
  // tail of the last round.
  int goff[KP], loff[KP];
#pragma unroll
  for (int u = 0; u < KP; ++u) {
    const int e = threadIdx.x + u * NT;
    goff[u] = -1;
    loff[u] = 0;
    if (e < NE) {
      const int ly = e / (NC * LX), r = e - ly * (NC * LX);
      const int lx = r / NC, cc = r - lx * NC;
      int gx = x0 - R + lx, gy = y0 - R + ly;
      gx = gx < 0 ? 0 : (gx >= nx ? nx - 1 : gx);
      gy = gy < 0 ? 0 : (gy >= ny ? ny - 1 : gy);
      goff[u] = (gy * nx + gx) * NC + cc;
      loff[u] = (cc * LY + ly) * LXP + lx;
    }
  }
  float* tile_flat = &tile[0][0][0];
  auto fetch = [&](float* dstv, int zq) {
    const int zs = zq < 0 ? 0 : (zq >= nz ? nz - 1 : zq);
    const float* __restrict__ srcz = src + static_cast<int64_t>(zs) * plane * NC;
#pragma unroll
    for (int u = 0; u < KP; ++u) dstv[u] = goff[u] >= 0 ? srcz[goff[u]] : 0.f;
  };
  const int zend = z1 + R;
#pragma unroll
  for (int qs = 0; qs < P; ++qs)
    if (z0 - R + qs < zend) fetch(pv[qs], z0 - R + qs);
  float nrm = 0.f;
  // one plane step: stage plane zz (prefetched in pv[qs]), convolve it along
  // x and y, push it into the z ring and emit output plane zz - R.  With
  // kSmoothRotate the z loop is unrolled W times and plane s lives in ring
  // slot PH = s % W (compile-time indices); otherwise the ring shifts.
  auto step = [&](auto PH, int qs, int zz) {
    constexpr int ph = decltype(PH)::value;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < KP; ++u)
      if (goff[u] >= 0) tile_flat[loff[u]] = pv[qs][u];
    __syncthreads();
    if (zz + P < zend) fetch(pv[qs], zz + P);
    // x pass, four outputs per task from aligned 16-byte LDS reads: task
    // (ly, q) reads tile[a][ly][4q .. 4q+15] and writes xs[a][ly][4q .. 4q+3];
    // the taps of output e use w[e .. e + 2R], all inside the staged row
    // (e + 2R <= TX - 1 + 2R < LX), so unstaged pitch padding is never used
    static_assert(TX % 4 == 0 && R >= 1 && R <= 6, "x tiling: one output quad reads 16 values");
    for (int task = threadIdx.x; task < LY * (TX / 4); task += NT) {
      const int ly = task / (TX / 4), q = task - ly * (TX / 4);
#pragma unroll
      for (int a = 0; a < NC; ++a) {
        float w[16];
#pragma unroll
        for (int h = 0; h < 4; ++h) {
          const float4 v4 = (4 * q + 4 * h < LXP) ? *reinterpret_cast<const float4*>(&tile[a][ly][4 * q + 4 * h])
                                                   : make_float4(0.f, 0.f, 0.f, 0.f);
          w[4 * h + 0] = v4.x;
          w[4 * h + 1] = v4.y;
          w[4 * h + 2] = v4.z;
          w[4 * h + 3] = v4.w;
        }
        float o[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          float acc = c[0] * w[e + R];
#pragma unroll
          for (int b = 1; b <= R; ++b) acc += c[b] * (w[e + R - b] + w[e + R + b]);
          o[e] = acc;
        }
        *reinterpret_cast<float4*>(&xs[a][ly][4 * q]) = make_float4(o[0], o[1], o[2], o[3]);
      }
    }
    __syncthreads();
    float p[NC];
    if constexpr (kSmoothYSeg > 0) {
      constexpr int S = kSmoothYSeg;
      static_assert(TY % S == 0, "y segments tile the block's rows");
      constexpr int NTASK = NC * TX * (TY / S);
      for (int task = threadIdx.x; task < NTASK; task += NT) {
        const int a = task / (TX * (TY / S));
        const int r = task - a * (TX * (TY / S));
        const int k = r / TX, cx = r - k * TX;
        const int yr = k * S;
        float w[S + 2 * R];
#pragma unroll
        for (int h = 0; h < S + 2 * R; ++h) w[h] = xs[a][yr + h][cx];
#pragma unroll
        for (int e = 0; e < S; ++e) {
          float acc = c[0] * w[e + R];
#pragma unroll
          for (int b = 1; b <= R; ++b) acc += c[b] * (w[e + R - b] + w[e + R + b]);
          ysg[a][yr + e][cx] = acc;
        }
      }
      __syncthreads();
#pragma unroll
      for (int a = 0; a < NC; ++a) p[a] = ysg[a][ty][tx];
    } else {
#pragma unroll
      for (int a = 0; a < NC; ++a) {
        float acc = c[0] * xs[a][ty + R][tx];
#pragma unroll
        for (int b = 1; b <= R; ++b) acc += c[b] * (xs[a][ty + R - b][tx] + xs[a][ty + R + b][tx]);
        p[a] = acc;
      }
    }
    // ring slot of plane zz - R + k, k = -R .. R (k = R is plane zz)
    auto slot = [](int k) { return kSmoothRotate ? (ph + W + k - R) % W : R + k; };
#pragma unroll
    for (int a = 0; a < NC; ++a) {
      if constexpr (kSmoothRotate) {
        ring[a][ph] = p[a];
      } else {
#pragma unroll
        for (int b = 0; b < W - 1; ++b) ring[a][b] = ring[a][b + 1];
        ring[a][W - 1] = p[a];
      }
    }
    const int zc = zz - R;
    if (active && zc >= z0 && zc < z1) {
      const int64_t i = static_cast<int64_t>(zc) * plane + static_cast<int64_t>(y) * nx + x;
      const bool bnd = boundary && (x == 0 || x == nx - 1 || y == 0 || y == ny - 1 || zc == 0 || zc == nz - 1);
      float s2 = 0.f;
#pragma unroll
      for (int a = 0; a < NC; ++a) {
        float acc = c[0] * ring[a][slot(0)];
#pragma unroll
        for (int b = 1; b <= R; ++b) acc += c[b] * (ring[a][slot(-b)] + ring[a][slot(b)]);
        const float v = bnd ? 0.f : (w2 != 0.f ? acc * w1 + org[i * NC + a] * w2 : acc * w1);
        dst[i * NC + a] = v;
#if ANTS_SMOOTH_NORM_MUL
        const float q = v * g.isp[a];
#else
        const float q = v / g.sp[a];
#endif
        s2 += q * q;
      }
      nrm = fmaxf(nrm, sqrtf(s2));
    }
  };
  if constexpr (kSmoothRotate) {
    static_assert(!kSmoothRotate || P == 1, "the rotating ring prefetches one plane ahead");
    auto step1 = [&](auto PH, int zz) { step(PH, 0, zz); };
    for (int zz = z0 - R; zz < zend;) unrolled_steps<0, W>(step1, zz, zend);
  } else {
    for (int zb0 = z0 - R; zb0 < zend; zb0 += P) {
#pragma unroll
      for (int qs = 0; qs < P; ++qs) {
        const int zz = zb0 + qs;
        if (zz >= zend) break;
        step(std::integral_constant<int, 0>{}, qs, zz);
      }
    }
  }
  if (maxbits) {
    double dummy = 0.0;
    block_sum_max<NT>(dummy, nrm);
    if (threadIdx.x == 0) atomicMax(&maxbits[f], __float_as_uint(nrm));
  }
}

// ---------------------------------------------------------------------------
// Block 0 also zeroes `clear` (nclear words; the iteration's statistics,
// which only later kernels of the iteration write).
template <int D, class T, bool MAP>
__global__ void __launch_bounds__(kBlock) k_warp_pair(GridDev gl, GridDev gF, const T* __restrict__ Fs,
                                                      const float* __restrict__ uF, float* __restrict__ Fw,
                                                      GridDev gM, const T* __restrict__ Ms,
                                                      const float* __restrict__ uM, float* __restrict__ Mw,
                                                      uint32_t* __restrict__ clear, int nclear, WarpMap mF,
                                                      WarpMap mM) {
  if (blockIdx.x == 0)
    for (int w = threadIdx.x; w < nclear; w += blockDim.x) clear[w] = 0u;
  const int64_t N = gl.count();
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= N) return;
  int x, y, z;
  unlin(gl, i, x, y, z);
  float u[3], c[3], v;
#pragma unroll
  for (int d = 0; d < D; ++d) u[d] = uF[i * D + d];
  if constexpr (MAP)
    map_cidx<D>(mF, x, y, z, u, c);
  else
    cidx_of<D>(gl, x, y, z, u, gF, c);
  v = 0.f;
  if (inside<D>(gF, c)) v = interp_img<D, T>(gF, Fs, c);
  Fw[i] = v;
#pragma unroll
  for (int d = 0; d < D; ++d) u[d] = uM[i * D + d];
  if constexpr (MAP)
    map_cidx<D>(mM, x, y, z, u, c);
  else
    cidx_of<D>(gl, x, y, z, u, gM, c);
  v = 0.f;
  if (inside<D>(gM, c)) v = interp_img<D, T>(gM, Ms, c);
  Mw[i] = v;
}

__global__ void __launch_bounds__(kBlock) k_to_bf16(const float* __restrict__ in, uint16_t* __restrict__ out,
                                                    int64_t n) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    out[i] = f2bf(in[i]);
}

// Plain 16-byte-vector copy: reads n4*16 bytes and writes as many, once each.
__global__ void __launch_bounds__(kBlock) k_copy_calibration(const float4* __restrict__ in, float4* __restrict__ out,
                                                             int64_t n4) {
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n4;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    out[i] = in[i];
}

template <int D>
__global__ void __launch_bounds__(kBlock) k_warp_image(GridDev gi, const float* __restrict__ img, GridDev go,
                                                       const float* __restrict__ u, float* __restrict__ out) {
  const int64_t N = go.count();
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= N) return;
  int x, y, z;
  unlin(go, i, x, y, z);
  float uu[3] = {0.f, 0.f, 0.f}, c[3], v = 0.f;
  if (u) {
#pragma unroll
    for (int d = 0; d < D; ++d) uu[d] = u[i * D + d];
  }
  cidx_of<D>(go, x, y, z, uu, gi, c);
  if (inside<D>(gi, c)) interp<D, 1>(gi, img, c, &v);
  out[i] = v;
}

// ---------------------------------------------------------------------------
// CentralDifferenceImageFunction component on grid values (half-voxel linear
// samples): 0 at the last index, a half difference at index 0.
// CentralDifferenceImageFunction with UseImageDirection: a derivative taken
// along the index axes becomes the physical vector dir * v.
template <int D>
__device__ __forceinline__ void rotate_to_physical(const GridDev& g, float v[3]) {
  float r[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int k = 0; k < D; ++k) r[a] += g.dir[a * 3 + k] * v[k];
#pragma unroll
  for (int a = 0; a < D; ++a) v[a] = r[a];
}

// ih = 1 / spacing (GridDev::isp): a product instead of an IEEE division
// (ten instructions), at most an ulp from (up - dn) / h.
#ifndef ANTS_CDIFF_MUL
#define ANTS_CDIFF_MUL 1
#endif
__device__ __forceinline__ float cdiff(float vm, float v0, float vp, int pos, int len, float h, float ih) {
  if (pos >= len - 1) return 0.f;
  const float up = v0 + (vp - v0) * 0.5f;
  const float dn = pos == 0 ? v0 : vm + (v0 - vm) * 0.5f;
#if ANTS_CDIFF_MUL
  (void)h;
  return (up - dn) * ih;
#else
  (void)ih;
  return (up - dn) / h;
#endif
}

// Per-voxel CC value and both SyN derivatives from window sums.  The sums
// and this formula are evaluated in double: products of float samples are
// exact in double, so sFF = sum f^2 - (sum f)^2 / n keeps its digits where a
// float evaluation cancels (flat, noisy windows).  FP64 vector FMA runs at
// the non-packed FP32 rate on CDNA4.
struct CCOut {
  double cc;
  float a1, a2;  // g_f2m = a1 * grad m ; g_m2m = a2 * grad f
};
// rc = 1 / (window count)
__device__ __forceinline__ CCOut cc_voxel_rc(double rc, double sf, double sm, double sff, double smm, double sfm,
                                             float f, float m, float eps) {
  // In double the algebraically equal short forms are as accurate as ITK's
  // sff - mf*sf - mf*sf + cnt*mf*mf, and one reciprocal replaces the divisions
  // by sFF*sMM, sMM and sFF (1/sMM = sFF/p, 1/sFF = sMM/p).
  const double mf = sf * rc, mm = sm * rc;
  const double sFF = sff - sf * mf;
  const double sMM = smm - sm * mm;
  const double sFM = sfm - sf * mm;
  const double fA = f - mf, mA = m - mm;
  CCOut o;
  o.cc = 1.0;
  o.a1 = 0.f;
  o.a2 = 0.f;
  if (sFF > eps && sMM > eps) {
    const double rp = 1.0 / (sFF * sMM);
    const double q = sFM * rp;
    o.cc = sFM * q;
    const double a = 2.0 * q;
    o.a1 = static_cast<float>(a * (fA - q * sFF * mA));
    o.a2 = static_cast<float>(a * (mA - q * sMM * fA));
  }
  return o;
}

__device__ __forceinline__ CCOut cc_voxel(double cnt, double sf, double sm, double sff, double smm, double sfm,
                                          float f, float m, float eps) {
  return cc_voxel_rc(1.0 / cnt, sf, sm, sff, smm, sfm, f, m, eps);
}

// Direct CC (any D): each thread sums its (2r+1)^D window from global memory.
// EXT: the moving gradient comes from gext (physical, at each voxel's mapped
// point: the v4 metric of a field transform); only g1 = a1 * gext is written.
template <int D, bool EXT = false>
__global__ void __launch_bounds__(kBlock) k_cc_direct(GridDev g, const float* __restrict__ F,
                                                      const float* __restrict__ M, int R, float eps,
                                                      float* __restrict__ g1, float* __restrict__ g2,
                                                      Partial* partials, const float* __restrict__ gext) {
  const int64_t N = g.count();
  double cc_sum = 0.0;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < N;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    int x, y, z;
    unlin(g, i, x, y, z);
    const int zlo = D == 3 ? max(z - R, 0) : 0, zhi = D == 3 ? min(z + R, g.n[2] - 1) : 0;
    const int ylo = max(y - R, 0), yhi = min(y + R, g.n[1] - 1);
    const int xlo = max(x - R, 0), xhi = min(x + R, g.n[0] - 1);
    double sf = 0.0, sm = 0.0, sff = 0.0, smm = 0.0, sfm = 0.0;
    for (int zz = zlo; zz <= zhi; ++zz)
      for (int yy = ylo; yy <= yhi; ++yy)
        for (int xx = xlo; xx <= xhi; ++xx) {
          const int64_t j = lin(g, xx, yy, zz);
          const double f = F[j], m = M[j];
          sf += f;
          sm += m;
          sff += f * f;
          smm += m * m;
          sfm += f * m;
        }
    const double cnt = static_cast<double>((zhi - zlo + 1) * (yhi - ylo + 1) * (xhi - xlo + 1));
    const float f = F[i], m = M[i];
    const CCOut o = cc_voxel(cnt, sf, sm, sff, smm, sfm, f, m, eps);
    cc_sum += o.cc;
    if constexpr (EXT) {
      if (g1) {
#pragma unroll
        for (int d = 0; d < D; ++d) g1[i * D + d] = o.a1 * gext[i * D + d];
      }
      continue;
    }
    const int pos[3] = {x, y, z};
    const int64_t st[3] = {1, g.n[0], static_cast<int64_t>(g.n[0]) * g.n[1]};
    float vm[3] = {0.f, 0.f, 0.f}, vf[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < D; ++d) {
      float gm = 0.f, gf = 0.f;
      if (o.a1 != 0.f || o.a2 != 0.f) {
        const bool lo = pos[d] > 0, hi = pos[d] < g.n[d] - 1;
        const float mm_ = lo ? M[i - st[d]] : m, mp = hi ? M[i + st[d]] : m;
        const float fm_ = lo ? F[i - st[d]] : f, fp = hi ? F[i + st[d]] : f;
        gm = cdiff(mm_, m, mp, pos[d], g.n[d], g.sp[d], g.isp[d]);
        gf = cdiff(fm_, f, fp, pos[d], g.n[d], g.sp[d], g.isp[d]);
      }
      vm[d] = o.a1 * gm;
      vf[d] = o.a2 * gf;
    }
    if (g.oblique) {
      rotate_to_physical<D>(g, vm);
      rotate_to_physical<D>(g, vf);
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
      g1[i * D + d] = vm[d];
      g2[i * D + d] = vf[d];
    }
  }
  block_partial<kBlock>(cc_sum, 0.f, partials + blockIdx.x);
}

// 3-D CC with LDS-staged planes and a register ring along z.
// Block = TX x TY output columns over a z-chunk.  Each step loads plane zz
// of Fw and Mw (with an R halo in x and y) into LDS, forms the (2R+1)^2
// in-plane box sums of f, m, f^2, m^2, f*m separably (x rows, then y) in
// double, and pushes them into a (2R+1)-deep register ring; the ring total
// is the window sum for the voxel R planes behind.  Outside-volume samples
// are 0 and the count is analytic, which equals ITK's in-bounds-only sums.
#ifndef ANTS_CC_ROTATE
#define ANTS_CC_ROTATE 1
#endif
constexpr bool kCCRotate = ANTS_CC_ROTATE != 0;
#ifndef ANTS_CC_WAVES
#define ANTS_CC_WAVES 1
#endif
constexpr int kCCWavesPerEU = ANTS_CC_WAVES;  // occupancy floor (waves per SIMD)
#ifndef ANTS_CC_WAVES_FLOAT
#define ANTS_CC_WAVES_FLOAT 1
#endif
constexpr int kCCWavesPerEUFloat = ANTS_CC_WAVES_FLOAT;  // the same for float box sums
template <class ST>
constexpr int cc_waves() {
  return std::is_same<ST, float>::value ? kCCWavesPerEUFloat : kCCWavesPerEU;
}
// y box sums: 0 = each thread sums its W rows (5 W LDS reads per voxel);
// S > 0 = a task walks S output rows of one column and quantity with a
// running sum (W + 2 (S - 1) reads per S outputs) into a staging buffer.
#ifndef ANTS_CC_YSEG
#define ANTS_CC_YSEG 8
#endif
constexpr int kCCYSeg = ANTS_CC_YSEG;
// x box sums: outputs per task (running sum along the row).
#ifndef ANTS_CC_XSEG
#define ANTS_CC_XSEG 2
#endif
constexpr int kCCXSeg = ANTS_CC_XSEG;
//
// ST is the type of the box sums.  double: products of float samples are
// exact and every sum keeps its digits.  float: the samples are stored
// shifted by a block constant (the block's centre sample of F and of M), so
// that sFF = sum f'^2 - (sum f')^2 / n, which is invariant to the shift,
// does not cancel in flat windows near that value; the z window sums are
// re-added from the ring in a fixed order every plane (no running sum to
// drift), and the per-voxel formula is evaluated in double.  Out-of-volume
// samples stay 0, so the shifted sums are the in-bounds sums of f - c.
// EXT: the moving gradient is read from gext (see k_cc_direct); the
// in-plane and z gradients are not formed and only g1 is written.
template <int R, int TX, int TY, class ST = double, bool EXT = false>
__global__ void __launch_bounds__(TX* TY, cc_waves<ST>()) k_cc3d(GridDev g, const float* __restrict__ F, const float* __restrict__ M,
                                                 float eps, int zchunk, float* __restrict__ g1,
                                                 float* __restrict__ g2, Partial* partials,
                                                 const float* __restrict__ gext) {
  constexpr int W = 2 * R + 1;
  constexpr int LX = TX + 2 * R, LY = TY + 2 * R;
  constexpr int NT = TX * TY;
  constexpr bool kShift = std::is_same<ST, float>::value;
  __shared__ float q[2][LY][LX];
  __shared__ ST sx[5][LY][TX];
  constexpr int SYR = kCCYSeg > 0 ? TY : 1;
  __shared__ ST sy[kCCYSeg > 0 ? 5 : 1][SYR][TX];
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  int btx, bty, btz;
  tile_of(ANTS_CC_XCD_MAP != 0, btx, bty, btz);
  const int x0 = btx * TX, y0 = bty * TY, z0 = btz * zchunk;
  const int z1 = min(z0 + zchunk, nz);
  const int x = x0 + tx, y = y0 + ty;
  const bool active = x < nx && y < ny;
  const double cxy = static_cast<double>((min(x + R, nx - 1) - max(x - R, 0) + 1) *
                                         (min(y + R, ny - 1) - max(y - R, 0) + 1));
  const double rc_in = 1.0 / (cxy * static_cast<double>(W));

  // z rings: double plane sums of the last W planes, and the raw centre
  // values and in-plane gradients of the planes still needed.  With
  // kCCRotate the z loop is unrolled W times and plane s lives in slot
  // s % W (compile-time indices, no register moves); otherwise the rings
  // shift by one slot per plane.
  constexpr int NRF = kCCRotate ? W : R + 2, NRG = kCCRotate ? W : R + 1;
  static_assert(kCCRotate || !kShift, "float sums need the rotating ring");
  ST ring[5][W];
  ST zsum[5] = {ST(0), ST(0), ST(0), ST(0), ST(0)};
#pragma unroll
  for (int a = 0; a < 5; ++a)
#pragma unroll
    for (int b = 0; b < W; ++b) ring[a][b] = ST(0);
  float rf[NRF], rm[NRF];          // raw centre values
  float rgf[NRG][2], rgm[NRG][2];  // in-plane gradients
#pragma unroll
  for (int b = 0; b < NRF; ++b) rf[b] = rm[b] = 0.f;
#pragma unroll
  for (int b = 0; b < NRG; ++b) rgf[b][0] = rgf[b][1] = rgm[b][0] = rgm[b][1] = 0.f;
  double cc_sum = 0.0;
  const int64_t plane = static_cast<int64_t>(nx) * ny;
  // block shift (float sums): the centre sample of the block's columns in
  // the middle plane of its chunk, the same address for every thread
  float cF = 0.f, cM = 0.f;
  if constexpr (kShift) {
    const int xc = min(x0 + TX / 2, nx - 1), yc = min(y0 + TY / 2, ny - 1);
    const int zc0 = min(z0 + (z1 - z0) / 2, nz - 1);
    const int64_t ic = static_cast<int64_t>(zc0) * plane + static_cast<int64_t>(yc) * nx + xc;
    cF = F[ic];
    cM = M[ic];
  }

  // Plane samples are prefetched into registers one plane ahead, so their
  // global-load latency overlaps the box sums of the current plane.
  constexpr int KP = (LX * LY + NT - 1) / NT;
  float pf[KP], pm[KP];
  // loop-invariant in-plane source offsets (-1: outside the image, reads 0)
  // and LDS slots (-1: unused tail of the last round)
  int goff[KP], loff[KP];
#pragma unroll
  for (int u = 0; u < KP; ++u) {
    const int k = threadIdx.x + u * NT;
    const int ly = k / LX, lx = k - ly * LX;
    const int gx = x0 - R + lx, gy = y0 - R + ly;
    loff[u] = k < LX * LY ? k : -1;
    goff[u] = (k < LX * LY && gx >= 0 && gx < nx && gy >= 0 && gy < ny) ? gy * nx + gx : -1;
  }
  auto fetch = [&](int zq) {
    const bool zin_q = zq >= 0 && zq < nz;
    const int64_t zo = static_cast<int64_t>(zq) * plane;
#pragma unroll
    for (int u = 0; u < KP; ++u) {
      float f = 0.f, m = 0.f;
      if (zin_q && goff[u] >= 0) {
        f = F[zo + goff[u]] - cF;
        m = M[zo + goff[u]] - cM;
      }
      pf[u] = f;
      pm[u] = m;
    }
  };
  fetch(z0 - R);
  const int zend = z1 + R;

  // one plane step; PH = (zz - (z0 - R)) % W when rotating
  auto step = [&](auto PH, int zz) {
    constexpr int ph = decltype(PH)::value;
    const bool zin = zz >= 0 && zz < nz;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < KP; ++u) {
      if (loff[u] >= 0) {
        (&q[0][0][0])[loff[u]] = pf[u];
        (&q[1][0][0])[loff[u]] = pm[u];
      }
    }
    __syncthreads();
    if (zz + 1 < zend) fetch(zz + 1);
    // x box sums, XS adjacent outputs per task: the window of output
    // xs + e is that of xs + e - 1 with its first sample dropped and one added
    constexpr int XS = kCCXSeg;
    static_assert(TX % XS == 0 && XS >= 2, "x sums run in segments");
    for (int task = threadIdx.x; task < LY * (TX / XS); task += NT) {
      const int ly = task / (TX / XS), xs = XS * (task - ly * (TX / XS));
      ST s0 = ST(0), s1 = ST(0), s2 = ST(0), s3 = ST(0), s4 = ST(0);
#pragma unroll
      for (int b = 0; b < W; ++b) {
        const ST f = q[0][ly][xs + b], m = q[1][ly][xs + b];
        s0 += f;
        s1 += m;
        s2 += f * f;
        s3 += m * m;
        s4 += f * m;
      }
      sx[0][ly][xs] = s0;
      sx[1][ly][xs] = s1;
      sx[2][ly][xs] = s2;
      sx[3][ly][xs] = s3;
      sx[4][ly][xs] = s4;
#pragma unroll
      for (int e = 1; e < XS; ++e) {
        const ST fo = q[0][ly][xs + e - 1], mo = q[1][ly][xs + e - 1];
        const ST fn = q[0][ly][xs + e - 1 + W], mn = q[1][ly][xs + e - 1 + W];
        s0 += fn - fo;
        s1 += mn - mo;
        s2 += fn * fn - fo * fo;
        s3 += mn * mn - mo * mo;
        s4 += fn * mn - fo * mo;
        sx[0][ly][xs + e] = s0;
        sx[1][ly][xs + e] = s1;
        sx[2][ly][xs + e] = s2;
        sx[3][ly][xs + e] = s3;
        sx[4][ly][xs + e] = s4;
      }
    }
    // centre values and in-plane gradients of plane zz at (x, y)
    const float fc = q[0][ty + R][tx + R], mc = q[1][ty + R][tx + R];
    float gfx = 0.f, gfy = 0.f, gmx = 0.f, gmy = 0.f;
    if (!EXT && active && zin) {
      gfx = cdiff(q[0][ty + R][tx + R - 1], fc, q[0][ty + R][tx + R + 1], x, nx, g.sp[0], g.isp[0]);
      gmx = cdiff(q[1][ty + R][tx + R - 1], mc, q[1][ty + R][tx + R + 1], x, nx, g.sp[0], g.isp[0]);
      gfy = cdiff(q[0][ty + R - 1][tx + R], fc, q[0][ty + R + 1][tx + R], y, ny, g.sp[1], g.isp[1]);
      gmy = cdiff(q[1][ty + R - 1][tx + R], mc, q[1][ty + R + 1][tx + R], y, ny, g.sp[1], g.isp[1]);
    }
    __syncthreads();
    ST p[5];
    if constexpr (kCCYSeg > 0) {
      constexpr int S = kCCYSeg;
      static_assert(TY % S == 0, "y segments tile the block's rows");
      constexpr int NTASK = 5 * TX * (TY / S);
      for (int task = threadIdx.x; task < NTASK; task += NT) {
        const int a = task / (TX * (TY / S));
        const int r = task - a * (TX * (TY / S));
        const int k = r / TX, cx = r - k * TX;
        const int yr = k * S;
        ST acc = ST(0);
#pragma unroll
        for (int b = 0; b < W; ++b) acc += sx[a][yr + b][cx];
        sy[a][yr][cx] = acc;
#pragma unroll
        for (int e = 1; e < S; ++e) {
          acc += sx[a][yr + e + W - 1][cx] - sx[a][yr + e - 1][cx];
          sy[a][yr + e][cx] = acc;
        }
      }
      __syncthreads();
#pragma unroll
      for (int a = 0; a < 5; ++a) p[a] = sy[a][ty][tx];
    } else {
#pragma unroll
      for (int a = 0; a < 5; ++a) {
        ST t = ST(0);
#pragma unroll
        for (int b = 0; b < W; ++b) t += sx[a][ty + b][tx];
        p[a] = t;
      }
    }
    // running z-window sums: add the new plane, drop the one W planes back;
    // slots of plane zc = zz - R and of its z neighbours
    int c0, cm, cp, cg;
    if constexpr (kCCRotate) {
#pragma unroll
      for (int a = 0; a < 5; ++a) {
        if constexpr (kShift) {
          // float: the window total re-added from the slots in slot order
          ring[a][ph] = p[a];
          ST t = ST(0);
#pragma unroll
          for (int b = 0; b < W; ++b) t += ring[a][b];
          zsum[a] = t;
        } else {
          zsum[a] += p[a] - ring[a][ph];
          ring[a][ph] = p[a];
        }
      }
      rf[ph] = fc;
      rm[ph] = mc;
      rgf[ph][0] = gfx;
      rgf[ph][1] = gfy;
      rgm[ph][0] = gmx;
      rgm[ph][1] = gmy;
      c0 = (ph + W - R) % W;
      cm = (ph + W - R - 1) % W;
      cp = (ph + W - R + 1) % W;
      cg = c0;
    } else {
#pragma unroll
      for (int a = 0; a < 5; ++a) {
        zsum[a] += p[a] - ring[a][0];
#pragma unroll
        for (int b = 0; b < W - 1; ++b) ring[a][b] = ring[a][b + 1];
        ring[a][W - 1] = p[a];
      }
#pragma unroll
      for (int b = 0; b < R + 1; ++b) {
        rf[b] = rf[b + 1];
        rm[b] = rm[b + 1];
      }
      rf[R + 1] = fc;
      rm[R + 1] = mc;
#pragma unroll
      for (int b = 0; b < R; ++b) {
        rgf[b][0] = rgf[b + 1][0];
        rgf[b][1] = rgf[b + 1][1];
        rgm[b][0] = rgm[b + 1][0];
        rgm[b][1] = rgm[b + 1][1];
      }
      rgf[R][0] = gfx;
      rgf[R][1] = gfy;
      rgm[R][0] = gmx;
      rgm[R][1] = gmy;
      // rf[R + 1] is plane zz = zc + R ... rf[1] = zc, rf[0] = zc - 1
      c0 = 1;
      cm = 0;
      cp = 2;
      cg = 0;
    }

    const int zc = zz - R;
    if (active && zc >= z0 && zc < z1) {
      const double cz = static_cast<double>(min(zc + R, nz - 1) - max(zc - R, 0) + 1);
      const float f = rf[c0], m = rm[c0];
      // 1 / count: the thread's interior value away from the z faces (the
      // branch is uniform over the block's plane)
      double rc;
      if (cz == static_cast<double>(W))
        rc = rc_in;
      else
        rc = 1.0 / (cxy * cz);
      const CCOut o = cc_voxel_rc(rc, static_cast<double>(zsum[0]), static_cast<double>(zsum[1]),
                                  static_cast<double>(zsum[2]), static_cast<double>(zsum[3]),
                                  static_cast<double>(zsum[4]), f, m, eps);
      cc_sum += o.cc;
      const int64_t i = static_cast<int64_t>(zc) * plane + static_cast<int64_t>(y) * nx + x;
      if constexpr (EXT) {
        if (g1) {
          g1[i * 3 + 0] = o.a1 * gext[i * 3 + 0];
          g1[i * 3 + 1] = o.a1 * gext[i * 3 + 1];
          g1[i * 3 + 2] = o.a1 * gext[i * 3 + 2];
        }
        return;
      }
      float gmz = 0.f, gfz = 0.f;
      if (o.a1 != 0.f || o.a2 != 0.f) {
        gmz = cdiff(zc > 0 ? rm[cm] : m, m, rm[cp], zc, nz, g.sp[2], g.isp[2]);
        gfz = cdiff(zc > 0 ? rf[cm] : f, f, rf[cp], zc, nz, g.sp[2], g.isp[2]);
      }
      float vm[3] = {o.a1 * rgm[cg][0], o.a1 * rgm[cg][1], o.a1 * gmz};
      float vf[3] = {o.a2 * rgf[cg][0], o.a2 * rgf[cg][1], o.a2 * gfz};
      if (g.oblique) {
        rotate_to_physical<3>(g, vm);
        rotate_to_physical<3>(g, vf);
      }
      g1[i * 3 + 0] = vm[0];
      g1[i * 3 + 1] = vm[1];
      g1[i * 3 + 2] = vm[2];
      g2[i * 3 + 0] = vf[0];
      g2[i * 3 + 1] = vf[1];
      g2[i * 3 + 2] = vf[2];
    }
  };
  if constexpr (kCCRotate) {
    for (int zz = z0 - R; zz < zend;) unrolled_steps<0, W>(step, zz, zend);
  } else {
    for (int zz = z0 - R; zz < zend; ++zz) step(std::integral_constant<int, 0>{}, zz);
  }
  const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  block_partial<NT>(cc_sum, 0.f, partials + bid);
}

// ---------------------------------------------------------------------------
// out = warp + scale * disp(x + warp(x)), 0 outside the buffer of disp.
template <int D>
__global__ void __launch_bounds__(kBlock) k_compose(GridDev g, PtrPack4 disp, PtrPack4 warp, MutPack4 out,
                                                    const unsigned int* maxbits, float rate) {
  const int k = blockIdx.y;
  const float* __restrict__ d = disp.p[k];
  const float* __restrict__ w = warp.p[k];
  float* __restrict__ o = out.p[k];
  const int64_t N = g.count();
  const int64_t i = gather_block() * blockDim.x + threadIdx.x;
  if (i >= N) return;
  float scale = 1.f;
  if (maxbits) {
    const float mx = __uint_as_float(maxbits[k]);
    scale = mx > 0.f ? rate / mx : rate;
  }
  int x, y, z;
  unlin(g, i, x, y, z);
  float wv[3] = {0.f, 0.f, 0.f}, c[3], v[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int dd = 0; dd < D; ++dd) wv[dd] = w[i * D + dd];
  cidx_of<D>(g, x, y, z, wv, g, c);
  if (inside<D>(g, c)) interp<D, D>(g, d, c, v);
#pragma unroll
  for (int dd = 0; dd < D; ++dd) o[i * D + dd] = wv[dd] + scale * v[dd];
}

// Inverse-consistency residual r(x) = fwd(x) + inv(x + fwd(x)) (inv linear,
// 0 outside its buffer) over the voxels at least `margin` voxels from every
// face: per-block partials of sum |r / spacing| in double and of its max.
template <int D>
__global__ void __launch_bounds__(kBlock) k_ic_residual(GridDev g, const float* __restrict__ fwd,
                                                        const float* __restrict__ inv, int margin,
                                                        Partial* partials) {
  const int64_t N = g.count();
  double sum = 0.0;
  float mx = 0.f;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < N;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    int x, y, z;
    unlin(g, i, x, y, z);
    const int pos[3] = {x, y, z};
    bool keep = true;
#pragma unroll
    for (int d = 0; d < D; ++d) keep = keep && pos[d] >= margin && pos[d] < g.n[d] - margin;
    if (!keep) continue;
    float wv[3] = {0.f, 0.f, 0.f}, c[3], v[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < D; ++d) wv[d] = fwd[i * D + d];
    cidx_of<D>(g, x, y, z, wv, g, c);
    if (inside<D>(g, c)) interp<D, D>(g, inv, c, v);
    double q = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double r = (static_cast<double>(wv[d]) + static_cast<double>(v[d])) / g.dsp[d];
      q += r * r;
    }
    q = sqrt(q);
    sum += q;
    mx = fmaxf(mx, static_cast<float>(q));
  }
  block_partial<kBlock>(sum, mx, partials + blockIdx.x);
}

// ---------------------------------------------------------------------------
// InvertDisplacementFieldImageFilter.  Norm pass j measures
// e_j = inv_j + field(x + inv_j) and its norms n_j; update j is
//   inv_{j+1} = inv_j - eps * e_j * min(1, eps*max_j/|e_j|)
// (eps 0.75 for the first update, 0.5 after; boundary voxels zeroed); the
// loop goes on while j < 20, max_j > 0.1 and mean_j > 0.001.  In place on
// inv: each voxel reads and writes only itself (the field is gathered).
// The max test carries a 1e-5 relative guard: with a zero initial inverse the
// first residual's max norm equals the learning rate, which for SyN[0.1,..]
// sits exactly on the 0.1 tolerance; the guard resolves that tie the same way
// in float and double.  ANTS_SYN_INV_STRICT_TIE=1 selects ITK's test as
// written instead (float max norm > 0.1f), which leaves the tie to rounding.
constexpr float kInvMaxTol = 0.1f * (1.0f + 1e-5f);
constexpr float kInvMaxTolStrict = 0.1f;
template <int D, class IT = int64_t>
__device__ __forceinline__ float inv_residual(const GridDev& g, const float* __restrict__ field, int x, int y, int z,
                                              const float iv[3], float e[3]) {
  float c[3], v[3] = {0.f, 0.f, 0.f};
  cidx_of<D>(g, x, y, z, iv, g, c);
  const bool in = inside<D>(g, c);
  interp<D, D, true, IT>(g, field, c, v);
  float s2 = 0.f;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    e[d] = iv[d] + (in ? v[d] : 0.f);
    const float q = e[d] * g.isp[d];
    s2 += q * q;
  }
  return sqrtf(s2);
}

// Inversion steps run large blocks over a bounded grid (kInvBlocks per
// problem): steps that have converged cost one cheap dispatch, active steps
// still fill the device, and the closing reduction reads one partial per
// block.  Block b (renumbered by slab_block) walks the contiguous voxel
// range [b*chunk, (b+1)*chunk), chunk a multiple of 64, one voxel per thread
// and round, so the 64 lanes of a wave cover one 64-voxel word of the
// candidate masks below, in every launch.  The next round's inverse (and
// mask word) are loaded while this round gathers the field.
//
// Speculative updates.  The update after norm pass j clamps e_j at
// eps_j * max_j, a global maximum known only after the pass.  Launch j
// therefore writes inv_{j+1} = inv_j - eps_j * e_j in place for every voxel
// whose |e_j| is at most L_j = eps_j * q_j, where q_j <= max_j is the
// largest of a sample of exact |e_j| (one word in 64; computed by the
// previous launch from its own speculative values, or by the sample launch
// j = -1 before the first step).  Such voxels are unclamped for certain.
// The others ("candidates") keep inv_j and set their bit in the mask of
// step j; launch j+1 recomputes e_j for them (the field is unchanged),
// applies the clamped update with max_j, and goes on from the result.  The
// voxel-wise arithmetic is that of the plain update, so the results are the
// same bits; with no sample (L = 0) every voxel is a candidate and the
// scheme falls back to recomputing the update in the next launch.
#ifndef ANTS_INV_THREADS
#define ANTS_INV_THREADS 512
#endif
constexpr int kInvThreads = ANTS_INV_THREADS;
static_assert(kInvThreads % 64 == 0, "waves cover whole mask words");
// Grid per problem: one block per kInvThreads * ANTS_INV_VPT voxels, at
// least 256 (while there are voxels) and at most ANTS_INV_MAX_BLOCKS (see
// inv_blocks).  Every block of a step first reduces the previous step's
// per-block partials, so fewer, fuller blocks shorten that prologue.
#ifndef ANTS_INV_MAX_BLOCKS
#define ANTS_INV_MAX_BLOCKS 768
#endif
constexpr int kInvBlocks = ANTS_INV_MAX_BLOCKS;
// Blocks of the sample launch (j = -1; 0: as many as the steps): it gathers
// one word in 64, so a smaller grid spends less on dispatch (measured per
// 192x224x192 pair: 41.19 -> 41.0 ms at 256; 128 and 512 alike).  The sample
// is the same words whatever the grid.
#ifndef ANTS_INV_SAMPLE_BLOCKS
#define ANTS_INV_SAMPLE_BLOCKS 256
#endif
__host__ __device__ constexpr int64_t inv_sample_blocks(int64_t step_blocks) {
  return ANTS_INV_SAMPLE_BLOCKS > 0 && step_blocks > ANTS_INV_SAMPLE_BLOCKS ? ANTS_INV_SAMPLE_BLOCKS : step_blocks;
}
#ifndef ANTS_INV_WAVES
#define ANTS_INV_WAVES 1
#endif
// Sampled words of the lower bound q_j (one in `every`).  Wave v of a block walks
// words base + v + 8r (base a multiple of 8), so the rule rotates the
// sampled rounds through the waves instead of giving them all to one.
__device__ __forceinline__ bool inv_sampled_word(int64_t w, int64_t every) {
  return (((w >> 3) + ((w & 7) << 1)) & (every - 1)) == 0;
}
// Density of the sample launch before step 0 (a power of two): one word in
// 64 measured cheaper than 16 (a shorter launch) and than 256 (a looser
// bound, more candidates): inversion 20.1 / 19.9 / 20.2 ms per pair.
// In-launch sample density for the next step's bound: 1/8, 1/16, 1/32, 1/64
// and 1/128 of the words measured 19.9, 19.6, 19.5, 19.5 and 19.45 ms of
// inversion per 192x224x192 pair (sparser: cheaper sampling; the bound
// stays close); 1/64 keeps some margin for broader residual spreads.
#ifndef ANTS_INV_SAMPLE_EVERY
#define ANTS_INV_SAMPLE_EVERY 64
#endif
#ifndef ANTS_INV_SAMPLE0_EVERY
#define ANTS_INV_SAMPLE0_EVERY 64
#endif

template <int NT>
__device__ __forceinline__ void reduce_partials_q(const Partial* p, int n, double& s, float& m, float& q) {
  __shared__ double r_s;
  __shared__ float r_m, r_q;
  double a = 0.0;
  float b = 0.f;
  float c = 0.f;
  for (int i = threadIdx.x; i < n; i += NT) {
    a += p[i].s;
    b = fmaxf(b, p[i].m);
    c = fmaxf(c, p[i].q);
  }
  block_sum_max<NT>(a, b);
  c = wave_max(c);
  __shared__ float wq[NT / 64];
  if ((threadIdx.x & 63) == 0) wq[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int w = 0; w < NT / 64; ++w) t = fmaxf(t, wq[w]);
    r_s = a;
    r_m = b;
    r_q = t;
  }
  __syncthreads();
  s = r_s;
  m = r_m;
  q = r_q;
}

template <int NT>
__device__ __forceinline__ void block_partial_q(double s, float m, float q, Partial* out) {
  __shared__ float wq[NT / 64];
  q = wave_max(q);
  if ((threadIdx.x & 63) == 0) wq[threadIdx.x >> 6] = q;
  block_sum_max<NT>(s, m);  // its barrier also publishes wq
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int w = 0; w < NT / 64; ++w) t = fmaxf(t, wq[w]);
    out->s = s;
    out->m = m;
    out->q = t;
  }
}

// Voxel range of this block (see above); chunk is a multiple of 64.
__device__ __forceinline__ void inv_range(int64_t N, int64_t& vbeg, int64_t& vend) {
  const int64_t nb = gridDim.x;
  const int64_t vb = slab_block(blockIdx.x, nb);
  const int64_t chunk = ((N + nb - 1) / nb + 63) / 64 * 64;
  vbeg = vb * chunk;
  vend = min(N, vbeg + chunk);
}

template <int D, class IT>
__global__ void __launch_bounds__(kInvThreads, ANTS_INV_WAVES) k_inv_step(GridDev g, PtrPack4 field, MutPack4 inv,
                                                                          SlotPack4 slots, Partial* partials,
                                                                          int64_t part_stride, uint64_t* masks,
                                                                          int64_t mask_stride, int j, int speculate,
                                                                          InvGate gate, float max_tol) {
  // gated launch: a no-op while the inversions it depends on still had work
  // pending after their last launched step (the host then reruns it)
  if (gate.step >= 0) {
    for (int q = 0; q < gate.n; ++q)
      if (gate.slots[q][gate.step].computed) return;
  }
  const int k = blockIdx.y;
  RedSlot* st = slots.p[k];
  // partials and masks of step jj: two parities x np problems
  auto part = [&](int jj) { return partials + ((jj & 1) * static_cast<int64_t>(gridDim.y) + k) * part_stride; };
  auto mask = [&](int jj) { return masks + ((jj & 1) * static_cast<int64_t>(gridDim.y) + k) * mask_stride; };
  const float* __restrict__ fld = field.p[k];
  float* __restrict__ iv = inv.p[k];
  const int64_t N = g.count();
  int64_t vbeg, vend;
  inv_range(N, vbeg, vend);
  const int lane = threadIdx.x & 63;
  constexpr int64_t stride = kInvThreads;

  if (j < 0) {
    // sample launch: exact |e_0| on the sampled words
    float qmax = 0.f;
    for (int64_t i = vbeg + threadIdx.x; i < vend; i += stride) {
      if (!inv_sampled_word((i - lane) >> 6, ANTS_INV_SAMPLE0_EVERY)) continue;  // uniform per wave
      int x, y, z;
      unlin(g, static_cast<IT>(i), x, y, z);
      float cur[3], e[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) cur[d] = d < D ? iv[static_cast<IT>(i) * D + d] : 0.f;
      qmax = fmaxf(qmax, inv_residual<D, IT>(g, fld, x, y, z, cur, e));
    }
    block_partial_q<kInvThreads>(0.0, 0.f, qmax, part(j) + blockIdx.x);
    return;
  }

  bool do_fix = false, do_norm = false;
  float maxprev = 0.f, eps_prev = 0.f, qb = 0.f;
  if (j == 0) {
    do_norm = true;
    if (speculate) {
      double ds;
      float dm;
      reduce_partials_q<kInvThreads>(part(-1), static_cast<int>(inv_sample_blocks(gridDim.x)), ds, dm, qb);
    }
  } else if (st[j - 1].computed) {
    // every block reduces the previous step's partials (same order, same bits)
    double psum;
    reduce_partials_q<kInvThreads>(part(j - 1), gridDim.x, psum, maxprev, qb);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st[j - 1].sum = psum;
      st[j - 1].maxbits = __float_as_uint(maxprev);
    }
    do_fix = true;
    const double mean = psum / static_cast<double>(N);
    eps_prev = (j - 1 == 0) ? 0.75f : 0.5f;
    do_norm = j < kInvMaxIter && maxprev > max_tol && mean > 0.001;
  }
  if (!do_fix && !do_norm) return;
  const float lim = eps_prev * maxprev;      // clamp of the pending update j-1
  const float eps = (j == 0) ? 0.75f : 0.5f;  // update j
  const float L = eps * qb;                   // <= eps * max_j: unclamped for sure
  // without speculation every voxel is pending: launch j recomputes and
  // applies update j-1 everywhere, then measures e_j (no masks, no samples)
  const uint64_t* __restrict__ mprev = do_fix && speculate ? mask(j - 1) : nullptr;
  uint64_t* __restrict__ mcur = mask(j);
  // the clamped update j-1 of a candidate voxel (recomputes e_{j-1})
  auto fix = [&](int x, int y, int z, bool bnd, float cur[3]) {
    float ep[3];
    const float nrm = inv_residual<D, IT>(g, fld, x, y, z, cur, ep);
    float sc = 1.f;
    if (nrm > lim) sc = lim / nrm;
#pragma unroll
    for (int d = 0; d < D; ++d) cur[d] = bnd ? 0.f : cur[d] + (-ep[d] * sc) * eps_prev;
  };

  if (!do_norm && !speculate) {
    // final launch of a plain inversion: update j-1 everywhere
    for (int64_t i = vbeg + threadIdx.x; i < vend; i += stride) {
      int x, y, z;
      unlin(g, static_cast<IT>(i), x, y, z);
      float cur[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) cur[d] = d < D ? iv[static_cast<IT>(i) * D + d] : 0.f;
      fix(x, y, z, on_boundary(g, x, y, z), cur);
#pragma unroll
      for (int d = 0; d < D; ++d) iv[static_cast<IT>(i) * D + d] = cur[d];
    }
    return;
  }
  if (!do_norm) {
    // final launch of an inversion: only the pending candidates.  The
    // wave's mask words are loaded 64 rounds at a time (lane r holds round
    // r's word) so that the empty ones cost no load latency each.
    const int64_t wave_first = vbeg + (threadIdx.x - lane);  // first voxel of the wave's round 0
    for (int64_t r0 = 0; wave_first + r0 * stride < vend; r0 += 64) {
      const int64_t my_first = wave_first + (r0 + lane) * stride;
      const uint64_t myw = my_first < vend ? mprev[my_first >> 6] : 0ull;
      const unsigned lo = static_cast<unsigned>(myw), hi = static_cast<unsigned>(myw >> 32);
      for (int r = 0; r < 64 && wave_first + (r0 + r) * stride < vend; ++r) {
        const uint64_t w = (static_cast<uint64_t>(static_cast<unsigned>(__shfl(static_cast<int>(hi), r))) << 32) |
                           static_cast<unsigned>(__shfl(static_cast<int>(lo), r));
        if (w == 0) continue;  // uniform per wave
        const int64_t i = wave_first + (r0 + r) * stride + lane;
        if (i < vend && ((w >> lane) & 1)) {
          int x, y, z;
          unlin(g, static_cast<IT>(i), x, y, z);
          float cur[3];
#pragma unroll
          for (int d = 0; d < 3; ++d) cur[d] = d < D ? iv[static_cast<IT>(i) * D + d] : 0.f;
          fix(x, y, z, on_boundary(g, x, y, z), cur);
#pragma unroll
          for (int d = 0; d < D; ++d) iv[static_cast<IT>(i) * D + d] = cur[d];
        }
      }
    }
    return;
  }

  double nsum = 0.0;
  float nmax = 0.f, qmax = 0.f;
  // next round's inverse and previous-step mask word (indices clamped into
  // the range: loads stay in bounds, values past it are unused)
  float pcur[3];
  uint64_t pword = 0;
  auto prefetch = [&](int64_t i) {
    const IT ic = static_cast<IT>(i < vend ? i : vend - 1);
#pragma unroll
    for (int d = 0; d < 3; ++d) pcur[d] = d < D ? iv[ic * D + d] : 0.f;
    if (mprev) pword = mprev[(i - lane) >> 6];
  };
  const int64_t i_first = vbeg + threadIdx.x;
  // a wave whose first voxel is in range runs; its tail lanes idle
  if (i_first - lane < vend) prefetch(i_first);
  for (int64_t i = i_first; i - lane < vend; i += stride) {
    const bool ok = i < vend;
    const IT ic = static_cast<IT>(ok ? i : vend - 1);
    int x, y, z;
    unlin(g, ic, x, y, z);
    const bool bnd = on_boundary(g, x, y, z);
    float cur[3] = {pcur[0], pcur[1], pcur[2]};
    const uint64_t wprev = pword;
    if (i + stride - lane < vend) prefetch(i + stride);
    const bool fixme = do_fix && ok && (!speculate || ((wprev >> lane) & 1));
    if (fixme) fix(x, y, z, bnd, cur);
    float e[3];
    const float nrm = inv_residual<D, IT>(g, fld, x, y, z, cur, e);
    const bool spec = speculate && ok && nrm <= L;
    float nxt[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < D; ++d) nxt[d] = bnd ? 0.f : cur[d] + (-e[d] * 1.f) * eps;
    if (ok) {
      nsum += nrm;
      nmax = fmaxf(nmax, nrm);
      if (spec || fixme) {
#pragma unroll
        for (int d = 0; d < D; ++d) iv[ic * D + d] = spec ? nxt[d] : cur[d];
      }
    }
    if (speculate) {
      const uint64_t bits = __ballot(ok && !spec);
      if (lane == 0) mcur[(i - lane) >> 6] = bits;
    }
    // exact |e_{j+1}| of sampled speculative voxels: the next step's bound
    if (inv_sampled_word((i - lane) >> 6, ANTS_INV_SAMPLE_EVERY) && spec) {
      float e2[3];
      qmax = fmaxf(qmax, inv_residual<D, IT>(g, fld, x, y, z, nxt, e2));
    }
  }
  block_partial_q<kInvThreads>(nsum, nmax, qmax, part(j) + blockIdx.x);
  if (blockIdx.x == 0 && threadIdx.x == 0) st[j].computed = 1;
}

// ---------------------------------------------------------------------------
template <int D>
__global__ void __launch_bounds__(kBlock) k_resample_fields(GridDev gi, GridDev go, PtrPack4 in, MutPack4 out) {
  const int k = blockIdx.y;
  const int64_t N = go.count();
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= N) return;
  int x, y, z;
  unlin(go, i, x, y, z);
  float c[3], v[3] = {0.f, 0.f, 0.f};
  cidx_of<D>(go, x, y, z, nullptr, gi, c);
  if (inside<D>(gi, c)) interp<D, D>(gi, in.p[k], c, v);
#pragma unroll
  for (int d = 0; d < D; ++d) out.p[k][i * D + d] = v[d];
}

// antsApplyTransforms NearestNeighbor: coordinates in double so that the
// rounding decision (floor(c + 0.5), half up) matches a double CPU path bit for bit.
template <int D>
__global__ void __launch_bounds__(kBlock) k_labels_nn(GridDev go, const float* __restrict__ u, GridDev gl,
                                                      const int32_t* __restrict__ lab, int32_t* __restrict__ out) {
#pragma clang fp contract(off)
  const int64_t N = go.count();
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= N) return;
  int x, y, z;
  unlin(go, i, x, y, z);
  const int ix[3] = {x, y, z};
  int idx[3] = {0, 0, 0};
  bool in = true;
  if (go.oblique | gl.oblique) {
    double q[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int r = 0; r < D; ++r) {
      double p = go.dorg[r];
#pragma unroll
      for (int k = 0; k < D; ++k) p += go.ddir[r * 3 + k] * (static_cast<double>(ix[k]) * go.dsp[k]);
      if (u) p += static_cast<double>(u[i * D + r]);
      q[r] = p - gl.dorg[r];
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < D; ++k) t += gl.ddir[k * 3 + d] * q[k];
      const double c = t / gl.dsp[d];
      if (!(c >= -0.5 && c < static_cast<double>(gl.n[d]) - 0.5)) in = false;
      idx[d] = static_cast<int>(floor(c + 0.5));
    }
  } else {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double p = go.dorg[d] + static_cast<double>(ix[d]) * go.dsp[d];
      if (u) p += static_cast<double>(u[i * D + d]);
      const double c = (p - gl.dorg[d]) / gl.dsp[d];
      if (!(c >= -0.5 && c < static_cast<double>(gl.n[d]) - 0.5)) in = false;
      idx[d] = static_cast<int>(floor(c + 0.5));
    }
  }
  int32_t v = 0;
  if (in) v = lab[idx[0] + static_cast<int64_t>(gl.n[0]) * (idx[1] + static_cast<int64_t>(gl.n[1]) * idx[2])];
  out[i] = v;
}

// ---------------------------------------------------------------------------
// antsApplyTransforms subset: ResampleImageFilter through a composite of
// affine and displacement-field transforms.  Coordinates are formed in
// double with contraction off and a fixed operation order, so that a double
// CPU path makes the same nearest-neighbour and label decisions.
#pragma clang fp contract(off)
template <int D>
__device__ __forceinline__ void geom_point(const GeomDev& g, const int ix[3], double p[3]) {
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    double s = g.org[r];
#pragma unroll
    for (int c = 0; c < D; ++c) s = s + g.m[r * 3 + c] * static_cast<double>(ix[c]);
    p[r] = r < D ? s : 0.0;
  }
}
template <int D>
__device__ __forceinline__ void geom_cidx(const GeomDev& g, const double p[3], double c[3]) {
  double q[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) q[k] = k < D ? p[k] - g.org[k] : 0.0;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) s = s + g.minv[r * 3 + k] * q[k];
    c[r] = r < D ? s : 0.0;
  }
}
template <int D>
__device__ __forceinline__ bool geom_inside(const GeomDev& g, const double c[3]) {
#pragma unroll
  for (int d = 0; d < D; ++d)
    if (!(c[d] >= -0.5 && c[d] < static_cast<double>(g.n[d]) - 0.5)) return false;
  return true;
}
// Neighbours and weights of linear interpolation at an inside point, the
// upper neighbour clamped to the buffer (ITK LinearInterpolateImageFunction).
template <int D>
__device__ __forceinline__ void lin_setup(const GeomDev& g, const double c[3], int64_t lo[3], int64_t hi[3],
                                          double w[3]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (d < D) {
      const double f = floor(c[d]);
      int64_t b = static_cast<int64_t>(f);
      double t = c[d] - f;
      if (b < 0) {
        b = 0;
        t = 0.0;
      }
      lo[d] = b;
      hi[d] = b + 1 < g.n[d] ? b + 1 : b;
      w[d] = t;
    } else {
      lo[d] = hi[d] = 0;
      w[d] = 0.0;
    }
  }
}
// corner k (bit 0: x, bit 1: y, bit 2: z) of the interpolation cell
template <int D>
__device__ __forceinline__ int64_t corner(const GeomDev& g, const int64_t lo[3], const int64_t hi[3], int k) {
  const int64_t x = (k & 1) ? hi[0] : lo[0], y = (k & 2) ? hi[1] : lo[1], z = (k & 4) ? hi[2] : lo[2];
  return x + static_cast<int64_t>(g.n[0]) * (y + (D == 3 ? static_cast<int64_t>(g.n[1]) * z : 0));
}
// lerp of corner values v[8] (x, then y, then z)
template <int D>
__device__ __forceinline__ double lerp_cell(const double v[8], const double w[3]) {
  const double a00 = v[0] + (v[1] - v[0]) * w[0];
  const double a10 = v[2] + (v[3] - v[2]) * w[0];
  const double b0 = a00 + (a10 - a00) * w[1];
  if (D == 2) return b0;
  const double a01 = v[4] + (v[5] - v[4]) * w[0];
  const double a11 = v[6] + (v[7] - v[6]) * w[0];
  const double b1 = a01 + (a11 - a01) * w[1];
  return b0 + (b1 - b0) * w[2];
}

// BSplineInterpolateImageFunction's support and weights for order Q at
// continuous index x: the Q + 1 samples from start = floor(x) - Q/2 (odd Q)
// or floor(x + 1/2) - Q/2 (even Q), weighted by the centred B-spline of
// degree Q at x - j, written per order as the polynomial pieces that the
// fractional offset w selects (Unser, IEEE SPM 16(6) 1999, eq. 8) so that no
// division is needed.
template <int Q>
__device__ __forceinline__ int64_t bspline_weights(double x, double wt[Q + 1]) {
  const double fl = floor((Q & 1) ? x : x + 0.5);
  const double w = x - fl;  // odd Q: [0, 1); even Q: [-1/2, 1/2)
  if constexpr (Q == 0) {
    wt[0] = 1.0;
  } else if constexpr (Q == 1) {
    wt[0] = 1.0 - w;
    wt[1] = w;
  } else if constexpr (Q == 2) {
    const double a = 0.5 - w, b = 0.5 + w;
    wt[0] = 0.5 * a * a;
    wt[1] = 0.75 - w * w;
    wt[2] = 0.5 * b * b;
  } else if constexpr (Q == 3) {
    const double a = 1.0 - w, w2 = w * w, a2 = a * a;
    wt[0] = (1.0 / 6.0) * a2 * a;
    wt[1] = (2.0 / 3.0) - w2 + 0.5 * w2 * w;
    wt[2] = (2.0 / 3.0) - a2 + 0.5 * a2 * a;
    wt[3] = (1.0 / 6.0) * w2 * w;
  } else if constexpr (Q == 4) {
    // |t| < 1/2: 115/192 - 5/8 t^2 + 1/4 t^4; 1/2 <= |t| < 3/2:
    // 55/96 + 5/24 t - 5/4 t^2 + 5/6 t^3 - 1/6 t^4; 3/2 <= |t| < 5/2: (5/2 - |t|)^4 / 24
    auto mid = [](double t) {
      return (55.0 / 96.0) + t * ((5.0 / 24.0) + t * ((-5.0 / 4.0) + t * ((5.0 / 6.0) + t * (-1.0 / 6.0))));
    };
    const double a = 0.5 - w, b = 0.5 + w, w2 = w * w, a2 = a * a, b2 = b * b;
    wt[0] = (1.0 / 24.0) * a2 * a2;
    wt[1] = mid(1.0 + w);
    wt[2] = (115.0 / 192.0) + w2 * ((-5.0 / 8.0) + 0.25 * w2);
    wt[3] = mid(1.0 - w);
    wt[4] = (1.0 / 24.0) * b2 * b2;
  } else {
    static_assert(Q == 5, "B-spline order 0..5");
    // |t| < 1: 11/20 - 1/2 t^2 + 1/4 t^4 - 1/12 |t|^5; 1 <= |t| < 2:
    // 17/40 + 5/8 t - 7/4 t^2 + 5/4 t^3 - 3/8 t^4 + 1/24 t^5; 2 <= |t| < 3: (3 - |t|)^5 / 120
    auto in = [](double t) {
      const double t2 = t * t;
      return (11.0 / 20.0) + t2 * (-0.5 + t2 * (0.25 + t * (-1.0 / 12.0)));
    };
    auto mid = [](double t) {
      return (17.0 / 40.0) +
             t * ((5.0 / 8.0) + t * ((-7.0 / 4.0) + t * ((5.0 / 4.0) + t * ((-3.0 / 8.0) + t * (1.0 / 24.0)))));
    };
    const double a = 1.0 - w, a2 = a * a, w2 = w * w;
    wt[0] = (1.0 / 120.0) * a2 * a2 * a;
    wt[1] = mid(1.0 + w);
    wt[2] = in(w);
    wt[3] = in(a);
    wt[4] = mid(2.0 - w);
    wt[5] = (1.0 / 120.0) * w2 * w2 * w;
  }
  return static_cast<int64_t>(fl) - Q / 2;
}
// Mirror (whole-sample symmetric) index into [0, len): period 2 (len - 1).
__device__ __forceinline__ int64_t mirror_index(int64_t j, int64_t len) {
  if (len == 1) return 0;
  const int64_t period = 2 * (len - 1);
  j = j < 0 ? -j : j;
  j = j % period;
  return j >= len ? period - j : j;
}

// A physical point through the transform list in order: affine
// y = A p + off, or displacement field y = p + u(p) (u linearly
// interpolated in the field's grid, zero outside it).
// WindowedSincInterpolateImageFunction, radius m = 3: the kernel
// window(x) * sinc(x), sinc(x) = sin(pi x) / (pi x); windows (Meijering,
// Niessen & Viergever, Med. Image Anal. 3(2) 1999): 0 cosine
// cos(pi x / 2m), 1 Hamming 0.54 + 0.46 cos(pi x / m), 2 Lanczos
// sinc(x / m), 3 Blackman 0.42 + 0.5 cos(pi x / m) + 0.08 cos(2 pi x / m),
// 4 Welch 1 - x^2 / m^2.
constexpr int kSincRadius = 3;
// The six weights at offsets x_k = dist + 2 - k (k = 0..5) for 0 < dist < 1,
// from one sincos per angle instead of one per tap: sin(pi x_k) =
// (-1)^k sin(pi dist), and the windows' cos / sin of pi x_k / 3 (pi x_k / 6
// for the cosine window) by angle addition from x_0 with the constants
// cos(k pi / 3), sin(k pi / 3) (and k pi / 6, 2 k pi / 3).
template <int W>
__device__ __forceinline__ void sinc_weights(double dist, double wt[6]) {
  constexpr double pi = 3.14159265358979323846;
  constexpr double h = 0.86602540378443864676;  // sin(pi / 3)
  constexpr double C3[6] = {1.0, 0.5, -0.5, -1.0, -0.5, 0.5};
  constexpr double S3[6] = {0.0, h, h, 0.0, -h, -h};
  const double x0 = dist + 2.0;
  const double s0 = sin(pi * dist);
  double st, ct;  // sin, cos of pi x0 / 3 (pi x0 / 6 for the cosine window)
  sincos(x0 * (W == 0 ? pi / 6.0 : pi / 3.0), &st, &ct);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const double xk = x0 - k;
    const double sinc = ((k & 1) ? -s0 : s0) / (pi * xk);
    double w;
    if constexpr (W == 0) {
      constexpr double C6[6] = {1.0, h, 0.5, 0.0, -0.5, -h};
      constexpr double S6[6] = {0.0, 0.5, h, 1.0, h, 0.5};
      w = ct * C6[k] + st * S6[k];  // cos(pi x_k / 6)
    } else if constexpr (W == 1) {
      w = 0.54 + 0.46 * (ct * C3[k] + st * S3[k]);
    } else if constexpr (W == 2) {
      w = (st * C3[k] - ct * S3[k]) / (xk * (pi / 3.0));  // sinc(x_k / 3)
    } else if constexpr (W == 3) {
      constexpr double C23[6] = {1.0, -0.5, -0.5, 1.0, -0.5, -0.5};
      constexpr double S23[6] = {0.0, h, -h, 0.0, h, -h};
      const double c2 = ct * ct - st * st, s2 = 2.0 * st * ct;  // of 2 pi x0 / 3
      w = 0.42 + 0.5 * (ct * C3[k] + st * S3[k]) + 0.08 * (c2 * C23[k] + s2 * S23[k]);
    } else {
      static_assert(W == 4, "window 0..4");
      w = 1.0 - xk * xk * (1.0 / 9.0);
    }
    wt[k] = w * sinc;
  }
}

template <int D>
__device__ __forceinline__ void apply_chain(const ApplyDev& P, double p[3]) {
  for (int t = 0; t < P.nt; ++t) {
    const XformDev& T = P.t[t];
    if (T.kind == 1) {
      double q[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) s = s + T.a[r * 3 + k] * p[k];
        q[r] = r < D ? s + T.off[r] : 0.0;
      }
#pragma unroll
      for (int r = 0; r < 3; ++r) p[r] = q[r];
    } else if (T.kind == 2) {
      double c[3];
      geom_cidx<D>(T.fg, p, c);
      if (geom_inside<D>(T.fg, c)) {
        int64_t lo[3], hi[3];
        double w[3];
        lin_setup<D>(T.fg, c, lo, hi, w);
        int64_t idx[8];
#pragma unroll
        for (int k = 0; k < (1 << D); ++k) idx[k] = corner<D>(T.fg, lo, hi, k);
#pragma unroll
        for (int d = 0; d < D; ++d) {
          double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
          for (int k = 0; k < (1 << D); ++k) v[k] = static_cast<double>(T.f[idx[k] * D + d]);
          p[d] = p[d] + lerp_cell<D>(v, w);
        }
      }
    }
  }
}

// TransformToDisplacementFieldFilter: the displacement T(p) - p of each
// reference point p, D interleaved physical components per voxel.
template <int D>
__global__ void __launch_bounds__(kBlock) k_compose_field(GridDev go, ApplyDev P, float* __restrict__ out) {
  const int64_t N = go.count();
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= N) return;
  int x, y, z;
  unlin(go, i, x, y, z);
  const int ix[3] = {x, y, z};
  double p0[3], p[3];
  geom_point<D>(P.ref, ix, p0);
#pragma unroll
  for (int d = 0; d < 3; ++d) p[d] = p0[d];
  apply_chain<D>(P, p);
#pragma unroll
  for (int d = 0; d < D; ++d) out[i * D + d] = static_cast<float>(p[d] - p0[d]);
}

// CompositeTransform::TransformPoint over a point list: point i (D
// doubles at in + i * D) through the transform list, to out (may equal in).
template <int D>
__global__ void __launch_bounds__(kBlock) k_transform_points(ApplyDev P, const double* in, double* out,
                                                             int64_t n) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  double p[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int d = 0; d < D; ++d) p[d] = in[i * D + d];
  apply_chain<D>(P, p);
#pragma unroll
  for (int d = 0; d < D; ++d) out[i * D + d] = p[d];
}

template <int D, int INTERP>
__global__ void __launch_bounds__(kBlock) k_apply_transforms(GridDev go, ApplyDev P, const float* __restrict__ img,
                                                             float* __restrict__ out) {
  const int64_t N = go.count();
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= N) return;
  int x, y, z;
  unlin(go, i, x, y, z);
  const int ix[3] = {x, y, z};
  double p[3];
  geom_point<D>(P.ref, ix, p);
  apply_chain<D>(P, p);
  double c[3];
  geom_cidx<D>(P.in, p, c);
  if (!geom_inside<D>(P.in, c)) {
    out[i] = P.defv;
    return;
  }
  if (INTERP == 1) {
    int64_t id[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < D; ++d) id[d] = static_cast<int64_t>(floor(c[d] + 0.5));
    out[i] = img[id[0] + static_cast<int64_t>(P.in.n[0]) * (id[1] + static_cast<int64_t>(P.in.n[1]) * id[2])];
    return;
  }
  if constexpr (INTERP >= 32) {
    // WindowedSincInterpolateImageFunction: the 2m samples from floor(c) -
    // m + 1 to floor(c) + m on each axis, weighted by window(x) sinc(x) of
    // their offset x, windows (Meijering, Niessen & Viergever, Med. Image
    // Anal. 3(2) 1999, m = 3): cosine cos(pi x / 2m), Hamming 0.54 +
    // 0.46 cos(pi x / m), Lanczos sinc(x / m), Blackman 0.42 +
    // 0.5 cos(pi x / m) + 0.08 cos(2 pi x / m), Welch 1 - x^2 / m^2; a delta
    // at floor(c) where c is whole; samples outside the buffer 0
    constexpr int Wn = INTERP - 32, K = 2 * kSincRadius;
    int64_t base[3] = {0, 0, 0};
    double wt[3][K];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
#pragma unroll
      for (int k = 0; k < K; ++k) wt[d][k] = 0.0;
      if (d >= D) continue;
      const double fl = floor(c[d]);
      base[d] = static_cast<int64_t>(fl) - (kSincRadius - 1);
      const double dist = c[d] - fl;
      if (dist == 0.0) wt[d][kSincRadius - 1] = 1.0;
      else sinc_weights<Wn>(dist, wt[d]);
    }
    // separable sum: rows outside the buffer are skipped, taps outside a
    // row read nothing (their weight is zeroed; the clamped index stays in
    // the row)
    int xo[K];
    const bool xin_all = base[0] >= 0 && base[0] + K <= P.in.n[0];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int64_t x = base[0] + k;
      const bool in = x >= 0 && x < P.in.n[0];
      xo[k] = in ? static_cast<int>(x) : 0;
      if (!in) wt[0][k] = 0.0;
    }
    constexpr int KZ = D == 3 ? K : 1;
    double acc = 0.0;
    for (int kz = 0; kz < KZ; ++kz) {
      const int64_t z = base[2] + kz;
      if (D == 3 && (z < 0 || z >= P.in.n[2])) continue;
      double plane = 0.0;
      for (int ky = 0; ky < K; ++ky) {
        const int64_t y = base[1] + ky;
        if (y < 0 || y >= P.in.n[1]) continue;
        const float* __restrict__ row = img + (z * P.in.n[1] + y) * P.in.n[0];
        double r = 0.0;
        if (xin_all) {
          // interior: the taps are contiguous from base[0]
          const float* __restrict__ r0 = row + base[0];
#pragma unroll
          for (int kx = 0; kx < K; ++kx) r = fma(wt[0][kx], static_cast<double>(r0[kx]), r);
        } else {
#pragma unroll
          for (int kx = 0; kx < K; ++kx) r = fma(wt[0][kx], static_cast<double>(row[xo[kx]]), r);
        }
        plane = fma(wt[1][ky], r, plane);
      }
      acc = D == 3 ? fma(wt[2][kz], plane, acc) : plane;
    }
    out[i] = static_cast<float>(acc);
    return;
  } else if constexpr (INTERP >= 16) {
    // BSplineInterpolateImageFunction of order Q over the coefficients,
    // indices mirrored into the buffer
    constexpr int Q = INTERP - 16;
    int64_t idx[3][Q + 1];
    double wt[3][Q + 1];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (d >= D) {
#pragma unroll
        for (int k = 0; k <= Q; ++k) {
          idx[d][k] = 0;
          wt[d][k] = 0.0;
        }
        continue;
      }
      const int64_t start = bspline_weights<Q>(c[d], wt[d]);
#pragma unroll
      for (int k = 0; k <= Q; ++k) idx[d][k] = mirror_index(start + k, P.in.n[d]);
    }
    constexpr int QZ = D == 3 ? Q : 0;
    if constexpr (D == 2) wt[2][0] = 1.0;
    double acc = 0.0;
#pragma unroll
    for (int kz = 0; kz <= QZ; ++kz) {
      const int64_t oz = idx[2][kz] * P.in.n[1];
#pragma unroll
      for (int ky = 0; ky <= Q; ++ky) {
        const double* __restrict__ row = P.coef + (oz + idx[1][ky]) * P.in.n[0];
        double r = 0.0;
#pragma unroll
        for (int kx = 0; kx <= Q; ++kx) r = r + wt[0][kx] * row[idx[0][kx]];
        acc = acc + (wt[2][kz] * wt[1][ky]) * r;
      }
    }
    out[i] = static_cast<float>(acc);
    return;
  } else {
  int64_t lo[3], hi[3];
  double w[3];
  lin_setup<D>(P.in, c, lo, hi, w);
  float lab[8];
#pragma unroll
  for (int k = 0; k < (1 << D); ++k) lab[k] = img[corner<D>(P.in, lo, hi, k)];
  if (INTERP == 0) {
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < (1 << D); ++k) v[k] = static_cast<double>(lab[k]);
    out[i] = static_cast<float>(lerp_cell<D>(v, w));
    return;
  }
  // generic label: the label whose indicator image interpolates highest,
  // the smaller label on ties (labels absent from the cell interpolate to 0)
  float best = lab[0];
  double bestv = -1.0;
#pragma unroll
  for (int k = 0; k < (1 << D); ++k) {
    const float l = lab[k];
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < (1 << D); ++q) v[q] = lab[q] == l ? 1.0 : 0.0;
    const double val = lerp_cell<D>(v, w);
    if (val > bestv || (val == bestv && l < best)) {
      bestv = val;
      best = l;
    }
  }
  out[i] = best;
  }
}

// BSplineDecompositionImageFilter along one axis: every line is multiplied
// by the gain prod (1 - z)(1 - 1/z) of the order's poles, then filtered per
// pole by a causal pass (initial value: the mirrored geometric sum,
// truncated where |z|^n falls below 1e-10) and an anticausal pass (Unser,
// Aldroubi & Eden, IEEE TSP 41(2) 1993; the recursive filter of Unser,
// IEEE Signal Processing Magazine 16(6) 1999).
struct DecompPoles {
  int np;
  int64_t horizon[2];
  double z[2];
  double gain;
};

template <typename I>
__device__ __forceinline__ void decompose_line(double* __restrict__ L, I len, I step, const DecompPoles& P) {
  for (int p = 0; p < P.np; ++p) {
    const double z = P.z[p];
    double c0;
    if (P.horizon[p] < static_cast<int64_t>(len)) {
      double zn = z, sum = L[0];
      for (I n = 1; n < static_cast<I>(P.horizon[p]); ++n) {
        sum = sum + zn * L[n * step];
        zn = zn * z;
      }
      c0 = sum;
    } else {
      double zn = z;
      const double iz = 1.0 / z;
      double z2n = pow(z, static_cast<double>(len - 1));
      double sum = L[0] + z2n * L[(len - 1) * step];
      z2n = z2n * z2n * iz;
      for (I n = 1; n <= len - 2; ++n) {
        sum = sum + (zn + z2n) * L[n * step];
        zn = zn * z;
        z2n = z2n * iz;
      }
      c0 = sum / (1.0 - zn * zn);
    }
    L[0] = c0;
    double prev = c0;
    I k = 1;
    for (; k + 4 <= len; k += 4) {
      const double a0 = L[k * step], a1 = L[(k + 1) * step], a2 = L[(k + 2) * step], a3 = L[(k + 3) * step];
      const double b0 = a0 + z * prev;
      const double b1 = a1 + z * b0;
      const double b2 = a2 + z * b1;
      const double b3 = a3 + z * b2;
      L[k * step] = b0;
      L[(k + 1) * step] = b1;
      L[(k + 2) * step] = b2;
      L[(k + 3) * step] = b3;
      prev = b3;
    }
    for (; k < len; ++k) {
      prev = L[k * step] + z * prev;
      L[k * step] = prev;
    }
    prev = (z / (z * z - 1.0)) * (z * L[(len - 2) * step] + prev);
    L[(len - 1) * step] = prev;
    k = len - 2;
    for (; k - 3 >= 0; k -= 4) {
      const double a0 = L[k * step], a1 = L[(k - 1) * step], a2 = L[(k - 2) * step], a3 = L[(k - 3) * step];
      const double b0 = z * (prev - a0);
      const double b1 = z * (b0 - a1);
      const double b2 = z * (b1 - a2);
      const double b3 = z * (b2 - a3);
      L[k * step] = b0;
      L[(k - 1) * step] = b1;
      L[(k - 2) * step] = b2;
      L[(k - 3) * step] = b3;
      prev = b3;
    }
    for (; k >= 0; --k) {
      prev = z * (prev - L[k * step]);
      L[k * step] = prev;
    }
  }
}

// A tile of TL lines staged in LDS: the block (one wave) loads the tile with
// coalesced reads, multiplied by the gain (from the float image on the first
// axis), each of the first TL lanes runs its line's recursions in LDS with
// the next four samples loaded ahead, and the wave writes the tile back.  X:
// lines along the contiguous axis (rows of the tile padded to len + 1
// doubles); else TL neighbouring lines interleaved, sample k of line r at
// k * TL + r.
template <bool X, int TL>
__global__ void __launch_bounds__(64) k_bspline_decompose_lds(const float* __restrict__ fin, double* __restrict__ c,
                                                              int64_t lines, int64_t len, int64_t stride,
                                                              int64_t outer_stride, DecompPoles P) {
  extern __shared__ double lds_d[];
  const int tid = threadIdx.x;
  int64_t base;
  int nr;
  const int n = static_cast<int>(len);  // a tile holds at most 8192 samples
  if constexpr (X) {
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * TL;
    nr = static_cast<int>(min(static_cast<int64_t>(TL), lines - r0));
    base = r0 * len;
  } else {
    const int64_t tpb = (stride + TL - 1) / TL;  // tiles across the lines of one outer index
    const int64_t b = blockIdx.x / tpb, a0 = (blockIdx.x % tpb) * TL;
    nr = static_cast<int>(min(static_cast<int64_t>(TL), stride - a0));
    base = b * outer_stride + a0;
  }
  static_assert((TL & (TL - 1)) == 0 && TL <= 64, "TL: a power of two up to 64");
  constexpr int sh = TL == 64 ? 6 : TL == 32 ? 5 : TL == 16 ? 4 : 3;
  // tile element e: X: row e / n, sample e % n (contiguous in memory);
  // else sample e >> sh of line e & (TL - 1)
  const int total = X ? nr * n : n * TL;
  auto gaddr = [&](int e, bool& ok) -> int64_t {
    if constexpr (X) {
      ok = e < total;
      return base + e;
    } else {
      const int k = e >> sh, r = e & (TL - 1);
      ok = e < total && r < nr;
      return base + k * stride + r;
    }
  };
  auto laddr = [&](int e) -> int {
    if constexpr (X) {
      const int r = e / n;
      return e + r;  // r * (n + 1) + (e - r * n)
    } else {
      return e;
    }
  };
  // eight independent loads in flight per lane
  for (int e0 = tid; e0 < total; e0 += 64 * 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      bool ok;
      const int64_t g = gaddr(e0 + 64 * u, ok);
      v[u] = ok ? (fin ? static_cast<double>(fin[g]) : c[g]) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      bool ok;
      (void)gaddr(e0 + 64 * u, ok);
      if (ok) lds_d[laddr(e0 + 64 * u)] = v[u] * P.gain;
    }
  }
  __syncthreads();
  if (tid < nr) {
    if constexpr (X) decompose_line<int>(lds_d + tid * (n + 1), n, 1, P);
    else decompose_line<int>(lds_d + tid, n, TL, P);
  }
  __syncthreads();
  for (int e = tid; e < total; e += 64) {
    bool ok;
    const int64_t g = gaddr(e, ok);
    if (ok) c[g] = lds_d[laddr(e)];
  }
}

// Lines too long for an LDS tile: one thread per line in global memory
// (line l: offset l % stride within the axes before this one, l / stride
// over the axes after it).
__global__ void __launch_bounds__(kBlock) k_bspline_decompose(double* __restrict__ c, int64_t lines, int64_t len,
                                                              int64_t stride, int64_t outer_stride, DecompPoles P) {
  const int64_t l = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (l >= lines) return;
  double* __restrict__ v = c + (l / stride) * outer_stride + l % stride;
  for (int64_t k = 0; k < len; ++k) v[k * stride] = v[k * stride] * P.gain;
  decompose_line<int64_t>(v, len, stride, P);
}

__global__ void __launch_bounds__(kBlock) k_to_double(const float* __restrict__ in, double* __restrict__ out,
                                                      int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = static_cast<double>(in[i]);
}
#pragma clang fp contract(on)

// ---- B-spline fit of displacement fields (BSplineSyN) ----
// A pass along one axis views its input as [outer O][axis Lin][inner I] and
// its output as [O][Lout][I] (I counts floats: the components and the axes
// before this one).  Fit: control point k gathers the axis positions of its
// range [lo_k, hi_k] with weights fw[p][k - span_p]; evaluation: position v
// sums the control points span_v .. span_v + q with weights bw[v][j].

constexpr int kBspKPT = 4;  // output positions per thread of a strided pass
#ifndef ANTS_BSP_EVAL_BLOCKS
#define ANTS_BSP_EVAL_BLOCKS 8192
#endif
constexpr int kBspEvalBlocks = ANTS_BSP_EVAL_BLOCKS;  // blocks per field of the x evaluation

// Passes along y or z (and x when a row does not fit in LDS): 64 lanes over
// the contiguous inner index, 4 x kBspKPT output positions per block, blockIdx.z =
// field * O + outer.  `last` (fit): the pass that completes the lattice
// divides by W_k; i = (ky m0 + kx) D + c (3-D) or kx D + c (2-D).
template <int D, bool FIT>
__global__ void __launch_bounds__(256) k_bsp_strided(BspAxisDev A, BspDev T, PtrPack4 in, MutPack4 out, int I, int Lin,
                                                     int Lout, int Oc, int o_base, int m0, int last) {
  const int f = blockIdx.z / Oc;
  const int o = o_base + static_cast<int>(blockIdx.z - f * Oc);
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= I) return;
  const float* __restrict__ src = in.p[f] + static_cast<int64_t>(o) * Lin * I + i;
  float* __restrict__ dst = out.p[f] + static_cast<int64_t>(o) * Lout * I + i;
  const int kb = (blockIdx.y * 4 + threadIdx.y) * kBspKPT;
  for (int k = kb; k < kb + kBspKPT && k < Lout; ++k) {
  float acc = 0.f;
  if (FIT) {
    const int g0 = A.goff[k], len = A.goff[k + 1] - g0;
    const float* sp = src + static_cast<int64_t>(A.lo[k]) * I;
    const float* w = A.gw + g0;
    int j = 0;
    for (; j + 4 <= len; j += 4) {  // four independent loads in flight
      const float v0 = sp[static_cast<int64_t>(j) * I], v1 = sp[static_cast<int64_t>(j + 1) * I];
      const float v2 = sp[static_cast<int64_t>(j + 2) * I], v3 = sp[static_cast<int64_t>(j + 3) * I];
      acc = fmaf(w[j], v0, acc);
      acc = fmaf(w[j + 1], v1, acc);
      acc = fmaf(w[j + 2], v2, acc);
      acc = fmaf(w[j + 3], v3, acc);
    }
    for (; j < len; ++j) acc = fmaf(w[j], sp[static_cast<int64_t>(j) * I], acc);
    if (last) {
      const int kxy = i / D;
      const int kk[3] = {D == 3 ? kxy % m0 : kxy, D == 3 ? kxy / m0 : k, k};
      double pa = 1.0, pi = 1.0;
      for (int d = 0; d < D; ++d) {
        pa *= T.ax[d].A[kk[d]];
        pi *= T.ax[d].Ai[kk[d]];
      }
      const double W = pa + (1.0e10 - 1.0) * (pa - pi);
      acc = W != 0.0 ? static_cast<float>(static_cast<double>(acc) / W) : 0.f;
    }
  } else {
    const float* s0 = src + static_cast<int64_t>(A.span[k]) * I;
    for (int j = 0; j <= T.q; ++j) acc = fmaf(A.bw[k * 4 + j], s0[static_cast<int64_t>(j) * I], acc);
  }
  dst[static_cast<int64_t>(k) * I] = acc;
  }
}

// Fit along x: a block stages the axis's fit weights and R consecutive rows
// (n0 voxels x D components each, contiguous in memory) in LDS, the rows with
// eight independent loads in flight per thread, then computes the m0 x D
// control-point values of each row from LDS.  LDS: gtot weights, 2 m0 + 1
// ints, R n0 D floats.
template <int D>
__global__ void __launch_bounds__(256) k_bsp_fit_x(BspAxisDev A, PtrPack4 in, MutPack4 out, int rows, int R) {
  extern __shared__ float lds[];
  const int n0 = A.n, m0 = A.m;
  float* wl = lds;
  int* lol = reinterpret_cast<int*>(lds + A.gtot);
  int* gol = lol + m0;
  float* rowbuf = reinterpret_cast<float*>(gol + m0 + 1);
  const int row0 = blockIdx.x * R;
  const int nr = min(R, rows - row0);
  const float* __restrict__ src = in.p[blockIdx.y] + static_cast<int64_t>(row0) * n0 * D;
  float* __restrict__ dst = out.p[blockIdx.y] + static_cast<int64_t>(row0) * m0 * D;
  const int ne = nr * n0 * D;
  for (int e0 = threadIdx.x; e0 < ne; e0 += 256 * 8) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + u * 256;
      v[u] = e < ne ? src[e] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + u * 256;
      if (e < ne) rowbuf[e] = v[u];
    }
  }
  for (int e = threadIdx.x; e < A.gtot; e += 256) wl[e] = A.gw[e];
  for (int e = threadIdx.x; e < m0; e += 256) lol[e] = A.lo[e];
  for (int e = threadIdx.x; e <= m0; e += 256) gol[e] = A.goff[e];
  __syncthreads();
  const int per_row = m0 * D;
  for (int t = threadIdx.x; t < nr * per_row; t += 256) {
    const int r = t / per_row;
    const int kc = t - r * per_row;
    const int k = kc / D;
    const int c = kc - k * D;
    const float* lr = rowbuf + (r * n0 + lol[k]) * D + c;
    const int g0 = gol[k], len = gol[k + 1] - g0;
    float acc = 0.f;
    for (int j = 0; j < len; ++j) acc = fmaf(wl[g0 + j], lr[j * D], acc);
    dst[t] = acc;
  }
}

// Evaluation along x (the last pass): one voxel per thread (consecutive
// lanes on consecutive voxels of a row), grid-stride over the voxels with
// enough blocks to keep the dependent table -> control-point loads in
// flight.  Each block stores its max |u / spacing| in part[field][block]
// (when part is set); k_max_partials folds them.  Same-address device
// atomics from every block serialise at the memory side (about 11 ns each
// measured here), so the blocks write plain partials instead.
template <int D>
__global__ void __launch_bounds__(256) k_bsp_eval_x(BspAxisDev A, PtrPack4 in, MutPack4 out, int nvox, int q,
                                                    float sp0, float sp1, float sp2, float* part) {
  const int n0 = A.n, m0 = A.m;
  const float isp[3] = {1.f / sp0, 1.f / sp1, 1.f / sp2};
  const float* __restrict__ src = in.p[blockIdx.y];
  float* __restrict__ dst = out.p[blockIdx.y];
  float nrm = 0.f;
  for (int v = blockIdx.x * 256 + threadIdx.x; v < nvox; v += gridDim.x * 256) {
    const int row = v / n0;
    const int x = v - row * n0;
    const float4 w = *reinterpret_cast<const float4*>(A.bw + 4 * x);
    const float* s0 = src + (static_cast<int64_t>(row) * m0 + A.span[x]) * D;
    const float wj[4] = {w.x, w.y, w.z, w.w};
    float acc[D];
#pragma unroll
    for (int c = 0; c < D; ++c) acc[c] = wj[0] * s0[c];
#pragma unroll
    for (int j = 1; j < 4; ++j)
      if (j <= q) {
#pragma unroll
        for (int c = 0; c < D; ++c) acc[c] = fmaf(wj[j], s0[j * D + c], acc[c]);
      }
    float s2 = 0.f;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      dst[static_cast<int64_t>(v) * D + c] = acc[c];
      const float t = acc[c] * isp[c];
      s2 += t * t;
    }
    nrm = fmaxf(nrm, sqrtf(s2));
  }
  if (part) {
    double dummy = 0.0;
    block_sum_max<256>(dummy, nrm);
    if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = nrm;
  }
}

// Fused evaluation along y and x (the last passes): a block owns voxel rows
// [y0, y0 + R) of one z plane; it stages the control-point rows they need
// (the y-evaluated input (m0, m1, n2): rows span(y0) .. span(y0+R-1) + q)
// and the x tables in LDS, evaluates the R x-rows of control points along y
// into LDS, then the R field rows along x, and stores the block's max
// |u / spacing| in part (same arithmetic and order as the separate passes).
template <int D>
__global__ void __launch_bounds__(256) k_bsp_eval_yx(BspAxisDev AX, BspAxisDev AY, PtrPack4 in, MutPack4 out,
                                                     int R, int q, float sp0, float sp1, float sp2, float* part) {
  extern __shared__ float lds[];
  const int n0 = AX.n, m0 = AX.m, n1 = AY.n, m1 = AY.m;
  const int per_row = m0 * D;
  float* bwl = lds;
  int* spl = reinterpret_cast<int*>(bwl + 4 * n0);
  float* t1 = reinterpret_cast<float*>(spl + n0);
  float* cprow = t1 + R * per_row;
  const float isp[3] = {1.f / sp0, 1.f / sp1, 1.f / sp2};
  const int y0 = blockIdx.x * R;
  const int nr = min(R, n1 - y0);
  const int z = blockIdx.y;
  const int ky_lo = AY.span[y0];
  const int nk = AY.span[y0 + nr - 1] + q - ky_lo + 1;
  const float* __restrict__ src = in.p[blockIdx.z] + (static_cast<int64_t>(z) * m1 + ky_lo) * per_row;
  float* __restrict__ dst = out.p[blockIdx.z] + (static_cast<int64_t>(z) * n1 + y0) * n0 * D;
  for (int e = threadIdx.x; e < 4 * n0; e += 256) bwl[e] = AX.bw[e];
  for (int e = threadIdx.x; e < n0; e += 256) spl[e] = AX.span[e];
  for (int e = threadIdx.x; e < nk * per_row; e += 256) cprow[e] = src[e];
  __syncthreads();
  for (int t = threadIdx.x; t < nr * per_row; t += 256) {
    const int r = t / per_row;
    const int kc = t - r * per_row;
    const int y = y0 + r;
    const float* s0 = cprow + (AY.span[y] - ky_lo) * per_row + kc;
    float a = 0.f;
    for (int j = 0; j <= q; ++j) a = fmaf(AY.bw[y * 4 + j], s0[j * per_row], a);
    t1[t] = a;
  }
  __syncthreads();
  float nrm = 0.f;
  for (int t = threadIdx.x; t < nr * n0; t += 256) {
    const int r = t / n0;
    const int x = t - r * n0;
    const float* s0 = t1 + (r * m0 + spl[x]) * D;
    const float4 w = *reinterpret_cast<const float4*>(bwl + 4 * x);
    const float wj[4] = {w.x, w.y, w.z, w.w};
    float acc[D];
#pragma unroll
    for (int c = 0; c < D; ++c) acc[c] = wj[0] * s0[c];
#pragma unroll
    for (int j = 1; j < 4; ++j)
      if (j <= q) {
#pragma unroll
        for (int c = 0; c < D; ++c) acc[c] = fmaf(wj[j], s0[j * D + c], acc[c]);
      }
    float s2 = 0.f;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      dst[static_cast<int64_t>(t) * D + c] = acc[c];
      const float v = acc[c] * isp[c];
      s2 += v * v;
    }
    nrm = fmaxf(nrm, sqrtf(s2));
  }
  if (part) {
    double dummy = 0.0;
    block_sum_max<256>(dummy, nrm);
    if (threadIdx.x == 0) part[blockIdx.z * (gridDim.x * gridDim.y) + blockIdx.y * gridDim.x + blockIdx.x] = nrm;
  }
}

// Image moments for CenteredTransformInitializer::MomentsOn (ImageMoments-
// Calculator's centre of gravity): per block, in double, the sums of I and
// of I times each index coordinate; out[block * 4 + {0, 1, 2, 3}].  The
// host sums the blocks in order and maps the index centroid to physical
// space (a linear map, so it commutes with the mean).
__global__ void __launch_bounds__(kBlock) k_moments(GridDev g, const float* __restrict__ img, double* out) {
  const int64_t N = g.count();
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < N;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    int x, y, z;
    unlin(g, i, x, y, z);
    const double v = img[i];
    a[0] += v;
    a[1] += v * x;
    a[2] += v * y;
    a[3] += v * z;
  }
  __shared__ double sh[4][kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double w = wave_sum(a[k]);
    if (lane == 0) sh[k][wid] = w;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double t = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) t += sh[threadIdx.x][w];
    out[blockIdx.x * 4 + threadIdx.x] = t;
  }
}

// One block per field: the max of nb partials into maxbits[field].
__global__ void __launch_bounds__(1024) k_max_partials(const float* __restrict__ part, int nb, unsigned int* maxbits) {
  float m = 0.f;
  for (int i = threadIdx.x; i < nb; i += 1024) m = fmaxf(m, part[blockIdx.x * nb + i]);
  double dummy = 0.0;
  block_sum_max<1024>(dummy, m);
  if (threadIdx.x == 0) atomicMax(&maxbits[blockIdx.x], __float_as_uint(m));
}

// ---------------------------------------------------------------------------
// Linear stages: one pass over the virtual sample points of a level gives the
// metric's sums at one set of stage parameters (ImageToImageMetricv4 with a
// Translation / Euler / Affine transform under the moving initial composite).
// Coordinates, interpolation weights and sums are double; the images float.
__device__ __forceinline__ bool lin_sample(const GeomDev& g, const float* __restrict__ img, const double p[3],
                                           double& v) {
  double c[3];
  if (g.D == 2) {
    geom_cidx<2>(g, p, c);
    if (!geom_inside<2>(g, c)) return false;
    int64_t lo[3], hi[3];
    double w[3];
    lin_setup<2>(g, c, lo, hi, w);
    double cv[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 4; ++k) cv[k] = static_cast<double>(img[corner<2>(g, lo, hi, k)]);
    v = lerp_cell<2>(cv, w);
    return true;
  }
  geom_cidx<3>(g, p, c);
  if (!geom_inside<3>(g, c)) return false;
  int64_t lo[3], hi[3];
  double w[3];
  lin_setup<3>(g, c, lo, hi, w);
  double cv[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) cv[k] = static_cast<double>(img[corner<3>(g, lo, hi, k)]);
  v = lerp_cell<3>(cv, w);
  return true;
}

// Centred cubic B-spline and its derivative (Parzen window of the Mattes metric).
__device__ __forceinline__ double bspline3(double u) {
  const double a = fabs(u);
  if (a < 1.0) return (4.0 - 6.0 * a * a + 3.0 * a * a * a) / 6.0;
  if (a < 2.0) return (2.0 - a) * (2.0 - a) * (2.0 - a) / 6.0;
  return 0.0;
}
__device__ __forceinline__ double bspline3_d(double u) {
  const double a = fabs(u);
  const double s = u > 0.0 ? 1.0 : (u < 0.0 ? -1.0 : 0.0);
  if (a < 1.0) return s * (-2.0 * a + 1.5 * a * a);
  if (a < 2.0) return -s * 0.5 * (2.0 - a) * (2.0 - a);
  return 0.0;
}

template <int MODE, bool DERIV, int P>
constexpr int lin_q() {
  return MODE == kLinMS ? (DERIV ? 2 + P : 2)
                        : MODE == kLinGC ? (DERIV ? 6 + 3 * P : 6) : MODE == kLinHist ? 1 : 1 + P;
}

template <int D, int MODE, bool DERIV, int P>
__global__ void __launch_bounds__(kLinThreads) k_lin_metric(LinDev L, LinJac J, double* __restrict__ partials,
                                                            unsigned long long* __restrict__ hist,
                                                            unsigned int* __restrict__ fcnt) {
  constexpr int Q = lin_q<MODE, DERIV, P>();
  constexpr bool GRAD = DERIV && MODE != kLinHist;
  constexpr int HB = MODE == kLinHist ? kLinLdsBins : 1;
  __shared__ unsigned long long hsh[HB * HB];
  __shared__ unsigned int csh[HB];
  const bool lds_hist = MODE == kLinHist && L.bins <= kLinLdsBins;
  if (MODE == kLinHist && lds_hist) {
    for (int k = threadIdx.x; k < L.bins * L.bins; k += blockDim.x) hsh[k] = 0ull;
    for (int k = threadIdx.x; k < L.bins; k += blockDim.x) csh[k] = 0u;
    __syncthreads();
  }
  double acc[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) acc[q] = 0.0;
  const int64_t n = L.pts ? L.npts : static_cast<int64_t>(L.vg.n[0]) * L.vg.n[1] * L.vg.n[2];
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    double x[3] = {0.0, 0.0, 0.0};
    if (L.pts) {
#pragma unroll
      for (int d = 0; d < D; ++d) x[d] = L.pts[i * D + d];
    } else {
      const int64_t nx = L.vg.n[0], ny = L.vg.n[1];
      const int ix[3] = {static_cast<int>(i % nx), static_cast<int>((i / nx) % ny), static_cast<int>(i / (nx * ny))};
      geom_point<D>(L.vg, ix, x);
    }
    double pf[3] = {0.0, 0.0, 0.0}, py[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int r = 0; r < D; ++r) {
      double s = 0.0, t = 0.0;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        s = s + L.fa[r * 3 + k] * x[k];
        t = t + L.ya[r * 3 + k] * x[k];
      }
      pf[r] = s + L.fb[r];
      py[r] = t + L.yb[r];
    }
    double f, m;
    if (!lin_sample(L.fg, L.fimg, pf, f)) continue;
    if (!lin_sample(L.mg, L.mimg, py, m)) continue;
    double dm[P > 0 ? P : 1];
    if (GRAD || MODE == kLinHistDeriv) {
      // CentralDifferenceImageFunction::Evaluate(point) at py: neighbours
      // -+ half a spacing along physical axis d, 0 when either leaves the
      // buffer
      double gr[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int d = 0; d < D; ++d) {
        double lo[3] = {py[0], py[1], py[2]}, hi[3] = {py[0], py[1], py[2]};
        const double off = 0.5 * L.msp[d];
        lo[d] = py[d] - off;
        hi[d] = py[d] + off;
        double vl, vh;
        if (!lin_sample(L.mg, L.mimg, lo, vl)) continue;
        if (!lin_sample(L.mg, L.mimg, hi, vh)) continue;
        const double delta = hi[d] - lo[d];
        gr[d] = delta > 10.0 * 2.220446049250313e-16 ? (vh - vl) / delta : 0.0;
      }
      // u = grad^T G (the moving initial chain's linear part), r = x - c
      double u[3] = {0.0, 0.0, 0.0}, rr[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int k = 0; k < D; ++k) {
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) s = s + gr[d] * L.G[d * 3 + k];
        u[k] = s;
        rr[k] = x[k] - L.c[k];
      }
#pragma unroll
      for (int p = 0; p < P; ++p) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) {
          double t = J.tau[p][k];
#pragma unroll
          for (int j = 0; j < D; ++j) t = t + J.L[p][k * 3 + j] * rr[j];
          s = s + u[k] * t;
        }
        dm[p] = s;
      }
    }
    if (MODE == kLinMS) {
      const double diff = f - m;
      acc[0] += 1.0;
      acc[1] += diff * diff;
      if (DERIV) {
#pragma unroll
        for (int p = 0; p < P; ++p) acc[2 + p] += 2.0 * diff * dm[p];
      }
    } else if (MODE == kLinGC) {
      acc[0] += 1.0;
      acc[1] += f;
      acc[2] += m;
      acc[3] += f * f;
      acc[4] += m * m;
      acc[5] += f * m;
      if (DERIV) {
#pragma unroll
        for (int p = 0; p < P; ++p) {
          acc[6 + p] += dm[p];
          acc[6 + P + p] += f * dm[p];
          acc[6 + 2 * P + p] += m * dm[p];
        }
      }
    } else {
      // Mattes: fixed bin (zero-order window), moving cubic window
      const int B = L.bins;
      int64_t fi = static_cast<int64_t>(f / L.fbs - L.fnm);
      fi = fi < 2 ? 2 : (fi > B - 3 ? B - 3 : fi);
      const double mt = m / L.mbs - L.mnm;
      int64_t mi = static_cast<int64_t>(mt);
      mi = mi < 2 ? 2 : (mi > B - 3 ? B - 3 : mi);
      acc[0] += 1.0;
      if (MODE == kLinHist) {
        if (lds_hist) atomicAdd(&csh[fi], 1u);
        else atomicAdd(&fcnt[fi], 1u);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int64_t kk = mi - 1 + k;
          const unsigned long long q = __double2ull_rn(bspline3(static_cast<double>(kk) - mt) * kLinHistScale);
          if (lds_hist) atomicAdd(&hsh[fi * B + kk], q);
          else atomicAdd(&hist[fi * B + kk], q);
        }
      } else {
        double w = 0.0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int64_t kk = mi - 1 + k;
          w += L.wtab[fi * B + kk] * (-bspline3_d(static_cast<double>(kk) - mt));
        }
#pragma unroll
        for (int p = 0; p < P; ++p) acc[1 + p] += w * dm[p];
      }
    }
  }
  // block sums in a fixed order
  constexpr int NW = kLinThreads / 64;
  __shared__ double red[Q][NW];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const double w = wave_sum(acc[q]);
    if (lane == 0) red[q][wid] = w;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < Q; q += blockDim.x) {
    double t = 0.0;
    for (int w = 0; w < NW; ++w) t += red[q][w];
    partials[static_cast<int64_t>(blockIdx.x) * Q + q] = t;
  }
  if (MODE == kLinHist && lds_hist) {
    for (int k = threadIdx.x; k < L.bins * L.bins; k += blockDim.x)
      if (hsh[k]) atomicAdd(&hist[k], hsh[k]);
    for (int k = threadIdx.x; k < L.bins; k += blockDim.x)
      if (csh[k]) atomicAdd(&fcnt[k], csh[k]);
  }
}

// Q sums of nb block partials, each in a fixed order (one block).
__global__ void __launch_bounds__(256) k_lin_reduce(const double* __restrict__ partials, int nb, int Q,
                                                     double* __restrict__ out) {
  __shared__ double sh[256];
  for (int q = 0; q < Q; ++q) {
    double t = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) t += partials[static_cast<int64_t>(b) * Q + q];
    sh[threadIdx.x] = t;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = sh[0];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Dense field-transform stages (GaussianDisplacementField, Exponential).
// The image at y = A (x + u(x)) + b for each voxel x of the virtual grid
// (u may be null), linear, 0 outside its buffer; with GRAD also the
// CentralDifferenceImageFunction gradient at y -- samples half a spacing
// either side along each physical axis, (hi - lo) / (their distance), 0 for
// an axis where either leaves the buffer -- taken back through A (A^T grad):
// the derivative of the metric by the displacement at x.  Double
// coordinates, contraction off (as the antsApplyTransforms kernel).
// With u2 the displacement is u + s2 u2, formed as the field update kernel
// forms it (the same float operations): a line-search trial is sampled
// without writing the trial field.
template <int D, bool GRAD>
__global__ void __launch_bounds__(kBlock) k_field_sample(GeomDev vg, GeomDev ig, const float* __restrict__ img,
                                                         const float* __restrict__ u, const float* __restrict__ u2,
                                                         float s2, FieldMap map, float* __restrict__ out,
                                                         float* __restrict__ grad) {
  const int64_t N = static_cast<int64_t>(vg.n[0]) * vg.n[1] * (D == 3 ? vg.n[2] : 1);
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int64_t nx = vg.n[0], ny = vg.n[1];
  const int ix[3] = {static_cast<int>(i % nx), static_cast<int>((i / nx) % ny), static_cast<int>(i / (nx * ny))};
  double x[3];
  geom_point<D>(vg, ix, x);
  if (u) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      float ud = u[i * D + d];
      if (u2) {
        const float t = s2 * u2[i * D + d];
        ud = ud + t;
      }
      x[d] = x[d] + static_cast<double>(ud);
    }
  }
  double y[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int r = 0; r < D; ++r) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) t = t + map.a[r * 3 + k] * x[k];
    y[r] = t + map.b[r];
  }
  double v = 0.0;
  if (!lin_sample(ig, img, y, v)) v = 0.0;
  out[i] = static_cast<float>(v);
  if constexpr (GRAD) {
    double gr[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double lo[3] = {y[0], y[1], y[2]}, hi[3] = {y[0], y[1], y[2]};
      const double off = 0.5 * map.sp[d];
      lo[d] = y[d] - off;
      hi[d] = y[d] + off;
      double vl, vh;
      if (!lin_sample(ig, img, lo, vl)) continue;
      if (!lin_sample(ig, img, hi, vh)) continue;
      const double delta = hi[d] - lo[d];
      gr[d] = delta > 10.0 * 2.220446049250313e-16 ? (vh - vl) / delta : 0.0;
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
      double t = 0.0;
#pragma unroll
      for (int d = 0; d < D; ++d) t = t + gr[d] * map.a[d * 3 + k];
      grad[i * D + k] = static_cast<float>(t);
    }
  }
}

// Per-block partials of a field pair: sum a.b, sum b.b, sum a.a, max |a|^2
// (b may be null: zeros), over n vectors of D components; fixed order
// within a block.
template <int D>
__global__ void __launch_bounds__(kLinThreads) k_field_dots(int64_t n, const float* __restrict__ a,
                                                            const float* __restrict__ b,
                                                            double* __restrict__ partials) {
  double ab = 0.0, bb = 0.0, aa = 0.0, am = 0.0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    double s2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double av = a[i * D + d];
      s2 += av * av;
      if (b) {
        const double bv = b[i * D + d];
        ab += av * bv;
        bb += bv * bv;
      }
    }
    aa += s2;
    am = fmax(am, s2);
  }
  constexpr int NW = kLinThreads / 64;
  __shared__ double red[4][NW];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  ab = wave_sum(ab);
  bb = wave_sum(bb);
  aa = wave_sum(aa);
  for (int o = 32; o > 0; o >>= 1) am = fmax(am, __shfl_xor(am, o));
  if (lane == 0) {
    red[0][wid] = ab;
    red[1][wid] = bb;
    red[2][wid] = aa;
    red[3][wid] = am;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, m = 0.0;
    for (int w = 0; w < NW; ++w) {
      s0 += red[0][w];
      s1 += red[1][w];
      s2 += red[2][w];
      m = fmax(m, red[3][w]);
    }
    partials[blockIdx.x * 4 + 0] = s0;
    partials[blockIdx.x * 4 + 1] = s1;
    partials[blockIdx.x * 4 + 2] = s2;
    partials[blockIdx.x * 4 + 3] = m;
  }
}

// One block: the block partials reduced in a fixed order (thread t takes
// partials t, t + 256, ... in turn, then a fixed tree over the threads):
// three sums and a max, the same bits on every run.
__global__ void __launch_bounds__(256) k_field_dots_final(const double* __restrict__ partials, int nb,
                                                          double* __restrict__ out) {
  __shared__ double sh[4][256];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, m = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) {
    s0 += partials[b * 4 + 0];
    s1 += partials[b * 4 + 1];
    s2 += partials[b * 4 + 2];
    m = fmax(m, partials[b * 4 + 3]);
  }
  sh[0][threadIdx.x] = s0;
  sh[1][threadIdx.x] = s1;
  sh[2][threadIdx.x] = s2;
  sh[3][threadIdx.x] = m;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + h];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + h];
      sh[2][threadIdx.x] += sh[2][threadIdx.x + h];
      sh[3][threadIdx.x] = fmax(sh[3][threadIdx.x], sh[3][threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) out[threadIdx.x] = sh[threadIdx.x][0];
}

// out = a + s * b elementwise (a null: s * b); out may alias a or b.
__global__ void __launch_bounds__(kBlock) k_field_axpy(int64_t n, const float* a, float sc, const float* b,
                                                       float* out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float t = sc * b[i];
  out[i] = a ? a[i] + t : t;
}

inline unsigned int blocks_for(int64_t n, int bs = kBlock) { return static_cast<unsigned int>((n + bs - 1) / bs); }
inline unsigned int red_blocks(int64_t n) {
  const unsigned int b = blocks_for(n);
  return b < static_cast<unsigned int>(kRedBlocks) ? (b < 1 ? 1 : b) : static_cast<unsigned int>(kRedBlocks);
}
inline unsigned int grid_stride_blocks(int64_t n) {
  int64_t b = (n + kBlock - 1) / kBlock;
  return static_cast<unsigned int>(b < kRedBlocks ? (b < 1 ? 1 : b) : kRedBlocks);
}

template <class A>
PtrPack4 pack(const A* const* p, int np) {
  PtrPack4 r{};
  for (int k = 0; k < np; ++k) r.p[k] = p[k];
  return r;
}
MutPack4 mpack(float* const* p, int np) {
  MutPack4 r{};
  for (int k = 0; k < np; ++k) r.p[k] = p[k];
  return r;
}

// CC tile shapes (3-D): TX x TY columns; the z-chunk depth is chosen per grid
// from the number of blocks the device holds at once (zchunk_for).
// Tile of the 3-D CC kernel: 16 x 16 columns per block measured fastest over
// the pyramid levels of a 192x224x192 registration (the x box sums are
// computed for TY + 2R rows per TY output rows, so square tiles waste less).
#ifndef ANTS_CC_TX
#define ANTS_CC_TX 16
#endif
#ifndef ANTS_CC_TY
#define ANTS_CC_TY 16
#endif
constexpr int kCCTX = ANTS_CC_TX, kCCTY = ANTS_CC_TY;
// z-chunk of the tiled CC and smoothing kernels.  ANTS_ZCHUNK_MODEL=0 selects
// the former rule (about k*_TARGET_BLOCKS blocks within [ZMIN, ZMAX] planes);
// the default picks the depth from the resident-block count (zchunk_for),
// measured 52.0 -> 49.9 ms per 192x224x192 pair (smoothing 9.45 -> 7.7 ms,
// CC 7.72 -> 7.25 ms): at full resolution 3 chunks of 64 planes fill the
// 512 resident blocks in one round, where 7 chunks of 31 took three.
#ifndef ANTS_CC_TARGET_BLOCKS
#define ANTS_CC_TARGET_BLOCKS 1024
#endif
#ifndef ANTS_CC_ZMIN
#define ANTS_CC_ZMIN 8
#endif
#ifndef ANTS_SMOOTH_TARGET_BLOCKS
#define ANTS_SMOOTH_TARGET_BLOCKS 1024
#endif
#ifndef ANTS_SMOOTH_ZMIN
#define ANTS_SMOOTH_ZMIN 16
#endif

#ifndef ANTS_ZCHUNK_MODEL
#define ANTS_ZCHUNK_MODEL 1
#endif
#ifndef ANTS_ZCHUNK_MIN
#define ANTS_ZCHUNK_MIN 4
#endif

}  // namespace

std::atomic<int> g_tune[kTuneKnobCount][kTuneBuckets];
std::atomic<uint64_t> g_tune_gen{1};

static int tune_bucket(int64_t voxels) {
  int b = 0;
  while (b + 1 < kTuneBuckets && (int64_t{1} << (b + 1)) <= voxels) ++b;
  return b;
}
int tune_get(int knob, int64_t voxels) {
  if (knob < 0 || knob >= kTuneKnobCount || voxels < 1) return 0;
  return g_tune[knob][tune_bucket(voxels)].load(std::memory_order_relaxed);
}
int tune_set(int knob, int bucket, int value) {
  if (knob < 0 || knob >= kTuneKnobCount || bucket < 0 || bucket >= kTuneBuckets || value < 0) return -1;
  if ((knob == kTuneCCTile || knob == kTuneSmoothTile) && value > 2) return -1;
  g_tune[knob][bucket].store(value, std::memory_order_relaxed);
  g_tune_gen.fetch_add(1);
  return 0;
}
void tune_clear() {
  for (auto& k : g_tune)
    for (auto& v : k) v.store(0, std::memory_order_relaxed);
  g_tune_gen.fetch_add(1);
}
uint64_t tune_generation() { return g_tune_gen.load(); }

namespace {

// Blocks of a kernel resident on the whole device at once: blocks per CU
// (from its register and LDS use) times the CU count.
// `per_cu` > 0 overrides the occupancy query.
int64_t device_slots(const void* kernel, int threads, int per_cu = 0) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess || per < 1) per = 1;
  if (per_cu > 0) per = per_cu;
  return static_cast<int64_t>(cus) * per;
}
#ifndef ANTS_CC_SLOTS_PER_CU
#define ANTS_CC_SLOTS_PER_CU 0
#endif
#ifndef ANTS_SMOOTH_SLOTS_PER_CU
#define ANTS_SMOOTH_SLOTS_PER_CU 0
#endif

// z-chunk depth of a kernel that streams `columns` block columns along z,
// reading `halo` warm-up planes per chunk: the depth that minimises
// rounds x planes per block, where rounds = ceil(blocks / slots) waves of
// resident blocks.  Ties go to the deeper chunk (fewer halo reads).
int zchunk_for(int nz, int64_t columns, int halo, int64_t slots) {
  int best_zc = nz;
  int64_t best = INT64_MAX;
  for (int zc = nz; zc >= 1; --zc) {
    if (zc < ANTS_ZCHUNK_MIN && zc < nz) break;
    const int64_t nzc = (nz + zc - 1) / zc;
    if (zc > 1 && (nz + zc - 2) / (zc - 1) == nzc) continue;  // same chunk count as a shallower depth
    const int64_t rounds = (columns * nzc + slots - 1) / slots;
    const int64_t cost = rounds * (zc + halo);
    if (cost < best) {
      best = cost;
      best_zc = zc;
    }
  }
  return best_zc;
}

// Box sums of the tiled CC: double unless the caller asks for float sums;
// ANTS_SYN_CC_SUMS=fp32|fp64 overrides (read per launch).  Measured per
// 192x224x192 pair: float sums 6.34 ms of CC against 5.88 ms in double (160
// against 226 VGPRs, 3 against 2 waves per SIMD, but the same LDS
// instruction count and the fixed-order z sums on top), so double stays the
// default.
bool cc_sums_fp64(bool fp32_sums) {
  const char* e = std::getenv("ANTS_SYN_CC_SUMS");
  if (e && std::strcmp(e, "fp64") == 0) return true;
  if (e && std::strcmp(e, "fp32") == 0) return false;
  return !fp32_sums;
}

// Launch shape of the tiled CC: tile (tuned, else 16 x 16) and z-chunk
// depth (tuned, else chosen by zchunk_for from the resident-block count).
// Float sums (the BF16 form) keep the 16 x 16 tile.
struct CCShape {
  int tx, ty, zc;
};
CCShape cc_shape(const GridDev& g, int radius, bool fp32_sums) {
  const bool d = cc_sums_fp64(fp32_sums);
  const int64_t N = g.count();
  const bool narrow = d && tune_get(kTuneCCTile, N) == 2;
  CCShape sh{kCCTX, narrow ? kCCTY / 2 : kCCTY, 0};
  const int64_t tiles = static_cast<int64_t>((g.n[0] + sh.tx - 1) / sh.tx) * ((g.n[1] + sh.ty - 1) / sh.ty);
  const int tz = tune_get(kTuneCCZChunk, N);
  if (tz > 0) {
    sh.zc = tz < g.n[2] ? tz : g.n[2];
    return sh;
  }
#if ANTS_ZCHUNK_MODEL
  static const int64_t slots4 = device_slots(reinterpret_cast<const void*>(&k_cc3d<4, kCCTX, kCCTY, double>),
                                             kCCTX * kCCTY, ANTS_CC_SLOTS_PER_CU);
  static const int64_t slots2 = device_slots(reinterpret_cast<const void*>(&k_cc3d<2, kCCTX, kCCTY, double>),
                                             kCCTX * kCCTY, ANTS_CC_SLOTS_PER_CU);
  static const int64_t slots4n = device_slots(reinterpret_cast<const void*>(&k_cc3d<4, kCCTX, kCCTY / 2, double>),
                                              kCCTX * kCCTY / 2, ANTS_CC_SLOTS_PER_CU);
  static const int64_t slots2n = device_slots(reinterpret_cast<const void*>(&k_cc3d<2, kCCTX, kCCTY / 2, double>),
                                              kCCTX * kCCTY / 2, ANTS_CC_SLOTS_PER_CU);
  static const int64_t slots4f = device_slots(reinterpret_cast<const void*>(&k_cc3d<4, kCCTX, kCCTY, float>),
                                              kCCTX * kCCTY, ANTS_CC_SLOTS_PER_CU);
  static const int64_t slots2f = device_slots(reinterpret_cast<const void*>(&k_cc3d<2, kCCTX, kCCTY, float>),
                                              kCCTX * kCCTY, ANTS_CC_SLOTS_PER_CU);
  const int64_t slots = radius == 4 ? (!d ? slots4f : narrow ? slots4n : slots4)
                                    : (!d ? slots2f : narrow ? slots2n : slots2);
  sh.zc = zchunk_for(g.n[2], tiles, 2 * radius, slots);
#else
  int64_t zc = (static_cast<int64_t>(g.n[2]) * tiles) / ANTS_CC_TARGET_BLOCKS;
  zc = zc < ANTS_CC_ZMIN ? ANTS_CC_ZMIN : (zc > 32 ? 32 : zc);
  sh.zc = static_cast<int>(zc);
#endif
  return sh;
}

}  // namespace

// ---------------------------------------------------------------------------
void launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_stride_blocks(n)), dim3(kBlock), 0, s, p, v, n);
}

void launch_minmax(const float* in, int64_t n, uint32_t* out2, hipStream_t s) {
  hipLaunchKernelGGL(k_minmax, dim3(grid_stride_blocks(n)), dim3(kBlock), 0, s, in, n, out2);
}

void launch_window(const float* in, float* out, int64_t n, float lo, float hi, double scale, double shift,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_window, dim3(grid_stride_blocks(n)), dim3(kBlock), 0, s, in, out, n, lo, hi, scale, shift);
}

void launch_gauss_axis(const GridDev& g, int ncomp, const float* in, float* out, const Taps& t, int axis,
                       hipStream_t s) {
  const dim3 grid(blocks_for(g.count()));
  if (ncomp == 1)
    hipLaunchKernelGGL(k_gauss_axis<1>, grid, dim3(kBlock), 0, s, g, in, out, t, axis);
  else if (ncomp == 2)
    hipLaunchKernelGGL(k_gauss_axis<2>, grid, dim3(kBlock), 0, s, g, in, out, t, axis);
  else
    hipLaunchKernelGGL(k_gauss_axis<3>, grid, dim3(kBlock), 0, s, g, in, out, t, axis);
}

void launch_gauss_field_final(const GridDev& g, int np, const float* const* in, const float* const* orig,
                              float* const* out, const Taps& t, int axis, float w1, float w2, bool boundary,
                              unsigned int* maxbits, hipStream_t s) {
  const dim3 grid(red_blocks(g.count()), np);
  if (g.D == 2)
    hipLaunchKernelGGL(k_gauss_field_final<2>, grid, dim3(kBlock), 0, s, g, pack(in, np), pack(orig, np),
                       mpack(out, np), t, axis, w1, w2, static_cast<int>(boundary), maxbits);
  else
    hipLaunchKernelGGL(k_gauss_field_final<3>, grid, dim3(kBlock), 0, s, g, pack(in, np), pack(orig, np),
                       mpack(out, np), t, axis, w1, w2, static_cast<int>(boundary), maxbits);
}

// The fused field smoothing with tile TX x TY and kernel radius R; z-chunk
// tuned (radius 6, the SyN default variance 3) or modelled.
template <int TX, int TY, int R = 6>
static void smooth3d_fields(const GridDev& g, int np, const float* const* in, const float* const* orig,
                            float* const* out, const Taps& t, float w1, float w2, bool boundary,
                            unsigned int* maxbits, hipStream_t s) {
  const int64_t tiles = static_cast<int64_t>((g.n[0] + TX - 1) / TX) * ((g.n[1] + TY - 1) / TY) * np;
  int64_t zc = R == 6 ? tune_get(kTuneSmoothZChunk, g.count()) : 0;
  if (zc > 0) {
    zc = zc < g.n[2] ? zc : g.n[2];
  } else {
#if ANTS_ZCHUNK_MODEL
    static const int64_t slots =
        device_slots(reinterpret_cast<const void*>(&k_smooth3d<R, TX, TY>), TX * TY, ANTS_SMOOTH_SLOTS_PER_CU);
    zc = zchunk_for(g.n[2], tiles, 2 * R, slots);
#else
    zc = (static_cast<int64_t>(g.n[2]) * tiles) / ANTS_SMOOTH_TARGET_BLOCKS;
    zc = zc < ANTS_SMOOTH_ZMIN ? ANTS_SMOOTH_ZMIN : (zc > 64 ? 64 : zc);
#endif
  }
  const int nzc = static_cast<int>((g.n[2] + zc - 1) / zc);
  const dim3 grid((g.n[0] + TX - 1) / TX, (g.n[1] + TY - 1) / TY, nzc * np);
  hipLaunchKernelGGL((k_smooth3d<R, TX, TY>), grid, dim3(TX * TY), 0, s, g, pack(in, np), pack(orig, np),
                     mpack(out, np), t, static_cast<int>(zc), nzc, w1, w2, static_cast<int>(boundary), maxbits);
}

bool launch_smooth3d(const GridDev& g, int np, const float* const* in, const float* const* orig, float* const* out,
                     const Taps& t, float w1, float w2, bool boundary, unsigned int* maxbits, hipStream_t s) {
  if (g.D != 3 || t.r < 1 || t.r > 6 || np > 4) return false;
  if (static_cast<int64_t>(g.n[0]) * g.n[1] * 3 >= (int64_t{1} << 31)) return false;  // 32-bit in-plane offsets
  constexpr int TX = kSmoothTX, TY = kSmoothTY;
  switch (t.r) {
    // other variances (field models' update / total / velocity smoothing,
    // SyN with another update variance): the 16 x 32 tile, modelled z-chunks
    case 1: smooth3d_fields<TX, TY, 1>(g, np, in, orig, out, t, w1, w2, boundary, maxbits, s); return true;
    case 2: smooth3d_fields<TX, TY, 2>(g, np, in, orig, out, t, w1, w2, boundary, maxbits, s); return true;
    case 3: smooth3d_fields<TX, TY, 3>(g, np, in, orig, out, t, w1, w2, boundary, maxbits, s); return true;
    case 4: smooth3d_fields<TX, TY, 4>(g, np, in, orig, out, t, w1, w2, boundary, maxbits, s); return true;
    case 5: smooth3d_fields<TX, TY, 5>(g, np, in, orig, out, t, w1, w2, boundary, maxbits, s); return true;
    default: break;
  }
  if (tune_get(kTuneSmoothTile, g.count()) == 2)
    smooth3d_fields<TX, TY / 2>(g, np, in, orig, out, t, w1, w2, boundary, maxbits, s);
  else
    smooth3d_fields<TX, TY>(g, np, in, orig, out, t, w1, w2, boundary, maxbits, s);
  return true;
}

template <int R>
static void smooth3d_image_r(const GridDev& g, const float* in, float* out, const Taps& t, hipStream_t s) {
  constexpr int TX = kSmoothTX, TY = kSmoothTY;
  const int64_t tiles = static_cast<int64_t>((g.n[0] + TX - 1) / TX) * ((g.n[1] + TY - 1) / TY);
  static const int64_t slots = device_slots(reinterpret_cast<const void*>(&k_smooth3d<R, TX, TY, 1>), TX * TY);
  const int64_t zc = zchunk_for(g.n[2], tiles, 2 * R, slots);
  const int nzc = static_cast<int>((g.n[2] + zc - 1) / zc);
  const dim3 grid((g.n[0] + TX - 1) / TX, (g.n[1] + TY - 1) / TY, nzc);
  const float* ins[1] = {in};
  float* outs[1] = {out};
  hipLaunchKernelGGL((k_smooth3d<R, TX, TY, 1>), grid, dim3(TX * TY), 0, s, g, pack(ins, 1), pack(ins, 1),
                     mpack(outs, 1), t, static_cast<int>(zc), nzc, 1.f, 0.f, 0, nullptr);
}

bool launch_smooth3d_image(const GridDev& g, const float* in, float* out, const Taps& t, hipStream_t s) {
  if (g.D != 3 || in == out) return false;
  if (static_cast<int64_t>(g.n[0]) * g.n[1] >= (int64_t{1} << 31)) return false;  // 32-bit in-plane offsets
  switch (t.r) {
    case 1: smooth3d_image_r<1>(g, in, out, t, s); return true;
    case 2: smooth3d_image_r<2>(g, in, out, t, s); return true;
    case 3: smooth3d_image_r<3>(g, in, out, t, s); return true;
    case 4: smooth3d_image_r<4>(g, in, out, t, s); return true;
    case 5: smooth3d_image_r<5>(g, in, out, t, s); return true;
    case 6: smooth3d_image_r<6>(g, in, out, t, s); return true;
    default: return false;
  }
}

template <class T>
static void warp_pair_t(const GridDev& gl, const GridDev& gF, const T* Fs, const float* uF, float* Fw,
                        const GridDev& gM, const T* Ms, const float* uM, float* Mw, hipStream_t s, uint32_t* clear,
                        int nclear, const WarpMap* maps) {
  const dim3 grid(blocks_for(gl.count()));
  const WarpMap none{};
  const WarpMap& mF = maps ? maps[0] : none;
  const WarpMap& mM = maps ? maps[1] : none;
  if (gl.D == 2) {
    if (maps)
      hipLaunchKernelGGL((k_warp_pair<2, T, true>), grid, dim3(kBlock), 0, s, gl, gF, Fs, uF, Fw, gM, Ms, uM, Mw,
                         clear, nclear, mF, mM);
    else
      hipLaunchKernelGGL((k_warp_pair<2, T, false>), grid, dim3(kBlock), 0, s, gl, gF, Fs, uF, Fw, gM, Ms, uM, Mw,
                         clear, nclear, mF, mM);
  } else {
    if (maps)
      hipLaunchKernelGGL((k_warp_pair<3, T, true>), grid, dim3(kBlock), 0, s, gl, gF, Fs, uF, Fw, gM, Ms, uM, Mw,
                         clear, nclear, mF, mM);
    else
      hipLaunchKernelGGL((k_warp_pair<3, T, false>), grid, dim3(kBlock), 0, s, gl, gF, Fs, uF, Fw, gM, Ms, uM, Mw,
                         clear, nclear, mF, mM);
  }
}

void launch_warp_pair(const GridDev& gl, const GridDev& gF, const float* Fs, const float* uF, float* Fw,
                      const GridDev& gM, const float* Ms, const float* uM, float* Mw, hipStream_t s,
                      uint32_t* clear, int nclear, const WarpMap* maps) {
  warp_pair_t<float>(gl, gF, Fs, uF, Fw, gM, Ms, uM, Mw, s, clear, nclear, maps);
}

void launch_warp_pair_bf16(const GridDev& gl, const GridDev& gF, const uint16_t* Fs, const float* uF, float* Fw,
                           const GridDev& gM, const uint16_t* Ms, const float* uM, float* Mw, hipStream_t s,
                           uint32_t* clear, int nclear, const WarpMap* maps) {
  warp_pair_t<uint16_t>(gl, gF, Fs, uF, Fw, gM, Ms, uM, Mw, s, clear, nclear, maps);
}

void launch_to_bf16(const float* in, uint16_t* out, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_to_bf16, dim3(grid_stride_blocks(n)), dim3(kBlock), 0, s, in, out, n);
}

void launch_copy_calibration(const void* in, void* out, int64_t n4, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_calibration, dim3(grid_stride_blocks(n4)), dim3(kBlock), 0, s,
                     static_cast<const float4*>(in), static_cast<float4*>(out), n4);
}

void launch_warp_image(const GridDev& gi, const float* img, const GridDev& go, const float* u, float* out,
                       hipStream_t s) {
  const dim3 grid(blocks_for(go.count()));
  if (go.D == 2)
    hipLaunchKernelGGL(k_warp_image<2>, grid, dim3(kBlock), 0, s, gi, img, go, u, out);
  else
    hipLaunchKernelGGL(k_warp_image<3>, grid, dim3(kBlock), 0, s, gi, img, go, u, out);
}

static bool cc_tiled(const GridDev& g, int radius) {
  // the tiled kernel keeps 32-bit in-plane offsets
  return g.D == 3 && (radius == 2 || radius == 4) && static_cast<int64_t>(g.n[0]) * g.n[1] < (int64_t{1} << 31);
}

int64_t cc_partials_needed(const GridDev& g, int radius, bool fp32_sums) {
  if (cc_tiled(g, radius)) {
    const CCShape sh = cc_shape(g, radius, fp32_sums);
    return static_cast<int64_t>((g.n[0] + sh.tx - 1) / sh.tx) * ((g.n[1] + sh.ty - 1) / sh.ty) *
           ((g.n[2] + sh.zc - 1) / sh.zc);
  }
  return red_blocks(g.count());
}

int launch_cc(const GridDev& g, const float* F, const float* M, int radius, float eps, float* gF2M, float* gM2M,
              Partial* partials, RedSlot* slot, hipStream_t s, bool reduce, bool fp32_sums, const float* gext) {
  if (!cc_tiled(g, radius))
    return launch_cc_direct(g, F, M, radius, eps, gF2M, gM2M, partials, slot, s, reduce, gext);
  const CCShape sh = cc_shape(g, radius, fp32_sums && !gext);
  const int zc = sh.zc;
  const dim3 grid((g.n[0] + sh.tx - 1) / sh.tx, (g.n[1] + sh.ty - 1) / sh.ty, (g.n[2] + zc - 1) / zc);
  const bool d = gext || cc_sums_fp64(fp32_sums);
  const bool narrow = sh.ty != kCCTY;
  constexpr int TY2 = kCCTY / 2;
  if (gext) {
    // field transforms: double sums, moving gradient from gext
    if (radius == 4 && narrow)
      hipLaunchKernelGGL((k_cc3d<4, kCCTX, TY2, double, true>), grid, dim3(kCCTX * TY2), 0, s, g, F, M, eps, zc, gF2M,
                         gM2M, partials, gext);
    else if (radius == 4)
      hipLaunchKernelGGL((k_cc3d<4, kCCTX, kCCTY, double, true>), grid, dim3(kCCTX * kCCTY), 0, s, g, F, M, eps, zc,
                         gF2M, gM2M, partials, gext);
    else if (narrow)
      hipLaunchKernelGGL((k_cc3d<2, kCCTX, TY2, double, true>), grid, dim3(kCCTX * TY2), 0, s, g, F, M, eps, zc, gF2M,
                         gM2M, partials, gext);
    else
      hipLaunchKernelGGL((k_cc3d<2, kCCTX, kCCTY, double, true>), grid, dim3(kCCTX * kCCTY), 0, s, g, F, M, eps, zc,
                         gF2M, gM2M, partials, gext);
  } else if (radius == 4 && d && narrow)
    hipLaunchKernelGGL((k_cc3d<4, kCCTX, TY2, double>), grid, dim3(kCCTX * TY2), 0, s, g, F, M, eps, zc, gF2M, gM2M,
                       partials, nullptr);
  else if (radius == 4 && d)
    hipLaunchKernelGGL((k_cc3d<4, kCCTX, kCCTY, double>), grid, dim3(kCCTX * kCCTY), 0, s, g, F, M, eps, zc, gF2M,
                       gM2M, partials, nullptr);
  else if (radius == 4)
    hipLaunchKernelGGL((k_cc3d<4, kCCTX, kCCTY, float>), grid, dim3(kCCTX * kCCTY), 0, s, g, F, M, eps, zc, gF2M,
                       gM2M, partials, nullptr);
  else if (d && narrow)
    hipLaunchKernelGGL((k_cc3d<2, kCCTX, TY2, double>), grid, dim3(kCCTX * TY2), 0, s, g, F, M, eps, zc, gF2M, gM2M,
                       partials, nullptr);
  else if (d)
    hipLaunchKernelGGL((k_cc3d<2, kCCTX, kCCTY, double>), grid, dim3(kCCTX * kCCTY), 0, s, g, F, M, eps, zc, gF2M,
                       gM2M, partials, nullptr);
  else
    hipLaunchKernelGGL((k_cc3d<2, kCCTX, kCCTY, float>), grid, dim3(kCCTX * kCCTY), 0, s, g, F, M, eps, zc, gF2M,
                       gM2M, partials, nullptr);
  const int np = static_cast<int>(grid.x * grid.y * grid.z);
  if (reduce) hipLaunchKernelGGL(k_reduce_slot<kBlock>, dim3(1), dim3(kBlock), 0, s, partials, np, slot);
  return np;
}

int launch_cc_direct(const GridDev& g, const float* F, const float* M, int radius, float eps, float* gF2M,
                     float* gM2M, Partial* partials, RedSlot* slot, hipStream_t s, bool reduce, const float* gext) {
  const dim3 grid(red_blocks(g.count()));
  if (g.D == 2 && gext)
    hipLaunchKernelGGL((k_cc_direct<2, true>), grid, dim3(kBlock), 0, s, g, F, M, radius, eps, gF2M, gM2M, partials,
                       gext);
  else if (gext)
    hipLaunchKernelGGL((k_cc_direct<3, true>), grid, dim3(kBlock), 0, s, g, F, M, radius, eps, gF2M, gM2M, partials,
                       gext);
  else if (g.D == 2)
    hipLaunchKernelGGL(k_cc_direct<2>, grid, dim3(kBlock), 0, s, g, F, M, radius, eps, gF2M, gM2M, partials, nullptr);
  else
    hipLaunchKernelGGL(k_cc_direct<3>, grid, dim3(kBlock), 0, s, g, F, M, radius, eps, gF2M, gM2M, partials, nullptr);
  const int np = static_cast<int>(grid.x);
  if (reduce) hipLaunchKernelGGL(k_reduce_slot<kBlock>, dim3(1), dim3(kBlock), 0, s, partials, np, slot);
  return np;
}

void launch_publish_stats(const Partial* cc_parts, int cc_n, RedSlot* energy, const uint32_t* src, uint32_t* dst,
                          int nwords, uint32_t* flag, hipStream_t s) {
  static_assert(kCCTX * kCCTY == kBlock, "CC partials are reduced in the order of k_reduce_slot<kBlock>");
  hipLaunchKernelGGL(k_publish_stats<kBlock>, dim3(1), dim3(kBlock), 0, s, cc_parts, cc_n, energy, src, dst, nwords,
                     flag);
}

// Strided pass along `axis` (any axis): see k_bsp_strided.
static void bsp_strided(const BspDev& T, int np, const float* const* in, float* const* out, const int dims[3],
                        int axis, bool fit, bool last, hipStream_t s) {
  const BspAxisDev& A = T.ax[axis];
  int64_t I = T.D, O = 1;
  for (int d = 0; d < axis; ++d) I *= dims[d];
  for (int d = axis + 1; d < T.D; ++d) O *= dims[d];
  const int Lin = dims[axis], Lout = fit ? A.m : A.n;
  const dim3 block(64, 4);
  const int m0 = T.ax[0].m;
  // blockIdx.z covers (field, outer index) pairs: outer chunks within the grid limit
  const int64_t chunk = std::max<int64_t>(1, 65535 / np);
  for (int64_t ob = 0; ob < O; ob += chunk) {
    const int oc = static_cast<int>(std::min<int64_t>(chunk, O - ob));
    const dim3 grid(static_cast<unsigned int>((I + 63) / 64),
                    static_cast<unsigned int>((Lout + 4 * kBspKPT - 1) / (4 * kBspKPT)),
                    static_cast<unsigned int>(oc * np));
    const int lst = static_cast<int>(last);
    if (T.D == 2) {
      if (fit)
        hipLaunchKernelGGL((k_bsp_strided<2, true>), grid, block, 0, s, A, T, pack(in, np), mpack(out, np),
                           static_cast<int>(I), Lin, Lout, oc, static_cast<int>(ob), m0, lst);
      else
        hipLaunchKernelGGL((k_bsp_strided<2, false>), grid, block, 0, s, A, T, pack(in, np), mpack(out, np),
                           static_cast<int>(I), Lin, Lout, oc, static_cast<int>(ob), m0, 0);
    } else {
      if (fit)
        hipLaunchKernelGGL((k_bsp_strided<3, true>), grid, block, 0, s, A, T, pack(in, np), mpack(out, np),
                           static_cast<int>(I), Lin, Lout, oc, static_cast<int>(ob), m0, lst);
      else
        hipLaunchKernelGGL((k_bsp_strided<3, false>), grid, block, 0, s, A, T, pack(in, np), mpack(out, np),
                           static_cast<int>(I), Lin, Lout, oc, static_cast<int>(ob), m0, 0);
    }
  }
}

void launch_bsp_fit_axis(const BspDev& T, int np, const float* const* in, float* const* out, const int dims[3],
                         int axis, bool last, hipStream_t s) {
  const BspAxisDev& A = T.ax[axis];
  const int64_t row_bytes = static_cast<int64_t>(A.n) * T.D * 4;
  const int64_t table_bytes = static_cast<int64_t>(A.gtot) * 4 + (2 * static_cast<int64_t>(A.m) + 1) * 4;
  if (axis != 0 || last || table_bytes + row_bytes > 65536) {
    bsp_strided(T, np, in, out, dims, axis, true, last, s);
    return;
  }
  const int rows = dims[1] * dims[2];
  const int R = static_cast<int>(std::max<int64_t>(
      1, std::min<int64_t>(8, (std::max<int64_t>(16384 - table_bytes, row_bytes)) / row_bytes)));
  const dim3 grid(static_cast<unsigned int>((rows + R - 1) / R), np);
  const size_t lds = static_cast<size_t>(table_bytes + R * row_bytes);
  if (T.D == 2)
    hipLaunchKernelGGL(k_bsp_fit_x<2>, grid, dim3(256), lds, s, A, pack(in, np), mpack(out, np), rows, R);
  else
    hipLaunchKernelGGL(k_bsp_fit_x<3>, grid, dim3(256), lds, s, A, pack(in, np), mpack(out, np), rows, R);
}

void launch_bsp_eval_axis(const BspDev& T, int np, const float* const* in, float* const* out, const int dims[3],
                          int axis, const float sp[3], unsigned int* maxbits, float* part, hipStream_t s) {
  if (axis != 0) {
    bsp_strided(T, np, in, out, dims, axis, false, false, s);
    return;
  }
  const BspAxisDev& A = T.ax[0];
  const int64_t nvox = static_cast<int64_t>(A.n) * dims[1] * dims[2];  // < 2^31 (checked with the tables)
  const unsigned int blocks = static_cast<unsigned int>(std::min<int64_t>((nvox + 255) / 256, kBspEvalBlocks));
  const dim3 grid(blocks, np);
  float* pp = maxbits ? part : nullptr;
  if (T.D == 2)
    hipLaunchKernelGGL(k_bsp_eval_x<2>, grid, dim3(256), 0, s, A, pack(in, np), mpack(out, np), static_cast<int>(nvox),
                       T.q, sp[0], sp[1], sp[2], pp);
  else
    hipLaunchKernelGGL(k_bsp_eval_x<3>, grid, dim3(256), 0, s, A, pack(in, np), mpack(out, np), static_cast<int>(nvox),
                       T.q, sp[0], sp[1], sp[2], pp);
  if (maxbits)
    hipLaunchKernelGGL(k_max_partials, dim3(np), dim3(1024), 0, s, part, static_cast<int>(blocks), maxbits);
}

int64_t bsp_partials_needed() { return kBspEvalBlocks; }

int launch_moments(const GridDev& g, const float* img, double* out, hipStream_t s) {
  const unsigned int blocks = red_blocks(g.count());
  hipLaunchKernelGGL(k_moments, dim3(blocks), dim3(kBlock), 0, s, g, img, out);
  return static_cast<int>(blocks);
}

bool launch_bsp_eval_yx(const BspDev& T, int np, const float* const* in, float* const* out, int n2,
                        const float sp[3], unsigned int* maxbits, float* part, hipStream_t s) {
  const BspAxisDev &AX = T.ax[0], &AY = T.ax[1];
  const int64_t per_row = static_cast<int64_t>(AX.m) * T.D;
  const int spans = AY.m - T.q;
  // enough rows per block that the partials fit; the control-point rows a block needs
  int R = 8;
  while ((static_cast<int64_t>(AY.n + R - 1) / R) * n2 > kBspEvalBlocks && R < AY.n) R *= 2;
  if ((static_cast<int64_t>(AY.n + R - 1) / R) * n2 > kBspEvalBlocks) return false;
  const int64_t KR = std::min<int64_t>(AY.m, (static_cast<int64_t>(R - 1) * spans) / std::max(AY.n - 1, 1) + T.q + 2);
  const int64_t lds = (5 * static_cast<int64_t>(AX.n) + (R + KR) * per_row) * 4;
  if (lds > 65536 || n2 > 65535) return false;
  const dim3 grid(static_cast<unsigned int>((AY.n + R - 1) / R), static_cast<unsigned int>(n2), np);
  float* pp = maxbits ? part : nullptr;
  if (T.D == 2)
    hipLaunchKernelGGL(k_bsp_eval_yx<2>, grid, dim3(256), lds, s, AX, AY, pack(in, np), mpack(out, np), R, T.q,
                       sp[0], sp[1], sp[2], pp);
  else
    hipLaunchKernelGGL(k_bsp_eval_yx<3>, grid, dim3(256), lds, s, AX, AY, pack(in, np), mpack(out, np), R, T.q,
                       sp[0], sp[1], sp[2], pp);
  if (maxbits)
    hipLaunchKernelGGL(k_max_partials, dim3(np), dim3(1024), 0, s, part, static_cast<int>(grid.x * grid.y), maxbits);
  return true;
}

void launch_compose(const GridDev& g, int np, const float* const* disp, const float* const* warp, float* const* out,
                    const unsigned int* maxbits, float rate, hipStream_t s) {
  const dim3 grid(blocks_for(g.count()), np);
  if (g.D == 2)
    hipLaunchKernelGGL(k_compose<2>, grid, dim3(kBlock), 0, s, g, pack(disp, np), pack(warp, np), mpack(out, np),
                       maxbits, rate);
  else
    hipLaunchKernelGGL(k_compose<3>, grid, dim3(kBlock), 0, s, g, pack(disp, np), pack(warp, np), mpack(out, np),
                       maxbits, rate);
}

int64_t ic_partials_needed(const GridDev& g) { return red_blocks(g.count()); }

void launch_ic_residual(const GridDev& g, const float* fwd, const float* inv, int margin, Partial* partials,
                        RedSlot* slot, hipStream_t s) {
  const dim3 grid(red_blocks(g.count()));
  if (g.D == 2)
    hipLaunchKernelGGL(k_ic_residual<2>, grid, dim3(kBlock), 0, s, g, fwd, inv, margin, partials);
  else
    hipLaunchKernelGGL(k_ic_residual<3>, grid, dim3(kBlock), 0, s, g, fwd, inv, margin, partials);
  hipLaunchKernelGGL(k_reduce_slot<kBlock>, dim3(1), dim3(kBlock), 0, s, partials, static_cast<int>(grid.x), slot);
}

// Measured per 192x224x192 pair: inversion 24.5 ms with one voxel per thread
// and 1024 blocks, 23.9 with 4 voxels, 23.7 with at most 768 blocks (the
// plain scheme); with speculation, 8 voxels per thread: at 96x112x96 the
// 2 x 252 blocks fit the 768 resident blocks (79 VGPRs: 3 blocks of 512
// threads per CU) in one round (19.9 -> 19.7 ms against 4 voxels).
#ifndef ANTS_INV_VPT
#define ANTS_INV_VPT 8
#endif
static unsigned int inv_blocks(int64_t n) {
  // a tuned block count for this grid size, when the table holds one
  const int tuned = tune_get(kTuneInvBlocks, n);
  if (tuned > 0) {
    const int64_t most = (n + 63) / 64;  // every block keeps at least one 64-voxel word
    int64_t b = tuned < 2048 ? tuned : 2048;
    b = b < most ? b : most;
    return static_cast<unsigned int>(b < 1 ? 1 : b);
  }
  int64_t b = (n + int64_t{kInvThreads} * ANTS_INV_VPT - 1) / (int64_t{kInvThreads} * ANTS_INV_VPT);
  // small grids keep up to 256 blocks (one per CU) at one voxel per thread
  const int64_t b1 = (n + kInvThreads - 1) / kInvThreads;
  const int64_t floor_b = b1 < 256 ? b1 : 256;
  b = b < floor_b ? floor_b : b;
  return static_cast<unsigned int>(b < 1 ? 1 : (b > kInvBlocks ? kInvBlocks : b));
}

int64_t inv_partials_needed(const GridDev& g) { return inv_blocks(g.count()); }

int64_t inv_mask_words(const GridDev& g) { return (g.count() + 63) / 64; }

// Speculative updates pay for the sample launch and the sampled gathers
// only on large grids (measured per 192x224x192 pair: the 192x224x192
// level's inversion 727 -> 555 us per iteration, the 96x112x96 level's
// 105 -> 113 us).
#ifndef ANTS_INV_SPEC_MIN_VOXELS
#define ANTS_INV_SPEC_MIN_VOXELS (int64_t{1} << 21)
#endif
// ANTS_SYN_INV_SPEC_MIN_VOXELS overrides the threshold (read per launch, so
// that a test can compare both schemes on one grid; results are the same
// bits either way).
bool inv_speculates(const GridDev& g) {
  const char* env = std::getenv("ANTS_SYN_INV_SPEC_MIN_VOXELS");
  const int64_t t = env ? std::atoll(env) : ANTS_INV_SPEC_MIN_VOXELS;
  return g.count() >= t;
}

void launch_inv_step(const GridDev& g, int np, const float* const* field, float* const* inv, RedSlot* const* slots,
                     Partial* partials, uint64_t* masks, int j, hipStream_t s, const InvGate& gate) {
  const dim3 grid(inv_blocks(g.count()), np);
  SlotPack4 sp{};
  for (int k = 0; k < np; ++k) sp.p[k] = slots[k];
  const int64_t stride = inv_partials_needed(g);
  const int64_t mstride = inv_mask_words(g);
  const int smp = inv_speculates(g) ? 1 : 0;
  if (j < 0 && !smp) return;  // no sample launch without speculation
  const dim3 grid_j = j < 0 ? dim3(static_cast<unsigned int>(inv_sample_blocks(grid.x)), np) : grid;
  const bool idx32 = g.count() * g.D < (int64_t{1} << 31);
  const char* strict = std::getenv("ANTS_SYN_INV_STRICT_TIE");
  const float tol = (strict && strict[0] == '1') ? kInvMaxTolStrict : kInvMaxTol;
  if (g.D == 2 && idx32)
    hipLaunchKernelGGL((k_inv_step<2, int>), grid_j, dim3(kInvThreads), 0, s, g, pack(field, np), mpack(inv, np), sp,
                       partials, stride, masks, mstride, j, smp, gate, tol);
  else if (g.D == 2)
    hipLaunchKernelGGL((k_inv_step<2, int64_t>), grid_j, dim3(kInvThreads), 0, s, g, pack(field, np), mpack(inv, np),
                       sp, partials, stride, masks, mstride, j, smp, gate, tol);
  else if (idx32)
    hipLaunchKernelGGL((k_inv_step<3, int>), grid_j, dim3(kInvThreads), 0, s, g, pack(field, np), mpack(inv, np), sp,
                       partials, stride, masks, mstride, j, smp, gate, tol);
  else
    hipLaunchKernelGGL((k_inv_step<3, int64_t>), grid_j, dim3(kInvThreads), 0, s, g, pack(field, np), mpack(inv, np),
                       sp, partials, stride, masks, mstride, j, smp, gate, tol);
}

void launch_resample_fields(const GridDev& gi, const GridDev& go, int np, const float* const* in, float* const* out,
                            hipStream_t s) {
  const dim3 grid(blocks_for(go.count()), np);
  if (go.D == 2)
    hipLaunchKernelGGL(k_resample_fields<2>, grid, dim3(kBlock), 0, s, gi, go, pack(in, np), mpack(out, np));
  else
    hipLaunchKernelGGL(k_resample_fields<3>, grid, dim3(kBlock), 0, s, gi, go, pack(in, np), mpack(out, np));
}

void launch_apply_transforms(const GridDev& go, const ApplyDev& P, int interp, const float* in, float* out,
                             hipStream_t s) {
  const dim3 grid(blocks_for(go.count()));
#define ANTS_APPLY_LAUNCH(DD, II) \
  hipLaunchKernelGGL((k_apply_transforms<DD, II>), grid, dim3(kBlock), 0, s, go, P, in, out)
#define ANTS_APPLY_DIM(DD)                   \
  switch (interp) {                          \
    case 0: ANTS_APPLY_LAUNCH(DD, 0); break;   \
    case 1: ANTS_APPLY_LAUNCH(DD, 1); break;   \
    case 2: ANTS_APPLY_LAUNCH(DD, 2); break;   \
    case 16: ANTS_APPLY_LAUNCH(DD, 16); break; \
    case 17: ANTS_APPLY_LAUNCH(DD, 17); break; \
    case 18: ANTS_APPLY_LAUNCH(DD, 18); break; \
    case 19: ANTS_APPLY_LAUNCH(DD, 19); break; \
    case 20: ANTS_APPLY_LAUNCH(DD, 20); break; \
    case 21: ANTS_APPLY_LAUNCH(DD, 21); break; \
    case 32: ANTS_APPLY_LAUNCH(DD, 32); break; \
    case 33: ANTS_APPLY_LAUNCH(DD, 33); break; \
    case 34: ANTS_APPLY_LAUNCH(DD, 34); break; \
    case 35: ANTS_APPLY_LAUNCH(DD, 35); break; \
    default: ANTS_APPLY_LAUNCH(DD, 36); break; \
  }
  if (go.D == 2) {
    ANTS_APPLY_DIM(2)
  } else {
    ANTS_APPLY_DIM(3)
  }
#undef ANTS_APPLY_DIM
#undef ANTS_APPLY_LAUNCH
}

// CreateJacobianDeterminantImage over a displacement field u (D physical
// components per voxel): F[i][j] = (-u_i(x+2e_j) + 8 u_i(x+e_j) - 8 u_i(x-e_j)
// + u_i(x-2e_j)) / 12 / spacing_j with neighbours clamped to the buffer
// (zero-flux Neumann), each row then rotated by the direction matrix
// (row_i <- R row_i), plus the identity; out = max(det F, 0), or
// log(max(det F, 0.001)) with `logj`; grad (nullable) receives F row-major.

template <int D>
__global__ void __launch_bounds__(kBlock) k_jacobian(GridDev g, JacobianDev J, const float* __restrict__ u,
                                                     int logj, float* __restrict__ out, float* __restrict__ grad) {
  const int64_t N = g.count();
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= N) return;
  int x, y, z;
  unlin(g, i, x, y, z);
  const int ix[3] = {x, y, z};
  double F[3][3];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    int64_t off[4];
    const int dd[4] = {2, 1, -1, -2};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      int q[3] = {ix[0], ix[1], ix[2]};
      q[j] = min(max(q[j] + dd[k], 0), J.n[j] - 1);
      off[k] = (static_cast<int64_t>(q[2]) * J.n[1] + q[1]) * J.n[0] + q[0];
    }
#pragma unroll
    for (int c = 0; c < D; ++c) {
      const double n2 = u[off[0] * D + c], n1 = u[off[1] * D + c], p1 = u[off[2] * D + c], p2 = u[off[3] * D + c];
      F[c][j] = J.isp[j] * (-n2 + 8.0 * n1 - 8.0 * p1 + p2) / 12.0;
    }
  }
  double G[3][3];
#pragma unroll
  for (int r = 0; r < D; ++r)
#pragma unroll
    for (int c = 0; c < D; ++c) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < D; ++k) acc = acc + J.R[c * 3 + k] * F[r][k];
      G[r][c] = acc + (r == c ? 1.0 : 0.0);
    }
  if (grad) {
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int c = 0; c < D; ++c) grad[i * D * D + r * D + c] = static_cast<float>(G[r][c]);
  }
  double det;
  if constexpr (D == 2) {
    det = G[0][0] * G[1][1] - G[0][1] * G[1][0];
  } else {
    det = G[0][0] * (G[1][1] * G[2][2] - G[1][2] * G[2][1]) - G[0][1] * (G[1][0] * G[2][2] - G[1][2] * G[2][0]) +
          G[0][2] * (G[1][0] * G[2][1] - G[1][1] * G[2][0]);
  }
  out[i] = static_cast<float>(logj ? log(fmax(fmax(det, 0.0), 0.001)) : fmax(det, 0.0));
}

// GeometricJacobianDeterminantImageFilter: the simplex around each voxel
// (triangle: index offsets (0, s), (-1/2, -s), (1/2, -s), s = sqrt(3)/4;
// tetrahedron: (1/2, 0, -t), (-1/2, 0, -t), (0, 1/2, t), (0, -1/2, t),
// t = 1/(2 sqrt 2)), each vertex moved by the field linearly interpolated
// there (corner indices clamped to the buffer), the moved simplex's
// |volume| over the unmoved one; log(max(., 0.001)) with `logj`.  Vertex
// positions are taken relative to the voxel's point (the common term
// cancels in the volume).
template <int D>
__global__ void __launch_bounds__(kBlock) k_jacobian_geometric(GridDev g, JacobianDev J, const float* __restrict__ u,
                                                               int logj, double undisplaced,
                                                               float* __restrict__ out) {
  const int64_t N = g.count();
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= N) return;
  int x, y, z;
  unlin(g, i, x, y, z);
  const int ix[3] = {x, y, z};
  constexpr int NV = D + 1;
  const double s3 = 0.25 * 1.7320508075688772935, t2 = 0.5 / 1.4142135623730950488;
  const double offs2[3][2] = {{0.0, s3}, {-0.5, -s3}, {0.5, -s3}};
  const double offs3[4][3] = {{0.5, 0.0, -t2}, {-0.5, 0.0, -t2}, {0.0, 0.5, t2}, {0.0, -0.5, t2}};
  double q[NV][3];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    double off[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < D; ++d) off[d] = D == 2 ? offs2[v][d] : offs3[v][d];
    // interpolate u at ix + off: per axis floor, fraction, clamped corners
    int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    double w[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double c = static_cast<double>(ix[d]) + off[d];
      const double fl = floor(c);
      w[d] = c - fl;
      const int64_t b = static_cast<int64_t>(fl);
      lo[d] = min(max(b, int64_t{0}), static_cast<int64_t>(J.n[d] - 1));
      hi[d] = min(max(b + 1, int64_t{0}), static_cast<int64_t>(J.n[d] - 1));
    }
    double disp[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < (1 << D); ++k) {
      double ov = 1.0;
      int64_t id[3] = {0, 0, 0};
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const bool up = (k >> d) & 1;
        ov = ov * (up ? w[d] : 1.0 - w[d]);
        id[d] = up ? hi[d] : lo[d];
      }
      const int64_t lin = (id[2] * J.n[1] + id[1]) * J.n[0] + id[0];
#pragma unroll
      for (int c = 0; c < D; ++c) disp[c] = disp[c] + ov * static_cast<double>(u[lin * D + c]);
    }
    // physical offset M off = R diag(spacing) off, plus the displacement
#pragma unroll
    for (int r = 0; r < D; ++r) {
      double m = 0.0;
#pragma unroll
      for (int c = 0; c < D; ++c) m = m + J.R[r * 3 + c] * (off[c] / J.isp[c]);
      q[v][r] = m + disp[r];
    }
  }
  double vol;
  if constexpr (D == 2) {
    const double abx = q[0][0] - q[1][0], aby = q[0][1] - q[1][1];
    const double acx = q[0][0] - q[2][0], acy = q[0][1] - q[2][1];
    vol = 0.5 * fabs(abx * acy - acx * aby);
  } else {
    double ad[3], bd[3], cd[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      ad[r] = q[0][r] - q[3][r];
      bd[r] = q[1][r] - q[3][r];
      cd[r] = q[2][r] - q[3][r];
    }
    const double cx = bd[1] * cd[2] - bd[2] * cd[1], cy = bd[2] * cd[0] - bd[0] * cd[2],
                 cz = bd[0] * cd[1] - bd[1] * cd[0];
    vol = fabs(ad[0] * cx + ad[1] * cy + ad[2] * cz) / 6.0;
  }
  const double ratio = vol / undisplaced;
  out[i] = static_cast<float>(logj ? log(fmax(ratio, 0.001)) : ratio);
}

void launch_jacobian_geometric(const GridDev& g, const JacobianDev& J, const float* u, int logj, double undisplaced,
                               float* out, hipStream_t s) {
  const dim3 grid(blocks_for(g.count()));
  if (g.D == 2)
    hipLaunchKernelGGL(k_jacobian_geometric<2>, grid, dim3(kBlock), 0, s, g, J, u, logj, undisplaced, out);
  else
    hipLaunchKernelGGL(k_jacobian_geometric<3>, grid, dim3(kBlock), 0, s, g, J, u, logj, undisplaced, out);
}

void launch_jacobian(const GridDev& g, const JacobianDev& J, const float* u, int logj, float* out, float* grad,
                     hipStream_t s) {
  const dim3 grid(blocks_for(g.count()));
  if (g.D == 2)
    hipLaunchKernelGGL(k_jacobian<2>, grid, dim3(kBlock), 0, s, g, J, u, logj, out, grad);
  else
    hipLaunchKernelGGL(k_jacobian<3>, grid, dim3(kBlock), 0, s, g, J, u, logj, out, grad);
}

void launch_transform_points(int D, const ApplyDev& P, const double* in, double* out, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  const dim3 grid(blocks_for(n));
  if (D == 2)
    hipLaunchKernelGGL(k_transform_points<2>, grid, dim3(kBlock), 0, s, P, in, out, n);
  else
    hipLaunchKernelGGL(k_transform_points<3>, grid, dim3(kBlock), 0, s, P, in, out, n);
}

void launch_compose_field(const GridDev& go, const ApplyDev& P, float* out, hipStream_t s) {
  const dim3 grid(blocks_for(go.count()));
  if (go.D == 2)
    hipLaunchKernelGGL(k_compose_field<2>, grid, dim3(kBlock), 0, s, go, P, out);
  else
    hipLaunchKernelGGL(k_compose_field<3>, grid, dim3(kBlock), 0, s, go, P, out);
}

void launch_bspline_coefficients(const int n[3], int D, const float* img, double* coef, int order, hipStream_t s) {
  const int64_t N = static_cast<int64_t>(n[0]) * n[1] * (D == 3 ? n[2] : 1);
  // poles of the B-spline of each order (Unser 1999, table 1)
  DecompPoles P{};
  switch (order) {
    case 2: P.z[0] = std::sqrt(8.0) - 3.0; P.np = 1; break;
    case 3: P.z[0] = std::sqrt(3.0) - 2.0; P.np = 1; break;
    case 4:
      P.z[0] = std::sqrt(664.0 - std::sqrt(438976.0)) + std::sqrt(304.0) - 19.0;
      P.z[1] = std::sqrt(664.0 + std::sqrt(438976.0)) - std::sqrt(304.0) - 19.0;
      P.np = 2;
      break;
    case 5:
      P.z[0] = std::sqrt(135.0 / 2.0 - std::sqrt(17745.0 / 4.0)) + std::sqrt(105.0 / 4.0) - 13.0 / 2.0;
      P.z[1] = std::sqrt(135.0 / 2.0 + std::sqrt(17745.0 / 4.0)) - std::sqrt(105.0 / 4.0) - 13.0 / 2.0;
      P.np = 2;
      break;
    default: break;  // orders 0 and 1 interpolate the samples themselves
  }
  P.gain = 1.0;
  for (int p = 0; p < P.np; ++p) {
    P.gain = P.gain * (1.0 - P.z[p]) * (1.0 - 1.0 / P.z[p]);
    // causal initialisation: geometric sum truncated where |z|^n < 1e-10
    P.horizon[p] = static_cast<int64_t>(std::ceil(std::log(1e-10) / std::log(std::fabs(P.z[p]))));
  }
  bool converted = false;
  for (int d = 0; d < D && P.np > 0; ++d) {
    const int64_t len = n[d];
    if (len < 2) continue;
    int64_t inner = 1;
    for (int k = 0; k < d; ++k) inner *= n[k];
    const int64_t lines = N / len;
    // widest tile of 64..8 lines within 64 KiB of LDS
    const int64_t row = d == 0 ? len + 1 : len;
    int TL = 64;
    while (TL >= 8 && TL * row * 8 > 65536) TL /= 2;
    const float* fin = converted ? nullptr : img;
    if (TL < 8) {
      // lines too long for LDS: one thread per line in global memory
      if (!converted)
        hipLaunchKernelGGL(k_to_double, dim3(blocks_for(N)), dim3(kBlock), 0, s, img, coef, N);
      hipLaunchKernelGGL(k_bspline_decompose, dim3(blocks_for(lines)), dim3(kBlock), 0, s, coef, lines, len, inner,
                         inner * len, P);
    } else {
      const int64_t tiles = d == 0 ? (lines + TL - 1) / TL : (N / (inner * len)) * ((inner + TL - 1) / TL);
      const dim3 grid(static_cast<unsigned int>(tiles));
      const size_t lds = static_cast<size_t>(TL * row * 8);
#define ANTS_DECOMP(XX, TT)                                                                                    \
  hipLaunchKernelGGL((k_bspline_decompose_lds<XX, TT>), grid, dim3(64), lds, s, fin, coef, lines, len,        \
                     XX ? int64_t{1} : inner, XX ? len : inner * len, P)
      if (d == 0) {
        if (TL == 64) ANTS_DECOMP(true, 64);
        else if (TL == 32) ANTS_DECOMP(true, 32);
        else if (TL == 16) ANTS_DECOMP(true, 16);
        else ANTS_DECOMP(true, 8);
      } else {
        if (TL == 64) ANTS_DECOMP(false, 64);
        else if (TL == 32) ANTS_DECOMP(false, 32);
        else if (TL == 16) ANTS_DECOMP(false, 16);
        else ANTS_DECOMP(false, 8);
      }
#undef ANTS_DECOMP
    }
    converted = true;
  }
  if (!converted) hipLaunchKernelGGL(k_to_double, dim3(blocks_for(N)), dim3(kBlock), 0, s, img, coef, N);
}

void launch_labels_nn(const GridDev& go, const float* u, const GridDev& gl, const int32_t* labels, int32_t* out,
                      hipStream_t s) {
  const dim3 grid(blocks_for(go.count()));
  if (go.D == 2)
    hipLaunchKernelGGL(k_labels_nn<2>, grid, dim3(kBlock), 0, s, go, u, gl, labels, out);
  else
    hipLaunchKernelGGL(k_labels_nn<3>, grid, dim3(kBlock), 0, s, go, u, gl, labels, out);
}

int lin_quantities(int mode, bool deriv, int P) {
  if (mode == kLinMS) return deriv ? 2 + P : 2;
  if (mode == kLinGC) return deriv ? 6 + 3 * P : 6;
  if (mode == kLinHist) return 1;
  return 1 + P;
}

int64_t lin_blocks(int64_t npts) {
  const int64_t b = (npts + kLinThreads - 1) / kLinThreads;
  return b < 1 ? 1 : (b > kLinMaxBlocks ? kLinMaxBlocks : b);
}

template <int D, int P>
static void lin_launch_p(const LinDev& L, const LinJac& J, int mode, bool deriv, int nb, double* partials,
                         unsigned long long* hist, unsigned int* fcnt, hipStream_t s) {
  const dim3 grid(nb), block(kLinThreads);
  if (!deriv) {
    if (mode == kLinMS) hipLaunchKernelGGL((k_lin_metric<D, kLinMS, false, 0>), grid, block, 0, s, L, J, partials, hist, fcnt);
    else if (mode == kLinGC) hipLaunchKernelGGL((k_lin_metric<D, kLinGC, false, 0>), grid, block, 0, s, L, J, partials, hist, fcnt);
    else hipLaunchKernelGGL((k_lin_metric<D, kLinHist, false, 0>), grid, block, 0, s, L, J, partials, hist, fcnt);
    return;
  }
  if (mode == kLinMS) hipLaunchKernelGGL((k_lin_metric<D, kLinMS, true, P>), grid, block, 0, s, L, J, partials, hist, fcnt);
  else if (mode == kLinGC) hipLaunchKernelGGL((k_lin_metric<D, kLinGC, true, P>), grid, block, 0, s, L, J, partials, hist, fcnt);
  else hipLaunchKernelGGL((k_lin_metric<D, kLinHistDeriv, true, P>), grid, block, 0, s, L, J, partials, hist, fcnt);
}

bool launch_lin_metric(const LinDev& L, const LinJac& J, int D, int P, int mode, bool deriv, double* partials,
                       double* out, unsigned long long* hist, unsigned int* fcnt, hipStream_t s) {
  if (mode == kLinHist) deriv = false;
  if (mode == kLinHistDeriv) deriv = true;
  const int64_t npts = L.pts ? L.npts : static_cast<int64_t>(L.vg.n[0]) * L.vg.n[1] * L.vg.n[2];
  const int nb = static_cast<int>(lin_blocks(npts));
  const int Q = lin_quantities(mode, deriv, P);
  if (D == 2) {
    if (!deriv || P == 2) lin_launch_p<2, 2>(L, J, mode, deriv, nb, partials, hist, fcnt, s);
    else if (P == 3) lin_launch_p<2, 3>(L, J, mode, deriv, nb, partials, hist, fcnt, s);
    else if (P == 6) lin_launch_p<2, 6>(L, J, mode, deriv, nb, partials, hist, fcnt, s);
    else return false;
  } else {
    if (!deriv || P == 3) lin_launch_p<3, 3>(L, J, mode, deriv, nb, partials, hist, fcnt, s);
    else if (P == 6) lin_launch_p<3, 6>(L, J, mode, deriv, nb, partials, hist, fcnt, s);
    else if (P == 12) lin_launch_p<3, 12>(L, J, mode, deriv, nb, partials, hist, fcnt, s);
    else return false;
  }
  hipLaunchKernelGGL(k_lin_reduce, dim3(1), dim3(256), 0, s, partials, nb, Q, out);
  return true;
}


void launch_field_sample(const GeomDev& vg, const GeomDev& ig, const float* img, const float* u, const FieldMap& map,
                         float* out, float* grad, hipStream_t s, const float* u2, float s2) {
  const int64_t n = static_cast<int64_t>(vg.n[0]) * vg.n[1] * (vg.D == 3 ? vg.n[2] : 1);
  const dim3 grid(blocks_for(n));
  if (vg.D == 2 && grad)
    hipLaunchKernelGGL((k_field_sample<2, true>), grid, dim3(kBlock), 0, s, vg, ig, img, u, u2, s2, map, out, grad);
  else if (vg.D == 2)
    hipLaunchKernelGGL((k_field_sample<2, false>), grid, dim3(kBlock), 0, s, vg, ig, img, u, u2, s2, map, out, grad);
  else if (grad)
    hipLaunchKernelGGL((k_field_sample<3, true>), grid, dim3(kBlock), 0, s, vg, ig, img, u, u2, s2, map, out, grad);
  else
    hipLaunchKernelGGL((k_field_sample<3, false>), grid, dim3(kBlock), 0, s, vg, ig, img, u, u2, s2, map, out, grad);
}

int field_dots_partials(int64_t n) { return static_cast<int>(lin_blocks(n)); }

void launch_field_dots(int D, int64_t n, const float* a, const float* b, double* partials, double* out,
                       hipStream_t s) {
  const int nb = field_dots_partials(n);
  if (D == 2)
    hipLaunchKernelGGL(k_field_dots<2>, dim3(nb), dim3(kLinThreads), 0, s, n, a, b, partials);
  else
    hipLaunchKernelGGL(k_field_dots<3>, dim3(nb), dim3(kLinThreads), 0, s, n, a, b, partials);
  hipLaunchKernelGGL(k_field_dots_final, dim3(1), dim3(256), 0, s, partials, nb, out);
}

void launch_field_axpy(int64_t n, const float* a, float sc, const float* b, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_field_axpy, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, a, sc, b, out);
}
}  // namespace ants
