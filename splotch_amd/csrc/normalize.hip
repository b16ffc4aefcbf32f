// particle_normalize on the GPU (splotch/scenemaker.cc:40-296, host branch):
// per particle the optional log10 / asinh of intensity and colour, per type
// the min / max of intensity, colour and smoothing length, then the clamp of
// Normalizer::normAndClamp (splotch/splotchutils.h:138-171).  Two passes over
// the SoA fields (transform + reduce, then clamp), both HBM-bound.
//
// The per-type extremes are kept as order-preserving 32-bit keys (sign-
// flipped float bits) so they reduce with integer atomics: per thread in
// registers, per wave by shuffles, per block in LDS, then one atomic per
// block and value.  The per-type switches are staged in LDS (a lane-indexed
// kernel argument would be copied to private memory).  NaN
// never enters (std::min / std::max never take it); of +0 and -0 the minimum
// takes -0 and the maximum +0 (the reference keeps whichever its thread order
// meets first).
#include "kernels.hpp"

#include <algorithm>

#include "workspace.hpp"

namespace splhip {
namespace {

__device__ __forceinline__ uint32_t fkey(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ float fdekey(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// log10 of a float, rounded from the double result.
__device__ __forceinline__ float log10_f(float v) {
  return static_cast<float>(log10(static_cast<double>(v)));
}

// my_asinh (splotchutils.cc:31-32): log(v + sqrt(1 + v v)) in double.
__device__ __forceinline__ float asinh_f(float v) {
  const double d = static_cast<double>(v);
  return static_cast<float>(log(d + sqrt(1. + d * d)));
}

// Slots of a type's keys: colour min / max, intensity min / max, size min / max.
enum { K_CMIN, K_CMAX, K_IMIN, K_IMAX, K_RMIN, K_RMAX, K_N };

__global__ __launch_bounds__(256) void k_norm_init(uint32_t *keys, int nt) {
  const int i = threadIdx.x;
  if (i < nt * K_N) keys[i] = (i % 2 == 0) ? fkey(1e37f) : fkey(-1e37f);
}

__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
#pragma unroll
  for (int d = 32; d; d >>= 1)
    v = min(v, static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), d, 64)));
  return v;
}

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
  for (int d = 32; d; d >>= 1)
    v = max(v, static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), d, 64)));
  return v;
}

// Pass 1: transforms in place and reduces the extremes of NT types.
template <int NT>
__global__ __launch_bounds__(256) void k_norm_reduce(NormArgs a, uint32_t *keys) {
  __shared__ int32_t s_flags[NT][4];
  if (threadIdx.x < NT * 4) s_flags[threadIdx.x / 4][threadIdx.x % 4] =
      a.flags[threadIdx.x / 4][threadIdx.x % 4];
  __syncthreads();
  uint32_t k[NT][K_N];
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int q = 0; q < K_N; ++q) k[t][q] = (q % 2 == 0) ? 0xffffffffu : 0u;
  // four particles per step, their loads issued before any store
  const int64_t stride = static_cast<int64_t>(gridDim.x) * 256;
  for (int64_t m0 = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; m0 < a.n;
       m0 += 4 * stride) {
    int tv[4];
    float Iv[4], ev[4], rv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t m = m0 + j * stride;
      tv[j] = NT;
      if (m < a.n) {
        tv[j] = a.type ? a.type[m] : 0;
        Iv[j] = a.I[m];
        ev[j] = a.er[m];
        rv[j] = a.r[m];
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t m = m0 + j * stride;
      const int t = tv[j];
      if (t >= NT) continue;  // past the end, or a type beyond ptypes (not rendered)
      const int f_il = s_flags[t][0], f_cl = s_flags[t][1], f_ca = s_flags[t][2],
                f_cv = s_flags[t][3];
      // extremes of this particle (min keys start high, max keys low)
      uint32_t cmin = 0xffffffffu, cmax = 0u, imin = 0xffffffffu, imax = 0u;
      const float r = rv[j];
      if (r == r) {
        const uint32_t kr = fkey(r);
#pragma unroll
        for (int u = 0; u < NT; ++u)
          if (u == t) {
            k[u][K_RMIN] = min(k[u][K_RMIN], kr);
            k[u][K_RMAX] = max(k[u][K_RMAX], kr);
          }
      }
      float I = Iv[j];
      bool icol = true;
      if (f_il) {
        if (I > 0) {
          I = log10_f(I);
        } else {
          I = -38.f;
          icol = false;
        }
        a.I[m] = I;
      }
      if (icol && I == I) imin = imax = fkey(I);
      float e = ev[j];
      bool ccol = true;
      if (f_cl) {
        if (e > 0) {
          e = log10_f(e);
        } else {
          e = -38.f;
          ccol = false;
        }
        a.er[m] = e;
      } else if (f_ca) {
        e = asinh_f(e);
        a.er[m] = e;
      }
      if (ccol && e == e) cmin = cmax = fkey(e);
      if (f_cv) {
        float g = a.eg[m], b = a.eb[m];
        if (f_cl) {
          g = log10_f(g);
          b = log10_f(b);
        }
        if (f_ca) {
          g = asinh_f(g);
          b = asinh_f(b);
        }
        if (f_cl || f_ca) {
          a.eg[m] = g;
          a.eb[m] = b;
        }
        if (g == g) {
          cmin = min(cmin, fkey(g));
          cmax = max(cmax, fkey(g));
        }
        if (b == b) {
          cmin = min(cmin, fkey(b));
          cmax = max(cmax, fkey(b));
        }
      }
#pragma unroll
      for (int u = 0; u < NT; ++u)
        if (u == t) {
          k[u][K_CMIN] = min(k[u][K_CMIN], cmin);
          k[u][K_CMAX] = max(k[u][K_CMAX], cmax);
          k[u][K_IMIN] = min(k[u][K_IMIN], imin);
          k[u][K_IMAX] = max(k[u][K_IMAX], imax);
        }
    }
  }
  // per wave by shuffles, per block through LDS, then one atomic per block
  // and value (thousands of atomics on a few addresses would serialize)
  __shared__ uint32_t s_k[4][NT * K_N];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int q = 0; q < K_N; ++q) {
      const uint32_t v = (q % 2 == 0) ? wave_min(k[t][q]) : wave_max(k[t][q]);
      if (lane == 0) s_k[wave][t * K_N + q] = v;
    }
  __syncthreads();
  const int i = threadIdx.x;
  if (i < NT * K_N) {
    const bool is_min = (i % K_N) % 2 == 0;
    uint32_t v = s_k[0][i];
    for (int w = 1; w < 4; ++w) v = is_min ? min(v, s_k[w][i]) : max(v, s_k[w][i]);
    if (is_min) {
      if (v != 0xffffffffu) atomicMin(&keys[i], v);
    } else {
      if (v != 0u) atomicMax(&keys[i], v);
    }
  }
}

// The range a type's values are clamped to (after the parameter overrides).
struct Range {
  float minv, maxv;
  __device__ __forceinline__ void clamp(float &v) const {
    if (maxv == minv) {
      v = 1.f;
    } else {
      const float lo = (v < maxv) ? v : maxv;   // std::min(maxv, v)
      const float hi = (minv < lo) ? lo : minv; // std::max(minv, lo)
      v = (hi - minv) / (maxv - minv);
    }
  }
};

// Pass 2: normAndClamp of intensity (unless NO_I_NORM) and colour; the
// block first stages every type's two ranges and vector flag in LDS.
__global__ __launch_bounds__(256) void k_norm_clamp(NormArgs a, const uint32_t *keys) {
  __shared__ Range s_int[kMaxTypes], s_col[kMaxTypes];
  __shared__ int32_t s_vec[kMaxTypes];
  const int i = threadIdx.x;
  if (i < a.ntypes) {
    Range I, C;
    I.minv = a.ovr_set[i][0] ? a.ovr[i][0] : fdekey(keys[i * K_N + K_IMIN]);
    I.maxv = a.ovr_set[i][1] ? a.ovr[i][1] : fdekey(keys[i * K_N + K_IMAX]);
    C.minv = a.ovr_set[i][2] ? a.ovr[i][2] : fdekey(keys[i * K_N + K_CMIN]);
    C.maxv = a.ovr_set[i][3] ? a.ovr[i][3] : fdekey(keys[i * K_N + K_CMAX]);
    s_int[i] = I;
    s_col[i] = C;
    s_vec[i] = a.flags[i][3];
  }
  __syncthreads();
  const int64_t stride = static_cast<int64_t>(gridDim.x) * 256;
  for (int64_t m0 = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; m0 < a.n;
       m0 += 4 * stride) {
    int tv[4];
    float Iv[4], ev[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {  // loads first
      const int64_t m = m0 + j * stride;
      tv[j] = kMaxTypes;
      if (m < a.n) {
        tv[j] = a.type ? a.type[m] : 0;
        Iv[j] = a.I[m];
        ev[j] = a.er[m];
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t m = m0 + j * stride;
      const int t = tv[j];
      if (t >= a.ntypes) continue;
      if (a.norm_int) {
        float I = Iv[j];
        s_int[t].clamp(I);
        a.I[m] = I;
      }
      const Range C = s_col[t];
      float e = ev[j];
      C.clamp(e);
      a.er[m] = e;
      if (s_vec[t]) {
        float g = a.eg[m], b = a.eb[m];
        C.clamp(g);
        C.clamp(b);
        a.eg[m] = g;
        a.eb[m] = b;
      }
    }
  }
}

template <int NT>
void launch_reduce(const NormArgs &a, uint32_t *keys, unsigned grid, hipStream_t s) {
  k_norm_reduce<NT><<<grid, 256, 0, s>>>(a, keys);
}

}  // namespace

void launch_normalize(const NormArgs &a, uint32_t *keys, hipStream_t s) {
  k_norm_init<<<1, 256, 0, s>>>(keys, a.ntypes);
  hip_check(hipGetLastError(), "k_norm_init launch");
  if (a.n > 0) {
    const unsigned grid =
        static_cast<unsigned>(std::min<int64_t>(std::max<int64_t>((a.n + 255) / 256, 1), 1024));
    switch (a.ntypes) {
      case 1: launch_reduce<1>(a, keys, grid, s); break;
      case 2: launch_reduce<2>(a, keys, grid, s); break;
      case 3: launch_reduce<3>(a, keys, grid, s); break;
      case 4: launch_reduce<4>(a, keys, grid, s); break;
      case 5: launch_reduce<5>(a, keys, grid, s); break;
      case 6: launch_reduce<6>(a, keys, grid, s); break;
      case 7: launch_reduce<7>(a, keys, grid, s); break;
      default: launch_reduce<8>(a, keys, grid, s); break;
    }
    hip_check(hipGetLastError(), "k_norm_reduce launch");
    k_norm_clamp<<<grid, 256, 0, s>>>(a, keys);
    hip_check(hipGetLastError(), "k_norm_clamp launch");
  }
}

}  // namespace splhip
