// Device helpers shared by the splat and adjoint kernels.
#pragma once

#include "kernels.hpp"

namespace splhip {

__device__ __forceinline__ uint32_t uniform(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
}

__device__ __forceinline__ float lane_f(float v, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kPad = 65;  // LDS row stride: conflict-free column writes and reads

// Where global unit u (an 8x8 block of one job's reference tile) lies: the
// job's image and flags, the tile origin and clipped size, the block's place
// in the tile.  For a wave-uniform u the lookups compile to scalar loads.
struct UnitGeom {
  float *pic;
  int yres, accumulate, need_empty;
  float grayabsorb;
  int x0s, y0s, w, h, gx, gy;
};

__device__ __forceinline__ UnitGeom unit_geom(const SplatArgs &a, uint32_t u) {
  const int j = find_job(a.ubase, a.nj, u);
  const JobSplat &J = a.jobs[j];
  const uint32_t local = u - a.ubase[j];
  const int ngm = J.ngm, T = J.T;
  const int per_tile = ngm * ngm;
  const int tile = static_cast<int>(local / per_tile);
  const int rem = static_cast<int>(local - static_cast<uint32_t>(tile) * per_tile);
  UnitGeom g;
  g.gy = rem / ngm;
  g.gx = rem - g.gy * ngm;
  g.x0s = (tile % J.ncx) * T;
  g.y0s = (tile / J.ncx) * T;
  g.w = min(g.x0s + T, J.xres) - g.x0s;
  g.h = min(g.y0s + T, J.yres) - g.y0s;
  g.pic = J.pic;
  g.yres = J.yres;
  g.accumulate = J.accumulate[a.pass];
  g.need_empty = J.need_empty[a.pass];
  g.grayabsorb = J.grayabsorb;
  return g;
}

// unit_geom for a wave-uniform unit, held in scalar registers.
__device__ __forceinline__ UnitGeom unit_geom_uniform(const SplatArgs &a, uint32_t u) {
  UnitGeom g = unit_geom(a, u);
  auto ui = [](int v) { return __builtin_amdgcn_readfirstlane(v); };
  const uint64_t pp = reinterpret_cast<uint64_t>(g.pic);
  const uint64_t lo = static_cast<uint32_t>(ui(static_cast<int>(pp & 0xffffffffu)));
  const uint64_t hi = static_cast<uint32_t>(ui(static_cast<int>(pp >> 32)));
  g.pic = reinterpret_cast<float *>(lo | (hi << 32));
  g.yres = ui(g.yres);
  g.accumulate = ui(g.accumulate);
  g.need_empty = ui(g.need_empty);
  g.grayabsorb = __int_as_float(ui(__float_as_int(g.grayabsorb)));
  g.x0s = ui(g.x0s);
  g.y0s = ui(g.y0s);
  g.w = ui(g.w);
  g.h = ui(g.h);
  g.gx = ui(g.gx);
  g.gy = ui(g.gy);
  return g;
}

}  // namespace splhip
