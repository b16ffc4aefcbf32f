// Shared host/device definitions of the gfx950 Splotch renderer.
//
// Float semantics: every translation unit is compiled with -ffp-contract=off
// and without fast-math, so each float/double operation below rounds exactly
// as the reference's x86 code does (no fused multiply-adds, IEEE division and
// square root).  Float->int conversions go through cvt_i32, which reproduces
// x86 cvttss2si (NaN and out-of-range give INT_MIN); gfx950's v_cvt_i32_f32
// saturates instead.
#pragma once

#include <hip/hip_runtime.h>
#include <climits>
#include <cstdint>

#include "../../include/splotch_hip.h"

namespace splhip {

constexpr int kMaxTypes = SPLHIP_MAX_TYPES;
constexpr int kThreads = 256;  // 4 wave64 per workgroup

// Projected particle in render order: 32 bytes, one dwordx4 pair per gather.
struct alignas(16) Rec {
  float x, y, r, e0, e1, e2, z, I;
};

// Constants of splotch_host.cc:44-49 and the exptable parameters of
// splotchutils.h:174-215, all computed on the host.
struct Consts {
  float expfac, taylorlimit, h2sigma, bfak, rfac;
};

// int(f) with x86 cvttss2si semantics.
__device__ __forceinline__ int cvt_i32(float f) {
  return (f >= -2147483648.0f && f < 2147483648.0f) ? static_cast<int>(f) : INT_MIN;
}

// linearMap<COLOUR>::getVal_const (linear_map.h:71-84) with interpol_helper
// (math_utils.h:186-202): X (double, ascending) and Y (float RGB) hold np
// nodes; the weights are double, each product is rounded to float.
__device__ __forceinline__ void palette_val(const double *X, const float *Y, int np, double xv,
                                            float &cr, float &cg, float &cb) {
  if (np == 1) {
    cr = Y[0]; cg = Y[1]; cb = Y[2];
  } else if (xv >= X[np - 1]) {
    cr = Y[3 * (np - 1)]; cg = Y[3 * (np - 1) + 1]; cb = Y[3 * (np - 1) + 2];
  } else if (xv <= X[0]) {
    cr = Y[0]; cg = Y[1]; cb = Y[2];
  } else {
    int lo = 0, len = np;
    while (len > 0) {  // lower_bound
      const int half = len >> 1;
      if (X[lo + half] < xv) {
        lo += half + 1;
        len -= half + 1;
      } else {
        len = half;
      }
    }
    int idx = lo;
    if (idx > 0) --idx;
    idx = min(np - 2, idx);
    const double frac = (xv - X[idx]) / (X[idx + 1] - X[idx]);
    const float *y0 = Y + 3 * idx, *y1 = Y + 3 * (idx + 1);
    cr = static_cast<float>((1. - frac) * static_cast<double>(y0[0])) +
         static_cast<float>(frac * static_cast<double>(y1[0]));
    cg = static_cast<float>((1. - frac) * static_cast<double>(y0[1])) +
         static_cast<float>(frac * static_cast<double>(y1[1]));
    cb = static_cast<float>((1. - frac) * static_cast<double>(y0[2])) +
         static_cast<float>(frac * static_cast<double>(y1[2]));
  }
}

// exptable<float32>::operator() (splotchutils.h:203-209); tab = tab1 | tab2.
__device__ __forceinline__ float xexp_op(const float *tab, float arg, float expfac) {
  const int iarg = cvt_i32(arg * expfac);
  if (iarg & ~((1 << 20) - 1)) return (iarg < 0) ? 1.f : 0.f;
  return tab[iarg >> 10] * tab[1024 + (iarg & 1023)];
}

// exptable<float32>::expm1 (splotchutils.h:210-214).
__device__ __forceinline__ float xexp_expm1(const float *tab, float arg, float expfac,
                                            float taylorlimit) {
  if (fabsf(arg) < taylorlimit) return arg;
  return xexp_op(tab, arg, expfac) - 1.f;
}

// Binning geometry: reference tiles of T px (tile id iy*ncx+ix) and, when
// G > 0, work regions of G x G px aligned to each tile's origin
// (region id = tile*ngm*ngm + gy*ngm + gx).
struct BinGeom {
  int T, xres, yres, ncx, ncy;
  int G, ngm;
  float rcell, cmid0, rfac;
  float grayabsorb;  // gray_absorption of the job (ordered pass: q = e / (e + grayabsorb))
};

// Per-job splat description: one entry per view of a batch (a plain render is
// a batch of one).  Units (8x8 blocks or tiles) of job j are numbered from
// unit_base[j] (a separate table) in the job's own tile / region order.
struct JobSplat {
  float *pic;
  int T, xres, yres, ncx, ngm;
  // per pass (0 ordered, 1 additive): 0 the pass writes its chains' values,
  // 1 it adds them to what an earlier pass wrote, 2 its chains continue from
  // the image (a later chunk of the same pass)
  int accumulate[2];
  int need_empty[2];  // per pass: the pass writes every pixel of the job (empty units too)
  float grayabsorb;
};

// Largest j with base[j] <= v, for v < base[nj] (base ascending, base[0] = 0;
// empty jobs repeat their successor's base and are skipped).
__device__ __forceinline__ int find_job(const uint32_t *base, int nj, uint32_t v) {
  int lo = 0, hi = nj;  // invariant: base[lo] <= v < base[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (base[mid] <= v) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Candidate tiles of a particle: [minx, maxx] x [miny, maxy] (empty when
// min > max), and the squared distance bound of the circle test.
struct BinRange {
  int minx, maxx, miny, maxy;
  float sumsq;
};

// A tile the particle is binned to, as the rectangle of its units:
// unit(gx, gy) = base + gy * ngm + gx for gx0 <= gx <= gx1, gy0 <= gy <= gy1
// (G == 0: the single unit base = tile).
struct BinRect {
  uint32_t base;
  int gx0, gx1, gy0, gy1;
};

// Tile rule: splotch_host.cc:367-383 (range + circle test, T = tile_size).
__device__ __forceinline__ BinRange bin_range(const Rec &R, const BinGeom &g) {
  const float rfacr = g.rfac * R.r;
  BinRange b;
  b.minx = max(0, cvt_i32(R.x - rfacr + 1) / g.T);
  b.maxx = min(g.ncx - 1, cvt_i32(R.x + rfacr) / g.T);
  b.miny = max(0, cvt_i32(R.y - rfacr + 1) / g.T);
  b.maxy = min(g.ncy - 1, cvt_i32(R.y + rfacr) / g.T);
  b.sumsq = (g.rcell + rfacr) * (g.rcell + rfacr);
  return b;
}

// Units of candidate tile (ix, iy); false when the particle is not binned
// there.  Region rule: the footprint box that render_new clips to the tile
// (splotch_host.cc:462-472), split into G x G regions.
__device__ __forceinline__ bool bin_tile(const Rec &R, const BinGeom &g, const BinRange &b, int ix,
                                         int iy, BinRect &rc) {
  const float cx = g.cmid0 + static_cast<float>(ix * g.T);
  const float cy = g.cmid0 + static_cast<float>(iy * g.T);
  const float rtot2 = (R.x - cx) * (R.x - cx) + (R.y - cy) * (R.y - cy);
  if (!(rtot2 < b.sumsq)) return false;
  const int tile = iy * g.ncx + ix;
  if (g.G == 0) {
    rc = BinRect{static_cast<uint32_t>(tile), 0, 0, 0, 0};
    return true;
  }
  const int x0s = ix * g.T, y0s = iy * g.T;
  const int w = min(x0s + g.T, g.xres) - x0s, h = min(y0s + g.T, g.yres) - y0s;
  const float rf2 = R.r * g.rfac;
  const float posx = R.x - static_cast<float>(x0s), posy = R.y - static_cast<float>(y0s);
  const int bx0 = max(cvt_i32(posx - rf2 + 1), 0), bx1 = min(cvt_i32(posx + rf2 + 1), w);
  const int by0 = max(cvt_i32(posy - rf2 + 1), 0), by1 = min(cvt_i32(posy + rf2 + 1), h);
  if (bx0 >= bx1 || by0 >= by1) return false;
  rc.base = static_cast<uint32_t>(tile) * g.ngm * g.ngm;
  rc.gx0 = bx0 / g.G;
  rc.gx1 = (bx1 - 1) / g.G;
  rc.gy0 = by0 / g.G;
  rc.gy1 = (by1 - 1) / g.G;
  return true;
}

}  // namespace splhip
