// Differentiable form of the splat: adjoint kernels and the projection chain.
//
// Model.  The float32 renderer's footprint is discrete (which pixels a
// particle touches follows integer truncation, splotch_host.cc:591-609), so the
// gradient is that of the smooth model on the FROZEN footprint: each
// (pixel, particle) pair the forward render visits contributes with the true
// exponential instead of the exp table,
//     att = exp(stp*(y-sy)^2) * exp(stp*(x-sx)^2),   stp = -1/(h2sigma*r)^2
//     additive:  P_c += -e_c * att                         (splotch_host.cc:615)
//     ordered :  P_c  = q_c + (P_c - q_c) * exp(-e_c * att), q_c = e_c/(e_c+g)
// (the limit of both branches of splotch_host.cc:648-658).  Pairs outside the
// footprint contribute nothing whatever the parameters.
//
// k_adjoint_add: additive mode, one wave per list segment of an 8x8 block,
// lane = particle: a particle's gradient is a sum over its own footprint
// pixels.  k_adjoint_ord: ordered mode, one wave per list segment, lane =
// pixel, particles in render order, in passes: per segment the composite map
// and optical depth of its particles (A), the segments joined per pixel
// (k_adjoint_join, B), then the gradients (C), which carry the composited
// value P_{k-1} and the prefix depth, from which the transmittance behind
// particle k, S_k = exp(-(tau - prefix_k)), follows without division.
// Per particle, the six screen-space partials d/d(sx, sy, r, e0, e1, e2) are
// summed over the block's 64 pixels and stored per list entry (float); the
// entries of each particle are then gathered in a fixed order (a stable sort
// of the entries by particle) and summed in FP64, so the gradient is the same
// bit for bit on every run.
//
// k_chain: per particle, d(screen state)/d(x, y, z, hsml, I, colour) of
// particle_project + particle_colorize (splotch_host.cc:111-178, 252-287),
// evaluated in double, applied to the accumulated screen-space gradient.
#include "device_util.hpp"
#include "kernels.hpp"
#include "workspace.hpp"

namespace splhip {
namespace {

constexpr int kPadA = 65;
constexpr uint32_t kSeg = kAdjSeg;  // entries per backward work item (list segment)

__device__ __forceinline__ int64_t gtid_() {
  return static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
}
__device__ __forceinline__ int64_t gstride_() {
  return static_cast<int64_t>(gridDim.x) * blockDim.x;
}

// Sum over the wave's 64 lanes with DPP row operations (quad permutes,
// half-row and row mirrors, then row broadcasts 15 and 31); the total ends in
// lane 63 and is returned to every lane.
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_add(float v) {
  const int s = __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xf, false);
  return v + __int_as_float(s);
}

__device__ __forceinline__ float wave_sum(float v) {
  v = dpp_add<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
  v = dpp_add<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
  v = dpp_add<0x141, 0xf>(v);  // row_half_mirror
  v = dpp_add<0x140, 0xf>(v);  // row_mirror: every lane of a row holds the row sum
  v = dpp_add<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v = dpp_add<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Additive adjoint, one lane per particle: the additive gradient of a particle
// is a sum over its own footprint pixels, so each lane walks the block's 64
// pixels for its particle (dL/dpic of the block staged in LDS) and needs no
// cross-lane reduction.
__global__ __launch_bounds__(256) void k_adjoint_add(SplatArgs a, const uint32_t *seg_unit,
                                                     const uint32_t *seg_first, uint32_t nseg,
                                                     const float *grad, float *part) {
  __shared__ float s_g[4][3][64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t sg = blockIdx.x * 4u + wave;
  if (sg >= nseg) return;  // waves only synchronise among their own lanes
  const uint32_t u = uniform(seg_unit[sg]);
  constexpr int chan = 3;
  const UnitGeom ug = unit_geom_uniform(a, u);
  const int x0s = ug.x0s, y0s = ug.y0s, w = ug.w, h = ug.h;
  const int bxo = ug.gx * 8, byo = ug.gy * 8;  // tile frame
  {
    const int cx = lane >> 3, cy = lane & 7;
    const int lx = bxo + cx, ly = byo + cy;
    const bool pix = lx < w && ly < h;
    const float *g = grad + (static_cast<int64_t>(x0s + lx) * ug.yres + (y0s + ly)) * 3;
#pragma unroll
    for (int c = 0; c < 3; ++c) s_g[wave][c][lane] = (pix && (chan == 3 || chan == c)) ? g[c] : 0.f;
  }
  wave_sync();
  const float(*G)[64] = s_g[wave];
  const uint32_t e0 = uniform(seg_first[sg]);
  const uint32_t e1 = uniform(min(e0 + kSeg, a.bin_off[u + 1]));
  for (uint32_t base = e0; base < e1; base += 64) {
    if (base + lane >= e1) break;
    const uint32_t s = a.ent[base + lane];
    const Rec R = a.recs[s];
    const float rfacr = R.r * a.rfac;
    const float posx = R.x - static_cast<float>(x0s), posy = R.y - static_cast<float>(y0s);
    const int bx0 = max(cvt_i32(posx - rfacr + 1), 0), bx1 = min(cvt_i32(posx + rfacr + 1), w);
    const int by0 = max(cvt_i32(posy - rfacr + 1), 0), by1 = min(cvt_i32(posy + rfacr + 1), h);
    const float radsq = rfacr * rfacr;
    const float sigma = a.h2sigma * R.r;
    const float stp = -1.f / (sigma * sigma);
    const float dstp = 2.f / (a.h2sigma * a.h2sigma * R.r * R.r * R.r);
    float ey[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const float dyf = static_cast<float>(byo + r) - posy;
      ey[r] = __expf(stp * dyf * dyf);
    }
    float gsx = 0.f, gsy = 0.f, gr = 0.f, ge0 = 0.f, ge1 = 0.f, ge2 = 0.f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int x = bxo + c;
      if (x < bx0 || x >= bx1) continue;
      const float dx = static_cast<float>(x) - posx;
      const float dxsq = dx * dx;
      const float dy = sqrtf(radsq - dxsq);
      const int lo = max(by0, cvt_i32(posy - dy + 1));
      const int hi = min(by1, cvt_i32(posy + dy + 1));
      const float fx = __expf(stp * dxsq);
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int y = byo + r;
        if (y < lo || y >= hi) continue;  // the forward's footprint (same float rules)
        const float att = ey[r] * fx;
        const float g0 = G[0][c * 8 + r], g1 = G[1][c * 8 + r], g2 = G[2][c * 8 + r];
        const float gatt = -(g0 * R.e0 + g1 * R.e1 + g2 * R.e2);
        const float dyr = static_cast<float>(y) - posy;
        gsx += gatt * (-2.f * stp * dx * att);
        gsy += gatt * (-2.f * stp * dyr * att);
        gr += gatt * att * (dxsq + dyr * dyr) * dstp;
        ge0 -= g0 * att;
        ge1 -= g1 * att;
        ge2 -= g2 * att;
      }
    }
    float *pp = part + 6 * static_cast<size_t>(base + lane);  // this entry's partials
    pp[0] = gsx;
    pp[1] = gsy;
    pp[2] = gr;
    pp[3] = ge0;
    pp[4] = ge1;
    pp[5] = ge2;
  }
}

// Backward work items are list segments of at most kSeg entries: (unit,
// first entry).  Additive segments are independent; ordered segments are
// joined per pixel by the affine composite map of each segment (passes A, B).

// Per-segment ordered state, SoA over [segment][channel][pixel] (doubles):
//   slot 0: pass A: A = exp(-D), the segment's transmittance; pass B: P at the
//           segment start
//   slot 1: pass A: B = the segment's composite of P = 0; pass B: D behind
//           the segment (optical depth of the later segments)
//   slot 2: D = optical depth of the segment
struct SegView {
  double *s0, *s1, *s2;
  __device__ size_t at(uint32_t seg, int c, int px) const {
    return (static_cast<size_t>(seg) * 3 + c) * 64 + px;
  }
};

__global__ __launch_bounds__(256) void k_seg_count(const uint32_t *bin_off, uint32_t nunits,
                                                   uint32_t *cnt) {
  for (int64_t u = gtid_(); u < nunits; u += gstride_())
    cnt[u] = (bin_off[u + 1] - bin_off[u] + kSeg - 1) / kSeg;
}

__global__ __launch_bounds__(256) void k_seg_emit(const uint32_t *bin_off, uint32_t nunits,
                                                  const uint32_t *soff, uint32_t *seg_unit,
                                                  uint32_t *seg_first) {
  for (int64_t u = gtid_(); u < nunits; u += gstride_()) {
    const uint32_t b = bin_off[u], e = bin_off[u + 1];
    uint32_t o = soff[u];
    for (uint32_t f = b; f < e; f += kSeg, ++o) {
      seg_unit[o] = static_cast<uint32_t>(u);
      seg_first[o] = f;
    }
  }
}

// One wave per ordered segment, lane = pixel of the 8x8 block.
//   PASS 0 (A): the segment's composite map and optical depth per pixel.
//   PASS 2 (C): gradients, from the segment's start value and the depth
//               behind it (pass B), particles in render order.
template <int PASS>
__global__ __launch_bounds__(256) void k_adjoint_ord(SplatArgs a, const uint32_t *seg_unit,
                                                     const uint32_t *seg_first, uint32_t nseg,
                                                     const float *grad, float *part, SegView sv) {
  __shared__ uint32_t s_col[4][8][kPadA];
  __shared__ float s_ex[4][8][kPadA];
  __shared__ float s_ey[4][8][kPadA];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t sg = blockIdx.x * 4u + wave;
  if (sg >= nseg) return;  // waves only synchronise among their own lanes
  const uint32_t u = uniform(seg_unit[sg]);
  const UnitGeom ug = unit_geom_uniform(a, u);
  const int x0s = ug.x0s, y0s = ug.y0s, w = ug.w, h = ug.h;
  const int bxo = ug.gx * 8, byo = ug.gy * 8;  // tile frame
  const int cx = lane >> 3, cy = lane & 7;
  const int lx = bxo + cx, ly = byo + cy;
  uint32_t(*col)[kPadA] = s_col[wave];
  float(*ex)[kPadA] = s_ex[wave];
  float(*ey)[kPadA] = s_ey[wave];
  const bool pix = lx < w && ly < h;
  float G[3] = {0.f, 0.f, 0.f};
  if (PASS == 2 && pix) {
    const float *g = grad + (static_cast<int64_t>(x0s + lx) * ug.yres + (y0s + ly)) * 3;
    G[0] = g[0];
    G[1] = g[1];
    G[2] = g[2];
  }
  const float gray = ug.grayabsorb;
  const float lxf = static_cast<float>(lx), lyf = static_cast<float>(ly);
  const uint32_t e0 = uniform(seg_first[sg]);
  const uint32_t e1 = uniform(min(e0 + kSeg, a.bin_off[u + 1]));

  // per-lane particle state of the current batch (lane = particle)
  float pe[3], pq[3], ppx, ppy, pstp, pdstp;
  // Phase A (lane = particle): the forward's footprint masks of this block's
  // columns (same float rules as the splat), smooth column / row factors.
  auto phase_a = [&](uint32_t base, int nb) {
    pe[0] = pe[1] = pe[2] = pq[0] = pq[1] = pq[2] = 0.f;
    ppx = ppy = pstp = pdstp = 0.f;
    if (lane < nb) {
      const uint32_t s = a.ent[base + lane];
      const Rec R = a.recs[s];
      const float rfacr = R.r * a.rfac;
      const float posx = R.x - static_cast<float>(x0s), posy = R.y - static_cast<float>(y0s);
      const int bx0 = max(cvt_i32(posx - rfacr + 1), 0), bx1 = min(cvt_i32(posx + rfacr + 1), w);
      const int by0 = max(cvt_i32(posy - rfacr + 1), 0), by1 = min(cvt_i32(posy + rfacr + 1), h);
      const float radsq = rfacr * rfacr;
      const float sigma = a.h2sigma * R.r;
      const float stp = -1.f / (sigma * sigma);
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int x = bxo + c;
        uint32_t m = 0;
        float fx = 0.f;
        if (x >= bx0 && x < bx1) {
          const float dxsq = (static_cast<float>(x) - posx) * (static_cast<float>(x) - posx);
          const float dy = sqrtf(radsq - dxsq);
          const int lo = max(by0, cvt_i32(posy - dy + 1));
          const int hi = min(by1, cvt_i32(posy + dy + 1));
          const int r0 = min(max(lo - byo, 0), 8), r1 = min(max(hi - byo, 0), 8);
          if (hi > lo && r1 > r0) m = ((1u << r1) - 1u) & ~((1u << r0) - 1u);
          fx = __expf(stp * dxsq);
        }
        col[c][lane] = m;
        ex[c][lane] = fx;
      }
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const float dyf = static_cast<float>(byo + r) - posy;
        ey[r][lane] = __expf(stp * dyf * dyf);
      }
      pe[0] = R.e0;
      pe[1] = R.e1;
      pe[2] = R.e2;
      pq[0] = R.e0 / (R.e0 + gray);
      pq[1] = R.e1 / (R.e1 + gray);
      pq[2] = R.e2 / (R.e2 + gray);
      ppx = posx;
      ppy = posy;
      pstp = stp;
      // d stp / d r = 2 / (h2sigma^2 r^3)
      pdstp = 2.f / (a.h2sigma * a.h2sigma * R.r * R.r * R.r);
    }
  };

  // running state in double: P - q and depth differences cancel for opaque pixels
  double P[3], D[3], Dseg[3], Dafter[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    D[c] = 0.0;
    if (PASS == 0) {
      P[c] = 0.0;
      Dseg[c] = Dafter[c] = 0.0;
    } else {
      P[c] = sv.s0[sv.at(sg, c, lane)];
      Dafter[c] = sv.s1[sv.at(sg, c, lane)];
      Dseg[c] = sv.s2[sv.at(sg, c, lane)];
    }
  }
  for (uint32_t base = e0; base < e1; base += 64) {
    const int nb = static_cast<int>(uniform(min<uint32_t>(64u, e1 - base)));
    phase_a(base, nb);
    wave_sync();
    for (int k = 0; k < nb; ++k) {
      const bool in = (col[cx][k] >> cy) & 1u;
      if (PASS == 0) {
        if (in) {
          const float att = ey[cy][k] * ex[cx][k];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const float ea = lane_f(pe[c], k) * att;
            const double q = lane_f(pq[c], k);
            const double T = 1.0 + static_cast<double>(expm1f(-ea));
            P[c] = q + (P[c] - q) * T;
            D[c] += static_cast<double>(ea);
          }
        }
        continue;
      }
      const uint64_t any = __ballot(in);
      float *pp = part + 6 * static_cast<size_t>(base + k);  // entry base + k's partials
      if (any == 0) {  // uniform: the particle misses this block's pixels
        if (lane < 6) pp[lane] = 0.f;
        continue;
      }
      const float att = ey[cy][k] * ex[cx][k];
      const float dx = lxf - lane_f(ppx, k), dy = lyf - lane_f(ppy, k);
      const float stp = lane_f(pstp, k);
      float gatt = 0.f, ge[3] = {0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (!in) continue;
        const float e = lane_f(pe[c], k);
        const float q = lane_f(pq[c], k);
        const float ea = e * att;
        const float omt = -expm1f(-ea);  // 1 - T without cancellation for small e*att
        const float T = 1.f - omt;
        D[c] += static_cast<double>(ea);
        // transmittance behind particle k: the rest of this segment and the later ones
        const float S = __expf(static_cast<float>(-((Dseg[c] - D[c]) + Dafter[c])));
        const float gk = G[c] * S;  // dL/dP_k
        const float dPdT = static_cast<float>(P[c] - static_cast<double>(q));
        const float dqde = gray / ((e + gray) * (e + gray));
        gatt += gk * dPdT * (-e * T);
        ge[c] = gk * (dPdT * (-att * T) + omt * dqde);
        P[c] = static_cast<double>(q) + (P[c] - static_cast<double>(q)) * static_cast<double>(T);
      }
      const float gsx = gatt * (-2.f * stp * dx * att);
      const float gsy = gatt * (-2.f * stp * dy * att);
      const float gr = gatt * att * (dx * dx + dy * dy) * lane_f(pdstp, k);
      const float s0 = wave_sum(gsx), s1 = wave_sum(gsy), s2 = wave_sum(gr);
      const float s3 = wave_sum(ge[0]), s4 = wave_sum(ge[1]), s5 = wave_sum(ge[2]);
      const float v = lane == 0 ? s0 : lane == 1 ? s1 : lane == 2 ? s2 :
                      lane == 3 ? s3 : lane == 4 ? s4 : s5;
      if (lane < 6) pp[lane] = v;
    }
    wave_sync();
  }
  if (PASS == 0) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      sv.s0[sv.at(sg, c, lane)] = exp(-D[c]);
      sv.s1[sv.at(sg, c, lane)] = P[c];
      sv.s2[sv.at(sg, c, lane)] = D[c];
    }
  }
}

// Pass B: per (unit, channel, pixel), join the unit's segments in order: the
// composite value at each segment's start (P_0 = 0, P_{j+1} = A_j P_j + B_j)
// and the optical depth behind each segment.
__global__ __launch_bounds__(192) void k_adjoint_join(const uint32_t *soff, uint32_t nunits,
                                                      SegView sv) {
  const uint32_t u = blockIdx.x;
  if (u >= nunits) return;
  const int c = threadIdx.x >> 6, px = threadIdx.x & 63;
  const uint32_t j0 = soff[u], j1 = soff[u + 1];
  double Pst = 0.0;
  for (uint32_t j = j0; j < j1; ++j) {
    const size_t i = sv.at(j, c, px);
    const double A = sv.s0[i], B = sv.s1[i];
    sv.s0[i] = Pst;
    Pst = A * Pst + B;
  }
  double Da = 0.0;
  for (uint32_t j = j1; j > j0; --j) {
    const size_t i = sv.at(j - 1, c, px);
    sv.s1[i] = Da;
    Da += sv.s2[i];
  }
}

// Chain rule through particle_project + particle_colorize for every particle
// of the job (original order).  Inactive particles get zero gradients.
__global__ __launch_bounds__(256) void k_chain(const ProjectArgs *__restrict__ jobs,
                                               const uint8_t *__restrict__ act,
                                               const double *__restrict__ gscr,
                                               splhip_grads out) {
  const ProjectArgs &a = jobs[0];
  for (int64_t i = gtid_(); i < a.n; i += gstride_()) {
    const double *g = gscr + 6 * static_cast<size_t>(i);
    float gx = 0.f, gy = 0.f, gz = 0.f, gh = 0.f, gI = 0.f, gc0 = 0.f, gc1 = 0.f, gc2 = 0.f;
    if (act[a.poff + i] != 0) {
      const int t = a.type ? a.type[i] : 0;
      const int tt = t < kMaxTypes ? t : kMaxTypes - 1;
      const float *m = a.cam.mat;
      const double X = a.x[i], Y = a.y[i], Z = a.z[i];
      const double vx = X * m[0] + Y * m[1] + Z * m[2] + m[3];
      const double vy = X * m[4] + Y * m[5] + Z * m[6] + m[7];
      const double vz = X * m[8] + Y * m[9] + Z * m[10] + m[11];
      const double hs = a.hsml[i], I = a.inten[i];
      const double res2 = a.cam.res2;
      const double xfac2 = a.cam.projection ? -res2 / (static_cast<double>(a.cam.fovfct) * vz)
                                            : static_cast<double>(a.cam.xfac);
      const double sm = a.smooth_fac[tt], mr = a.cam.minrad_pix;
      const double r2 = 2.0 * hs * xfac2 * sm;
      const double r = sqrt(r2 * r2 + mr * mr);
      const double k0 = 0.5 * static_cast<double>(a.bfak);
      const double I1 = I * k0 / hs;
      const double Ie = I1 * r2 / r;
      const double br = a.brightness[tt];
      // colour and its slope in the palette (linear_map.h:71-84)
      double col[3], slope[3] = {0.0, 0.0, 0.0};
      if (a.col_vector[tt]) {
        col[0] = a.col[i];
        col[1] = a.c1 ? a.c1[i] : 0.0;
        col[2] = a.c2 ? a.c2[i] : 0.0;
      } else {
        const double *Xp = a.pal_x + static_cast<int64_t>(tt) * a.pal_stride;
        const float *Yp = a.pal_rgb + static_cast<int64_t>(tt) * a.pal_stride * 3;
        const int np = a.pal_n[tt];
        const double xv = static_cast<float>(a.col[i]);
        if (np == 1 || xv <= Xp[0]) {
          for (int c = 0; c < 3; ++c) col[c] = Yp[c];
        } else if (xv >= Xp[np - 1]) {
          for (int c = 0; c < 3; ++c) col[c] = Yp[3 * (np - 1) + c];
        } else {
          int lo = 0, len = np;
          while (len > 0) {
            const int half = len >> 1;
            if (Xp[lo + half] < xv) {
              lo += half + 1;
              len -= half + 1;
            } else {
              len = half;
            }
          }
          int idx = lo;
          if (idx > 0) --idx;
          idx = min(np - 2, idx);
          const double dxp = Xp[idx + 1] - Xp[idx];
          const double frac = (xv - Xp[idx]) / dxp;
          for (int c = 0; c < 3; ++c) {
            const double y0 = Yp[3 * idx + c], y1 = Yp[3 * (idx + 1) + c];
            col[c] = (1.0 - frac) * y0 + frac * y1;
            slope[c] = (y1 - y0) / dxp;
          }
        }
      }
      const double gsx = g[0], gsy = g[1], gr = g[2];
      double gIe = 0.0;
      for (int c = 0; c < 3; ++c) gIe += g[3 + c] * col[c] * br;
      // r = sqrt(r2^2 + mr^2), Ie = I1 * r2 / r
      const double gr2 = gr * (r2 / r) + gIe * (I1 * mr * mr / (r * r * r));
      gh = static_cast<float>(gr2 * (2.0 * xfac2 * sm) + gIe * (-Ie / hs));
      gI = static_cast<float>(gIe * (k0 / hs) * (r2 / r));
      if (a.col_vector[tt]) {
        gc0 = static_cast<float>(g[3] * Ie * br);
        gc1 = static_cast<float>(g[4] * Ie * br);
        gc2 = static_cast<float>(g[5] * Ie * br);
      } else {
        double s = 0.0;
        for (int c = 0; c < 3; ++c) s += g[3 + c] * slope[c] * Ie * br;
        gc0 = static_cast<float>(s);
      }
      double gvx, gvy, gvz;
      if (a.cam.projection) {
        gvx = gsx * xfac2;
        gvy = gsy * xfac2;
        gvz = -(gsx * vx + gsy * vy) * xfac2 / vz - gr2 * r2 / vz;
      } else {
        gvx = gsx * xfac2;
        gvy = gsy * xfac2;
        gvz = 0.0;
      }
      gx = static_cast<float>(m[0] * gvx + m[4] * gvy + m[8] * gvz);
      gy = static_cast<float>(m[1] * gvx + m[5] * gvy + m[9] * gvz);
      gz = static_cast<float>(m[2] * gvx + m[6] * gvy + m[10] * gvz);
    }
    out.x[i] = gx;
    out.y[i] = gy;
    out.z[i] = gz;
    out.hsml[i] = gh;
    out.inten[i] = gI;
    out.col[i] = gc0;
    if (out.c1) out.c1[i] = gc1;
    if (out.c2) out.c2[i] = gc2;
    if (out.screen)
      for (int q = 0; q < 6; ++q) out.screen[6 * i + q] = static_cast<float>(g[q]);
  }
}

__global__ __launch_bounds__(256) void k_zero_f64(double *p, int64_t n) {
  for (int64_t i = gtid_(); i < n; i += gstride_()) p[i] = 0.0;
}

unsigned grid_of(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > (1 << 20)) g = 1 << 20;
  return static_cast<unsigned>(g);
}


__global__ __launch_bounds__(256) void k_iota(uint32_t *v, uint32_t n) {
  for (int64_t i = gtid_(); i < n; i += gstride_()) v[i] = static_cast<uint32_t>(i);
}

// Per particle s of a pass (render position): the sum in FP64 of its list
// entries' partials, the entries in ascending list position (jv[poff[s] ..
// poff[s+1]), sorted stably by particle), added to the particle's gradient.
__global__ __launch_bounds__(256) void k_reduce_parts(const uint32_t *poff, const uint32_t *jv,
                                                      const float *part, const uint32_t *order,
                                                      uint32_t m, double *gscr) {
  for (int64_t s = gtid_(); s < m; s += gstride_()) {
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t t = poff[s]; t < poff[s + 1]; ++t) {
      const float *pp = part + 6 * static_cast<size_t>(jv[t]);
#pragma unroll
      for (int i = 0; i < 6; ++i) acc[i] += static_cast<double>(pp[i]);
    }
    double *g = gscr + 6 * static_cast<size_t>(order[s]);
#pragma unroll
    for (int i = 0; i < 6; ++i) g[i] += acc[i];
  }
}
}  // namespace

void launch_seg_count(const uint32_t *bin_off, uint32_t nunits, uint32_t *cnt, hipStream_t s) {
  k_seg_count<<<grid_of(nunits), 256, 0, s>>>(bin_off, nunits, cnt);
  hip_check(hipGetLastError(), "k_seg_count launch");
}

void launch_seg_emit(const uint32_t *bin_off, uint32_t nunits, const uint32_t *soff,
                     uint32_t *seg_unit, uint32_t *seg_first, hipStream_t s) {
  k_seg_emit<<<grid_of(nunits), 256, 0, s>>>(bin_off, nunits, soff, seg_unit, seg_first);
  hip_check(hipGetLastError(), "k_seg_emit launch");
}

void launch_adjoint(bool additive, const uint32_t *seg_unit, const uint32_t *seg_first,
                    uint32_t nseg, const uint32_t *soff, uint32_t nunits, const SplatArgs &a,
                    const float *grad, float *part, double *segstate, hipStream_t s) {
  if (nseg == 0) return;
  const unsigned grid = (nseg + 3) / 4;
  if (additive) {
    k_adjoint_add<<<grid, 256, 0, s>>>(a, seg_unit, seg_first, nseg, grad, part);
    hip_check(hipGetLastError(), "k_adjoint_add launch");
    return;
  }
  const size_t per = static_cast<size_t>(nseg) * 3 * 64;
  const SegView sv{segstate, segstate + per, segstate + 2 * per};
  k_adjoint_ord<0><<<grid, 256, 0, s>>>(a, seg_unit, seg_first, nseg, grad, part, sv);
  hip_check(hipGetLastError(), "k_adjoint_ord (segment maps) launch");
  k_adjoint_join<<<nunits, 192, 0, s>>>(soff, nunits, sv);
  hip_check(hipGetLastError(), "k_adjoint_join launch");
  k_adjoint_ord<2><<<grid, 256, 0, s>>>(a, seg_unit, seg_first, nseg, grad, part, sv);
  hip_check(hipGetLastError(), "k_adjoint_ord (gradients) launch");
}

void launch_iota(uint32_t *v, uint32_t n, hipStream_t s) {
  if (n == 0) return;
  k_iota<<<grid_of(n), 256, 0, s>>>(v, n);
  hip_check(hipGetLastError(), "k_iota launch");
}

void launch_reduce_parts(const uint32_t *poff, const uint32_t *jv, const float *part,
                         const uint32_t *order, uint32_t m, double *gscr, hipStream_t s) {
  if (m == 0) return;
  k_reduce_parts<<<grid_of(m), 256, 0, s>>>(poff, jv, part, order, m, gscr);
  hip_check(hipGetLastError(), "k_reduce_parts launch");
}

void launch_chain(const ProjectArgs *jobs, int64_t n, const uint8_t *act, const double *gscr,
                  const splhip_grads &out, hipStream_t s) {
  if (n == 0) return;
  k_chain<<<grid_of(n), 256, 0, s>>>(jobs, act, gscr, out);
  hip_check(hipGetLastError(), "k_chain launch");
}

void launch_zero_f64(double *p, int64_t n, hipStream_t s) {
  if (n == 0) return;
  k_zero_f64<<<grid_of(n), 256, 0, s>>>(p, n);
  hip_check(hipGetLastError(), "k_zero_f64 launch");
}

}  // namespace splhip
