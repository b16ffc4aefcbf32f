// Wave-per-block splat (8 x 8 pixel blocks) with heaviest-first scheduling.
//
// A work unit is one 8x8 pixel block of a reference tile (blocks aligned to
// the tile origin); its particle list holds exactly the particles whose
// footprint box (splotch_host.cc:462-472, tile frame) overlaps the block, in
// render order.  One wave owns one unit: lane = pixel.  Per batch of 64
// particles, lane l tabulates particle l's 8 column ranges + Gaussian column
// factors and 8 row factors (splotch_host.cc:596-607) into the wave's LDS
// slice; then the wave walks the batch in order and every lane updates its
// pixel exactly as render_new does, so each pixel sees the reference's
// summation / compositing order and rounding.
//
// Scheduling: units are issued heaviest first (work items sorted by list
// length, descending).  In the ordered mode the per-pixel chain is serial, so a
// long list is the critical path; units with at least `split` entries are
// issued as three items, one per colour channel (the channels' chains are
// independent), which cuts the serial work per lane to about a third.
#include "kernels.hpp"

#include "device_util.hpp"
#include "workspace.hpp"

namespace splhip {
namespace {

__device__ __forceinline__ int64_t gtid() {
  return static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
}
__device__ __forceinline__ int64_t gstride() {
  return static_cast<int64_t>(gridDim.x) * blockDim.x;
}

// exptable op with the x86 conversion semantics folded in: p = arg*expfac;
// p >= 2^31 or NaN -> INT_MIN -> 1; p in [2^20, 2^31) -> 0; p <= -1 -> negative
// index -> 1; otherwise the table product at trunc(p).
__device__ __forceinline__ float xop(const float *tab, float arg, float expfac) {
  const float p = arg * expfac;
  if (p < 1048576.f) {
    if (p > -1.f) {
      const int i = static_cast<int>(p);
      return tab[i >> 10] * tab[1024 + (i & 1023)];
    }
    return 1.f;
  }
  return (p < 2147483648.f) ? 0.f : 1.f;
}

__device__ __forceinline__ float xexpm1(const float *tab, float arg, float expfac, float tl) {
  if (fabsf(arg) < tl) return arg;
  return xop(tab, arg, expfac) - 1.f;
}

// Branch-free xexpm1 (same results as xop/xexpm1): the table is always read
// at an in-range index so that several lookups can be in flight at once.
__device__ __forceinline__ float xexpm1_bf(const float *tab, float arg, float expfac, float tl) {
  const float p = arg * expfac;
  const bool hi_ok = p < 1048576.f;
  const bool ok = hi_ok && (p > -1.f);
  const int i = static_cast<int>(ok ? p : 0.f);
  const float prod = tab[i >> 10] * tab[1024 + (i & 1023)];
  const float opv = ok ? prod : ((hi_ok || !(p < 2147483648.f)) ? 1.f : 0.f);
  return (fabsf(arg) < tl) ? arg : opv - 1.f;
}

// exptable::expm1 (splotchutils.h:203-214) split in two for pipelining:
// xexp_index gives the table index for p = arg*expfac (0 where the table does
// not apply), and xexpm1_finish combines the two table values t1, t2 fetched
// at that index with the out-of-range and Taylor cases.
__device__ __forceinline__ float xexpm1_finish(float arg, float p, float t1, float t2, float tl) {
  const bool hi_ok = p < 1048576.f;
  const bool ok = hi_ok && (p > -1.f);
  const float opv = ok ? t1 * t2 : ((hi_ok || !(p < 2147483648.f)) ? 1.f : 0.f);
  return (fabsf(arg) < tl) ? arg : opv - 1.f;
}

__device__ __forceinline__ int xexp_index(float p) {
  const bool ok = (p < 1048576.f) && (p > -1.f);
  return static_cast<int>(ok ? p : 0.f);
}

// Chunked walk of one batch for one channel class CH (0..2 one channel, 3 all):
// for J particles at a time, phase 1 issues every LDS read of the chunk (column
// masks, Gaussian factors, then the exp-table entries of each fragment) and
// phase 2 applies the J updates in list order.  No loaded value is carried
// across chunks, so the LDS latency is paid twice per chunk, not per particle.
// The code is branch-free: loads past the batch end read a valid slot and are
// masked, and a pixel outside a footprint adds +0.  The accumulators start at
// +0 and a sum of floats is -0 only if both terms are -0, so they never hold
// -0 and l + 0 == l bit for bit: the result equals the reference's skip.
template <int ADD, int CH, int J>
__device__ __forceinline__ void walk_chunks(int nb, int cx, int cy, const uint32_t (*cmask)[kPad],
                                            const float (*pre2)[kPad], const float (*pre1)[kPad],
                                            const float *tab, float va0, float va1, float va2,
                                            float vq0, float vq1, float vq2, float vmax,
                                            float expfac, float tl, float &l0, float &l1,
                                            float &l2) {
  constexpr bool C0 = (CH == 3 || CH == 0), C1 = (CH == 3 || CH == 1), C2 = (CH == 3 || CH == 2);
  for (int k0 = 0; k0 < nb; k0 += J) {
    float att[J], g0[J], g1[J], g2[J], p0[J], p1[J], p2[J];
    float u0[J], w0[J], u1[J], w1[J], u2[J], w2[J];
    uint32_t msk[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int k = min(k0 + j, nb - 1);  // uniform; out-of-batch slots masked below
      const uint32_t mk = cmask[cx][k];
      msk[j] = (k0 + j < nb) ? mk : 0u;
      att[j] = pre1[cy][k] * pre2[cx][k];
    }
    if (!ADD) {
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int k = min(k0 + j, nb - 1);
        if (C0) { g0[j] = att[j] * lane_f(va0, k); p0[j] = g0[j] * expfac; const int i = xexp_index(p0[j]); u0[j] = tab[i >> 10]; w0[j] = tab[1024 + (i & 1023)]; }
        if (C1) { g1[j] = att[j] * lane_f(va1, k); p1[j] = g1[j] * expfac; const int i = xexp_index(p1[j]); u1[j] = tab[i >> 10]; w1[j] = tab[1024 + (i & 1023)]; }
        if (C2) { g2[j] = att[j] * lane_f(va2, k); p2[j] = g2[j] * expfac; const int i = xexp_index(p2[j]); u2[j] = tab[i >> 10]; w2[j] = tab[1024 + (i & 1023)]; }
      }
    }
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int k = min(k0 + j, nb - 1);
      const bool in = (msk[j] >> cy) & 1u;
      const float at = att[j];
      if (ADD) {
        if (C0) { const float t = lane_f(va0, k) * at; l0 = l0 + (in ? t : 0.f); }
        if (C1) { const float t = lane_f(va1, k) * at; l1 = l1 + (in ? t : 0.f); }
        if (C2) { const float t = lane_f(va2, k) * at; l2 = l2 + (in ? t : 0.f); }
      } else {
        const bool tay = (lane_f(vmax, k) * at) < tl;
        if (C0) {
          const float em = xexpm1_finish(g0[j], p0[j], u0[j], w0[j], tl);
          const float d = l0 - lane_f(vq0, k);
          const float ta = (d * lane_f(va0, k)) * at, te = em * d;
          l0 = l0 + (in ? (tay ? ta : te) : 0.f);
        }
        if (C1) {
          const float em = xexpm1_finish(g1[j], p1[j], u1[j], w1[j], tl);
          const float d = l1 - lane_f(vq1, k);
          const float ta = (d * lane_f(va1, k)) * at, te = em * d;
          l1 = l1 + (in ? (tay ? ta : te) : 0.f);
        }
        if (C2) {
          const float em = xexpm1_finish(g2[j], p2[j], u2[j], w2[j], tl);
          const float d = l2 - lane_f(vq2, k);
          const float ta = (d * lane_f(va2, k)) * at, te = em * d;
          l2 = l2 + (in ? (tay ? ta : te) : 0.f);
        }
      }
    }
  }
}

// Software-pipelined walk of one batch for one channel class CH (0..2 one
// channel, 3 all).  Iteration k updates particle k with table values fetched in
// iteration k-1, computes particle k+1's attenuation and issues its table
// reads, and loads particle k+2's column mask and Gaussian factors, so that the
// only serial dependency per step is the pixel's own update.
template <int ADD, int CH>
__device__ __forceinline__ void walk_batch(int nb, int cx, int cy, const uint32_t (*cmask)[kPad],
                                           const float (*pre2)[kPad], const float (*pre1)[kPad],
                                           const float *tab, float va0, float va1, float va2,
                                           float vq0, float vq1, float vq2, float vmax,
                                           float expfac, float tl, float &l0, float &l1,
                                           float &l2) {
  constexpr bool C0 = (CH == 3 || CH == 0), C1 = (CH == 3 || CH == 1), C2 = (CH == 3 || CH == 2);
  // stage inputs
  uint32_t m_n = cmask[cx][0];                               // mask of k (current)
  float att_n = pre1[cy][0] * pre2[cx][0];                    // att of k
  uint32_t m_nn = 0;
  float p1_nn = 0.f, p2_nn = 0.f;
  if (nb > 1) {
    m_nn = cmask[cx][1];
    p1_nn = pre1[cy][1];
    p2_nn = pre2[cx][1];
  }
  float arg0 = 0.f, arg1 = 0.f, arg2 = 0.f, pp0 = 0.f, pp1 = 0.f, pp2 = 0.f;
  float t10 = 0.f, t20 = 0.f, t11 = 0.f, t21 = 0.f, t12 = 0.f, t22 = 0.f;
  if (!ADD) {
    const float e_ = expfac;
    if (C0) { arg0 = att_n * lane_f(va0, 0); pp0 = arg0 * e_; const int i = xexp_index(pp0); t10 = tab[i >> 10]; t20 = tab[1024 + (i & 1023)]; }
    if (C1) { arg1 = att_n * lane_f(va1, 0); pp1 = arg1 * e_; const int i = xexp_index(pp1); t11 = tab[i >> 10]; t21 = tab[1024 + (i & 1023)]; }
    if (C2) { arg2 = att_n * lane_f(va2, 0); pp2 = arg2 * e_; const int i = xexp_index(pp2); t12 = tab[i >> 10]; t22 = tab[1024 + (i & 1023)]; }
  }
  for (int k = 0; k < nb; ++k) {
    // current particle k: values computed / fetched in the previous step
    const uint32_t m = m_n;
    const float att = att_n;
    const float carg0 = arg0, carg1 = arg1, carg2 = arg2;
    const float cp0 = pp0, cp1 = pp1, cp2 = pp2;
    const float c10 = t10, c20 = t20, c11 = t11, c21 = t21, c12 = t12, c22 = t22;
    // particle k+1: attenuation from the prefetched factors, issue its table reads
    if (k + 1 < nb) {
      m_n = m_nn;
      att_n = p1_nn * p2_nn;
      if (!ADD) {
        if (C0) { arg0 = att_n * lane_f(va0, k + 1); pp0 = arg0 * expfac; const int i = xexp_index(pp0); t10 = tab[i >> 10]; t20 = tab[1024 + (i & 1023)]; }
        if (C1) { arg1 = att_n * lane_f(va1, k + 1); pp1 = arg1 * expfac; const int i = xexp_index(pp1); t11 = tab[i >> 10]; t21 = tab[1024 + (i & 1023)]; }
        if (C2) { arg2 = att_n * lane_f(va2, k + 1); pp2 = arg2 * expfac; const int i = xexp_index(pp2); t12 = tab[i >> 10]; t22 = tab[1024 + (i & 1023)]; }
      }
      // particle k+2: prefetch column mask and factors
      if (k + 2 < nb) {
        m_nn = cmask[cx][k + 2];
        p1_nn = pre1[cy][k + 2];
        p2_nn = pre2[cx][k + 2];
      }
    }
    // update of particle k (splotch_host.cc:615 / 648-658)
    const bool in = (m >> cy) & 1u;
    if (ADD) {
      if (C0) l0 = in ? l0 + lane_f(va0, k) * att : l0;
      if (C1) l1 = in ? l1 + lane_f(va1, k) * att : l1;
      if (C2) l2 = in ? l2 + lane_f(va2, k) * att : l2;
    } else {
      const bool tay = (lane_f(vmax, k) * att) < tl;
      if (C0) {
        const float em = xexpm1_finish(carg0, cp0, c10, c20, tl);
        const float d = l0 - lane_f(vq0, k);
        const float t = tay ? (d * lane_f(va0, k)) * att : em * d;
        l0 = in ? l0 + t : l0;
      }
      if (C1) {
        const float em = xexpm1_finish(carg1, cp1, c11, c21, tl);
        const float d = l1 - lane_f(vq1, k);
        const float t = tay ? (d * lane_f(va1, k)) * att : em * d;
        l1 = in ? l1 + t : l1;
      }
      if (C2) {
        const float em = xexpm1_finish(carg2, cp2, c12, c22, tl);
        const float d = l2 - lane_f(vq2, k);
        const float t = tay ? (d * lane_f(va2, k)) * att : em * d;
        l2 = in ? l2 + t : l2;
      }
    }
  }
}

template <int ADD, int VAR>
__global__ __launch_bounds__(256) void k_splat8(SplatArgs a, const uint32_t *items,
                                                uint32_t nitems) {
  __shared__ float s_tab[2048];
  __shared__ uint32_t s_col[4][8][kPad];
  __shared__ float s_pre2[4][8][kPad];
  __shared__ float s_pre1[4][8][kPad];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int i = tid; i < 2048; i += 256) s_tab[i] = a.exptab[i];
  __syncthreads();
  const uint32_t it = blockIdx.x * 4u + wave;
  if (it >= nitems) return;
  const uint64_t t_start = a.trace ? __builtin_amdgcn_s_memrealtime() : 0;
  // wave-uniform values are moved to scalar registers so that loop bounds and
  // channel selection compile to scalar branches
  const uint32_t item = uniform(items[it]);
  const uint32_t u = item >> 2;
  const int chan = static_cast<int>(item & 3u);  // 3: all channels
  const UnitGeom ug = unit_geom_uniform(a, u);
  const int gy = ug.gy, gx = ug.gx, x0s = ug.x0s, y0s = ug.y0s, w = ug.w, h = ug.h;
  const int bxo = gx * 8, byo = gy * 8;  // tile frame
  const int cx = lane >> 3, cy = lane & 7;
  const int lx = bxo + cx, ly = byo + cy;
  uint32_t(*col)[kPad] = s_col[wave];
  float(*pre2)[kPad] = s_pre2[wave];
  float(*pre1)[kPad] = s_pre1[wave];
  const float expfac = a.expfac, tl = a.taylorlimit;

  float l0 = 0.f, l1 = 0.f, l2 = 0.f;
  const uint32_t e0 = uniform(a.bin_off[u]), e1 = uniform(a.bin_off[u + 1]);
  // long lists are the critical path: let their waves issue ahead of the rest
  if (e1 - e0 >= a.prio_len) __builtin_amdgcn_s_setprio(3);
  for (uint32_t base = e0; base < e1; base += 64) {
    const int nb = static_cast<int>(uniform(min<uint32_t>(64u, e1 - base)));
    // Phase A: lane = particle (splotch_host.cc:462-476 + 596-607, tile frame)
    float va0 = 0.f, va1 = 0.f, va2 = 0.f, vq0 = 0.f, vq1 = 0.f, vq2 = 0.f, vmax = 0.f;
    if (lane < nb) {
      const Rec R = a.recs[a.ent[base + lane]];
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
        uint32_t lohi = 0;
        float p2 = 0.f;
        if (x >= bx0 && x < bx1) {
          const float dxsq = (static_cast<float>(x) - posx) * (static_cast<float>(x) - posx);
          const float dy = sqrtf(radsq - dxsq);
          const int lo = max(by0, cvt_i32(posy - dy + 1));
          const int hi = min(by1, cvt_i32(posy + dy + 1));
          if (hi > lo) {
            if (VAR >= 3) {  // bit r: block row byo + r lies in [lo, hi)
              const int r0 = min(max(lo - byo, 0), 8), r1 = min(max(hi - byo, 0), 8);
              lohi = (r1 > r0) ? (((1u << r1) - 1u) & ~((1u << r0) - 1u)) : 0u;
            } else {
              lohi = static_cast<uint32_t>(lo) | (static_cast<uint32_t>(hi) << 16);
            }
          }
          p2 = xop(s_tab, stp * dxsq, expfac);
        }
        col[c][lane] = lohi;
        pre2[c][lane] = p2;
      }
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int y = byo + r;
        float p1 = 0.f;
        if (y >= by0 && y < by1) {
          const float dyf = static_cast<float>(y) - posy;
          p1 = xop(s_tab, stp * dyf * dyf, expfac);
        }
        pre1[r][lane] = p1;
      }
      va0 = -R.e0;
      va1 = -R.e1;
      va2 = -R.e2;
      if (!ADD) {
        const float g = ug.grayabsorb;
        vq0 = R.e0 / (R.e0 + g);
        vq1 = R.e1 / (R.e1 + g);
        vq2 = R.e2 / (R.e2 + g);
        vmax = fmaxf(fabsf(va0), fmaxf(fabsf(va1), fabsf(va2)));
      }
    }
    wave_sync();
    // Phase B: lane = pixel, particles in list order.
    if (VAR == 4) {
      switch (chan) {
        case 0: walk_chunks<ADD, 0, 8>(nb, cx, cy, col, pre2, pre1, s_tab, va0, va1, va2, vq0, vq1, vq2, vmax, expfac, tl, l0, l1, l2); break;
        case 1: walk_chunks<ADD, 1, 8>(nb, cx, cy, col, pre2, pre1, s_tab, va0, va1, va2, vq0, vq1, vq2, vmax, expfac, tl, l0, l1, l2); break;
        case 2: walk_chunks<ADD, 2, 8>(nb, cx, cy, col, pre2, pre1, s_tab, va0, va1, va2, vq0, vq1, vq2, vmax, expfac, tl, l0, l1, l2); break;
        default: walk_chunks<ADD, 3, (ADD ? 8 : 4)>(nb, cx, cy, col, pre2, pre1, s_tab, va0, va1, va2, vq0, vq1, vq2, vmax, expfac, tl, l0, l1, l2); break;
      }
      wave_sync();
      continue;
    }
    if (VAR == 3) {
      switch (chan) {
        case 0: walk_batch<ADD, 0>(nb, cx, cy, col, pre2, pre1, s_tab, va0, va1, va2, vq0, vq1, vq2, vmax, expfac, tl, l0, l1, l2); break;
        case 1: walk_batch<ADD, 1>(nb, cx, cy, col, pre2, pre1, s_tab, va0, va1, va2, vq0, vq1, vq2, vmax, expfac, tl, l0, l1, l2); break;
        case 2: walk_batch<ADD, 2>(nb, cx, cy, col, pre2, pre1, s_tab, va0, va1, va2, vq0, vq1, vq2, vmax, expfac, tl, l0, l1, l2); break;
        default: walk_batch<ADD, 3>(nb, cx, cy, col, pre2, pre1, s_tab, va0, va1, va2, vq0, vq1, vq2, vmax, expfac, tl, l0, l1, l2); break;
      }
      wave_sync();
      continue;
    }
    if (VAR == 1) {
      for (int k = 0; k < nb; ++k) {
        const uint32_t lohi = col[cx][k];
        if (ly >= static_cast<int>(lohi & 0xffffu) && ly < static_cast<int>(lohi >> 16)) {
          const float att = pre1[cy][k] * pre2[cx][k];
          if (ADD) {
            if (chan == 3 || chan == 0) l0 += lane_f(va0, k) * att;
            if (chan == 3 || chan == 1) l1 += lane_f(va1, k) * att;
            if (chan == 3 || chan == 2) l2 += lane_f(va2, k) * att;
          } else {
            const float a0 = lane_f(va0, k), a1 = lane_f(va1, k), a2 = lane_f(va2, k);
            const float q0 = lane_f(vq0, k), q1 = lane_f(vq1, k), q2 = lane_f(vq2, k);
            if ((lane_f(vmax, k) * att) < tl) {
              if (chan == 3 || chan == 0) l0 += ((l0 - q0) * a0) * att;
              if (chan == 3 || chan == 1) l1 += ((l1 - q1) * a1) * att;
              if (chan == 3 || chan == 2) l2 += ((l2 - q2) * a2) * att;
            } else {
              if (chan == 3 || chan == 0) l0 += xexpm1(s_tab, att * a0, expfac, tl) * (l0 - q0);
              if (chan == 3 || chan == 1) l1 += xexpm1(s_tab, att * a1, expfac, tl) * (l1 - q1);
              if (chan == 3 || chan == 2) l2 += xexpm1(s_tab, att * a2, expfac, tl) * (l2 - q2);
            }
          }
        }
      }
      wave_sync();
      continue;
    }
    // 4 particles per step: the table reads of a step are issued together and
    // only the update chain is serial.
    for (int k0 = 0; k0 < nb; k0 += 4) {
      float att[4];
      bool in[4];
      int kk[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        kk[j] = min(k0 + j, nb - 1);
        const uint32_t lohi = col[cx][kk[j]];
        in[j] = (k0 + j < nb) && ly >= static_cast<int>(lohi & 0xffffu) &&
                ly < static_cast<int>(lohi >> 16);
        att[j] = pre1[cy][kk[j]] * pre2[cx][kk[j]];
      }
      if (ADD) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const float a0 = lane_f(va0, kk[j]), a1 = lane_f(va1, kk[j]), a2 = lane_f(va2, kk[j]);
          if (chan == 3 || chan == 0) l0 = in[j] ? l0 + a0 * att[j] : l0;
          if (chan == 3 || chan == 1) l1 = in[j] ? l1 + a1 * att[j] : l1;
          if (chan == 3 || chan == 2) l2 = in[j] ? l2 + a2 * att[j] : l2;
        }
      } else {
        float em0[4], em1[4], em2[4];
        bool tay[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int k = kk[j];
          tay[j] = (lane_f(vmax, k) * att[j]) < tl;
          if (chan == 3 || chan == 0) em0[j] = xexpm1_bf(s_tab, att[j] * lane_f(va0, k), expfac, tl);
          if (chan == 3 || chan == 1) em1[j] = xexpm1_bf(s_tab, att[j] * lane_f(va1, k), expfac, tl);
          if (chan == 3 || chan == 2) em2[j] = xexpm1_bf(s_tab, att[j] * lane_f(va2, k), expfac, tl);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int k = kk[j];
          const float at = att[j];
          if (chan == 3 || chan == 0) {
            const float a0 = lane_f(va0, k), q0 = lane_f(vq0, k);
            const float t = tay[j] ? ((l0 - q0) * a0) * at : em0[j] * (l0 - q0);
            l0 = in[j] ? l0 + t : l0;
          }
          if (chan == 3 || chan == 1) {
            const float a1 = lane_f(va1, k), q1 = lane_f(vq1, k);
            const float t = tay[j] ? ((l1 - q1) * a1) * at : em1[j] * (l1 - q1);
            l1 = in[j] ? l1 + t : l1;
          }
          if (chan == 3 || chan == 2) {
            const float a2 = lane_f(va2, k), q2 = lane_f(vq2, k);
            const float t = tay[j] ? ((l2 - q2) * a2) * at : em2[j] * (l2 - q2);
            l2 = in[j] ? l2 + t : l2;
          }
        }
      }
    }
    wave_sync();
  }
  if (a.trace && lane == 0) {
    uint64_t *tr = a.trace + 4 * static_cast<uint64_t>(it);
    tr[0] = t_start;
    tr[1] = __builtin_amdgcn_s_memrealtime();
    tr[2] = static_cast<uint64_t>(e1 - e0) | (static_cast<uint64_t>(chan) << 32);
    tr[3] = static_cast<uint64_t>(__smid());
  }
  if (lx < min(bxo + 8, w) && ly < min(byo + 8, h)) {
    float *d = ug.pic + (static_cast<int64_t>(x0s + lx) * ug.yres + (y0s + ly)) * 3;
    if (chan == 3 || chan == 0) d[0] = ug.accumulate ? d[0] + l0 : l0;
    if (chan == 3 || chan == 1) d[1] = ug.accumulate ? d[1] + l1 : l1;
    if (chan == 3 || chan == 2) d[2] = ug.accumulate ? d[2] + l2 : l2;
  }
}

// ---------------------------------------------------------------------------
// Heavy blocks: producer / consumer workgroup (16 waves) per 8x8 block.
//
// A block whose list holds hundreds of thousands of particles is the critical
// path of the frame: its pixels' update chains are serial.  Here the serial
// part is all a consumer wave does: waves 0..2 own channel 0..2 of the block's
// 64 pixels and, per particle, apply
//     l += (l - q) * m1 * m2       (two roundings, left to right)
// with multipliers prepared by the producers so that this reproduces
// render_new's update bit for bit (splotch_host.cc:615, 648-658):
//   Taylor branch  m1 = a,  m2 = att     ->  ((l - q) * a) * att
//   expm1 branch   m1 = expm1(att * a), m2 = 1  ->  (l - q) * em   (x*1 == x)
//   additive       l += a * m2, m2 = att
//   outside        m2 = -1: no update.
// Waves 3..15 (832 lanes) run a three-stage pipeline ahead of the consumers:
// particle set-up three batches ahead (tile-frame position, box, Gaussian
// width, emission / absorption; splotch_host.cc:462-476, 591-632), the
// separable column / row tables two batches ahead (splotch_host.cc:596-607),
// and the per-fragment multipliers one batch ahead.
constexpr int kHB = 32;  // particles per batch
constexpr int kHWaves = 16;
constexpr int kHProd = (kHWaves - 3) * 64;

struct HeavyPart {
  float posx, posy, radsq, stp;
  int bx0, bx1, by0, by1;
};

template <int ADD>
__global__ __launch_bounds__(1024) void k_splat_heavy(SplatArgs a, const uint32_t *items,
                                                      uint32_t nheavy) {
  __shared__ float s_tab[2048];
  __shared__ HeavyPart s_part[4][kHB];
  __shared__ float4 s_cst[4][3][kHB];  // [slot][channel][k] = (a, q, maxa, 0)
  __shared__ uint32_t s_cm[2][kHB][8];  // column row-masks
  __shared__ float s_p2[2][kHB][8];     // column factors
  __shared__ float s_p1[2][kHB][8];     // row factors
  __shared__ float s_m2[2][kHB][64];
  __shared__ float s_m1[ADD ? 1 : 2][ADD ? 1 : 3][ADD ? 1 : kHB][64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool producer = wave >= 3;
  const int ptid = tid - 192;
  for (int i = tid; i < 2048; i += 1024) s_tab[i] = a.exptab[i];
  const uint32_t it = blockIdx.x;
  const uint64_t t_start = a.trace ? __builtin_amdgcn_s_memrealtime() : 0;
  const uint32_t u = uniform(items[it]) >> 2;
  const UnitGeom ug = unit_geom_uniform(a, u);
  const int gy = ug.gy, gx = ug.gx, x0s = ug.x0s, y0s = ug.y0s, w = ug.w, h = ug.h;
  const int bxo = gx * 8, byo = gy * 8;
  const uint32_t e0 = uniform(a.bin_off[u]), e1 = uniform(a.bin_off[u + 1]);
  const int n = static_cast<int>(e1 - e0);
  const int nbat = (n + kHB - 1) / kHB;
  const float expfac = a.expfac, tl = a.taylorlimit;
  __builtin_amdgcn_s_setprio(3);

  auto count = [&](int b) { return min(kHB, n - b * kHB); };
  auto setup = [&](int b) {  // particle b-batch set-up into slot b % 4
    const int slot = b & 3;
    if (ptid < count(b)) {
      const Rec R = a.recs[a.ent[e0 + b * kHB + ptid]];
      const float rfacr = R.r * a.rfac;
      const float posx = R.x - static_cast<float>(x0s), posy = R.y - static_cast<float>(y0s);
      HeavyPart P;
      P.posx = posx;
      P.posy = posy;
      P.bx0 = max(cvt_i32(posx - rfacr + 1), 0);
      P.bx1 = min(cvt_i32(posx + rfacr + 1), w);
      P.by0 = max(cvt_i32(posy - rfacr + 1), 0);
      P.by1 = min(cvt_i32(posy + rfacr + 1), h);
      P.radsq = rfacr * rfacr;
      const float sigma = a.h2sigma * R.r;
      P.stp = -1.f / (sigma * sigma);
      s_part[slot][ptid] = P;
      const float a0 = -R.e0, a1 = -R.e1, a2 = -R.e2;
      float q0 = 0.f, q1 = 0.f, q2 = 0.f, mx = 0.f;
      if (!ADD) {
        const float g = ug.grayabsorb;
        q0 = R.e0 / (R.e0 + g);
        q1 = R.e1 / (R.e1 + g);
        q2 = R.e2 / (R.e2 + g);
        mx = fmaxf(fabsf(a0), fmaxf(fabsf(a1), fabsf(a2)));
      }
      s_cst[slot][0][ptid] = make_float4(a0, q0, mx, 0.f);
      s_cst[slot][1][ptid] = make_float4(a1, q1, mx, 0.f);
      s_cst[slot][2][ptid] = make_float4(a2, q2, mx, 0.f);
    }
  };
  auto tables = [&](int b) {  // column ranges / factors and row factors into b & 1
    const int slot = b & 3, ts = b & 1;
    if (ptid < count(b) * 16) {
      const int k = ptid >> 4, j = ptid & 15;
      const HeavyPart P = s_part[slot][k];
      if (j < 8) {
        const int x = bxo + j;
        uint32_t m = 0;
        float p2 = 0.f;
        if (x >= P.bx0 && x < P.bx1) {
          const float dxsq = (static_cast<float>(x) - P.posx) * (static_cast<float>(x) - P.posx);
          const float dy = sqrtf(P.radsq - dxsq);
          const int lo = max(P.by0, cvt_i32(P.posy - dy + 1));
          const int hi = min(P.by1, cvt_i32(P.posy + dy + 1));
          const int r0 = min(max(lo - byo, 0), 8), r1 = min(max(hi - byo, 0), 8);
          if (hi > lo && r1 > r0) m = ((1u << r1) - 1u) & ~((1u << r0) - 1u);
          p2 = xop(s_tab, P.stp * dxsq, expfac);
        }
        s_cm[ts][k][j] = m;
        s_p2[ts][k][j] = p2;
      } else {
        const int y = byo + (j - 8);
        float p1 = 0.f;
        if (y >= P.by0 && y < P.by1) {
          const float dyf = static_cast<float>(y) - P.posy;
          p1 = xop(s_tab, P.stp * dyf * dyf, expfac);
        }
        s_p1[ts][k][j - 8] = p1;
      }
    }
  };
  auto fragments = [&](int b) {  // per-fragment multipliers into b & 1
    const int slot = b & 3, ts = b & 1;
    const int cnt = count(b);
    for (int f = ptid; f < cnt * 64; f += kHProd) {
      const int k = f >> 6, p = f & 63, c = p >> 3, r = p & 7;
      const bool in = (s_cm[ts][k][c] >> r) & 1u;
      const float att = s_p1[ts][k][r] * s_p2[ts][k][c];
      if (ADD) {
        s_m2[ts][k][p] = in ? att : -1.f;
      } else {
        const float4 c0 = s_cst[slot][0][k], c1 = s_cst[slot][1][k], c2 = s_cst[slot][2][k];
        const bool tay = (c0.z * att) < tl;
        s_m2[ts][k][p] = in ? (tay ? att : 1.f) : -1.f;
        const float e0v = xexpm1_bf(s_tab, att * c0.x, expfac, tl);
        const float e1v = xexpm1_bf(s_tab, att * c1.x, expfac, tl);
        const float e2v = xexpm1_bf(s_tab, att * c2.x, expfac, tl);
        s_m1[ADD ? 0 : ts][0][k][p] = tay ? c0.x : e0v;
        s_m1[ADD ? 0 : ts][1][k][p] = tay ? c1.x : e1v;
        s_m1[ADD ? 0 : ts][2][k][p] = tay ? c2.x : e2v;
      }
    }
  };

  float l = 0.f;  // consumer: channel `wave`, pixel `lane`
  for (int s = -3; s < nbat; ++s) {
    if (producer) {
      if (s + 3 < nbat) setup(s + 3);
      if (s + 2 >= 0 && s + 2 < nbat) tables(s + 2);
      if (s + 1 >= 0 && s + 1 < nbat) fragments(s + 1);
    } else if (s >= 0) {
      const int slot = s & 3, fb = s & 1, c = wave;
      const int cnt = count(s);
      const float *m2b = &s_m2[fb][0][0];
      const float *m1b = &s_m1[ADD ? 0 : fb][ADD ? 0 : c][0][0];
      const float4 *cst = s_cst[slot][c];
      for (int k0 = 0; k0 < cnt; k0 += 8) {
        float m1[8], m2[8], cv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int k = min(k0 + j, cnt - 1);
          const float v = m2b[k * 64 + lane];
          m2[j] = (k0 + j < cnt) ? v : -1.f;
          if (!ADD) m1[j] = m1b[k * 64 + lane];
          cv[j] = ADD ? cst[k].x : cst[k].y;  // a (additive) or q
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const bool in = m2[j] >= 0.f;
          if (ADD) {
            const float t = cv[j] * m2[j];
            l = l + (in ? t : 0.f);
          } else {
            const float t = ((l - cv[j]) * m1[j]) * m2[j];
            l = l + (in ? t : 0.f);
          }
        }
      }
    }
    __syncthreads();
  }
  if (a.trace && tid == 0) {
    uint64_t *tr = a.trace + 4 * static_cast<uint64_t>(it);
    tr[0] = t_start;
    tr[1] = __builtin_amdgcn_s_memrealtime();
    tr[2] = static_cast<uint64_t>(n) | (uint64_t(3) << 32);
    tr[3] = static_cast<uint64_t>(__smid());
  }
  if (!producer) {
    const int lx = bxo + (lane >> 3), ly = byo + (lane & 7);
    if (lx < min(bxo + 8, w) && ly < min(byo + 8, h)) {
      float *d = ug.pic + (static_cast<int64_t>(x0s + lx) * ug.yres + (y0s + ly)) * 3 + wave;
      *d = ug.accumulate ? *d + l : l;
    }
  }
}

// Work items per unit: none for units outside their tile; none for empty units
// when accumulating; three (one per channel) for long ordered lists; else one.
// (need_empty: the job's pass writes its whole image, so empty blocks are
// written too.)
__device__ __forceinline__ uint32_t unit_items(const SplatArgs &a, uint32_t u, int ordered,
                                               uint32_t split, uint32_t &len) {
  len = a.bin_off[u + 1] - a.bin_off[u];
  const UnitGeom g = unit_geom(a, u);
  if (g.gx * 8 >= g.w || g.gy * 8 >= g.h) return 0;
  if (len == 0) return g.need_empty ? 1u : 0u;
  if (len >= a.heavy_len) return 1u;  // one producer/consumer workgroup for all channels
  return (len >= (ordered ? split : 2 * split)) ? 3u : 1u;
}

__global__ __launch_bounds__(256) void k_items_count(SplatArgs a, uint32_t nunits, int ordered,
                                                     uint32_t split, uint32_t *cnt,
                                                     uint32_t *heavy_cnt) {
  for (int64_t u = gtid(); u < nunits; u += gstride()) {
    uint32_t len;
    const uint32_t k = unit_items(a, static_cast<uint32_t>(u), ordered, split, len);
    cnt[u] = k;
    if (heavy_cnt && k == 1 && len >= a.heavy_len) atomicAdd(heavy_cnt, 1u);
  }
}

__global__ __launch_bounds__(256) void k_items_emit(SplatArgs a, uint32_t nunits, int ordered,
                                                    uint32_t split, const uint32_t *off,
                                                    uint32_t *keys, uint32_t *vals) {
  for (int64_t u = gtid(); u < nunits; u += gstride()) {
    uint32_t len;
    const uint32_t k = unit_items(a, static_cast<uint32_t>(u), ordered, split, len);
    const uint32_t o = off[u];
    for (uint32_t c = 0; c < k; ++c) {
      keys[o + c] = ~len;  // ascending sort of ~len = heaviest first
      vals[o + c] = (static_cast<uint32_t>(u) << 2) | (k == 3 ? c : 3u);
    }
  }
}

unsigned grid_for(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > (1 << 20)) g = 1 << 20;
  return static_cast<unsigned>(g);
}

}  // namespace

void launch_items_count(const uint32_t *bin_off, uint32_t nunits, const SplatArgs &a,
                        int ordered, uint32_t split, uint32_t *cnt, uint32_t *heavy_cnt,
                        hipStream_t s) {
  if (nunits == 0) return;
  SplatArgs b = a;
  b.bin_off = bin_off;
  k_items_count<<<grid_for(nunits), 256, 0, s>>>(b, nunits, ordered, split, cnt, heavy_cnt);
  hip_check(hipGetLastError(), "k_items_count launch");
}

void launch_items_emit(const uint32_t *bin_off, uint32_t nunits, const SplatArgs &a, int ordered,
                       uint32_t split, const uint32_t *off, uint32_t *keys, uint32_t *vals,
                       hipStream_t s) {
  if (nunits == 0) return;
  SplatArgs b = a;
  b.bin_off = bin_off;
  k_items_emit<<<grid_for(nunits), 256, 0, s>>>(b, nunits, ordered, split, off, keys, vals);
  hip_check(hipGetLastError(), "k_items_emit launch");
}

void launch_splat_heavy(bool additive, const uint32_t *items, uint32_t nheavy, const SplatArgs &a,
                        hipStream_t s) {
  if (nheavy == 0) return;
  if (additive) k_splat_heavy<1><<<nheavy, 1024, 0, s>>>(a, items, nheavy);
  else k_splat_heavy<0><<<nheavy, 1024, 0, s>>>(a, items, nheavy);
  hip_check(hipGetLastError(), "k_splat_heavy launch");
}

void launch_splat8(bool additive, int variant, const uint32_t *items, uint32_t nitems,
                   const SplatArgs &a, hipStream_t s) {
  if (nitems == 0) return;
  const unsigned grid = (nitems + 3) / 4;
  if (variant == 4) {
    if (additive) k_splat8<1, 4><<<grid, 256, 0, s>>>(a, items, nitems);
    else k_splat8<0, 4><<<grid, 256, 0, s>>>(a, items, nitems);
  } else if (variant == 3) {
    if (additive) k_splat8<1, 3><<<grid, 256, 0, s>>>(a, items, nitems);
    else k_splat8<0, 3><<<grid, 256, 0, s>>>(a, items, nitems);
  } else if (variant == 2) {
    if (additive) k_splat8<1, 2><<<grid, 256, 0, s>>>(a, items, nitems);
    else k_splat8<0, 2><<<grid, 256, 0, s>>>(a, items, nitems);
  } else {
    if (additive) k_splat8<1, 1><<<grid, 256, 0, s>>>(a, items, nitems);
    else k_splat8<0, 1><<<grid, 256, 0, s>>>(a, items, nitems);
  }
  hip_check(hipGetLastError(), "k_splat8 launch");
}

}  // namespace splhip
