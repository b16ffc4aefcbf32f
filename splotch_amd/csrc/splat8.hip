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

#include <algorithm>
#include <type_traits>

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

// xop in two steps, so that several lookups can be in flight: xop_reads
// reads the two table entries for p = arg * expfac (at an in-range index in
// every case), xop_value gives xop's result from them.
__device__ __forceinline__ void xop_reads(const float *tab, float p, float &t1, float &t2) {
  const bool ok = (p < 1048576.f) && (p > -1.f);
  const int i = static_cast<int>(ok ? p : 0.f);
  t1 = tab[i >> 10];
  t2 = tab[1024 + (i & 1023)];
}
__device__ __forceinline__ float xop_value(float p, float t1, float t2) {
  const bool hi_ok = p < 1048576.f;
  const bool ok = hi_ok && (p > -1.f);
  return ok ? t1 * t2 : ((hi_ok || !(p < 2147483648.f)) ? 1.f : 0.f);
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

// Absorption range of the fast table path: att * a * expfac stays in [0, 2^20)
// for att in [0, 1] (expfac = 2^20 / -20).  A fast particle also has finite
// q_c (e_c / (e_c + gray_absorption) is not when gray_absorption <= 0 meets
// e_c = -gray_absorption): the fast paths apply it outside the footprint as
// l + ((l - q) * m1) * 0, which must be l.
constexpr float kFastA = -19.5f;

__device__ __forceinline__ bool finite3(float x, float y, float z) {
  const float inf = __builtin_inff();
  return fabsf(x) < inf && fabsf(y) < inf && fabsf(z) < inf;
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

// Table entries of the fast path, where p = arg * expfac lies in [0, 2^20):
// i4 = trunc(arg * 4 expfac) = 4 trunc(p) + (0..3) (scaling by 4 is exact), so
// the byte offsets of tab1[trunc(p) >> 10] and tab2[trunc(p) & 1023] are two
// masks of i4.
__device__ __forceinline__ void xtab_fast(const float *tab, float arg, float expfac4, float &t1,
                                          float &t2) {
  const uint32_t i4 = static_cast<uint32_t>(static_cast<int>(arg * expfac4));
  const char *b = reinterpret_cast<const char *>(tab);
  t1 = *reinterpret_cast<const float *>(b + ((i4 >> 10) & 0xffcu));
  t2 = *reinterpret_cast<const float *>(b + 4096 + (i4 & 0xffcu));
}

__device__ __forceinline__ int xexp_index(float p) {
  const bool ok = (p < 1048576.f) && (p > -1.f);
  return static_cast<int>(ok ? p : 0.f);
}

// x where row cy of the column mask m is set (the pixel is in the particle's
// footprint), else +0: the row's bit, sign-extended, masks x's bits.
__device__ __forceinline__ float in_footprint(float x, uint32_t m, int cy) {
  const uint32_t all = static_cast<uint32_t>(__builtin_amdgcn_sbfe(static_cast<int>(m), cy, 1));
  return __uint_as_float(__float_as_uint(x) & all);
}

// One fast chunk of the ordered walk (all particles of the chunk have every
// a_c in [kFastA, 0] and finite q_c) for channel class CH: J particles from
// slot k0, their constants in cst; bit j of TB set: particle k0 + j takes the
// Taylor branch at every pixel of its footprint in the block, so its update
// is l += ((l - q) * a) * att (att replaced by +0 outside the footprint) with
// no table work.
template <int CH, int J, uint32_t TB>
__device__ __forceinline__ void fast_chunk(int k0, int cx, int cy, const uint32_t (*cmask)[kPad],
                                           const float (*pre2)[kPad],
                                           const float (*pre1)[kPad], const float *tab,
                                           const float4 *cst, float expfac4, float tl,
                                           float &l0, float &l1, float &l2) {
  constexpr bool C0 = (CH == 3 || CH == 0), C1 = (CH == 3 || CH == 1), C2 = (CH == 3 || CH == 2);
  float m2[J], g0[J], g1[J], g2[J], u0[J], w0[J], u1[J], w1[J], u2[J], w2[J];
  float4 ca[J], cq[J];
  bool tay[J];
  // slots past the batch hold finite values of earlier batches (or
  // zeros) and are masked; table offsets are masked into the tables
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int k = k0 + j;
    ca[j] = cst[2 * k];
    cq[j] = cst[2 * k + 1];
  }
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int k = k0 + j;
    const float at = pre1[cy][k] * pre2[cx][k];
    if ((TB >> j) & 1u) {
      m2[j] = in_footprint(at, cmask[cx][k], cy);
      continue;
    }
    tay[j] = (ca[j].w * at) < tl;
    m2[j] = in_footprint(tay[j] ? at : 1.f, cmask[cx][k], cy);
    if (C0) { g0[j] = at * ca[j].x; xtab_fast(tab, g0[j], expfac4, u0[j], w0[j]); }
    if (C1) { g1[j] = at * ca[j].y; xtab_fast(tab, g1[j], expfac4, u1[j], w1[j]); }
    if (C2) { g2[j] = at * ca[j].z; xtab_fast(tab, g2[j], expfac4, u2[j], w2[j]); }
  }
#pragma unroll
  for (int j = 0; j < J; ++j) {
    if ((TB >> j) & 1u) {
      if (C0) l0 = l0 + ((l0 - cq[j].x) * ca[j].x) * m2[j];
      if (C1) l1 = l1 + ((l1 - cq[j].y) * ca[j].y) * m2[j];
      if (C2) l2 = l2 + ((l2 - cq[j].z) * ca[j].z) * m2[j];
      continue;
    }
    // the table product is formed before the select: written inside
    // it, the compiler may move the table reads into a branch that
    // waits for every LDS read in flight
    if (C0) {
      const float ex = u0[j] * w0[j] - 1.f;
      const float em = (fabsf(g0[j]) < tl) ? g0[j] : ex;
      const float m1 = tay[j] ? ca[j].x : em;
      l0 = l0 + ((l0 - cq[j].x) * m1) * m2[j];
    }
    if (C1) {
      const float ex = u1[j] * w1[j] - 1.f;
      const float em = (fabsf(g1[j]) < tl) ? g1[j] : ex;
      const float m1 = tay[j] ? ca[j].y : em;
      l1 = l1 + ((l1 - cq[j].y) * m1) * m2[j];
    }
    if (C2) {
      const float ex = u2[j] * w2[j] - 1.f;
      const float em = (fabsf(g2[j]) < tl) ? g2[j] : ex;
      const float m1 = tay[j] ? ca[j].z : em;
      l2 = l2 + ((l2 - cq[j].z) * m1) * m2[j];
    }
  }
}

// Chunked walk of one batch for one channel class CH (0..2 one channel, 3 all):
// for J particles at a time, phase 1 issues every LDS read of the chunk (column
// masks, Gaussian factors, then the exp-table entries of each fragment) and
// phase 2 applies the J updates in list order.  No loaded value is carried
// across chunks, so the LDS latency is paid twice per chunk, not per particle.
// The code is branch-free: loads past the batch end read slots whose column
// masks are 0 (phase A clears them), and a pixel outside a footprint adds +0.
// The accumulators start at +0 and a sum of floats is -0 only if both terms
// are -0, so they never hold -0 and l + 0 == l bit for bit: the result equals
// the reference's skip.
template <int ADD, int CH, int J>
__device__ __forceinline__ void walk_chunks(int nb, int cx, int cy, const uint32_t (*cmask)[kPad],
                                            const float (*pre2)[kPad], const float (*pre1)[kPad],
                                            const float *tab, float va0, float va1, float va2,
                                            float vq0, float vq1, float vq2, float vmax,
                                            const float4 *cst, uint64_t fastmask,
                                            uint64_t taymask, float expfac, float tl,
                                            float &l0, float &l1,
                                            float &l2) {
  constexpr bool C0 = (CH == 3 || CH == 0), C1 = (CH == 3 || CH == 1), C2 = (CH == 3 || CH == 2);
  static_assert(64 % J == 0, "chunks tile a batch: slot k0 + j < 64");
  const float expfac4 = 4.f * expfac;  // exact
  for (int k0 = 0; k0 < nb; k0 += J) {
    if (!ADD) {
      // Fast chunk: every particle of the chunk has all absorptions a_c in
      // [kFastA, 0], so arg = att*a_c (att in [0, 1]) gives p = arg*expfac in
      // [0, 2^20): the table index is trunc(p) with no range cases.  The
      // update is written l += ((l - q) * m1) * m2 with
      //   Taylor branch: m1 = a, m2 = att;  expm1 branch: m1 = em, m2 = 1
      //   (x * 1 == x);  outside the footprint: m2 = 0 (l + +-0 == l, as l is
      //   never -0),
      // which is render_new's update bit for bit (splotch_host.cc:648-658).
      // The particles' constants come from LDS (cst: (a0, a1, a2, max |a|),
      // (q0, q1, q2, 0)) as vector operands.
      const int cnt = min(J, nb - k0);
      const uint64_t need = ((1ull << cnt) - 1ull) << k0;
      if ((fastmask & need) == need) {
        // bit j: particle k0 + j takes the Taylor branch at every pixel of
        // its footprint in the block (wave-uniform): no table work for it.
        // Three-channel chunks of 4 have a straight-line body per pattern.
        const uint32_t tb = static_cast<uint32_t>(taymask >> k0);
        if (CH == 3 && J == 4) {
          switch (tb & 15u) {
#define SPLHIP_FAST(T) case T: fast_chunk<CH, J, T>(k0, cx, cy, cmask, pre2, pre1, tab, cst, expfac4, tl, l0, l1, l2); break;
            SPLHIP_FAST(0) SPLHIP_FAST(1) SPLHIP_FAST(2) SPLHIP_FAST(3)
            SPLHIP_FAST(4) SPLHIP_FAST(5) SPLHIP_FAST(6) SPLHIP_FAST(7)
            SPLHIP_FAST(8) SPLHIP_FAST(9) SPLHIP_FAST(10) SPLHIP_FAST(11)
            SPLHIP_FAST(12) SPLHIP_FAST(13) SPLHIP_FAST(14) SPLHIP_FAST(15)
#undef SPLHIP_FAST
          }
        } else {
          fast_chunk<CH, J, 0u>(k0, cx, cy, cmask, pre2, pre1, tab, cst, expfac4, tl, l0, l1, l2);
        }
        continue;
      }
    } else {
      // Additive: the particles' a_c come from LDS (cst: (a0, a1, a2, 0)) as
      // vector operands, and slots past the batch (stale or uninitialised
      // values) are read at their own offsets and masked by the select, so
      // no address or lane read is computed per particle.
      float at[J];
      float4 ca[J];
      uint32_t mv[J];
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int k = k0 + j;
        ca[j] = cst[k];
        mv[j] = cmask[cx][k];
        at[j] = pre1[cy][k] * pre2[cx][k];
      }
#pragma unroll
      for (int j = 0; j < J; ++j) {
        if (C0) l0 = l0 + in_footprint(ca[j].x * at[j], mv[j], cy);
        if (C1) l1 = l1 + in_footprint(ca[j].y * at[j], mv[j], cy);
        if (C2) l2 = l2 + in_footprint(ca[j].z * at[j], mv[j], cy);
      }
      continue;
    }
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

// Phase A of a batch, lane = particle (splotch_host.cc:462-476 + 596-607, tile
// frame): the lane's particle (record R) tabulates, for the block's 8
// columns, its row range (BITS: bit r set when block row byo + r lies in the
// column's [lo, hi); else lo | hi << 16) and Gaussian column factor, and for
// the 8 rows its Gaussian row factor, into column `lane` of the wave's LDS
// slices.  att_bound: the largest row factor times the largest column factor
// (+inf if one is NaN); float multiplication being monotone, every pixel's
// att = pre1 * pre2 is at most it, and vmax * att at most vmax * att_bound.
template <bool BITS>
__device__ __forceinline__ void tabulate(const SplatArgs &a, const Rec &R, int lane, int x0s,
                                        int y0s, int w, int h, int bxo, int byo,
                                        const float *s_tab, uint32_t (*col)[kPad],
                                        float (*pre2)[kPad], float (*pre1)[kPad],
                                        float expfac, float &att_bound) {
  const float rfacr = R.r * a.rfac;
  const float posx = R.x - static_cast<float>(x0s), posy = R.y - static_cast<float>(y0s);
  const int bx0 = max(cvt_i32(posx - rfacr + 1), 0), bx1 = min(cvt_i32(posx + rfacr + 1), w);
  const int by0 = max(cvt_i32(posy - rfacr + 1), 0), by1 = min(cvt_i32(posy + rfacr + 1), h);
  const float radsq = rfacr * rfacr;
  const float sigma = a.h2sigma * R.r;
  const float stp = -1.f / (sigma * sigma);
  // branch-free: every column and row is computed (outside the box the
  // values are discarded, e.g. the square root of a negative number), the
  // table reads of all 16 lookups are issued before any of them is used,
  // and the LDS stores follow
  uint32_t lohi[8];
  bool cin[8], rin[8];
  float pc[8], pr[8], t1[16], t2[16];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int x = bxo + c;
    cin[c] = x >= bx0 && x < bx1;
    const float dxsq = (static_cast<float>(x) - posx) * (static_cast<float>(x) - posx);
    const float dy = sqrtf(radsq - dxsq);
    const int lo = max(by0, cvt_i32(posy - dy + 1));
    const int hi = min(by1, cvt_i32(posy + dy + 1));
    uint32_t lh;
    if (BITS) {
      const int r0 = min(max(lo - byo, 0), 8), r1 = min(max(hi - byo, 0), 8);
      lh = (r1 > r0) ? (((1u << r1) - 1u) & ~((1u << r0) - 1u)) : 0u;
    } else {
      lh = static_cast<uint32_t>(lo) | (static_cast<uint32_t>(hi) << 16);
    }
    lohi[c] = (cin[c] && hi > lo) ? lh : 0u;
    pc[c] = (stp * dxsq) * expfac;
    xop_reads(s_tab, pc[c], t1[c], t2[c]);
  }
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int y = byo + r;
    rin[r] = y >= by0 && y < by1;
    const float dyf = static_cast<float>(y) - posy;
    pr[r] = (stp * dyf * dyf) * expfac;
    xop_reads(s_tab, pr[r], t1[8 + r], t2[8 + r]);
  }
  // every read is used here, so that none is moved into a branch of the
  // selects below
#pragma unroll
  for (int k = 0; k < 16; ++k) asm volatile("" : "+v"(t1[k]), "+v"(t2[k]));
  float b2 = 0.f, b1 = 0.f;
  bool nan = false;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const float p2 = cin[c] ? xop_value(pc[c], t1[c], t2[c]) : 0.f;
    b2 = fmaxf(b2, p2);
    nan |= p2 != p2;
    col[c][lane] = lohi[c];
    pre2[c][lane] = p2;
  }
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const float p1 = rin[r] ? xop_value(pr[r], t1[8 + r], t2[8 + r]) : 0.f;
    b1 = fmaxf(b1, p1);
    nan |= p1 != p1;
    pre1[r][lane] = p1;
  }
  att_bound = nan ? __builtin_inff() : b1 * b2;
}

// Work items of a launch: items[0, n), or with device counts (dc: the pass's
// item total and long-list total) the items after the first min(dc[1], cap)
// (those go to the producer/consumer kernel).
__device__ __forceinline__ uint32_t splat_items(const uint32_t *&items, uint32_t n,
                                                const ItemCounts &dc) {
  if (!dc.items) return n;
  const uint32_t tot = *dc.items, nh = min(min(*dc.heavy, dc.heavy_cap), tot);
  items += nh;
  return tot - nh;
}

// PPW (variant 4): particles per chunk of the walk (1, 2, 4 or 8; 0: 8 for
// one channel, 8 additive / 4 ordered for all three).
template <int ADD, int VAR, int PPW>
__global__ __launch_bounds__(256) void k_splat8(SplatArgs a, const uint32_t *items,
                                                uint32_t nitems_, ItemCounts dc,
                                                uint32_t nsplit_) {
  constexpr int J1 = PPW ? PPW : 8, J3 = PPW ? PPW : (ADD ? 8 : 4);
  const uint32_t nreal = splat_items(items, nitems_, dc);
  // the first nsplit items are long lists served as three one-channel items
  const uint32_t nsplit = min(dc.nsplit ? *dc.nsplit : nsplit_, nreal);
  const uint32_t nitems = nreal + 2u * nsplit;
  __shared__ float s_tab[2048];
  __shared__ uint32_t s_col[4][8][kPad];
  __shared__ float s_pre2[4][8][kPad];
  __shared__ float s_pre1[4][8][kPad];
  // per-particle constants: ordered (a0, a1, a2, max |a|), (q0, q1, q2, 0);
  // additive (a0, a1, a2, 0)
  __shared__ float4 s_cst[4][64 * (ADD ? 1 : 2)];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // a workgroup past the items (a grid sized for a capacity) leaves at once
  if (blockIdx.x * 4u >= nitems) return;
  for (int i = tid; i < 2048; i += 256) s_tab[i] = a.exptab[i];
  __syncthreads();
  const uint32_t it = blockIdx.x * 4u + wave;
  if (it >= nitems) return;
  const uint64_t t_start = a.trace ? __builtin_amdgcn_s_memrealtime() : 0;
  // wave-uniform values are moved to scalar registers so that loop bounds and
  // channel selection compile to scalar branches
  const uint32_t item = it < 3u * nsplit ? (uniform(items[it / 3u]) & ~3u) | (it % 3u)
                                          : uniform(items[it - 2u * nsplit]);
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
  float4 *wcst = s_cst[wave];
  if (!ADD) {
    // the fast chunks read every slot of a chunk, also past a short batch
    // (masked there): start from finite values
    for (int i = lane; i < 8 * kPad; i += 64) {
      (&col[0][0])[i] = 0u;
      (&pre2[0][0])[i] = 0.f;
      (&pre1[0][0])[i] = 0.f;
    }
    wcst[2 * lane] = make_float4(0.f, 0.f, 0.f, 0.f);
    wcst[2 * lane + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  const float expfac = a.expfac, tl = a.taylorlimit;

  float l0 = 0.f, l1 = 0.f, l2 = 0.f;
  const bool pix_in = lx < min(bxo + 8, w) && ly < min(byo + 8, h);
  if (ug.accumulate == 2 && pix_in) {  // a later chunk: continue the chain
    const float *d = ug.pic + (static_cast<int64_t>(x0s + lx) * ug.yres + (y0s + ly)) * 3;
    l0 = d[0];
    l1 = d[1];
    l2 = d[2];
  }
  const uint32_t e0 = uniform(a.bin_off[u]), e1 = uniform(a.bin_off[u + 1]);
  // long lists are near the critical path: let their waves issue ahead of the
  // rest (priority 2; the producer/consumer workgroups, the critical path, run at 3)
  if (e1 - e0 >= a.prio_len) __builtin_amdgcn_s_setprio(2);
  // The records of a batch are gathered one batch ahead and their slots two
  // ahead, so that the walk of a batch covers the latency of the next gather.
  Rec rn{};
  uint32_t sn = 0;
  if (e0 + lane < e1) rn = a.recs[a.ent[e0 + lane]];
  if (e0 + 64u + lane < e1) sn = a.ent[e0 + 64u + lane];
  for (uint32_t base = e0; base < e1; base += 64) {
    const int nb = static_cast<int>(uniform(min<uint32_t>(64u, e1 - base)));
    const Rec R = rn;
    if (base + 64u + lane < e1) rn = a.recs[sn];
    if (base + 128u + lane < e1) sn = a.ent[base + 128u + lane];
    // Phase A: lane = particle (splotch_host.cc:462-476 + 596-607, tile frame)
    float va0 = 0.f, va1 = 0.f, va2 = 0.f, vq0 = 0.f, vq1 = 0.f, vq2 = 0.f, vmax = 0.f;
    bool vfast = false, vtay = false;
    if (lane < nb) {
      float ab;
      tabulate<VAR >= 3>(a, R, lane, x0s, y0s, w, h, bxo, byo, s_tab, col, pre2, pre1, expfac,
                         ab);
      va0 = -R.e0;
      va1 = -R.e1;
      va2 = -R.e2;
      if (!ADD) {
        const float g = ug.grayabsorb;
        vq0 = R.e0 / (R.e0 + g);
        vq1 = R.e1 / (R.e1 + g);
        vq2 = R.e2 / (R.e2 + g);
        vmax = fmaxf(fabsf(va0), fmaxf(fabsf(va1), fabsf(va2)));
        vfast = va0 >= kFastA && va0 <= 0.f && va1 >= kFastA && va1 <= 0.f && va2 >= kFastA &&
                va2 <= 0.f && finite3(vq0, vq1, vq2);
        // Taylor particle: vmax * att < tl at every pixel of the block
        // (splotch_host.cc:648), shown by the bound on att
        vtay = vfast && (vmax * ab) < tl;
        // only the fast chunks read these, also for masked slots past the
        // batch, so a particle that is not fast (possibly a non-finite a or
        // q) leaves zeros, which a masked slot applies as l + 0
        const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
        wcst[2 * lane] = vfast ? make_float4(va0, va1, va2, vmax) : z4;
        wcst[2 * lane + 1] = vfast ? make_float4(vq0, vq1, vq2, 0.f) : z4;
      } else {
        wcst[lane] = make_float4(va0, va1, va2, 0.f);
      }
    } else {
      // a slot past a short batch: no pixel in its footprint
#pragma unroll
      for (int c = 0; c < 8; ++c) col[c][lane] = 0u;
    }
    wave_sync();
    const uint64_t fastmask = __ballot(vfast), taymask = __ballot(vtay);
    // Phase B: lane = pixel, particles in list order.
    if (VAR == 4) {
      switch (chan) {
        case 0: walk_chunks<ADD, 0, J1>(nb, cx, cy, col, pre2, pre1, s_tab, va0, va1, va2, vq0, vq1, vq2, vmax, wcst, fastmask, taymask, expfac, tl, l0, l1, l2); break;
        case 1: walk_chunks<ADD, 1, J1>(nb, cx, cy, col, pre2, pre1, s_tab, va0, va1, va2, vq0, vq1, vq2, vmax, wcst, fastmask, taymask, expfac, tl, l0, l1, l2); break;
        case 2: walk_chunks<ADD, 2, J1>(nb, cx, cy, col, pre2, pre1, s_tab, va0, va1, va2, vq0, vq1, vq2, vmax, wcst, fastmask, taymask, expfac, tl, l0, l1, l2); break;
        default: walk_chunks<ADD, 3, J3>(nb, cx, cy, col, pre2, pre1, s_tab, va0, va1, va2, vq0, vq1, vq2, vmax, wcst, fastmask, taymask, expfac, tl, l0, l1, l2); break;
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
  if (pix_in) {
    float *d = ug.pic + (static_cast<int64_t>(x0s + lx) * ug.yres + (y0s + ly)) * 3;
    const bool add = ug.accumulate == 1;
    if (chan == 3 || chan == 0) d[0] = add ? d[0] + l0 : l0;
    if (chan == 3 || chan == 1) d[1] = add ? d[1] + l1 : l1;
    if (chan == 3 || chan == 2) d[2] = add ? d[2] + l2 : l2;
  }
}

// ---------------------------------------------------------------------------
// Long lists: producer / consumer workgroups, one per 8x8 block, all three
// colour channels.
//
// A block whose list holds thousands to hundreds of thousands of particles is
// near the critical path of the frame: its pixels' update chains are serial.
// Such a block gets a 768-thread workgroup.  Wave 0 is the consumer: lane =
// pixel, and per particle it only applies, per channel c,
//     l_c += ((l_c - q_c) * m1_c) * m2     (ordered)     l_c += t_c   (additive)
// with operands that reproduce render_new's update bit for bit
// (splotch_host.cc:615, 648-658):
//   Taylor branch   m1 = a,  m2 = att           -> ((l - q) * a) * att
//   expm1 branch    m1 = expm1(att * a), m2 = 1 -> (l - q) * em      (x*1 == x)
//   additive        t = a * att
//   outside         m1 = m2 = t = 0             -> l + (+-0) == l
// (l starts at +0 and a float sum is -0 only if both terms are -0, so l is
// never -0 and adding a zero leaves it unchanged, as the reference's skip.
// For finite records l - q stays finite, so (l - q) * 0 is a zero.)
// Waves 1..11 run a pipeline ahead of the consumer: particle set-up three
// steps ahead (tile-frame position, box, Gaussian width, emission /
// absorption; splotch_host.cc:462-476, 591-632), the separable column / row
// tables two steps ahead (splotch_host.cc:596-607), and the operands one step
// ahead, with lane = pixel and each producer wave on its own four particles
// of the step: the producers do the table work of the lane-per-pixel kernel
// once for all three channels, the consumer four operations per channel.
// The consumer runs at a higher wave priority than the producers, which
// share its SIMD.
// Records reach the set-up through LDS in super-batches that wave 1 loads
// about five steps before their use, so the gather latency stays off the
// critical path.  Three waves per SIMD hide most of the LDS latency of the
// stages; a workgroup occupies its CU (LDS and registers).
constexpr int kPW = 11;                // producer waves (three waves per SIMD)
constexpr int kPB = 4;                 // particles per producer wave and step
constexpr int kB = kPW * kPB;          // particles per step
constexpr int kQ = 4;                  // operand stage: particles per group of LDS reads
static_assert(kPB % kQ == 0, "groups tile a producer's particles");
constexpr int kPThreads = 64 * (kPW + 1);
constexpr int kPProd = 64 * kPW;       // producer lanes
constexpr int kSB = 5;                 // steps per record super-batch
constexpr int kSR = kSB * kB;          // records per super-batch
constexpr int kSRL = (kSR + 63) / 64;  // records per wave-1 lane and super-batch
static_assert(kSB >= 5, "a super-batch is stored four steps before its first use");
constexpr int kSRP = kSRL * 64;        // staging row: whole wave stores (tail unused)
static_assert(kB <= 64, "fast-particle mask is 64 bits");

struct PcPart {
  float posx, posy, radsq, stp;
  int bx0, bx1, by0, by1;
};

__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int ADD>
__global__ __launch_bounds__(kPThreads) void k_splat_pc(SplatArgs a, const uint32_t *items,
                                                        uint32_t npc, ItemCounts dc) {
  // device counts: *dc.npc items from *dc.first on; the rest of the grid
  // (sized for a capacity) has no work (uniform per workgroup, before any
  // barrier)
  if (dc.npc && blockIdx.x >= *dc.npc) return;
  // One LDS block with the exp table at address 0, so that a table entry's
  // address is its byte offset (no base to add per lookup).
  struct Lds {
    float tab[2048];
    float4 op[2][kB][64];   // consumer operands (m1_0, m1_1, m1_2, m2) / (t0, t1, t2, 0)
    float4 ca[4][kB];       // (a0, a1, a2, max |a|)
    float4 cq[4][kB];       // (q0, q1, q2, 0)
    Rec rec[2][kSRP];       // staged records, two super-batches
    PcPart part[2][kB];
    uint64_t fast[4];       // bit k: particle k has every a_c in [kFastA, 0]
    uint32_t cm[2][kB][8];  // column row-masks
    float p2[2][kB][8];     // column factors
    float p1[2][kB][8];     // row factors
  };
  __shared__ Lds L;
  float(&s_tab)[2048] = L.tab;
  Rec(&s_rec)[2][kSRP] = L.rec;
  PcPart(&s_part)[2][kB] = L.part;
  float4(&s_ca)[4][kB] = L.ca;
  float4(&s_cq)[4][kB] = L.cq;
  uint64_t(&s_fast)[4] = L.fast;
  uint32_t(&s_cm)[2][kB][8] = L.cm;
  float(&s_p2)[2][kB][8] = L.p2;
  float(&s_p1)[2][kB][8] = L.p1;
  float4(&s_op)[2][kB][64] = L.op;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = static_cast<int>(uniform(static_cast<uint32_t>(tid >> 6)));
  const bool producer = wave >= 1;
  // table task rank: waves 4 and 8 share SIMD 0 with the consumer (waves go
  // to the SIMDs in turn) and take the last, cheaper row tasks
  const int trank = (wave & 3) == 0 ? 8 + (wave >> 2) : wave - 1 - (wave >> 2);
  const int ptid = trank * 64 + lane;
  for (int i = tid; i < 2048; i += kPThreads) s_tab[i] = a.exptab[i];
  const uint32_t it = blockIdx.x + (dc.first ? uniform(*dc.first) : 0u);
  const uint64_t t_start = a.trace ? __builtin_amdgcn_s_memrealtime() : 0;
  const uint32_t u = uniform(items[it]) >> 2;
  const UnitGeom ug = unit_geom_uniform(a, u);
  const int gy = ug.gy, gx = ug.gx, x0s = ug.x0s, y0s = ug.y0s, w = ug.w, h = ug.h;
  const int bxo = gx * 8, byo = gy * 8;
  const uint32_t e0 = uniform(a.bin_off[u]), e1 = uniform(a.bin_off[u + 1]);
  const int n = static_cast<int>(e1 - e0);
  const int nbat = (n + kB - 1) / kB;
  const float expfac = a.expfac, tl = a.taylorlimit;
  const float expfac4 = 4.f * expfac;  // exact
  const int cx = lane >> 3, cy = lane & 7;  // producers and consumer: lane = pixel
  if (producer) __builtin_amdgcn_s_setprio(2);
  else __builtin_amdgcn_s_setprio(3);

  auto count = [&](int b) { return min(kB, n - b * kB); };
  // particle set-up of batch b (wave 1, lane = particle).  Lanes past the
  // batch (and batches past the list) use clamped records, which the
  // operand stage masks.
  auto setup = [&](int b, const Rec &R) {
    const float rfacr = R.r * a.rfac;
    const float posx = R.x - static_cast<float>(x0s), posy = R.y - static_cast<float>(y0s);
    PcPart P;
    P.posx = posx;
    P.posy = posy;
    P.bx0 = max(cvt_i32(posx - rfacr + 1), 0);
    P.bx1 = min(cvt_i32(posx + rfacr + 1), w);
    P.by0 = max(cvt_i32(posy - rfacr + 1), 0);
    P.by1 = min(cvt_i32(posy + rfacr + 1), h);
    P.radsq = rfacr * rfacr;
    const float sigma = a.h2sigma * R.r;
    P.stp = -1.f / (sigma * sigma);
    const float a0 = -R.e0, a1 = -R.e1, a2 = -R.e2;
    float q0 = 0.f, q1 = 0.f, q2 = 0.f, mx = 0.f;
    bool fast = false;
    if (!ADD) {
      const float g = ug.grayabsorb;
      q0 = R.e0 / (R.e0 + g);
      q1 = R.e1 / (R.e1 + g);
      q2 = R.e2 / (R.e2 + g);
      mx = fmaxf(fabsf(a0), fmaxf(fabsf(a1), fabsf(a2)));
      fast = a0 >= kFastA && a0 <= 0.f && a1 >= kFastA && a1 <= 0.f && a2 >= kFastA &&
             a2 <= 0.f && finite3(q0, q1, q2);
    }
    const uint64_t fb = __ballot(fast);
    if (lane < kB) {
      s_part[b & 1][lane] = P;
      s_ca[b & 3][lane] = make_float4(a0, a1, a2, mx);
      s_cq[b & 3][lane] = make_float4(q0, q1, q2, 0.f);
    }
    if (lane == 0) s_fast[b & 3] = fb;  // bits past the batch are not read
  };
  // Column ranges / factors and row factors of batch b into slot b & 1: task
  // (particle k, j) is column j (j < 8) or row j - 8.  Unrolled over a lane's
  // tasks with every LDS read issued before the arithmetic that uses it.
  auto tables = [&](int b) {
    const int ts = b & 1, cnt = count(b);
    const PcPart(&part)[kB] = s_part[b & 1];
    constexpr int NT = (kB * 16 + kPProd - 1) / kPProd;
    PcPart P[NT];
    int kt[NT], jt[NT];
    bool vt[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      // tasks [0, 8 kB) are columns, [8 kB, 16 kB) rows, so a wave runs
      // one kind (the column path is the longer one)
      const int t = ptid + i * kPProd;
      const bool row = t >= kB * 8;
      const int tr = row ? t - kB * 8 : t;
      vt[i] = t < kB * 16 && (tr >> 3) < cnt;
      kt[i] = vt[i] ? tr >> 3 : 0;
      jt[i] = vt[i] ? (tr & 7) + (row ? 8 : 0) : 0;
      P[i] = part[kt[i]];
    }
    float arg[NT], dd[NT];
    bool on[NT];
    uint32_t mk[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      const int j = jt[i];
      mk[i] = 0;
      if (j < 8) {
        const int x = bxo + j;
        on[i] = x >= P[i].bx0 && x < P[i].bx1;
        const float dxsq = (static_cast<float>(x) - P[i].posx) * (static_cast<float>(x) - P[i].posx);
        const float dy = sqrtf(P[i].radsq - dxsq);
        const int lo = max(P[i].by0, cvt_i32(P[i].posy - dy + 1));
        const int hi = min(P[i].by1, cvt_i32(P[i].posy + dy + 1));
        const int r0 = min(max(lo - byo, 0), 8), r1 = min(max(hi - byo, 0), 8);
        if (on[i] && hi > lo && r1 > r0) mk[i] = ((1u << r1) - 1u) & ~((1u << r0) - 1u);
        arg[i] = P[i].stp * dxsq;
      } else {
        const int y = byo + (j - 8);
        on[i] = y >= P[i].by0 && y < P[i].by1;
        const float dyf = static_cast<float>(y) - P[i].posy;
        arg[i] = P[i].stp * dyf * dyf;
      }
      dd[i] = on[i] ? xop(s_tab, arg[i], expfac) : 0.f;
    }
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      if (!vt[i]) continue;
      const int k = kt[i], j = jt[i];
      if (j < 8) {
        s_cm[ts][k][j] = mk[i];
        s_p2[ts][k][j] = dd[i];
      } else {
        s_p1[ts][k][j - 8] = dd[i];
      }
    }
  };
  // Consumer operands of batch b: producer wave pw covers particles
  // [pw * kPB, pw * kPB + kPB) of the batch, lane = pixel, in groups of kQ
  // with every LDS read of a group issued first.  Particles past the batch
  // get zero operands.
  // FAST: every particle of the group has all a_c in [kFastA, 0], so the
  // exp-table index is trunc(p) with p in [0, 2^20).
  auto operands_group = [&](int b, int kb, auto fast_tag) {  // particles [kb, kb + kQ)
    constexpr bool FAST = decltype(fast_tag)::value;
    const int ts = b & 1, cs = b & 3, cnt = count(b);
    uint32_t mk[kQ];
    float p1[kQ], p2[kQ];
    float4 ca[kQ];
#pragma unroll
    for (int j = 0; j < kQ; ++j) {
      const int k = kb + j;
      mk[j] = s_cm[ts][k][cx];
      p1[j] = s_p1[ts][k][cy];
      p2[j] = s_p2[ts][k][cx];
      ca[j] = s_ca[cs][k];
    }
    float4 op[kQ];
    if (ADD) {
#pragma unroll
      for (int j = 0; j < kQ; ++j) {
        const float at = p1[j] * p2[j];
        const bool in = (kb + j < cnt) && ((mk[j] >> cy) & 1u);
        op[j] = in ? make_float4(ca[j].x * at, ca[j].y * at, ca[j].z * at, 0.f)
                   : make_float4(0.f, 0.f, 0.f, 0.f);
      }
#pragma unroll
      for (int j = 0; j < kQ; ++j) s_op[ts][kb + j][lane] = op[j];
      return;
    }
    float at[kQ];
    bool in[kQ], tay[kQ];
    uint32_t tb = 0;  // bit j: every pixel of particle j takes the Taylor branch or is outside
#pragma unroll
    for (int j = 0; j < kQ; ++j) {
      at[j] = p1[j] * p2[j];
      in[j] = (kb + j < cnt) && ((mk[j] >> cy) & 1u);
      tay[j] = (ca[j].w * at[j]) < tl;
      if (__ballot(tay[j] || !in[j]) == ~0ull) tb |= 1u << j;
    }
    // One straight-line body per pattern: the all-Taylor particles skip the
    // exp-table work, the others' table reads are all issued before their
    // arithmetic.  op = (a_c, att) for a Taylor pixel, (expm1(att a_c), 1)
    // otherwise, zero outside.
    auto body = [&](auto tb_tag) {
      constexpr uint32_t TB = decltype(tb_tag)::value;
      float g0[kQ], g1[kQ], g2[kQ], pg0[kQ], pg1[kQ], pg2[kQ];
      float u0[kQ], w0[kQ], u1[kQ], w1[kQ], u2[kQ], w2[kQ];
#pragma unroll
      for (int j = 0; j < kQ; ++j) {
        if ((TB >> j) & 1u) continue;
        g0[j] = at[j] * ca[j].x;
        g1[j] = at[j] * ca[j].y;
        g2[j] = at[j] * ca[j].z;
        if (FAST) {
          xtab_fast(s_tab, g0[j], expfac4, u0[j], w0[j]);
          xtab_fast(s_tab, g1[j], expfac4, u1[j], w1[j]);
          xtab_fast(s_tab, g2[j], expfac4, u2[j], w2[j]);
        } else {
          pg0[j] = g0[j] * expfac;
          pg1[j] = g1[j] * expfac;
          pg2[j] = g2[j] * expfac;
          const int i0 = xexp_index(pg0[j]), i1 = xexp_index(pg1[j]), i2 = xexp_index(pg2[j]);
          u0[j] = s_tab[i0 >> 10];
          w0[j] = s_tab[1024 + (i0 & 1023)];
          u1[j] = s_tab[i1 >> 10];
          w1[j] = s_tab[1024 + (i1 & 1023)];
          u2[j] = s_tab[i2 >> 10];
          w2[j] = s_tab[1024 + (i2 & 1023)];
        }
      }
#pragma unroll
      for (int j = 0; j < kQ; ++j) {
        if ((TB >> j) & 1u) {
          op[j] = in[j] ? make_float4(ca[j].x, ca[j].y, ca[j].z, at[j])
                        : make_float4(0.f, 0.f, 0.f, 0.f);
          continue;
        }
        float em0, em1, em2;
        if (FAST) {
          em0 = (fabsf(g0[j]) < tl) ? g0[j] : u0[j] * w0[j] - 1.f;
          em1 = (fabsf(g1[j]) < tl) ? g1[j] : u1[j] * w1[j] - 1.f;
          em2 = (fabsf(g2[j]) < tl) ? g2[j] : u2[j] * w2[j] - 1.f;
        } else {
          em0 = xexpm1_finish(g0[j], pg0[j], u0[j], w0[j], tl);
          em1 = xexpm1_finish(g1[j], pg1[j], u1[j], w1[j], tl);
          em2 = xexpm1_finish(g2[j], pg2[j], u2[j], w2[j], tl);
        }
        op[j] = in[j] ? make_float4(tay[j] ? ca[j].x : em0, tay[j] ? ca[j].y : em1,
                                    tay[j] ? ca[j].z : em2, tay[j] ? at[j] : 1.f)
                      : make_float4(0.f, 0.f, 0.f, 0.f);
      }
#pragma unroll
      for (int j = 0; j < kQ; ++j) s_op[ts][kb + j][lane] = op[j];
    };
    static_assert(kQ == 4, "one body per pattern of four particles");
    switch (uniform(tb)) {
#define SPLHIP_PCT(T) case T: body(std::integral_constant<uint32_t, T>{}); break;
      SPLHIP_PCT(0) SPLHIP_PCT(1) SPLHIP_PCT(2) SPLHIP_PCT(3)
      SPLHIP_PCT(4) SPLHIP_PCT(5) SPLHIP_PCT(6) SPLHIP_PCT(7)
      SPLHIP_PCT(8) SPLHIP_PCT(9) SPLHIP_PCT(10) SPLHIP_PCT(11)
      SPLHIP_PCT(12) SPLHIP_PCT(13) SPLHIP_PCT(14) SPLHIP_PCT(15)
#undef SPLHIP_PCT
    }
  };
  auto operands = [&](int b, int pw) {
    const uint64_t f64 = s_fast[b & 3];
    const uint64_t fb = (static_cast<uint64_t>(uniform(static_cast<uint32_t>(f64 >> 32))) << 32) |
                        uniform(static_cast<uint32_t>(f64));
#pragma unroll
    for (int g = 0; g < kPB; g += kQ) {
      const int kb = pw * kPB + g;
      const uint64_t want = ((uint64_t(1) << kQ) - 1u) << kb;
      if (!ADD && (fb & want) == want)
        operands_group(b, kb, std::true_type{});
      else
        operands_group(b, kb, std::false_type{});
    }
  };

  // Records through LDS: wave 1 holds the next super-batch's records and the
  // one after that's list indices in registers (positions clamped to the list
  // end), and stores a super-batch into its half of s_rec four steps before
  // its first set-up.
  auto entry = [&](int J, int r) { return a.ent[e0 + min(J * kSR + r * 64 + lane, n - 1)]; };
  // records as two 4-float halves in native vector registers (the arrays
  // must stay in registers: struct copies would keep them in memory)
  typedef float v4f __attribute__((ext_vector_type(4)));
  v4f RA[kSRL], RB[kSRL];
  uint32_t I[kSRL];
  auto load_recs = [&]() {
#pragma unroll
    for (int r = 0; r < kSRL; ++r) {
      const v4f *src = reinterpret_cast<const v4f *>(a.recs + I[r]);
      RA[r] = src[0];
      RB[r] = src[1];
    }
  };
  auto store_recs = [&](int slot) {
#pragma unroll
    for (int r = 0; r < kSRL; ++r) {
      v4f *dst = reinterpret_cast<v4f *>(&s_rec[slot][r * 64 + lane]);
      dst[0] = RA[r];
      dst[1] = RB[r];
    }
  };
  if (wave == 1) {
#pragma unroll
    for (int r = 0; r < kSRL; ++r) I[r] = entry(0, r);
    load_recs();
    store_recs(0);
#pragma unroll
    for (int r = 0; r < kSRL; ++r) I[r] = entry(1, r);
    load_recs();
#pragma unroll
    for (int r = 0; r < kSRL; ++r) I[r] = entry(2, r);
  }

  float l0 = 0.f, l1 = 0.f, l2 = 0.f;  // consumer: pixel lane
  const int hlx = bxo + cx, hly = byo + cy;
  const bool hpix = hlx < min(bxo + 8, w) && hly < min(byo + 8, h);
  float *hdst = ug.pic + (static_cast<int64_t>(x0s + hlx) * ug.yres + (y0s + hly)) * 3;
  if (!producer && ug.accumulate == 2 && hpix) {  // a later chunk: continue the chains
    l0 = hdst[0];
    l1 = hdst[1];
    l2 = hdst[2];
  }
  __syncthreads();  // s_tab, s_rec[0]
  for (int s = -3; s < nbat; ++s) {
    if (producer) {
      if (wave == 1) {
        const int b = s + 3;
        setup(b, s_rec[(b / kSB) & 1][(b % kSB) * kB + min(lane, kB - 1)]);
        if ((s + 4) % kSB == 0 && s + 4 > 0) {  // super-batch J starts next step
          const int J = (s + 4) / kSB;
          store_recs(J & 1);
          load_recs();
#pragma unroll
          for (int r = 0; r < kSRL; ++r) I[r] = entry(J + 2, r);
        }
      }
      if (s + 2 >= 0 && s + 2 < nbat) tables(s + 2);
      if (s + 1 >= 0 && s + 1 < nbat) operands(s + 1, wave - 1);
    } else if (s >= 0) {
      const float4(*op)[64] = s_op[s & 1];
      const float4 *cq = s_cq[s & 3];
      const int cnt = count(s);
      constexpr int C = 4;  // particles per group of reads
      static_assert(kB % C == 0, "groups tile the batch");
#pragma unroll
      for (int k0 = 0; k0 < kB; k0 += C) {
        if (k0 >= cnt) break;
        float4 m[C], q[C];
#pragma unroll
        for (int j = 0; j < C; ++j) {
          m[j] = op[k0 + j][lane];
          if (!ADD) q[j] = cq[k0 + j];
        }
#pragma unroll
        for (int j = 0; j < C; ++j) {
          if (ADD) {
            l0 = l0 + m[j].x;
            l1 = l1 + m[j].y;
            l2 = l2 + m[j].z;
          } else {
            l0 = l0 + ((l0 - q[j].x) * m[j].x) * m[j].w;
            l1 = l1 + ((l1 - q[j].y) * m[j].y) * m[j].w;
            l2 = l2 + ((l2 - q[j].z) * m[j].z) * m[j].w;
          }
        }
      }
    }
    lds_barrier();
  }
  if (a.trace && tid == 0) {
    uint64_t *tr = a.trace + 4 * static_cast<uint64_t>(it);
    tr[0] = t_start;
    tr[1] = __builtin_amdgcn_s_memrealtime();
    tr[2] = static_cast<uint64_t>(n) | (uint64_t(3) << 32);
    tr[3] = static_cast<uint64_t>(__smid());
  }
  if (!producer && hpix) {
    const bool add = ug.accumulate == 1;
    hdst[0] = add ? hdst[0] + l0 : l0;
    hdst[1] = add ? hdst[1] + l1 : l1;
    hdst[2] = add ? hdst[2] + l2 : l2;
  }
}

// ---------------------------------------------------------------------------
// Staged long lists.  A long list's pixel chains are serial, and in a
// lane-per-pixel wave (or a producer/consumer workgroup) the per-particle
// table work, not the chain, sets their pace.  The staged path moves that
// work off the chain: k_stage_ops (every CU) writes each staged list's
// per-pixel operands to HBM, and k_stage_chain streams them back in list
// order, one wave per list, channel and 16-pixel slice, which applies only
// the chain's own operations (splotch_host.cc:615, 648-658):
//   additive  l_c = l_c + t_c                 t_c = a_c * att
//   ordered   l_c = l_c + ((l_c - q_c) * m1_c) * m2
//             Taylor branch m1 = a, m2 = att; expm1 branch m1 = expm1(att a),
//             m2 = 1 (x * 1 == x)
// A pixel outside a particle's footprint gets t = +0, or m1 = m2 = 0 (l + +-0
// == l: l starts at +0 and a float sum is -0 only if both terms are -0; for
// finite records l - q stays finite, so (l - q) * 0 is a zero, and ordered
// lists are staged only when every q_c of the pass is finite), so every
// pixel sees the reference's sums and compositing, rounded as it rounds them.
//
// Operand layout: the lists' entries in groups of 4, a list's groups padded
// with zero operands to a multiple of kStageLoad (the groups of one load of
// the chain kernel); list i owns groups [cbase[i], cbase[i+1]).  Group g of
// plane c is 64 float4 (lane = pixel cx * 8 + cy, components = the group's 4
// entries) at ops + (g * P + c) * 64, with P = 3 planes t_c (additive) or 4
// (m1_0, m1_1, m1_2, m2; ordered).  Ordered lists also have q_c per entry:
// qv[(g / kStageQ) * 48 + c * 16 + e] for entry e of the kStageQ groups
// holding g.  A list's groups are padded to a multiple of kStageLoad.
__host__ __device__ constexpr uint32_t stage_groups(uint32_t len) {
  return (len + 4 * kStageLoad - 1) / (4 * kStageLoad) * kStageLoad;
}
template <int ADD>
constexpr uint32_t stage_planes() {
  return ADD ? 3u : 4u;
}

// List offsets of the staged items (one workgroup): cbase (groups of 4
// entries) and cbat (batches of 64 entries), exclusive, totals at [n].
__global__ __launch_bounds__(256) void k_stage_base(SplatArgs a, const uint32_t *items,
                                                    uint32_t n, uint32_t *cbase, uint32_t *cbat,
                                                    const uint32_t *dn) {
  if (dn) n = min(n, *dn);
  __shared__ uint32_t s_g[256], s_b[256];
  __shared__ uint32_t s_carry[2];
  const int tid = threadIdx.x;
  if (tid == 0) s_carry[0] = s_carry[1] = 0;
  __syncthreads();
  for (uint32_t i0 = 0; i0 < n; i0 += 256) {
    const uint32_t i = i0 + tid;
    uint32_t len = 0;
    if (i < n) {
      const uint32_t u = items[i] >> 2;
      len = a.bin_off[u + 1] - a.bin_off[u];
    }
    const uint32_t g = stage_groups(len), b = (len + 63) / 64;
    s_g[tid] = g;
    s_b[tid] = b;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {  // inclusive scan (Hillis-Steele)
      const uint32_t vg = tid >= d ? s_g[tid - d] : 0u, vb = tid >= d ? s_b[tid - d] : 0u;
      __syncthreads();
      s_g[tid] += vg;
      s_b[tid] += vb;
      __syncthreads();
    }
    if (i < n) {
      cbase[i] = s_carry[0] + s_g[tid] - g;
      cbat[i] = s_carry[1] + s_b[tid] - b;
    }
    __syncthreads();
    if (tid == 0) {
      s_carry[0] += s_g[255];
      s_carry[1] += s_b[255];
    }
    __syncthreads();
  }
  if (tid == 0) {
    cbase[n] = s_carry[0];
    cbat[n] = s_carry[1];
  }
}

// Operands of staged lists [i0, i1): their batches of 64 entries (cbat[i0]
// to cbat[i1]) split evenly over the launch's waves, each wave walking a
// contiguous run (phase A as the 8x8 splat's, then lane = pixel for the
// batch's groups).
template <int ADD>
__global__ __launch_bounds__(256) void k_stage_ops(SplatArgs a, const uint32_t *items,
                                                   uint32_t i0, uint32_t i1,
                                                   const uint32_t *cbase, const uint32_t *cbat,
                                                   float4 *ops, float *qv, const uint32_t *dn) {
  constexpr uint32_t P = stage_planes<ADD>();
  if (dn) i1 = min(i1, *dn);
  if (i1 <= i0) return;  // the whole grid (device count)
  // the chains wait for these operands: issue ahead of the 8x8 splat's waves
  __builtin_amdgcn_s_setprio(3);
  __shared__ float s_tab[2048];
  __shared__ uint32_t s_col[4][8][kPad];
  __shared__ float s_pre2[4][8][kPad];
  __shared__ float s_pre1[4][8][kPad];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int i = tid; i < 2048; i += 256) s_tab[i] = a.exptab[i];
  __syncthreads();
  const uint32_t first = uniform(cbat[i0]), total = uniform(cbat[i1]) - first;
  const uint32_t nw = gridDim.x * 4u, w = blockIdx.x * 4u + static_cast<uint32_t>(wave);
  const uint32_t span = (total + nw - 1) / nw;
  uint32_t gb = first + w * span;
  const uint32_t gend = first + min(total, w * span + span);
  if (gb >= gend) return;
  // the staged item holding batch gb: cbat[lo] <= gb < cbat[lo + 1]
  uint32_t lo = i0, hi = i1;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (uniform(cbat[mid]) <= gb) lo = mid;
    else hi = mid;
  }
  uint32_t (*col)[kPad] = s_col[wave];
  float(*pre2)[kPad] = s_pre2[wave];
  float(*pre1)[kPad] = s_pre1[wave];
  const float expfac = a.expfac, tl = a.taylorlimit;
  const int cx = lane >> 3, cy = lane & 7;
  uint32_t i = lo, b = gb - uniform(cbat[lo]);
  for (; gb < gend; ++gb, ++b) {
    while (gb >= uniform(cbat[i + 1])) {  // next list (uniform)
      ++i;
      b = 0;
    }
    const uint32_t u = uniform(items[i]) >> 2;
    const UnitGeom ug = unit_geom_uniform(a, u);
    const uint32_t e0 = uniform(a.bin_off[u]), e1 = uniform(a.bin_off[u + 1]);
    const uint32_t base = e0 + 64u * b;
    const int nb = static_cast<int>(uniform(min<uint32_t>(64u, e1 - base)));
    const uint32_t g0 = uniform(cbase[i]) + 16u * b;  // the batch's first group
    const int ng = static_cast<int>(stage_groups(nb));  // 64 entries: whole loads
    float va0 = 0.f, va1 = 0.f, va2 = 0.f, vmx = 0.f;
    float vq[3] = {0.f, 0.f, 0.f};
    if (lane < nb) {
      const Rec R = a.recs[a.ent[base + lane]];
      float ab;
      tabulate<true>(a, R, lane, ug.x0s, ug.y0s, ug.w, ug.h, ug.gx * 8, ug.gy * 8, s_tab, col, pre2,
                     pre1, expfac, ab);
      va0 = -R.e0;
      va1 = -R.e1;
      va2 = -R.e2;
      if (!ADD) {
        const float gray = ug.grayabsorb;
        vq[0] = R.e0 / (R.e0 + gray);
        vq[1] = R.e1 / (R.e1 + gray);
        vq[2] = R.e2 / (R.e2 + gray);
        vmx = fmaxf(fabsf(va0), fmaxf(fabsf(va1), fabsf(va2)));
      }
    }
    if (!ADD && lane < ng * 4) {  // q_c per entry (0 past the list), lane = entry
      float *qd = qv + (g0 / kStageQ + static_cast<uint32_t>(lane) / 16u) * 48u + lane % 16;
      qd[0] = vq[0];
      qd[16] = vq[1];
      qd[32] = vq[2];
    }
    wave_sync();
    float4 *dst = ops + static_cast<size_t>(g0) * P * 64u + lane;
    for (int q = 0; q < ng; ++q) {
      float t0[4], t1[4], t2[4], t3[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k0 = 4 * q + j, k = min(k0, nb - 1);  // uniform; slots past the list are 0
        const bool in = k0 < nb && ((col[cx][k] >> cy) & 1u);
        const float at = pre1[cy][k] * pre2[cx][k];
        const float a0 = lane_f(va0, k), a1 = lane_f(va1, k), a2 = lane_f(va2, k);
        if (ADD) {
          const float x0 = a0 * at, x1 = a1 * at, x2 = a2 * at;
          t0[j] = in ? x0 : 0.f;
          t1[j] = in ? x1 : 0.f;
          t2[j] = in ? x2 : 0.f;
        } else {
          const bool tay = (lane_f(vmx, k) * at) < tl;
          const float em0 = xexpm1_bf(s_tab, at * a0, expfac, tl);
          const float em1 = xexpm1_bf(s_tab, at * a1, expfac, tl);
          const float em2 = xexpm1_bf(s_tab, at * a2, expfac, tl);
          t0[j] = in ? (tay ? a0 : em0) : 0.f;
          t1[j] = in ? (tay ? a1 : em1) : 0.f;
          t2[j] = in ? (tay ? a2 : em2) : 0.f;
          t3[j] = in ? (tay ? at : 1.f) : 0.f;
        }
      }
      float4 *g = dst + static_cast<size_t>(q) * P * 64u;
      g[0] = make_float4(t0[0], t0[1], t0[2], t0[3]);
      g[64] = make_float4(t1[0], t1[1], t1[2], t1[3]);
      g[128] = make_float4(t2[0], t2[1], t2[2], t2[3]);
      if (!ADD) g[192] = make_float4(t3[0], t3[1], t3[2], t3[3]);
    }
    wave_sync();
  }
}

// The chains of the staged lists.  A wave applies channel c of list i to a
// slice of kPix pixels of the block: workgroup (3 i + c) kS + h takes pixels
// [h kPix, h kPix + kPix).  One load instruction brings kS consecutive groups
// of a plane (lane part q: group g + q of the slice's pixels, 16 bytes per
// lane), so the loads in flight per wave cover kS groups each; lanes
// [0, kPix) run the chains and read the other parts' groups back through
// LDS one step ahead.  A chain's wave is bound by its bytes in flight and
// its serial operations, not by issue, so narrower slices (more waves per
// list, fewer bytes per entry and wave) run a list faster: 8 pixels for
// additive chains, 16 for ordered ones (whose three operand planes would not
// fit the registers at 8).
template <int ADD> constexpr int chain_groups() { return ADD ? kStageLoad : kStageQ; }
static_assert(kStageLoad % kStageQ == 0, "chain loads tile the padded groups");
template <int ADD>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_stage_chain(
    SplatArgs a, const uint32_t *items, uint32_t i0, const uint32_t *cbase,
    const float4 *__restrict__ ops, const float *__restrict__ qv, const uint32_t *dn) {
  constexpr int kS = chain_groups<ADD>(), kPix = 64 / kS;
  constexpr uint32_t P = stage_planes<ADD>();
  constexpr int kR = ADD ? 32 : 10;  // steps in flight (ordered: three loads per step)
  __shared__ float4 xs[3][64];
  const uint32_t i = i0 + blockIdx.x / (3u * kS);
  if (dn && i >= *dn) return;  // past the device count (uniform per workgroup)
  const uint32_t r3 = blockIdx.x % (3u * kS);
  const int c = static_cast<int>(r3 / kS), h = static_cast<int>(r3 % kS);
  const int lane = threadIdx.x, qo = lane / kPix, pl = lane % kPix;
  const int px = h * kPix + pl;  // pixel cx * 8 + cy
  const uint64_t t_start = a.trace ? __builtin_amdgcn_s_memrealtime() : 0;
  __builtin_amdgcn_s_setprio(3);
  const uint32_t u = uniform(items[i]) >> 2;
  const UnitGeom ug = unit_geom_uniform(a, u);
  const uint32_t len = uniform(a.bin_off[u + 1]) - uniform(a.bin_off[u]);
  const uint32_t G = stage_groups(len);  // the list is not empty
  const uint32_t gb0 = uniform(cbase[i]);
  const int cx = px >> 3, cy = px & 7;
  const int lx = ug.gx * 8 + cx, ly = ug.gy * 8 + cy;
  const bool pix_in = lane < kPix && lx < min(ug.gx * 8 + 8, ug.w) &&
                      ly < min(ug.gy * 8 + 8, ug.h);
  float *d = ug.pic + (static_cast<int64_t>(ug.x0s + lx) * ug.yres + (ug.y0s + ly)) * 3 + c;
  float l = 0.f;
  if (ug.accumulate == 2 && pix_in) l = *d;  // a later chunk: continue the chain
  // l is in its register before the operand loads are issued: a load still
  // in flight at the loop would make the compiler wait for every load there
  asm volatile("" : "+v"(l));
  // group g of plane c and this pixel: src[g * P * 64]; lane quarter qo reads g + qo
  const float4 *src = ops + static_cast<size_t>(gb0) * P * 64u + static_cast<size_t>(c) * 64u + px;
  const float4 *srcw = ops + static_cast<size_t>(gb0) * P * 64u + 3u * 64u + px;  // m2 (ordered)
  // q_c of the entries of group g + qo (4 floats; the same for a part's lanes;
  // ordered chains load kS = kStageQ groups, one block of q_c)
  const float4 *qsrc = reinterpret_cast<const float4 *>(qv + static_cast<size_t>(gb0 / kStageQ) * 48u +
                                                       static_cast<size_t>(c) * 16u) + qo;
  auto at_load = [&](uint32_t g0) {  // groups [g0, g0 + kS), the last ones again past the list
    return static_cast<size_t>(min<uint32_t>(g0, G - kS) + qo) * P * 64u;
  };
  auto at_q = [&](uint32_t g0) {  // that load's q_c (48 floats per kStageQ groups)
    return static_cast<size_t>(min<uint32_t>(g0, G - kS) / kStageQ) * 12u;
  };
  // one load's 4 kS entries for lanes [0, kPix): its own part, then the
  // others read back through LDS
  struct Q {
    float4 v[kS];
  };
  auto exchange = [&](int slot, const float4 &v) {
    xs[slot][lane] = v;
    wave_sync();
    Q r;
    r.v[0] = v;
#pragma unroll
    for (int k = 1; k < kS; ++k) r.v[k] = xs[slot][pl + k * kPix];
    return r;
  };
  struct Step {
    Q m, w, q;  // ordered: m1_c, m2 and q_c (the same in every lane) of the load's entries
  };
  auto step_of = [&](const float4 &vm, const float4 &vw, const float4 &vq) {
    Step st;
    st.m = exchange(0, vm);
    if (!ADD) {
      st.w = exchange(1, vw);
      st.q = exchange(2, vq);
    }
    return st;
  };
  auto apply = [&](const Step &st) {
    if (ADD) {
#pragma unroll
      for (int k = 0; k < kS; ++k) {  // load part k: entries 4k .. 4k + 3
        l = l + st.m.v[k].x;
        l = l + st.m.v[k].y;
        l = l + st.m.v[k].z;
        l = l + st.m.v[k].w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < kS; ++k) {
        const float4 &m4 = st.m.v[k], &w4 = st.w.v[k], &q4 = st.q.v[k];
        l = l + ((l - q4.x) * m4.x) * w4.x;
        l = l + ((l - q4.y) * m4.y) * w4.y;
        l = l + ((l - q4.z) * m4.z) * w4.z;
        l = l + ((l - q4.w) * m4.w) * w4.w;
      }
    }
  };
  float4 bm[kR], bw[ADD ? 1 : kR], bq[ADD ? 1 : kR];
#pragma unroll
  for (int r = 0; r < kR; ++r) {  // issued in the order they are consumed
    const size_t o = at_load(r * kS);
    bm[r] = src[o];
    if (!ADD) {
      bw[r] = srcw[o];
      bq[r] = qsrc[at_q(r * kS)];
    }
    asm volatile("" ::: "memory");
  }
  // Software pipeline over the loads: step r of a round applies the entries
  // of load r (exchanged a step earlier), exchanges load r + 1 meanwhile, and
  // refills that load's registers with the groups kR loads further on.
  Step cur = step_of(bm[0], bw[0], bq[0]);
  asm volatile("" ::: "memory");
  {
    const size_t o = at_load(kR * kS);
    bm[0] = src[o];
    if (!ADD) {
      bw[0] = srcw[o];
      bq[0] = qsrc[at_q(kR * kS)];
    }
  }
  // whole rounds first (every step applies: one straight-line block per
  // round, a scheduling region per step, in which the exchange and the
  // refill sit among the dependent additions), then the last, partial round
  auto round = [&](uint32_t g, bool whole) {
#pragma unroll
    for (int r = 0; r < kR; ++r) {
      const int rn = (r + 1) % kR;
      const uint32_t g0 = g + r * kS;
      const Step nxt = step_of(bm[rn], bw[ADD ? 0 : rn], bq[ADD ? 0 : rn]);
      if (whole || g0 < G) apply(cur);
      cur = nxt;
      // the barriers order the exchange and sums before the refill (and the
      // next exchange's LDS writes after these reads), keep the load from
      // sinking to its use a round later, so kR loads stay in flight, and
      // keep the scheduler from moving an exchange next to its use
      asm volatile("" : "+v"(l) : : "memory");
      __builtin_amdgcn_sched_barrier(0);
      const size_t o = at_load(g + (r + 1) * kS + kR * kS);
      bm[rn] = src[o];
      if (!ADD) {
        bw[rn] = srcw[o];
        bq[rn] = qsrc[at_q(g + (r + 1) * kS + kR * kS)];
      }
    }
  };
  uint32_t g = 0;
  for (; g + kR * kS <= G; g += kR * kS) round(g, true);
  if (g < G) round(g, false);
  if (a.trace && c == 0 && h == 0 && lane == 0) {
    uint64_t *tr = a.trace + 4 * static_cast<uint64_t>(i);
    tr[0] = t_start;
    tr[1] = __builtin_amdgcn_s_memrealtime();
    tr[2] = static_cast<uint64_t>(len) | (uint64_t(3) << 32);
    tr[3] = static_cast<uint64_t>(__smid());
  }
  if (pix_in) *d = ug.accumulate == 1 ? *d + l : l;
}

// Work items per unit: none for units outside their tile; none for empty units
// when accumulating; one producer/consumer item for lists of at least
// heavy_len; otherwise three (one per channel) for long ordered lists, else one.
// (need_empty: the job's pass writes its whole image, so empty blocks are
// written too.)
// The list length from which a unit goes to the producer/consumer kernel:
// none when the pass has a particle with a non-finite q_c.
__device__ __forceinline__ uint32_t heavy_len_of(const SplatArgs &a) {
  return (a.qbad && *a.qbad) ? 0xffffffffu : a.heavy_len;
}

// The list length from which a unit is staged: none when the pass has a
// particle with a non-finite q_c (a staged ordered list applies zero operands
// outside the footprints, as the producer/consumer kernel does).
__device__ __forceinline__ uint32_t stage_len_of(const SplatArgs &a) {
  return (a.qbad && *a.qbad) ? 0xffffffffu : a.stage_len;
}

// Long lists (one item each, sorted first): staged lists, then the
// producer/consumer kernel's.
__device__ __forceinline__ uint32_t long_len_of(const SplatArgs &a) {
  return min(heavy_len_of(a), stage_len_of(a));
}

__device__ __forceinline__ uint32_t unit_items(const SplatArgs &a, uint32_t u, int ordered,
                                               uint32_t split, uint32_t &len) {
  len = a.bin_off[u + 1] - a.bin_off[u];
  const UnitGeom g = unit_geom(a, u);
  if (g.gx * 8 >= g.w || g.gy * 8 >= g.h) return 0;
  if (len == 0) return g.need_empty ? 1u : 0u;
  if (len >= long_len_of(a)) return 1u;  // staged, or one producer/consumer workgroup
  return (len >= (ordered ? split : 2 * split)) ? 3u : 1u;
}

__global__ __launch_bounds__(256) void k_items_count(SplatArgs a, uint32_t nunits, int ordered,
                                                     uint32_t split, uint32_t *cnt,
                                                     uint32_t *heavy_cnt, uint32_t *stage_cnt,
                                                     uint32_t *stage_grp, uint32_t *lmax) {
  for (int64_t u = gtid(); u < nunits; u += gstride()) {
    uint32_t len;
    const uint32_t k = unit_items(a, static_cast<uint32_t>(u), ordered, split, len);
    cnt[u] = k;
    if (heavy_cnt && k > 0 && len >= heavy_len_of(a)) atomicAdd(heavy_cnt, k);
    const uint32_t sl = stage_len_of(a);
    if (stage_cnt && k > 0 && len >= sl) {  // counts per level of sl
      for (int j = 0; j < kStageLevels; ++j) {
        if (static_cast<uint64_t>(len) < stage_level(sl, j)) break;
        atomicAdd(stage_cnt + j, 1u);
        atomicAdd(stage_grp + j, stage_groups(len));
      }
    }
    if (lmax && k > 0) atomicMax(lmax, len);
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
                        hipStream_t s, uint32_t *stage_cnt, uint32_t *stage_grp,
                        uint32_t *lmax) {
  if (nunits == 0) return;
  SplatArgs b = a;
  b.bin_off = bin_off;
  k_items_count<<<grid_for(nunits), 256, 0, s>>>(b, nunits, ordered, split, cnt, heavy_cnt,
                                                 stage_cnt, stage_grp, lmax);
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

uint64_t stage_floats(bool additive, uint64_t groups) {
  return groups * (additive ? 3u : 4u) * 256u;
}

uint64_t stage_q_floats(bool additive, uint64_t groups) {
  return additive ? 0 : groups / kStageQ * 48u;
}

void launch_stage_base(const uint32_t *items, uint32_t n, uint32_t *cbase, uint32_t *cbat,
                       const SplatArgs &a, hipStream_t s, const uint32_t *dn) {
  if (n == 0) return;
  k_stage_base<<<1, 256, 0, s>>>(a, items, n, cbase, cbat, dn);
  hip_check(hipGetLastError(), "k_stage_base launch");
}

// Workgroups of kernel k (256 threads) resident at once on the current device.
unsigned resident_grid(const void *k, int slot) {
  static unsigned cache[2][16] = {};
  int dev = 0;
  hip_check(hipGetDevice(&dev), "get device");
  unsigned &r = cache[slot][dev & 15];
  if (!r) {
    int cus = 0, per = 0;
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
    hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, 256, 0), "occupancy");
    r = static_cast<unsigned>(std::max(1, cus * per));
  }
  return r;
}

void launch_stage_ops(bool additive, const uint32_t *items, uint32_t i0, uint32_t i1,
                      uint64_t groups, const uint32_t *cbase, const uint32_t *cbat, float4 *ops,
                      float *qv, const SplatArgs &a, hipStream_t s, const uint32_t *dn) {
  if (i1 <= i0) return;
  // batches: at most groups / 16 + lists; about four per wave when there are
  // many, in one round of resident workgroups (a second round would queue
  // behind the 8x8 splat's workgroups, and the chains wait for all of them)
  const uint64_t nbat = groups / 16 + (i1 - i0);
  const unsigned res = additive ? resident_grid(reinterpret_cast<const void *>(k_stage_ops<1>), 1)
                                : resident_grid(reinterpret_cast<const void *>(k_stage_ops<0>), 0);
  const unsigned grid =
      std::max(1u, static_cast<unsigned>(std::min<uint64_t>(res, (nbat + 15) / 16)));
  if (additive) k_stage_ops<1><<<grid, 256, 0, s>>>(a, items, i0, i1, cbase, cbat, ops, qv, dn);
  else k_stage_ops<0><<<grid, 256, 0, s>>>(a, items, i0, i1, cbase, cbat, ops, qv, dn);
  hip_check(hipGetLastError(), "k_stage_ops launch");
}

void launch_stage_chain(bool additive, const uint32_t *items, uint32_t i0, uint32_t i1,
                        const uint32_t *cbase, const float4 *ops, const float *qv,
                        const SplatArgs &a, hipStream_t s, const uint32_t *dn) {
  if (i1 <= i0) return;
  const unsigned grid = 3u * static_cast<unsigned>(additive ? chain_groups<1>() : chain_groups<0>()) *
                        (i1 - i0);
  if (additive) k_stage_chain<1><<<grid, 64, 0, s>>>(a, items, i0, cbase, ops, qv, dn);
  else k_stage_chain<0><<<grid, 64, 0, s>>>(a, items, i0, cbase, ops, qv, dn);
  hip_check(hipGetLastError(), "k_stage_chain launch");
}

__global__ void k_partition(PartitionArgs a, uint32_t *out) {
  uint32_t ns = 0;
  if (a.len0 != 0xffffffffu) {
    const int j = stage_rule(a.len0, a.ordered_auto != 0, *a.lmax, *a.etot, a.span, a.ratio,
                             a.crit);
    if (j >= 0) {
      ns = a.stage_cnt[j];
      if (ns > a.ns_cap || a.stage_grp[j] > a.groups_cap) ns = 0;
    }
  }
  const uint32_t nh = *a.heavy_cnt;
  const uint32_t npc = (a.heavy_max == 0 || nh <= ns) ? 0u : min(nh - ns, a.heavy_max);
  const uint32_t nlong = a.len0 <= a.heavy_len ? a.stage_cnt[0] : nh;
  out[0] = ns;
  out[1] = npc;
  out[2] = ns + npc;
  out[3] = split_lists(a.len0, a.stage_cnt, nlong, ns + npc, *a.etot, a.split);
}

void launch_partition(const PartitionArgs &pa, uint32_t *out, hipStream_t s) {
  k_partition<<<1, 1, 0, s>>>(pa, out);
  hip_check(hipGetLastError(), "k_partition launch");
}

void launch_splat_pc(bool additive, const uint32_t *items, uint32_t npc, const SplatArgs &a,
                     hipStream_t s, const ItemCounts &dc) {
  if (npc == 0) return;
  if (additive) k_splat_pc<1><<<npc, kPThreads, 0, s>>>(a, items, npc, dc);
  else k_splat_pc<0><<<npc, kPThreads, 0, s>>>(a, items, npc, dc);
  hip_check(hipGetLastError(), "k_splat_pc launch");
}

void launch_splat8(bool additive, int variant, const uint32_t *items, uint32_t nitems,
                   const SplatArgs &a, hipStream_t s, const ItemCounts &dc, int ppw,
                   uint32_t nsplit) {
  if (nitems == 0) return;
  // device counts: nitems, a capacity of three items per unit, also bounds
  // the items with the long lists served by channel
  if (dc.items) nsplit = 0;
  const unsigned grid = static_cast<unsigned>((uint64_t(nitems) + 2u * uint64_t(nsplit) + 3) / 4);
#define SPLAT8(ADDV, VARV, PPWV) k_splat8<ADDV, VARV, PPWV><<<grid, 256, 0, s>>>(a, items, nitems, dc, nsplit)
  if (variant == 4) {
    switch (ppw) {
      case 1: if (additive) SPLAT8(1, 4, 1); else SPLAT8(0, 4, 1); break;
      case 2: if (additive) SPLAT8(1, 4, 2); else SPLAT8(0, 4, 2); break;
      case 4: if (additive) SPLAT8(1, 4, 4); else SPLAT8(0, 4, 4); break;
      case 8: if (additive) SPLAT8(1, 4, 8); else SPLAT8(0, 4, 8); break;
      default: if (additive) SPLAT8(1, 4, 0); else SPLAT8(0, 4, 0); break;
    }
  } else if (variant == 3) {
    if (additive) SPLAT8(1, 3, 0); else SPLAT8(0, 3, 0);
  } else if (variant == 2) {
    if (additive) SPLAT8(1, 2, 0); else SPLAT8(0, 2, 0);
  } else {
    if (additive) SPLAT8(1, 1, 0); else SPLAT8(0, 1, 0);
  }
#undef SPLAT8
  hip_check(hipGetLastError(), "k_splat8 launch");
}

}  // namespace splhip
