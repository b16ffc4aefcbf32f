// gfx950 kernels of the Splotch particle-splat hot path.
//
//   k_project      particle_project + particle_colorize (splotch_host.cc:58-287)
//                  over SoA streams; FP64 transform with the float32 matrix
//                  exactly as TRANSFORM::TransPoint (kernel/transform.cxx:212-219).
//   k_compact      stream compaction of one pass's particles (+ depth-sort keys).
//   k_gather_count gather into render order + per-particle bin count.
//   k_emit         (bin, particle) entries in particle order, balanced over
//                  the wave in two levels (tiles, then units).
//   k_bin_offsets  CSR offsets from the sorted bin keys (binary search).
//   k_splat        one workgroup per work region: the region's particles are
//                  staged in LDS in batches, per-particle column ranges and the
//                  separable Gaussian factors are tabulated in LDS
//                  (splotch_host.cc:596-607), and each lane owns one pixel of an
//                  8x8 block and accumulates in registers in list order, i.e.
//                  in the reference's per-pixel summation order.
#include "kernels.hpp"

#include <algorithm>

#include "workspace.hpp"

namespace splhip {
namespace {

__device__ __forceinline__ int64_t gtid() {
  return static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
}

__device__ __forceinline__ int64_t gstride() {
  return static_cast<int64_t>(gridDim.x) * blockDim.x;
}

__device__ __forceinline__ void wave_sync_() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

unsigned grid_for(int64_t n, int per_block = 256, int64_t cap = 1 << 20) {
  int64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return static_cast<unsigned>(g);
}

// ---------------------------------------------------------------- project

// The job of a block is wave-uniform; with the tables marked restrict (the
// kernel writes only rec/act) its fields load through the scalar cache.
__global__ __launch_bounds__(256) void k_project(const ProjectArgs *__restrict__ jobs,
                                                 const uint32_t *__restrict__ blk, int nj,
                                                 Rec *__restrict__ rec,
                                                 uint8_t *__restrict__ act) {
  const int j = __builtin_amdgcn_readfirstlane(find_job(blk, nj, blockIdx.x));
  const ProjectArgs &a = jobs[j];
  const int64_t i0 = static_cast<int64_t>(blockIdx.x - blk[j]) * kPerBlock;
  const int64_t iend = min(i0 + kPerBlock, a.n);
  for (int64_t i = i0 + threadIdx.x; i < iend; i += blockDim.x) {
    const int t = a.type ? a.type[i] : 0;
    const int tt = t < kMaxTypes ? t : kMaxTypes - 1;
    float px = a.x[i], py = a.y[i], pz = a.z[i];
    float R = a.hsml[i], II = a.inten[i];
    float e0 = a.col[i];
    float e1 = a.c1 ? a.c1[i] : 0.f;
    float e2 = a.c2 ? a.c2[i] : 0.f;
    const float *m = a.cam.mat;
    // TransPoint: double accumulation, left to right, no contraction.
    const double dx = px, dy = py, dz = pz;
    const double vx = dx * static_cast<double>(m[0]) + dy * static_cast<double>(m[1]) +
                      dz * static_cast<double>(m[2]) + static_cast<double>(m[3]);
    const double vy = dx * static_cast<double>(m[4]) + dy * static_cast<double>(m[5]) +
                      dz * static_cast<double>(m[6]) + static_cast<double>(m[7]);
    const double vz = dx * static_cast<double>(m[8]) + dy * static_cast<double>(m[9]) +
                      dz * static_cast<double>(m[10]) + static_cast<double>(m[11]);
    px = static_cast<float>(vx);
    py = static_cast<float>(vy);
    pz = static_cast<float>(vz);
    uint8_t code = 0;
    bool live = (t < a.ntypes) && !(-pz <= a.cam.zminval) && !(-pz >= a.cam.zmaxval);
    if (live) {
      float xfac2 = a.cam.xfac;
      const float res2 = a.cam.res2;
      if (!a.cam.projection) {
        px = res2 + px * xfac2;
        py = res2 + a.cam.ycorr + py * xfac2;
      } else {
        xfac2 = -res2 / (a.cam.fovfct * pz);
        px = static_cast<float>(static_cast<double>(res2) + a.cam.xshift +
                                static_cast<double>(px * xfac2));
        py = static_cast<float>(static_cast<double>(res2) + a.cam.yshift +
                                static_cast<double>(a.cam.ycorr) +
                                static_cast<double>(py * xfac2));
      }
      II *= (0.5f * a.bfak) / R;
      R *= 2;
      R *= xfac2 * a.smooth_fac[tt];
      const float mr = a.cam.minrad_pix;
      const float rcorr = sqrtf(R * R + mr * mr) / R;
      R *= rcorr;
      II /= rcorr;
      const float rfacr = a.rfac * R;
      const int xres = a.cam.xres, yres = a.cam.yres;
      int minx = cvt_i32(px - rfacr + 1);
      int maxx = cvt_i32(px + rfacr + 1);
      int miny = cvt_i32(py - rfacr + 1);
      int maxy = cvt_i32(py + rfacr + 1);
      bool act = !(minx >= xres);
      minx = max(minx, 0);
      act = act && !(maxx <= 0);
      maxx = min(maxx, xres);
      act = act && !(minx >= maxx);
      act = act && !(miny >= yres);
      miny = max(miny, 0);
      act = act && !(maxy <= 0);
      maxy = min(maxy, yres);
      act = act && !(miny >= maxy);
      if (act) {
        // particle_colorize: palette lookup (linear_map.h:71-84, math_utils.h:186-202)
        if (!a.col_vector[tt]) {
          const double *X = a.pal_x + static_cast<int64_t>(tt) * a.pal_stride;
          const float *Y = a.pal_rgb + static_cast<int64_t>(tt) * a.pal_stride * 3;
          float cr, cg, cb;
          palette_val(X, Y, a.pal_n[tt], static_cast<double>(e0), cr, cg, cb);
          e0 = cr; e1 = cg; e2 = cb;
        }
        const float fac = II * a.brightness[tt];
        e0 *= fac; e1 *= fac; e2 *= fac;
        code = static_cast<uint8_t>(a.pass_code[tt]);
      }
    }
    Rec o;
    o.x = px; o.y = py; o.r = R; o.e0 = e0; o.e1 = e1; o.e2 = e2; o.z = pz; o.I = II;
    rec[a.poff + i] = o;
    act[a.poff + i] = code;
  }
}

// ---------------------------------------------------------------- compaction

// Order-preserving float key (ascending); sort types of splotchutils.h:74-96.
__device__ __forceinline__ uint32_t sort_key(const Rec &R, int sort_type) {
  if (sort_type == 0) return 0u;  // unsorted: the stable sort keeps index order
  float f = (sort_type == 1) ? R.z : (sort_type == 4 ? R.r : R.e0);
  if (f == 0.f) f = 0.f;  // -0 and +0 compare equal
  uint32_t u = __float_as_uint(f);
  u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  if (sort_type == 2 || sort_type == 4) u = ~u;  // descending
  return u;
}

__global__ __launch_bounds__(256) void k_compact(const ProjectArgs *__restrict__ jobs,
                                                 const uint32_t *__restrict__ blk, int nj,
                                                 const uint8_t *__restrict__ act, uint8_t code,
                                                 const uint32_t *__restrict__ pos,
                                                 const Rec *__restrict__ rec,
                                                 uint32_t *__restrict__ keys,
                                                 uint32_t *__restrict__ vals) {
  const int j = __builtin_amdgcn_readfirstlane(find_job(blk, nj, blockIdx.x));
  const int64_t n = jobs[j].n, poff = jobs[j].poff;
  const int sort_type = jobs[j].sort_type;
  const int64_t i0 = static_cast<int64_t>(blockIdx.x - blk[j]) * kPerBlock;
  const int64_t iend = min(i0 + kPerBlock, n);
  for (int64_t k = i0 + threadIdx.x; k < iend; k += blockDim.x) {
    const int64_t i = poff + k;
    if (!(act[i] & code)) continue;
    const uint32_t p = pos[i];
    vals[p] = static_cast<uint32_t>(i);
    if (keys) keys[p] = sort_key(rec[i], sort_type);
  }
}

__global__ __launch_bounds__(256) void k_pass_offsets(const uint32_t *pos, const uint32_t *pbase,
                                                      int nj, int64_t n, const uint32_t *total,
                                                      uint32_t *mo) {
  for (int64_t j = gtid(); j <= nj; j += gstride())
    mo[j] = (j < nj && pbase[j] < n) ? pos[pbase[j]] : *total;
}

// The count of a launch sized for `cap` items: cap, or the device count *dn
// when it is given (at most cap).
__device__ __forceinline__ uint32_t dcount(uint32_t cap, const uint32_t *dn) {
  return dn ? min(cap, *dn) : cap;
}

__global__ __launch_bounds__(256) void k_job_of_index(const uint32_t *vals, uint32_t cap,
                                                      const uint32_t *dm, const uint32_t *pbase,
                                                      int nj, uint32_t *keys) {
  const uint32_t m = dcount(cap, dm);
  for (int64_t i = gtid(); i < m; i += gstride())
    keys[i] = static_cast<uint32_t>(find_job(pbase, nj, vals[i]));
}

__global__ __launch_bounds__(256) void k_gather_count(const uint32_t *order, uint32_t cap,
                                                      const uint32_t *dm, const Rec *rec,
                                                      Rec *recs, uint32_t *cnt,
                                                      uint32_t *cnt16, int rsh,
                                                      const BinGeom *geoms, const uint32_t *mo,
                                                      int nj, uint32_t *qbad) {
  const uint32_t m = dcount(cap, dm);
  for (int64_t s = gtid(); s < m; s += gstride()) {
    const BinGeom g = geoms[find_job(mo, nj, static_cast<uint32_t>(s))];
    const Rec R = rec[order[s]];
    recs[s] = R;
    if (qbad) {  // ordered pass: a non-finite q_c = e_c / (e_c + gray_absorption)
      const float q0 = R.e0 / (R.e0 + g.grayabsorb), q1 = R.e1 / (R.e1 + g.grayabsorb),
                  q2 = R.e2 / (R.e2 + g.grayabsorb);
      const float inf = __builtin_inff();
      if (!(fabsf(q0) < inf && fabsf(q1) < inf && fabsf(q2) < inf)) atomicOr(qbad, 1u);
    }
    const BinRange b = bin_range(R, g);
    uint32_t c = 0, c16 = 0;  // units of the binned tiles; their regions (2^rsh units square)
    for (int ix = b.minx; ix <= b.maxx; ++ix)
      for (int iy = b.miny; iy <= b.maxy; ++iy) {
        BinRect rc;
        if (bin_tile(R, g, b, ix, iy, rc)) {
          c += static_cast<uint32_t>((rc.gx1 - rc.gx0 + 1) * (rc.gy1 - rc.gy0 + 1));
          c16 += static_cast<uint32_t>(((rc.gx1 >> rsh) - (rc.gx0 >> rsh) + 1) *
                                       ((rc.gy1 >> rsh) - (rc.gy0 >> rsh) + 1));
        }
      }
    cnt[s] = c;
    if (cnt16) cnt16[s] = c16;
  }
}

// Entries of 64 consecutive particles (one wave) are one contiguous range of
// the output (off is their exclusive scan); a particle's entries go tile
// column by column (ix outer, iy inner), each tile's units row by row.
// Per-particle entry counts vary by orders of magnitude, so the work is
// split in two balanced levels: each lane walks its particle's
// candidate tiles (one per step) and queues every binned tile as an item (a
// rectangle of units and its output position) in LDS; once 64 items are
// queued, the wave writes the items' entries with one entry per lane.
constexpr int kEmitQ = 128;  // a step queues at most 64 items; drained from 64 on

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  return v;
}

struct EmitQueue {
  uint32_t pos[kEmitQ], key[kEmitQ], val[kEmitQ], nx_ngm[kEmitQ], n[kEmitQ], pre[kEmitQ];
  uint32_t rect[kEmitQ];  // R == 2: the tile's 8x8-block rectangle gx0 | gx1 | gy0 | gy1 bytes
  float px[kEmitQ], py[kEmitQ], rf[kEmitQ];  // R == 2: tile-frame centre, footprint radius
};

// Mask of the units (S GX + sx, S GY + sy) of region (GX, GY), S = 2^rsh, bit
// S sy + sx, that lie in the rectangle rc = gx0 | gx1 << 8 | gy0 << 16 | gy1 << 24
// (the region lies in the rectangle's region range, so each axis has one).
__device__ __forceinline__ uint32_t region_mask(uint32_t rc, int GX, int GY, int rsh) {
  const int gx0 = rc & 0xff, gx1 = (rc >> 8) & 0xff, gy0 = (rc >> 16) & 0xff, gy1 = rc >> 24;
  const int S = 1 << rsh;
  const int lx = max(gx0 - S * GX, 0), hx = min(gx1 - S * GX, S - 1);
  const int ly = max(gy0 - S * GY, 0), hy = min(gy1 - S * GY, S - 1);
  const uint32_t xm = ((2u << hx) - 1u) & ~((1u << lx) - 1u);
  uint32_t m = 0;
  for (int sy = ly; sy <= hy; ++sy) m |= xm << (sy * S);
  return m;
}

// Blocks (bit S sy + sx) of region (GX, GY), S = 2^rsh, that the footprint of
// radius rf about tile-frame point (px, py) can reach.  render_new updates
// pixel column x only for x in (posx - rfacr, posx + rfacr] and, within it,
// rows y with |y - posy| <= sqrtf(radsq - (x - posx)^2) (splotch_host.cc:
// 591-607), so every updated pixel lies within rfacr of the centre up to the
// float rounding of those bounds, which stays far below the 2 px and 2^-16
// relative margins used here.  A block none of whose pixels [8 b, 8 b + 7]
// per axis is that close gets no update from the particle: its entry would
// add +0 to every pixel, and leaving it out of the block's list keeps the
// image bit for bit.  NaN or infinite operands keep every block.
__device__ __forceinline__ uint32_t region_reach(float px, float py, float rf, int GX, int GY,
                                                 int rsh) {
  const int S = 1 << rsh;
  const float lim = rf + 2.f;
  const float thr = (lim * lim) * 1.0000153f;
  float dy2[4];
#pragma unroll
  for (int sy = 0; sy < 4; ++sy) {
    const float y0 = static_cast<float>(8 * (S * GY + sy));
    const float d = fmaxf(fmaxf(y0 - py, py - (y0 + 7.f)), 0.f);
    dy2[sy] = d * d;
  }
  uint32_t m = 0;
#pragma unroll
  for (int sx = 0; sx < 4; ++sx) {
    if (sx >= S) break;
    const float x0 = static_cast<float>(8 * (S * GX + sx));
    const float d = fmaxf(fmaxf(x0 - px, px - (x0 + 7.f)), 0.f);
    const float dx2 = d * d;
#pragma unroll
    for (int sy = 0; sy < 4; ++sy)
      if (sy < S && !(dx2 + dy2[sy] > thr)) m |= 1u << (sy * S + sx);
  }
  return m;
}

// Write the entries of items [0, qn) of the queue: item k covers output
// [pos, pos + n), entry i = (gy, gx) = (i / nx, i % nx) is unit key + gy*ngm + gx
// (R == 1: 8x8 blocks; R == 2: regions of 2^rsh x 2^rsh blocks, key | mask
// << rbits of the region's blocks in the rectangle that the footprint
// reaches; a region whose mask comes out empty adds to no block list).
template <int R>
__device__ __forceinline__ void emit_drain(EmitQueue &Q, int qn, int lane, uint32_t *keys,
                                           uint32_t *vals, int rsh, int rbits) {
  const uint32_t c0 = lane < qn ? Q.n[lane] : 0u, c1 = lane + 64 < qn ? Q.n[lane + 64] : 0u;
  const uint32_t i0 = wave_incl_scan(c0, lane), i1 = wave_incl_scan(c1, lane);
  const uint32_t t0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(i0), 63));
  const uint32_t tot = t0 + static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(i1), 63));
  Q.pre[lane] = i0 - c0;
  Q.pre[lane + 64] = t0 + i1 - c1;
  wave_sync_();
  for (uint32_t t = 0; t < tot; t += 64) {
    const uint32_t q = t + lane;
    if (q < tot) {
      int k = 0;  // last item with pre <= q (items have n > 0, so pre ascends strictly)
#pragma unroll
      for (int step = 64; step; step >>= 1)
        if (k + step < qn && Q.pre[k + step] <= q) k += step;
      const uint32_t i = q - Q.pre[k];
      const uint32_t nn = Q.nx_ngm[k];
      const int nx = static_cast<int>(nn & 0xffffu), ngm = static_cast<int>(nn >> 16);
      int gy = static_cast<int>(static_cast<float>(i) * __builtin_amdgcn_rcpf(static_cast<float>(nx)));
      int gx = static_cast<int>(i) - gy * nx;
      if (gx < 0) {
        --gy;
        gx += nx;
      } else if (gx >= nx) {
        ++gy;
        gx -= nx;
      }
      const uint32_t o = Q.pos[k] + i;
      uint32_t key = Q.key[k] + static_cast<uint32_t>(gy * ngm + gx);
      if (R == 2) {
        const uint32_t rc = Q.rect[k];
        const int GX = (static_cast<int>(rc & 0xff) >> rsh) + gx;
        const int GY = (static_cast<int>((rc >> 16) & 0xff) >> rsh) + gy;
        key |= (region_mask(rc, GX, GY, rsh) & region_reach(Q.px[k], Q.py[k], Q.rf[k], GX, GY,
                                                            rsh))
               << rbits;
      }
      keys[o] = key;
      vals[o] = Q.val[k];
    }
  }
  wave_sync_();
}

// R == 1: (8x8 block, particle) entries.  R == 2: (region | unit mask << rbits,
// particle) entries for regions of S x S blocks (S = 2^rsh), the mask naming
// the region's blocks the particle is binned to; regions of job j are numbered
// from ubase[j], ceil(ngm / S) per tile row.
template <int R>
__global__ __launch_bounds__(256) void k_emit(const Rec *recs, uint32_t mcap, const uint32_t *dm,
                                              const uint32_t *off, uint32_t ecap,
                                              const uint32_t *de, const BinGeom *geoms,
                                              const uint32_t *ubase, const uint32_t *mo, int nj,
                                              uint32_t *keys, uint32_t *vals, int rsh,
                                              int rbits) {
  __shared__ EmitQueue s_q[4];
  const uint32_t m = dcount(mcap, dm), e = dcount(ecap, de);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  EmitQueue &Q = s_q[wave];
  const int64_t nwaves = static_cast<int64_t>(gridDim.x) * 4;
  for (int64_t w = static_cast<int64_t>(blockIdx.x) * 4 + wave; w * 64 < m; w += nwaves) {
    const int64_t s = w * 64 + lane;
    const bool valid = s < m;
    BinGeom g{};
    Rec P{};
    BinRange br{0, -1, 0, -1, 0.f};
    uint32_t ub = 0, q = 0, end = 0;
    if (valid) {
      const int j = find_job(mo, nj, static_cast<uint32_t>(s));
      g = geoms[j];
      ub = ubase[j];
      P = recs[s];
      q = off[s];
      end = (s + 1 < m) ? off[s + 1] : e;  // same enumeration as the count
      br = bin_range(P, g);
    }
    bool more = br.minx <= br.maxx && br.miny <= br.maxy;
    int ix = br.minx, iy = br.miny;
    int qn = 0;  // uniform
    while (__ballot(more)) {
      bool have = false;
      uint32_t ipos = 0, ikey = 0, inn = 0, in = 0, irect = 0;
      float ipx = 0.f, ipy = 0.f, irf = 0.f;
      if (more) {
        BinRect rc;
        if (bin_tile(P, g, br, ix, iy, rc)) {
          uint32_t nx, cnt;
          if (R == 1) {
            nx = static_cast<uint32_t>(rc.gx1 - rc.gx0 + 1);
            cnt = nx * static_cast<uint32_t>(rc.gy1 - rc.gy0 + 1);
            ikey = ub + rc.base + static_cast<uint32_t>(rc.gy0 * g.ngm + rc.gx0);
            inn = nx | (static_cast<uint32_t>(g.ngm) << 16);
          } else {
            const int nr = (g.ngm + (1 << rsh) - 1) >> rsh;
            const uint32_t tile = static_cast<uint32_t>(iy * g.ncx + ix);
            nx = static_cast<uint32_t>((rc.gx1 >> rsh) - (rc.gx0 >> rsh) + 1);
            cnt = nx * static_cast<uint32_t>((rc.gy1 >> rsh) - (rc.gy0 >> rsh) + 1);
            ikey = ub + tile * static_cast<uint32_t>(nr * nr) +
                   static_cast<uint32_t>((rc.gy0 >> rsh) * nr + (rc.gx0 >> rsh));
            inn = nx | (static_cast<uint32_t>(nr) << 16);
            irect = static_cast<uint32_t>(rc.gx0) | (static_cast<uint32_t>(rc.gx1) << 8) |
                    (static_cast<uint32_t>(rc.gy0) << 16) | (static_cast<uint32_t>(rc.gy1) << 24);
            ipx = P.x - static_cast<float>(ix * g.T);  // as bin_tile's posx, posy, rfacr
            ipy = P.y - static_cast<float>(iy * g.T);
            irf = P.r * g.rfac;
          }
          in = min(cnt, end - q);
          have = in > 0;
          ipos = q;
          q += in;
        }
        if (++iy > br.maxy) {
          iy = br.miny;
          more = ++ix <= br.maxx;
        }
      }
      const uint64_t hb = __ballot(have);
      if (have) {
        const int slot = qn + static_cast<int>(__builtin_amdgcn_mbcnt_hi(
                                  static_cast<uint32_t>(hb >> 32),
                                  __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(hb), 0u)));
        Q.pos[slot] = ipos;
        Q.key[slot] = ikey;
        Q.val[slot] = static_cast<uint32_t>(s);
        Q.nx_ngm[slot] = inn;
        Q.n[slot] = in;
        Q.rect[slot] = irect;
        if (R == 2) {
          Q.px[slot] = ipx;
          Q.py[slot] = ipy;
          Q.rf[slot] = irf;
        }
      }
      qn += __builtin_popcountll(hb);
      if (qn >= 64) {
        wave_sync_();
        emit_drain<R>(Q, qn, lane, keys, vals, rsh, rbits);
        qn = 0;
      }
    }
    if (qn) {
      wave_sync_();
      emit_drain<R>(Q, qn, lane, keys, vals, rsh, rbits);
    }
  }
}

// A region (id r of S x S blocks, job j = the one whose region range holds r)
// and its blocks: unit ids (or ~0u where the block lies past the tile's ngm),
// u[sy S + sx].
template <int S>
__device__ __forceinline__ void region_units(uint32_t r, const uint32_t *rbase,
                                             const uint32_t *ubase, const BinGeom *geoms, int nj,
                                             uint32_t (&u)[S * S]) {
  const int j = find_job(rbase, nj, r);
  const int ngm = geoms[j].ngm, nr = (ngm + S - 1) / S;
  const uint32_t local = r - rbase[j];
  const uint32_t per = static_cast<uint32_t>(nr * nr);
  const uint32_t tile = local / per;
  const int rem = static_cast<int>(local - tile * per);
  const int GY = rem / nr, GX = rem - GY * nr;
#pragma unroll
  for (int b = 0; b < S * S; ++b) {
    const int gx = S * GX + (b % S), gy = S * GY + (b / S);
    u[b] = (gx < ngm && gy < ngm)
               ? ubase[j] + tile * static_cast<uint32_t>(ngm * ngm) +
                     static_cast<uint32_t>(gy * ngm + gx)
               : ~0u;
  }
}

// Splitting the sorted region lists into block lists.  An entry's key is
// region | mask << rbits.  S_b[i] is the number of entries before position i
// of the sorted array whose mask has bit b; an entry i of region r with bit b
// goes to position
//   boff[unit(r, b)] + S_b[i] - S_b[roff[r]]
// of its block's list, which keeps list order.  S_b comes from per-chunk bit
// counts (chunks of kSplitChunk entries, one wave each, so a long region list
// is spread over many waves) scanned as one array laid out [bit][chunk]: the
// bits' offsets cancel in every difference.
constexpr int kSplitChunk = 1024;

__device__ __forceinline__ uint32_t mbcnt64(uint64_t bal) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(bal >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(bal), 0u));
}

// The value of v[b] on lane b (b < NB), as a lane select (no private-array
// indexing by lane).
template <int NB>
__device__ __forceinline__ uint32_t lane_pick(const uint32_t (&v)[NB], int lane) {
  uint32_t x = 0;
#pragma unroll
  for (int b = 0; b < NB; ++b) x = lane == b ? v[b] : x;
  return x;
}

template <int NB>
__global__ __launch_bounds__(256) void k_split_chunks(const uint32_t *keys, uint32_t ecap,
                                                      const uint32_t *de, uint32_t nchunks,
                                                      int rbits, uint32_t *ccnt) {
  const uint32_t e16 = dcount(ecap, de);
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = static_cast<int64_t>(gridDim.x) * 4;
  for (int64_t c = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6); c < nchunks;
       c += nwaves) {
    const uint32_t i0 = static_cast<uint32_t>(c) * kSplitChunk;
    const uint32_t i1 = min(i0 + kSplitChunk, e16);
    uint32_t n[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) n[b] = 0;
    for (uint32_t i = i0; i < i1; i += 64) {
      const uint32_t mk = (i + lane < i1) ? (keys[i + lane] >> rbits) : 0u;
#pragma unroll
      for (int b = 0; b < NB; ++b)
        n[b] += static_cast<uint32_t>(__popcll(__ballot((mk >> b) & 1u)));
    }
    const uint32_t v = lane_pick<NB>(n, lane);
    if (lane < NB) ccnt[static_cast<int64_t>(lane) * nchunks + c] = v;
  }
}

// S_b at position x (x <= e16) for every bit b, from the scanned chunk counts.
template <int NB>
__device__ __forceinline__ void split_prefix(const uint32_t *keys, const uint32_t *cpre,
                                             uint32_t nchunks, uint32_t x, int rbits, int lane,
                                             uint32_t (&n)[NB]) {
  const uint32_t c = x / kSplitChunk;
#pragma unroll
  for (int b = 0; b < NB; ++b) n[b] = cpre[static_cast<int64_t>(b) * nchunks + c];
  for (uint32_t i = c * kSplitChunk; i < x; i += 64) {
    const uint32_t mk = (i + lane < x) ? (keys[i + lane] >> rbits) : 0u;
#pragma unroll
    for (int b = 0; b < NB; ++b)
      n[b] += static_cast<uint32_t>(__popcll(__ballot((mk >> b) & 1u)));
  }
}

// One wave per region: S_b at the region's start (sx[NB r + b]) and its
// blocks' list lengths (ucnt; every unit is written once).
template <int S>
__global__ __launch_bounds__(256) void k_split_bounds(const uint32_t *roff, const uint32_t *keys,
                                                      uint32_t nregions, const uint32_t *cpre,
                                                      uint32_t nchunks, int rbits,
                                                      const uint32_t *rbase,
                                                      const uint32_t *ubase,
                                                      const BinGeom *geoms, int nj,
                                                      uint32_t *sx, uint32_t *ucnt) {
  constexpr int NB = S * S;
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = static_cast<int64_t>(gridDim.x) * 4;
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6); r < nregions;
       r += nwaves) {
    const uint32_t x0 = roff[r], x1 = roff[r + 1];
    uint32_t s0[NB], s1[NB], u[NB];
    split_prefix<NB>(keys, cpre, nchunks, x0, rbits, lane, s0);
    if (x1 == x0) {
#pragma unroll
      for (int b = 0; b < NB; ++b) s1[b] = s0[b];
    } else {
      split_prefix<NB>(keys, cpre, nchunks, x1, rbits, lane, s1);
    }
    region_units<S>(static_cast<uint32_t>(r), rbase, ubase, geoms, nj, u);
    const uint32_t uu = lane_pick<NB>(u, lane);
    const uint32_t a0 = lane_pick<NB>(s0, lane), a1 = lane_pick<NB>(s1, lane);
    if (lane < NB) {
      sx[NB * r + lane] = a0;
      if (uu != ~0u) ucnt[uu] = a1 - a0;
    }
  }
}

// Per region and block: boff[unit] - S_b[region start] (0 for blocks past
// ngm, which no entry names).
template <int S>
__global__ __launch_bounds__(256) void k_split_base(uint32_t nregions, const uint32_t *rbase,
                                                    const uint32_t *ubase, const BinGeom *geoms,
                                                    int nj, const uint32_t *sx,
                                                    const uint32_t *boff, uint32_t *base) {
  constexpr int NB = S * S;
  for (int64_t r = gtid(); r < nregions; r += gstride()) {
    uint32_t u[NB];
    region_units<S>(static_cast<uint32_t>(r), rbase, ubase, geoms, nj, u);
    uint32_t o[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) o[b] = u[b] != ~0u ? boff[u[b]] - sx[NB * r + b] : 0u;
#pragma unroll
    for (int b = 0; b < NB; ++b) base[NB * r + b] = o[b];
  }
}

// One wave per chunk: every entry to its blocks' lists.
template <int S>
__global__ __launch_bounds__(256) void k_split_scatter(const uint32_t *keys, const uint32_t *vals,
                                                       uint32_t ecap, const uint32_t *de,
                                                       uint32_t nchunks, int rbits,
                                                       const uint32_t *cpre,
                                                       const uint32_t *base, uint32_t *ent) {
  const uint32_t e16 = dcount(ecap, de);
  constexpr int NB = S * S;
  const int lane = threadIdx.x & 63;
  const uint32_t rmask = (1u << rbits) - 1u;
  const int64_t nwaves = static_cast<int64_t>(gridDim.x) * 4;
  for (int64_t c = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6); c < nchunks;
       c += nwaves) {
    const uint32_t i0 = static_cast<uint32_t>(c) * kSplitChunk;
    const uint32_t i1 = min(i0 + kSplitChunk, e16);
    uint32_t n[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) n[b] = cpre[static_cast<int64_t>(b) * nchunks + c];
    // four batches of 64 entries at a time, block by block: a block's four
    // stores are one contiguous run of its list, issued back to back, so
    // they reach L2 as whole lines rather than 16 interleaved partial ones
    for (uint32_t i = i0; i < i1; i += 4 * 64) {
      uint32_t mk[4], idx[4];
      const uint32_t *bs[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t t = i + 64u * j + lane;
        const bool in = t < i1;
        const uint32_t k = in ? keys[t] : 0u;
        idx[j] = in ? vals[t] : 0u;
        mk[j] = in ? k >> rbits : 0u;  // no block for entries past the chunk
        bs[j] = base + static_cast<int64_t>(NB) * (k & rmask);
      }
#pragma unroll
      for (int b = 0; b < NB; ++b) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const bool on = (mk[j] >> b) & 1u;
          const uint64_t bal = __ballot(on);
          if (on) ent[bs[j][b] + n[b] + mbcnt64(bal)] = idx[j];
          n[b] += static_cast<uint32_t>(__popcll(bal));
        }
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_bin_offsets(const uint32_t *keys, uint32_t ecap,
                                                     const uint32_t *de, uint32_t nbins,
                                                     uint32_t kmask, uint32_t *bin_off) {
  const uint32_t e = dcount(ecap, de);
  for (int64_t r = gtid(); r <= nbins; r += gstride()) {
    uint32_t lo = 0, hi = e;
    while (lo < hi) {
      const uint32_t mid = lo + ((hi - lo) >> 1);
      if ((keys[mid] & kmask) < r) lo = mid + 1;
      else hi = mid;
    }
    bin_off[r] = lo;
  }
}

__global__ __launch_bounds__(256) void k_map(const uint32_t *idx, const uint32_t *table,
                                             uint32_t *out, int64_t n) {
  for (int64_t i = gtid(); i < n; i += gstride()) out[i] = table[idx[i]];
}

__global__ __launch_bounds__(256) void k_fill(float *p, float v, int64_t n) {
  for (int64_t i = gtid(); i < n; i += gstride()) p[i] = v;
}

// 64-bit sum of n u32 values (block partial sums added atomically).
__global__ __launch_bounds__(256) void k_sum_u64(const uint32_t *v, int64_t cap,
                                                 const uint32_t *dn, unsigned long long *out) {
  const int64_t n = dn ? min(cap, static_cast<int64_t>(*dn)) : cap;
  __shared__ unsigned long long part[256];
  unsigned long long acc = 0;
  for (int64_t i = gtid(); i < n; i += gstride()) acc += v[i];
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, part[0]);
}

__global__ __launch_bounds__(256) void k_add_images(float *pic, const float *add, int64_t n) {
  for (int64_t i = gtid(); i < n; i += gstride()) pic[i] = pic[i] + add[i];
}

// Bounded wait of `ticks` s_memrealtime ticks (100 MHz): one wave, then exit.
__global__ __launch_bounds__(64) void k_delay(uint32_t ticks) {
  const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
  while (__builtin_amdgcn_s_memrealtime() - t0 < ticks) __builtin_amdgcn_s_sleep(8);
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t *p, uint32_t v, int64_t n) {
  for (int64_t i = gtid(); i < n; i += gstride()) p[i] = v;
}

// Per-job pass flags from the pass offsets mo0 / mo1 (null: pass absent):
// the ordered pass writes the image of a job that has ordered particles; the
// additive pass adds to it there and writes the image elsewhere.
__global__ __launch_bounds__(64) void k_job_flags(const uint32_t *mo0, const uint32_t *mo1,
                                                  int nj, JobSplat *js) {
  for (int64_t j = gtid(); j < nj; j += gstride()) {
    const bool w0 = mo0 && mo0[j + 1] > mo0[j];
    const bool w1 = mo1 && mo1[j + 1] > mo1[j];
    JobSplat &J = js[j];
    J.accumulate[0] = 0;
    J.accumulate[1] = w0 ? 1 : 0;
    J.need_empty[0] = w0 ? 1 : 0;
    J.need_empty[1] = (w1 && !w0) ? 1 : 0;
  }
}

// Zeros into the image of every job with no particle in either pass
// (grid.y = job).
__global__ __launch_bounds__(256) void k_fill_empty_jobs(const uint32_t *mo0, const uint32_t *mo1,
                                                         const JobSplat *js) {
  const int j = blockIdx.y;
  if ((mo0 && mo0[j + 1] > mo0[j]) || (mo1 && mo1[j + 1] > mo1[j])) return;
  const JobSplat &J = js[j];
  const int64_t n = static_cast<int64_t>(J.xres) * J.yres * 3;
  for (int64_t i = gtid(); i < n; i += gstride()) J.pic[i] = 0.f;
}

// Capacity check of a pass binned with device counts: if its entries (u64
// total e64) exceed ecap, or its region entries (e16) exceed e16cap, the
// pass's particle and entry counts are set to zero (nothing is emitted or
// splatted, so nothing is written past a buffer) and bit `bit` of *ovf is set.
__global__ void k_cap_check(uint32_t *m, uint32_t *e, const unsigned long long *e64,
                            uint32_t *e16, uint64_t ecap, uint32_t e16cap, uint32_t bit,
                            uint32_t *ovf) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const bool over = *e64 > ecap || (e16 && *e16 > e16cap);
  if (over) {
    *m = 0;
    *e = 0;
    if (e16) *e16 = 0;
    *ovf |= bit;
  }
}

__global__ __launch_bounds__(256) void k_unpack_records(const char *stage, int64_t n, int64_t C,
                                                        int ncol, float *soa, int64_t nn,
                                                        uint8_t *type) {
  const int64_t rb = 4 * ncol + 1;
  for (int64_t i = gtid(); i < n; i += gstride()) {
    const int64_t k = i / C, j = i - k * C, cnt = min(C, n - k * C);
    const char *base = stage + k * C * rb;
    const float *col = reinterpret_cast<const float *>(base);
    for (int q = 0; q < ncol; ++q) soa[q * nn + i] = col[q * cnt + j];
    type[i] = static_cast<uint8_t>(base[4 * ncol * cnt + j]);
  }
}

// ---------------------------------------------------------------- splat

// Workgroup = one work region (G x G px, aligned to its reference tile's
// origin) = (G/8)^2 pixel blocks of 8x8; wave w owns blocks w, w+4, ...
// B particles are staged per batch.
template <int G, int ADD, int B>
__global__ __launch_bounds__(256) void k_splat(SplatArgs a) {
  constexpr int NB = G / 8;
  constexpr int MAXBLK = (NB * NB + 3) / 4;
  __shared__ float s_tab[2048];
  __shared__ float s_px[B], s_py[B], s_rad[B], s_stp[B];
  __shared__ float s_a[3][B];
  __shared__ float s_q[3][B];
  __shared__ float s_maxa[B];
  __shared__ uint32_t s_bx[B], s_by[B];  // footprint box, lo | hi << 16 (tile frame)
  __shared__ uint32_t s_col[B][G];       // column y-range, lo | hi << 16
  __shared__ float s_pre2[B][G];
  __shared__ float s_pre1[B][G];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t rid = blockIdx.x;  // unit (G x G region) of job j
  const int j = find_job(a.ubase, a.nj, rid);
  const JobSplat &J = a.jobs[j];
  const uint32_t local = rid - a.ubase[j];
  const int ngm = J.ngm, T = J.T, yres = J.yres;
  const int accumulate = J.accumulate[a.pass];
  const float grayabsorb = J.grayabsorb;
  float *const pic = J.pic;
  const int per_tile = ngm * ngm;
  const int tile = static_cast<int>(local / per_tile);
  const int rem = static_cast<int>(local - static_cast<uint32_t>(tile) * per_tile);
  const int gy = rem / ngm, gx = rem - gy * ngm;
  const int cix = tile % J.ncx, ciy = tile / J.ncx;
  const int x0s = cix * T, y0s = ciy * T;
  const int w = min(x0s + T, J.xres) - x0s, h = min(y0s + T, yres) - y0s;
  const int rx0 = gx * G, ry0 = gy * G;
  if (rx0 >= w || ry0 >= h) return;
  const int rw = min(G, w - rx0), rh = min(G, h - ry0);
  const int nbx = (rw + 7) >> 3, nby = (rh + 7) >> 3, nblk = nbx * nby;

  for (int i = tid; i < 2048; i += 256) s_tab[i] = a.exptab[i];

  float acc[MAXBLK][3];
#pragma unroll
  for (int j = 0; j < MAXBLK; ++j) acc[j][0] = acc[j][1] = acc[j][2] = 0.f;

  const uint32_t e0 = a.bin_off[rid], e1 = a.bin_off[rid + 1];
  const float expfac = a.expfac;
  for (uint32_t base = e0; base < e1; base += B) {
    const int nb = static_cast<int>(min<uint32_t>(B, e1 - base));
    __syncthreads();
    if (tid < nb) {
      const Rec R = a.recs[a.ent[base + tid]];
      // splotch_host.cc:462-476 (tile frame)
      const float rfacr = R.r * a.rfac;
      const float posx = R.x - static_cast<float>(x0s), posy = R.y - static_cast<float>(y0s);
      int bx0 = max(cvt_i32(posx - rfacr + 1), 0), bx1 = min(cvt_i32(posx + rfacr + 1), w);
      int by0 = max(cvt_i32(posy - rfacr + 1), 0), by1 = min(cvt_i32(posy + rfacr + 1), h);
      bx1 = max(bx1, bx0);
      by1 = max(by1, by0);
      const float sigma = a.h2sigma * R.r;
      s_px[tid] = posx;
      s_py[tid] = posy;
      s_rad[tid] = rfacr * rfacr;
      s_stp[tid] = -1.f / (sigma * sigma);
      s_bx[tid] = static_cast<uint32_t>(bx0) | (static_cast<uint32_t>(bx1) << 16);
      s_by[tid] = static_cast<uint32_t>(by0) | (static_cast<uint32_t>(by1) << 16);
      s_a[0][tid] = -R.e0;
      s_a[1][tid] = -R.e1;
      s_a[2][tid] = -R.e2;
      if (!ADD) {
        const float g = grayabsorb;
        s_q[0][tid] = R.e0 / (R.e0 + g);
        s_q[1][tid] = R.e1 / (R.e1 + g);
        s_q[2][tid] = R.e2 / (R.e2 + g);
        s_maxa[tid] = fmaxf(fabsf(-R.e0), fmaxf(fabsf(-R.e1), fabsf(-R.e2)));
      }
    }
    __syncthreads();
    // Phase A: per particle, G column entries and G row entries.
    for (int it = tid; it < nb * 2 * G; it += 256) {
      const int k = it / (2 * G);
      const int c = it - k * (2 * G);
      const uint32_t bx = s_bx[k], by = s_by[k];
      const float posx = s_px[k], posy = s_py[k], stp = s_stp[k];
      if (c < G) {
        const int x = rx0 + c;
        uint32_t lohi = 0;
        float pre2 = 0.f;
        if (x >= static_cast<int>(bx & 0xffff) && x < static_cast<int>(bx >> 16)) {
          const float dxsq = (static_cast<float>(x) - posx) * (static_cast<float>(x) - posx);
          const float dy = sqrtf(s_rad[k] - dxsq);
          const int lo = max(static_cast<int>(by & 0xffff), cvt_i32(posy - dy + 1));
          const int hi = min(static_cast<int>(by >> 16), cvt_i32(posy + dy + 1));
          if (hi > lo) lohi = static_cast<uint32_t>(lo) | (static_cast<uint32_t>(hi) << 16);
          pre2 = xexp_op(s_tab, stp * dxsq, expfac);
        }
        s_col[k][c] = lohi;
        s_pre2[k][c] = pre2;
      } else {
        const int y = ry0 + (c - G);
        float pre1 = 0.f;
        if (y >= static_cast<int>(by & 0xffff) && y < static_cast<int>(by >> 16)) {
          const float dyf = static_cast<float>(y) - posy;
          pre1 = xexp_op(s_tab, stp * dyf * dyf, expfac);
        }
        s_pre1[k][c - G] = pre1;
      }
    }
    __syncthreads();
    // Phase B: accumulate in list order.
#pragma unroll
    for (int j = 0; j < MAXBLK; ++j) {
      const int blk = wave + 4 * j;
      if (blk >= nblk) break;
      const int bxo = rx0 + (blk % nbx) * 8, byo = ry0 + (blk / nbx) * 8;
      const int lx = bxo + (lane >> 3), ly = byo + (lane & 7);
      const int cx = lx - rx0, cy = ly - ry0;
      float l0 = acc[j][0], l1 = acc[j][1], l2 = acc[j][2];
      for (int k = 0; k < nb; ++k) {
        const uint32_t bx = s_bx[k], by = s_by[k];
        if (static_cast<int>(bx & 0xffff) >= bxo + 8 || static_cast<int>(bx >> 16) <= bxo ||
            static_cast<int>(by & 0xffff) >= byo + 8 || static_cast<int>(by >> 16) <= byo)
          continue;
        const uint32_t lohi = s_col[k][cx];
        if (ly < static_cast<int>(lohi & 0xffff) || ly >= static_cast<int>(lohi >> 16)) continue;
        const float att = s_pre1[k][cy] * s_pre2[k][cx];
        if (ADD) {
          l0 += s_a[0][k] * att;
          l1 += s_a[1][k] * att;
          l2 += s_a[2][k] * att;
        } else {
          const float a0 = s_a[0][k], a1 = s_a[1][k], a2 = s_a[2][k];
          const float q0 = s_q[0][k], q1 = s_q[1][k], q2 = s_q[2][k];
          if ((s_maxa[k] * att) < a.taylorlimit) {
            l0 += ((l0 - q0) * a0) * att;
            l1 += ((l1 - q1) * a1) * att;
            l2 += ((l2 - q2) * a2) * att;
          } else {
            l0 += xexp_expm1(s_tab, att * a0, expfac, a.taylorlimit) * (l0 - q0);
            l1 += xexp_expm1(s_tab, att * a1, expfac, a.taylorlimit) * (l1 - q1);
            l2 += xexp_expm1(s_tab, att * a2, expfac, a.taylorlimit) * (l2 - q2);
          }
        }
      }
      acc[j][0] = l0;
      acc[j][1] = l1;
      acc[j][2] = l2;
    }
  }
  // write back (x-major RGB)
#pragma unroll
  for (int j = 0; j < MAXBLK; ++j) {
    const int blk = wave + 4 * j;
    if (blk >= nblk) break;
    const int bxo = rx0 + (blk % nbx) * 8, byo = ry0 + (blk / nbx) * 8;
    const int lx = bxo + (lane >> 3), ly = byo + (lane & 7);
    if (lx < rx0 + rw && ly < ry0 + rh) {
      const int64_t X = x0s + lx, Y = y0s + ly;
      float *d = pic + (X * yres + Y) * 3;
      if (accumulate) {
        d[0] = d[0] + acc[j][0];
        d[1] = d[1] + acc[j][1];
        d[2] = d[2] + acc[j][2];
      } else {
        d[0] = acc[j][0];
        d[1] = acc[j][1];
        d[2] = acc[j][2];
      }
    }
  }
}

}  // namespace

void launch_project(const ProjectArgs *jobs, const uint32_t *blk, int nj, uint32_t nblocks,
                    Rec *rec, uint8_t *act, hipStream_t s) {
  if (nblocks == 0) return;
  k_project<<<nblocks, 256, 0, s>>>(jobs, blk, nj, rec, act);
  hip_check(hipGetLastError(), "k_project launch");
}

void launch_compact(const ProjectArgs *jobs, const uint32_t *blk, int nj, uint32_t nblocks,
                    const uint8_t *act, uint8_t code, const uint32_t *pos, const Rec *rec,
                    uint32_t *keys, uint32_t *vals, hipStream_t s) {
  if (nblocks == 0) return;
  k_compact<<<nblocks, 256, 0, s>>>(jobs, blk, nj, act, code, pos, rec, keys, vals);
  hip_check(hipGetLastError(), "k_compact launch");
}

void launch_pass_offsets(const uint32_t *pos, const uint32_t *pbase, int nj, int64_t n,
                         const uint32_t *total, uint32_t *mo, hipStream_t s) {
  k_pass_offsets<<<grid_for(nj + 1), 256, 0, s>>>(pos, pbase, nj, n, total, mo);
  hip_check(hipGetLastError(), "k_pass_offsets launch");
}

void launch_job_of_index(const uint32_t *vals, uint32_t m, const uint32_t *pbase, int nj,
                         uint32_t *keys, hipStream_t s, const uint32_t *dm) {
  k_job_of_index<<<grid_for(m), 256, 0, s>>>(vals, m, dm, pbase, nj, keys);
  hip_check(hipGetLastError(), "k_job_of_index launch");
}

void launch_gather_count(const uint32_t *order, uint32_t m, const Rec *rec, Rec *recs,
                         uint32_t *cnt, uint32_t *cnt16, int rsh, const BinGeom *geoms,
                         const uint32_t *mo, int nj, uint32_t *qbad, hipStream_t s,
                         const uint32_t *dm) {
  k_gather_count<<<grid_for(m), 256, 0, s>>>(order, m, dm, rec, recs, cnt, cnt16, rsh, geoms,
                                             mo, nj, qbad);
  hip_check(hipGetLastError(), "k_gather_count launch");
}

void launch_emit(int rsh, int rbits, const Rec *recs, uint32_t m, const uint32_t *off,
                 uint32_t e, const BinGeom *geoms, const uint32_t *ubase, const uint32_t *mo,
                 int nj, uint32_t *keys, uint32_t *vals, hipStream_t s, const uint32_t *dm,
                 const uint32_t *de) {
  if (rsh > 0)
    k_emit<2><<<grid_for(m, 256, 1 << 16), 256, 0, s>>>(recs, m, dm, off, e, de, geoms, ubase,
                                                        mo, nj, keys, vals, rsh, rbits);
  else
    k_emit<1><<<grid_for(m, 256, 1 << 16), 256, 0, s>>>(recs, m, dm, off, e, de, geoms, ubase,
                                                        mo, nj, keys, vals, 0, 0);
  hip_check(hipGetLastError(), "k_emit launch");
}

uint32_t split_chunks(uint32_t e16) {
  return (e16 + kSplitChunk - 1) / kSplitChunk;
}

void launch_split_chunks(int S, const uint32_t *keys, uint32_t e16, int rbits, uint32_t *ccnt,
                         hipStream_t s, const uint32_t *de16) {
  const uint32_t nc = split_chunks(e16);
  if (nc == 0) return;
  const unsigned g = grid_for(nc, 4, 1 << 16);
  if (S == 4) k_split_chunks<16><<<g, 256, 0, s>>>(keys, e16, de16, nc, rbits, ccnt);
  else k_split_chunks<4><<<g, 256, 0, s>>>(keys, e16, de16, nc, rbits, ccnt);
  hip_check(hipGetLastError(), "k_split_chunks launch");
}

void launch_split_bounds(int S, const uint32_t *roff, const uint32_t *keys, uint32_t nregions,
                         const uint32_t *cpre, uint32_t e16, int rbits, const uint32_t *rbase,
                         const uint32_t *ubase, const BinGeom *geoms, int nj, uint32_t *sx,
                         uint32_t *ucnt, hipStream_t s) {
  if (nregions == 0) return;
  const unsigned g = grid_for(nregions, 4, 1 << 16);
  const uint32_t nc = split_chunks(e16);
  if (S == 4)
    k_split_bounds<4><<<g, 256, 0, s>>>(roff, keys, nregions, cpre, nc, rbits, rbase, ubase,
                                         geoms, nj, sx, ucnt);
  else
    k_split_bounds<2><<<g, 256, 0, s>>>(roff, keys, nregions, cpre, nc, rbits, rbase, ubase,
                                         geoms, nj, sx, ucnt);
  hip_check(hipGetLastError(), "k_split_bounds launch");
}

void launch_split_base(int S, uint32_t nregions, const uint32_t *rbase, const uint32_t *ubase,
                       const BinGeom *geoms, int nj, const uint32_t *sx, const uint32_t *boff,
                       uint32_t *base, hipStream_t s) {
  if (nregions == 0) return;
  if (S == 4)
    k_split_base<4><<<grid_for(nregions), 256, 0, s>>>(nregions, rbase, ubase, geoms, nj, sx,
                                                        boff, base);
  else
    k_split_base<2><<<grid_for(nregions), 256, 0, s>>>(nregions, rbase, ubase, geoms, nj, sx,
                                                        boff, base);
  hip_check(hipGetLastError(), "k_split_base launch");
}

void launch_split_scatter(int S, const uint32_t *keys, const uint32_t *vals, uint32_t e16,
                          int rbits, const uint32_t *cpre, const uint32_t *base, uint32_t *ent,
                          hipStream_t s, const uint32_t *de16) {
  const uint32_t nc = split_chunks(e16);
  if (nc == 0) return;
  const unsigned g = grid_for(nc, 4, 1 << 16);
  if (S == 4)
    k_split_scatter<4><<<g, 256, 0, s>>>(keys, vals, e16, de16, nc, rbits, cpre, base, ent);
  else
    k_split_scatter<2><<<g, 256, 0, s>>>(keys, vals, e16, de16, nc, rbits, cpre, base, ent);
  hip_check(hipGetLastError(), "k_split_scatter launch");
}

void launch_bin_offsets(const uint32_t *sorted_keys, uint32_t e, uint32_t nbins,
                        uint32_t *bin_off, hipStream_t s, uint32_t kmask, const uint32_t *de) {
  k_bin_offsets<<<grid_for(static_cast<int64_t>(nbins) + 1), 256, 0, s>>>(sorted_keys, e, de,
                                                                           nbins, kmask, bin_off);
  hip_check(hipGetLastError(), "k_bin_offsets launch");
}

void launch_map(const uint32_t *idx, const uint32_t *table, uint32_t *out, int64_t n,
                hipStream_t s) {
  k_map<<<grid_for(n), 256, 0, s>>>(idx, table, out, n);
  hip_check(hipGetLastError(), "k_map launch");
}

void launch_fill(float *p, float v, int64_t n, hipStream_t s) {
  k_fill<<<grid_for(n), 256, 0, s>>>(p, v, n);
  hip_check(hipGetLastError(), "k_fill launch");
}

void launch_sum_u64(const uint32_t *v, int64_t n, uint64_t *out, hipStream_t s,
                    const uint32_t *dn) {
  k_fill_u32<<<1, 64, 0, s>>>(reinterpret_cast<uint32_t *>(out), 0u, 2);
  if (n > 0)
    k_sum_u64<<<grid_for(n, 256, 2048), 256, 0, s>>>(v, n, dn,
                                                     reinterpret_cast<unsigned long long *>(out));
  hip_check(hipGetLastError(), "k_sum_u64 launch");
}

void launch_add_images(float *pic, const float *add, int64_t n, hipStream_t s) {
  k_add_images<<<grid_for(n), 256, 0, s>>>(pic, add, n);
  hip_check(hipGetLastError(), "k_add_images launch");
}

void launch_delay(uint32_t us, hipStream_t s) {
  k_delay<<<1, 64, 0, s>>>(us * 100u);
  hip_check(hipGetLastError(), "k_delay launch");
}

void launch_fill_u32(uint32_t *p, uint32_t v, int64_t n, hipStream_t s) {
  k_fill_u32<<<grid_for(n), 256, 0, s>>>(p, v, n);
  hip_check(hipGetLastError(), "k_fill_u32 launch");
}

void launch_job_flags(const uint32_t *mo0, const uint32_t *mo1, int nj, JobSplat *js,
                      hipStream_t s) {
  k_job_flags<<<(nj + 63) / 64, 64, 0, s>>>(mo0, mo1, nj, js);
  hip_check(hipGetLastError(), "k_job_flags launch");
}

void launch_fill_empty_jobs(const uint32_t *mo0, const uint32_t *mo1, int nj, const JobSplat *js,
                            hipStream_t s) {
  if (nj == 0) return;
  k_fill_empty_jobs<<<dim3(64, nj), 256, 0, s>>>(mo0, mo1, js);
  hip_check(hipGetLastError(), "k_fill_empty_jobs launch");
}

void launch_cap_check(uint32_t *m, uint32_t *e, const uint64_t *e64, uint32_t *e16,
                      uint64_t ecap, uint32_t e16cap, uint32_t bit, uint32_t *ovf,
                      hipStream_t s) {
  k_cap_check<<<1, 64, 0, s>>>(m, e, reinterpret_cast<const unsigned long long *>(e64), e16,
                               ecap, e16cap, bit, ovf);
  hip_check(hipGetLastError(), "k_cap_check launch");
}


void launch_unpack_records(const char *stage, int64_t n, int64_t C, int ncol, float *soa,
                           int64_t nn, uint8_t *type, hipStream_t s) {
  if (n <= 0) return;
  k_unpack_records<<<grid_for(n), 256, 0, s>>>(stage, n, C, ncol, soa, nn, type);
  hip_check(hipGetLastError(), "k_unpack_records launch");
}

void launch_splat(int G, bool additive, uint32_t nbins, const SplatArgs &a, hipStream_t s) {
  if (nbins == 0) return;
  const dim3 grid(nbins), block(256);
#define SPLAT_CASE(GG, BB)                                         \
  if (G == GG) {                                                   \
    if (additive) k_splat<GG, 1, BB><<<grid, block, 0, s>>>(a);    \
    else k_splat<GG, 0, BB><<<grid, block, 0, s>>>(a);             \
    hip_check(hipGetLastError(), "k_splat launch");                \
    return;                                                        \
  }
  SPLAT_CASE(16, 64)
  SPLAT_CASE(32, 64)
  SPLAT_CASE(64, 32)
#undef SPLAT_CASE
  throw HipError{hipErrorInvalidValue, "unsupported work region size"};
}

}  // namespace splhip
