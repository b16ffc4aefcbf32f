// Device-wide exclusive scan and stable LSD radix sort for wave64 (gfx950).
//
// Scan: reduce-then-scan over 4096-element tiles (256 threads x 16 items).
// Each thread scans its 16 items serially, the 64 lanes of a wave combine
// their sums with a log-step shuffle scan, and the 4 wave totals are combined
// through LDS.  Tile sums are scanned recursively.
//
// Radix sort: per digit of up to 9 bits, (1) a per-tile digit histogram
// written digit-major, (2) an exclusive scan of that histogram (digit-major
// order makes the result stable), (3) a scatter: each wave ranks its keys in
// input order with a 64-lane "match any" on the digit (one ballot per bit), the
// rank among equal digits from a popcount under the lane mask, and a
// wave-private running count per digit; the tile is then staged in LDS in
// digit order and written out as contiguous per-digit runs.
#include "scan_sort.hpp"

#include <algorithm>

namespace splhip {
namespace {

constexpr int kTile = 4096;
constexpr int kItems = 16;
constexpr int kMaxBits = 9;

struct LoadU32 {
  const uint32_t *p;
  __device__ __forceinline__ uint32_t operator()(int64_t i) const { return p[i]; }
};

struct LoadU8Bit {
  const uint8_t *p;
  uint8_t v;
  __device__ __forceinline__ uint32_t operator()(int64_t i) const { return (p[i] & v) ? 1u : 0u; }
};

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

__device__ __forceinline__ uint32_t wave_reduce(uint32_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *s_w) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t inc = wave_incl_scan(v, lane);
  if (lane == 63) s_w[wave] = inc;
  __syncthreads();
  uint32_t woff = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w)
    if (w < wave) woff += s_w[w];
  __syncthreads();
  return woff + inc - v;
}

// The element count of a launch sized for `cap` elements: cap, or the device
// count *dn when it is given (at most cap).
__device__ __forceinline__ int64_t count_of(int64_t cap, const uint32_t *dn) {
  return dn ? min(cap, static_cast<int64_t>(*dn)) : cap;
}

template <class In>
__global__ __launch_bounds__(256) void k_scan_reduce(In in, int64_t cap, const uint32_t *dn,
                                                     uint32_t *partial) {
  __shared__ uint32_t s_w[4];
  const int64_t n = count_of(cap, dn);
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kTile;
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const int64_t idx = base + i * 256 + threadIdx.x;
    if (idx < n) s += in(idx);
  }
  s = wave_reduce(s);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

template <class In>
__global__ __launch_bounds__(256) void k_scan_apply(In in, uint32_t *out, int64_t cap,
                                                    const uint32_t *dn, const uint32_t *partial,
                                                    uint32_t *total) {
  __shared__ uint32_t s_w[4];
  const int64_t n = count_of(cap, dn);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kTile + threadIdx.x * kItems;
  uint32_t v[kItems];
  uint32_t sum = 0;
  if (base + kItems <= n) {
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      v[i] = in(base + i);
      sum += v[i];
    }
  } else {
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      v[i] = (base + i < n) ? in(base + i) : 0u;
      sum += v[i];
    }
  }
  const uint32_t inc = wave_incl_scan(sum, lane);
  if (lane == 63) s_w[wave] = inc;
  __syncthreads();
  uint32_t woff = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w)
    if (w < wave) woff += s_w[w];
  uint32_t prefix = (partial ? partial[blockIdx.x] : 0u) + woff + inc - sum;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    if (base + i < n) out[base + i] = prefix;
    prefix += v[i];
  }
  if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 255) *total = prefix;
}

__global__ void k_set_u32(uint32_t *p, uint32_t v) { *p = v; }

// n: the element count, or with dn the capacity the launches are sized for
// (the count itself then read from *dn on the device).
template <class In>
void scan_impl(In in, uint32_t *out, int64_t n, uint32_t *total, Workspace &ws, hipStream_t s,
               int level, const uint32_t *dn) {
  if (n <= 0) {
    if (total) k_set_u32<<<1, 1, 0, s>>>(total, 0u);
    return;
  }
  const int64_t nb = (n + kTile - 1) / kTile;
  if (nb == 1) {
    k_scan_apply<In><<<1, 256, 0, s>>>(in, out, n, dn, nullptr, total);
    return;
  }
  if (level >= 4) throw HipError{hipErrorInvalidValue, "scan too deep"};
  uint32_t *partial = ws.get_as<uint32_t>(S_SCAN0 + level, static_cast<size_t>(nb));
  k_scan_reduce<In><<<static_cast<unsigned>(nb), 256, 0, s>>>(in, n, dn, partial);
  scan_impl(LoadU32{partial}, partial, nb, nullptr, ws, s, level + 1, nullptr);
  k_scan_apply<In><<<static_cast<unsigned>(nb), 256, 0, s>>>(in, out, n, dn, partial, total);
}

template <int BITS>
__global__ __launch_bounds__(256) void k_rs_hist(const uint32_t *keys, int64_t cap,
                                                 const uint32_t *dn, int shift, uint32_t *hist,
                                                 int64_t nblocks) {
  const int64_t n = count_of(cap, dn);
  constexpr int R = 1 << BITS;
  constexpr uint32_t M = R - 1;
  __shared__ uint32_t h[4][R];
  for (int i = threadIdx.x; i < 4 * R; i += 256) (&h[0][0])[i] = 0;
  __syncthreads();
  const int wave = threadIdx.x >> 6;
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kTile;
#pragma unroll 4
  for (int i = 0; i < kItems; ++i) {
    const int64_t idx = base + i * 256 + threadIdx.x;
    if (idx < n) atomicAdd(&h[wave][(keys[idx] >> shift) & M], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < R; d += 256)
    hist[static_cast<int64_t>(d) * nblocks + blockIdx.x] = h[0][d] + h[1][d] + h[2][d] + h[3][d];
}

template <int BITS>
__global__ __launch_bounds__(256) void k_rs_scatter(const uint32_t *kin, const uint32_t *vin,
                                                    uint32_t *kout, uint32_t *vout, int64_t cap,
                                                    const uint32_t *dn, int shift,
                                                    const uint32_t *hist, int64_t nblocks) {
  const int64_t n = count_of(cap, dn);
  constexpr int R = 1 << BITS;
  constexpr uint32_t M = R - 1;
  constexpr int DPT = (R + 255) / 256;  // digits per thread in the block scan
  __shared__ uint32_t wh[4][R];         // per-wave digit counts -> offsets
  __shared__ uint32_t dstart[R];        // block-local start of each digit run
  __shared__ uint32_t gbase[R];         // global start of (digit, this tile)
  __shared__ uint32_t sk[kTile], sv[kTile];
  __shared__ uint32_t s_w[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < 4 * R; i += 256) (&wh[0][0])[i] = 0;
  for (int d = threadIdx.x; d < R; d += 256)
    gbase[d] = hist[static_cast<int64_t>(d) * nblocks + blockIdx.x];
  __syncthreads();
  const int64_t tbase = static_cast<int64_t>(blockIdx.x) * kTile;
  const int64_t base = tbase + wave * (kTile / 4);
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t kk[kItems], vv[kItems], rk[kItems];
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const int64_t idx = base + it * 64 + lane;
    const bool valid = idx < n;
    const uint32_t key = valid ? kin[idx] : 0u;
    const uint32_t d = (key >> shift) & M;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      m &= bit ? bb : ~bb;
    }
    const uint32_t rank_in = __popcll(m & lt);
    const uint32_t cnt = __popcll(m);
    // the lowest lane of each peer group reserves cnt slots; peers read its base
    const int leader = __ffsll(static_cast<long long>(m)) - 1;
    uint32_t old = 0;
    if (valid && rank_in == 0) old = atomicAdd(&wh[wave][d], cnt);
    old = __shfl(old, leader & 63, 64);
    kk[it] = key;
    vv[it] = valid ? vin[idx] : 0u;
    rk[it] = old + rank_in;
  }
  __syncthreads();
  // digit totals of the tile -> wave offsets within each digit, digit starts
  uint32_t tot[DPT];
  uint32_t tsum = 0;
#pragma unroll
  for (int j = 0; j < DPT; ++j) {
    const int d = threadIdx.x * DPT + j;
    uint32_t acc = 0;
    if (d < R) {
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        const uint32_t t = wh[w][d];
        wh[w][d] = acc;
        acc += t;
      }
    }
    tot[j] = acc;
    tsum += acc;
  }
  uint32_t start = block_excl_scan(tsum, s_w);
#pragma unroll
  for (int j = 0; j < DPT; ++j) {
    const int d = threadIdx.x * DPT + j;
    if (d < R) dstart[d] = start;
    start += tot[j];
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const int64_t idx = base + it * 64 + lane;
    if (idx < n) {
      const uint32_t d = (kk[it] >> shift) & M;
      const uint32_t lp = dstart[d] + wh[wave][d] + rk[it];
      sk[lp] = kk[it];
      sv[lp] = vv[it];
    }
  }
  __syncthreads();
  const int cnt = static_cast<int>(min<int64_t>(kTile, n - tbase));
  for (int i = threadIdx.x; i < cnt; i += 256) {
    const uint32_t key = sk[i];
    const uint32_t d = (key >> shift) & M;
    const uint32_t pos = gbase[d] + (static_cast<uint32_t>(i) - dstart[d]);
    kout[pos] = key;
    vout[pos] = sv[i];
  }
}

template <int BITS>
void rs_pass(const uint32_t *sk, const uint32_t *sv, uint32_t *dk, uint32_t *dv, int64_t n,
             const uint32_t *dn, int shift, int64_t nb, uint32_t *hist, Workspace &ws,
             hipStream_t s) {
  k_rs_hist<BITS><<<static_cast<unsigned>(nb), 256, 0, s>>>(sk, n, dn, shift, hist, nb);
  scan_u32(hist, hist, (int64_t(1) << BITS) * nb, nullptr, ws, s, 0);
  k_rs_scatter<BITS><<<static_cast<unsigned>(nb), 256, 0, s>>>(sk, sv, dk, dv, n, dn, shift,
                                                               hist, nb);
  hip_check(hipGetLastError(), "radix sort launch");
}

}  // namespace

void scan_u32(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *total, Workspace &ws,
              hipStream_t s, int level, const uint32_t *dn) {
  scan_impl(LoadU32{in}, out, n, total, ws, s, level, dn);
  hip_check(hipGetLastError(), "scan_u32 launch");
}

void scan_u8bit(const uint8_t *flags, uint8_t bits, uint32_t *out, int64_t n, uint32_t *total,
                Workspace &ws, hipStream_t s) {
  scan_impl(LoadU8Bit{flags, bits}, out, n, total, ws, s, 0, nullptr);
  hip_check(hipGetLastError(), "scan_u8bit launch");
}

int radix_sort_pairs(uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1, int64_t n,
                     int begin_bit, int end_bit, Workspace &ws, hipStream_t s,
                     const uint32_t *dn) {
  if (n <= 1 || end_bit <= begin_bit) return 0;
  const int64_t nb = (n + kTile - 1) / kTile;
  const int bits = end_bit - begin_bit;
  const int passes = (bits + kMaxBits - 1) / kMaxBits;
  const int width = (bits + passes - 1) / passes;  // <= 9
  uint32_t *hist = ws.get_as<uint32_t>(S_HIST, (size_t(1) << width) * static_cast<size_t>(nb));
  int cur = 0;
  for (int shift = begin_bit; shift < end_bit; shift += width) {
    const int w = std::min(width, end_bit - shift);
    const uint32_t *sk = cur ? k1 : k0;
    const uint32_t *sv = cur ? v1 : v0;
    uint32_t *dk = cur ? k0 : k1;
    uint32_t *dv = cur ? v0 : v1;
    switch (w) {
      case 1: rs_pass<1>(sk, sv, dk, dv, n, dn, shift, nb, hist, ws, s); break;
      case 2: rs_pass<2>(sk, sv, dk, dv, n, dn, shift, nb, hist, ws, s); break;
      case 3: rs_pass<3>(sk, sv, dk, dv, n, dn, shift, nb, hist, ws, s); break;
      case 4: rs_pass<4>(sk, sv, dk, dv, n, dn, shift, nb, hist, ws, s); break;
      case 5: rs_pass<5>(sk, sv, dk, dv, n, dn, shift, nb, hist, ws, s); break;
      case 6: rs_pass<6>(sk, sv, dk, dv, n, dn, shift, nb, hist, ws, s); break;
      case 7: rs_pass<7>(sk, sv, dk, dv, n, dn, shift, nb, hist, ws, s); break;
      case 8: rs_pass<8>(sk, sv, dk, dv, n, dn, shift, nb, hist, ws, s); break;
      default: rs_pass<9>(sk, sv, dk, dv, n, dn, shift, nb, hist, ws, s); break;
    }
    cur ^= 1;
  }
  return cur;
}

}  // namespace splhip
