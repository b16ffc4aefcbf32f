// SoA ingest of Splotch's block binary files (reader/bin_reader.cc:147-190,
// bin_reader_block): the file holds one float column per quantity
// (xxxx...yyyy...), so each wanted column is read straight from the file into
// pinned host memory and copied into its device array; the columns' byte
// order and scale are then fixed on the GPU in place.  No 36-byte
// particle_sim record is ever built.
#include "kernels.hpp"

#include <algorithm>

#include "workspace.hpp"

namespace splhip {
namespace {

// In place: byte swap (swap_endian), then v = scale * v for the smoothing
// length column (the float product smooth_factor * buffer[i]; a NaN operand
// comes out quiet with its payload, as x86 mulss returns it).  The other
// columns are copied bit for bit (1.0 * buffer[i] folds to a copy).
__global__ __launch_bounds__(256) void k_ingest(float *d, int64_t n, int swap, int scaled,
                                                float scale) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * 256) {
    uint32_t b = __float_as_uint(d[i]);
    if (swap) b = __builtin_bswap32(b);
    float o = __uint_as_float(b);
    if (scaled) o = (o != o) ? __uint_as_float(b | 0x00400000u) : scale * o;
    d[i] = o;
  }
}

__global__ __launch_bounds__(256) void k_fill_u8(uint8_t *p, uint8_t v, int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * 256)
    p[i] = v;
}

unsigned grid_of(int64_t n) {
  return static_cast<unsigned>(std::min<int64_t>(std::max<int64_t>((n + 255) / 256, 1), 8192));
}

}  // namespace

void launch_ingest_fix(float *d, int64_t n, int swap, int scaled, float scale, hipStream_t s) {
  if (n <= 0) return;
  k_ingest<<<grid_of(n), 256, 0, s>>>(d, n, swap, scaled, scale);
  hip_check(hipGetLastError(), "k_ingest launch");
}

void launch_fill_u8(uint8_t *p, uint8_t v, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  k_fill_u8<<<grid_of(n), 256, 0, s>>>(p, v, n);
  hip_check(hipGetLastError(), "k_fill_u8 launch");
}

}  // namespace splhip
