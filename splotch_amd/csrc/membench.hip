// Memory microbenchmarks for gfx950 (measurement only; not part of the render
// library): the chip's practical peaks that the bench sets kernels against,
// and known-size kernels for scaling profiler counters.
//
// LDS read bandwidth: every workgroup fills 32 KiB of LDS, then each lane
// reads 4-, 8- or 16-byte words (ds_read_b32 / b64 / b128) at conflict-free
// addresses: a wave's 64 lanes read consecutive bytes per instruction.  The
// reads are folded into an integer checksum (at most two VALU operations per
// read, below the VALU issue limit), which is written once per lane so that
// no read can be removed.  Bytes per launch = grid * 256 lanes * iters * 8
// reads * the word size.
//
// HBM copy: dst[i] = src[i] over n 16-byte vectors, grid-stride, each lane
// moving four vectors per step with all four loads issued before the stores.
// Bytes per launch = 2 * 16 * n (read + write).
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kWords = 2048;  // uint4 slots: 32 KiB

// V: 32-bit words per lane and read (1: ds_read_b32, 2: b64, 4: b128).
template <int V>
__global__ __launch_bounds__(256) void k_lds_read(uint32_t *out, int iters) {
  __shared__ uint4 buf[kWords];
  const int t = threadIdx.x;
  for (int i = t; i < kWords; i += 256)
    buf[i] = make_uint4(i, i * 3u + 1u, i * 5u + 2u, i * 7u + 3u);
  __syncthreads();
  uint32_t acc = 0;
  int base = t;
  constexpr int kSlots = kWords * 4 / V;  // V-word slots in the buffer
  for (int k = 0; k < iters; ++k) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int slot = (base + j * 256) & (kSlots - 1);
      if (V == 4) {
        const uint4 v = buf[slot];
        acc ^= v.x + v.w;
      } else if (V == 2) {
        const uint2 v = reinterpret_cast<const uint2 *>(buf)[slot];
        acc ^= v.x + v.y;
      } else {
        acc += reinterpret_cast<const uint32_t *>(buf)[slot];
      }
    }
    base = (base + 64) & (kSlots - 1);  // the next reads start a wave further on
  }
  out[static_cast<int64_t>(blockIdx.x) * 256 + t] = acc;
}

__global__ __launch_bounds__(256) void k_copy(const uint4 *__restrict__ src,
                                              uint4 *__restrict__ dst, int64_t n) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * 256 * 4;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 * 4 + threadIdx.x; i < n; i += stride) {
    uint4 v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t j = i + k * 256;
      v[k] = j < n ? src[j] : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t j = i + k * 256;
      if (j < n) dst[j] = v[k];
    }
  }
}

}  // namespace

extern "C" {

// Times one launch of `grid` workgroups x `iters` iterations of reads of
// `words` 32-bit words per lane (1, 2 or 4), after one warm-up launch, with
// HIP events on a stream of its own.  Returns 0 and writes the milliseconds
// and the bytes read from LDS; a HIP error code otherwise.
int lds_peak_run(int device, int grid, int iters, int words, float *ms, double *bytes) {
  if (grid < 1 || iters < 1 || !ms || !bytes || (words != 1 && words != 2 && words != 4))
    return -1;
  if (hipSetDevice(device) != hipSuccess) return -2;
  uint32_t *out = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  if (hipMalloc(&out, static_cast<size_t>(grid) * 256 * 4) != hipSuccess) return -2;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = -2;
  } else {
    auto launch = [&]() {
      if (words == 4) k_lds_read<4><<<grid, 256, 0, s>>>(out, iters);
      else if (words == 2) k_lds_read<2><<<grid, 256, 0, s>>>(out, iters);
      else k_lds_read<1><<<grid, 256, 0, s>>>(out, iters);
    };
    launch();
    (void)hipEventRecord(e0, s);
    launch();
    (void)hipEventRecord(e1, s);
    if (hipEventSynchronize(e1) != hipSuccess || hipGetLastError() != hipSuccess) {
      rc = -2;
    } else {
      (void)hipEventElapsedTime(ms, e0, e1);
      *bytes = static_cast<double>(grid) * 256.0 * iters * 8.0 * 4.0 * words;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  (void)hipFree(out);
  return rc;
}

// Copies `bytes` (a multiple of 16) between two device buffers it allocates:
// one warm-up launch, then `reps` launches timed with HIP events; writes the
// average milliseconds per launch.  grid = 0: 8 workgroups per CU.
int hbm_copy_run(int device, int64_t bytes, int reps, int grid, float *ms) {
  if (bytes <= 0 || (bytes & 15) || reps < 1 || !ms) return -1;
  if (hipSetDevice(device) != hipSuccess) return -2;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return -2;
  if (grid <= 0) grid = 8 * prop.multiProcessorCount;
  uint4 *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  const int64_t n = bytes / 16;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess ||
      hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = -2;
  } else {
    k_copy<<<grid, 256, 0, s>>>(a, b, n);
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < reps; ++r) k_copy<<<grid, 256, 0, s>>>(a, b, n);
    (void)hipEventRecord(e1, s);
    if (hipEventSynchronize(e1) != hipSuccess || hipGetLastError() != hipSuccess) {
      rc = -2;
    } else {
      (void)hipEventElapsedTime(ms, e0, e1);
      *ms /= reps;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

}  // extern "C"
