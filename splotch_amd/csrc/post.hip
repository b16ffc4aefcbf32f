// Post-processing and 8-bit output of a rendered frame on the GPU: the
// master's epilogue of splotch.cc:272-310 fused into one pass over the image.
//
//   a_eq_e:        c = -xexp.expm1(c)                     splotch.cc:275-284
//   colour_adjust: c = float(contrast * pow(double(c), gamma) + brightness),
//                  only when gamma != 1, brightness != 0 or contrast != 1
//                                                         splotchutils.cc:230-250
//   add_colorbar:  bar k (the k-th type with a scalar palette) covers rows
//                  yres-1-10k-y, y < 10, with the palette colour at
//                  ix / double(xres)                      splotchutils.cc:35-56
//   Colour8:       byte = max(0, min(255, int(c * 256))), x86 truncation
//                                                         ls_image.h:65-75
//
// The float image is read once; the bytes go out either in the image's own
// x-major order (LS_Image's pixel[ix][iy]) or as the body of the TGA file
// LS_Image::write_TGA writes (rows iy ascending, pixels ix ascending, B G R;
// ls_image.cc:178-202), which is a transpose of the image: that form is
// staged through LDS in 64 x 16 pixel tiles so both the float reads and the
// byte writes are contiguous.
#include "kernels.hpp"

#include <algorithm>

#include "workspace.hpp"

namespace splhip {
namespace {

// The value of channel c of pixel (ix, iy) after the float steps.
__device__ __forceinline__ float post_value(const PostArgs &a, float v, int ix, int iy, int c) {
  if (a.a_eq_e) v = -xexp_expm1(a.exptab, v, a.expfac, a.taylorlimit);
  if (a.adjust) {
    const double d = static_cast<double>(v);
    // pow(x, 1) is x for every x (C99 F.9.4.4), so gamma 1 skips the call
    const double pw = a.gamma == 1.0 ? d : pow(d, a.gamma);
    v = static_cast<float>(a.contrast * pw + a.brightness);
  }
  const int k = (a.yres - 1 - iy) / 10;
  if (k < a.nbars) {
    const int t = a.bar_type[k];
    float r, g, b;
    palette_val(a.pal_x + static_cast<int64_t>(t) * a.pal_stride,
                a.pal_rgb + static_cast<int64_t>(t) * a.pal_stride * 3, a.pal_n[t],
                static_cast<double>(ix) / static_cast<double>(a.xres), r, g, b);
    v = c == 0 ? r : c == 1 ? g : b;
  }
  return v;
}

__device__ __forceinline__ uint8_t colour8(float v) {
  return static_cast<uint8_t>(max(0, min(255, cvt_i32(v * 256.f))));
}

// Bytes in the image's order (or only the float image).
__global__ __launch_bounds__(256) void k_post_xmajor(PostArgs a) {
  // 3 xres yres < 2^31 (checked by the caller)
  const uint32_t yres = static_cast<uint32_t>(a.yres);
  const uint32_t n = static_cast<uint32_t>(a.xres) * yres * 3;
  for (uint32_t e = blockIdx.x * 256 + threadIdx.x; e < n; e += gridDim.x * 256) {
    const uint32_t p = e / 3;
    const int c = static_cast<int>(e - p * 3);
    const uint32_t ix = p / yres, iy = p - ix * yres;
    const float v = post_value(a, a.pic[e], static_cast<int>(ix), static_cast<int>(iy), c);
    if (a.out) a.out[e] = v;
    if (a.rgb8) a.rgb8[e] = colour8(v);
  }
}

// TGA body through LDS: block (bx, by) owns pixels ix in [64 bx, 64 bx + 64),
// iy in [16 by, 16 by + 16).
constexpr int kTX = 64, kTY = 16;
constexpr int kRow = 3 * kTY + 4;  // LDS bytes per pixel column (padded)

__global__ __launch_bounds__(256) void k_post_tga(PostArgs a) {
  __shared__ uint8_t tile[kTX * kRow];
  const int ix0 = blockIdx.x * kTX, iy0 = blockIdx.y * kTY;
  const int nx = min(kTX, a.xres - ix0), ny = min(kTY, a.yres - iy0);
  // read: per pixel column ix, the 3 ny floats of iy0.. are contiguous; full
  // columns of a 4-aligned image go as float4 (48 floats = 12 float4), all
  // of a thread's loads issued before its stores (out may be pic)
  if (ny == kTY && a.vec_in) {
    constexpr int kPer = kTX * 12 / 256;  // float4 per thread
    float4 v4[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int f = threadIdx.x + 256 * k, r = f / 12, w0 = 4 * (f - r * 12);
      if (r < nx)
        v4[k] = *reinterpret_cast<const float4 *>(
            a.pic + (static_cast<int64_t>(ix0 + r) * a.yres + iy0) * 3 + w0);
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int f = threadIdx.x + 256 * k, r = f / 12, w0 = 4 * (f - r * 12);
      if (r < nx) {
        const int ix = ix0 + r;
        float o[4] = {v4[k].x, v4[k].y, v4[k].z, v4[k].w};
        int il = w0 / 3, c = w0 - 3 * il;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          o[q] = post_value(a, o[q], ix, iy0 + il, c);
          tile[r * kRow + w0 + q] = colour8(o[q]);
          if (++c == 3) {
            c = 0;
            ++il;
          }
        }
        if (a.out)
          *reinterpret_cast<float4 *>(a.out + (static_cast<int64_t>(ix) * a.yres + iy0) * 3 +
                                      w0) = make_float4(o[0], o[1], o[2], o[3]);
      }
    }
  } else {
    for (int q = threadIdx.x; q < kTX * 3 * kTY; q += 256) {
      const int r = q / (3 * kTY), w = q - r * (3 * kTY);
      const int il = w / 3, c = w - 3 * il;
      if (r < nx && il < ny) {
        const int ix = ix0 + r, iy = iy0 + il;
        const int64_t e = (static_cast<int64_t>(ix) * a.yres + iy) * 3 + c;
        const float v = post_value(a, a.pic[e], ix, iy, c);
        if (a.out) a.out[e] = v;
        tile[r * kRow + w] = colour8(v);
      }
    }
  }
  __syncthreads();
  // write: per row iy, the 3 nx bytes of ix0.. are contiguous (B, G, R);
  // full rows of a 4-aligned output go as 32-bit words (192 bytes = 48)
  if (nx == kTX && a.vec_out) {
    for (int f = threadIdx.x; f < kTY * 48; f += 256) {
      const int il = f / 48, w0 = 4 * (f - il * 48);
      if (il < ny) {
        uint32_t word = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int w = w0 + k, r = w / 3, cb = w - 3 * r;
          word |= static_cast<uint32_t>(tile[r * kRow + il * 3 + (2 - cb)]) << (8 * k);
        }
        const int64_t o = (static_cast<int64_t>(iy0 + il) * a.xres + ix0) * 3 + w0;
        *reinterpret_cast<uint32_t *>(a.rgb8 + o) = word;
      }
    }
  } else {
    for (int q = threadIdx.x; q < kTY * 3 * kTX; q += 256) {
      const int il = q / (3 * kTX), w = q - il * (3 * kTX);
      const int r = w / 3, cb = w - 3 * r;
      if (il < ny && r < nx) {
        const int64_t o = (static_cast<int64_t>(iy0 + il) * a.xres + ix0) * 3 + w;
        a.rgb8[o] = tile[r * kRow + il * 3 + (2 - cb)];
      }
    }
  }
}

}  // namespace

void launch_post(const PostArgs &a, int layout, hipStream_t s) {
  if (layout == 1 && a.rgb8) {
    const dim3 grid((a.xres + kTX - 1) / kTX, (a.yres + kTY - 1) / kTY);
    PostArgs b = a;
    // float4 columns need every column start (ix yres + iy0) 3 on a 16-byte
    // boundary; 32-bit rows need (iy xres + ix0) 3 on a 4-byte boundary
    b.vec_in = a.yres % 4 == 0 && reinterpret_cast<uintptr_t>(a.pic) % 16 == 0 &&
               reinterpret_cast<uintptr_t>(a.out) % 16 == 0;
    b.vec_out = a.xres % 4 == 0 && reinterpret_cast<uintptr_t>(a.rgb8) % 4 == 0;
    k_post_tga<<<grid, 256, 0, s>>>(b);
    hip_check(hipGetLastError(), "k_post_tga launch");
  } else {
    const int64_t n = static_cast<int64_t>(a.xres) * a.yres * 3;
    const int64_t g = std::min<int64_t>((n + 255) / 256, 1 << 16);
    k_post_xmajor<<<static_cast<unsigned>(std::max<int64_t>(g, 1)), 256, 0, s>>>(a);
    hip_check(hipGetLastError(), "k_post_xmajor launch");
  }
}

}  // namespace splhip
