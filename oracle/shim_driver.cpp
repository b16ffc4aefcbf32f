/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * C entry points over the Splotch-side binding (splotch_amd/shim/hip_splotch.cc)
 * called with GENUINE Splotch objects: a paramfile built from "key=value"
 * text, the colour maps loaded by the genuine get_colourmaps
 * (splotch/splotchutils.cc:60-123) from the palette{t} files, particle_sim
 * records.  Beside it, the genuine host_rendering (splotch_host.cc:711-834) on
 * the same objects.  Compiled with the reference objects by oracle/Makefile
 * (target `shim`) into oracle/_ref/libsplshim.so.
 */
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "splotch/splotch_host.h"
#include "splotch/splotchutils.h"
#include "cxxsupport/paramfile.h"
#include "hip_splotch.h"

namespace {

std::map<std::string, std::string> kv_map(const char *kv) {
  std::map<std::string, std::string> m;
  std::string s(kv ? kv : "");
  size_t pos = 0;
  while (pos < s.size()) {
    size_t nl = s.find('\n', pos);
    if (nl == std::string::npos) nl = s.size();
    const std::string line = s.substr(pos, nl - pos);
    const size_t eq = line.find('=');
    if (eq != std::string::npos) m[line.substr(0, eq)] = line.substr(eq + 1);
    pos = nl + 1;
  }
  return m;
}

vec3 v3(const double *p) { return vec3(p[0], p[1], p[2]); }

void copy_pic(arr2<COLOUR> &pic, int xres, int yres, float *out) {
  for (int ix = 0; ix < xres; ++ix)
    for (int iy = 0; iy < yres; ++iy) {
      float *d = out + (static_cast<size_t>(ix) * yres + iy) * 3;
      d[0] = pic[ix][iy].r;
      d[1] = pic[ix][iy].g;
      d[2] = pic[ix][iy].b;
    }
}

}  // namespace

extern "C" {

/* hip_setup on genuine objects: the render keys, camera and per-type palette
   nodes (pal_n[t] nodes at pal_x[t * stride], pal_rgb[3 (t * stride + i)]). */
int shim_setup(const char *kv, const double *campos, const double *lookat, const double *sky,
               const double *centerpos, float b_brightness, splhip_params *prm,
               splhip_camera *cam, int *pal_n, double *pal_x, float *pal_rgb, int stride) {
  try {
    paramfile params(kv_map(kv), false);
    std::vector<COLOURMAP> amap;
    get_colourmaps(params, amap);
    hip_frame_setup f;
    hip_setup(params, amap, v3(campos), v3(centerpos), v3(lookat), v3(sky), b_brightness, f);
    *prm = f.prm;
    *cam = f.cam;
    for (int t = 0; t < f.prm.ptypes; ++t) {
      const int n = static_cast<int>(f.pal_x[t].size());
      if (n > stride) return -2;
      pal_n[t] = n;
      for (int i = 0; i < n; ++i) {
        pal_x[t * stride + i] = f.pal_x[t][i];
        for (int c = 0; c < 3; ++c) pal_rgb[3 * (t * stride + i) + c] = f.pal_rgb[t][3 * i + c];
      }
    }
  } catch (...) {
    return -1;
  }
  return 0;
}

/* One frame of n particle_sim records (splhip_particle_sim layout) with the
   .par keys kv: through hip_rendering (genuine=0) or the genuine
   host_rendering (genuine=1); pic_out receives the x-major RGB image.  The
   records passed in are left as they are (host_rendering works on a copy). */
int shim_render(const char *kv, int64_t n, const void *records, const double *campos,
                const double *lookat, const double *sky, const double *centerpos,
                float b_brightness, int xres, int yres, int genuine, float *pic_out) {
  try {
    paramfile params(kv_map(kv), false);
    std::vector<COLOURMAP> amap;
    get_colourmaps(params, amap);
    std::vector<particle_sim> p(static_cast<size_t>(n));
    if (n > 0) std::memcpy(p.data(), records, sizeof(particle_sim) * static_cast<size_t>(n));
    if (genuine) {
      render_context rc;
      rc.campos = v3(campos);
      rc.lookat = v3(lookat);
      rc.sky = v3(sky);
      rc.centerpos = v3(centerpos);
      rc.amap = amap;
      rc.pic.alloc(xres, yres);
      rc.npart = n;
      rc.npart_all = n;
      rc.b_brightness = b_brightness;
      rc.xres = xres;
      rc.yres = yres;
      host_rendering(params, p, rc);
      copy_pic(rc.pic, xres, yres, pic_out);
    } else {
      arr2<COLOUR> pic;
      vec3 sk = v3(sky);
      hip_rendering(0, pic, p, v3(campos), v3(centerpos), v3(lookat), sk, amap, b_brightness,
                    params);
      if (static_cast<int>(pic.size1()) != xres || static_cast<int>(pic.size2()) != yres)
        return -3;
      copy_pic(pic, xres, yres, pic_out);
      // the GPU path leaves the records unmodified (cuda_rendering semantics)
      if (n > 0 && std::memcmp(p.data(), records, sizeof(particle_sim) * static_cast<size_t>(n)))
        return -4;
    }
  } catch (...) {
    return -1;
  }
  return 0;
}

void shim_release() { hip_release(); }

}  // extern "C"
