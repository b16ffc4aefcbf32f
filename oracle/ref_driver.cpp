/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * C entry points over the GENUINE Splotch host renderer, compiled from the
 * reference sources where they lie (see oracle/Makefile, target `ref`).  This
 * file is this project's own code: it builds reference objects (particle_sim,
 * paramfile, COLOURMAP, TRANSFORM) and calls the reference functions
 *   host_funct::particle_project   splotch/splotch_host.cc:58-246
 *   host_funct::particle_colorize  splotch/splotch_host.cc:252-287
 *   host_funct::particle_sort      splotch/splotch_host.cc:289-321
 *   host_funct::render_new         splotch/splotch_host.cc:330-697
 *   host_rendering                 splotch/splotch_host.cc:711-834
 *   exptable<float32>              splotch/splotchutils.h:174-215
 *   colour_adjust, add_colorbar    splotch/splotchutils.cc:35-56,230-250
 *   sceneMaker::particle_normalize splotch/scenemaker.cc:40-296
 *   bin_reader_block               reader/bin_reader.cc:147-190
 *   sceneMaker scene list          splotch/scenemaker.cc:586-673
 *   LS_Image::put_pixel/write_TGA  cxxsupport/ls_image.h:65-90,134-139, ls_image.cc:178-202
 * It is used to write the golden fixtures under tests/golden/ (in the build
 * container) and as the "reference" CPU baseline in bench.py.
 */
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#include <map>

#include "splotch/splotch_host.h"
#include "splotch/splotchutils.h"
#include "kernel/transform.h"
#include "cxxsupport/paramfile.h"
#include "cxxsupport/walltimer.h"
#include "cxxsupport/lsconstants.h"
#include "cxxsupport/ls_image.h"

#include <cstdio>
#include <fstream>
#include <iterator>
#include <sstream>
#include <iostream>
#include <algorithm>
#include <cmath>
#include <math.h>

// particle_normalize is a private member of sceneMaker.  This translation unit
// (test infrastructure only) opens the class's access so it can call the
// genuine function; the reference objects are compiled from their sources
// unchanged, and access does not change the class layout.
#define private public
#include "splotch/scenemaker.h"
#undef private
#include "reader/reader.h"

namespace host_funct {
// Definitions in splotch_host.cc (the header declares different signatures).
void particle_sort(particle_sim *p, int npart, int sort_type, bool verbose);
void render_new(particle_sim *p, int npart, arr2<COLOUR> &pic, bool a_eq_e, float32 grayabsorb,
                int tile_size);
}  // namespace host_funct

namespace {

std::map<std::string, std::string> parse_kv(const char *kv) {
  std::map<std::string, std::string> m;
  std::string s(kv ? kv : "");
  size_t pos = 0;
  while (pos < s.size()) {
    size_t nl = s.find('\n', pos);
    if (nl == std::string::npos) nl = s.size();
    std::string line = s.substr(pos, nl - pos);
    size_t eq = line.find('=');
    if (eq != std::string::npos) m[line.substr(0, eq)] = line.substr(eq + 1);
    pos = nl + 1;
  }
  return m;
}

void to_aos(int64_t n, const float *x, const float *y, const float *z, const float *r,
            const float *I, const float *e0, const float *e1, const float *e2,
            const uint8_t *type, const uint8_t *active, std::vector<particle_sim> &p) {
  p.resize(n);
  for (int64_t i = 0; i < n; ++i) {
    p[i].e = COLOUR(e0[i], e1 ? e1[i] : 0.f, e2 ? e2[i] : 0.f);
    p[i].x = x[i]; p[i].y = y[i]; p[i].z = z[i]; p[i].r = r[i]; p[i].I = I[i];
    p[i].type = type ? type[i] : 0;
    p[i].active = active ? (active[i] != 0) : true;
  }
}

void from_aos(const std::vector<particle_sim> &p, float *x, float *y, float *z, float *r,
              float *I, float *e0, float *e1, float *e2, uint8_t *type, uint8_t *active) {
  for (size_t i = 0; i < p.size(); ++i) {
    x[i] = p[i].x; y[i] = p[i].y; z[i] = p[i].z; r[i] = p[i].r; I[i] = p[i].I;
    e0[i] = p[i].e.r; e1[i] = p[i].e.g; e2[i] = p[i].e.b;
    if (type) type[i] = (uint8_t)p[i].type;
    if (active) active[i] = p[i].active ? 1 : 0;
  }
}

void build_maps(int ntypes, const int *pal_n, int pal_stride, const double *pal_x,
                const float *pal_rgb, std::vector<COLOURMAP> &amap) {
  amap.resize(ntypes);
  for (int t = 0; t < ntypes; ++t) {
    for (int i = 0; i < pal_n[t]; ++i) {
      const float *c = pal_rgb + 3 * ((int64_t)t * pal_stride + i);
      amap[t].addVal(pal_x[(int64_t)t * pal_stride + i], COLOUR(c[0], c[1], c[2]));
    }
    amap[t].sortMap();
  }
}

}  // namespace

extern "C" {

/* Genuine exptable<float32>(-20): op or expm1 over an argument array. */
void ref_exp_op(int64_t n, const float *arg, float *out, int expm1) {
  exptable<float32> xexp(-20.);
  for (int64_t i = 0; i < n; ++i) out[i] = expm1 ? xexp.expm1(arg[i]) : xexp(arg[i]);
}

float ref_taylor_limit() {
  exptable<float32> xexp(-20.);
  return xexp.taylorLimit();
}

/* The camera matrix sequence of splotch_host.cc:83-104, using the genuine
   vec3 / TRANSMAT / TRANSFORM code.  out12 = trans.Matrix().p, out_d = {xshift,
   yshift}, out_f = {res2, ycorr, xfac, fovfct, dist}. */
void ref_camera(const double *campos, const double *lookat, const double *sky_in,
                const double *centerpos, int xres, int yres, float fov, float *out12,
                double *out_d, float *out_f) {
  vec3 cam(campos[0], campos[1], campos[2]), look(lookat[0], lookat[1], lookat[2]);
  vec3 sky(sky_in[0], sky_in[1], sky_in[2]), cen(centerpos[0], centerpos[1], centerpos[2]);
  float32 ycorr = .5f * (yres - xres);
  float32 res2 = 0.5f * xres;
  float32 fovfct = tan(fov * 0.5f * degr2rad);
  sky.Normalize();
  vec3 zaxis = (cen - look).Norm();
  vec3 xaxis = crossprod(sky, zaxis).Norm();
  vec3 yaxis = crossprod(zaxis, xaxis);
  TRANSFORM trans;
  trans.Make_General_Transform(TRANSMAT(xaxis.x, xaxis.y, xaxis.z, yaxis.x, yaxis.y, yaxis.z,
                                        zaxis.x, zaxis.y, zaxis.z, 0, 0, 0));
  trans.Invert();
  TRANSFORM trans2;
  trans2.Make_Translation_Transform(-cam);
  trans2.Add_Transform(trans);
  trans = trans2;
  vec3 tcpos = trans.TransPoint(cen);
  float32 dist = (cen - look).Length();
  float32 xfac = res2 / (fovfct * dist);
  std::memcpy(out12, trans.Matrix().p, 12 * sizeof(float));
  out_d[0] = -xfac * tcpos.x;
  out_d[1] = -xfac * tcpos.y;
  out_f[0] = res2; out_f[1] = ycorr; out_f[2] = xfac; out_f[3] = fovfct; out_f[4] = dist;
}

/* particle_project + particle_colorize on SoA arrays (updated in place).
   kv: "key=value\n" lines for the paramfile. */
int ref_project_colorize(const char *kv, int64_t n, float *x, float *y, float *z, float *r,
                         float *I, float *e0, float *e1, float *e2, uint8_t *type,
                         uint8_t *active, const double *campos, const double *lookat,
                         const double *sky, const double *centerpos, int ntypes,
                         const int *pal_n, int pal_stride, const double *pal_x,
                         const float *pal_rgb, float b_brightness, int do_colorize) {
  try {
    paramfile params(parse_kv(kv), false);
    std::vector<particle_sim> p;
    to_aos(n, x, y, z, r, I, e0, e1, e2, type, nullptr, p);
    vec3 cam(campos[0], campos[1], campos[2]), look(lookat[0], lookat[1], lookat[2]);
    vec3 sk(sky[0], sky[1], sky[2]), cen(centerpos[0], centerpos[1], centerpos[2]);
    host_funct::particle_project(params, p, cam, look, sk, cen);
    if (do_colorize) {
      std::vector<COLOURMAP> amap;
      build_maps(ntypes, pal_n, pal_stride, pal_x, pal_rgb, amap);
      host_funct::particle_colorize(params, p, amap, b_brightness);
    }
    from_aos(p, x, y, z, r, I, e0, e1, e2, type, active);
  } catch (...) {
    return -1;
  }
  return 0;
}

/* particle_sort over SoA arrays.  The original index of each record is
   carried in the bits of its I field (no comparator reads I) and returned in
   perm. */
int ref_sort(int64_t n, const float *z, const float *er, const float *r, int sort_type,
             uint32_t *perm) {
  try {
    std::vector<particle_sim> p(n);
    for (int64_t i = 0; i < n; ++i) {
      p[i].e = COLOUR(er[i], 0.f, 0.f);
      p[i].x = 0.f; p[i].y = 0.f; p[i].z = z[i]; p[i].r = r[i];
      uint32_t u = (uint32_t)i;
      std::memcpy(&p[i].I, &u, 4);
      p[i].type = 0; p[i].active = true;
    }
    host_funct::particle_sort(p.data(), (int)n, sort_type, false);
    for (int64_t i = 0; i < n; ++i) std::memcpy(&perm[i], &p[i].I, 4);
  } catch (...) {
    return -1;
  }
  return 0;
}

/* render_new on projected + coloured SoA particles; pic x-major RGB. */
int ref_render_new(int64_t n, const float *x, const float *y, const float *z, const float *r,
                   const float *I, const float *e0, const float *e1, const float *e2,
                   const uint8_t *active, int xres, int yres, int a_eq_e, float grayabsorb,
                   int tile, float *pic) {
  try {
    std::vector<particle_sim> p;
    to_aos(n, x, y, z, r, I, e0, e1, e2, nullptr, active, p);
    arr2<COLOUR> img(xres, yres);
    host_funct::render_new(p.data(), (int)n, img, a_eq_e != 0, grayabsorb, tile);
    for (int ix = 0; ix < xres; ++ix)
      for (int iy = 0; iy < yres; ++iy) {
        float *d = &pic[(size_t(ix) * yres + iy) * 3];
        d[0] = img[ix][iy].r; d[1] = img[ix][iy].g; d[2] = img[ix][iy].b;
      }
  } catch (...) {
    return -1;
  }
  return 0;
}

/* The whole host_rendering() path (project, colorize, [sort], render_new).
   Particles are copied (host_rendering mutates its vector).  Returns 0/-1. */
int ref_host_rendering(const char *kv, int64_t n, const float *x, const float *y,
                       const float *z, const float *r, const float *I, const float *e0,
                       const float *e1, const float *e2, const uint8_t *type,
                       const double *campos, const double *lookat, const double *sky,
                       const double *centerpos, int ntypes, const int *pal_n, int pal_stride,
                       const double *pal_x, const float *pal_rgb, float b_brightness,
                       int xres, int yres, float *pic) {
  try {
    paramfile params(parse_kv(kv), false);
    render_context rc;
    rc.campos = vec3(campos[0], campos[1], campos[2]);
    rc.lookat = vec3(lookat[0], lookat[1], lookat[2]);
    rc.sky = vec3(sky[0], sky[1], sky[2]);
    rc.centerpos = vec3(centerpos[0], centerpos[1], centerpos[2]);
    build_maps(ntypes, pal_n, pal_stride, pal_x, pal_rgb, rc.amap);
    rc.pic.alloc(xres, yres);
    rc.npart = n;
    rc.npart_all = n;
    rc.b_brightness = b_brightness;
    rc.xres = xres;
    rc.yres = yres;
    std::vector<particle_sim> p;
    to_aos(n, x, y, z, r, I, e0, e1, e2, type, nullptr, p);
    host_rendering(params, p, rc);
    for (int ix = 0; ix < xres; ++ix)
      for (int iy = 0; iy < yres; ++iy) {
        float *d = &pic[(size_t(ix) * yres + iy) * 3];
        d[0] = rc.pic[ix][iy].r; d[1] = rc.pic[ix][iy].g; d[2] = rc.pic[ix][iy].b;
      }
  } catch (...) {
    return -1;
  }
  return 0;
}

/* The master's post-processing of splotch.cc:272-284 and the image output of
   splotch.cc:298-310, over a raw x-major RGB image (updated in place): for
   a_eq_e (the .par key) -xexp.expm1 per channel (the three lines of
   splotch.cc:275-284, restated here over the genuine exptable since splotch.cc
   holds main()), then the genuine colour_adjust and, with the colorbar key,
   add_colorbar; then the genuine LS_Image::put_pixel quantization and
   write_TGA to tga_path.  tga_out (18 + 3 xres yres bytes) receives the file.
   Returns 0/-1. */
int ref_postprocess(const char *kv, int xres, int yres, float *pic, int ntypes,
                    const int *pal_n, int pal_stride, const double *pal_x,
                    const float *pal_rgb, const char *tga_path, uint8_t *tga_out) {
  try {
    paramfile params(parse_kv(kv), false);
    render_context rc;
    build_maps(ntypes, pal_n, pal_stride, pal_x, pal_rgb, rc.amap);
    rc.pic.alloc(xres, yres);
    rc.xres = xres;
    rc.yres = yres;
    for (int ix = 0; ix < xres; ++ix)
      for (int iy = 0; iy < yres; ++iy) {
        const float *d = &pic[(size_t(ix) * yres + iy) * 3];
        rc.pic[ix][iy] = COLOUR(d[0], d[1], d[2]);
      }
    if (params.find<bool>("a_eq_e", true)) {
      exptable<float32> xexp(-20.0);
      for (int ix = 0; ix < xres; ++ix)
        for (int iy = 0; iy < yres; ++iy) {
          COLOUR &c = rc.pic[ix][iy];
          c.r = -xexp.expm1(c.r);
          c.g = -xexp.expm1(c.g);
          c.b = -xexp.expm1(c.b);
        }
    }
    colour_adjust(params, rc);
    if (params.find<bool>("colorbar", false)) add_colorbar(params, rc.pic, rc.amap);
    LS_Image img(xres, yres);
    for (int ix = 0; ix < xres; ++ix)
      for (int iy = 0; iy < yres; ++iy) {
        const COLOUR &c = rc.pic[ix][iy];
        img.put_pixel(ix, iy, Colour(c.r, c.g, c.b));
        float *d = &pic[(size_t(ix) * yres + iy) * 3];
        d[0] = c.r; d[1] = c.g; d[2] = c.b;
      }
    img.write_TGA(tga_path);
    std::ifstream in(tga_path, std::ios::binary);
    std::vector<char> bytes((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    if (bytes.size() != 18 + size_t(3) * xres * yres) return -1;
    std::memcpy(tga_out, bytes.data(), bytes.size());
    std::remove(tga_path);
  } catch (...) {
    return -1;
  }
  return 0;
}

/* The scene list of the genuine sceneMaker constructor for the .par keys kv
   (scene_file, scene_start ...), with each scene's camera taken as
   getNextScene takes it (scenemaker.cc:1140-1166: the scene's keys applied to
   the parameters, which keep them for later scenes).  Per scene, out[16 i ..]
   = campos, lookat, sky, center, fidx, keep_particles.  *n = the number of
   scenes (at most cap are written).  Returns 0/-1. */
int ref_scene_file(const char *kv, double *out, int cap, int *n) {
  try {
    auto m = parse_kv(kv);
    if (!m.count("outfile")) m["outfile"] = "scene_probe";
    paramfile params(m, false);
    sceneMaker sm(params);
    *n = (int)sm.scenes.size();
    for (int i = 0; i < *n && i < cap; ++i) {
      const auto sp = sm.scenes[i].sceneParameters.getParams();
      for (auto it = sp.begin(); it != sp.end(); ++it) params.setParam(it->first, it->second);
      double *o = out + 16 * i;
      o[0] = params.find<double>("camera_x"); o[1] = params.find<double>("camera_y");
      o[2] = params.find<double>("camera_z");
      o[3] = params.find<double>("lookat_x"); o[4] = params.find<double>("lookat_y");
      o[5] = params.find<double>("lookat_z");
      o[6] = params.find<double>("sky_x", 0); o[7] = params.find<double>("sky_y", 0);
      o[8] = params.find<double>("sky_z", 1);
      if (params.param_present("center_x")) {
        o[9] = params.find<double>("center_x"); o[10] = params.find<double>("center_y");
        o[11] = params.find<double>("center_z");
      } else {
        o[9] = o[0]; o[10] = o[1]; o[11] = o[2];
      }
      o[12] = params.find<double>("fidx", 0);
      o[13] = sm.scenes[i].keep_particles ? 1 : 0;
    }
  } catch (...) {
    return -1;
  }
  return 0;
}

/* The genuine bin_reader_block over the .par keys kv (infile, num_blocks,
   x ... C3, swap_endian, smooth_factor): the particles it returns, as SoA
   (n_cap entries at most; *n = the count).  Returns 0/-1. */
int ref_bin_block(const char *kv, int64_t n_cap, float *x, float *y, float *z, float *r, float *I,
                  float *er, float *eg, float *eb, uint8_t *type, int64_t *n) {
  try {
    paramfile params(parse_kv(kv), false);
    std::vector<particle_sim> p;
    bin_reader_block(params, p);
    if ((int64_t)p.size() > n_cap) return -1;
    *n = (int64_t)p.size();
    for (size_t i = 0; i < p.size(); ++i) {
      x[i] = p[i].x; y[i] = p[i].y; z[i] = p[i].z; r[i] = p[i].r; I[i] = p[i].I;
      er[i] = p[i].e.r; eg[i] = p[i].e.g; eb[i] = p[i].e.b;
      type[i] = (uint8_t)p[i].type;
    }
  } catch (...) {
    return -1;
  }
  return 0;
}

/* The genuine sceneMaker::particle_normalize (host branch: logs, asinh,
   per-type min/max, clamping) over n particles given as SoA (updated in
   place: I, er, eg, eb; r is read).  kv: .par keys (ptypes, intensity_log{t},
   color_log{t}, color_asinh{t}, color_is_vector{t}, intensity_min{t} ...).
   ranges[t*4 + k] = the parameters the function sets: range_min{t},
   range_max{t}, clamp_min{t}, clamp_max{t}.  Returns 0/-1. */
int ref_normalize(const char *kv, int64_t n, float *I, const float *r, float *er, float *eg,
                  float *eb, const uint8_t *type, int ntypes, double *ranges) {
  try {
    auto m = parse_kv(kv);
    if (!m.count("outfile")) m["outfile"] = "normalize_probe";
    paramfile params(m, false);
    sceneMaker sm(params);
    std::vector<particle_sim> p(n);
    for (int64_t i = 0; i < n; ++i) {
      p[i].e = COLOUR(er[i], eg[i], eb[i]);
      p[i].x = p[i].y = p[i].z = 0.f;
      p[i].r = r[i];
      p[i].I = I[i];
      p[i].type = type ? type[i] : 0;
      p[i].active = true;
    }
    sm.particle_normalize(p, false);
    for (int64_t i = 0; i < n; ++i) {
      I[i] = p[i].I; er[i] = p[i].e.r; eg[i] = p[i].e.g; eb[i] = p[i].e.b;
    }
    for (int t = 0; t < ntypes; ++t) {
      const std::string ts = dataToString(t);
      ranges[4 * t + 0] = sm.params.find<double>("range_min" + ts);
      ranges[4 * t + 1] = sm.params.find<double>("range_max" + ts);
      ranges[4 * t + 2] = sm.params.find<double>("clamp_min" + ts);
      ranges[4 * t + 3] = sm.params.find<double>("clamp_max" + ts);
    }
  } catch (...) {
    return -1;
  }
  return 0;
}

}  // extern "C"
