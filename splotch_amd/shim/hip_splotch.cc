// hip_splotch.cc — the Splotch-side binding of the gfx950 renderer (see
// hip_splotch.h): Splotch's paramfile, COLOURMAP, vec3 and particle_sim
// objects converted to the C-ABI's keys, camera, palettes and records.
#include "hip_splotch.h"

#include <cstddef>
#include <sstream>

#include "cxxsupport/error_handling.h"

namespace {

splhip_ctx *g_ctx = nullptr;  // one context per MPI rank (one device per rank)

// particle_sim (splotch/splotchutils.h:52-72) and splhip_particle_sim share
// their layout, so the host records go to splhip_render_host as they are
static_assert(sizeof(particle_sim) == sizeof(splhip_particle_sim), "record size");
static_assert(offsetof(particle_sim, x) == offsetof(splhip_particle_sim, x), "record x");
static_assert(offsetof(particle_sim, I) == offsetof(splhip_particle_sim, I), "record I");
static_assert(offsetof(particle_sim, type) == offsetof(splhip_particle_sim, type), "record type");
static_assert(offsetof(particle_sim, active) == offsetof(splhip_particle_sim, active),
              "record active");

}  // namespace

void hip_rendercontext_init(paramfile &params, render_context &) {
  if (g_ctx) return;
  const int dev = params.find<int>("hip_device", 0);
  if (splhip_create(dev, &g_ctx) != 0) planck_fail("splhip_create failed");
}

// The .par keys as "key=value" lines (paramfile.h getParams), the text form
// splhip_params_from_par and the other *_from_par calls read.
std::string hip_par_text(const paramfile &params) {
  std::ostringstream s;
  for (const auto &kv : params.getParams()) s << kv.first << "=" << kv.second << "\n";
  return s.str();
}

// A colour map's nodes (sorted, as getX/getY return them) as the positions
// and RGB triples splhip_set_palette takes.
void hip_palette(COLOURMAP &map, std::vector<double> &x, std::vector<float> &rgb) {
  const size_t n = map.size();
  x.resize(n);
  rgb.resize(3 * n);
  for (size_t i = 0; i < n; ++i) {
    x[i] = map.getX(i);
    const COLOUR c = map.getY(i);
    rgb[3 * i] = c.r;
    rgb[3 * i + 1] = c.g;
    rgb[3 * i + 2] = c.b;
  }
}

void hip_setup(paramfile &params, std::vector<COLOURMAP> &amap, const vec3 &campos,
               const vec3 &centerpos, const vec3 &lookat, const vec3 &sky, float b_brightness,
               hip_frame_setup &out) {
  out.par_text = hip_par_text(params);
  float sc[7];  // xres, yres, fov, zmin, zmax, projection, minrad_pix
  if (splhip_params_from_par(out.par_text.c_str(), b_brightness, &out.prm, sc) != 0)
    planck_fail("bad render keys");
  const double cp[3] = {campos.x, campos.y, campos.z};
  const double la[3] = {lookat.x, lookat.y, lookat.z};
  const double sk[3] = {sky.x, sky.y, sky.z};
  const double ce[3] = {centerpos.x, centerpos.y, centerpos.z};
  if (splhip_camera_setup(cp, la, sk, ce, static_cast<int>(sc[0]), static_cast<int>(sc[1]), sc[2],
                          sc[3], sc[4], static_cast<int>(sc[5]), sc[6], &out.cam) != 0)
    planck_fail("degenerate matrix in TRANSMAT::Inverse()");
  const int nt = out.prm.ptypes;
  out.pal_x.assign(nt, std::vector<double>());
  out.pal_rgb.assign(nt, std::vector<float>());
  for (int t = 0; t < nt && t < static_cast<int>(amap.size()); ++t)
    if (!out.prm.col_vector[t]) hip_palette(amap[t], out.pal_x[t], out.pal_rgb[t]);
}

void hip_rendering(int mydevID, arr2<COLOUR> &pic, std::vector<particle_sim> &particle,
                   const vec3 &campos, const vec3 &centerpos, const vec3 &lookat, vec3 &sky,
                   std::vector<COLOURMAP> &amap, float b_brightness, paramfile &params) {
  if (!g_ctx && splhip_create(mydevID >= 0 ? mydevID : 0, &g_ctx) != 0)
    planck_fail("splhip_create failed");
  hip_frame_setup f;
  hip_setup(params, amap, campos, centerpos, lookat, sky, b_brightness, f);
  for (int t = 0; t < f.prm.ptypes; ++t)
    if (!f.pal_x[t].empty() &&
        splhip_set_palette(g_ctx, t, static_cast<int32_t>(f.pal_x[t].size()), f.pal_x[t].data(),
                           f.pal_rgb[t].data()) != 0)
      planck_fail(splhip_last_error(g_ctx));
  pic.alloc(f.cam.xres, f.cam.yres);
  // arr2<COLOUR> is x-major RGB floats, the image layout of splhip_render_host
  if (splhip_render_host(g_ctx, reinterpret_cast<const splhip_particle_sim *>(particle.data()),
                         static_cast<int64_t>(particle.size()), &f.cam, &f.prm,
                         reinterpret_cast<float *>(&pic[0][0]), nullptr) != 0)
    planck_fail(splhip_last_error(g_ctx));
}

void hip_release() {
  splhip_destroy(g_ctx);
  g_ctx = nullptr;
}
