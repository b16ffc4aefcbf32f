/*
 * hip_splotch.h — the Splotch-side binding of the gfx950 renderer: what a
 * maintainer adds to Splotch's tree (as hip/hip_splotch.{h,cc}) to call this
 * repository's C-ABI (include/splotch_hip.h) from the render dispatch
 * splotch/splotch.cc:248-264 (and server/server.cc:158-171) under #ifdef HIP,
 * where cuda_rendering is called under #ifdef CUDA:
 *
 *   #ifdef HIP
 *     hip_rendering(0, context.pic, *pData, context.campos, context.centerpos,
 *                   context.lookat, context.sky, context.amap,
 *                   context.b_brightness, params);
 *   #else
 *     host_rendering(params, *pData, context);
 *   #endif
 *
 * It compiles against Splotch's headers (splotchutils.h, paramfile.h); it is
 * not part of this repository's library build.
 */
#ifndef HIP_SPLOTCH_H
#define HIP_SPLOTCH_H

#include <string>
#include <vector>

#include "splotch/splotchutils.h"
#include "cxxsupport/paramfile.h"
#include "splotch_hip.h"

/* Replaces cuda_rendercontext_init (cuda/cuda_splotch.h:33): binds device
   hip_device (.par key, default 0) to the rank's context, once. */
void hip_rendercontext_init(paramfile &params, render_context &context);

/* The frame's render setup from Splotch's objects, without device work: the
   .par keys as text, the render keys (splhip_params_from_par), the camera
   (splhip_camera_setup of campos / lookat / sky / centerpos and the xres,
   yres, fov, zmin, zmax, projection, minrad_pix keys) and each type's palette
   nodes (COLOURMAP = linearMap<COLOUR>, cxxsupport/linear_map.h:91-93). */
struct hip_frame_setup {
  std::string par_text;
  splhip_params prm;
  splhip_camera cam;
  std::vector<std::vector<double> > pal_x;  // per type: node positions
  std::vector<std::vector<float> > pal_rgb; // per type: RGB per node
};
std::string hip_par_text(const paramfile &params);
void hip_palette(COLOURMAP &map, std::vector<double> &x, std::vector<float> &rgb);
void hip_setup(paramfile &params, std::vector<COLOURMAP> &amap, const vec3 &campos,
               const vec3 &centerpos, const vec3 &lookat, const vec3 &sky, float b_brightness,
               hip_frame_setup &out);

/* Replaces cuda_rendering (cuda/cuda_splotch.h:35, cuda_splotch.cc:121-192):
   the raw x-major RGB image of the particles into pic (allocated xres x
   yres), as host_rendering (splotch/splotch_host.cc:711-834) leaves it; the
   particles are not modified.  Errors go through planck_fail, as the
   reference reports them. */
void hip_rendering(int mydevID, arr2<COLOUR> &pic, std::vector<particle_sim> &particle,
                   const vec3 &campos, const vec3 &centerpos, const vec3 &lookat, vec3 &sky,
                   std::vector<COLOURMAP> &amap, float b_brightness, paramfile &params);

/* Releases the rank's context (at the end of the run). */
void hip_release();

#endif
