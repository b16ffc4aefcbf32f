// Launch wrappers of the render kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "common.hpp"

namespace splhip {

struct ProjectArgs {
  const float *x, *y, *z, *hsml, *inten, *col, *c1, *c2;
  const uint8_t *type;
  int64_t n;
  splhip_camera cam;
  float smooth_fac[kMaxTypes];
  float brightness[kMaxTypes];
  int32_t col_vector[kMaxTypes];
  int32_t pass_code[kMaxTypes];  // 1: ordered pass, 2: additive pass, 0: not rendered
  int32_t pal_n[kMaxTypes];
  int32_t pal_stride;
  int32_t ntypes;
  const double *pal_x;
  const float *pal_rgb;
  float bfak, rfac;
  Rec *rec;
  uint8_t *act;
};

struct SplatArgs {
  const uint32_t *bin_off;  // [nbins+1]
  const uint32_t *ent;      // entries: slot in recs
  const Rec *recs;
  const float *exptab;      // 2048 floats
  float *pic;
  int T, xres, yres, ncx, ngm;
  int accumulate;
  float rfac, h2sigma, expfac, taylorlimit, grayabsorb;
  uint32_t prio_len;  // splat8: lists at least this long run at raised wave priority
  uint64_t *trace;    // splat8: optional per-item {start, end, len | chan << 32, hw id}
  uint32_t heavy_len; // splat8: lists at least this long go to the producer/consumer kernel
};

void launch_project(const ProjectArgs &a, hipStream_t s);
void launch_compact(const uint8_t *act, uint8_t code, const uint32_t *pos, int64_t n,
                    const Rec *rec, int sort_type, uint32_t *keys, uint32_t *vals, hipStream_t s);
void launch_gather_count(const uint32_t *order, uint32_t m, const Rec *rec, Rec *recs,
                         uint32_t *cnt, const BinGeom &g, hipStream_t s);
void launch_emit(const Rec *recs, uint32_t m, const uint32_t *off, uint32_t e, const BinGeom &g,
                 uint32_t *keys, uint32_t *vals, hipStream_t s);
void launch_bin_offsets(const uint32_t *sorted_keys, uint32_t e, uint32_t nbins,
                        uint32_t *bin_off, hipStream_t s);
void launch_map(const uint32_t *idx, const uint32_t *table, uint32_t *out, int64_t n,
                hipStream_t s);
void launch_fill(float *p, float v, int64_t n, hipStream_t s);
void launch_fill_u32(uint32_t *p, uint32_t v, int64_t n, hipStream_t s);
void launch_aos_to_soa(const splhip_particle_sim *aos, int64_t n, float *soa /*8 planes*/,
                       uint8_t *type, hipStream_t s);
// Wave-per-8x8-block splat (splat8.hip): work items (unit << 2 | channel,
// channel 3 = all), heaviest first.
void launch_items_count(const uint32_t *bin_off, uint32_t nunits, const SplatArgs &a,
                        int ordered, uint32_t split, int need_empty, uint32_t *cnt,
                        hipStream_t s);
void launch_items_emit(const uint32_t *bin_off, uint32_t nunits, const SplatArgs &a, int ordered,
                       uint32_t split, int need_empty, const uint32_t *off, uint32_t *keys,
                       uint32_t *vals, hipStream_t s);
void launch_splat8(bool additive, int variant, const uint32_t *items, uint32_t nitems,
                   const SplatArgs &a, hipStream_t s);
void launch_count_heavy(const uint32_t *sorted_keys, uint32_t nitems, uint32_t heavy,
                        uint32_t *out, hipStream_t s);
void launch_splat_heavy(bool additive, const uint32_t *items, uint32_t nheavy, const SplatArgs &a,
                        hipStream_t s);
// G: region size (16, 32 or 64 px); additive: a_eq_e mode.
void launch_splat(int G, bool additive, uint32_t nbins, const SplatArgs &a, hipStream_t s);

}  // namespace splhip
