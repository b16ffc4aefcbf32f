// Launch wrappers of the render kernels (kernels.hip, splat8.hip).
//
// Every stage works on a batch of jobs (views); a plain render is a batch of
// one.  Particles of all jobs are concatenated (job j owns global indices
// [pbase[j], pbase[j+1])), a pass's particles stay job-major through
// compaction and the depth sort (job j owns pass positions [mo[j], mo[j+1])),
// and the units (tiles or 8x8 blocks) of job j are numbered from ubase[j].
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "common.hpp"

namespace splhip {

// Per-job projection inputs (device table, one entry per job).
struct ProjectArgs {
  const float *x, *y, *z, *hsml, *inten, *col, *c1, *c2;
  const uint8_t *type;
  int64_t n;
  int64_t poff;  // first global particle index of the job
  splhip_camera cam;
  float smooth_fac[kMaxTypes];
  float brightness[kMaxTypes];
  int32_t col_vector[kMaxTypes];
  int32_t pass_code[kMaxTypes];  // 1: ordered pass, 2: additive pass, 0: not rendered
  int32_t pal_n[kMaxTypes];
  int32_t pal_stride;
  int32_t ntypes;
  int32_t sort_type;
  const double *pal_x;
  const float *pal_rgb;
  float bfak, rfac;
};

// Particle-parallel launches over jobs: jobs get whole blocks of kPerBlock
// particles; blk[j] is the first block of job j (blk[nj] = grid size).
constexpr int kPerBlock = 1024;

struct SplatArgs {
  const uint32_t *bin_off;  // [nunits+1]
  const uint32_t *ent;      // entries: slot in recs
  const Rec *recs;
  const float *exptab;      // 2048 floats
  const JobSplat *jobs;     // per-job geometry, image and flags
  const uint32_t *ubase;    // [nj+1] first unit of each job
  int nj;
  int pass;                 // 0 ordered, 1 additive
  float rfac, h2sigma, expfac, taylorlimit;
  uint32_t prio_len;  // splat8: lists at least this long run at raised wave priority
  uint64_t *trace;    // splat8: optional per-item {start, end, len | chan << 32, hw id}
  uint32_t heavy_len; // splat8: lists at least this long go to the producer/consumer kernel
  uint32_t stage_len; // lists at least this long are staged (0xffffffff: none)
  // (device, optional) nonzero: the pass has a particle with a non-finite q_c,
  // so no list goes to the producer/consumer kernel (it applies every particle
  // to the whole block with zero operands, and (l - q) * 0 is NaN there)
  const uint32_t *qbad;
};

void launch_project(const ProjectArgs *jobs, const uint32_t *blk, int nj, uint32_t nblocks,
                    Rec *rec, uint8_t *act, hipStream_t s);
// Compaction of one pass (act & code) into pass order; keys (optional): the
// depth-sort key of each job's sort_type.
void launch_compact(const ProjectArgs *jobs, const uint32_t *blk, int nj, uint32_t nblocks,
                    const uint8_t *act, uint8_t code, const uint32_t *pos, const Rec *rec,
                    uint32_t *keys, uint32_t *vals, hipStream_t s);
// Device counts: a launcher given an optional `const uint32_t *d*` pointer
// sizes its grid for the host count (then a capacity) and reads the actual
// count from that device word (at most the capacity), so that no host wait
// is needed between the stage that counts and the stage that consumes.
//
// mo[j] = pos[pbase[j]] (pass particles before job j), mo[nj] = *total.
void launch_pass_offsets(const uint32_t *pos, const uint32_t *pbase, int nj, int64_t n,
                         const uint32_t *total, uint32_t *mo, hipStream_t s);
// keys[i] = job of global particle index vals[i].
void launch_job_of_index(const uint32_t *vals, uint32_t m, const uint32_t *pbase, int nj,
                         uint32_t *keys, hipStream_t s, const uint32_t *dm = nullptr);
// cnt: entries per particle (units of the binned tiles); cnt16 (optional):
// the regions of 2^rsh x 2^rsh of those 8x8 blocks they cover.  qbad
// (optional, ordered pass): set to 1 when a particle's q_c = e_c / (e_c +
// gray_absorption) is not finite.
void launch_gather_count(const uint32_t *order, uint32_t m, const Rec *rec, Rec *recs,
                         uint32_t *cnt, uint32_t *cnt16, int rsh, const BinGeom *geoms,
                         const uint32_t *mo, int nj, uint32_t *qbad, hipStream_t s,
                         const uint32_t *dm = nullptr);
// rsh == 0: (unit, particle) entries; rsh > 0: (region | unit mask << rbits,
// particle) entries for regions of 2^rsh x 2^rsh blocks, regions of job j
// numbered from ubase[j] (then ubase holds the region bases).
void launch_emit(int rsh, int rbits, const Rec *recs, uint32_t m, const uint32_t *off,
                 uint32_t e, const BinGeom *geoms, const uint32_t *ubase, const uint32_t *mo,
                 int nj, uint32_t *keys, uint32_t *vals, hipStream_t s,
                 const uint32_t *dm = nullptr, const uint32_t *de = nullptr);
// Block lists from the sorted region lists (keys: region | mask << rbits,
// vals: particles; S x S blocks per region, S = 2 or 4; roff: region CSR;
// rbase / ubase: first region / unit of each job):
//   launch_split_chunks   per-chunk mask-bit counts ccnt[S^2 * split_chunks(e16)]
//   scan_u32(ccnt -> cpre, S^2 * nchunks, total at cpre[S^2 * nchunks])
//   launch_split_bounds   sx[S^2 * nregions] and the block list lengths ucnt
//   scan_u32(ucnt -> boff, nunits, total at boff[nunits])
//   launch_split_base     base[S^2 * nregions] from boff and sx (may be sx)
//   launch_split_scatter  the block lists ent, in list order
uint32_t split_chunks(uint32_t e16);
void launch_split_chunks(int S, const uint32_t *keys, uint32_t e16, int rbits, uint32_t *ccnt,
                         hipStream_t s, const uint32_t *de16 = nullptr);
void launch_split_bounds(int S, const uint32_t *roff, const uint32_t *keys, uint32_t nregions,
                         const uint32_t *cpre, uint32_t e16, int rbits, const uint32_t *rbase,
                         const uint32_t *ubase, const BinGeom *geoms, int nj, uint32_t *sx,
                         uint32_t *ucnt, hipStream_t s);
void launch_split_base(int S, uint32_t nregions, const uint32_t *rbase, const uint32_t *ubase,
                       const BinGeom *geoms, int nj, const uint32_t *sx, const uint32_t *boff,
                       uint32_t *base, hipStream_t s);
void launch_split_scatter(int S, const uint32_t *keys, const uint32_t *vals, uint32_t e16,
                          int rbits, const uint32_t *cpre, const uint32_t *base, uint32_t *ent,
                          hipStream_t s, const uint32_t *de16 = nullptr);
// kmask: the bits of a key that hold the bin (the rest ride along unsorted).
void launch_bin_offsets(const uint32_t *sorted_keys, uint32_t e, uint32_t nbins,
                        uint32_t *bin_off, hipStream_t s, uint32_t kmask = ~0u,
                        const uint32_t *de = nullptr);
void launch_map(const uint32_t *idx, const uint32_t *table, uint32_t *out, int64_t n,
                hipStream_t s);
void launch_fill(float *p, float v, int64_t n, hipStream_t s);
void launch_fill_u32(uint32_t *p, uint32_t v, int64_t n, hipStream_t s);
// *out (device, 8 bytes) = sum of v[0..n) in 64 bits.
void launch_sum_u64(const uint32_t *v, int64_t n, uint64_t *out, hipStream_t s,
                    const uint32_t *dn = nullptr);
// Device-count pipeline helpers: per-job pass flags from the pass offsets
// (null: pass absent), zeros into the images of jobs with no particle in
// either pass, and the capacity check of a pass (counts zeroed, bit set in
// *ovf when e64 > ecap or *e16 > e16cap).
void launch_job_flags(const uint32_t *mo0, const uint32_t *mo1, int nj, JobSplat *js,
                      hipStream_t s);
void launch_fill_empty_jobs(const uint32_t *mo0, const uint32_t *mo1, int nj, const JobSplat *js,
                            hipStream_t s);
void launch_cap_check(uint32_t *m, uint32_t *e, const uint64_t *e64, uint32_t *e16,
                      uint64_t ecap, uint32_t e16cap, uint32_t bit, uint32_t *ovf,
                      hipStream_t s);
// pic[i] += add[i]
void launch_add_images(float *pic, const float *add, int64_t n, hipStream_t s);
// A one-wave kernel that returns after `us` microseconds (bounded wait).
void launch_delay(uint32_t us, hipStream_t s);
// Host records staged chunk-major (chunk k: records [k C, k C + cnt_k), its
// ncol float columns of cnt_k values back to back, then cnt_k type bytes)
// into the SoA columns soa[q nn + i] and type[i].
void launch_unpack_records(const char *stage, int64_t n, int64_t C, int ncol, float *soa,
                           int64_t nn, uint8_t *type, hipStream_t s);

// Staging thresholds: level j = 0..kStageLevels-1 of a base length sl is
// sl * 2^(j/2) (odd levels times 181/128, about sqrt 2, rounded down).
constexpr int kStageLevels = 16;
__host__ __device__ inline uint64_t stage_level(uint32_t sl, int j) {
  const uint64_t b = static_cast<uint64_t>(sl) << (j >> 1);
  return (j & 1) ? (b * 181u) >> 7 : b;
}

// Wave-per-8x8-block splat (splat8.hip): work items (unit << 2 | channel,
// channel 3 = all), heaviest first.  heavy_cnt (device, optional): incremented
// once per unit whose list reaches a.heavy_len (one item); stage_cnt /
// stage_grp (optional, [kStageLevels] each): [j] once per unit whose list
// reaches stage_level(a.stage_len, j), and by its groups of 4 entries (padded
// to whole chain loads); lmax (optional): the longest list of a unit with
// items.
void launch_items_count(const uint32_t *bin_off, uint32_t nunits, const SplatArgs &a,
                        int ordered, uint32_t split, uint32_t *cnt, uint32_t *heavy_cnt,
                        hipStream_t s, uint32_t *stage_cnt = nullptr,
                        uint32_t *stage_grp = nullptr, uint32_t *lmax = nullptr);
void launch_items_emit(const uint32_t *bin_off, uint32_t nunits, const SplatArgs &a, int ordered,
                       uint32_t split, const uint32_t *off, uint32_t *keys, uint32_t *vals,
                       hipStream_t s);
// Device item counts of a pass (all null: the host counts are exact).
struct ItemCounts {
  const uint32_t *items = nullptr;  // work items of the pass
  const uint32_t *heavy = nullptr;  // of which the long lists (sorted first)
  uint32_t heavy_cap = 0;           // the 8x8 splat skips min(*heavy, heavy_cap) items
  const uint32_t *first = nullptr;  // producer/consumer items: from *first on,
  const uint32_t *npc = nullptr;    // *npc of them (a grid sized for a capacity)
  const uint32_t *nsplit = nullptr; // long lists the 8x8 splat serves by channel
};
// With dc set, launch_splat8 is given the pass's whole item array and a grid
// for nitems (a capacity) and skips the first min(*dc.heavy, heavy_cap) items;
// launch_splat_pc is given the whole item array and npc = its capacity.  ppw
// (variant 4): particles per chunk of the walk, 1, 2, 4 or 8 (0: 8 for one
// channel, 8 additive / 4 ordered for all three).  nsplit (with dc:
// *dc.nsplit): the first nsplit of the 8x8 splat's items are long lists (one
// all-channel item each) left over by the staged and producer/consumer
// kernels; the 8x8 splat serves each as three one-channel items, so that
// three waves share its chain (a grid for nitems + 2 nsplit items).
void launch_splat8(bool additive, int variant, const uint32_t *items, uint32_t nitems,
                   const SplatArgs &a, hipStream_t s, const ItemCounts &dc = ItemCounts{},
                   int ppw = 0, uint32_t nsplit = 0);
// Staged lists (splat8.hip), the first n items of a pass (groups: their
// groups of 4 entries, padded): list offsets into cbase / cbat ([n + 1] each);
// then, for lists [i0, i1), the per-pixel operands into ops (stage_floats
// floats for all n) and, ordered, the entries' q_c into qv (stage_q_floats
// floats), and the chains, one wave per list, channel and pixel slice (8
// pixels additive, 16 ordered), which write the lists' blocks.  dn (device,
// optional): the staged lists are min(n, *dn) (n, i1 are capacities).
constexpr int kStageLoad = 8;  // a list's groups of 4 entries: a multiple of kStageLoad
constexpr int kStageQ = 4;     // groups per block of q_c (48 floats: 3 channels x 16)
uint64_t stage_floats(bool additive, uint64_t groups);
uint64_t stage_q_floats(bool additive, uint64_t groups);
void launch_stage_base(const uint32_t *items, uint32_t n, uint32_t *cbase, uint32_t *cbat,
                       const SplatArgs &a, hipStream_t s, const uint32_t *dn = nullptr);
void launch_stage_ops(bool additive, const uint32_t *items, uint32_t i0, uint32_t i1,
                      uint64_t groups, const uint32_t *cbase, const uint32_t *cbat, float4 *ops,
                      float *qv, const SplatArgs &a, hipStream_t s,
                      const uint32_t *dn = nullptr);
void launch_stage_chain(bool additive, const uint32_t *items, uint32_t i0, uint32_t i1,
                        const uint32_t *cbase, const float4 *ops, const float *qv,
                        const SplatArgs &a, hipStream_t s, const uint32_t *dn = nullptr);
// Device-count partition of a pass's long lists (one thread), as the host
// does it with exact counts (stage_rule): the staged lists stage_cnt[j] at
// the rule's level j (none past ns_cap or groups_cap), then at most heavy_max
// of the other lists from heavy_len on (the count heavy_cnt).  out[0] staged,
// out[1] producer/consumer, out[2] both, out[3] the first of the other long
// lists (one item each: from min(len0, heavy_len) entries on), which the 8x8
// splat serves by channel (split_lists).
struct PartitionArgs {
  const uint32_t *stage_cnt, *stage_grp, *heavy_cnt, *lmax;
  const uint64_t *etot;
  uint32_t len0;  // 0xffffffff: no staging
  int ordered_auto;
  uint32_t ns_cap;
  uint64_t groups_cap;
  uint32_t heavy_max;  // 0: no producer/consumer workgroups
  uint32_t heavy_len;  // lists from this length on go to producer/consumer workgroups
  uint32_t span, ratio, crit, split;
};
// The long lists after the first nheavy items that the 8x8 splat serves by
// channel: those from the greatest level of len0 with level * split <= etot
// on (level 0 when none is; none without a stage length), of the nlong long
// lists; count[j]: the lists from level j on.
__host__ __device__ inline uint32_t split_lists(uint32_t len0, const uint32_t *count,
                                                uint32_t nlong, uint32_t nheavy, uint64_t etot,
                                                uint32_t split) {
  if (len0 == 0xffffffffu) return 0u;
  int j = 0;
  while (j < kStageLevels - 1 && stage_level(len0, j + 1) * split <= etot) ++j;
  const uint32_t n = count[j] < nlong ? count[j] : nlong;
  return n > nheavy ? n - nheavy : 0u;
}
// The staging level of a pass: 0, or for an ordered pass by default (auto)
// - when lmax * ratio >= etot (bound by its longest lists): the least level
//   with level * span >= lmax,
// - else when lmax * crit >= etot (its longest lists would end after the
//   rest): the greatest level with level * crit <= etot (none when even the
//   top level is below etot / crit),
// - else none (-1).
__host__ __device__ inline int stage_rule(uint32_t len0, bool ordered_auto, uint64_t lmax,
                                          uint64_t etot, uint32_t span, uint32_t ratio,
                                          uint32_t crit) {
  if (!ordered_auto) return 0;
  int j = 0;
  if (lmax * ratio >= etot) {
    while (j < kStageLevels - 1 && stage_level(len0, j) * span < lmax) ++j;
    return j;
  }
  if (lmax * crit < etot || stage_level(len0, kStageLevels - 1) * crit < etot) return -1;
  while (j < kStageLevels - 1 && stage_level(len0, j + 1) * crit <= etot) ++j;
  return j;
}
void launch_partition(const PartitionArgs &pa, uint32_t *out, hipStream_t s);
// Producer/consumer workgroups (splat8.hip), one per item: the first npc
// items, the units whose lists reach a.heavy_len.
void launch_splat_pc(bool additive, const uint32_t *items, uint32_t npc, const SplatArgs &a,
                     hipStream_t s, const ItemCounts &dc = ItemCounts{});
// Differentiable form (adjoint.hip): screen-space gradient accumulation over
// the forward's work items (gscr: n x 6 doubles, zeroed by the caller), the
// chain to the particle fields of job 0, and an FP64 fill.
// Backward work items: list segments of at most kAdjSeg entries per unit
// (cnt[u] segments; soff = exclusive scan of cnt).  A segment is one wave's
// serial walk, so short segments shorten the backward's critical path (500k
// ordered particles at 512^2, forward + backward: 8.41 ms with 2048 entries,
// 4.80 with 512, 4.50 with 256) for 18 B of FP64 segment state per entry.
constexpr uint32_t kAdjSeg = 256;
void launch_seg_count(const uint32_t *bin_off, uint32_t nunits, uint32_t *cnt, hipStream_t s);
void launch_seg_emit(const uint32_t *bin_off, uint32_t nunits, const uint32_t *soff,
                     uint32_t *seg_unit, uint32_t *seg_first, hipStream_t s);
// part: 6 floats per list entry (the particle's screen-space partials over
// the entry's block); segstate: ordered mode only, nseg * 3 * 64 * 3 doubles.
void launch_adjoint(bool additive, const uint32_t *seg_unit, const uint32_t *seg_first,
                    uint32_t nseg, const uint32_t *soff, uint32_t nunits, const SplatArgs &a,
                    const float *grad, float *part, double *segstate, hipStream_t s);
// v[i] = i
void launch_iota(uint32_t *v, uint32_t n, hipStream_t s);
// gscr[order[s]] += sum over t in [poff[s], poff[s+1]) of part[jv[t]] (FP64,
// in that order), for the m particles of a pass.
void launch_reduce_parts(const uint32_t *poff, const uint32_t *jv, const float *part,
                         const uint32_t *order, uint32_t m, double *gscr, hipStream_t s);
void launch_chain(const ProjectArgs *jobs, int64_t n, const uint8_t *act, const double *gscr,
                  const splhip_grads &out, hipStream_t s);
void launch_zero_f64(double *p, int64_t n, hipStream_t s);
// G: region size (16, 32 or 64 px); additive: a_eq_e mode.  Single job.
void launch_splat(int G, bool additive, uint32_t nbins, const SplatArgs &a, hipStream_t s);

// Post-processing of a frame (post.hip): -expm1 when a_eq_e, colour_adjust
// when adjust, nbars colorbars of types bar_type[], then Colour8 bytes.
struct PostArgs {
  const float *pic;  // x-major RGB image
  float *out;        // post-processed float image (may be pic, or null)
  uint8_t *rgb8;     // bytes (null: none)
  int32_t xres, yres;
  int32_t a_eq_e, adjust, nbars;
  int32_t vec_in, vec_out;  // float4 image columns / 32-bit byte rows (set by launch_post)
  int32_t bar_type[kMaxTypes];
  int32_t pal_n[kMaxTypes];
  int32_t pal_stride;
  double gamma, brightness, contrast;
  const double *pal_x;
  const float *pal_rgb;
  const float *exptab;
  float expfac, taylorlimit;
};
// layout 0: bytes in the image's x-major RGB order; 1: the TGA file body
// (rows iy ascending, B G R).
void launch_post(const PostArgs &a, int layout, hipStream_t s);

// particle_normalize over SoA fields (normalize.hip): flags[t] = intensity_log,
// color_log, color_asinh, color_is_vector; ovr_set / ovr[t] = intensity_min,
// intensity_max, color_min, color_max given as parameters.
struct NormArgs {
  int64_t n;
  float *I, *er, *eg, *eb;
  const float *r;
  const uint8_t *type;
  int32_t ntypes, norm_int;
  int32_t flags[kMaxTypes][4];
  int32_t ovr_set[kMaxTypes][4];
  float ovr[kMaxTypes][4];
};
// Block-file ingest (ingest.hip): in place, byte swap when swap, then
// scale * v when scaled (else a multiply by one: only NaNs are quieted).
void launch_ingest_fix(float *d, int64_t n, int swap, int scaled, float scale, hipStream_t s);
void launch_fill_u8(uint8_t *p, uint8_t v, int64_t n, hipStream_t s);

// keys: 6 ntypes u32 of scratch; after the call they hold the per-type
// extremes as order-preserving keys (colour, intensity, size; min then max).
void launch_normalize(const NormArgs &a, uint32_t *keys, hipStream_t s);

}  // namespace splhip
