// C-ABI of the gfx950 Splotch renderer: context, host-side camera and .par
// setup, and the frame pipeline (project -> [depth sort] -> bin -> splat).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "scan_sort.hpp"
#include "workspace.hpp"

using namespace splhip;

struct FwdState;  // the last single-view render, kept for splhip_render_backward
struct Reservation;  // capacities and workspace of a batch rendered under stream capture

// Scratch of splhip_normalize_noalloc: the per-type extremes of
// particle_normalize (6 order-preserving u32 keys per type), in a file-scope
// device array so that the call allocates nothing.  Each GPU holds its own
// copy; its address comes from hipGetSymbolAddress at context creation.
constexpr int kNoAllocWords = 64;
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

struct splhip_ctx {
  FwdState *fwd = nullptr;
  uint32_t *scratch_ptr = nullptr;  // this device's copy of `scratch`
  Reservation *rv = nullptr;
  bool split_hint = false;  // the last capacity error can be avoided by smaller groups
  // the 8x8 splat decomposition found by the last inner autotuner sweep:
  // variant, ppw, split, heavy, heavy_max
  uint32_t inner_choice[5] = {4, 0, 32768, 32768, 192};
  int device = 0;
  std::string err;
  Workspace ws;
  float *d_exptab = nullptr;
  Consts consts{};
  std::vector<double> pal_x[kMaxTypes];
  std::vector<float> pal_rgb[kMaxTypes];
  double *d_palx = nullptr;
  float *d_palrgb = nullptr;
  int pal_stride = 0;
  size_t pal_cap = 0;
  bool pal_dirty = true;
  uint32_t *h_tot = nullptr;  // pinned, device totals read back
  char *h_stage = nullptr;    // pinned, table uploads and offset read-backs
  size_t h_stage_cap = 0;
  hipEvent_t ev_stage = nullptr;  // last copy that used h_stage
  // stage timing (splhip_set_profiling): begin/end event pairs; a stage may
  // consist of several intervals of the stream
  int prof = 0;
  struct Mark {
    int stage;
    hipEvent_t b, e;
  };
  std::vector<Mark> marks;
  std::vector<hipEvent_t> ev_pool;
  int nmarks = 0;
  bool used[SPLHIP_NSTAGES] = {};
  float stage_ms[SPLHIP_NSTAGES] = {};
  int64_t last_entries[2] = {};
  int64_t last_active[2] = {};
  // work-item trace of the 8x8 splat (splhip_set_trace), per pass
  int trace = 0;
  std::vector<uint64_t> trace_items[2];
  std::vector<uint32_t> h_mo;
  // autotuner: the choice per batch bucket (opaque, see render_tuned)
  std::map<uint64_t, std::vector<uint32_t>> tuned;
  // variant, split, heavy cap (0: none), tuned, work region, ppw
  uint32_t last_tuning[6] = {0, 0, 0, 0, 0, 0};
  // staged lists and their groups of the last render's passes (capture_reserve)
  uint32_t last_stage_n[2] = {0, 0};
  uint64_t last_stage_groups[2] = {0, 0};
  int64_t last_regions[2] = {0, 0};  // region entries binned per pass (0: per block)
  // auxiliary stream (default priority) for the heavy-block kernel
  // auxiliary streams of a pass's splat (set 1: the additive pass's when it
  // runs beside the ordered pass's): producer/consumer workgroups (aux), the
  // staged lists (aux3: the longest lists' operands and chains; aux4: the
  // other lists'), and their fork and join events
  struct AuxSet {
    hipStream_t aux = nullptr, aux3 = nullptr, aux4 = nullptr;
    hipEvent_t fork = nullptr, join = nullptr, fork3 = nullptr, join3 = nullptr;
    hipEvent_t fork4 = nullptr, join4 = nullptr;
  };
  AuxSet ax[2];
  // second stream for the additive pass's splat, which runs beside the
  // ordered pass's, with its own temporaries
  hipStream_t aux2 = nullptr;
  hipEvent_t ev_fork2 = nullptr, ev_join2 = nullptr;
  // recorded by the ordered splat once its long-list workgroups have been
  // launched and given their head start (see stage_splat)
  hipEvent_t ev_placed = nullptr;
  Workspace ws2;
  // block-file ingest: two pinned buffers, each freed for reuse by its event
  char *h_ing[2] = {nullptr, nullptr};
  hipEvent_t ev_ing[2] = {nullptr, nullptr};
  // host-record upload: per worker thread two pinned buffers and their events
  std::vector<char *> h_rec;
  std::vector<hipEvent_t> ev_rec;
};

namespace {

// ------------------------------------------------------------------ constants

// splotch_host.cc:44-49 and exptable<float32>(-20) (splotchutils.h:189-199),
// computed on the host with the C library, as the reference does.
void host_exptable(float *tab, Consts &c) {
  const float maxexp = -20.f;
  const float expfac = 1048576 / maxexp;
  for (int m = 0; m < 1024; ++m) {
    tab[m] = std::exp(m * 1024 / expfac);
    tab[1024 + m] = std::exp(m / expfac);
  }
  c.expfac = expfac;
  c.taylorlimit = std::sqrt(float(2) * std::abs(maxexp) / 1048576);
  const double pi = 3.141592653589793238462643383279502884197;
  c.h2sigma = static_cast<float>(0.5 * std::pow(pi, -1. / 6.));
  c.bfak = static_cast<float>(0.5 * std::pow(pi, -5. / 6.));
  c.rfac = 0.75f;
}

int fail(splhip_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

int hip_fail(splhip_ctx *c, const HipError &e) {
  std::string m = std::string(e.what) + ": " + hipGetErrorString(e.code);
  return fail(c, SPLHIP_E_HIP, m);
}

// ------------------------------------------------------------------ camera (a1)

struct V3 {
  double x, y, z;
};
V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 vnorm(V3 a) {  // vec3.h:72-77
  const double l = 1.0 / std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
  return {a.x * l, a.y * l, a.z * l};
}
double vlen(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
V3 vcross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

struct M34 {
  float e[3][4];
};

// TRANSMAT::Inverse (kernel/transform.cxx:64-112)
bool m_inverse(const M34 &a, M34 &t) {
  t.e[0][0] = a.e[1][1] * a.e[2][2] - a.e[2][1] * a.e[1][2];
  t.e[0][1] = a.e[0][1] * a.e[2][2] - a.e[2][1] * a.e[0][2];
  t.e[0][2] = a.e[0][1] * a.e[1][2] - a.e[1][1] * a.e[0][2];
  t.e[1][0] = a.e[1][0] * a.e[2][2] - a.e[2][0] * a.e[1][2];
  t.e[1][1] = a.e[0][0] * a.e[2][2] - a.e[2][0] * a.e[0][2];
  t.e[1][2] = a.e[0][0] * a.e[1][2] - a.e[1][0] * a.e[0][2];
  t.e[2][0] = a.e[1][0] * a.e[2][1] - a.e[2][0] * a.e[1][1];
  t.e[2][1] = a.e[0][0] * a.e[2][1] - a.e[2][0] * a.e[0][1];
  t.e[2][2] = a.e[0][0] * a.e[1][1] - a.e[1][0] * a.e[0][1];
  double d = (a.e[0][0] * t.e[0][0] - a.e[1][0] * t.e[0][1] + a.e[2][0] * t.e[0][2]);
  if (!(std::abs(d) > 1e-10)) return false;
  d = 1. / d;
  t.e[0][0] *= d;
  t.e[2][0] *= d;
  t.e[1][1] *= d;
  t.e[0][2] *= d;
  t.e[2][2] *= d;
  d = -d;
  t.e[1][0] *= d;
  t.e[0][1] *= d;
  t.e[2][1] *= d;
  t.e[1][2] *= d;
  for (int j = 0; j < 3; ++j)
    t.e[j][3] = -(t.e[j][0] * a.e[0][3] + t.e[j][1] * a.e[1][3] + t.e[j][2] * a.e[2][3]);
  return true;
}

// TRANSMAT::operator*= (kernel/transform.cxx:50-62)
void m_mul_assign(M34 &self, const M34 &b) {
  const M34 a = self;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j)
      self.e[j][i] = a.e[0][i] * b.e[j][0] + a.e[1][i] * b.e[j][1] + a.e[2][i] * b.e[j][2];
  self.e[0][3] += b.e[0][3];
  self.e[1][3] += b.e[1][3];
  self.e[2][3] += b.e[2][3];
}

// ------------------------------------------------------------------ .par parsing

std::string trim(const std::string &s) {
  size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

// paramfile text format (cxxsupport/string_utils.cc:172-210)
std::map<std::string, std::string> parse_par(const char *text) {
  std::map<std::string, std::string> d;
  std::istringstream in(text ? text : "");
  std::string line;
  while (std::getline(in, line)) {
    line = line.substr(0, line.find('\r'));
    line = line.substr(0, line.find('#'));
    line = trim(line);
    if (line.empty()) continue;
    size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    std::string k = trim(line.substr(0, eq)), v = trim(line.substr(eq + 1));
    if (!k.empty()) d[k] = v;
  }
  return d;
}

struct ParError {
  std::string msg;
};

template <class T>
T par_num(const std::map<std::string, std::string> &d, const std::string &k, T deflt) {
  auto it = d.find(k);
  if (it == d.end()) return deflt;
  std::istringstream s(it->second);
  T v;
  s >> v;
  std::string rest;
  if (!s || (s >> rest, !rest.empty())) throw ParError{"conversion error for key " + k};
  return v;
}

// stringToData<bool> (cxxsupport/string_utils.cc:131-140)
bool par_bool(const std::map<std::string, std::string> &d, const std::string &k, bool deflt) {
  auto it = d.find(k);
  if (it == d.end()) return deflt;
  std::string v = it->second;
  for (auto &ch : v) ch = static_cast<char>(std::tolower(ch));
  if (v == "f" || v == "n" || v == "false" || v == ".false.") return false;
  if (v == "t" || v == "y" || v == "true" || v == ".true.") return true;
  throw ParError{"conversion error in bool key " + k};
}

// ------------------------------------------------------------------ pipeline

// Stage markers (HIP events on the render stream) when profiling is enabled.
void mark(splhip_ctx *c, int stage, bool begin, hipStream_t s) {
  if (!c->prof) return;
  const size_t need = 2 * static_cast<size_t>(c->nmarks + 1);
  while (c->ev_pool.size() < need) {
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    c->ev_pool.push_back(e);
  }
  if (begin) {
    if (c->marks.size() < static_cast<size_t>(c->nmarks + 1)) c->marks.resize(c->nmarks + 1);
    splhip_ctx::Mark &m = c->marks[c->nmarks];
    m.stage = stage;
    m.b = c->ev_pool[2 * c->nmarks];
    m.e = c->ev_pool[2 * c->nmarks + 1];
    hip_check(hipEventRecord(m.b, s), "hipEventRecord");
    ++c->nmarks;
  } else {
    for (int k = c->nmarks - 1; k >= 0; --k)
      if (c->marks[k].stage == stage) {
        hip_check(hipEventRecord(c->marks[k].e, s), "hipEventRecord");
        break;
      }
    c->used[stage] = true;
  }
}

void collect_marks(splhip_ctx *c) {
  if (!c->prof) return;
  for (int k = 0; k < SPLHIP_NSTAGES; ++k) c->stage_ms[k] = 0.f;
  for (int k = 0; k < c->nmarks; ++k) {
    const splhip_ctx::Mark &m = c->marks[k];
    float ms = 0.f;
    hip_check(hipEventSynchronize(m.e), "hipEventSynchronize");
    hip_check(hipEventElapsedTime(&ms, m.b, m.e), "hipEventElapsedTime");
    c->stage_ms[m.stage] += ms;
  }
}

void ensure_aux(splhip_ctx *c, int k = 0) {
  splhip_ctx::AuxSet &X = c->ax[k];
  if (X.aux) return;
  hip_check(hipStreamCreateWithFlags(&X.aux, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipEventCreateWithFlags(&X.fork, hipEventDisableTiming), "hipEventCreate");
  hip_check(hipEventCreateWithFlags(&X.join, hipEventDisableTiming), "hipEventCreate");
}

void ensure_aux2(splhip_ctx *c) {
  if (c->aux2) return;
  hip_check(hipStreamCreateWithFlags(&c->aux2, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipEventCreateWithFlags(&c->ev_fork2, hipEventDisableTiming), "hipEventCreate");
  hip_check(hipEventCreateWithFlags(&c->ev_join2, hipEventDisableTiming), "hipEventCreate");
  hip_check(hipEventCreateWithFlags(&c->ev_placed, hipEventDisableTiming), "hipEventCreate");
}

void ensure_aux3(splhip_ctx *c, int k = 0) {
  splhip_ctx::AuxSet &X = c->ax[k];
  if (X.aux3) return;
  hip_check(hipStreamCreateWithFlags(&X.aux3, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipStreamCreateWithFlags(&X.aux4, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipEventCreateWithFlags(&X.fork3, hipEventDisableTiming), "hipEventCreate");
  hip_check(hipEventCreateWithFlags(&X.join3, hipEventDisableTiming), "hipEventCreate");
  hip_check(hipEventCreateWithFlags(&X.fork4, hipEventDisableTiming), "hipEventCreate");
  hip_check(hipEventCreateWithFlags(&X.join4, hipEventDisableTiming), "hipEventCreate");
}


void upload_palettes(splhip_ctx *c) {
  if (!c->pal_dirty) return;
  int stride = 1;
  for (int t = 0; t < kMaxTypes; ++t) stride = std::max<int>(stride, c->pal_x[t].size());
  const size_t need = static_cast<size_t>(stride) * kMaxTypes;
  if (need > c->pal_cap) {
    if (c->d_palx) hip_check(hipFree(c->d_palx), "hipFree palette");
    if (c->d_palrgb) hip_check(hipFree(c->d_palrgb), "hipFree palette");
    hip_check(hipMalloc(&c->d_palx, need * sizeof(double)), "hipMalloc palette");
    hip_check(hipMalloc(&c->d_palrgb, need * 3 * sizeof(float)), "hipMalloc palette");
    c->pal_cap = need;
  }
  std::vector<double> hx(need, 0.0);
  std::vector<float> hrgb(need * 3, 0.f);
  for (int t = 0; t < kMaxTypes; ++t) {
    std::copy(c->pal_x[t].begin(), c->pal_x[t].end(), hx.begin() + static_cast<size_t>(t) * stride);
    std::copy(c->pal_rgb[t].begin(), c->pal_rgb[t].end(),
              hrgb.begin() + static_cast<size_t>(t) * stride * 3);
  }
  hip_check(hipMemcpy(c->d_palx, hx.data(), need * sizeof(double), hipMemcpyHostToDevice),
            "upload palette");
  hip_check(hipMemcpy(c->d_palrgb, hrgb.data(), need * 3 * sizeof(float), hipMemcpyHostToDevice),
            "upload palette");
  c->pal_stride = stride;
  c->pal_dirty = false;
}

int check_inputs(splhip_ctx *c, const splhip_particles *P, const splhip_camera *cam,
                 const splhip_params *prm) {
  if (!c || !P || !cam || !prm) return fail(c, SPLHIP_E_ARG, "null argument");
  if (P->n < 0) return fail(c, SPLHIP_E_ARG, "negative particle count");
  if (P->n > 0 && (!P->x || !P->y || !P->z || !P->hsml || !P->inten || !P->col))
    return fail(c, SPLHIP_E_ARG, "missing particle field");
  if (P->n >= (int64_t(1) << 32) - 1) return fail(c, SPLHIP_E_CAPACITY, "too many particles");
  if (cam->xres <= 0 || cam->yres <= 0) return fail(c, SPLHIP_E_ARG, "bad image size");
  if (prm->ptypes < 1 || prm->ptypes > kMaxTypes) return fail(c, SPLHIP_E_ARG, "bad ptypes");
  if (prm->tile_size < 1 || prm->tile_size > 4096) return fail(c, SPLHIP_E_ARG, "bad tile_size");
  if (prm->sort_type < 0 || prm->sort_type > 4)
    return fail(c, SPLHIP_E_ARG, "unknown sorting chosen ...");
  if (prm->hip_ppw != 0 && prm->hip_ppw != 1 && prm->hip_ppw != 2 && prm->hip_ppw != 4 &&
      prm->hip_ppw != 8)
    return fail(c, SPLHIP_E_ARG, "hip_ppw must be 1, 2, 4 or 8 (0 = auto)");
  for (int t = 0; t < prm->ptypes; ++t) {
    if (!prm->col_vector[t] && c->pal_x[t].empty())
      return fail(c, SPLHIP_E_ARG, "trying to access an empty map (palette " + std::to_string(t) + ")");
    if (prm->col_vector[t] && (!P->c1 || !P->c2))
      return fail(c, SPLHIP_E_ARG, "vector colour type without c1/c2 fields");
  }
  return 0;
}

int region_size(const splhip_params *prm) {
  int G = prm->hip_region;
  if (G == 0) G = 8;
  return G;
}

BinGeom make_geom(const splhip_camera *cam, const splhip_params *prm, int G, const Consts &k) {
  BinGeom g;
  g.T = prm->tile_size;
  g.xres = cam->xres;
  g.yres = cam->yres;
  g.ncx = (cam->xres + g.T - 1) / g.T;
  g.ncy = (cam->yres + g.T - 1) / g.T;
  g.G = G;
  g.ngm = G > 0 ? (g.T + G - 1) / G : 1;
  // splotch_host.cc:342-343
  const int chunkdim = g.T;
  g.rcell = std::sqrt(2.f) * (chunkdim * 0.5f - 0.5f);
  g.cmid0 = 0.5f * (chunkdim - 1);
  g.rfac = k.rfac;
  g.grayabsorb = prm->gray_absorption;
  return g;
}

int ceil_log2(uint64_t v) {
  int b = 0;
  while ((uint64_t(1) << b) < v) ++b;
  return std::max(b, 1);
}

// ------------------------------------------------------------------ batch plan

// One view as the entry points hand it over.
struct JobIn {
  const splhip_particles *P;
  const splhip_camera *cam;
  const splhip_params *prm;
  float *pic;
};

// Host plan of a batch and the device copies of its tables.
struct Plan {
  int nj = 0, G = 8;
  int64_t ntot = 0;
  uint32_t nblocks = 0, nunits = 0;
  uint32_t nregions = 0;             // G == 8: regions of 2^rsh x 2^rsh units
  int rsh = 0, rbits = 0;            // rsh 0: no regions; rbits: bits of a region id
  bool present[2] = {false, false};  // some job has types of the pass
  bool sort = false;                 // some job depth-sorts its ordered types
  std::vector<ProjectArgs> pa;
  std::vector<uint32_t> blk, pbase, ubase, rbase;
  std::vector<BinGeom> geom;
  std::vector<JobSplat> js;
  // device tables (inside the S_JOBS slot)
  const ProjectArgs *d_pa = nullptr;
  const uint32_t *d_blk = nullptr, *d_pbase = nullptr, *d_ubase = nullptr, *d_rbase = nullptr;
  const BinGeom *d_geom = nullptr;
  JobSplat *d_js = nullptr;
  // per pass
  uint32_t m[2] = {0, 0}, e[2] = {0, 0}, nitems[2] = {0, 0}, nheavy[2] = {0, 0};
  // long lists after the first nheavy items (one item each), which the 8x8
  // splat serves by channel
  uint32_t nsplit[2] = {0, 0};
  // staged lists (the first nstage items; nheavy counts them too) and their
  // groups of 4 entries (padded to whole chain loads)
  uint32_t nstage[2] = {0, 0};
  uint64_t sgroups[2] = {0, 0};
  uint64_t etot[2] = {0, 0};  // (block, particle) entries of each pass (or chunk)
  uint32_t e16[2] = {0, 0};         // (region, particle) entries
  bool r16[2] = {false, false};     // the pass's region counts are scanned (e16 valid)
  uint64_t used16[2] = {0, 0};      // region entries of the bins built through regions
  std::vector<uint32_t> mo[2];  // host copies of the job offsets in pass order
  const uint32_t *order[2] = {nullptr, nullptr};
  const Rec *recs[2] = {nullptr, nullptr};
  const uint32_t *ent[2] = {nullptr, nullptr};
  const uint32_t *bin_off[2] = {nullptr, nullptr};
  const uint32_t *d_mo[2] = {nullptr, nullptr};
  const uint32_t *items[2] = {nullptr, nullptr};  // sorted work items (8 px regions)
  const uint32_t *qbad = nullptr;  // device flag: the ordered pass has a non-finite q
  // workspace of the buffers (the context's, or a capture reservation's)
  Workspace *ws = nullptr;
  // device counts: sizes stay on the device (m, e, e16, items and heavy items
  // are capacities on the host; kernels read the counts from tot), so no
  // stage waits on the host
  bool dc = false;
  uint32_t *tot = nullptr;
  bool use16[2] = {false, false};  // dc: the pass bins through regions
  uint32_t ovf_bit = 0;
};

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Host staging (pinned) of at least `bytes`, free for writing: waits for the
// last copy that used it (marked by staged()).
char *staging(splhip_ctx *c, size_t bytes) {
  if (c->ev_stage) hip_check(hipEventSynchronize(c->ev_stage), "wait staging");
  if (c->h_stage_cap < bytes) {
    if (c->h_stage) hip_check(hipHostFree(c->h_stage), "hipHostFree");
    c->h_stage = nullptr;
    hip_check(hipHostMalloc(reinterpret_cast<void **>(&c->h_stage), bytes, hipHostMallocDefault),
              "hipHostMalloc");
    c->h_stage_cap = bytes;
  }
  return c->h_stage;
}

// Mark the end of the copies queued on s that use the staging buffer.
void staged(splhip_ctx *c, hipStream_t s) {
  if (!c->ev_stage)
    hip_check(hipEventCreateWithFlags(&c->ev_stage, hipEventDisableTiming), "hipEventCreate");
  hip_check(hipEventRecord(c->ev_stage, s), "hipEventRecord");
}

// Host tables of a batch (no device work besides a palette upload).
void build_plan(splhip_ctx *c, int nj, const JobIn *jobs, int G, Plan &pl) {
  upload_palettes(c);
  pl.ws = &c->ws;
  pl.nj = nj;
  pl.G = G;
  pl.pa.resize(nj);
  pl.geom.resize(nj);
  pl.js.resize(nj);
  pl.blk.assign(nj + 1, 0);
  pl.pbase.assign(nj + 1, 0);
  pl.ubase.assign(nj + 1, 0);
  pl.rbase.assign(nj + 1, 0);
  std::vector<uint32_t> rbase4(nj + 1, 0);
  uint64_t units = 0, blocks = 0, regions = 0, regions4 = 0;
  int64_t ntot = 0;
  for (int j = 0; j < nj; ++j) {
    const splhip_particles *P = jobs[j].P;
    const splhip_camera *cam = jobs[j].cam;
    const splhip_params *prm = jobs[j].prm;
    ProjectArgs &a = pl.pa[j];
    a = ProjectArgs{};
    a.x = P->x; a.y = P->y; a.z = P->z; a.hsml = P->hsml; a.inten = P->inten; a.col = P->col;
    a.c1 = P->c1; a.c2 = P->c2; a.type = P->type; a.n = P->n;
    a.poff = ntot;
    a.cam = *cam;
    for (int t = 0; t < kMaxTypes; ++t) {
      a.smooth_fac[t] = prm->smooth_fac[t];
      a.brightness[t] = prm->brightness[t];
      a.col_vector[t] = prm->col_vector[t];
      // pass bits: 1 ordered, 2 additive, 3 both (a_eq_e 2)
      a.pass_code[t] = prm->a_eq_e[t] == 2 ? 3 : (prm->a_eq_e[t] ? 2 : 1);
      a.pal_n[t] = static_cast<int32_t>(c->pal_x[t].size());
    }
    for (int t = 0; t < prm->ptypes; ++t) {
      if (prm->a_eq_e[t] != 1) pl.present[0] = true;
      if (prm->a_eq_e[t] != 0) pl.present[1] = true;
    }
    a.pal_stride = c->pal_stride;
    a.ntypes = prm->ptypes;
    a.sort_type = prm->sort_type;
    if (prm->sort_type != 0) pl.sort = true;
    a.pal_x = c->d_palx;
    a.pal_rgb = c->d_palrgb;
    a.bfak = c->consts.bfak;
    a.rfac = c->consts.rfac;
    pl.pbase[j] = static_cast<uint32_t>(ntot);
    pl.blk[j] = static_cast<uint32_t>(blocks);
    pl.ubase[j] = static_cast<uint32_t>(units);
    pl.rbase[j] = static_cast<uint32_t>(regions);
    rbase4[j] = static_cast<uint32_t>(regions4);
    ntot += P->n;
    blocks += (static_cast<uint64_t>(P->n) + kPerBlock - 1) / kPerBlock;
    const BinGeom g = make_geom(cam, prm, G, c->consts);
    pl.geom[j] = g;
    units += static_cast<uint64_t>(g.ncx) * g.ncy * (G > 0 ? g.ngm * g.ngm : 1);
    if (G == 8) {
      const uint64_t n2 = (static_cast<uint64_t>(g.ngm) + 1) / 2;
      const uint64_t n4 = (static_cast<uint64_t>(g.ngm) + 3) / 4;
      regions += static_cast<uint64_t>(g.ncx) * g.ncy * n2 * n2;
      regions4 += static_cast<uint64_t>(g.ncx) * g.ncy * n4 * n4;
    }
    JobSplat &J = pl.js[j];
    J = JobSplat{};
    J.pic = jobs[j].pic;
    J.T = g.T;
    J.xres = cam->xres;
    J.yres = cam->yres;
    J.ncx = g.ncx;
    J.ngm = g.ngm;
    J.grayabsorb = prm->gray_absorption;
  }
  if (ntot >= (int64_t(1) << 32) - 1) throw HipError{hipErrorInvalidValue, "too many particles"};
  if (units >= (uint64_t(1) << 30)) throw HipError{hipErrorInvalidValue, "too many bins"};
  pl.ntot = ntot;
  pl.nblocks = static_cast<uint32_t>(blocks);
  pl.nunits = static_cast<uint32_t>(units);
  pl.blk[nj] = pl.nblocks;
  pl.pbase[nj] = static_cast<uint32_t>(ntot);
  pl.ubase[nj] = pl.nunits;
  pl.rbase[nj] = static_cast<uint32_t>(regions);
  rbase4[nj] = static_cast<uint32_t>(regions4);
  // Binning through regions: the largest group (4 x 4, then 2 x 2 units)
  // whose region ids and unit masks fit a 32-bit key (hip_bin_group forces
  // one, 1 = none).
  pl.rsh = 0;
  pl.nregions = 0;
  if (G == 8) {
    const int bg = jobs[0].prm->hip_bin_group;
    for (int S : {4, 2}) {
      if (bg == 1 || (bg > 0 && bg != S)) continue;
      const uint64_t nr = S == 4 ? regions4 : regions;
      const int rbits = ceil_log2(nr);
      if (nr == 0 || nr >= (uint64_t(1) << 31) || rbits + S * S > 32) continue;
      pl.rsh = S == 4 ? 2 : 1;
      pl.rbits = rbits;
      pl.nregions = static_cast<uint32_t>(nr);
      if (S == 4) pl.rbase = rbase4;
      break;
    }
  }
}

// The plan's device tables as one block: [ProjectArgs][blk][pbase][ubase]
// [rbase][BinGeom][JobSplat]; offsets in o[0..6], total size returned.
size_t plan_tables(const Plan &pl, std::vector<char> *out, size_t o[7]) {
  const int nj = pl.nj;
  const size_t o_pa = 0;
  const size_t o_blk = align256(o_pa + nj * sizeof(ProjectArgs));
  const size_t o_pb = align256(o_blk + (nj + 1) * 4);
  const size_t o_ub = align256(o_pb + (nj + 1) * 4);
  const size_t o_rb = align256(o_ub + (nj + 1) * 4);
  const size_t o_g = align256(o_rb + (nj + 1) * 4);
  const size_t o_js = align256(o_g + nj * sizeof(BinGeom));
  const size_t total = align256(o_js + nj * sizeof(JobSplat));
  o[0] = o_pa; o[1] = o_blk; o[2] = o_pb; o[3] = o_ub; o[4] = o_rb; o[5] = o_g; o[6] = o_js;
  if (out) {
    out->assign(total, 0);
    char *h = out->data();
    std::memcpy(h + o_pa, pl.pa.data(), nj * sizeof(ProjectArgs));
    std::memcpy(h + o_blk, pl.blk.data(), (nj + 1) * 4);
    std::memcpy(h + o_pb, pl.pbase.data(), (nj + 1) * 4);
    std::memcpy(h + o_ub, pl.ubase.data(), (nj + 1) * 4);
    std::memcpy(h + o_rb, pl.rbase.data(), (nj + 1) * 4);
    std::memcpy(h + o_g, pl.geom.data(), nj * sizeof(BinGeom));
    std::memcpy(h + o_js, pl.js.data(), nj * sizeof(JobSplat));
  }
  return total;
}

// Point the plan at device tables laid out by plan_tables.
void plan_bind(Plan &pl, char *d, const size_t o[7]) {
  pl.d_pa = reinterpret_cast<const ProjectArgs *>(d + o[0]);
  pl.d_blk = reinterpret_cast<const uint32_t *>(d + o[1]);
  pl.d_pbase = reinterpret_cast<const uint32_t *>(d + o[2]);
  pl.d_ubase = reinterpret_cast<const uint32_t *>(d + o[3]);
  pl.d_rbase = reinterpret_cast<const uint32_t *>(d + o[4]);
  pl.d_geom = reinterpret_cast<const BinGeom *>(d + o[5]);
  pl.d_js = reinterpret_cast<JobSplat *>(d + o[6]);
}

// Host plan of a batch and the upload of its tables into the context's
// workspace (through the pinned staging buffer).
void make_plan(splhip_ctx *c, int nj, const JobIn *jobs, int G, Plan &pl, hipStream_t s) {
  build_plan(c, nj, jobs, G, pl);
  size_t o[7];
  const size_t total = plan_tables(pl, nullptr, o);
  char *h = staging(c, total);
  std::vector<char> tab;
  plan_tables(pl, &tab, o);
  std::memcpy(h, tab.data(), total);
  char *d = c->ws.get_as<char>(S_JOBS, total);
  hip_check(hipMemcpyAsync(d, h, total, hipMemcpyHostToDevice, s), "upload job tables");
  staged(c, s);
  plan_bind(pl, d, o);
}

// Totals slot layout (device u32): [0..1] pass sizes, [2..3] bin entries,
// [4..5] work items, [6..7] heavy items, [8..11] bin entries as u64,
// [12..13] region entries, [14] the ordered pass has a non-finite q,
// [15] capacity overflow bits of a device-count render (bit p: pass p),
// [16..47] lists from stage_level(stage_len, j) entries on, [48..79] their
// groups of 4 entries (padded), j = 0..15 ([16 + 16 p + j]: pass p),
// [80..81] the longest list, [82..89] a device-count render's long-list
// partition (4 per pass: staged, producer/consumer, both).
enum {
  T_M = 0, T_E = 2, T_ITEMS = 4, T_HEAVY = 6, T_E64 = 8, T_E16 = 12, T_QBAD = 14, T_OVF = 15,
  T_STAGE = 16, T_SGRP = 48, T_LMAX = 80, T_PART = 82, T_N = 90
};
constexpr int kTotalsWords = 128;  // the totals buffer, device and pinned host
static_assert(T_N <= kTotalsWords, "totals fit their buffers");
static_assert(T_SGRP - T_STAGE == 2 * kStageLevels, "one level count per pass");

// Project all jobs, then the compaction scan of each present pass and its job
// offsets in pass order; one host sync reads the pass sizes and offsets.
void stage_project_compact(splhip_ctx *c, Plan &pl, Rec *rec, uint8_t *act, uint32_t *tot,
                           hipStream_t s) {
  Workspace &ws = *pl.ws;
  mark(c, 0, true, s);
  launch_project(pl.d_pa, pl.d_blk, pl.nj, pl.nblocks, rec, act, s);
  mark(c, 0, false, s);
  launch_fill_u32(tot, 0, T_N, s);
  uint32_t *mo[2] = {nullptr, nullptr};
  for (int p = 0; p < 2; ++p) {
    if (!pl.present[p] || pl.ntot == 0) continue;
    mark(c, p == 0 ? 1 : 4, true, s);
    uint32_t *pos = ws.get_as<uint32_t>(pslot(p, P_POS), static_cast<size_t>(pl.ntot));
    scan_u8bit(act, static_cast<uint8_t>(p + 1), pos, pl.ntot, tot + T_M + p, ws, s);
    mo[p] = ws.get_as<uint32_t>(pslot(p, P_MO), pl.nj + 1);
    launch_pass_offsets(pos, pl.d_pbase, pl.nj, pl.ntot, tot + T_M + p, mo[p], s);
    mark(c, p == 0 ? 1 : 4, false, s);
  }
  if (pl.dc) {  // the pass sizes stay on the device; capacities: every particle
    for (int p = 0; p < 2; ++p) {
      pl.m[p] = mo[p] ? static_cast<uint32_t>(pl.ntot) : 0u;
      pl.d_mo[p] = mo[p];
    }
    return;
  }
  // sync 1: pass sizes and per-job offsets
  uint32_t *h = c->h_tot;
  const size_t nmo = static_cast<size_t>(pl.nj) + 1;
  std::vector<uint32_t> &hbv = c->h_mo;  // pageable: read after the sync below
  hbv.resize(2 * nmo);
  uint32_t *hb = hbv.data();
  hip_check(hipMemcpyAsync(h, tot, T_N * 4, hipMemcpyDeviceToHost, s), "read totals");
  for (int p = 0; p < 2; ++p)
    if (mo[p])
      hip_check(hipMemcpyAsync(hb + p * nmo, mo[p], nmo * 4, hipMemcpyDeviceToHost, s),
                "read pass offsets");
  hip_check(hipStreamSynchronize(s), "sync pass sizes");
  for (int p = 0; p < 2; ++p) {
    pl.m[p] = mo[p] ? h[T_M + p] : 0;
    pl.d_mo[p] = mo[p];
    if (mo[p]) pl.mo[p].assign(hb + p * nmo, hb + (p + 1) * nmo);
    else pl.mo[p].assign(nmo, 0);
  }
}

// Per-job pass flags from the pass sizes: the ordered pass writes the image
// of every job that has ordered particles; the additive pass adds to it there
// and writes it elsewhere (render_new renders each mode from zero).
void upload_job_flags(splhip_ctx *c, Plan &pl, hipStream_t s) {
  for (int j = 0; j < pl.nj; ++j) {
    const bool w0 = pl.mo[0][j + 1] > pl.mo[0][j];
    const bool w1 = pl.mo[1][j + 1] > pl.mo[1][j];
    JobSplat &J = pl.js[j];
    J.accumulate[0] = 0;
    J.accumulate[1] = w0 ? 1 : 0;
    J.need_empty[0] = w0 ? 1 : 0;
    J.need_empty[1] = (w1 && !w0) ? 1 : 0;
  }
  const size_t bytes = pl.nj * sizeof(JobSplat);
  char *h = staging(c, bytes);
  std::memcpy(h, pl.js.data(), bytes);
  hip_check(hipMemcpyAsync(pl.d_js, h, bytes, hipMemcpyHostToDevice, s), "upload job flags");
  staged(c, s);
}

// Render order of pass p: compaction (job-major, index order), then for the
// ordered pass the depth sort, stable, within each job.
void stage_order(splhip_ctx *c, Plan &pl, int p, const Rec *rec, const uint8_t *act,
                 hipStream_t s) {
  Workspace &ws = *pl.ws;
  const uint32_t m = pl.m[p];
  const bool do_sort = (p == 0) && pl.sort;
  mark(c, p == 0 ? 1 : 4, true, s);
  const uint32_t *dm = pl.dc ? pl.tot + T_M + p : nullptr;  // device count (m: capacity)
  uint32_t *k0 = do_sort ? ws.get_as<uint32_t>(pslot(p, P_K0), m) : nullptr;
  uint32_t *v0 = ws.get_as<uint32_t>(pslot(p, P_V0), m);
  const uint32_t *pos = ws.get_as<uint32_t>(pslot(p, P_POS), static_cast<size_t>(pl.ntot));
  launch_compact(pl.d_pa, pl.d_blk, pl.nj, pl.nblocks, act, static_cast<uint8_t>(p + 1), pos, rec,
                 k0, v0, s);
  const uint32_t *order = v0;
  if (do_sort) {
    uint32_t *k1 = ws.get_as<uint32_t>(pslot(p, P_K1), m);
    uint32_t *v1 = ws.get_as<uint32_t>(pslot(p, P_V1), m);
    const int r = radix_sort_pairs(k0, v0, k1, v1, m, 0, 32, ws, s, dm);
    uint32_t *kk = r ? k1 : k0, *vv = r ? v1 : v0;
    uint32_t *ko = r ? k0 : k1, *vo = r ? v0 : v1;
    order = vv;
    if (pl.nj > 1) {  // regroup by job, stable: (job, key, index) order
      launch_job_of_index(vv, m, pl.d_pbase, pl.nj, kk, s, dm);
      const int r2 = radix_sort_pairs(kk, vv, ko, vo, m, 0, ceil_log2(pl.nj), ws, s, dm);
      order = r2 ? vo : vv;
    }
  }
  pl.order[p] = order;
  mark(c, p == 0 ? 1 : 4, false, s);
}

// Gather into render order and count each particle's units; the entry total
// lands in tot[T_E + p].  With 8 px units and a region size chosen for the
// frame (region ids and block masks fit one 32-bit key) the entries can be
// binned through regions: their counts are scanned too (total in
// tot[T_E16 + p]).
void stage_count(splhip_ctx *c, Plan &pl, int p, const Rec *rec, uint32_t *tot, hipStream_t s) {
  Workspace &ws = *pl.ws;
  const uint32_t m = pl.m[p];
  mark(c, p == 0 ? 2 : 5, true, s);
  pl.r16[p] = pl.rsh > 0;
  Rec *recs = ws.get_as<Rec>(pslot(p, P_RECS), m);
  uint32_t *cnt = ws.get_as<uint32_t>(pslot(p, P_CNT), m);
  uint32_t *off = ws.get_as<uint32_t>(pslot(p, P_OFF), m);
  uint32_t *cnt16 = pl.r16[p] ? ws.get_as<uint32_t>(pslot(p, P_CNT16), m) : nullptr;
  const uint32_t *dm = pl.dc ? tot + T_M + p : nullptr;  // device count (m: capacity)
  launch_gather_count(pl.order[p], m, rec, recs, cnt, cnt16, pl.rsh, pl.d_geom, pl.d_mo[p], pl.nj,
                      p == 0 ? tot + T_QBAD : nullptr, s, dm);
  if (p == 0) pl.qbad = tot + T_QBAD;
  scan_u32(cnt, off, m, tot + T_E + p, ws, s, 0, dm);
  launch_sum_u64(cnt, m, reinterpret_cast<uint64_t *>(tot + T_E64 + 2 * p), s, dm);
  if (cnt16)
    scan_u32(cnt16, ws.get_as<uint32_t>(pslot(p, P_OFF16), m), m, tot + T_E16 + p, ws, s, 0,
             dm);
  if (pl.dc)  // entries past the reserved capacity: the pass is dropped and flagged
    launch_cap_check(tot + T_M + p, tot + T_E + p,
                     reinterpret_cast<const uint64_t *>(tot + T_E64 + 2 * p),
                     pl.use16[p] ? tot + T_E16 + p : nullptr, pl.e[p], pl.e16[p], 1u << p,
                     tot + T_OVF, s);
  pl.recs[p] = recs;
  mark(c, p == 0 ? 2 : 5, false, s);
}

// Emit the (unit, particle) entries, sort them by unit (stable, so each list
// keeps render order) and build the unit offsets.  Through regions of S x S
// blocks: emit (region | block mask, particle) entries, the mask naming the
// blocks the footprint can reach, sort them by region, and split each
// region's list into its 8x8 blocks' lists (render order in both forms; the
// region lists leave out entries that would add +0 to every pixel).
// wt: workspace of the temporaries (sorts and scans; default the plan's).
void stage_bins(splhip_ctx *c, Plan &pl, int p, hipStream_t s, Workspace *wt = nullptr) {
  Workspace &ws = *pl.ws;
  Workspace &tmp = wt ? *wt : ws;
  const uint32_t m = pl.m[p], e = pl.e[p];
  mark(c, p == 0 ? 2 : 5, true, s);
  // device counts (m, e, e16 are capacities then)
  const uint32_t *dm = pl.dc ? pl.tot + T_M + p : nullptr;
  const uint32_t *de = pl.dc ? pl.tot + T_E + p : nullptr;
  const uint32_t *de16 = pl.dc ? pl.tot + T_E16 + p : nullptr;
  // regions pay off where they merge entries (large footprints); small
  // footprints keep one entry per block either way
  if (pl.dc ? pl.use16[p] : (pl.r16[p] && 3ull * pl.e16[p] <= 2ull * e)) {
    const uint32_t e16 = pl.e16[p], nr = pl.nregions, nu = pl.nunits;
    const uint32_t *off16 = ws.get_as<uint32_t>(pslot(p, P_OFF16), m);
    uint32_t *rk0 = ws.get_as<uint32_t>(pslot(p, P_EK0), e16);
    uint32_t *rv0 = ws.get_as<uint32_t>(pslot(p, P_EV0), e16);
    uint32_t *rk1 = ws.get_as<uint32_t>(pslot(p, P_EK1), e16);
    uint32_t *rv1 = ws.get_as<uint32_t>(pslot(p, P_EV1), e16);
    const int S = 1 << pl.rsh, rbits = pl.rbits;
    launch_emit(pl.rsh, rbits, pl.recs[p], m, off16, e16, pl.d_geom, pl.d_rbase, pl.d_mo[p],
                pl.nj, rk0, rv0, s, dm, de16);
    // the sort orders the region bits only: the unit masks above them ride
    // along, and each region's list keeps render order
    const int r = radix_sort_pairs(rk0, rv0, rk1, rv1, e16, 0, rbits, tmp, s, de16);
    const uint32_t *rk = r ? rk1 : rk0, *rv = r ? rv1 : rv0;
    uint32_t *roff = ws.get_as<uint32_t>(pslot(p, P_ROFF), static_cast<size_t>(nr) + 1);
    launch_bin_offsets(rk, e16, nr, roff, s, (1u << rbits) - 1u, de16);
    const size_t ncb = static_cast<size_t>(S * S) * split_chunks(e16);
    uint32_t *ccnt = ws.get_as<uint32_t>(pslot(p, P_CCNT), ncb + 1);
    uint32_t *cpre = ws.get_as<uint32_t>(pslot(p, P_CPRE), ncb + 1);
    uint32_t *sx = ws.get_as<uint32_t>(pslot(p, P_SX), static_cast<size_t>(S * S) * (nr + 1));
    uint32_t *ucnt = ws.get_as<uint32_t>(pslot(p, P_UCNT), nu);
    uint32_t *bin_off = ws.get_as<uint32_t>(pslot(p, P_BINOFF), static_cast<size_t>(nu) + 1);
    uint32_t *ent = ws.get_as<uint32_t>(pslot(p, P_ENT), e);
    launch_split_chunks(S, rk, e16, rbits, ccnt, s, de16);
    scan_u32(ccnt, cpre, static_cast<int64_t>(ncb), cpre + ncb, tmp, s);
    launch_split_bounds(S, roff, rk, nr, cpre, e16, rbits, pl.d_rbase, pl.d_ubase, pl.d_geom,
                        pl.nj, sx, ucnt, s);
    scan_u32(ucnt, bin_off, nu, bin_off + nu, tmp, s);
    launch_split_base(S, nr, pl.d_rbase, pl.d_ubase, pl.d_geom, pl.nj, sx, bin_off, sx, s);
    launch_split_scatter(S, rk, rv, e16, rbits, cpre, sx, ent, s, de16);
    pl.ent[p] = ent;
    pl.bin_off[p] = bin_off;
    pl.used16[p] += e16;
    mark(c, p == 0 ? 2 : 5, false, s);
    return;
  }
  const uint32_t *off = ws.get_as<uint32_t>(pslot(p, P_OFF), m);
  uint32_t *ek0 = ws.get_as<uint32_t>(pslot(p, P_EK0), e);
  uint32_t *ev0 = ws.get_as<uint32_t>(pslot(p, P_EV0), e);
  uint32_t *ek1 = ws.get_as<uint32_t>(pslot(p, P_EK1), e);
  uint32_t *ev1 = ws.get_as<uint32_t>(pslot(p, P_EV1), e);
  launch_emit(0, 0, pl.recs[p], m, off, e, pl.d_geom, pl.d_ubase, pl.d_mo[p], pl.nj, ek0, ev0, s,
              dm, de);
  const int r = radix_sort_pairs(ek0, ev0, ek1, ev1, e, 0, ceil_log2(pl.nunits), tmp, s, de);
  uint32_t *bin_off = ws.get_as<uint32_t>(pslot(p, P_BINOFF), static_cast<size_t>(pl.nunits) + 1);
  launch_bin_offsets(r ? ek1 : ek0, e, pl.nunits, bin_off, s, ~0u, de);
  pl.ent[p] = r ? ev1 : ev0;
  pl.bin_off[p] = bin_off;
  mark(c, p == 0 ? 2 : 5, false, s);
}

struct Tuning {
  int variant, ppw;
  uint32_t split, heavy, heavy_max;
  uint64_t chunk;  // most (block, particle) entries of one pass chunk
  uint32_t stage, stage_ord;  // lists from this length on are staged, additive / ordered
                              // pass (0xffffffff: none)
  bool stage_ord_auto;        // ordered staging only for a pass bound by its longest list
};

// Staged operands of a pass (3 KiB per group of 4 entries additive, 4 KiB +
// 48 B ordered); a pass whose staged lists need more renders them unstaged.
constexpr uint64_t kStageBudget = uint64_t(32) << 30;
// Default ordered staging only for a pass bound by its longest list: one
// whose longest list, at about 108 clocks per entry in a producer/consumer
// workgroup, takes at least twice as long as the pass's entries at about 0.25
// clocks each over all CUs (the 8x8 splat's VALU work), i.e. longest list x
// kStageOrderedRatio >= entries.  A pass below keeps every CU's VALU busy for
// about as long as its longest lists take, and the operand kernel's extra
// VALU and HBM work would only slow it (16M particles at 2048^2: 837k-entry
// list, 301M entries, 43.3-53.7 ms staged from 786432-393216 entries on
// against 43.7 ms unstaged; 64 views: 65.6M ordered entries, 68.6 -> 48.1 ms;
// 500k at 512^2: 121k-entry list, 1.3M entries, 6.21 -> 3.58 ms).
constexpr uint64_t kStageOrderedRatio = 216;
// Below that, a pass whose longest lists would still end after its other
// entries (longest list x kStageOrderedCrit >= entries, at 108 / 0.25 clocks
// per entry) stages only those: the lists from entries / kStageOrderedCrit
// on (rounded down to a level), whose operands then compete least with the
// rest (16M particles at 2048^2: 813k-entry list, 331M entries; the 4 lists
// from 741376 on staged, 41.8 ms against 42.8 unstaged).
constexpr uint64_t kStageOrderedCrit = 432;
// Long lists left to the 8x8 splat that would outlast the pass, walked by
// one wave at about 230 ns per entry against about 0.1 ns per entry of the
// pass over all CUs (16M particles at 2048^2), are served by channel, three
// waves each: those from entries / kSplitRatio on (rounded down to a level).
// Serving every such list by channel costs more than it saves (41.2 ->
// 48.2 ms at 16M / 2048^2): three one-channel walks take more VALU work than
// one three-channel walk.
constexpr uint32_t kSplitRatio = 2048;
// Staged lists whose chains start right after their own operands (the
// longest; the others' operands and chains follow beside them).  The 64-view
// batch: 34.9 ms with 64, 33.7 with 16, 34.1 with 4; the 1M / 800^2 frame and
// the 500k ordered one within 1 % of each other.
constexpr uint32_t kStagePhase1 = 16;
// Default ordered staging threshold: about the longest list / 16 (the
// 64-view batch: 40.4 ms staged from 65536 entries on, 43.9 from 8192; 500k
// at 512^2, longest list 121k: 2.67 ms from 8192, 3.39 from 65536).
constexpr uint64_t kStageOrderedSpan = 16;

Tuning tuning_of(const splhip_params *prm) {
  Tuning t;
  t.variant = prm->hip_variant > 0 ? prm->hip_variant : 4;
  t.ppw = prm->hip_ppw;
  // lists from hip_heavy on go to producer/consumer workgroups (up to the cap)
  // or stay whole: splitting by channel pays only without those workgroups
  t.split = prm->hip_split > 0 ? static_cast<uint32_t>(prm->hip_split) : 32768u;
  t.heavy = prm->hip_heavy < 0 ? 0xffffffffu
            : prm->hip_heavy > 0 ? static_cast<uint32_t>(prm->hip_heavy) : 32768u;
  // a producer/consumer workgroup occupies a CU: leave the splat8 items
  // about a quarter of the CUs while the longest lists run
  t.heavy_max = prm->hip_heavy_max > 0 ? static_cast<uint32_t>(prm->hip_heavy_max) : 192u;
  // the largest int32 key value in whole MiB: fewer chunks, each a full
  // pipeline with a host sync and a splat tail; offsets stay below 2^31
  t.chunk = prm->hip_chunk_entries > 0 ? static_cast<uint64_t>(prm->hip_chunk_entries)
                                       : (uint64_t(1) << 31) - (uint64_t(1) << 20);
  // additive lists from 8192 entries on: a lane-per-pixel wave walks such a
  // list at about 100 ns per entry
  t.stage = prm->hip_stage < 0 ? 0xffffffffu
            : prm->hip_stage > 0 ? static_cast<uint32_t>(prm->hip_stage) : 8192u;
  // ordered lists from 8192 entries on (for a pass bound by its longest
  // list, see kStageOrderedRatio)
  t.stage_ord = prm->hip_stage_ordered < 0 ? 0xffffffffu
                : prm->hip_stage_ordered > 0 ? static_cast<uint32_t>(prm->hip_stage_ordered)
                                             : 8192u;
  t.stage_ord_auto = prm->hip_stage_ordered == 0;
  return t;
}

SplatArgs splat_args(splhip_ctx *c, const Plan &pl, int p, const Tuning &tn) {
  SplatArgs sa{};
  sa.bin_off = pl.bin_off[p];
  sa.ent = pl.ent[p];
  sa.recs = pl.recs[p];
  sa.exptab = c->d_exptab;
  sa.jobs = pl.d_js;
  sa.ubase = pl.d_ubase;
  sa.nj = pl.nj;
  sa.pass = p;
  sa.rfac = c->consts.rfac;
  sa.h2sigma = c->consts.h2sigma;
  sa.expfac = c->consts.expfac;
  sa.taylorlimit = c->consts.taylorlimit;
  sa.prio_len = tn.split;
  sa.heavy_len = tn.heavy;
  // device counts: staging only where the reservation holds room for it
  sa.stage_len = (pl.dc && pl.nstage[p] == 0) ? 0xffffffffu : p == 1 ? tn.stage : tn.stage_ord;
  sa.qbad = p == 0 ? pl.qbad : nullptr;
  return sa;
}

// Work items of pass p: count per unit (heavy units counted too), scan.
void stage_items_count(splhip_ctx *c, Plan &pl, int p, const Tuning &tn, uint32_t *tot,
                       hipStream_t s, Workspace *wt = nullptr) {
  Workspace &ws = *pl.ws;
  mark(c, p == 0 ? 2 : 5, true, s);
  const SplatArgs sa = splat_args(c, pl, p, tn);
  uint32_t *icnt = ws.get_as<uint32_t>(pslot(p, P_ICNT), pl.nunits);
  uint32_t *ioff = ws.get_as<uint32_t>(pslot(p, P_IOFF), pl.nunits);
  // the count of this call's long-list items only (chunks and autotuning
  // candidates count again)
  launch_fill_u32(tot + T_HEAVY + p, 0u, 1, s);
  launch_fill_u32(tot + T_STAGE + kStageLevels * p, 0u, kStageLevels, s);
  launch_fill_u32(tot + T_SGRP + kStageLevels * p, 0u, kStageLevels, s);
  launch_fill_u32(tot + T_LMAX + p, 0u, 1, s);
  launch_items_count(pl.bin_off[p], pl.nunits, sa, p == 0 ? 1 : 0, tn.split, icnt,
                     tot + T_HEAVY + p, s, tot + T_STAGE + kStageLevels * p,
                     tot + T_SGRP + kStageLevels * p,
                     tot + T_LMAX + p);
  scan_u32(icnt, ioff, pl.nunits, tot + T_ITEMS + p, wt ? *wt : ws, s);
  mark(c, p == 0 ? 2 : 5, false, s);
}

// Item counts of pass p from c->h_tot read from tot + T_ITEMS on (T_N -
// T_ITEMS words): work items; long lists (one item each, sorted first) = the
// staged ones (when their operands fit kStageBudget, else none), then at
// most heavy_max of the others from heavy_len on for the producer/consumer
// kernel; the rest go to the 8x8 splat.  Ordered passes by default stage the lists from
// stage_len << k entries on, k the least with stage_len << k >= the longest
// list / kStageOrderedSpan: shorter lists are off the critical path, and
// staging them costs more than the 8x8 splat does.
void take_items(splhip_ctx *c, Plan &pl, int p, const Tuning &tn) {
  const uint32_t *h = c->h_tot;
  pl.nitems[p] = pl.m[p] ? h[T_ITEMS - T_ITEMS + p] : 0u;
  const uint32_t nheavy_lists = pl.m[p] ? h[T_HEAVY - T_ITEMS + p] : 0u;
  const uint64_t lmax = h[T_LMAX - T_ITEMS + p];
  const uint32_t len0 = p == 1 ? tn.stage : tn.stage_ord;
  const int j = stage_rule(len0, p == 0 && tn.stage_ord_auto, lmax, pl.etot[p],
                           kStageOrderedSpan, kStageOrderedRatio, kStageOrderedCrit);
  uint32_t ns = pl.m[p] && j >= 0 ? h[T_STAGE - T_ITEMS + kStageLevels * p + j] : 0u;
  const uint64_t groups = j >= 0 ? h[T_SGRP - T_ITEMS + kStageLevels * p + j] : 0u;
  const uint64_t bytes = 4 * (stage_floats(p == 1, groups) + stage_q_floats(p == 1, groups));
  if (bytes > kStageBudget) ns = 0;
  pl.nstage[p] = ns;
  pl.sgroups[p] = ns ? groups : 0;
  c->last_stage_n[p] = ns;
  c->last_stage_groups[p] = pl.sgroups[p];
  // staged lists from a threshold at least heavy_len are all among those
  // from heavy_len on; from a lower one they include all of those
  const uint32_t npc = tn.heavy == 0xffffffffu || nheavy_lists <= ns
                           ? 0u : std::min(nheavy_lists - ns, tn.heavy_max);
  // the heaviest items (sorted first); a unit's channel items are
  // independent, so the cut may fall between them
  pl.nheavy[p] = std::min(ns + npc, pl.nitems[p]);
  // lists from min(stage length, heavy_len) entries on have one item each
  const uint32_t len0s = p == 1 ? tn.stage : tn.stage_ord;
  const uint32_t nlong = !pl.m[p] ? 0u
                         : len0s <= tn.heavy ? h[T_STAGE - T_ITEMS + kStageLevels * p]
                                             : nheavy_lists;
  pl.nsplit[p] = pl.m[p] ? split_lists(len0s, h + (T_STAGE - T_ITEMS + kStageLevels * p),
                                       std::min(nlong, pl.nitems[p]), pl.nheavy[p], pl.etot[p],
                                       kSplitRatio)
                         : 0u;
}

// Splat of pass p: items heaviest first; heavy units on the auxiliary stream.
// wt: workspace of the temporaries (the item sort); jobs: a job table in
// place of the plan's; fork: the long-list workgroups run on the auxiliary
// stream beside splat8 (else on s, first); placed: recorded on s just before
// the 8x8 splat's launch, after the long-list workgroups' head start.
void stage_splat(splhip_ctx *c, Plan &pl, int p, const Tuning &tn, hipStream_t s,
                 Workspace *wt = nullptr, JobSplat *jobs = nullptr, bool fork = true,
                 int set = 0, hipEvent_t placed = nullptr) {
  Workspace &ws = *pl.ws;
  splhip_ctx::AuxSet &X = c->ax[set];
  SplatArgs sa0 = splat_args(c, pl, p, tn);
  if (jobs) sa0.jobs = jobs;
  mark(c, p == 0 ? 3 : 6, true, s);
  if (pl.G != 8) {
    launch_splat(pl.G, p == 1, pl.nunits, sa0, s);
    mark(c, p == 0 ? 3 : 6, false, s);
    return;
  }
  const uint32_t nitems = pl.nitems[p], nheavy = std::min(pl.nheavy[p], nitems);
  const uint32_t *ioff = ws.get_as<uint32_t>(pslot(p, P_IOFF), pl.nunits);
  uint32_t *ik0 = ws.get_as<uint32_t>(pslot(p, P_IK0), nitems);
  uint32_t *iv0 = ws.get_as<uint32_t>(pslot(p, P_IV0), nitems);
  uint32_t *ik1 = ws.get_as<uint32_t>(pslot(p, P_IK1), nitems);
  uint32_t *iv1 = ws.get_as<uint32_t>(pslot(p, P_IV1), nitems);
  launch_items_emit(pl.bin_off[p], pl.nunits, sa0, p == 0 ? 1 : 0, tn.split, ioff, ik0, iv0, s);
  // device counts: nitems and nheavy are capacities; the kernels read the
  // pass's item and long-list totals and skip past the long lists themselves
  ItemCounts dc;
  uint32_t *part = nullptr;  // device counts: staged, producer/consumer, both
  if (pl.dc) {
    // the long-list partition from the device counts, as take_items makes it
    part = pl.tot + T_PART + 4 * p;
    PartitionArgs pa;
    pa.stage_cnt = pl.tot + T_STAGE + kStageLevels * p;
    pa.stage_grp = pl.tot + T_SGRP + kStageLevels * p;
    pa.heavy_cnt = pl.tot + T_HEAVY + p;
    pa.lmax = pl.tot + T_LMAX + p;
    pa.etot = reinterpret_cast<const uint64_t *>(pl.tot + T_E64 + 2 * p);
    pa.len0 = pl.nstage[p] ? (p == 1 ? tn.stage : tn.stage_ord) : 0xffffffffu;
    pa.ordered_auto = p == 0 && tn.stage_ord_auto;
    pa.ns_cap = pl.nstage[p];
    pa.groups_cap = pl.sgroups[p];
    pa.heavy_max = tn.heavy == 0xffffffffu ? 0u : tn.heavy_max;
    pa.heavy_len = tn.heavy;
    pa.span = static_cast<uint32_t>(kStageOrderedSpan);
    pa.ratio = static_cast<uint32_t>(kStageOrderedRatio);
    pa.split = kSplitRatio;
    pa.crit = static_cast<uint32_t>(kStageOrderedCrit);
    launch_partition(pa, part, s);
    dc.items = pl.tot + T_ITEMS + p;
    dc.heavy = part + 2;
    dc.heavy_cap = 0xffffffffu;
    dc.first = part;
    dc.npc = part + 1;
    dc.nsplit = part + 3;
  }
  const int rs = radix_sort_pairs(ik0, iv0, ik1, iv1, nitems, 0, 32, wt ? *wt : ws, s, dc.items);
  const uint32_t *items = rs ? iv1 : iv0;
  pl.items[p] = items;
  SplatArgs sa = sa0;
  // the trace holds a record per item, the long lists served by channel three
  const size_t ntrace = static_cast<size_t>(nitems) + 2 * static_cast<size_t>(pl.nsplit[p]);
  if (c->trace && !pl.dc)
    sa.trace = ws.get_as<uint64_t>(pslot(p, P_TRACE), ntrace * 4);
  SplatArgs sn = sa;  // normal items follow the heavy ones
  if (sn.trace) sn.trace += 4 * static_cast<size_t>(nheavy);
  // normal items: after the long lists (with device counts the kernel skips them)
  const uint32_t *nit = pl.dc ? items : items + nheavy;
  const uint32_t nn = pl.dc ? nitems : nitems - nheavy;
  // long lists: the staged ones first, then the producer/consumer workgroups'
  // (device counts: capacities; the kernels take the partition from part)
  const uint32_t nstage = pl.dc ? pl.nstage[p] : std::min(pl.nstage[p], nheavy);
  const uint32_t npc = pl.dc ? (tn.heavy == 0xffffffffu ? 0u : tn.heavy_max) : nheavy - nstage;
  SplatArgs sp = sa;
  if (sp.trace) sp.trace += 4 * static_cast<size_t>(nstage);
  const uint32_t *pit = pl.dc ? items : items + nstage;
  if (nstage > 0 || (npc > 0 && fork)) {
    // heavy blocks run on second streams, beside the normal items: the
    // producer/consumer workgroups at once, the staged chains after the
    // operand kernel (on s, across the GPU)
    if (npc > 0 && fork) {
      ensure_aux(c, set);
      hip_check(hipEventRecord(X.fork, s), "record fork");
      hip_check(hipStreamWaitEvent(X.aux, X.fork, 0), "wait fork");
      mark(c, p == 0 ? 8 : 10, true, X.aux);
      launch_splat_pc(p == 1, pit, npc, sp, X.aux, dc);
      mark(c, p == 0 ? 8 : 10, false, X.aux);
      hip_check(hipEventRecord(X.join, X.aux), "record join");
    } else if (npc > 0) {
      mark(c, p == 0 ? 8 : 10, true, s);
      launch_splat_pc(p == 1, pit, npc, sp, s, dc);
      mark(c, p == 0 ? 8 : 10, false, s);
    }
    if (nstage > 0) {
      // the operand kernel and then the chains on a third stream, beside the
      // normal items
      const bool add = p == 1;
      uint32_t *cb = ws.get_as<uint32_t>(pslot(p, P_SBASE), nstage + 1);
      uint32_t *cbt = ws.get_as<uint32_t>(pslot(p, P_SBAT), nstage + 1);
      float4 *ops = ws.get_as<float4>(pslot(p, P_SOPS), stage_floats(add, pl.sgroups[p]) / 4);
      float *qv = add ? nullptr
                      : ws.get_as<float>(pslot(p, P_SQ), stage_q_floats(add, pl.sgroups[p]));
      // the kStagePhase1 longest lists first: their chains start after their
      // own operands; the other lists' operands and chains follow on a fourth
      // stream, beside those chains
      const uint32_t n1 = std::min(nstage, kStagePhase1);
      ensure_aux3(c, set);
      hip_check(hipEventRecord(X.fork3, s), "record fork");
      hip_check(hipStreamWaitEvent(X.aux3, X.fork3, 0), "wait fork");
      mark(c, p == 0 ? 11 : 12, true, X.aux3);
      launch_stage_base(items, nstage, cb, cbt, sa, X.aux3, part);
      launch_stage_ops(add, items, 0, n1, pl.sgroups[p], cb, cbt, ops, qv, sa, X.aux3, part);
      if (n1 < nstage) {
        hip_check(hipEventRecord(X.fork4, X.aux3), "record fork");
        hip_check(hipStreamWaitEvent(X.aux4, X.fork4, 0), "wait fork");
        launch_stage_ops(add, items, n1, nstage, pl.sgroups[p], cb, cbt, ops, qv, sa, X.aux4,
                         part);
        launch_stage_chain(add, items, n1, nstage, cb, ops, qv, sa, X.aux4, part);
        hip_check(hipEventRecord(X.join4, X.aux4), "record join");
      }
      launch_stage_chain(add, items, 0, n1, cb, ops, qv, sa, X.aux3, part);
      mark(c, p == 0 ? 11 : 12, false, X.aux3);
      hip_check(hipEventRecord(X.join3, X.aux3), "record join");
    }
    // The normal items fill every CU; the heavy workgroups (the critical
    // path) must get their slots first, so the normal launch waits briefly.
    launch_delay(40, s);
    if (placed) hip_check(hipEventRecord(placed, s), "record placed");
    mark(c, p == 0 ? 7 : 9, true, s);
    launch_splat8(p == 1, tn.variant, nit, nn, sn, s, dc, tn.ppw, pl.nsplit[p]);
    mark(c, p == 0 ? 7 : 9, false, s);
    if (npc > 0 && fork) hip_check(hipStreamWaitEvent(s, X.join, 0), "wait join");
    if (nstage > 0) hip_check(hipStreamWaitEvent(s, X.join3, 0), "wait join");
    if (nstage > kStagePhase1) hip_check(hipStreamWaitEvent(s, X.join4, 0), "wait join");
  } else if (npc > 0) {
    mark(c, p == 0 ? 8 : 10, true, s);
    launch_splat_pc(p == 1, pit, npc, sp, s, dc);
    mark(c, p == 0 ? 8 : 10, false, s);
    if (placed) hip_check(hipEventRecord(placed, s), "record placed");
    mark(c, p == 0 ? 7 : 9, true, s);
    launch_splat8(p == 1, tn.variant, nit, nn, sn, s, dc, tn.ppw, pl.nsplit[p]);
    mark(c, p == 0 ? 7 : 9, false, s);
  } else {
    if (placed) hip_check(hipEventRecord(placed, s), "record placed");
    mark(c, p == 0 ? 7 : 9, true, s);
    launch_splat8(p == 1, tn.variant, items, nitems, sn, s, dc, tn.ppw, pl.nsplit[p]);
    mark(c, p == 0 ? 7 : 9, false, s);
  }
  mark(c, p == 0 ? 3 : 6, false, s);
  if (c->trace && !pl.dc) {
    std::vector<uint64_t> &tv = c->trace_items[p];
    tv.resize(ntrace * 4);
    if (nitems)
      hip_check(hipMemcpyAsync(tv.data(), sa.trace, tv.size() * 8, hipMemcpyDeviceToHost, s),
                "D2H trace");
    hip_check(hipStreamSynchronize(s), "sync trace");
  }
}

// No view of the plan has particles in both passes (the passes write
// disjoint images).
bool disjoint_passes(const Plan &pl) {
  for (int j = 0; j < pl.nj; ++j)
    if (pl.mo[0][j + 1] > pl.mo[0][j] && pl.mo[1][j + 1] > pl.mo[1][j]) return false;
  return true;
}

// Read n consecutive device totals into c->h_tot (one sync).
void read_totals(splhip_ctx *c, const uint32_t *d, int n, hipStream_t s) {
  hip_check(hipMemcpyAsync(c->h_tot, d, n * 4, hipMemcpyDeviceToHost, s), "read totals");
  hip_check(hipStreamSynchronize(s), "sync totals");
}

}  // namespace

struct FwdState {
  splhip_particles P;
  splhip_camera cam;
  splhip_params prm;
  Plan pl;
  Tuning tn;
};

namespace {

void drop_fwd(splhip_ctx *c) {
  delete c->fwd;
  c->fwd = nullptr;
}

// A single-view pass whose bin entries exceed the chunk size: consecutive
// ranges of its render order are binned and splatted one after another, each
// continuing every pixel's chain from the value the previous range left in
// the image (accumulate mode 2), which is the unchunked chain bit for bit.
// An additive pass that adds to an ordered pass renders its chains into a
// scratch image first and adds it at the end (pic = P_ordered + P_additive).
void render_chunked_pass(splhip_ctx *c, Plan &pl, int p, const Tuning &tn, uint32_t *tot,
                         hipStream_t s) {
  Workspace &ws = c->ws;
  JobSplat &J = pl.js[0];
  const JobSplat saved = J;
  const uint32_t m = pl.m[p];
  const Rec *recs = pl.recs[p];
  uint32_t *cnt = ws.get_as<uint32_t>(pslot(p, P_CNT), m);
  uint32_t *off = ws.get_as<uint32_t>(pslot(p, P_OFF), m);
  uint32_t *cnt16 = pl.r16[p] ? ws.get_as<uint32_t>(pslot(p, P_CNT16), m) : nullptr;
  uint32_t *off16 = pl.r16[p] ? ws.get_as<uint32_t>(pslot(p, P_OFF16), m) : nullptr;
  const int64_t npix = static_cast<int64_t>(J.xres) * J.yres * 3;
  const bool scratch = p == 1 && saved.accumulate[1] == 1;
  float *target = scratch ? ws.get_as<float>(S_SCRATCH, static_cast<size_t>(npix)) : saved.pic;
  uint64_t *e64 = reinterpret_cast<uint64_t *>(tot + T_E64 + 2 * p);
  uint32_t s0 = 0;
  for (int chunk = 0; s0 < m; ++chunk) {
    uint32_t n = m - s0;
    uint32_t e = 0;
    uint32_t e16 = 0;
    uint64_t ec_last = 0;
    for (;;) {  // shrink the range until its entries fit a chunk
      scan_u32(cnt + s0, off, n, tot + T_E + p, ws, s);
      launch_sum_u64(cnt + s0, n, e64, s);
      if (pl.r16[p]) scan_u32(cnt16 + s0, off16, n, tot + T_E16 + p, ws, s);
      read_totals(c, tot, T_N, s);
      uint64_t ec;
      std::memcpy(&ec, c->h_tot + T_E64 + 2 * p, 8);
      ec_last = ec;
      e = c->h_tot[T_E + p];
      e16 = c->h_tot[T_E16 + p];
      if (ec <= tn.chunk || n == 1) break;
      n = static_cast<uint32_t>(std::max<uint64_t>(1, static_cast<uint64_t>(n) * tn.chunk / ec * 15 / 16));
    }
    J.pic = target;
    J.accumulate[p] = chunk == 0 ? (scratch ? 0 : saved.accumulate[p]) : 2;
    J.need_empty[p] = chunk == 0 ? (scratch ? 1 : saved.need_empty[p]) : 0;
    char *h = staging(c, sizeof(JobSplat));
    std::memcpy(h, &J, sizeof(JobSplat));
    hip_check(hipMemcpyAsync(pl.d_js, h, sizeof(JobSplat), hipMemcpyHostToDevice, s),
              "upload chunk flags");
    staged(c, s);
    // this range as the pass: render positions are local to the range
    pl.m[p] = n;
    pl.e[p] = e;
    pl.etot[p] = ec_last;
    pl.e16[p] = e16;
    pl.recs[p] = recs + s0;
    if (e > 0) {
      stage_bins(c, pl, p, s);
      stage_items_count(c, pl, p, tn, tot, s);
      read_totals(c, tot + T_ITEMS, T_N - T_ITEMS, s);
      take_items(c, pl, p, tn);
      stage_splat(c, pl, p, tn, s);
    } else if (chunk == 0 && J.need_empty[p]) {
      launch_fill(target, 0.f, npix, s);  // no entries at all: the pass leaves zeros
    }
    s0 += n;
  }
  if (scratch) launch_add_images(saved.pic, target, npix, s);
  J = saved;
  char *h = staging(c, sizeof(JobSplat));
  std::memcpy(h, &J, sizeof(JobSplat));
  hip_check(hipMemcpyAsync(pl.d_js, h, sizeof(JobSplat), hipMemcpyHostToDevice, s),
            "restore flags");
  staged(c, s);
  pl.m[p] = m;
  pl.recs[p] = recs;
}

// Autotuner, inner part (8 px work regions): times the splat of every pass
// that can be repeated (no view adds to an earlier pass) on the frame's own
// bins under candidate decompositions, one knob at a time: the cap on
// producer/consumer workgroups (hip_heavy_max, 0 = none), then the list
// length from which a block is split per channel (hip_split), then the 8x8
// splat's inner loop (hip_variant 3: software pipelined; 4: chunks of hip_ppw
// particles).  All candidates give the same images.
Tuning autotune(splhip_ctx *c, Plan &pl, const Tuning &base, uint32_t *tot, hipStream_t s) {
  bool repeatable[2];
  for (int p = 0; p < 2; ++p) {
    repeatable[p] = pl.m[p] > 0;
    for (int j = 0; j < pl.nj; ++j)
      if (pl.js[j].accumulate[p] != 0) repeatable[p] = false;
  }
  if (!repeatable[0] && !repeatable[1]) return base;
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  auto time_of = [&](const Tuning &t) {
    float ms = 0.f;
    for (int p = 0; p < 2; ++p) {
      if (!repeatable[p]) continue;
      stage_items_count(c, pl, p, t, tot, s);
      read_totals(c, tot + T_ITEMS, T_N - T_ITEMS, s);
      take_items(c, pl, p, t);
      hip_check(hipEventRecord(e0, s), "hipEventRecord");
      stage_splat(c, pl, p, t, s);
      hip_check(hipEventRecord(e1, s), "hipEventRecord");
      hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
      float x = 0.f;
      hip_check(hipEventElapsedTime(&x, e0, e1), "hipEventElapsedTime");
      ms += x;
    }
    return ms;
  };
  Tuning best = base;
  best.variant = 4;
  best.ppw = 0;
  float best_ms = 1e30f;
  for (uint32_t hmax : {0u, 96u, 192u, 256u}) {
    Tuning t = best;
    t.heavy_max = hmax == 0 ? 1u : hmax;
    t.heavy = hmax == 0 ? 0xffffffffu : base.heavy;
    if (hmax != 0 && t.heavy == 0xffffffffu) t.heavy = 32768u;
    const float ms = time_of(t);
    if (ms < best_ms) { best_ms = ms; best = t; }
  }
  const Tuning b1 = best;
  for (uint32_t split : {4096u, 8192u, 32768u}) {
    Tuning t = b1;
    t.split = split;
    const float ms = time_of(t);
    if (ms < best_ms) { best_ms = ms; best = t; }
  }
  const Tuning b2 = best;
  const int cand[6][2] = {{4, 0}, {4, 1}, {4, 2}, {4, 4}, {4, 8}, {3, 0}};
  for (const auto &vp : cand) {
    Tuning t = b2;
    t.variant = vp[0];
    t.ppw = vp[1];
    const float ms = time_of(t);
    if (ms < best_ms) { best_ms = ms; best = t; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

// The frame pipeline over a batch of jobs; three host syncs per batch (pass
// sizes, bin entries, work items).
// A choice of the autotuner: work region and the 8x8 splat's decomposition.
struct Choice {
  int G;
  Tuning tn;
};

// force: a tuning chosen by the autotuner (null: the params' keys; with
// hip_autotune and 8 px regions the splat's decomposition is then swept on
// this frame, and the choice left in c->inner_choice).
int render_jobs(splhip_ctx *c, int nj, const JobIn *jobs, hipStream_t s, bool keep_fwd,
                const Choice *force) {
  const int G = force ? force->G : region_size(jobs[0].prm);
  if (G != 8 && G != 16 && G != 32 && G != 64)
    return fail(c, SPLHIP_E_ARG, "hip_region must be 8, 16, 32 or 64");
  drop_fwd(c);
  for (int k = 0; k < SPLHIP_NSTAGES; ++k) c->used[k] = false;
  c->nmarks = 0;
  c->last_entries[0] = c->last_entries[1] = c->last_active[0] = c->last_active[1] = 0;
  c->trace_items[0].clear();
  c->trace_items[1].clear();
  Tuning tn = force ? force->tn : tuning_of(jobs[0].prm);
  Plan pl;
  make_plan(c, nj, jobs, G, pl, s);
  Workspace &ws = c->ws;
  Rec *rec = ws.get_as<Rec>(S_REC, static_cast<size_t>(std::max<int64_t>(pl.ntot, 1)));
  uint8_t *act = ws.get_as<uint8_t>(S_ACT, static_cast<size_t>(std::max<int64_t>(pl.ntot, 1)));
  uint32_t *tot = ws.get_as<uint32_t>(S_TOTALS, kTotalsWords);
  stage_project_compact(c, pl, rec, act, tot, s);
  upload_job_flags(c, pl, s);
  for (int p = 0; p < 2; ++p)
    if (pl.m[p]) {
      stage_order(c, pl, p, rec, act, s);
      stage_count(c, pl, p, rec, tot, s);
    }
  bool chunked[2] = {false, false};
  if (pl.m[0] || pl.m[1]) {
    read_totals(c, tot + T_E, T_N - T_E, s);  // sync 2
    uint64_t e64[2];
    for (int p = 0; p < 2; ++p) {
      std::memcpy(&e64[p], c->h_tot + (T_E64 - T_E) + 2 * p, 8);
      pl.e[p] = pl.m[p] ? c->h_tot[p] : 0;
      pl.e16[p] = pl.m[p] ? c->h_tot[(T_E16 - T_E) + p] : 0;
      chunked[p] = pl.m[p] && e64[p] > tn.chunk;
    }
    // A single unchunked view with both passes: the additive splat runs
    // beside the ordered pass on a second stream, where its long lists are
    // off the critical path and staged operands' HBM traffic would slow the
    // ordered pass's binning: no staging unless hip_stage asks for it.
    if (nj == 1 && pl.m[0] && pl.m[1] && !chunked[0] && !chunked[1] &&
        jobs[0].prm->hip_stage == 0)
      tn.stage = 0xffffffffu;
    for (int p = 0; p < 2; ++p) pl.etot[p] = e64[p];
    if (chunked[0] || chunked[1]) {
      c->split_hint = nj > 1;  // the caller may render the views in smaller groups
      if (nj != 1 || G != 8)
        return fail(c, SPLHIP_E_CAPACITY,
                    "a batched (or 16-64 px region) pass exceeds hip_chunk_entries bin entries; "
                    "render its views one at a time");
      for (int p = 0; p < 2; ++p) {
        if (!pl.m[p]) continue;
        if (chunked[p]) {
          render_chunked_pass(c, pl, p, tn, tot, s);
          c->last_entries[p] = static_cast<int64_t>(e64[p]);
          continue;
        }
        stage_bins(c, pl, p, s);
        stage_items_count(c, pl, p, tn, tot, s);
        read_totals(c, tot + T_ITEMS, T_N - T_ITEMS, s);
        take_items(c, pl, p, tn);
        stage_splat(c, pl, p, tn, s);
      }
    }
  }
  bool tuned = false;
  // A single view with both passes (no autotuning): the additive pass is
  // binned on a second stream while the ordered pass bins (memory-bound),
  // and splatted there, into a cleared scratch image, while the ordered pass
  // splats; the scratch image is then added to the image: pic + l per pixel,
  // the additive pass's own update (pixels it does not touch add +0, which
  // leaves any value but -0 as it is; the ordered pass never writes -0).
  // Its longest lists stay in the 8x8 kernel: producer/consumer workgroups
  // would hold whole CUs beside the ordered pass's.
  const bool sweep = !force && G == 8 && jobs[0].prm->hip_autotune;
  const bool early = !chunked[0] && !chunked[1] && nj == 1 && G == 8 && pl.m[0] && pl.m[1] &&
                     pl.js[0].accumulate[1] == 1 && !sweep;
  // A batch in which no view has particles of both modes (the passes write
  // disjoint images): the additive pass bins and splats on a second stream,
  // with its own temporaries and auxiliary streams, beside the ordered pass;
  // the ordered splat is launched first (its long lists are the critical
  // path).
  const bool side = !early && !sweep && !chunked[0] && !chunked[1] && G == 8 && pl.m[0] &&
                    pl.m[1] && disjoint_passes(pl);
  if (side) {
    ensure_aux2(c);
    hip_check(hipEventRecord(c->ev_fork2, s), "record fork");
    hip_check(hipStreamWaitEvent(c->aux2, c->ev_fork2, 0), "wait fork");
    stage_bins(c, pl, 1, c->aux2, &c->ws2);
    stage_items_count(c, pl, 1, tn, tot, c->aux2, &c->ws2);
    stage_bins(c, pl, 0, s);
    stage_items_count(c, pl, 0, tn, tot, s);
    read_totals(c, tot + T_ITEMS, T_N - T_ITEMS, s);  // waits for stream s only
    take_items(c, pl, 0, tn);
    stage_splat(c, pl, 0, tn, s);
    read_totals(c, tot + T_ITEMS, T_N - T_ITEMS, c->aux2);  // the second stream only
    take_items(c, pl, 1, tn);
    stage_splat(c, pl, 1, tn, c->aux2, &c->ws2, nullptr, true, 1);
    hip_check(hipEventRecord(c->ev_join2, c->aux2), "record join");
    hip_check(hipStreamWaitEvent(s, c->ev_join2, 0), "wait join");
  } else if (early) {
    Tuning ta = tn;
    ta.heavy = 0xffffffffu;
    ta.heavy_max = 1;
    const JobSplat &J0 = pl.js[0];
    const int64_t npix = static_cast<int64_t>(J0.xres) * J0.yres * 3;
    float *scratch = ws.get_as<float>(S_SCRATCH, static_cast<size_t>(npix));
    JobSplat *d_alt = ws.get_as<JobSplat>(S_JOBS2, 1);
    // The additive pass bins on the second stream, beside the ordered pass's
    // binning (its sort and scan temporaries in the second workspace), and
    // its counts are read as soon as that stream has them.
    ensure_aux2(c);
    hip_check(hipEventRecord(c->ev_fork2, s), "record fork");
    hip_check(hipStreamWaitEvent(c->aux2, c->ev_fork2, 0), "wait fork");
    stage_bins(c, pl, 1, c->aux2, &c->ws2);
    stage_items_count(c, pl, 1, ta, tot, c->aux2, &c->ws2);
    stage_bins(c, pl, 0, s);
    stage_items_count(c, pl, 0, tn, tot, s);
    read_totals(c, tot + T_ITEMS, T_N - T_ITEMS, c->aux2);  // waits for the second stream only
    take_items(c, pl, 1, ta);
    // same flags as the items were counted with, but written to the scratch
    // image (blocks without items stay zero)
    JobSplat alt = J0;
    alt.pic = scratch;
    alt.accumulate[1] = 0;
    char *h = staging(c, sizeof(JobSplat));
    std::memcpy(h, &alt, sizeof(JobSplat));
    hip_check(hipMemcpyAsync(d_alt, h, sizeof(JobSplat), hipMemcpyHostToDevice, c->aux2),
              "upload scratch job");
    staged(c, c->aux2);
    launch_fill(scratch, 0.f, npix, c->aux2);
    read_totals(c, tot + T_ITEMS, T_N - T_ITEMS, s);  // waits for stream s only
    take_items(c, pl, 0, tn);
    // The ordered splat first: a producer/consumer workgroup needs a whole
    // CU, which it gets at once only before the 8x8 grids fill every CU, so
    // the additive splat (a full grid) waits for those workgroups' head
    // start.  Launched before them, it won that race or lost it depending on
    // launch timing (33 or 55 ms for the 16M frame's longest list).
    stage_splat(c, pl, 0, tn, s, nullptr, nullptr, true, 0, c->ev_placed);
    hip_check(hipStreamWaitEvent(c->aux2, c->ev_placed, 0), "wait placed");
    stage_splat(c, pl, 1, ta, c->aux2, &c->ws2, d_alt, false);
    hip_check(hipEventRecord(c->ev_join2, c->aux2), "record join");
    hip_check(hipStreamWaitEvent(s, c->ev_join2, 0), "wait join");
    launch_add_images(J0.pic, scratch, npix, s);
  } else if ((pl.m[0] || pl.m[1]) && !chunked[0] && !chunked[1]) {
    for (int p = 0; p < 2; ++p)
      if (pl.m[p]) stage_bins(c, pl, p, s);
    if (sweep) {
      tn = autotune(c, pl, tn, tot, s);
      const uint32_t ic[5] = {static_cast<uint32_t>(tn.variant), static_cast<uint32_t>(tn.ppw),
                              tn.split, tn.heavy, tn.heavy_max};
      std::memcpy(c->inner_choice, ic, sizeof(ic));
      tuned = true;
    }
    for (int p = 0; p < 2; ++p)
      if (pl.m[p] && G == 8) stage_items_count(c, pl, p, tn, tot, s);
    if (G == 8) {
      read_totals(c, tot + T_ITEMS, T_N - T_ITEMS, s);  // sync 3
      for (int p = 0; p < 2; ++p) take_items(c, pl, p, tn);
    }
    // A single view with both passes: the additive pass's splat renders into
    // a scratch image on a second stream, beside the ordered splat, and is
    // then added to the image: pic + l per pixel, as the pass's add mode
    // (pixels it does not touch add +0, which leaves any value but -0 as it
    // is; the ordered pass never writes -0).
    const bool overlap = nj == 1 && G == 8 && pl.m[0] && pl.m[1] &&
                         pl.js[0].accumulate[1] == 1;
    if (overlap) {
      const JobSplat &J0 = pl.js[0];
      const int64_t npix = static_cast<int64_t>(J0.xres) * J0.yres * 3;
      float *scratch = ws.get_as<float>(S_SCRATCH, static_cast<size_t>(npix));
      JobSplat *d_alt = ws.get_as<JobSplat>(S_JOBS2, 1);
      // same flags as the items were counted with, but written to a cleared
      // scratch image (blocks without items stay zero)
      JobSplat alt = J0;
      alt.pic = scratch;
      alt.accumulate[1] = 0;
      char *h = staging(c, sizeof(JobSplat));
      std::memcpy(h, &alt, sizeof(JobSplat));
      hip_check(hipMemcpyAsync(d_alt, h, sizeof(JobSplat), hipMemcpyHostToDevice, s),
                "upload scratch job");
      staged(c, s);
      ensure_aux2(c);
      hip_check(hipEventRecord(c->ev_fork2, s), "record fork");
      stage_splat(c, pl, 0, tn, s);
      hip_check(hipStreamWaitEvent(c->aux2, c->ev_fork2, 0), "wait fork");
      launch_fill(scratch, 0.f, npix, c->aux2);
      stage_splat(c, pl, 1, tn, c->aux2, &c->ws2, d_alt, false);
      hip_check(hipEventRecord(c->ev_join2, c->aux2), "record join");
      hip_check(hipStreamWaitEvent(s, c->ev_join2, 0), "wait join");
      launch_add_images(J0.pic, scratch, npix, s);
    } else if (G == 8 && pl.m[0] && pl.m[1] && disjoint_passes(pl)) {
      // No view has particles of both modes: the passes write disjoint
      // images, and the additive pass's splat runs on a second stream (its
      // own auxiliary streams and temporaries) beside the ordered pass's.
      ensure_aux2(c);
      hip_check(hipEventRecord(c->ev_fork2, s), "record fork");
      hip_check(hipStreamWaitEvent(c->aux2, c->ev_fork2, 0), "wait fork");
      stage_splat(c, pl, 0, tn, s);
      stage_splat(c, pl, 1, tn, c->aux2, &c->ws2, nullptr, true, 1);
      hip_check(hipEventRecord(c->ev_join2, c->aux2), "record join");
      hip_check(hipStreamWaitEvent(s, c->ev_join2, 0), "wait join");
    } else {
      for (int p = 0; p < 2; ++p)
        if (pl.m[p]) stage_splat(c, pl, p, tn, s);
    }
  }
  // jobs with nothing to splat in either pass
  for (int j = 0; j < nj; ++j) {
    const bool w = pl.mo[0][j + 1] > pl.mo[0][j] || pl.mo[1][j + 1] > pl.mo[1][j];
    if (!w)
      launch_fill(jobs[j].pic, 0.f, static_cast<int64_t>(jobs[j].cam->xres) * jobs[j].cam->yres * 3, s);
  }
  for (int p = 0; p < 2; ++p) {
    if (!chunked[p]) c->last_entries[p] = pl.e[p];
    c->last_active[p] = pl.m[p];
  }
  collect_marks(c);
  c->last_tuning[0] = static_cast<uint32_t>(tn.variant);
  c->last_tuning[1] = tn.split;
  c->last_tuning[2] = tn.heavy == 0xffffffffu ? 0u : tn.heavy_max;
  c->last_tuning[3] = (tuned || force) ? 1u : 0u;
  c->last_tuning[4] = static_cast<uint32_t>(G);
  c->last_tuning[5] = static_cast<uint32_t>(tn.ppw);
  c->last_regions[0] = static_cast<int64_t>(pl.used16[0]);
  c->last_regions[1] = static_cast<int64_t>(pl.used16[1]);
  if (keep_fwd && nj == 1 && G == 8 && !chunked[0] && !chunked[1]) {
    c->fwd = new FwdState{*jobs[0].P, *jobs[0].cam, *jobs[0].prm, pl, tn};
  }
  return 0;
}

// Autotuner, outer part: on the first render of a batch bucket (views,
// particles and pixels each by factors of sqrt(2) or 2, the pass mix and the
// reference tile size), the 8 px work regions with the inner sweep of the
// splat's decomposition, then 16, 32 and 64 px regions, each timed over one
// whole render; the fastest is reused for the bucket.  Every candidate gives
// the same images.  Renders that keep the backward's state (keep_fwd) and
// renders with hip_region set keep their work region.
uint64_t bucket_of(int nj, const JobIn *jobs) {
  int64_t n = 0;
  double px = 0.0;
  bool mode[2] = {false, false};
  for (int j = 0; j < nj; ++j) {
    n += jobs[j].P->n;
    px += static_cast<double>(jobs[j].cam->xres) * jobs[j].cam->yres;
    for (int t = 0; t < jobs[j].prm->ptypes; ++t) {
      if (jobs[j].prm->a_eq_e[t] != 1) mode[0] = true;
      if (jobs[j].prm->a_eq_e[t] != 0) mode[1] = true;
    }
  }
  const uint64_t bn = static_cast<uint64_t>(std::floor(2.0 * std::log2(std::max<double>(n, 1))));
  const uint64_t bp = static_cast<uint64_t>(std::floor(2.0 * std::log2(std::max(px, 1.0))));
  const uint64_t bj = static_cast<uint64_t>(std::floor(std::log2(static_cast<double>(nj))));
  const uint64_t T = static_cast<uint64_t>(jobs[0].prm->tile_size) & 0xfff;
  return bn | (bp << 8) | (bj << 16) | (static_cast<uint64_t>(mode[0]) << 24) |
         (static_cast<uint64_t>(mode[1]) << 25) | (T << 26);
}

int render_tuned(splhip_ctx *c, int nj, const JobIn *jobs, hipStream_t s, bool keep_fwd,
                 bool region8 = false) {
  const splhip_params *prm = jobs[0].prm;
  if (!prm->hip_autotune) return render_jobs(c, nj, jobs, s, keep_fwd, nullptr);
  const uint64_t key = bucket_of(nj, jobs) | ((keep_fwd || region8) ? uint64_t(1) << 40 : 0) |
                       (static_cast<uint64_t>(prm->hip_region & 0x7f) << 41);
  auto decode = [](const std::vector<uint32_t> &v, Choice &ch) {
    ch.G = static_cast<int>(v[0]);
    ch.tn.variant = static_cast<int>(v[1]);
    ch.tn.ppw = static_cast<int>(v[2]);
    ch.tn.split = v[3];
    ch.tn.heavy = v[4];
    ch.tn.heavy_max = v[5];
  };
  auto it = c->tuned.find(key);
  if (it != c->tuned.end()) {
    Choice ch{8, tuning_of(prm)};
    decode(it->second, ch);
    return render_jobs(c, nj, jobs, s, keep_fwd, &ch);
  }
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  auto timed = [&](const Choice *ch, float &ms) {
    hip_check(hipEventRecord(e0, s), "hipEventRecord");
    const int rc = render_jobs(c, nj, jobs, s, false, ch);
    hip_check(hipEventRecord(e1, s), "hipEventRecord");
    hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    return rc;
  };
  const Tuning base = tuning_of(prm);
  const int G0 = region_size(prm);
  Choice best{G0, base};
  float best_ms = 1e30f, ms = 0.f;
  int rc = 0;
  if (G0 == 8) {  // the inner sweep on this frame, then its choice timed over a render
    const uint32_t ic[5] = {static_cast<uint32_t>(base.variant), static_cast<uint32_t>(base.ppw),
                            base.split, base.heavy, base.heavy_max};
    std::memcpy(c->inner_choice, ic, sizeof(ic));
    rc = render_jobs(c, nj, jobs, s, false, nullptr);
    best.tn.variant = static_cast<int>(c->inner_choice[0]);
    best.tn.ppw = static_cast<int>(c->inner_choice[1]);
    best.tn.split = c->inner_choice[2];
    best.tn.heavy = c->inner_choice[3];
    best.tn.heavy_max = c->inner_choice[4];
  }
  if (rc == 0) rc = timed(&best, best_ms);
  if (rc == 0 && !keep_fwd && !region8 && prm->hip_region == 0) {
    for (int G : {16, 32, 64}) {
      const Choice ch{G, base};
      if (timed(&ch, ms) == 0 && ms < best_ms) {
        best_ms = ms;
        best = ch;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  c->tuned[key] = {static_cast<uint32_t>(best.G), static_cast<uint32_t>(best.tn.variant),
                   static_cast<uint32_t>(best.tn.ppw), best.tn.split, best.tn.heavy,
                   best.tn.heavy_max};
  return render_jobs(c, nj, jobs, s, keep_fwd, &best);
}

// A batch in groups of consecutive views of at most kGroupParticles
// particles (a larger view alone), each group one pipeline; a group with a
// pass past the chunk size (hip_chunk_entries) is halved until its views
// fit or render alone, chunked.  Every image equals its view's plain render.
constexpr int64_t kGroupParticles = int64_t(1) << 28;

int render_groups(splhip_ctx *c, int nj, const JobIn *jobs, hipStream_t s) {
  int j0 = 0;
  while (j0 < nj) {
    int j1 = j0 + 1;
    int64_t n = jobs[j0].P->n;
    while (j1 < nj && n + jobs[j1].P->n <= kGroupParticles) n += jobs[j1++].P->n;
    std::vector<std::pair<int, int>> todo{{j0, j1}};
    while (!todo.empty()) {
      const auto g = todo.back();
      todo.pop_back();
      c->split_hint = false;
      const int rc = render_tuned(c, g.second - g.first, jobs + g.first, s, false);
      if (rc == SPLHIP_E_CAPACITY && c->split_hint && g.second - g.first > 1) {
        const int mid = (g.first + g.second) / 2;
        todo.push_back({mid, g.second});
        todo.push_back({g.first, mid});
        continue;
      }
      if (rc) return rc;
    }
    j0 = j1;
  }
  return 0;
}

// ------------------------------------------------------------------ capture

}  // namespace

// A batch reserved for rendering with device counts (no host waits, no
// allocation), e.g. under stream capture: its device tables, the binning
// form and entry capacities of each pass, the tuning, and a workspace that
// only this batch uses and that is frozen once sized.
struct Reservation {
  std::vector<char> tables;  // host copy of the tables (a captured batch must match)
  size_t o[7] = {};
  char *d_tables = nullptr;  // in ws (S_JOBS)
  int nj = 0;
  bool use16[2] = {false, false};
  uint32_t ecap[2] = {0, 0}, e16cap[2] = {0, 0};
  uint32_t ns_cap[2] = {0, 0};        // staged lists (0: no staging)
  uint64_t groups_cap[2] = {0, 0};    // and their groups of 4 entries
  Tuning tn{};
  Workspace ws;
};

namespace {

// The frame pipeline with device counts over a reserved batch: every launch
// is sized from the reservation's capacities and reads the counts from the
// device, the job flags and empty images are set by kernels, and nothing
// waits on the host, allocates or copies from the host.  A pass whose
// entries exceed its capacity is dropped and flagged (splhip_capture_status).
int render_jobs_dc(splhip_ctx *c, int nj, const JobIn *jobs, hipStream_t s) {
  Reservation &rv = *c->rv;
  if (region_size(jobs[0].prm) != 8)
    return fail(c, SPLHIP_E_ARG, "device-count renders use 8 px work regions");
  if (c->pal_dirty)
    return fail(c, SPLHIP_E_STATE, "palettes changed since the batch was reserved");
  Plan pl;
  build_plan(c, nj, jobs, 8, pl);
  std::vector<char> tab;
  size_t o[7];
  plan_tables(pl, &tab, o);
  if (nj != rv.nj || tab != rv.tables)
    return fail(c, SPLHIP_E_STATE,
                "the batch differs from the one reserved by splhip_capture_reserve");
  plan_bind(pl, rv.d_tables, o);
  pl.ws = &rv.ws;
  pl.dc = true;
  Workspace &ws = rv.ws;
  const Tuning tn = rv.tn;
  const int prof = c->prof;
  c->prof = 0;  // no stage events (they would need host waits)
  Rec *rec = ws.get_as<Rec>(S_REC, static_cast<size_t>(std::max<int64_t>(pl.ntot, 1)));
  uint8_t *act = ws.get_as<uint8_t>(S_ACT, static_cast<size_t>(std::max<int64_t>(pl.ntot, 1)));
  uint32_t *tot = ws.get_as<uint32_t>(S_TOTALS, kTotalsWords);
  pl.tot = tot;
  for (int p = 0; p < 2; ++p) {
    pl.use16[p] = rv.use16[p];
    pl.e[p] = rv.ecap[p];
    pl.e16[p] = rv.e16cap[p];
  }
  stage_project_compact(c, pl, rec, act, tot, s);
  launch_job_flags(pl.d_mo[0], pl.d_mo[1], nj, pl.d_js, s);
  for (int p = 0; p < 2; ++p) {
    pl.nstage[p] = rv.ns_cap[p];
    pl.sgroups[p] = rv.groups_cap[p];
  }
  for (int p = 0; p < 2; ++p)
    if (pl.m[p]) {
      stage_order(c, pl, p, rec, act, s);
      stage_count(c, pl, p, rec, tot, s);
    }
  for (int p = 0; p < 2; ++p)
    if (pl.m[p]) stage_bins(c, pl, p, s);
  for (int p = 0; p < 2; ++p)
    if (pl.m[p]) {
      stage_items_count(c, pl, p, tn, tot, s);
      pl.nitems[p] = 3u * pl.nunits;  // at most three items per unit
      pl.nheavy[p] = pl.nstage[p] + (tn.heavy == 0xffffffffu ? 0u : tn.heavy_max);
    }
  for (int p = 0; p < 2; ++p)
    if (pl.m[p]) stage_splat(c, pl, p, tn, s);
  launch_fill_empty_jobs(pl.d_mo[0], pl.d_mo[1], nj, pl.d_js, s);
  c->prof = prof;
  return 0;
}

bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
  return st == hipStreamCaptureStatusActive;
}

// Render the batch once with host waits, size the capacities from its counts
// (entries x headroom), then render it once with device counts into a fresh
// workspace, which is frozen afterwards.
int capture_reserve(splhip_ctx *c, int nj, const JobIn *jobs, float headroom, hipStream_t s) {
  if (region_size(jobs[0].prm) != 8)
    return fail(c, SPLHIP_E_ARG, "device-count renders use 8 px work regions");
  if (nj > 65535) return fail(c, SPLHIP_E_CAPACITY, "at most 65535 views per reserved batch");
  if (!(headroom >= 1.f)) return fail(c, SPLHIP_E_ARG, "headroom must be at least 1");
  delete c->rv;
  c->rv = nullptr;
  if (int rc = render_tuned(c, nj, jobs, s, false, true)) return rc;
  Tuning tn = tuning_of(jobs[0].prm);
  if (c->last_entries[0] > static_cast<int64_t>(tn.chunk) ||
      c->last_entries[1] > static_cast<int64_t>(tn.chunk))
    return fail(c, SPLHIP_E_CAPACITY,
                "a pass exceeds hip_chunk_entries bin entries: not renderable without host waits");
  tn.variant = static_cast<int>(c->last_tuning[0]);
  tn.ppw = static_cast<int>(c->last_tuning[5]);
  tn.split = c->last_tuning[1];
  if (c->last_tuning[2] == 0) {
    tn.heavy = 0xffffffffu;
    tn.heavy_max = 1;
  } else {
    tn.heavy_max = c->last_tuning[2];
  }
  Reservation *rv = new Reservation;
  rv->tn = tn;
  rv->nj = nj;
  auto cap = [&](int64_t v) {
    const double x = static_cast<double>(v) * headroom + 4096.0;
    return static_cast<uint32_t>(std::min(x, 4294967295.0 - 65536.0));
  };
  for (int p = 0; p < 2; ++p) {
    rv->use16[p] = c->last_regions[p] > 0;
    rv->ecap[p] = cap(c->last_entries[p]);
    rv->e16cap[p] = cap(c->last_regions[p]);
    // staging where the sizing render staged
    if (c->last_stage_n[p] > 0) {
      rv->ns_cap[p] = static_cast<uint32_t>(std::min(
          static_cast<double>(c->last_stage_n[p]) * headroom + 16.0, 1048576.0));
      const double g = static_cast<double>(c->last_stage_groups[p]) * headroom + 1024.0;
      const uint64_t gc = static_cast<uint64_t>(g / kStageLoad + 1) * kStageLoad;
      if (4 * (stage_floats(p == 1, gc) + stage_q_floats(p == 1, gc)) <= kStageBudget) {
        rv->groups_cap[p] = gc;
      } else {
        rv->ns_cap[p] = 0;
      }
    }
  }
  Plan pl;
  build_plan(c, nj, jobs, 8, pl);
  const size_t total = plan_tables(pl, &rv->tables, rv->o);
  rv->d_tables = rv->ws.get_as<char>(S_JOBS, total);
  hip_check(hipStreamSynchronize(s), "sync before the table upload");
  hip_check(hipMemcpy(rv->d_tables, rv->tables.data(), total, hipMemcpyHostToDevice),
            "upload reserved tables");
  ensure_aux(c);
  ensure_aux3(c);
  c->rv = rv;
  if (int rc = render_jobs_dc(c, nj, jobs, s)) {
    delete c->rv;
    c->rv = nullptr;
    return rc;
  }
  uint32_t ovf = 0;
  const uint32_t *d_tot = rv->ws.get_as<uint32_t>(S_TOTALS, kTotalsWords);
  hip_check(hipMemcpyAsync(c->h_tot, d_tot + T_OVF, 4, hipMemcpyDeviceToHost, s),
            "read overflow flags");
  hip_check(hipStreamSynchronize(s), "sync reserve");
  ovf = c->h_tot[0];
  rv->ws.freeze(true);
  if (ovf != 0) {
    delete c->rv;
    c->rv = nullptr;
    return fail(c, SPLHIP_E_CAPACITY, "reserved capacities overflowed on the sizing render");
  }
  return 0;
}

// Gradient of the last single-view render (see splhip_render_backward).
int backward_impl(splhip_ctx *c, const float *grad, const splhip_grads *out, hipStream_t s) {
  FwdState &F = *c->fwd;
  Plan &pl = F.pl;
  const int64_t n = pl.ntot;
  Workspace &ws = c->ws;
  double *gscr = ws.get_as<double>(S_GRAD0, static_cast<size_t>(std::max<int64_t>(n, 1)) * 6);
  launch_zero_f64(gscr, n * 6, s);
  uint32_t *tot = ws.get_as<uint32_t>(S_TOTALS, kTotalsWords);
  for (int p = 0; p < 2; ++p) {
    if (!pl.m[p]) continue;
    // work items: list segments of at most kAdjSeg entries
    const uint32_t nu = pl.nunits;
    uint32_t *cnt = ws.get_as<uint32_t>(S_GRAD1, 2 * (static_cast<size_t>(nu) + 1));
    uint32_t *soff = cnt + nu + 1;
    launch_seg_count(pl.bin_off[p], nu, cnt, s);
    scan_u32(cnt, soff, nu, soff + nu, ws, s);
    hip_check(hipMemcpyAsync(tot, soff + nu, 4, hipMemcpyDeviceToDevice, s), "segment total");
    hip_check(hipMemcpyAsync(tot + 1, pl.bin_off[p] + nu, 4, hipMemcpyDeviceToDevice, s),
              "entry total");
    read_totals(c, tot, 2, s);
    const uint32_t nseg = c->h_tot[0], ne = c->h_tot[1];
    if (nseg == 0) continue;
    uint32_t *seg = ws.get_as<uint32_t>(S_GRAD3, 2 * static_cast<size_t>(nseg));
    launch_seg_emit(pl.bin_off[p], nu, soff, seg, seg + nseg, s);
    double *state = p == 0 ? ws.get_as<double>(S_GRAD4, static_cast<size_t>(nseg) * 3 * 64 * 3)
                           : nullptr;
    // per-entry partials, then per particle their sum in a fixed order: the
    // entries sorted stably by particle (render position), summed in FP64
    float *part = ws.get_as<float>(S_GRAD2, static_cast<size_t>(ne) * 6);
    const SplatArgs sa = splat_args(c, pl, p, F.tn);
    launch_adjoint(p == 1, seg, seg + nseg, nseg, soff, nu, sa, grad, part, state, s);
    uint32_t *k0 = ws.get_as<uint32_t>(S_GRAD5, ne), *v0 = ws.get_as<uint32_t>(S_GRAD6, ne);
    uint32_t *k1 = ws.get_as<uint32_t>(S_GRAD7, ne), *v1 = ws.get_as<uint32_t>(S_GRAD8, ne);
    uint32_t *poff = ws.get_as<uint32_t>(S_GRAD9, static_cast<size_t>(pl.m[p]) + 1);
    hip_check(hipMemcpyAsync(k0, pl.ent[p], static_cast<size_t>(ne) * 4, hipMemcpyDeviceToDevice,
                             s), "entry particles");
    launch_iota(v0, ne, s);
    const int r = radix_sort_pairs(k0, v0, k1, v1, ne, 0, ceil_log2(pl.m[p]), ws, s);
    launch_bin_offsets(r ? k1 : k0, ne, pl.m[p], poff, s);
    launch_reduce_parts(poff, r ? v1 : v0, part, pl.order[p], pl.m[p], gscr, s);
  }
  const uint8_t *act = ws.get_as<uint8_t>(S_ACT, static_cast<size_t>(std::max<int64_t>(n, 1)));
  launch_chain(pl.d_pa, n, act, gscr, *out, s);
  return 0;
}

}  // namespace

extern "C" {

int splhip_exptable(float tab[2048], float consts[5]) {
  if (!tab || !consts) return SPLHIP_E_ARG;
  Consts k;
  host_exptable(tab, k);
  consts[0] = k.expfac;
  consts[1] = k.taylorlimit;
  consts[2] = k.h2sigma;
  consts[3] = k.bfak;
  consts[4] = k.rfac;
  return 0;
}

int splhip_create(int device, splhip_ctx **out) {
  if (!out) return SPLHIP_E_ARG;
  *out = nullptr;
  splhip_ctx *c = new (std::nothrow) splhip_ctx();
  if (!c) return SPLHIP_E_CAPACITY;
  try {
    int ndev = 0;
    hip_check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
    if (device < 0 || device >= ndev) {
      delete c;
      return SPLHIP_E_ARG;
    }
    hip_check(hipSetDevice(device), "hipSetDevice");
    c->device = device;
    std::vector<float> tab(2048);
    host_exptable(tab.data(), c->consts);
    hip_check(hipMalloc(&c->d_exptab, 2048 * sizeof(float)), "hipMalloc exptable");
    hip_check(hipMemcpy(c->d_exptab, tab.data(), 2048 * sizeof(float), hipMemcpyHostToDevice),
              "upload exptable");
    hip_check(hipHostMalloc(reinterpret_cast<void **>(&c->h_tot),
                            kTotalsWords * sizeof(uint32_t), hipHostMallocDefault),
              "hipHostMalloc");
    // the device address of this GPU's `scratch` (the symbol's name in host
    // code is not a device address)
    hip_check(hipGetSymbolAddress(reinterpret_cast<void **>(&c->scratch_ptr), HIP_SYMBOL(scratch)),
              "hipGetSymbolAddress scratch");
  } catch (const HipError &e) {
    delete c;
    return SPLHIP_E_HIP;
  }
  *out = c;
  return 0;
}

int splhip_set_profiling(splhip_ctx *c, int32_t enable) {
  if (!c) return SPLHIP_E_ARG;
  c->prof = enable ? 1 : 0;
  return 0;
}

int splhip_last_tuning(splhip_ctx *c, int64_t out[6]) {
  if (!c || !out) return SPLHIP_E_ARG;
  for (int k = 0; k < 6; ++k) out[k] = c->last_tuning[k];
  return 0;
}

int splhip_last_binning(splhip_ctx *c, int64_t out[2]) {
  if (!c || !out) return SPLHIP_E_ARG;
  out[0] = c->last_regions[0];
  out[1] = c->last_regions[1];
  return 0;
}

int splhip_set_trace(splhip_ctx *c, int32_t enable) {
  if (!c) return SPLHIP_E_ARG;
  c->trace = enable ? 1 : 0;
  return 0;
}

int splhip_trace_items(splhip_ctx *c, int32_t ordered, uint64_t *out, int64_t cap, int64_t *n) {
  if (!c || !n) return SPLHIP_E_ARG;
  const std::vector<uint64_t> &tv = c->trace_items[ordered ? 0 : 1];
  *n = static_cast<int64_t>(tv.size() / 4);
  if (out) std::copy(tv.begin(), tv.begin() + std::min<size_t>(tv.size(), cap * 4), out);
  return 0;
}

int splhip_stage_times(splhip_ctx *c, float *ms, int64_t *counts) {
  if (!c || !ms) return SPLHIP_E_ARG;
  for (int k = 0; k < SPLHIP_NSTAGES; ++k) ms[k] = c->used[k] ? c->stage_ms[k] : 0.f;
  if (counts) {
    counts[0] = c->last_active[0];
    counts[1] = c->last_entries[0];
    counts[2] = c->last_active[1];
    counts[3] = c->last_entries[1];
  }
  return 0;
}

void splhip_destroy(splhip_ctx *c) {
  if (!c) return;
  delete c->fwd;
  (void)hipSetDevice(c->device);
  (void)hipDeviceSynchronize();
  delete c->rv;
  for (auto &e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->ev_stage) (void)hipEventDestroy(c->ev_stage);
  if (c->h_stage) (void)hipHostFree(c->h_stage);
  for (int k = 0; k < 2; ++k) {
    if (c->h_ing[k]) (void)hipHostFree(c->h_ing[k]);
    if (c->ev_ing[k]) (void)hipEventDestroy(c->ev_ing[k]);
  }
  for (char *h : c->h_rec)
    if (h) (void)hipHostFree(h);
  for (hipEvent_t e : c->ev_rec)
    if (e) (void)hipEventDestroy(e);
  for (splhip_ctx::AuxSet &X : c->ax) {
    for (hipEvent_t e : {X.fork, X.join, X.fork3, X.join3, X.fork4, X.join4})
      if (e) (void)hipEventDestroy(e);
    for (hipStream_t t : {X.aux, X.aux3, X.aux4})
      if (t) (void)hipStreamDestroy(t);
  }
  if (c->ev_fork2) (void)hipEventDestroy(c->ev_fork2);
  if (c->ev_join2) (void)hipEventDestroy(c->ev_join2);
  if (c->ev_placed) (void)hipEventDestroy(c->ev_placed);
  if (c->aux2) (void)hipStreamDestroy(c->aux2);
  c->ws.release();
  if (c->d_exptab) (void)hipFree(c->d_exptab);
  if (c->d_palx) (void)hipFree(c->d_palx);
  if (c->d_palrgb) (void)hipFree(c->d_palrgb);
  if (c->h_tot) (void)hipHostFree(c->h_tot);
  delete c;
}

const char *splhip_last_error(const splhip_ctx *c) { return c ? c->err.c_str() : "null context"; }

int splhip_camera_setup(const double campos[3], const double lookat[3], const double sky_in[3],
                        const double centerpos[3], int32_t xres, int32_t yres, float fov,
                        float zmin, float zmax, int32_t projection, float minrad_pix,
                        splhip_camera *out) {
  if (!campos || !lookat || !sky_in || !centerpos || !out) return SPLHIP_E_ARG;
  // splotch_host.cc:62-109
  const V3 cam{campos[0], campos[1], campos[2]}, look{lookat[0], lookat[1], lookat[2]};
  const V3 cen{centerpos[0], centerpos[1], centerpos[2]};
  V3 sky{sky_in[0], sky_in[1], sky_in[2]};
  float zmaxval = zmax, zminval = zmin;
  zminval = std::max(0.f, zminval);
  const float ycorr = .5f * (yres - xres);
  const float res2 = 0.5f * xres;
  const double degr2rad = 3.141592653589793238462643383279502884197 / 180.0;
  const float fovfct = std::tan(fov * 0.5f * degr2rad);
  sky = vnorm(sky);
  const V3 zaxis = vnorm(vsub(cen, look));
  const V3 xaxis = vnorm(vcross(sky, zaxis));
  const V3 yaxis = vcross(zaxis, xaxis);
  M34 axes;
  const float ax[3][3] = {{static_cast<float>(xaxis.x), static_cast<float>(xaxis.y),
                           static_cast<float>(xaxis.z)},
                          {static_cast<float>(yaxis.x), static_cast<float>(yaxis.y),
                           static_cast<float>(yaxis.z)},
                          {static_cast<float>(zaxis.x), static_cast<float>(zaxis.y),
                           static_cast<float>(zaxis.z)}};
  // TRANSMAT ctor stores its arguments transposed (transform.cxx:28-48)
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k) axes.e[k][r] = ax[r][k];
  axes.e[0][3] = axes.e[1][3] = axes.e[2][3] = 0.f;
  M34 rot;
  if (!m_inverse(axes, rot)) return SPLHIP_E_DEGENERATE;
  M34 tr{};
  tr.e[0][0] = tr.e[1][1] = tr.e[2][2] = 1.f;
  tr.e[0][3] = static_cast<float>(-cam.x);
  tr.e[1][3] = static_cast<float>(-cam.y);
  tr.e[2][3] = static_cast<float>(-cam.z);
  m_mul_assign(tr, rot);
  std::memcpy(out->mat, &tr.e[0][0], sizeof(out->mat));
  const float *p = out->mat;
  const double tcx = cen.x * p[0] + cen.y * p[1] + cen.z * p[2] + p[3];
  const double tcy = cen.x * p[4] + cen.y * p[5] + cen.z * p[6] + p[7];
  const float dist = vlen(vsub(cen, look));
  const float xfac = res2 / (fovfct * dist);
  out->xshift = -xfac * tcx;
  out->yshift = -xfac * tcy;
  out->res2 = res2;
  out->ycorr = ycorr;
  out->xfac = xfac;
  out->fovfct = fovfct;
  out->dist = dist;
  out->zminval = zminval;
  out->zmaxval = zmaxval;
  out->minrad_pix = minrad_pix;
  out->xres = xres;
  out->yres = yres;
  out->projection = projection ? 1 : 0;
  return 0;
}

int splhip_params_default(splhip_params *p) {
  if (!p) return SPLHIP_E_ARG;
  std::memset(p, 0, sizeof(*p));
  p->ptypes = 1;
  for (int t = 0; t < kMaxTypes; ++t) {
    p->smooth_fac[t] = 1.f;
    p->brightness[t] = 1.f;
    p->col_vector[t] = 0;
    p->a_eq_e[t] = 1;
  }
  p->sort_type = 1;
  p->gray_absorption = 0.2f;
  p->tile_size = 50;
  p->hip_region = 0;
  p->hip_split = 0;
  p->hip_variant = 0;
  p->hip_heavy = 0;
  p->hip_heavy_max = 0;
  p->hip_chunk_entries = 0;
  p->hip_autotune = 0;
  p->hip_bin_group = 0;
  p->hip_ppw = 0;
  p->hip_stage = 0;
  p->hip_stage_ordered = 0;
  return 0;
}

int splhip_params_from_par(const char *par_text, float b_brightness, splhip_params *p,
                           float scalars_out[7]) {
  if (!p || !scalars_out) return SPLHIP_E_ARG;
  try {
    const auto d = parse_par(par_text);
    splhip_params_default(p);
    const int xres = par_num<int>(d, "xres", 800);
    const int yres = par_num<int>(d, "yres", xres);
    scalars_out[0] = static_cast<float>(xres);
    scalars_out[1] = static_cast<float>(yres);
    scalars_out[2] = par_num<float>(d, "fov", 45.f);
    scalars_out[3] = par_num<float>(d, "zmin", 0.f);
    scalars_out[4] = par_num<float>(d, "zmax", 1.e23f);
    scalars_out[5] = par_bool(d, "projection", true) ? 1.f : 0.f;
    scalars_out[6] = par_num<float>(d, "minrad_pix", 1.f);
    const int nt = par_num<int>(d, "ptypes", 1);
    if (nt < 1 || nt > kMaxTypes) return SPLHIP_E_ARG;
    p->ptypes = nt;
    const bool a_eq_e = par_bool(d, "a_eq_e", true);
    for (int t = 0; t < nt; ++t) {
      const std::string ts = std::to_string(t);
      p->smooth_fac[t] = par_num<float>(d, "smooth_factor" + ts, 1.f);
      // brightness[t] = brightness{t} * b_brightness (splotch_host.cc:262-263)
      p->brightness[t] = par_num<float>(d, "brightness" + ts, 1.f) * b_brightness;
      p->col_vector[t] = par_bool(d, "color_is_vector" + ts, false) ? 1 : 0;
      p->a_eq_e[t] = par_bool(d, "a_eq_e" + ts, a_eq_e) ? 1 : 0;
    }
    p->sort_type = par_num<int>(d, "sort_type", 1);
    p->gray_absorption = par_num<float>(d, "gray_absorption", 0.2f);
    p->tile_size = static_cast<int>(par_num<float>(d, "tile_size_cpu", 50.f));
    p->hip_region = par_num<int>(d, "hip_region", 0);
    p->hip_split = par_num<int>(d, "hip_split", 0);
    p->hip_variant = par_num<int>(d, "hip_variant", 0);
    p->hip_heavy = par_num<int>(d, "hip_heavy", 0);
    p->hip_heavy_max = par_num<int>(d, "hip_heavy_max", 0);
    p->hip_chunk_entries = par_num<int>(d, "hip_chunk_entries", 0);
    p->hip_autotune = par_num<int>(d, "hip_autotune", 0);
    p->hip_bin_group = par_num<int>(d, "hip_bin_group", 0);
    p->hip_ppw = par_num<int>(d, "hip_ppw", 0);
    p->hip_stage = par_num<int>(d, "hip_stage", 0);
    p->hip_stage_ordered = par_num<int>(d, "hip_stage_ordered", 0);
  } catch (const ParError &) {
    return SPLHIP_E_ARG;
  }
  return 0;
}

int splhip_post_from_par(const char *par_text, splhip_post *out) {
  if (!out) return SPLHIP_E_ARG;
  try {
    const auto d = parse_par(par_text);
    std::memset(out, 0, sizeof(*out));
    // defaults: splotch.cc:186 (a_eq_e), :307 (colorbar), splotchutils.cc:232-234
    out->a_eq_e = par_bool(d, "a_eq_e", true) ? 1 : 0;
    out->gamma = par_num<double>(d, "pic_gamma", 1.0);
    out->brightness = par_num<double>(d, "pic_brightness", 0.0);
    out->contrast = par_num<double>(d, "pic_contrast", 1.0);
    out->colorbar = par_bool(d, "colorbar", false) ? 1 : 0;
    const int nt = par_num<int>(d, "ptypes", 1);
    if (nt < 1 || nt > kMaxTypes) return SPLHIP_E_ARG;
    out->ptypes = nt;
    for (int t = 0; t < nt; ++t)
      out->col_vector[t] = par_bool(d, "color_is_vector" + std::to_string(t), false) ? 1 : 0;
  } catch (const ParError &) {
    return SPLHIP_E_ARG;
  }
  return 0;
}

int splhip_postprocess(splhip_ctx *c, const float *pic, int32_t xres, int32_t yres,
                       const splhip_post *post, float *pic_out, uint8_t *rgb8, int32_t layout,
                       void *stream) {
  if (!c || !pic || !post) return fail(c, SPLHIP_E_ARG, "null argument");
  if (xres < 1 || yres < 1) return fail(c, SPLHIP_E_ARG, "bad image size");
  if (static_cast<int64_t>(xres) * yres * 3 >= (int64_t(1) << 31))
    return fail(c, SPLHIP_E_CAPACITY, "image too large for one post-processing pass");
  if (layout != 0 && layout != 1) return fail(c, SPLHIP_E_ARG, "layout must be 0 or 1");
  if (!pic_out && !rgb8) return fail(c, SPLHIP_E_ARG, "no output");
  if (post->ptypes < 1 || post->ptypes > kMaxTypes) return fail(c, SPLHIP_E_ARG, "bad ptypes");
  try {
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    hipStream_t s = static_cast<hipStream_t>(stream);
    PostArgs a{};
    a.pic = pic;
    a.out = pic_out;
    a.rgb8 = rgb8;
    a.xres = xres;
    a.yres = yres;
    a.a_eq_e = post->a_eq_e ? 1 : 0;
    a.adjust = (post->gamma != 1.0 || post->brightness != 0.0 || post->contrast != 1.0) ? 1 : 0;
    a.gamma = post->gamma;
    a.brightness = post->brightness;
    a.contrast = post->contrast;
    if (post->colorbar)
      for (int t = 0; t < post->ptypes; ++t) {
        if (post->col_vector[t]) continue;
        if (c->pal_x[t].empty()) return fail(c, SPLHIP_E_ARG, "colorbar of a type without palette");
        a.bar_type[a.nbars++] = t;
      }
    if (a.nbars) upload_palettes(c);
    for (int t = 0; t < kMaxTypes; ++t) a.pal_n[t] = static_cast<int32_t>(c->pal_x[t].size());
    a.pal_stride = c->pal_stride;
    a.pal_x = c->d_palx;
    a.pal_rgb = c->d_palrgb;
    a.exptab = c->d_exptab;
    a.expfac = c->consts.expfac;
    a.taylorlimit = c->consts.taylorlimit;
    launch_post(a, layout, s);
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return 0;
}

int splhip_norm_from_par(const char *par_text, splhip_norm *out) {
  if (!out) return SPLHIP_E_ARG;
  try {
    const auto d = parse_par(par_text);
    std::memset(out, 0, sizeof(*out));
    const int nt = par_num<int>(d, "ptypes", 1);
    if (nt < 1 || nt > kMaxTypes) return SPLHIP_E_ARG;
    out->ptypes = nt;
    out->norm_intensity = 1;
    for (int t = 0; t < nt; ++t) {  // scenemaker.cc:51-55, 223-233
      const std::string ts = std::to_string(t);
      out->intensity_log[t] = par_bool(d, "intensity_log" + ts, false) ? 1 : 0;
      out->color_log[t] = par_bool(d, "color_log" + ts, false) ? 1 : 0;
      out->color_asinh[t] = par_bool(d, "color_asinh" + ts, false) ? 1 : 0;
      out->col_vector[t] = par_bool(d, "color_is_vector" + ts, false) ? 1 : 0;
      struct { const char *key; int32_t *has; float *val; } ov[4] = {
          {"intensity_min", out->has_intensity_min, out->intensity_min},
          {"intensity_max", out->has_intensity_max, out->intensity_max},
          {"color_min", out->has_color_min, out->color_min},
          {"color_max", out->has_color_max, out->color_max}};
      for (auto &o : ov)
        if (d.count(o.key + ts)) {
          o.has[t] = 1;
          o.val[t] = par_num<float>(d, o.key + ts, 0.f);
        }
    }
  } catch (const ParError &) {
    return SPLHIP_E_ARG;
  }
  return 0;
}

namespace {

// particle_normalize with the per-type extremes in `keys` (device, 6 ptypes
// u32): the shared body of splhip_normalize and splhip_normalize_noalloc.
int normalize_impl(splhip_ctx *c, int64_t n, float *inten, float *col, float *c1, float *c2,
                   const float *hsml, const uint8_t *type, const splhip_norm *norm,
                   splhip_norm_ranges *ranges, void *stream, bool noalloc) {
  if (!c || !norm) return fail(c, SPLHIP_E_ARG, "null argument");
  if (n < 0) return fail(c, SPLHIP_E_ARG, "negative particle count");
  const int nt = norm->ptypes;
  if (nt < 1 || nt > kMaxTypes) return fail(c, SPLHIP_E_ARG, "bad ptypes");
  if (n > 0 && (!inten || !col || !hsml)) return fail(c, SPLHIP_E_ARG, "null field");
  bool vec = false;
  for (int t = 0; t < nt; ++t) vec = vec || norm->col_vector[t];
  if (n > 0 && vec && (!c1 || !c2))
    return fail(c, SPLHIP_E_ARG, "vector colour without c1/c2");
  try {
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    hipStream_t s = static_cast<hipStream_t>(stream);
    NormArgs a{};
    a.n = n;
    a.I = inten;
    a.er = col;
    a.eg = c1;
    a.eb = c2;
    a.r = hsml;
    a.type = type;
    a.ntypes = nt;
    a.norm_int = norm->norm_intensity ? 1 : 0;
    for (int t = 0; t < nt; ++t) {
      a.flags[t][0] = norm->intensity_log[t] ? 1 : 0;
      a.flags[t][1] = norm->color_log[t] ? 1 : 0;
      a.flags[t][2] = norm->color_asinh[t] ? 1 : 0;
      a.flags[t][3] = norm->col_vector[t] ? 1 : 0;
      a.ovr_set[t][0] = norm->has_intensity_min[t] ? 1 : 0;
      a.ovr_set[t][1] = norm->has_intensity_max[t] ? 1 : 0;
      a.ovr_set[t][2] = norm->has_color_min[t] ? 1 : 0;
      a.ovr_set[t][3] = norm->has_color_max[t] ? 1 : 0;
      a.ovr[t][0] = norm->intensity_min[t];
      a.ovr[t][1] = norm->intensity_max[t];
      a.ovr[t][2] = norm->color_min[t];
      a.ovr[t][3] = norm->color_max[t];
    }
    uint32_t *keys = noalloc ? c->scratch_ptr : c->ws.get_as<uint32_t>(S_NORM, 6 * kMaxTypes);
    launch_normalize(a, keys, s);
    if (ranges) {
      uint32_t h[6 * kMaxTypes];
      hip_check(hipMemcpyAsync(h, keys, sizeof(uint32_t) * 6 * nt, hipMemcpyDeviceToHost, s),
                "read ranges");
      hip_check(hipStreamSynchronize(s), "sync ranges");
      auto dek = [](uint32_t k) {
        const uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
        float f;
        std::memcpy(&f, &b, 4);
        return f;
      };
      for (int t = 0; t < nt; ++t) {
        splhip_norm_ranges &R = ranges[t];
        R.range_min = dek(h[6 * t + 0]);
        R.range_max = dek(h[6 * t + 1]);
        R.clamp_min = norm->has_color_min[t] ? norm->color_min[t] : R.range_min;
        R.clamp_max = norm->has_color_max[t] ? norm->color_max[t] : R.range_max;
        R.intensity_min = norm->has_intensity_min[t] ? norm->intensity_min[t] : dek(h[6 * t + 2]);
        R.intensity_max = norm->has_intensity_max[t] ? norm->intensity_max[t] : dek(h[6 * t + 3]);
        R.size_min = dek(h[6 * t + 4]);
        R.size_max = dek(h[6 * t + 5]);
      }
    }
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return 0;
}

}  // namespace

int splhip_normalize(splhip_ctx *c, int64_t n, float *inten, float *col, float *c1, float *c2,
                     const float *hsml, const uint8_t *type, const splhip_norm *norm,
                     splhip_norm_ranges *ranges, void *stream) {
  return normalize_impl(c, n, inten, col, c1, c2, hsml, type, norm, ranges, stream, false);
}

int splhip_normalize_noalloc(splhip_ctx *c, int64_t n, float *inten, float *col, float *c1,
                             float *c2, const float *hsml, const uint8_t *type,
                             const splhip_norm *norm, splhip_norm_ranges *ranges, void *stream) {
  if (!c || !norm) return fail(c, SPLHIP_E_ARG, "null argument");
  if (!c->scratch_ptr) return fail(c, SPLHIP_E_STATE, "no scratch address on this context");
  if (norm->ptypes < 1 || norm->ptypes > kMaxTypes || 6 * norm->ptypes > kNoAllocWords)
    return fail(c, SPLHIP_E_CAPACITY, "ptypes past the no-alloc scratch (at most 8 types)");
  return normalize_impl(c, n, inten, col, c1, c2, hsml, type, norm, ranges, stream, true);
}

namespace {

// bin_reader_prep (reader/bin_reader.cc:34-88): keys and column numbers.
struct BinBlock {
  std::string file;
  int nfields = 0;
  int col[8];  // 0-based column of x, y, z, r, I, C1, C2, C3, or -1
  bool swap = true;
  float smooth = 1.f;
  int64_t npart = 0;
};

int bin_block_keys(const char *par_text, BinBlock &b, std::string &err) {
  try {
    const auto d = parse_par(par_text);
    if (!d.count("infile") || !d.count("num_blocks") || !d.count("x") || !d.count("y") ||
        !d.count("z") || !d.count("C1")) {
      err = "block file keys infile, num_blocks, x, y, z and C1 are required";
      return SPLHIP_E_ARG;
    }
    b.file = d.at("infile");
    b.nfields = par_num<int>(d, "num_blocks", 0);
    b.swap = par_bool(d, "swap_endian", true);
    b.smooth = par_num<float>(d, "smooth_factor", 1.f);
    const char *keys[8] = {"x", "y", "z", "r", "I", "C1", "C2", "C3"};
    if (b.nfields < 1) {
      err = "num_blocks must be positive";
      return SPLHIP_E_ARG;
    }
    // a column number <= 0 names no column: the quantity takes its default
    // value, as qty_idx < 0 does (bin_reader.cc:60-67,166-178)
    for (int q = 0; q < 8; ++q) {
      b.col[q] = d.count(keys[q]) ? par_num<int>(d, keys[q], 0) - 1 : -1;
      if (b.col[q] < 0) b.col[q] = -1;
      if (b.col[q] >= b.nfields) {
        err = std::string("column of ") + keys[q] + " past num_blocks";
        return SPLHIP_E_ARG;
      }
    }
  } catch (const ParError &e) {
    err = e.msg;
    return SPLHIP_E_ARG;
  }
  FILE *f = std::fopen(b.file.c_str(), "rb");
  if (!f) {
    err = "could not open input file '" + b.file + "'";
    return SPLHIP_E_ARG;
  }
  std::fseek(f, 0, SEEK_END);
  const int64_t size = static_cast<int64_t>(std::ftell(f));
  std::fclose(f);
  b.npart = size / (4 * static_cast<int64_t>(b.nfields));
  return 0;
}

}  // namespace

int splhip_bin_block_count(const char *par_text, int64_t *npart) {
  if (!npart) return SPLHIP_E_ARG;
  BinBlock b;
  std::string err;
  if (int rc = bin_block_keys(par_text, b, err)) return rc;
  *npart = b.npart;
  return 0;
}

int splhip_read_bin_block(splhip_ctx *c, const char *par_text, int64_t n, float *x, float *y,
                          float *z, float *hsml, float *inten, float *col, float *c1, float *c2,
                          uint8_t *type, void *stream) {
  if (!c) return SPLHIP_E_ARG;
  BinBlock b;
  std::string err;
  if (int rc = bin_block_keys(par_text, b, err)) return fail(c, rc, err.c_str());
  if (n != b.npart) return fail(c, SPLHIP_E_ARG, "n differs from the file's particle count");
  FILE *f = std::fopen(b.file.c_str(), "rb");
  if (!f) return fail(c, SPLHIP_E_ARG, "could not open input file");
  try {
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    hipStream_t s = static_cast<hipStream_t>(stream);
    constexpr int64_t kChunk = int64_t(1) << 22;  // floats per staging buffer (16 MiB)
    for (int k = 0; k < 2; ++k)
      if (!c->h_ing[k]) {
        hip_check(hipHostMalloc(reinterpret_cast<void **>(&c->h_ing[k]), kChunk * 4,
                                hipHostMallocDefault),
                  "hipHostMalloc ingest");
        hip_check(hipEventCreateWithFlags(&c->ev_ing[k], hipEventDisableTiming),
                  "hipEventCreate");
      }
    float *dst[8] = {x, y, z, hsml, inten, col, c1, c2};
    // the reader's values for a missing column (CASEMACRO__ noval)
    const float noval[8] = {0.f, 0.f, 0.f, b.smooth, 0.5f, 1.f, 0.f, 0.f};
    int k = 0;
    for (int q = 0; q < 8; ++q) {
      if (!dst[q] || n == 0) continue;
      if (b.col[q] < 0) {
        launch_fill(dst[q], noval[q], n, s);
        continue;
      }
      const int64_t base = static_cast<int64_t>(b.col[q]) * b.npart;
      for (int64_t i0 = 0; i0 < n; i0 += kChunk) {
        const int64_t cnt = std::min<int64_t>(kChunk, n - i0);
        hip_check(hipEventSynchronize(c->ev_ing[k]), "wait ingest buffer");
        if (std::fseek(f, static_cast<long>(4 * (base + i0)), SEEK_SET) != 0 ||
            std::fread(c->h_ing[k], 4, static_cast<size_t>(cnt), f) != static_cast<size_t>(cnt)) {
          std::fclose(f);
          return fail(c, SPLHIP_E_ARG, "short read of the input file");
        }
        hip_check(hipMemcpyAsync(dst[q] + i0, c->h_ing[k], cnt * 4, hipMemcpyHostToDevice, s),
                  "H2D column");
        hip_check(hipEventRecord(c->ev_ing[k], s), "hipEventRecord");
        k ^= 1;
      }
      launch_ingest_fix(dst[q], n, b.swap ? 1 : 0, q == 3 ? 1 : 0, b.smooth, s);
    }
    if (type && n > 0) launch_fill_u8(type, 0, n, s);
  } catch (const HipError &e) {
    std::fclose(f);
    return hip_fail(c, e);
  }
  std::fclose(f);
  return 0;
}

int splhip_scene_file_parse(const char *par_text, const char *scene_text, splhip_scene *out,
                            int32_t cap, int32_t *n) {
  if (!n || (!out && cap > 0)) return SPLHIP_E_ARG;
  try {
    auto params = parse_par(par_text);
    const int start = par_num<int>(params, "scene_start", 0);
    const int incr = par_num<int>(params, "scene_incr", 1);
    const int last = par_num<int>(params, "last_scene", 1000000);
    if (incr < 1) return SPLHIP_E_ARG;
    std::vector<std::string> lines;
    {
      std::istringstream in(scene_text ? scene_text : "");
      std::string l;
      while (std::getline(in, l)) lines.push_back(l);
    }
    auto split = [](const std::string &l) {
      std::istringstream s(l);
      std::vector<std::string> w;
      std::string x;
      while (s >> x) w.push_back(x);
      return w;
    };
    size_t pos = 0;
    std::vector<std::string> keys;
    if (!lines.empty()) {
      double t[10];
      if (std::sscanf(lines[0].c_str(), "%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", &t[0], &t[1],
                      &t[2], &t[3], &t[4], &t[5], &t[6], &t[7], &t[8], &t[9]) == 10) {
        keys = {"camera_x", "camera_y", "camera_z", "lookat_x", "lookat_y", "lookat_z",
                "sky_x", "sky_y", "sky_z", "fidx"};  // old format: the line is data
      } else {
        keys = split(lines[0]);
        pos = 1;
      }
    }
    pos += static_cast<size_t>(std::max(start, 0));
    int scene = start, count = 0;
    double fidx_prev = 0.;
    const double fidx_par = par_num<double>(params, "fidx", 0.);
    for (; pos < lines.size() && scene < last; pos += static_cast<size_t>(incr), scene += incr) {
      const std::vector<std::string> vals = split(lines[pos]);
      if (vals.size() != keys.size()) return SPLHIP_E_ARG;  // "ERROR in scene file detected"
      // getNextScene applies the scene's keys to the parameters (they stay
      // set for later scenes; every scene line sets the same keys)
      for (size_t i = 0; i < keys.size(); ++i) params[keys[i]] = vals[i];
      auto num = [&](const char *k, const double *deflt) {
        if (!params.count(k)) {
          if (deflt) return *deflt;
          throw ParError{std::string("missing key ") + k};
        }
        return par_num<double>(params, k, 0.);
      };
      const double zero = 0., one = 1.;
      splhip_scene S{};
      S.campos[0] = num("camera_x", nullptr);
      S.campos[1] = num("camera_y", nullptr);
      S.campos[2] = num("camera_z", nullptr);
      S.lookat[0] = num("lookat_x", nullptr);
      S.lookat[1] = num("lookat_y", nullptr);
      S.lookat[2] = num("lookat_z", nullptr);
      S.sky[0] = num("sky_x", &zero);
      S.sky[1] = num("sky_y", &zero);
      S.sky[2] = num("sky_z", &one);
      if (params.count("center_x")) {
        S.center[0] = num("center_x", nullptr);
        S.center[1] = num("center_y", nullptr);
        S.center[2] = num("center_z", nullptr);
      } else {
        std::memcpy(S.center, S.campos, sizeof(S.center));
      }
      S.fidx = params.count("fidx") ? par_num<double>(params, "fidx", 0.) : fidx_par;
      S.line = static_cast<int32_t>(pos);
      // approx(fidx, fidxold): |a - b| <= 1e-5 |b| (math_utils.h:43-47)
      if (count > 0 && std::fabs(S.fidx - fidx_prev) <= 1e-5 * std::fabs(fidx_prev) &&
          count - 1 < cap)
        out[count - 1].keep_particles = 1;
      if (count < cap) out[count] = S;
      fidx_prev = S.fidx;
      ++count;
    }
    *n = count;
  } catch (const ParError &) {
    return SPLHIP_E_ARG;
  }
  return 0;
}

int splhip_set_palette(splhip_ctx *c, int32_t type, int32_t n, const double *x,
                       const float *rgb) {
  if (!c) return SPLHIP_E_ARG;
  if (type < 0 || type >= kMaxTypes || n < 1 || n > 65536 || !x || !rgb)
    return fail(c, SPLHIP_E_ARG, "bad palette");
  for (int i = 1; i < n; ++i)
    if (!(x[i] > x[i - 1])) return fail(c, SPLHIP_E_ARG, "palette nodes must ascend");
  c->pal_x[type].assign(x, x + n);
  c->pal_rgb[type].assign(rgb, rgb + 3 * n);
  c->pal_dirty = true;
  return 0;
}

int splhip_render(splhip_ctx *c, const splhip_particles *P, const splhip_camera *cam,
                  const splhip_params *prm, float *pic, void *stream) {
  if (int rc = check_inputs(c, P, cam, prm)) return rc;
  if (!pic) return fail(c, SPLHIP_E_ARG, "null image");
  try {
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    const JobIn job{P, cam, prm, pic};
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (stream_capturing(s)) {
      if (!c->rv) return fail(c, SPLHIP_E_STATE, "capture needs splhip_capture_reserve first");
      return render_jobs_dc(c, 1, &job, s);
    }
    return render_tuned(c, 1, &job, s, true);
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
}

namespace {

constexpr int64_t kRecChunk = int64_t(1) << 17;  // records per pinned buffer
constexpr int kRecWorkers = 8;

// Host particle_sim records into the device SoA (soa: 8 columns of nn floats;
// typ: u8): worker w gathers the fields of its chunks into its two pinned
// buffers in turn (each reused once its copy is done), a chunk's columns
// packed back to back, and queues one copy per chunk into the device staging
// area (chunk-major); one kernel then unpacks the columns.  The colour's
// second and third components travel only when a type has a vector colour.
// Types above 255 become 255 (no rendered type).
void upload_records(splhip_ctx *c, const splhip_particle_sim *hp, int64_t n, float *soa,
                    int64_t nn, uint8_t *typ, bool colour3, hipStream_t s) {
  constexpr size_t kBuf = kRecChunk * (8 * sizeof(float) + 1);
  if (c->h_rec.empty()) {
    c->h_rec.assign(2 * kRecWorkers, nullptr);
    c->ev_rec.assign(2 * kRecWorkers, nullptr);
    for (int k = 0; k < 2 * kRecWorkers; ++k) {
      hip_check(hipHostMalloc(reinterpret_cast<void **>(&c->h_rec[k]), kBuf, hipHostMallocDefault),
                "hipHostMalloc records");
      hip_check(hipEventCreateWithFlags(&c->ev_rec[k], hipEventDisableTiming), "hipEventCreate");
    }
  }
  const int ncol = colour3 ? 8 : 6;
  const int64_t rb = 4 * ncol + 1;  // staged bytes per record
  char *stage = c->ws.get_as<char>(S_RECSTAGE, static_cast<size_t>(n * rb));
  const int64_t nchunk = (n + kRecChunk - 1) / kRecChunk;
  std::vector<std::string> errs(kRecWorkers);
  auto work = [&](int w) {
    try {
      hip_check(hipSetDevice(c->device), "hipSetDevice");
      int k = 0;
      for (int64_t ch = w; ch < nchunk; ch += kRecWorkers) {
        const int64_t i0 = ch * kRecChunk, cnt = std::min(kRecChunk, n - i0);
        const int slot = 2 * w + k;
        hip_check(hipEventSynchronize(c->ev_rec[slot]), "wait record buffer");
        float *col = reinterpret_cast<float *>(c->h_rec[slot]);
        uint8_t *tb = reinterpret_cast<uint8_t *>(col + ncol * cnt);
        const splhip_particle_sim *p = hp + i0;
        for (int64_t i = 0; i < cnt; ++i) {
          col[0 * cnt + i] = p[i].x;
          col[1 * cnt + i] = p[i].y;
          col[2 * cnt + i] = p[i].z;
          col[3 * cnt + i] = p[i].r;
          col[4 * cnt + i] = p[i].I;
          col[5 * cnt + i] = p[i].er;
          if (colour3) {
            col[6 * cnt + i] = p[i].eg;
            col[7 * cnt + i] = p[i].eb;
          }
          tb[i] = static_cast<uint8_t>(std::min<uint32_t>(p[i].type, 255u));
        }
        hip_check(hipMemcpyAsync(stage + i0 * rb, col, static_cast<size_t>(cnt * rb),
                                 hipMemcpyHostToDevice, s),
                  "H2D records");
        hip_check(hipEventRecord(c->ev_rec[slot], s), "hipEventRecord");
        k ^= 1;
      }
    } catch (const HipError &e) {
      errs[w] = e.what;
    }
  };
  std::vector<std::thread> pool;
  for (int w = 1; w < kRecWorkers; ++w) pool.emplace_back(work, w);
  work(0);
  for (auto &t : pool) t.join();
  for (const auto &e : errs)
    if (!e.empty()) throw HipError{hipErrorUnknown, "record upload failed"};
  launch_unpack_records(stage, n, kRecChunk, ncol, soa, nn, typ, s);
}

}  // namespace

int splhip_render_host(splhip_ctx *c, const splhip_particle_sim *hp, int64_t n,
                       const splhip_camera *cam, const splhip_params *prm, float *host_pic,
                       void *stream) {
  if (!c || !cam || !prm || !host_pic || (n > 0 && !hp)) return fail(c, SPLHIP_E_ARG, "null argument");
  // the image and parameter checks before anything is allocated or uploaded
  if (n < 0) return fail(c, SPLHIP_E_ARG, "negative particle count");
  if (n >= (int64_t(1) << 32) - 1) return fail(c, SPLHIP_E_CAPACITY, "too many particles");
  if (cam->xres <= 0 || cam->yres <= 0) return fail(c, SPLHIP_E_ARG, "bad image size");
  if (prm->ptypes < 1 || prm->ptypes > kMaxTypes) return fail(c, SPLHIP_E_ARG, "bad ptypes");
  try {
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t nn = std::max<int64_t>(n, 1);
    float *soa = c->ws.get_as<float>(S_SOA, static_cast<size_t>(nn) * 8);
    uint8_t *typ = c->ws.get_as<uint8_t>(S_AUX0, static_cast<size_t>(nn));
    const size_t npix = static_cast<size_t>(cam->xres) * cam->yres * 3;
    float *pic = c->ws.get_as<float>(S_PIC, npix);
    bool colour3 = false;
    for (int t = 0; t < prm->ptypes && t < kMaxTypes; ++t) colour3 = colour3 || prm->col_vector[t];
    if (n > 0) upload_records(c, hp, n, soa, nn, typ, colour3, s);
    splhip_particles P{};
    P.x = soa; P.y = soa + nn; P.z = soa + 2 * nn; P.hsml = soa + 3 * nn;
    P.inten = soa + 4 * nn; P.col = soa + 5 * nn;
    P.c1 = colour3 ? soa + 6 * nn : nullptr;
    P.c2 = colour3 ? soa + 7 * nn : nullptr;
    P.type = typ;
    P.n = n;
    if (int rc = check_inputs(c, &P, cam, prm)) return rc;
    const JobIn job{&P, cam, prm, pic};
    if (int rc = render_tuned(c, 1, &job, s, false)) return rc;
    hip_check(hipMemcpyAsync(host_pic, pic, npix * sizeof(float), hipMemcpyDeviceToHost, s),
              "D2H image");
    hip_check(hipStreamSynchronize(s), "sync");
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return 0;
}

int splhip_project(splhip_ctx *c, const splhip_particles *P, const splhip_camera *cam,
                   const splhip_params *prm, float *rec_out, uint8_t *active_out, void *stream) {
  if (int rc = check_inputs(c, P, cam, prm)) return rc;
  if (!rec_out || !active_out) return fail(c, SPLHIP_E_ARG, "null output");
  try {
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    hipStream_t s = static_cast<hipStream_t>(stream);
    drop_fwd(c);
    const JobIn job{P, cam, prm, nullptr};
    Plan pl;
    make_plan(c, 1, &job, 0, pl, s);
    launch_project(pl.d_pa, pl.d_blk, 1, pl.nblocks, reinterpret_cast<Rec *>(rec_out), active_out,
                   s);
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return 0;
}

int splhip_bins(splhip_ctx *c, const splhip_particles *P, const splhip_camera *cam,
                const splhip_params *prm, int32_t ordered, int64_t *tile_off, uint32_t *idx,
                int64_t idx_cap, int64_t *n_entries, uint32_t *order_out, int64_t *n_order,
                void *stream) {
  if (int rc = check_inputs(c, P, cam, prm)) return rc;
  try {
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int p = ordered ? 0 : 1;
    drop_fwd(c);
    const JobIn job{P, cam, prm, nullptr};
    Plan pl;
    make_plan(c, 1, &job, 0, pl, s);  // units = reference tiles
    Workspace &ws = c->ws;
    Rec *rec = ws.get_as<Rec>(S_REC, static_cast<size_t>(std::max<int64_t>(pl.ntot, 1)));
    uint8_t *act = ws.get_as<uint8_t>(S_ACT, static_cast<size_t>(std::max<int64_t>(pl.ntot, 1)));
    uint32_t *tot = ws.get_as<uint32_t>(S_TOTALS, kTotalsWords);
    c->nmarks = 0;
    stage_project_compact(c, pl, rec, act, tot, s);
    if (pl.m[p]) {
      stage_order(c, pl, p, rec, act, s);
      stage_count(c, pl, p, rec, tot, s);
      read_totals(c, tot + T_E + p, 1, s);
      pl.e[p] = c->h_tot[0];
      stage_bins(c, pl, p, s);
    }
    const uint32_t m = pl.m[p], e = pl.e[p];
    if (n_entries) *n_entries = e;
    if (n_order) *n_order = m;
    const uint64_t ntiles = pl.nunits;
    std::vector<uint32_t> off(ntiles + 1, 0);
    if (m > 0)
      hip_check(hipMemcpyAsync(off.data(), pl.bin_off[p], (ntiles + 1) * 4, hipMemcpyDeviceToHost, s),
                "D2H offsets");
    if (order_out && m > 0)
      hip_check(hipMemcpyAsync(order_out, pl.order[p], static_cast<size_t>(m) * 4,
                               hipMemcpyDeviceToHost, s),
                "D2H order");
    if (idx && e > 0) {
      if (idx_cap < e) return fail(c, SPLHIP_E_CAPACITY, "idx buffer too small");
      uint32_t *mapped = ws.get_as<uint32_t>(S_MAP, e);
      launch_map(pl.ent[p], pl.order[p], mapped, e, s);
      hip_check(hipMemcpyAsync(idx, mapped, static_cast<size_t>(e) * 4, hipMemcpyDeviceToHost, s),
                "D2H entries");
    }
    hip_check(hipStreamSynchronize(s), "sync");
    if (tile_off)
      for (uint64_t t = 0; t <= ntiles; ++t) tile_off[t] = off[t];
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return 0;
}

int splhip_render_backward(splhip_ctx *c, const float *dL_dpic, const splhip_grads *out,
                           void *stream) {
  if (!c) return SPLHIP_E_ARG;
  if (!c->fwd) return fail(c, SPLHIP_E_STATE, "no single-view render (8 px regions) to differentiate");
  if (!dL_dpic || !out || !out->x || !out->y || !out->z || !out->hsml || !out->inten || !out->col)
    return fail(c, SPLHIP_E_ARG, "null gradient buffer");
  const splhip_params &prm = c->fwd->prm;
  for (int t = 0; t < prm.ptypes; ++t)
    if (prm.col_vector[t] && (!out->c1 || !out->c2))
      return fail(c, SPLHIP_E_ARG, "vector colour type without c1/c2 gradient buffers");
  try {
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    return backward_impl(c, dL_dpic, out, static_cast<hipStream_t>(stream));
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
}

int splhip_render_batched(splhip_ctx *c, int32_t njobs, const splhip_job *jobs, void *stream) {
  if (!c || njobs < 0 || (njobs > 0 && !jobs)) return fail(c, SPLHIP_E_ARG, "null argument");
  if (njobs == 0) return 0;
  if (njobs > (1 << 20)) return fail(c, SPLHIP_E_CAPACITY, "too many jobs");
  std::vector<JobIn> in(njobs);
  for (int j = 0; j < njobs; ++j) {
    if (int rc = check_inputs(c, &jobs[j].parts, &jobs[j].cam, &jobs[j].prm)) return rc;
    if (!jobs[j].pic) return fail(c, SPLHIP_E_ARG, "null image");
    in[j] = JobIn{&jobs[j].parts, &jobs[j].cam, &jobs[j].prm, jobs[j].pic};
  }
  try {
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (stream_capturing(s)) {
      if (!c->rv) return fail(c, SPLHIP_E_STATE, "capture needs splhip_capture_reserve first");
      return render_jobs_dc(c, njobs, in.data(), s);
    }
    return render_groups(c, njobs, in.data(), s);
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
}

int splhip_capture_reserve(splhip_ctx *c, int32_t njobs, const splhip_job *jobs, float headroom,
                           void *stream) {
  if (!c || njobs < 1 || !jobs) return fail(c, SPLHIP_E_ARG, "null argument");
  std::vector<JobIn> in(njobs);
  for (int j = 0; j < njobs; ++j) {
    if (int rc = check_inputs(c, &jobs[j].parts, &jobs[j].cam, &jobs[j].prm)) return rc;
    if (!jobs[j].pic) return fail(c, SPLHIP_E_ARG, "null image");
    in[j] = JobIn{&jobs[j].parts, &jobs[j].cam, &jobs[j].prm, jobs[j].pic};
  }
  try {
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (stream_capturing(s)) return fail(c, SPLHIP_E_STATE, "reserve outside stream capture");
    return capture_reserve(c, njobs, in.data(), headroom, s);
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
}

int splhip_capture_status(splhip_ctx *c, int32_t *overflow) {
  if (!c || !overflow) return fail(c, SPLHIP_E_ARG, "null argument");
  if (!c->rv) return fail(c, SPLHIP_E_STATE, "no reserved batch");
  try {
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    uint32_t *tot = c->rv->ws.get_as<uint32_t>(S_TOTALS, kTotalsWords);
    hip_check(hipDeviceSynchronize(), "sync status");
    hip_check(hipMemcpy(c->h_tot, tot + T_OVF, 4, hipMemcpyDeviceToHost), "read overflow flags");
    *overflow = static_cast<int32_t>(c->h_tot[0]);
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
  return 0;
}

}  // extern "C"
