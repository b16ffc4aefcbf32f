// C-ABI of the gfx950 Splotch renderer: context, host-side camera and .par
// setup, and the frame pipeline (project -> [depth sort] -> bin -> splat).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "scan_sort.hpp"
#include "workspace.hpp"

using namespace splhip;

struct splhip_ctx {
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
  uint32_t *h_tot = nullptr;  // pinned
  // stage timing (splhip_set_profiling): begin/end events per stage
  int prof = 0;
  hipEvent_t ev[2 * SPLHIP_NSTAGES] = {};
  bool used[SPLHIP_NSTAGES] = {};
  float stage_ms[SPLHIP_NSTAGES] = {};
  int64_t last_entries[2] = {};
  int64_t last_active[2] = {};
  // work-item trace of the 8x8 splat (splhip_set_trace), per pass
  int trace = 0;
  std::vector<uint64_t> trace_items[2];
  // auxiliary stream (default priority) for the heavy-block kernel
  hipStream_t aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
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
  hipEvent_t &e = c->ev[2 * stage + (begin ? 0 : 1)];
  if (!e) hip_check(hipEventCreate(&e), "hipEventCreate");
  hip_check(hipEventRecord(e, s), "hipEventRecord");
  if (!begin) c->used[stage] = true;
}

void collect_marks(splhip_ctx *c) {
  if (!c->prof) return;
  for (int k = 0; k < SPLHIP_NSTAGES; ++k) {
    c->stage_ms[k] = 0.f;
    if (!c->used[k]) continue;
    hip_check(hipEventSynchronize(c->ev[2 * k + 1]), "hipEventSynchronize");
    hip_check(hipEventElapsedTime(&c->stage_ms[k], c->ev[2 * k], c->ev[2 * k + 1]),
              "hipEventElapsedTime");
  }
}

void ensure_aux(splhip_ctx *c) {
  if (c->aux) return;
  hip_check(hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming), "hipEventCreate");
  hip_check(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming), "hipEventCreate");
}

uint32_t read_u32(splhip_ctx *c, const uint32_t *d, hipStream_t s) {
  hip_check(hipMemcpyAsync(c->h_tot, d, sizeof(uint32_t), hipMemcpyDeviceToHost, s),
            "read total");
  hip_check(hipStreamSynchronize(s), "sync total");
  return c->h_tot[0];
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
  return g;
}

void run_project(splhip_ctx *c, const splhip_particles *P, const splhip_camera *cam,
                 const splhip_params *prm, Rec *rec, uint8_t *act, hipStream_t s) {
  upload_palettes(c);
  ProjectArgs a{};
  a.x = P->x; a.y = P->y; a.z = P->z; a.hsml = P->hsml; a.inten = P->inten; a.col = P->col;
  a.c1 = P->c1; a.c2 = P->c2; a.type = P->type; a.n = P->n;
  a.cam = *cam;
  for (int t = 0; t < kMaxTypes; ++t) {
    a.smooth_fac[t] = prm->smooth_fac[t];
    a.brightness[t] = prm->brightness[t];
    a.col_vector[t] = prm->col_vector[t];
    a.pass_code[t] = prm->a_eq_e[t] ? 2 : 1;
    a.pal_n[t] = static_cast<int32_t>(c->pal_x[t].size());
  }
  a.pal_stride = c->pal_stride;
  a.ntypes = prm->ptypes;
  a.pal_x = c->d_palx;
  a.pal_rgb = c->d_palrgb;
  a.bfak = c->consts.bfak;
  a.rfac = c->consts.rfac;
  a.rec = rec;
  a.act = act;
  if (P->n > 0) launch_project(a, s);
}

struct PassOut {
  uint32_t m = 0, e = 0, nbins = 0;
  const uint32_t *order = nullptr;  // render order -> original index
  const Rec *recs = nullptr;
  const uint32_t *ent = nullptr;
  const uint32_t *bin_off = nullptr;
};

int ceil_log2(uint64_t v) {
  int b = 0;
  while ((uint64_t(1) << b) < v) ++b;
  return std::max(b, 1);
}

// One render pass (code 1: ordered types, 2: additive types): compaction,
// optional depth sort, gather, binning into CSR.
void run_pass(splhip_ctx *c, int code, int64_t n, const Rec *rec, const uint8_t *act,
              const splhip_params *prm, const BinGeom &g, hipStream_t s, PassOut &out) {
  Workspace &ws = c->ws;
  uint32_t *tot = ws.get_as<uint32_t>(S_TOTALS, 64);
  uint32_t *pos = ws.get_as<uint32_t>(S_POS, static_cast<size_t>(n));
  scan_u8eq(act, static_cast<uint8_t>(code), pos, n, tot, ws, s);
  const uint32_t m = read_u32(c, tot, s);
  out.m = m;
  if (m == 0) return;
  const bool do_sort = (code == 1) && prm->sort_type != 0;
  uint32_t *k0 = do_sort ? ws.get_as<uint32_t>(S_K0, m) : nullptr;
  uint32_t *v0 = ws.get_as<uint32_t>(S_V0, m);
  launch_compact(act, static_cast<uint8_t>(code), pos, n, rec, prm->sort_type, k0, v0, s);
  const uint32_t *order = v0;
  if (do_sort) {
    uint32_t *k1 = ws.get_as<uint32_t>(S_K1, m);
    uint32_t *v1 = ws.get_as<uint32_t>(S_V1, m);
    order = radix_sort_pairs(k0, v0, k1, v1, m, 0, 32, ws, s) ? v1 : v0;
  }
  const int sbase = code == 1 ? 1 : 4;
  mark(c, sbase, false, s);
  mark(c, sbase + 1, true, s);
  Rec *recs = ws.get_as<Rec>(S_RECS, m);
  uint32_t *cnt = ws.get_as<uint32_t>(S_CNT, m);
  uint32_t *off = ws.get_as<uint32_t>(S_OFF, m);
  launch_gather_count(order, m, rec, recs, cnt, g, s);
  scan_u32(cnt, off, m, tot + 1, ws, s);
  const uint32_t e = read_u32(c, tot + 1, s);
  const uint64_t nbins = static_cast<uint64_t>(g.ncx) * g.ncy * g.ngm * g.ngm;
  if (nbins >= (uint64_t(1) << 31)) throw HipError{hipErrorInvalidValue, "too many bins"};
  uint32_t *ek0 = ws.get_as<uint32_t>(S_EK0, e);
  uint32_t *ev0 = ws.get_as<uint32_t>(S_EV0, e);
  uint32_t *ek1 = ws.get_as<uint32_t>(S_EK1, e);
  uint32_t *ev1 = ws.get_as<uint32_t>(S_EV1, e);
  launch_emit(recs, m, off, e, g, ek0, ev0, s);
  const int r = radix_sort_pairs(ek0, ev0, ek1, ev1, e, 0, ceil_log2(nbins), ws, s);
  const uint32_t *keys = r ? ek1 : ek0;
  uint32_t *bin_off = ws.get_as<uint32_t>(S_REGOFF, nbins + 1);
  launch_bin_offsets(keys, e, static_cast<uint32_t>(nbins), bin_off, s);
  mark(c, sbase + 1, false, s);
  out.e = e;
  out.nbins = static_cast<uint32_t>(nbins);
  out.order = order;
  out.recs = recs;
  out.ent = r ? ev1 : ev0;
  out.bin_off = bin_off;
}

bool pass_present(const splhip_params *prm, int code) {
  for (int t = 0; t < prm->ptypes; ++t)
    if ((prm->a_eq_e[t] ? 2 : 1) == code) return true;
  return false;
}

int render_impl(splhip_ctx *c, const splhip_particles *P, const splhip_camera *cam,
                const splhip_params *prm, float *pic, hipStream_t s) {
  const int64_t n = P->n;
  const int G = region_size(prm);
  if (G != 8 && G != 16 && G != 32 && G != 64)
    return fail(c, SPLHIP_E_ARG, "hip_region must be 8, 16, 32 or 64");
  Rec *rec = c->ws.get_as<Rec>(S_REC, static_cast<size_t>(n));
  uint8_t *act = c->ws.get_as<uint8_t>(S_ACT, static_cast<size_t>(n));
  for (int k = 0; k < SPLHIP_NSTAGES; ++k) c->used[k] = false;
  c->last_entries[0] = c->last_entries[1] = c->last_active[0] = c->last_active[1] = 0;
  mark(c, 0, true, s);
  run_project(c, P, cam, prm, rec, act, s);
  mark(c, 0, false, s);
  const BinGeom g = make_geom(cam, prm, G, c->consts);
  bool wrote = false;
  for (int code = 1; code <= 2; ++code) {
    if (n == 0 || !pass_present(prm, code)) continue;
    PassOut po;
    const int sbase = code == 1 ? 1 : 4;
    mark(c, sbase, true, s);
    run_pass(c, code, n, rec, act, prm, g, s, po);
    if (po.m == 0) continue;
    mark(c, sbase + 2, true, s);
    SplatArgs sa{};
    sa.bin_off = po.bin_off;
    sa.ent = po.ent;
    sa.recs = po.recs;
    sa.exptab = c->d_exptab;
    sa.pic = pic;
    sa.T = g.T;
    sa.xres = cam->xres;
    sa.yres = cam->yres;
    sa.ncx = g.ncx;
    sa.ngm = g.ngm;
    sa.accumulate = wrote ? 1 : 0;
    sa.rfac = c->consts.rfac;
    sa.h2sigma = c->consts.h2sigma;
    sa.expfac = c->consts.expfac;
    sa.taylorlimit = c->consts.taylorlimit;
    sa.grayabsorb = prm->gray_absorption;
    sa.prio_len = prm->hip_split > 0 ? static_cast<uint32_t>(prm->hip_split) : 4096u;
    if (G == 8) {
      // work items, heaviest unit first; long ordered lists split by channel
      const uint32_t split = prm->hip_split > 0 ? static_cast<uint32_t>(prm->hip_split) : 4096u;
      const int ordered = code == 1 ? 1 : 0;
      const int need_empty = wrote ? 0 : 1;
      // lists at least this long get a producer/consumer workgroup (0 = auto, < 0 = never)
      sa.heavy_len = prm->hip_heavy < 0 ? 0xffffffffu
                     : prm->hip_heavy > 0 ? static_cast<uint32_t>(prm->hip_heavy) : 32768u;
      uint32_t *tot = c->ws.get_as<uint32_t>(S_TOTALS, 64);
      uint32_t *icnt = c->ws.get_as<uint32_t>(S_AUX1, po.nbins);
      uint32_t *ioff = c->ws.get_as<uint32_t>(S_AUX2, po.nbins);
      launch_items_count(po.bin_off, po.nbins, sa, ordered, split, need_empty, icnt, s);
      scan_u32(icnt, ioff, po.nbins, tot + 2, c->ws, s);
      const uint32_t nitems = read_u32(c, tot + 2, s);
      uint32_t *ik0 = c->ws.get_as<uint32_t>(S_K0, nitems);
      uint32_t *iv0 = c->ws.get_as<uint32_t>(S_V0, nitems);
      uint32_t *ik1 = c->ws.get_as<uint32_t>(S_K1, nitems);
      uint32_t *iv1 = c->ws.get_as<uint32_t>(S_V1, nitems);
      launch_items_emit(po.bin_off, po.nbins, sa, ordered, split, need_empty, ioff, ik0, iv0, s);
      const int rs = radix_sort_pairs(ik0, iv0, ik1, iv1, nitems, 0, 32, c->ws, s);
      const uint32_t *items = rs ? iv1 : iv0;
      launch_count_heavy(rs ? ik1 : ik0, nitems, sa.heavy_len, tot + 3, s);
      const uint32_t nheavy = read_u32(c, tot + 3, s);
      if (c->trace) sa.trace = c->ws.get_as<uint64_t>(S_AUX3, static_cast<size_t>(nitems) * 4);
      mark(c, sbase + 2, true, s);
      SplatArgs sn = sa;  // normal items follow the heavy ones
      if (sn.trace) sn.trace += 4 * static_cast<size_t>(nheavy);
      const int variant = prm->hip_variant > 0 ? prm->hip_variant : 4;
      if (nheavy > 0) {
        // heavy blocks run on a second stream, beside the normal items
        ensure_aux(c);
        hip_check(hipEventRecord(c->ev_fork, s), "record fork");
        hip_check(hipStreamWaitEvent(c->aux, c->ev_fork, 0), "wait fork");
        launch_splat_heavy(code == 2, items, nheavy, sa, c->aux);
        hip_check(hipEventRecord(c->ev_join, c->aux), "record join");
        launch_splat8(code == 2, variant, items + nheavy, nitems - nheavy, sn, s);
        hip_check(hipStreamWaitEvent(s, c->ev_join, 0), "wait join");
      } else {
        launch_splat8(code == 2, variant, items, nitems, sn, s);
      }
      if (c->trace) {
        std::vector<uint64_t> &tv = c->trace_items[code - 1];
        tv.resize(static_cast<size_t>(nitems) * 4);
        hip_check(hipMemcpyAsync(tv.data(), sa.trace, tv.size() * 8, hipMemcpyDeviceToHost, s),
                  "D2H trace");
        hip_check(hipStreamSynchronize(s), "sync trace");
      }
    } else {
      launch_splat(G, code == 2, po.nbins, sa, s);
    }
    mark(c, sbase + 2, false, s);
    c->last_entries[code - 1] = po.e;
    c->last_active[code - 1] = po.m;
    wrote = true;
  }
  if (!wrote)
    launch_fill(pic, 0.f, static_cast<int64_t>(cam->xres) * cam->yres * 3, s);
  collect_marks(c);
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
    hip_check(hipHostMalloc(reinterpret_cast<void **>(&c->h_tot), 64 * sizeof(uint32_t),
                            hipHostMallocDefault),
              "hipHostMalloc");
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
  (void)hipSetDevice(c->device);
  (void)hipDeviceSynchronize();
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->aux) (void)hipStreamDestroy(c->aux);
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
    return render_impl(c, P, cam, prm, pic, static_cast<hipStream_t>(stream));
  } catch (const HipError &e) {
    return hip_fail(c, e);
  }
}

int splhip_render_host(splhip_ctx *c, const splhip_particle_sim *hp, int64_t n,
                       const splhip_camera *cam, const splhip_params *prm, float *host_pic,
                       void *stream) {
  if (!c || !cam || !prm || !host_pic || (n > 0 && !hp)) return fail(c, SPLHIP_E_ARG, "null argument");
  try {
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int64_t nn = std::max<int64_t>(n, 1);
    auto *aos = c->ws.get_as<splhip_particle_sim>(S_AOS, static_cast<size_t>(nn));
    float *soa = c->ws.get_as<float>(S_SOA, static_cast<size_t>(nn) * 8);
    uint8_t *typ = c->ws.get_as<uint8_t>(S_AUX0, static_cast<size_t>(nn));
    const size_t npix = static_cast<size_t>(cam->xres) * cam->yres * 3;
    float *pic = c->ws.get_as<float>(S_PIC, npix);
    if (n > 0) {
      hip_check(hipMemcpyAsync(aos, hp, n * sizeof(splhip_particle_sim), hipMemcpyHostToDevice, s),
                "H2D particles");
      launch_aos_to_soa(aos, n, soa, typ, s);
    }
    splhip_particles P{};
    P.x = soa; P.y = soa + nn; P.z = soa + 2 * nn; P.hsml = soa + 3 * nn;
    P.inten = soa + 4 * nn; P.col = soa + 5 * nn; P.c1 = soa + 6 * nn; P.c2 = soa + 7 * nn;
    P.type = typ;
    P.n = n;
    if (int rc = check_inputs(c, &P, cam, prm)) return rc;
    if (int rc = render_impl(c, &P, cam, prm, pic, s)) return rc;
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
    run_project(c, P, cam, prm, reinterpret_cast<Rec *>(rec_out), active_out, s);
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
    const int64_t n = P->n;
    Rec *rec = c->ws.get_as<Rec>(S_REC, static_cast<size_t>(n));
    uint8_t *act = c->ws.get_as<uint8_t>(S_ACT, static_cast<size_t>(n));
    run_project(c, P, cam, prm, rec, act, s);
    const BinGeom g = make_geom(cam, prm, 0, c->consts);
    const uint64_t ntiles = static_cast<uint64_t>(g.ncx) * g.ncy;
    PassOut po;
    if (n > 0) run_pass(c, ordered ? 1 : 2, n, rec, act, prm, g, s, po);
    if (n_entries) *n_entries = po.e;
    if (n_order) *n_order = po.m;
    std::vector<uint32_t> off(ntiles + 1, 0);
    if (po.m > 0)
      hip_check(hipMemcpyAsync(off.data(), po.bin_off, (ntiles + 1) * 4, hipMemcpyDeviceToHost, s),
                "D2H offsets");
    if (order_out && po.m > 0)
      hip_check(hipMemcpyAsync(order_out, po.order, static_cast<size_t>(po.m) * 4,
                               hipMemcpyDeviceToHost, s),
                "D2H order");
    if (idx && po.e > 0) {
      if (idx_cap < po.e) return fail(c, SPLHIP_E_CAPACITY, "idx buffer too small");
      uint32_t *mapped = c->ws.get_as<uint32_t>(S_MAP, po.e);
      launch_map(po.ent, po.order, mapped, po.e, s);
      hip_check(hipMemcpyAsync(idx, mapped, static_cast<size_t>(po.e) * 4, hipMemcpyDeviceToHost, s),
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

}  // extern "C"
