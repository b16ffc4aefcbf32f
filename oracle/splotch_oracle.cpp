/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the Splotch host renderer (the particle-splat hot path),
 * used as the checker for the HIP library and as the "port" CPU baseline.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library; the product path never does.
 *
 * Every function cites the reference code it restates (paths relative to the
 * Splotch source tree).  Build flags MUST include -ffp-contract=off and must
 * not include -ffast-math: the reference's float arithmetic is reproduced
 * operation by operation (no fused multiply-adds, IEEE division and sqrt).
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   orc_exptable      closed form (tab == expf at nodes, |rel err| <= 1.91e-5,
 *                     op(0)=1, op(x>0)=1, op(x<=-20)=0) + genuine-build fixture
 *   orc_camera        genuine-build fixture (matrix, xfac, shifts) + M*M^-1=I,
 *                     lookat -> screen centre
 *   orc_project       genuine-build fixtures, bit-exact
 *   orc_colorize      genuine-build fixtures, bit-exact; node / midpoint closed forms
 *   orc_sort_perm     genuine-build fixture (std::sort order); stable mode vs
 *                     numpy lexsort
 *   orc_bins          brute force over every (tile, particle) + render parity
 *   orc_render        genuine-build fixtures (raw pic) bit-exact, both modes;
 *                     single-particle and two-particle closed forms
 *   orc_fragments     a count of orc_render's per-pixel updates (measurement
 *                     only, not a parity function): the same footprint loops
 *                     without the arithmetic; checked against a per-pixel
 *                     brute-force count of the same rule on small fields
 *   orc_render64      FP64 twin, checked against orc_render within the exptable
 *                     bound and by additivity
 *   orc_postprocess   genuine-build fixtures (exptable -expm1, colour_adjust,
 *                     add_colorbar, LS_Image quantization and the TGA file, bit
 *                     for bit); closed forms (bar rows = palette nodes, 8-bit
 *                     code of k/256 is k, clamps, x86 conversion of NaN/inf)
 *   orc_normalize     genuine-build fixtures (sceneMaker::particle_normalize,
 *                     bit for bit); closed forms (min -> 0, max -> 1, one value
 *                     -> 1, log10 of powers of ten, -38 for non-positive logs)
 *   orc_adjoint64     FP64 adjoint; with pyoracle.chain64 (the FP64 chain rule
 *                     through pyoracle.project64), checked against FP64 central
 *                     differences (h = 1e-4) of orc_render64 after project64 on
 *                     200 random particles, both modes (test_oracle_adjoint.py)
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <omp.h>

namespace {

// cxxsupport/lsconstants.h:35,51
const double kPi = 3.141592653589793238462643383279502884197;
const double kDegr2rad = kPi / 180.0;

// splotch/splotch_host.cc:44-49 (SPLOTCH_CLASSIC)
const float kH2sigma = 0.5 * std::pow(kPi, -1. / 6.);
const float kBfak = 0.5 * std::pow(kPi, -5. / 6.);
const float kRfac = 0.75;

// splotch/splotchutils.h:174-215, exptable<float32>(-20)
struct ExpTable {
  float expfac, taylorlimit;
  float tab1[1024], tab2[1024];
  explicit ExpTable(float maxexp) {
    expfac = 1048576 / maxexp;
    for (int m = 0; m < 1024; ++m) {
      tab1[m] = std::exp(m * 1024 / expfac);
      tab2[m] = std::exp(m / expfac);
    }
    taylorlimit = std::sqrt(float(2) * std::abs(maxexp) / 1048576);
  }
  float op(float arg) const {
    int iarg = int(arg * expfac);
    if (iarg & ~((1 << 20) - 1)) return (iarg < 0) ? 1.f : 0.f;
    return tab1[iarg >> 10] * tab2[iarg & 1023];
  }
  float expm1(float arg) const {
    if (std::abs(arg) < taylorlimit) return arg;
    return op(arg) - 1.f;
  }
};

const ExpTable &xexp() {
  static ExpTable t(-20.f);
  return t;
}

// linearMap<COLOUR>::getVal_const, cxxsupport/linear_map.h:71-84 with
// interpol_helper, cxxsupport/math_utils.h:186-202: px (double, ascending)
// and py (float RGB) hold np nodes.
void pal_val(const double *px, const float *py, int np, double xv, float out[3]) {
  if (np == 1) {
    out[0] = py[0]; out[1] = py[1]; out[2] = py[2];
  } else if (xv >= px[np - 1]) {
    out[0] = py[3 * (np - 1)]; out[1] = py[3 * (np - 1) + 1]; out[2] = py[3 * (np - 1) + 2];
  } else if (xv <= px[0]) {
    out[0] = py[0]; out[1] = py[1]; out[2] = py[2];
  } else {
    int64_t idx = std::lower_bound(px, px + np, xv) - px;
    if (idx > 0) --idx;
    idx = std::min<int64_t>(np - 2, idx);
    const double frac = (xv - px[idx]) / (px[idx + 1] - px[idx]);
    const float *y0 = py + 3 * idx, *y1 = py + 3 * (idx + 1);
    // (1.-frac)*y[index] + frac*y[index+1], each product rounded to float
    for (int c = 0; c < 3; ++c) out[c] = float((1. - frac) * y0[c]) + float(frac * y1[c]);
  }
}

struct V3 { double x, y, z; };
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
// cxxsupport/vec3.h:72-77 (Normalize multiplies by 1/sqrt)
V3 norm(V3 a) {
  double l = 1.0 / std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
  return {a.x * l, a.y * l, a.z * l};
}
double length(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
// cxxsupport/vec3.h:129-131
V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// kernel/transform.h:33-40: float entry[3][4], aliased as p[12]
struct Mat { float e[3][4]; };

Mat mat_identity() {
  Mat m;
  for (int j = 0; j < 3; ++j) for (int i = 0; i < 4; ++i) m.e[j][i] = 0.f;
  m.e[0][0] = m.e[1][1] = m.e[2][2] = 1.f;
  return m;
}

// kernel/transform.cxx:28-48 (ctor stores the arguments transposed)
Mat mat_from_axes(float a00, float a01, float a02, float a10, float a11, float a12,
                  float a20, float a21, float a22) {
  Mat m;
  m.e[0][0] = a00; m.e[1][0] = a01; m.e[2][0] = a02;
  m.e[0][1] = a10; m.e[1][1] = a11; m.e[2][1] = a12;
  m.e[0][2] = a20; m.e[1][2] = a21; m.e[2][2] = a22;
  m.e[0][3] = 0.f; m.e[1][3] = 0.f; m.e[2][3] = 0.f;
  return m;
}

// kernel/transform.cxx:64-112
bool mat_inverse(const Mat &a, Mat &t) {
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
  t.e[0][0] *= d; t.e[2][0] *= d; t.e[1][1] *= d; t.e[0][2] *= d; t.e[2][2] *= d;
  d = -d;
  t.e[1][0] *= d; t.e[0][1] *= d; t.e[2][1] *= d; t.e[1][2] *= d;
  t.e[0][3] = -(t.e[0][0] * a.e[0][3] + t.e[0][1] * a.e[1][3] + t.e[0][2] * a.e[2][3]);
  t.e[1][3] = -(t.e[1][0] * a.e[0][3] + t.e[1][1] * a.e[1][3] + t.e[1][2] * a.e[2][3]);
  t.e[2][3] = -(t.e[2][0] * a.e[0][3] + t.e[2][1] * a.e[1][3] + t.e[2][2] * a.e[2][3]);
  return true;
}

// kernel/transform.cxx:50-62
void mat_mul_assign(Mat &self, const Mat &b) {
  Mat a = self;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j)
      self.e[j][i] = a.e[0][i] * b.e[j][0] + a.e[1][i] * b.e[j][1] + a.e[2][i] * b.e[j][2];
  self.e[0][3] += b.e[0][3];
  self.e[1][3] += b.e[1][3];
  self.e[2][3] += b.e[2][3];
}

// kernel/transform.cxx:212-219 (double accumulation with float matrix)
V3 trans_point(const float *p, V3 v) {
  return {v.x * p[0] + v.y * p[1] + v.z * p[2] + p[3],
          v.x * p[4] + v.y * p[5] + v.z * p[6] + p[7],
          v.x * p[8] + v.y * p[9] + v.z * p[10] + p[11]};
}

}  // namespace

extern "C" {

/* Camera parameters: the per-frame output of splotch_host.cc:62-109. */
struct OrcCamera {
  float mat[12];
  double xshift, yshift;
  float res2, ycorr, xfac, fovfct, dist;
  float zminval, zmaxval, minrad_pix;
  int xres, yres, projection;
};

/* splotch/splotchutils.h:174-215 — expose the tables and limits. */
void orc_exptable(float *tab1, float *tab2, float *expfac, float *taylorlimit) {
  const ExpTable &t = xexp();
  std::memcpy(tab1, t.tab1, sizeof(t.tab1));
  std::memcpy(tab2, t.tab2, sizeof(t.tab2));
  *expfac = t.expfac;
  *taylorlimit = t.taylorlimit;
}

void orc_exp_op(int64_t n, const float *arg, float *out, int expm1) {
  const ExpTable &t = xexp();
  for (int64_t i = 0; i < n; ++i) out[i] = expm1 ? t.expm1(arg[i]) : t.op(arg[i]);
}

/* splotch/splotch_host.cc:62-109 + kernel/transform.cxx:28-219.
   Returns 0, or -1 for the degenerate matrix (transform.cxx:84). */
int orc_camera(const double *campos, const double *lookat, const double *sky_in,
               const double *centerpos, int xres, int yres, float fov, float zmax,
               float zmin, int projection, float minrad_pix, OrcCamera *out) {
  V3 cam{campos[0], campos[1], campos[2]};
  V3 look{lookat[0], lookat[1], lookat[2]};
  V3 sky{sky_in[0], sky_in[1], sky_in[2]};
  V3 cen{centerpos[0], centerpos[1], centerpos[2]};
  float zmaxval = zmax, zminval = zmin;
  zminval = std::max(0.f, zminval);
  float ycorr = .5f * (yres - xres);
  float res2 = 0.5f * xres;
  float fovfct = std::tan(fov * 0.5f * kDegr2rad);
  sky = norm(sky);
  V3 zaxis = norm(sub(cen, look));
  V3 xaxis = norm(cross(sky, zaxis));
  V3 yaxis = cross(zaxis, xaxis);
  // Make_General_Transform then Invert(): matrix = Inverse(axes), inverse = axes
  Mat axes = mat_from_axes(xaxis.x, xaxis.y, xaxis.z, yaxis.x, yaxis.y, yaxis.z,
                           zaxis.x, zaxis.y, zaxis.z);
  Mat axes_inv;
  if (!mat_inverse(axes, axes_inv)) return -1;
  Mat rot = axes_inv;  // trans.matrix after Invert()
  // trans2: translation by -campos; trans2.Add_Transform(trans)
  Mat tr = mat_identity();
  tr.e[0][3] = -cam.x; tr.e[1][3] = -cam.y; tr.e[2][3] = -cam.z;
  mat_mul_assign(tr, rot);
  std::memcpy(out->mat, &tr.e[0][0], sizeof(out->mat));
  V3 tcpos = trans_point(out->mat, cen);
  float dist = length(sub(cen, look));
  float xfac = res2 / (fovfct * dist);
  out->xshift = -xfac * tcpos.x;
  out->yshift = -xfac * tcpos.y;
  out->res2 = res2; out->ycorr = ycorr; out->xfac = xfac; out->fovfct = fovfct;
  out->dist = dist;
  out->zminval = zminval; out->zmaxval = zmaxval; out->minrad_pix = minrad_pix;
  out->xres = xres; out->yres = yres; out->projection = projection;
  return 0;
}

/* splotch/splotch_host.cc:111-178.  In/out SoA; x,y,z,r,I are updated in place
   (the reference mutates its particle vector). active is 0/1. */
void orc_project(const OrcCamera *c, int64_t n, float *x, float *y, float *z, float *r,
                 float *I, const uint8_t *type, const float *smooth_fac, uint8_t *active) {
  const float res2 = c->res2, ycorr = c->ycorr, fovfct = c->fovfct;
  const double xshift = c->xshift, yshift = c->yshift;
  const int xres = c->xres, yres = c->yres;
  const float minrad_pix = c->minrad_pix;
#pragma omp parallel for schedule(guided, 1000)
  for (int64_t m = 0; m < n; ++m) {
    float xfac2 = c->xfac;
    V3 v{x[m], y[m], z[m]};
    v = trans_point(c->mat, v);
    x[m] = v.x; y[m] = v.y; z[m] = v.z;
    active[m] = 0;
    if (-z[m] <= c->zminval) continue;
    if (-z[m] >= c->zmaxval) continue;
    if (!c->projection) {
      x[m] = res2 + x[m] * xfac2;
      y[m] = res2 + ycorr + y[m] * xfac2;
    } else {
      xfac2 = -res2 / (fovfct * z[m]);
      x[m] = res2 + xshift + x[m] * xfac2;
      y[m] = res2 + yshift + ycorr + y[m] * xfac2;
    }
    I[m] *= 0.5f * kBfak / r[m];
    r[m] *= 2;
    r[m] *= xfac2 * smooth_fac[type ? type[m] : 0];
    float rcorr = std::sqrt(r[m] * r[m] + minrad_pix * minrad_pix) / r[m];
    r[m] *= rcorr;
    I[m] /= rcorr;
    float posx = x[m], posy = y[m];
    float rfacr = kRfac * r[m];
    int minx = int(posx - rfacr + 1);
    if (minx >= xres) continue;
    minx = std::max(minx, 0);
    int maxx = int(posx + rfacr + 1);
    if (maxx <= 0) continue;
    maxx = std::min(maxx, xres);
    if (minx >= maxx) continue;
    int miny = int(posy - rfacr + 1);
    if (miny >= yres) continue;
    miny = std::max(miny, 0);
    int maxy = int(posy + rfacr + 1);
    if (maxy <= 0) continue;
    maxy = std::min(maxy, yres);
    if (miny >= maxy) continue;
    active[m] = 1;
  }
}

/* splotch/splotch_host.cc:252-287 with cxxsupport/linear_map.h:71-84 and
   cxxsupport/math_utils.h:186-202.  pal_x[t*pal_stride + i] (double) and
   pal_rgb[(t*pal_stride + i)*3 + c] (float) hold palette t with pal_n[t] entries.
   e0,e1,e2 in: colour fields (e0 = scalar colour when !col_vector); out: e. */
void orc_colorize(int64_t n, float *e0, float *e1, float *e2, const float *I,
                  const uint8_t *type, const uint8_t *active, const float *brightness,
                  const int *col_vector, const int *pal_n, int pal_stride,
                  const double *pal_x, const float *pal_rgb) {
#pragma omp parallel for schedule(guided, 1000)
  for (int64_t m = 0; m < n; ++m) {
    if (!active[m]) continue;
    int t = type ? type[m] : 0;
    if (!col_vector[t]) {
      float e[3];
      pal_val(pal_x + (int64_t)t * pal_stride, pal_rgb + (int64_t)t * pal_stride * 3, pal_n[t],
              double(e0[m]), e);
      e0[m] = e[0]; e1[m] = e[1]; e2[m] = e[2];
    }
    float fac = I[m] * brightness[t];
    e0[m] *= fac; e1[m] *= fac; e2[m] *= fac;
  }
}

/* splotch/splotch_host.cc:289-321 + comparators splotchutils.h:74-96.
   Writes the permutation: out slot s holds the original index of the record
   that ends up at position s.  stable=0 uses std::sort exactly like the
   reference (libstdc++ introsort; its result depends only on comparison
   outcomes, so sorting (key,index) records with a key-only comparator yields
   the same permutation as sorting particle_sim records); stable=1 uses
   std::stable_sort, the canonical (key, index) order.  Returns -1 for an
   unknown sort_type (planck_fail in the reference). */
int orc_sort_perm(int64_t n, int sort_type, const float *z, const float *er, const float *r,
                  int stable, uint32_t *perm) {
  struct Rec { float k; uint32_t i; };
  std::vector<Rec> v(n);
  const float *key = nullptr;
  switch (sort_type) {
    case 0: for (int64_t i = 0; i < n; ++i) perm[i] = i; return 0;
    case 1: key = z; break;
    case 2: case 3: key = er; break;
    case 4: key = r; break;
    default: return -1;
  }
  for (int64_t i = 0; i < n; ++i) v[i] = {key[i], (uint32_t)i};
  auto asc = [](const Rec &a, const Rec &b) { return a.k < b.k; };
  auto desc = [](const Rec &a, const Rec &b) { return a.k > b.k; };
  bool descending = (sort_type == 2 || sort_type == 4);
  if (stable) {
    if (descending) std::stable_sort(v.begin(), v.end(), desc);
    else std::stable_sort(v.begin(), v.end(), asc);
  } else {
    if (descending) std::sort(v.begin(), v.end(), desc);
    else std::sort(v.begin(), v.end(), asc);
  }
  for (int64_t i = 0; i < n; ++i) perm[i] = v[i].i;
  return 0;
}

/* splotch/splotch_host.cc:337-406: per-tile index lists in ascending particle
   index.  Tile id = iy*ncx + ix (the chunk number of work_distributor,
   splotchutils.h:217-241).  If tile_off is null only the total count is
   returned; otherwise tile_off[ntiles+1] and idx[] are filled. */
int64_t orc_bins(int64_t n, const float *x, const float *y, const float *r,
                 const uint8_t *active, int xres, int yres, int tile, int32_t *tile_cnt,
                 int64_t *tile_off, uint32_t *idx) {
  const int chunkdim = tile;
  const int ncx = (xres + chunkdim - 1) / chunkdim, ncy = (yres + chunkdim - 1) / chunkdim;
  const float rcell = std::sqrt(2.f) * (chunkdim * 0.5f - 0.5f);
  const float cmid0 = 0.5f * (chunkdim - 1);
  std::vector<std::vector<uint32_t>> lists(size_t(ncx) * ncy);
  for (int64_t i = 0; i < n; ++i) {
    float rfacr = kRfac * r[i];
    if (!active[i]) continue;
    int minx = std::max(0, int(x[i] - rfacr + 1) / chunkdim);
    int maxx = std::min(ncx - 1, int(x[i] + rfacr) / chunkdim);
    int miny = std::max(0, int(y[i] - rfacr + 1) / chunkdim);
    int maxy = std::min(ncy - 1, int(y[i] + rfacr) / chunkdim);
    float sumsq = (rcell + rfacr) * (rcell + rfacr);
    for (int ix = minx; ix <= maxx; ++ix) {
      float cx = cmid0 + ix * chunkdim;
      for (int iy = miny; iy <= maxy; ++iy) {
        float cy = cmid0 + iy * chunkdim;
        float rtot2 = (x[i] - cx) * (x[i] - cx) + (y[i] - cy) * (y[i] - cy);
        if (rtot2 < sumsq) lists[size_t(iy) * ncx + ix].push_back((uint32_t)i);
      }
    }
  }
  int64_t total = 0;
  for (size_t t = 0; t < lists.size(); ++t) {
    if (tile_cnt) tile_cnt[t] = (int32_t)lists[t].size();
    if (tile_off) tile_off[t] = total;
    if (idx) std::copy(lists[t].begin(), lists[t].end(), idx + total);
    total += lists[t].size();
  }
  if (tile_off) tile_off[lists.size()] = total;
  return total;
}

/* splotch/splotch_host.cc:330-697 (render_new, no MPI_A_NEQ_E, __SSE2__ branch
   structure).  Particles are the projected + coloured SoA; pic is x-major RGB
   (pic[(x*yres+y)*3+c]).  The V4sf lanes of the reference are restated as three
   independent float lanes. */
void orc_render(int64_t n, const float *px, const float *py, const float *pr, const float *e0,
                const float *e1, const float *e2, const uint8_t *active, int xres, int yres,
                int a_eq_e, float grayabsorb, int tile, float *pic) {
  const ExpTable &xe = xexp();
  const float taylorlimit = xe.taylorlimit;
  const int chunkdim = tile;
  const int ncx = (xres + chunkdim - 1) / chunkdim, ncy = (yres + chunkdim - 1) / chunkdim;
  int64_t total = orc_bins(n, px, py, pr, active, xres, yres, tile, nullptr, nullptr, nullptr);
  std::vector<int64_t> off(size_t(ncx) * ncy + 1);
  std::vector<uint32_t> idx(total);
  orc_bins(n, px, py, pr, active, xres, yres, tile, nullptr, off.data(), idx.data());
  const int nchunks = ncx * ncy;
#pragma omp parallel
  {
    std::vector<float> pre1(chunkdim);
    std::vector<float> lpic(size_t(chunkdim) * chunkdim * 3);
#pragma omp for schedule(dynamic, 1)
    for (int chunk = 0; chunk < nchunks; ++chunk) {
      // work_distributor::chunk_info
      int cix = chunk % ncx, ciy = chunk / ncx;
      int x0 = cix * chunkdim, x1 = std::min(x0 + chunkdim, xres);
      int y0 = ciy * chunkdim, y1 = std::min(y0 + chunkdim, yres);
      int x0s = x0, y0s = y0;
      x1 -= x0; x0 = 0; y1 -= y0; y0 = 0;
      std::fill(lpic.begin(), lpic.end(), 0.f);
      auto L = [&](int xx, int yy) { return &lpic[(size_t(xx) * y1 + yy) * 3]; };
      for (int64_t m = off[chunk]; m < off[chunk + 1]; ++m) {
        const uint32_t i = idx[m];
        float rfacr = pr[i] * kRfac;
        float posx = px[i], posy = py[i];
        posx -= x0s; posy -= y0s;
        int minx = int(posx - rfacr + 1);
        minx = std::max(minx, x0);
        int maxx = int(posx + rfacr + 1);
        maxx = std::min(maxx, x1);
        int miny = int(posy - rfacr + 1);
        miny = std::max(miny, y0);
        int maxy = int(posy + rfacr + 1);
        maxy = std::min(maxy, y1);
        float radsq = rfacr * rfacr;
        float sigma = kH2sigma * pr[i];
        float stp = -1.f / (sigma * sigma);
        float a[3] = {-e0[i], -e1[i], -e2[i]};
        for (int yy = miny; yy < maxy; ++yy) pre1[yy] = xe.op(stp * (yy - posy) * (yy - posy));
        if (a_eq_e) {
          for (int xx = minx; xx < maxx; ++xx) {
            float dxsq = (xx - posx) * (xx - posx);
            float dy = std::sqrt(radsq - dxsq);
            int miny2 = std::max(miny, int(posy - dy + 1)), maxy2 = std::min(maxy, int(posy + dy + 1));
            float pre2 = xe.op(stp * dxsq);
            for (int yy = miny2; yy < maxy2; ++yy) {
              float att = pre1[yy] * pre2;
              float *l = L(xx, yy);
              l[0] += a[0] * att; l[1] += a[1] * att; l[2] += a[2] * att;
            }
          }
        } else {
          float q[3] = {e0[i] / (e0[i] + grayabsorb), e1[i] / (e1[i] + grayabsorb),
                        e2[i] / (e2[i] + grayabsorb)};
          float maxa = std::max(std::abs(a[0]), std::max(std::abs(a[1]), std::abs(a[2])));
          for (int xx = minx; xx < maxx; ++xx) {
            float dxsq = (xx - posx) * (xx - posx);
            float dy = std::sqrt(radsq - dxsq);
            int miny2 = std::max(miny, int(posy - dy + 1)), maxy2 = std::min(maxy, int(posy + dy + 1));
            float pre2 = xe.op(stp * dxsq);
            for (int yy = miny2; yy < maxy2; ++yy) {
              float att = pre1[yy] * pre2;
              float *l = L(xx, yy);
              if ((maxa * att) < taylorlimit) {
                for (int c = 0; c < 3; ++c) l[c] += ((l[c] - q[c]) * a[c]) * att;
              } else {
                for (int c = 0; c < 3; ++c) l[c] += xe.expm1(att * a[c]) * (l[c] - q[c]);
              }
            }
          }
        }
      }
      for (int ix = 0; ix < x1; ++ix)
        for (int iy = 0; iy < y1; ++iy) {
          float *dst = &pic[(size_t(ix + x0s) * yres + (iy + y0s)) * 3];
          const float *l = L(ix, iy);
          dst[0] = l[0]; dst[1] = l[1]; dst[2] = l[2];
        }
    }
  }
}

/* The number of per-pixel updates render_new makes (its fragments): the
   footprint loops of splotch_host.cc:461-609 per (tile, particle) of the
   tile lists, counting the rows of each column. */
int64_t orc_fragments(int64_t n, const float *px, const float *py, const float *pr,
                      const uint8_t *active, int xres, int yres, int tile) {
  const int chunkdim = tile;
  const int ncx = (xres + chunkdim - 1) / chunkdim, ncy = (yres + chunkdim - 1) / chunkdim;
  int64_t total = orc_bins(n, px, py, pr, active, xres, yres, tile, nullptr, nullptr, nullptr);
  std::vector<int64_t> off(size_t(ncx) * ncy + 1);
  std::vector<uint32_t> idx(total);
  orc_bins(n, px, py, pr, active, xres, yres, tile, nullptr, off.data(), idx.data());
  const int nchunks = ncx * ncy;
  int64_t frags = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : frags)
  for (int chunk = 0; chunk < nchunks; ++chunk) {
    const int cix = chunk % ncx, ciy = chunk / ncx;
    const int x0s = cix * chunkdim, y0s = ciy * chunkdim;
    const int x1 = std::min(x0s + chunkdim, xres) - x0s, y1 = std::min(y0s + chunkdim, yres) - y0s;
    for (int64_t m = off[chunk]; m < off[chunk + 1]; ++m) {
      const uint32_t i = idx[m];
      const float rfacr = pr[i] * kRfac;
      const float posx = px[i] - x0s, posy = py[i] - y0s;
      const int minx = std::max(int(posx - rfacr + 1), 0), maxx = std::min(int(posx + rfacr + 1), x1);
      const int miny = std::max(int(posy - rfacr + 1), 0), maxy = std::min(int(posy + rfacr + 1), y1);
      const float radsq = rfacr * rfacr;
      for (int xx = minx; xx < maxx; ++xx) {
        const float dxsq = (xx - posx) * (xx - posx);
        const float dy = std::sqrt(radsq - dxsq);
        const int lo = std::max(miny, int(posy - dy + 1)), hi = std::min(maxy, int(posy + dy + 1));
        if (hi > lo) frags += hi - lo;
      }
    }
  }
  return frags;
}

/* Brute force (no binning): every pixel loops over ALL active particles in
   ascending index order, with each particle's footprint evaluated in the frame
   of the pixel's tile exactly as splotch_host.cc:461-622 does, and the
   tile-membership rule of :370-383 applied per (tile, particle).  Used only to
   pin orc_bins' conservativeness on small inputs (additive mode). */
void orc_render_bruteforce(int64_t n, const float *px, const float *py, const float *pr,
                           const float *e0, const float *e1, const float *e2,
                           const uint8_t *active, int xres, int yres, int tile, float *pic) {
  const ExpTable &xe = xexp();
  const int T = tile;
  const int ncx = (xres + T - 1) / T, ncy = (yres + T - 1) / T;
  const float rcell = std::sqrt(2.f) * (T * 0.5f - 0.5f);
  const float cmid0 = 0.5f * (T - 1);
#pragma omp parallel for schedule(dynamic, 1)
  for (int X = 0; X < xres; ++X) {
    for (int Y = 0; Y < yres; ++Y) {
      int cix = X / T, ciy = Y / T;
      int x0s = cix * T, y0s = ciy * T;
      int x1 = std::min(x0s + T, xres) - x0s, y1 = std::min(y0s + T, yres) - y0s;
      int xl = X - x0s, yl = Y - y0s;
      float acc[3] = {0.f, 0.f, 0.f};
      for (int64_t i = 0; i < n; ++i) {
        if (!active[i]) continue;
        float rfacr = kRfac * pr[i];
        int minxt = std::max(0, int(px[i] - rfacr + 1) / T);
        int maxxt = std::min(ncx - 1, int(px[i] + rfacr) / T);
        int minyt = std::max(0, int(py[i] - rfacr + 1) / T);
        int maxyt = std::min(ncy - 1, int(py[i] + rfacr) / T);
        if (cix < minxt || cix > maxxt || ciy < minyt || ciy > maxyt) continue;
        float sumsq = (rcell + rfacr) * (rcell + rfacr);
        float cx = cmid0 + cix * T, cy = cmid0 + ciy * T;
        float rtot2 = (px[i] - cx) * (px[i] - cx) + (py[i] - cy) * (py[i] - cy);
        if (!(rtot2 < sumsq)) continue;
        float posx = px[i] - x0s, posy = py[i] - y0s;
        int minx = std::max(int(posx - rfacr + 1), 0), maxx = std::min(int(posx + rfacr + 1), x1);
        int miny = std::max(int(posy - rfacr + 1), 0), maxy = std::min(int(posy + rfacr + 1), y1);
        if (xl < minx || xl >= maxx) continue;
        float radsq = rfacr * rfacr;
        float sigma = kH2sigma * pr[i];
        float stp = -1.f / (sigma * sigma);
        float dxsq = (xl - posx) * (xl - posx);
        float dy = std::sqrt(radsq - dxsq);
        int miny2 = std::max(miny, int(posy - dy + 1)), maxy2 = std::min(maxy, int(posy + dy + 1));
        if (yl < miny2 || yl >= maxy2) continue;
        float pre1 = xe.op(stp * (yl - posy) * (yl - posy));
        float pre2 = xe.op(stp * dxsq);
        float att = pre1 * pre2;
        acc[0] += -e0[i] * att; acc[1] += -e1[i] * att; acc[2] += -e2[i] * att;
      }
      float *d = &pic[(size_t(X) * yres + Y) * 3];
      d[0] = acc[0]; d[1] = acc[1]; d[2] = acc[2];
    }
  }
}

/* The master's post-processing and image output, splotch.cc:272-310, over a
   raw x-major RGB image pic[(ix*yres + iy)*3 + c] (updated in place):
     a_eq_e:   c = -xexp.expm1(c)                      splotch.cc:275-284
     colour_adjust (only if gamma != 1, brightness != 0 or contrast != 1):
               c = float(contrast * pow(double(c), gamma) + brightness)
                                                       splotchutils.cc:230-250
     colorbar: for each type t < ptypes with col_vector[t] == 0 (bar k = the
               k-th such type), rows iy = yres-1-10k-y, y < 10, take the
               palette colour at ix/double(xres)  splotchutils.cc:35-56
               (rows below 0 are skipped; the reference writes out of bounds)
   then the Colour8 conversion of LS_Image::put_pixel, ls_image.h:65-75:
     byte = max(0, min(255, int(c * 256))) with x86 truncation (NaN and values
     outside the int range give INT_MIN, i.e. 0),
   and the file of LS_Image::write_TGA, ls_image.cc:178-202: an 18-byte header
   then rows iy = 0..yres-1 of pixels ix = 0..xres-1 as B, G, R.
   tga: 18 + 3*xres*yres bytes (may be NULL). */
static int x86_trunc(float v) {
  if (!(v >= -2147483648.f && v < 2147483648.f)) return INT32_MIN;
  return int(v);
}

void orc_postprocess(int xres, int yres, float *pic, int a_eq_e, double gamma, double brightness,
                     double contrast, int colorbar, int ptypes, const int *col_vector,
                     const int *pal_n, int pal_stride, const double *pal_x,
                     const float *pal_rgb, uint8_t *tga) {
  const int64_t npix = int64_t(xres) * yres;
  const ExpTable &X = xexp();
  if (a_eq_e)
    for (int64_t i = 0; i < 3 * npix; ++i) pic[i] = -X.expm1(pic[i]);
  if (gamma != 1.0 || brightness != 0.0 || contrast != 1.0)
    for (int64_t i = 0; i < 3 * npix; ++i)
      pic[i] = float(contrast * std::pow(double(pic[i]), gamma) + brightness);
  if (colorbar) {
    int offset = 0;
    for (int t = 0; t < ptypes; ++t) {
      if (col_vector[t]) continue;
      for (int ix = 0; ix < xres; ++ix) {
        float e[3];
        pal_val(pal_x + int64_t(t) * pal_stride, pal_rgb + int64_t(t) * pal_stride * 3, pal_n[t],
                ix / double(xres), e);
        for (int y = 0; y < 10; ++y) {
          const int iy = yres - offset - 1 - y;
          if (iy < 0) continue;
          float *d = pic + (int64_t(ix) * yres + iy) * 3;
          d[0] = e[0]; d[1] = e[1]; d[2] = e[2];
        }
      }
      offset += 10;
    }
  }
  if (!tga) return;
  const uint8_t hdr[18] = {0, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0, uint8_t(xres % 256),
                           uint8_t(xres / 256), uint8_t(yres % 256), uint8_t(yres / 256), 24, 32};
  std::memcpy(tga, hdr, 18);
  for (int iy = 0; iy < yres; ++iy)
    for (int ix = 0; ix < xres; ++ix) {
      const float *d = pic + (int64_t(ix) * yres + iy) * 3;
      uint8_t *o = tga + 18 + (int64_t(iy) * xres + ix) * 3;
      for (int c = 0; c < 3; ++c)
        o[2 - c] = uint8_t(std::max(0, std::min(255, x86_trunc(d[c] * 256))));
    }
}

/* sceneMaker::particle_normalize, host branch (splotch/scenemaker.cc:40-296),
   over SoA fields updated in place (I, er, eg, eb; r read), with Normalizer
   (splotch/splotchutils.h:138-171) semantics: collect keeps min/max through
   std::min/std::max (NaN never enters), normAndClamp maps
   (max(minv, min(maxv, v)) - minv) / (maxv - minv), or 1 when maxv == minv.
   Per type t: flags[t*4 + 0..3] = intensity_log, color_log, color_asinh,
   color_is_vector; ovr_set[t*4 + k] / ovr[t*4 + k] = intensity_min,
   intensity_max, color_min, color_max given in the parameters.  norm_int: 0
   for a build with NO_I_NORM.  out[t*6 + 0..5] = colour range of the data,
   colour range used, intensity range used (before normAndClamp); size[t*2]
   = the smoothing-length range.  Single-threaded: the reference's OpenMP
   reduction order only decides which of +0 / -0 survives a tie. */
static double my_asinh(double v) { return std::log(v + std::sqrt(1. + v * v)); }

void orc_normalize(int64_t n, float *I, const float *r, float *er, float *eg, float *eb,
                   const uint8_t *type, int ntypes, const int *flags, const int *ovr_set,
                   const float *ovr, int norm_int, float *out, float *size) {
  struct Norm {
    float minv = 1e37f, maxv = -1e37f;
    void collect(float v) { minv = std::min(minv, v); maxv = std::max(maxv, v); }
    void clamp(float &v) const {
      if (maxv == minv) v = 1.f;
      else v = (std::max(minv, std::min(maxv, v)) - minv) / (maxv - minv);
    }
  };
  std::vector<Norm> in(ntypes), cn(ntypes), rn(ntypes);
  for (int64_t m = 0; m < n; ++m) {
    const int t = type ? type[m] : 0;
    const int *f = flags + 4 * t;
    rn[t].collect(r[m]);
    if (f[0]) {
      if (I[m] > 0) {
        I[m] = std::log10(I[m]);
        in[t].collect(I[m]);
      } else {
        I[m] = -38;
      }
    } else {
      in[t].collect(I[m]);
    }
    if (f[1]) {
      if (er[m] > 0) {
        er[m] = std::log10(er[m]);
        cn[t].collect(er[m]);
      } else {
        er[m] = -38;
      }
    } else {
      if (f[2]) er[m] = float(my_asinh(er[m]));
      cn[t].collect(er[m]);
    }
    if (f[3]) {
      if (f[1]) {
        eg[m] = std::log10(eg[m]);
        eb[m] = std::log10(eb[m]);
      }
      if (f[2]) {
        eg[m] = float(my_asinh(eg[m]));
        eb[m] = float(my_asinh(eb[m]));
      }
      cn[t].collect(eg[m]);
      cn[t].collect(eb[m]);
    }
  }
  for (int t = 0; t < ntypes; ++t) {
    out[6 * t + 0] = cn[t].minv;
    out[6 * t + 1] = cn[t].maxv;
    if (ovr_set[4 * t + 0]) in[t].minv = ovr[4 * t + 0];
    if (ovr_set[4 * t + 1]) in[t].maxv = ovr[4 * t + 1];
    if (ovr_set[4 * t + 2]) cn[t].minv = ovr[4 * t + 2];
    if (ovr_set[4 * t + 3]) cn[t].maxv = ovr[4 * t + 3];
    out[6 * t + 2] = cn[t].minv;
    out[6 * t + 3] = cn[t].maxv;
    out[6 * t + 4] = in[t].minv;
    out[6 * t + 5] = in[t].maxv;
    size[2 * t] = rn[t].minv;
    size[2 * t + 1] = rn[t].maxv;
  }
  for (int64_t m = 0; m < n; ++m) {
    const int t = type ? type[m] : 0;
    if (norm_int) in[t].clamp(I[m]);
    cn[t].clamp(er[m]);
    if (flags[4 * t + 3]) {
      cn[t].clamp(eg[m]);
      cn[t].clamp(eb[m]);
    }
  }
}

/* ---------------------------------------------------------------------------
   FP64 twin and FP64 adjoint.

   The float32 renderer's footprint is discrete (integer pixel ranges, tile
   membership).  The FP64 twin keeps that footprint FROZEN — each (pixel,
   particle) pair contributes iff the float32 renderer would visit it — and
   evaluates the contribution with double arithmetic and the true exponential:
     att = exp(stp*(Y-sy)^2) * exp(stp*(X-sx)^2),  stp = -1/(h2sigma*r)^2
     additive:  P += -e*att
     ordered :  P  = q + (P - q)*exp(-e*att),  q = e/(e+g)
   (the limit of splotch_host.cc:596-671 without table quantization).
   The screen-space particle state (sx, sy, r, e) is passed in double.
   Footprint: fp_x0/fp_x1/fp_y0... are not stored; instead the caller passes
   the float32 projected particles and the frozen pairs are recomputed here
   from them with the float rules.
   --------------------------------------------------------------------------- */

}  // extern "C"

/* Enumerate the frozen footprint of particle i inside tile `chunk` with the
   float rules of splotch_host.cc:461-609, calling f(X, Y) per pixel. */
template <class F>
static void footprint_in_tile(float fx, float fy, float fr, int x0s, int y0s, int x1, int y1,
                              F &&f) {
  float rfacr = fr * kRfac;
  float posx = fx - x0s, posy = fy - y0s;
  int minx = std::max(int(posx - rfacr + 1), 0), maxx = std::min(int(posx + rfacr + 1), x1);
  int miny = std::max(int(posy - rfacr + 1), 0), maxy = std::min(int(posy + rfacr + 1), y1);
  float radsq = rfacr * rfacr;
  for (int xx = minx; xx < maxx; ++xx) {
    float dxsq = (xx - posx) * (xx - posx);
    float dy = std::sqrt(radsq - dxsq);
    int miny2 = std::max(miny, int(posy - dy + 1)), maxy2 = std::min(maxy, int(posy + dy + 1));
    for (int yy = miny2; yy < maxy2; ++yy) f(xx + x0s, yy + y0s);
  }
}

extern "C" {

/* FP64 forward on the frozen footprint.  f* = float32 projected particles
   (define the footprint, in render order), d* = double screen-space state
   (sx, sy, r, e0..e2) in the same order.  pic64 x-major RGB double. */
void orc_render64(int64_t n, const float *fx, const float *fy, const float *fr,
                  const uint8_t *active, const double *sx, const double *sy, const double *sr,
                  const double *de0, const double *de1, const double *de2, int xres, int yres,
                  int a_eq_e, double grayabsorb, int tile, double *pic64) {
  const int T = tile;
  const int ncx = (xres + T - 1) / T, ncy = (yres + T - 1) / T;
  int64_t total = orc_bins(n, fx, fy, fr, active, xres, yres, T, nullptr, nullptr, nullptr);
  std::vector<int64_t> off(size_t(ncx) * ncy + 1);
  std::vector<uint32_t> idx(total);
  orc_bins(n, fx, fy, fr, active, xres, yres, T, nullptr, off.data(), idx.data());
  const double h2s = 0.5 * std::pow(kPi, -1. / 6.);
  std::fill(pic64, pic64 + size_t(xres) * yres * 3, 0.0);
#pragma omp parallel for schedule(dynamic, 1)
  for (int chunk = 0; chunk < ncx * ncy; ++chunk) {
    int cix = chunk % ncx, ciy = chunk / ncx;
    int x0s = cix * T, y0s = ciy * T;
    int x1 = std::min(x0s + T, xres) - x0s, y1 = std::min(y0s + T, yres) - y0s;
    for (int64_t m = off[chunk]; m < off[chunk + 1]; ++m) {
      const uint32_t i = idx[m];
      double sigma = h2s * sr[i];
      double stp = -1.0 / (sigma * sigma);
      double e[3] = {de0[i], de1[i], de2[i]};
      footprint_in_tile(fx[i], fy[i], fr[i], x0s, y0s, x1, y1, [&](int X, int Y) {
        double dx = X - sx[i], dy = Y - sy[i];
        double att = std::exp(stp * dy * dy) * std::exp(stp * dx * dx);
        double *P = &pic64[(size_t(X) * yres + Y) * 3];
        for (int c = 0; c < 3; ++c) {
          if (a_eq_e) P[c] += -e[c] * att;
          else {
            double q = e[c] / (e[c] + grayabsorb);
            P[c] = q + (P[c] - q) * std::exp(-e[c] * att);
          }
        }
      });
    }
  }
}

/* FP64 adjoint of orc_render64 with respect to the screen-space state:
   given G = dL/dP (x-major RGB double), returns per particle
   g_sx, g_sy, g_r, g_e0, g_e1, g_e2 (accumulated, in render order).
   Ordered mode: transmittance is recomputed per pixel — first a forward sweep
   for the total optical depth tau_c = sum_k e_k,c att_k, then a second forward
   sweep that carries P_{k-1} and the prefix depth, from which the suffix
   transmittance S_k = exp(-(tau - prefix_k)) follows without division. */
void orc_adjoint64(int64_t n, const float *fx, const float *fy, const float *fr,
                   const uint8_t *active, const double *sx, const double *sy, const double *sr,
                   const double *de0, const double *de1, const double *de2, int xres, int yres,
                   int a_eq_e, double grayabsorb, int tile, const double *G, double *g_sx,
                   double *g_sy, double *g_r, double *g_e0, double *g_e1, double *g_e2) {
  const int T = tile;
  const int ncx = (xres + T - 1) / T, ncy = (yres + T - 1) / T;
  int64_t total = orc_bins(n, fx, fy, fr, active, xres, yres, T, nullptr, nullptr, nullptr);
  std::vector<int64_t> off(size_t(ncx) * ncy + 1);
  std::vector<uint32_t> idx(total);
  orc_bins(n, fx, fy, fr, active, xres, yres, T, nullptr, off.data(), idx.data());
  const double h2s = 0.5 * std::pow(kPi, -1. / 6.);
  for (int64_t i = 0; i < n; ++i) g_sx[i] = g_sy[i] = g_r[i] = g_e0[i] = g_e1[i] = g_e2[i] = 0;
  // per-entry partials, reduced per particle afterwards in a fixed order
  std::vector<double> part(size_t(total) * 6, 0.0);
#pragma omp parallel for schedule(dynamic, 1)
  for (int chunk = 0; chunk < ncx * ncy; ++chunk) {
    int cix = chunk % ncx, ciy = chunk / ncx;
    int x0s = cix * T, y0s = ciy * T;
    int x1 = std::min(x0s + T, xres) - x0s, y1 = std::min(y0s + T, yres) - y0s;
    std::vector<double> tau, Pprev, pref;
    if (!a_eq_e) {
      tau.assign(size_t(x1) * y1 * 3, 0.0);
      Pprev.assign(size_t(x1) * y1 * 3, 0.0);
      pref.assign(size_t(x1) * y1 * 3, 0.0);
      for (int64_t m = off[chunk]; m < off[chunk + 1]; ++m) {
        const uint32_t i = idx[m];
        double sigma = h2s * sr[i], stp = -1.0 / (sigma * sigma);
        footprint_in_tile(fx[i], fy[i], fr[i], x0s, y0s, x1, y1, [&](int X, int Y) {
          double dx = X - sx[i], dy = Y - sy[i];
          double att = std::exp(stp * dy * dy) * std::exp(stp * dx * dx);
          double *t = &tau[(size_t(X - x0s) * y1 + (Y - y0s)) * 3];
          t[0] += de0[i] * att; t[1] += de1[i] * att; t[2] += de2[i] * att;
        });
      }
    }
    for (int64_t m = off[chunk]; m < off[chunk + 1]; ++m) {
      const uint32_t i = idx[m];
      double sigma = h2s * sr[i];
      double stp = -1.0 / (sigma * sigma);
      // d stp / d r = 2 / (h2s^2 r^3)
      double dstp_dr = 2.0 / (h2s * h2s * sr[i] * sr[i] * sr[i]);
      double e[3] = {de0[i], de1[i], de2[i]};
      double acc[6] = {0, 0, 0, 0, 0, 0};
      footprint_in_tile(fx[i], fy[i], fr[i], x0s, y0s, x1, y1, [&](int X, int Y) {
        double dx = X - sx[i], dy = Y - sy[i];
        double ex = std::exp(stp * dx * dx), ey = std::exp(stp * dy * dy);
        double att = ey * ex;
        // d att / d sx = att * (-2 stp dx) * (-1)... : d(dx)/d sx = -1
        double datt_dsx = att * (2.0 * stp * dx) * -1.0;
        double datt_dsy = att * (2.0 * stp * dy) * -1.0;
        double datt_dr = att * (dx * dx + dy * dy) * dstp_dr;
        const double *g = &G[(size_t(X) * yres + Y) * 3];
        double gatt = 0.0;
        if (a_eq_e) {
          for (int c = 0; c < 3; ++c) {
            gatt += g[c] * -e[c];
            acc[3 + c] += g[c] * -att;
          }
        } else {
          size_t li = (size_t(X - x0s) * y1 + (Y - y0s)) * 3;
          for (int c = 0; c < 3; ++c) {
            double q = e[c] / (e[c] + grayabsorb);
            double Tk = std::exp(-e[c] * att);
            double Pm = Pprev[li + c];
            pref[li + c] += e[c] * att;
            double S = std::exp(-(tau[li + c] - pref[li + c]));  // prod_{j>k} T_j
            double gk = g[c] * S;                                  // dL/dP_k
            // P_k = q + (P_{k-1} - q) T_k
            double dP_dT = Pm - q;
            double dP_dq = 1.0 - Tk;
            double dT_de = -att * Tk, dT_datt = -e[c] * Tk;
            double dq_de = grayabsorb / ((e[c] + grayabsorb) * (e[c] + grayabsorb));
            gatt += gk * dP_dT * dT_datt;
            acc[3 + c] += gk * (dP_dT * dT_de + dP_dq * dq_de);
            Pprev[li + c] = q + (Pm - q) * Tk;
          }
        }
        acc[0] += gatt * datt_dsx;
        acc[1] += gatt * datt_dsy;
        acc[2] += gatt * datt_dr;
      });
      for (int k = 0; k < 6; ++k) part[size_t(m) * 6 + k] = acc[k];
    }
  }
  // fixed-order reduction: tiles in chunk order
  for (int chunk = 0; chunk < ncx * ncy; ++chunk)
    for (int64_t m = off[chunk]; m < off[chunk + 1]; ++m) {
      const uint32_t i = idx[m];
      g_sx[i] += part[size_t(m) * 6 + 0];
      g_sy[i] += part[size_t(m) * 6 + 1];
      g_r[i] += part[size_t(m) * 6 + 2];
      g_e0[i] += part[size_t(m) * 6 + 3];
      g_e1[i] += part[size_t(m) * 6 + 4];
      g_e2[i] += part[size_t(m) * 6 + 5];
    }
}

int orc_num_threads() { return omp_get_max_threads(); }

}  // extern "C"
