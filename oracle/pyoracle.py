"""ORACLE — TEST INFRASTRUCTURE ONLY.

ctypes bindings for
  * ``oracle/liboracle.so``   — this project's C++ restatement of the Splotch
    host renderer (splotch_oracle.cpp), and
  * ``oracle/_ref/libsplref.so`` — the genuine reference host renderer compiled
    from its sources (Makefile target ``ref``), when it was built.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg use
this module, as the checker / CPU baseline.  The product path never does.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libsplref.so")

_lib = None
_ref = None


class OrcCamera(C.Structure):
    _fields_ = [("mat", C.c_float * 12), ("xshift", C.c_double), ("yshift", C.c_double),
                ("res2", C.c_float), ("ycorr", C.c_float), ("xfac", C.c_float),
                ("fovfct", C.c_float), ("dist", C.c_float), ("zminval", C.c_float),
                ("zmaxval", C.c_float), ("minrad_pix", C.c_float), ("xres", C.c_int),
                ("yres", C.c_int), ("projection", C.c_int)]


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.c_void_p)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        _lib = C.CDLL(ORACLE_SO)
        _lib.orc_bins.restype = C.c_int64
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("genuine reference library not built (oracle/_ref)")
        _ref = C.CDLL(REF_SO)
        _ref.ref_taylor_limit.restype = C.c_float
    return _ref


def _v3(t):
    return np.ascontiguousarray(np.asarray(t, dtype=np.float64).reshape(3))


# ------------------------------------------------------------------ exptable

def exptable():
    t1 = np.zeros(1024, np.float32)
    t2 = np.zeros(1024, np.float32)
    ef = C.c_float()
    tl = C.c_float()
    lib().orc_exptable(_p(t1), _p(t2), C.byref(ef), C.byref(tl))
    return t1, t2, ef.value, tl.value


def exp_op(arg, expm1=False, use_ref=False):
    arg = np.ascontiguousarray(arg, dtype=np.float32)
    out = np.zeros_like(arg)
    L = ref() if use_ref else lib()
    fn = L.ref_exp_op if use_ref else L.orc_exp_op
    fn(C.c_int64(arg.size), _p(arg), _p(out), C.c_int(int(expm1)))
    return out


# ------------------------------------------------------------------ camera

def camera(view, zmax=1e23, zmin=0.0, projection=True, minrad_pix=1.0):
    cam = OrcCamera()
    rc = lib().orc_camera(_p(_v3(view.campos)), _p(_v3(view.lookat)), _p(_v3(view.sky)),
                          _p(_v3(view.center)), C.c_int(view.xres), C.c_int(view.yres),
                          C.c_float(view.fov), C.c_float(zmax), C.c_float(zmin),
                          C.c_int(int(projection)), C.c_float(minrad_pix), C.byref(cam))
    if rc != 0:
        raise ValueError("degenerate matrix in TRANSMAT::Inverse()")
    return cam


def ref_camera(view):
    m = np.zeros(12, np.float32)
    d = np.zeros(2, np.float64)
    f = np.zeros(5, np.float32)
    ref().ref_camera(_p(_v3(view.campos)), _p(_v3(view.lookat)), _p(_v3(view.sky)),
                     _p(_v3(view.center)), C.c_int(view.xres), C.c_int(view.yres),
                     C.c_float(view.fov), _p(m), _p(d), _p(f))
    return m, d, f


# ------------------------------------------------------------------ project / colorize

class Projected:
    """SoA particle state after particle_project (+ particle_colorize)."""

    def __init__(self, n):
        self.x = np.zeros(n, np.float32)
        self.y = np.zeros(n, np.float32)
        self.z = np.zeros(n, np.float32)
        self.r = np.zeros(n, np.float32)
        self.I = np.zeros(n, np.float32)
        self.e0 = np.zeros(n, np.float32)
        self.e1 = np.zeros(n, np.float32)
        self.e2 = np.zeros(n, np.float32)
        self.type = np.zeros(n, np.uint8)
        self.active = np.zeros(n, np.uint8)

    @property
    def n(self):
        return self.x.shape[0]

    def take(self, perm):
        o = Projected(0)
        for k in ("x", "y", "z", "r", "I", "e0", "e1", "e2", "type", "active"):
            setattr(o, k, np.ascontiguousarray(getattr(self, k)[perm]))
        return o

    def fields(self):
        return {k: getattr(self, k) for k in ("x", "y", "z", "r", "I", "e0", "e1", "e2",
                                              "active")}


def _init_proj(field):
    n = field.n
    p = Projected(n)
    p.x[:] = field.x; p.y[:] = field.y; p.z[:] = field.z
    p.r[:] = field.hsml; p.I[:] = field.inten; p.e0[:] = field.col
    if field.c1 is not None:
        p.e1[:] = field.c1
    if field.c2 is not None:
        p.e2[:] = field.c2
    if field.type is not None:
        p.type[:] = field.type
    return p


def _pal_arrays(palettes):
    """palettes: list of (x float64[n], rgb float32[n,3]) per type."""
    nt = len(palettes)
    stride = max(len(x) for x, _ in palettes)
    px = np.zeros((nt, stride), np.float64)
    prgb = np.zeros((nt, stride, 3), np.float32)
    pn = np.zeros(nt, np.int32)
    for t, (x, rgb) in enumerate(palettes):
        px[t, :len(x)] = x
        prgb[t, :len(x)] = rgb
        pn[t] = len(x)
    return pn, stride, np.ascontiguousarray(px), np.ascontiguousarray(prgb)


def project_colorize(field, view, palettes, smooth_fac=None, brightness=None,
                     col_vector=None, b_brightness=1.0, zmax=1e23, zmin=0.0, projection=True,
                     minrad_pix=1.0, colorize=True):
    nt = len(palettes)
    sf = np.ones(8, np.float32) if smooth_fac is None else np.asarray(
        list(smooth_fac) + [1.0] * (8 - len(smooth_fac)), np.float32)
    br = np.ones(8, np.float32) if brightness is None else np.asarray(
        list(brightness) + [1.0] * (8 - len(brightness)), np.float32)
    # brightness{t} * b_brightness (splotch_host.cc:262-263), float32
    br = (br * np.float32(b_brightness)).astype(np.float32)
    cv = np.zeros(8, np.int32) if col_vector is None else np.asarray(
        list(col_vector) + [0] * (8 - len(col_vector)), np.int32)
    cam = camera(view, zmax, zmin, projection, minrad_pix)
    p = _init_proj(field)
    typ = _p(p.type) if field.type is not None else None
    lib().orc_project(C.byref(cam), C.c_int64(p.n), _p(p.x), _p(p.y), _p(p.z), _p(p.r),
                      _p(p.I), typ, _p(sf), _p(p.active))
    if colorize:
        pn, stride, px, prgb = _pal_arrays(palettes)
        lib().orc_colorize(C.c_int64(p.n), _p(p.e0), _p(p.e1), _p(p.e2), _p(p.I), typ,
                           _p(p.active), _p(br), _p(cv), _p(pn), C.c_int(stride), _p(px),
                           _p(prgb))
    return p


def kv_text(params: dict) -> str:
    return "".join("%s=%s\n" % (k, v) for k, v in params.items())


def ref_project_colorize(field, view, palettes, params: dict, b_brightness=1.0, colorize=True):
    p = _init_proj(field)
    pn, stride, px, prgb = _pal_arrays(palettes)
    rc = ref().ref_project_colorize(
        kv_text(params).encode(), C.c_int64(p.n), _p(p.x), _p(p.y), _p(p.z), _p(p.r), _p(p.I),
        _p(p.e0), _p(p.e1), _p(p.e2), _p(p.type), _p(p.active), _p(_v3(view.campos)),
        _p(_v3(view.lookat)), _p(_v3(view.sky)), _p(_v3(view.center)), C.c_int(len(palettes)),
        _p(pn), C.c_int(stride), _p(px), _p(prgb), C.c_float(b_brightness),
        C.c_int(int(colorize)))
    if rc != 0:
        raise RuntimeError("reference particle_project/colorize failed")
    return p


# ------------------------------------------------------------------ sort / bins / render

def sort_perm(p: Projected, sort_type=1, stable=True):
    perm = np.zeros(p.n, np.uint32)
    rc = lib().orc_sort_perm(C.c_int64(p.n), C.c_int(sort_type), _p(p.z), _p(p.e0), _p(p.r),
                             C.c_int(int(stable)), _p(perm))
    if rc != 0:
        raise ValueError("unknown sorting chosen ...")
    return perm


def ref_sort_perm(p: Projected, sort_type=1):
    perm = np.zeros(p.n, np.uint32)
    rc = ref().ref_sort(C.c_int64(p.n), _p(p.z), _p(p.e0), _p(p.r), C.c_int(sort_type),
                        _p(perm))
    if rc != 0:
        raise ValueError("reference particle_sort failed")
    return perm


def bins(p: Projected, xres, yres, tile):
    ncx = (xres + tile - 1) // tile
    ncy = (yres + tile - 1) // tile
    cnt = np.zeros(ncx * ncy, np.int32)
    total = lib().orc_bins(C.c_int64(p.n), _p(p.x), _p(p.y), _p(p.r), _p(p.active),
                           C.c_int(xres), C.c_int(yres), C.c_int(tile), _p(cnt), None, None)
    off = np.zeros(ncx * ncy + 1, np.int64)
    idx = np.zeros(max(total, 1), np.uint32)
    lib().orc_bins(C.c_int64(p.n), _p(p.x), _p(p.y), _p(p.r), _p(p.active), C.c_int(xres),
                   C.c_int(yres), C.c_int(tile), _p(cnt), _p(off), _p(idx))
    return off, idx[:off[-1]]


def render(p: Projected, xres, yres, a_eq_e=True, grayabsorb=0.2, tile=50):
    pic = np.zeros((xres, yres, 3), np.float32)
    lib().orc_render(C.c_int64(p.n), _p(p.x), _p(p.y), _p(p.r), _p(p.e0), _p(p.e1), _p(p.e2),
                     _p(p.active), C.c_int(xres), C.c_int(yres), C.c_int(int(a_eq_e)),
                     C.c_float(grayabsorb), C.c_int(tile), _p(pic))
    return pic


def fragments(p: Projected, xres, yres, tile=50):
    """The number of per-pixel updates render_new makes (orc_fragments)."""
    lib().orc_fragments.restype = C.c_int64
    return int(lib().orc_fragments(C.c_int64(p.n), _p(p.x), _p(p.y), _p(p.r), _p(p.active),
                                   C.c_int(xres), C.c_int(yres), C.c_int(tile)))


def render_bruteforce(p: Projected, xres, yres, tile=50):
    pic = np.zeros((xres, yres, 3), np.float32)
    lib().orc_render_bruteforce(C.c_int64(p.n), _p(p.x), _p(p.y), _p(p.r), _p(p.e0),
                                _p(p.e1), _p(p.e2), _p(p.active), C.c_int(xres),
                                C.c_int(yres), C.c_int(tile), _p(pic))
    return pic


def ref_render(p: Projected, xres, yres, a_eq_e=True, grayabsorb=0.2, tile=50):
    pic = np.zeros((xres, yres, 3), np.float32)
    rc = ref().ref_render_new(C.c_int64(p.n), _p(p.x), _p(p.y), _p(p.z), _p(p.r), _p(p.I),
                              _p(p.e0), _p(p.e1), _p(p.e2), _p(p.active), C.c_int(xres),
                              C.c_int(yres), C.c_int(int(a_eq_e)), C.c_float(grayabsorb),
                              C.c_int(tile), _p(pic))
    if rc != 0:
        raise RuntimeError("reference render_new failed")
    return pic


def ref_host_rendering(field, view, palettes, params: dict, b_brightness=1.0):
    pn, stride, px, prgb = _pal_arrays(palettes)
    n = field.n
    typ = field.type if field.type is not None else np.zeros(n, np.uint8)
    e1 = field.c1 if field.c1 is not None else np.zeros(n, np.float32)
    e2 = field.c2 if field.c2 is not None else np.zeros(n, np.float32)
    pic = np.zeros((view.xres, view.yres, 3), np.float32)
    rc = ref().ref_host_rendering(
        kv_text(params).encode(), C.c_int64(n), _p(field.x), _p(field.y), _p(field.z),
        _p(field.hsml), _p(field.inten), _p(field.col), _p(e1), _p(e2), _p(typ),
        _p(_v3(view.campos)), _p(_v3(view.lookat)), _p(_v3(view.sky)), _p(_v3(view.center)),
        C.c_int(len(palettes)), _p(pn), C.c_int(stride), _p(px), _p(prgb),
        C.c_float(b_brightness), C.c_int(view.xres), C.c_int(view.yres), _p(pic))
    if rc != 0:
        raise RuntimeError("reference host_rendering failed")
    return pic


def host_rendering(field, view, palettes, a_eq_e=True, sort_type=1, grayabsorb=0.2, tile=50,
                   stable=True, **kw):
    """Oracle restatement of splotch_host.cc:711-834 (project, colorize, sort, render).

    Returns (pic, projected particles in render order)."""
    p = project_colorize(field, view, palettes, **kw)
    if not a_eq_e:
        perm = sort_perm(p, sort_type, stable)
        p = p.take(perm)
    return render(p, view.xres, view.yres, a_eq_e, grayabsorb, tile), p


def postprocess(pic, palettes=(), a_eq_e=True, gamma=1.0, brightness=0.0, contrast=1.0,
                colorbar=False, col_vector=None):
    """Oracle restatement of the post-processing and TGA output of
    splotch.cc:272-310 (orc_postprocess).  pic: raw (xres, yres, 3) float32
    image.  Returns (post-processed float image, TGA file bytes)."""
    out = np.ascontiguousarray(pic, np.float32).copy()
    xres, yres = out.shape[:2]
    nt = len(palettes)
    cv = np.zeros(8, np.int32) if col_vector is None else np.asarray(
        list(col_vector) + [0] * (8 - len(col_vector)), np.int32)
    if nt:
        pn, stride, px, prgb = _pal_arrays(palettes)
    else:
        pn, stride, px, prgb = np.zeros(1, np.int32), 1, np.zeros(1), np.zeros(3, np.float32)
    tga = np.zeros(18 + 3 * xres * yres, np.uint8)
    lib().orc_postprocess(C.c_int(xres), C.c_int(yres), _p(out), C.c_int(int(a_eq_e)),
                          C.c_double(gamma), C.c_double(brightness), C.c_double(contrast),
                          C.c_int(int(colorbar)), C.c_int(nt), _p(cv), _p(pn), C.c_int(stride),
                          _p(px), _p(prgb), _p(tga))
    return out, tga


def ref_postprocess(pic, palettes, params: dict, tga_path):
    """The genuine post-processing (ref_postprocess): params are .par keys
    (a_eq_e, pic_gamma, pic_brightness, pic_contrast, colorbar, ptypes,
    color_is_vector{t}).  Returns (float image, TGA file bytes)."""
    out = np.ascontiguousarray(pic, np.float32).copy()
    xres, yres = out.shape[:2]
    pn, stride, px, prgb = _pal_arrays(palettes)
    tga = np.zeros(18 + 3 * xres * yres, np.uint8)
    rc = ref().ref_postprocess(kv_text(params).encode(), C.c_int(xres), C.c_int(yres), _p(out),
                               C.c_int(len(palettes)), _p(pn), C.c_int(stride), _p(px),
                               _p(prgb), str(tga_path).encode(), _p(tga))
    if rc != 0:
        raise RuntimeError("reference post-processing failed")
    return out, tga


BIN_FIELDS = ("x", "y", "z", "r", "I", "er", "eg", "eb")


def read_bin_block(kv: dict):
    """Restatement of bin_reader_block (reader/bin_reader.cc:34-88,147-190) in
    numpy: column q of the file (1-based numbers in kv) holds npart floats,
    npart = file size // (4 num_blocks); swap_endian (default true) reads them
    big-endian; r = smooth_factor * value (float product; a NaN comes out
    quiet with its payload) or smooth_factor; I = value or 0.5; C1 = value or
    1; C2, C3 = value or 0 (bit copies: 1.0 * value folds to a copy); type 0.
    A column number <= 0 counts as absent.  Returns a dict of float32 arrays (+ type uint8)."""
    truth = ("t", "y", "true", ".true.")
    nfields = int(kv["num_blocks"])
    swap = str(kv.get("swap_endian", "true")).lower() in truth
    smooth = np.float32(float(kv.get("smooth_factor", 1.0)))
    raw = np.fromfile(kv["infile"], dtype=">u4" if swap else "<u4")
    npart = raw.size // nfields
    keys = ("x", "y", "z", "r", "I", "C1", "C2", "C3")
    noval = (0.0, 0.0, 0.0, smooth, 0.5, 1.0, 0.0, 0.0)
    out = {}
    for q, (key, name) in enumerate(zip(keys, BIN_FIELDS)):
        # a column number <= 0 names no column (qty_idx < 0, bin_reader.cc:60-67)
        c = int(kv[key]) - 1 if key in kv else -1
        if c >= 0:
            b = raw[c * npart:(c + 1) * npart].astype(np.uint32)
            v = b.view(np.float32)
            if q == 3:
                nan = np.isnan(v)
                quiet = (b | np.uint32(0x00400000)).view(np.float32)
                with np.errstate(all="ignore"):
                    v = np.where(nan, quiet, (smooth * v).astype(np.float32))
            out[name] = np.ascontiguousarray(v, np.float32)
        else:
            out[name] = np.full(npart, noval[q], np.float32)
    out["type"] = np.zeros(npart, np.uint8)
    return out


def ref_read_bin_block(kv: dict, n_cap: int):
    """The genuine bin_reader_block (ref_bin_block).  Returns a dict as
    read_bin_block."""
    out = {k: np.zeros(n_cap, np.float32) for k in BIN_FIELDS}
    out["type"] = np.zeros(n_cap, np.uint8)
    n = C.c_int64(0)
    with open(os.devnull, "w"):
        rc = ref().ref_bin_block(kv_text(kv).encode(), C.c_int64(n_cap),
                                 *[_p(out[k]) for k in BIN_FIELDS], _p(out["type"]),
                                 C.byref(n))
    if rc != 0:
        raise RuntimeError("reference bin_reader_block failed")
    return {k: v[:n.value] for k, v in out.items()}


def ref_scene_file(kv: dict, cap: int = 4096):
    """The genuine sceneMaker scene list (ref_scene_file): per scene a dict
    campos, lookat, sky, center (3-tuples), fidx, keep_particles."""
    out = np.zeros((cap, 16), np.float64)
    n = C.c_int(0)
    rc = ref().ref_scene_file(kv_text(kv).encode(), _p(out), C.c_int(cap), C.byref(n))
    if rc != 0:
        raise RuntimeError("reference sceneMaker failed")
    res = []
    for o in out[:min(n.value, cap)]:
        res.append(dict(campos=tuple(o[0:3]), lookat=tuple(o[3:6]), sky=tuple(o[6:9]),
                        center=tuple(o[9:12]), fidx=float(o[12]), keep_particles=bool(o[13])))
    return res


NORM_FLAGS = ("intensity_log", "color_log", "color_asinh", "color_is_vector")
NORM_OVR = ("intensity_min", "intensity_max", "color_min", "color_max")


def norm_tables(kv: dict, ntypes: int):
    """Per-type flags / overrides of the normalization keys of a .par dict."""
    truth = ("t", "y", "true", ".true.")
    flags = np.zeros((ntypes, 4), np.int32)
    ovr_set = np.zeros((ntypes, 4), np.int32)
    ovr = np.zeros((ntypes, 4), np.float32)
    for t in range(ntypes):
        for k, key in enumerate(NORM_FLAGS):
            flags[t, k] = str(kv.get("%s%d" % (key, t), "false")).lower() in truth
        for k, key in enumerate(NORM_OVR):
            if "%s%d" % (key, t) in kv:
                ovr_set[t, k] = 1
                ovr[t, k] = np.float32(float(kv["%s%d" % (key, t)]))
    return flags, ovr_set, ovr


def normalize(fields: dict, kv: dict, ntypes: int, norm_int=True):
    """Oracle restatement of sceneMaker::particle_normalize (orc_normalize).
    fields: I, r, er, eg, eb (float32), type (uint8).  Returns (new fields,
    ranges float32[ntypes, 6], size ranges float32[ntypes, 2])."""
    f = {k: np.ascontiguousarray(fields[k], np.float32).copy() for k in ("I", "r", "er", "eg", "eb")}
    typ = np.ascontiguousarray(fields["type"], np.uint8)
    flags, ovr_set, ovr = norm_tables(kv, ntypes)
    out = np.zeros((ntypes, 6), np.float32)
    size = np.zeros((ntypes, 2), np.float32)
    lib().orc_normalize(C.c_int64(len(typ)), _p(f["I"]), _p(f["r"]), _p(f["er"]), _p(f["eg"]),
                        _p(f["eb"]), _p(typ), C.c_int(ntypes), _p(flags), _p(ovr_set), _p(ovr),
                        C.c_int(int(norm_int)), _p(out), _p(size))
    f["type"] = typ
    return f, out, size


def ref_normalize(fields: dict, kv: dict, ntypes: int):
    """The genuine sceneMaker::particle_normalize (ref_normalize).  Returns
    (new fields, float64[ntypes, 4] of the range_min, range_max, clamp_min,
    clamp_max parameters it sets)."""
    f = {k: np.ascontiguousarray(fields[k], np.float32).copy() for k in ("I", "r", "er", "eg", "eb")}
    typ = np.ascontiguousarray(fields["type"], np.uint8)
    ranges = np.zeros((ntypes, 4), np.float64)
    kv = dict(kv, ptypes=ntypes)
    rc = ref().ref_normalize(kv_text(kv).encode(), C.c_int64(len(typ)), _p(f["I"]), _p(f["r"]),
                             _p(f["er"]), _p(f["eg"]), _p(f["eb"]), _p(typ), C.c_int(ntypes),
                             _p(ranges))
    if rc != 0:
        raise RuntimeError("reference particle_normalize failed")
    f["type"] = typ
    return f, ranges


def render64(p: Projected, s, xres, yres, a_eq_e=True, grayabsorb=0.2, tile=50):
    """FP64 twin on the frozen float32 footprint; s: dict sx, sy, r, e0, e1, e2 (float64)."""
    pic = np.zeros((xres, yres, 3), np.float64)
    d = {k: np.ascontiguousarray(s[k], np.float64) for k in ("sx", "sy", "r", "e0", "e1", "e2")}
    lib().orc_render64(C.c_int64(p.n), _p(p.x), _p(p.y), _p(p.r), _p(p.active), _p(d["sx"]),
                       _p(d["sy"]), _p(d["r"]), _p(d["e0"]), _p(d["e1"]), _p(d["e2"]),
                       C.c_int(xres), C.c_int(yres), C.c_int(int(a_eq_e)),
                       C.c_double(grayabsorb), C.c_int(tile), _p(pic))
    return pic


def adjoint64(p: Projected, s, G, xres, yres, a_eq_e=True, grayabsorb=0.2, tile=50):
    d = {k: np.ascontiguousarray(s[k], np.float64) for k in ("sx", "sy", "r", "e0", "e1", "e2")}
    G = np.ascontiguousarray(G, np.float64)
    out = {k: np.zeros(p.n, np.float64) for k in ("sx", "sy", "r", "e0", "e1", "e2")}
    lib().orc_adjoint64(C.c_int64(p.n), _p(p.x), _p(p.y), _p(p.r), _p(p.active), _p(d["sx"]),
                        _p(d["sy"]), _p(d["r"]), _p(d["e0"]), _p(d["e1"]), _p(d["e2"]),
                        C.c_int(xres), C.c_int(yres), C.c_int(int(a_eq_e)),
                        C.c_double(grayabsorb), C.c_int(tile), _p(G), _p(out["sx"]),
                        _p(out["sy"]), _p(out["r"]), _p(out["e0"]), _p(out["e1"]),
                        _p(out["e2"]))
    return out


# ------------------------------------------------------------------ FP64 projection and chain

def _type_tables(field, palettes, smooth_fac, brightness, col_vector, b_brightness):
    def pad(v, d, dt):
        return np.asarray(list(v) + [d] * (8 - len(v)), dt) if v is not None else np.full(8, d, dt)
    sf = pad(smooth_fac, 1.0, np.float32)
    br = (pad(brightness, 1.0, np.float32) * np.float32(b_brightness)).astype(np.float32)
    cv = pad(col_vector, 0, np.int32)
    t = field.type.astype(np.int64) if field.type is not None else np.zeros(field.n, np.int64)
    return sf, br, cv, np.minimum(t, 7)


def _palette64(palettes, t, col):
    """Palette colour (linear_map.h:71-84, interpolation in double) and its
    slope d colour / d x at x = col, per particle; (n, 3) each."""
    n = len(col)
    rgb = np.zeros((n, 3))
    slope = np.zeros((n, 3))
    for tt, (xs, ys) in enumerate(palettes):
        sel = t == tt
        if not sel.any():
            continue
        xs = np.asarray(xs, np.float64)
        ys = np.asarray(ys, np.float32).reshape(-1, 3).astype(np.float64)
        xv = np.asarray(col[sel], np.float64)
        if len(xs) == 1:
            rgb[sel] = ys[0]
            continue
        idx = np.searchsorted(xs, xv, side="left") - 1  # lower_bound - 1
        idx = np.clip(idx, 0, len(xs) - 2)
        dxp = xs[idx + 1] - xs[idx]
        frac = (xv - xs[idx]) / dxp
        c = (1.0 - frac)[:, None] * ys[idx] + frac[:, None] * ys[idx + 1]
        s = (ys[idx + 1] - ys[idx]) / dxp[:, None]
        lo, hi = xv <= xs[0], xv >= xs[-1]
        c[lo] = ys[0]
        c[hi] = ys[-1]
        s[lo | hi] = 0.0
        rgb[sel] = c
        slope[sel] = s
    return rgb, slope


def project64(field, view, palettes, smooth_fac=None, brightness=None, col_vector=None,
              b_brightness=1.0, zmax=1e23, zmin=0.0, projection=True, minrad_pix=1.0):
    """particle_project + particle_colorize (splotch_host.cc:111-178, 252-287)
    evaluated in double for every particle (no culling): the screen-space state
    {sx, sy, r, e0, e1, e2} of the differentiable form, original order.  The
    camera (float32 matrix, double shifts) is the float renderer's."""
    cam = camera(view, zmax, zmin, projection, minrad_pix)
    m = np.array(cam.mat[:], np.float32).astype(np.float64)
    sf, br, cv, t = _type_tables(field, palettes, smooth_fac, brightness, col_vector, b_brightness)
    X, Y, Z = (np.asarray(a, np.float64) for a in (field.x, field.y, field.z))
    vx = X * m[0] + Y * m[1] + Z * m[2] + m[3]
    vy = X * m[4] + Y * m[5] + Z * m[6] + m[7]
    vz = X * m[8] + Y * m[9] + Z * m[10] + m[11]
    res2 = float(cam.res2)
    if projection:
        xfac2 = -res2 / (float(cam.fovfct) * vz)
        sx = res2 + cam.xshift + vx * xfac2
        sy = res2 + cam.yshift + float(cam.ycorr) + vy * xfac2
    else:
        xfac2 = np.full_like(vx, float(cam.xfac))
        sx = res2 + vx * xfac2
        sy = res2 + float(cam.ycorr) + vy * xfac2
    hs = np.asarray(field.hsml, np.float64)
    I = np.asarray(field.inten, np.float64)
    bfak = 0.5 * math.pi ** (-5.0 / 6.0)
    r2 = 2.0 * hs * xfac2 * sf[t].astype(np.float64)
    mr = float(np.float32(minrad_pix))
    r = np.sqrt(r2 * r2 + mr * mr)
    Ie = I * (0.5 * float(np.float32(bfak))) / hs * r2 / r
    vec = cv[t] != 0
    col = np.asarray(field.col, np.float64)
    rgb, _ = _palette64(palettes, t, np.asarray(field.col))
    if vec.any():
        rgb[vec, 0] = col[vec]
        rgb[vec, 1] = np.asarray(field.c1, np.float64)[vec]
        rgb[vec, 2] = np.asarray(field.c2, np.float64)[vec]
    e = rgb * (Ie * br[t].astype(np.float64))[:, None]
    return {"sx": sx, "sy": sy, "r": r, "e0": e[:, 0], "e1": e[:, 1], "e2": e[:, 2]}


def chain64(field, view, palettes, g, smooth_fac=None, brightness=None, col_vector=None,
            b_brightness=1.0, zmax=1e23, zmin=0.0, projection=True, minrad_pix=1.0):
    """Chain rule of project64: world-space gradients (x, y, z, hsml, I, col
    [, c1, c2]) from the screen-space gradient g (dict of sx, sy, r, e0..e2,
    original order), in double."""
    cam = camera(view, zmax, zmin, projection, minrad_pix)
    m = np.array(cam.mat[:], np.float32).astype(np.float64)
    sf, br, cv, t = _type_tables(field, palettes, smooth_fac, brightness, col_vector, b_brightness)
    X, Y, Z = (np.asarray(a, np.float64) for a in (field.x, field.y, field.z))
    vx = X * m[0] + Y * m[1] + Z * m[2] + m[3]
    vy = X * m[4] + Y * m[5] + Z * m[6] + m[7]
    vz = X * m[8] + Y * m[9] + Z * m[10] + m[11]
    res2 = float(cam.res2)
    xfac2 = -res2 / (float(cam.fovfct) * vz) if projection else np.full_like(vx, float(cam.xfac))
    hs = np.asarray(field.hsml, np.float64)
    I = np.asarray(field.inten, np.float64)
    bfak = 0.5 * math.pi ** (-5.0 / 6.0)
    k0 = 0.5 * float(np.float32(bfak))
    sm = sf[t].astype(np.float64)
    r2 = 2.0 * hs * xfac2 * sm
    mr = float(np.float32(minrad_pix))
    r = np.sqrt(r2 * r2 + mr * mr)
    I1 = I * k0 / hs
    Ie = I1 * r2 / r
    b = br[t].astype(np.float64)
    rgb, slope = _palette64(palettes, t, np.asarray(field.col))
    vec = cv[t] != 0
    if vec.any():
        rgb[vec, 0] = np.asarray(field.col, np.float64)[vec]
        rgb[vec, 1] = np.asarray(field.c1, np.float64)[vec]
        rgb[vec, 2] = np.asarray(field.c2, np.float64)[vec]
        slope[vec] = 0.0
    ge = np.stack([g["e0"], g["e1"], g["e2"]], axis=1)
    # e_c = rgb_c * Ie * b
    gIe = (ge * rgb).sum(axis=1) * b
    gr2 = g["r"] * (r2 / r) + gIe * (I1 * mr * mr / r ** 3)
    out = {"hsml": gr2 * (2.0 * xfac2 * sm) + gIe * (-Ie / hs),
           "inten": gIe * (k0 / hs) * (r2 / r),
           "col": np.where(vec, ge[:, 0] * Ie * b, (ge * slope).sum(axis=1) * Ie * b)}
    if vec.any():
        out["c1"] = np.where(vec, ge[:, 1] * Ie * b, 0.0)
        out["c2"] = np.where(vec, ge[:, 2] * Ie * b, 0.0)
    gvx = g["sx"] * xfac2
    gvy = g["sy"] * xfac2
    if projection:
        gvz = -(g["sx"] * vx + g["sy"] * vy) * xfac2 / vz - gr2 * r2 / vz
    else:
        gvz = np.zeros_like(gvx)
    out["x"] = m[0] * gvx + m[4] * gvy + m[8] * gvz
    out["y"] = m[1] * gvx + m[5] * gvy + m[9] * gvz
    out["z"] = m[2] * gvx + m[6] * gvy + m[10] * gvz
    return out


def frozen_model(field, view, palettes, a_eq_e=True, sort_type=1, **kw):
    """Float32 projected particles in render order (the frozen footprint) and
    the render-order permutation of the original particles."""
    p = project_colorize(field, view, palettes, **kw)
    if a_eq_e:
        return p, np.arange(p.n)
    perm = sort_perm(p, sort_type, True)
    return p.take(perm), perm


# ------------------------------------------------------------------ Splotch-side binding

SHIM_SO = os.path.join(HERE, "_ref", "libsplshim.so")
_shim = None


def shim_available() -> bool:
    return os.path.exists(SHIM_SO)


def shim():
    """The Splotch-side binding (splotch_amd/shim/hip_splotch.cc) compiled
    with the genuine reference objects (shim_driver.cpp)."""
    global _shim
    if _shim is None:
        if not shim_available():
            raise RuntimeError("binding test library not built (make -C oracle shim)")
        _shim = C.CDLL(SHIM_SO)
    return _shim


def shim_setup(kv: dict, view, b_brightness=1.0, stride=1024):
    """hip_setup on a genuine paramfile (kv, with palette{t} file paths) and
    the colour maps the genuine get_colourmaps loads: (params struct bytes,
    camera struct bytes, per-type palettes [(x, rgb)])."""
    from splotch_amd import capi
    prm, cam = capi.Params(), capi.Camera()
    nt = int(kv.get("ptypes", 1))
    pal_n = np.zeros(8, np.int32)
    pal_x = np.zeros(8 * stride, np.float64)
    pal_rgb = np.zeros(8 * stride * 3, np.float32)
    rc = shim().shim_setup(kv_text(kv).encode(), _p(_v3(view.campos)), _p(_v3(view.lookat)),
                           _p(_v3(view.sky)), _p(_v3(view.center)), C.c_float(b_brightness),
                           C.byref(prm), C.byref(cam), _p(pal_n), _p(pal_x), _p(pal_rgb),
                           C.c_int(stride))
    if rc != 0:
        raise RuntimeError("shim_setup failed (%d)" % rc)
    pals = [(pal_x[t * stride:t * stride + pal_n[t]].copy(),
             pal_rgb[3 * t * stride:3 * (t * stride + pal_n[t])].reshape(-1, 3).copy())
            for t in range(nt)]
    return prm, cam, pals


def shim_render(kv: dict, records: np.ndarray, view, genuine: bool, b_brightness=1.0):
    """One frame of particle_sim records through hip_rendering (the GPU
    renderer) or the genuine host_rendering, both called with genuine
    Splotch objects.  Returns the x-major RGB image."""
    pic = np.zeros((view.xres, view.yres, 3), np.float32)
    rec = np.ascontiguousarray(records)
    rc = shim().shim_render(kv_text(kv).encode(), C.c_int64(len(rec)), _p(rec),
                            _p(_v3(view.campos)), _p(_v3(view.lookat)), _p(_v3(view.sky)),
                            _p(_v3(view.center)), C.c_float(b_brightness), C.c_int(view.xres),
                            C.c_int(view.yres), C.c_int(int(genuine)), _p(pic))
    if rc != 0:
        raise RuntimeError("shim_render failed (%d)" % rc)
    return pic
