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
