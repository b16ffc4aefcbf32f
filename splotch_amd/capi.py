"""ctypes binding of the C-ABI in include/splotch_hip.h (libsplotch_hip.so).

The Python layer mirrors the reference's render interface for this path
(host_rendering / cuda_rendering, splotch/splotch_host.cc:711-834 and
cuda/cuda_splotch.cc:121-192): a context bound to one device, palettes per
particle type, a camera derived from campos/lookat/sky/centerpos and the .par
keys, and a render call that writes the raw x-major RGB image.  Device memory
and streams come from torch; the work is done by the HIP kernels.  If the
library is missing, every entry point raises: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsplotch_hip.so")
MAX_TYPES = 8

_lib = None


class SplhipError(RuntimeError):
    pass


class Camera(C.Structure):
    _fields_ = [("mat", C.c_float * 12), ("xshift", C.c_double), ("yshift", C.c_double),
                ("res2", C.c_float), ("ycorr", C.c_float), ("xfac", C.c_float),
                ("fovfct", C.c_float), ("dist", C.c_float), ("zminval", C.c_float),
                ("zmaxval", C.c_float), ("minrad_pix", C.c_float), ("xres", C.c_int32),
                ("yres", C.c_int32), ("projection", C.c_int32)]


class Params(C.Structure):
    _fields_ = [("ptypes", C.c_int32), ("smooth_fac", C.c_float * MAX_TYPES),
                ("brightness", C.c_float * MAX_TYPES), ("col_vector", C.c_int32 * MAX_TYPES),
                ("a_eq_e", C.c_int32 * MAX_TYPES), ("sort_type", C.c_int32),
                ("gray_absorption", C.c_float), ("tile_size", C.c_int32),
                ("hip_region", C.c_int32), ("hip_split", C.c_int32),
                ("hip_variant", C.c_int32), ("hip_heavy", C.c_int32),
                ("hip_heavy_max", C.c_int32), ("hip_chunk_entries", C.c_int32),
                ("hip_autotune", C.c_int32), ("hip_bin_group", C.c_int32),
                ("hip_ppw", C.c_int32), ("hip_stage", C.c_int32),
                ("hip_stage_ordered", C.c_int32)]


class Post(C.Structure):
    """splhip_post: the epilogue switches (splotch.cc:272-310)."""
    _fields_ = [("a_eq_e", C.c_int32), ("colorbar", C.c_int32), ("ptypes", C.c_int32),
                ("col_vector", C.c_int32 * MAX_TYPES), ("gamma", C.c_double),
                ("brightness", C.c_double), ("contrast", C.c_double)]


class Norm(C.Structure):
    """splhip_norm: particle_normalize switches (scenemaker.cc:40-296)."""
    _fields_ = [("ptypes", C.c_int32)] + [
        (k, C.c_int32 * MAX_TYPES) for k in (
            "intensity_log", "color_log", "color_asinh", "col_vector", "has_intensity_min",
            "has_intensity_max", "has_color_min", "has_color_max")] + [
        (k, C.c_float * MAX_TYPES) for k in (
            "intensity_min", "intensity_max", "color_min", "color_max")] + [
        ("norm_intensity", C.c_int32)]


class NormRanges(C.Structure):
    """splhip_norm_ranges: per-type ranges found and applied."""
    _fields_ = [(k, C.c_float) for k in ("range_min", "range_max", "clamp_min", "clamp_max",
                                         "intensity_min", "intensity_max", "size_min",
                                         "size_max")]


class Scene(C.Structure):
    """splhip_scene: one scene of a scene file."""
    _fields_ = [("campos", C.c_double * 3), ("lookat", C.c_double * 3), ("sky", C.c_double * 3),
                ("center", C.c_double * 3), ("fidx", C.c_double), ("keep_particles", C.c_int32),
                ("line", C.c_int32)]


class Particles(C.Structure):
    _fields_ = [("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("hsml", C.c_void_p),
                ("inten", C.c_void_p), ("col", C.c_void_p), ("c1", C.c_void_p),
                ("c2", C.c_void_p), ("type", C.c_void_p), ("n", C.c_int64)]


class Job(C.Structure):
    """splhip_job: one view of a batch."""
    _fields_ = [("parts", Particles), ("cam", Camera), ("prm", Params), ("pic", C.c_void_p)]


class Grads(C.Structure):
    """splhip_grads: device gradient buffers."""
    _fields_ = [("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("hsml", C.c_void_p),
                ("inten", C.c_void_p), ("col", C.c_void_p), ("c1", C.c_void_p),
                ("c2", C.c_void_p), ("screen", C.c_void_p)]


# numpy dtype of the reference particle_sim record (splotch/splotchutils.h:52-72)
PARTICLE_SIM = np.dtype([("er", "<f4"), ("eg", "<f4"), ("eb", "<f4"), ("x", "<f4"),
                         ("y", "<f4"), ("z", "<f4"), ("r", "<f4"), ("I", "<f4"),
                         ("type", "<u2"), ("active", "u1"), ("pad_", "u1")])

EXPORTS = ["splhip_create", "splhip_destroy", "splhip_last_error", "splhip_camera_setup",
           "splhip_params_default", "splhip_params_from_par", "splhip_set_palette",
           "splhip_render", "splhip_render_host", "splhip_project", "splhip_bins",
           "splhip_exptable", "splhip_set_profiling", "splhip_stage_times", "splhip_set_trace",
           "splhip_trace_items", "splhip_render_batched", "splhip_render_backward",
           "splhip_last_tuning", "splhip_last_binning", "splhip_post_from_par",
           "splhip_postprocess", "splhip_norm_from_par", "splhip_normalize",
           "splhip_bin_block_count", "splhip_read_bin_block", "splhip_scene_file_parse",
           "splhip_capture_reserve", "splhip_capture_status", "splhip_normalize_noalloc"]

STAGES = ("project", "ordered_sort", "ordered_bin", "ordered_splat", "additive_compact",
          "additive_bin", "additive_splat", "ordered_k_splat8", "ordered_k_splat_heavy",
          "additive_k_splat8", "additive_k_splat_heavy", "ordered_staged", "additive_staged")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SplhipError("libsplotch_hip.so is not built (python -m splotch_amd.build)")
        L = C.CDLL(LIB_PATH)
        L.splhip_last_error.restype = C.c_char_p
        L.splhip_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
        L.splhip_destroy.argtypes = [C.c_void_p]
        L.splhip_last_error.argtypes = [C.c_void_p]
        L.splhip_set_palette.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
                                         C.c_void_p]
        L.splhip_render.argtypes = [C.c_void_p, C.POINTER(Particles), C.POINTER(Camera),
                                    C.POINTER(Params), C.c_void_p, C.c_void_p]
        L.splhip_render_host.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(Camera),
                                         C.POINTER(Params), C.c_void_p, C.c_void_p]
        L.splhip_project.argtypes = [C.c_void_p, C.POINTER(Particles), C.POINTER(Camera),
                                     C.POINTER(Params), C.c_void_p, C.c_void_p, C.c_void_p]
        L.splhip_bins.argtypes = [C.c_void_p, C.POINTER(Particles), C.POINTER(Camera),
                                  C.POINTER(Params), C.c_int32, C.c_void_p, C.c_void_p,
                                  C.c_int64, C.POINTER(C.c_int64), C.c_void_p,
                                  C.POINTER(C.c_int64), C.c_void_p]
        L.splhip_render_batched.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]
        L.splhip_capture_reserve.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_float,
                                             C.c_void_p]
        L.splhip_capture_status.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
        L.splhip_render_backward.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(Grads),
                                             C.c_void_p]
        L.splhip_set_profiling.argtypes = [C.c_void_p, C.c_int32]
        L.splhip_stage_times.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def _np(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _v3(t):
    return np.ascontiguousarray(np.asarray(t, dtype=np.float64).reshape(3))


def exptable():
    tab = np.zeros(2048, np.float32)
    k = np.zeros(5, np.float32)
    rc = lib().splhip_exptable(_np(tab), _np(k))
    if rc != 0:
        raise SplhipError("splhip_exptable failed")
    return tab, k


def camera_setup(view, zmin=0.0, zmax=1e23, projection=True, minrad_pix=1.0) -> Camera:
    """Camera of splotch_host.cc:62-109 for a synth.View-like object."""
    cam = Camera()
    rc = lib().splhip_camera_setup(_np(_v3(view.campos)), _np(_v3(view.lookat)),
                                   _np(_v3(view.sky)), _np(_v3(view.center)),
                                   C.c_int32(view.xres), C.c_int32(view.yres),
                                   C.c_float(view.fov), C.c_float(zmin), C.c_float(zmax),
                                   C.c_int32(int(projection)), C.c_float(minrad_pix),
                                   C.byref(cam))
    if rc == -3:
        raise SplhipError("degenerate matrix in TRANSMAT::Inverse()")
    if rc != 0:
        raise SplhipError("splhip_camera_setup failed (%d)" % rc)
    return cam


def params(ptypes=1, smooth_fac=None, brightness=None, col_vector=None, a_eq_e=True,
           sort_type=1, gray_absorption=0.2, tile_size=50, hip_region=0, hip_split=0, hip_variant=0,
           hip_heavy=0, hip_heavy_max=0, hip_chunk_entries=0, hip_autotune=0, hip_bin_group=0,
           hip_ppw=0, hip_stage=0, hip_stage_ordered=0,
           b_brightness=1.0) -> Params:
    """Render parameters; brightness is brightness{t} (times b_brightness, as
    splotch_host.cc:262-263 does, in float32).  a_eq_e per type: True / 1
    additive, False / 0 ordered, 2 both passes."""
    p = Params()
    lib().splhip_params_default(C.byref(p))
    p.ptypes = ptypes
    for t in range(MAX_TYPES):
        if smooth_fac is not None and t < len(smooth_fac):
            p.smooth_fac[t] = smooth_fac[t]
        b = brightness[t] if brightness is not None and t < len(brightness) else 1.0
        p.brightness[t] = float(np.float32(b) * np.float32(b_brightness))
        if col_vector is not None and t < len(col_vector):
            p.col_vector[t] = int(bool(col_vector[t]))
        # 1 additive, 0 ordered, 2 both passes
        mode = lambda v: 2 if (not isinstance(v, bool) and v == 2) else int(bool(v))
        if isinstance(a_eq_e, (list, tuple)):
            p.a_eq_e[t] = mode(a_eq_e[t]) if t < len(a_eq_e) else 1
        else:
            p.a_eq_e[t] = mode(a_eq_e)
    p.sort_type = sort_type
    p.gray_absorption = gray_absorption
    p.tile_size = tile_size
    p.hip_region = hip_region
    p.hip_split = hip_split
    p.hip_variant = hip_variant
    p.hip_heavy = hip_heavy
    p.hip_heavy_max = hip_heavy_max
    p.hip_chunk_entries = hip_chunk_entries
    p.hip_autotune = hip_autotune
    p.hip_bin_group = hip_bin_group
    p.hip_ppw = hip_ppw
    p.hip_stage = hip_stage
    p.hip_stage_ordered = hip_stage_ordered
    return p


def params_from_par(text: str, b_brightness: float = 1.0):
    """Read the hot-path .par keys (see include/splotch_hip.h)."""
    p = Params()
    sc = np.zeros(7, np.float32)
    rc = lib().splhip_params_from_par(text.encode(), C.c_float(b_brightness), C.byref(p),
                                      _np(sc))
    if rc != 0:
        raise SplhipError("bad parameter text")
    keys = ("xres", "yres", "fov", "zmin", "zmax", "projection", "minrad_pix")
    scal = dict(zip(keys, sc.tolist()))
    scal["xres"] = int(scal["xres"])
    scal["yres"] = int(scal["yres"])
    scal["projection"] = bool(scal["projection"])
    return p, scal


def post(a_eq_e=True, gamma=1.0, brightness=0.0, contrast=1.0, colorbar=False, ptypes=1,
         col_vector=None) -> Post:
    """splhip_post from keyword values (the .par defaults)."""
    q = Post()
    q.a_eq_e = int(bool(a_eq_e))
    q.gamma, q.brightness, q.contrast = float(gamma), float(brightness), float(contrast)
    q.colorbar = int(bool(colorbar))
    q.ptypes = int(ptypes)
    for t, v in enumerate(col_vector or []):
        q.col_vector[t] = int(v)
    return q


def post_from_par(text: str) -> Post:
    """The post-processing keys of a .par text (splhip_post_from_par)."""
    q = Post()
    if lib().splhip_post_from_par(text.encode(), C.byref(q)) != 0:
        raise SplhipError("bad parameter text")
    return q


def norm_from_par(text: str) -> Norm:
    """The particle_normalize keys of a .par text (splhip_norm_from_par)."""
    q = Norm()
    if lib().splhip_norm_from_par(text.encode(), C.byref(q)) != 0:
        raise SplhipError("bad parameter text")
    return q


def bin_block_count(text: str) -> int:
    """Particles of the block file a .par text names (splhip_bin_block_count)."""
    n = C.c_int64(0)
    if lib().splhip_bin_block_count(text.encode(), C.byref(n)) != 0:
        raise SplhipError("bad block-file keys or unreadable file")
    return n.value


def scene_file_parse(par_text: str, scene_text: str):
    """The scenes of a scene file (splhip_scene_file_parse): list of dicts
    campos, lookat, sky, center, fidx, keep_particles, line."""
    n = C.c_int32(0)
    if lib().splhip_scene_file_parse(par_text.encode(), scene_text.encode(), None, 0,
                                     C.byref(n)) != 0:
        raise SplhipError("bad scene file or parameters")
    arr = (Scene * max(n.value, 1))()
    if lib().splhip_scene_file_parse(par_text.encode(), scene_text.encode(), arr, n.value,
                                     C.byref(n)) != 0:
        raise SplhipError("bad scene file or parameters")
    return [dict(campos=tuple(s.campos), lookat=tuple(s.lookat), sky=tuple(s.sky),
                 center=tuple(s.center), fidx=s.fidx, keep_particles=bool(s.keep_particles),
                 line=s.line) for s in arr[:n.value]]


def scene_views(par_text: str, scene_text: str, b_brightness: float = 1.0):
    """Per scene of a scene file: (Camera, Params, fidx, keep_particles), the
    scene's keys applied over par_text as getNextScene applies them."""
    import types
    scenes = scene_file_parse(par_text, scene_text)
    lines = scene_text.splitlines()
    first = lines[0].split() if lines else []
    try:
        [float(w) for w in first]
        keys = ("camera_x camera_y camera_z lookat_x lookat_y lookat_z sky_x sky_y sky_z "
                "fidx").split() if len(first) == 10 else first
    except ValueError:
        keys = first
    text = par_text
    out = []
    for s in scenes:
        vals = lines[s["line"]].split()
        text = text + "\n" + "\n".join("%s=%s" % kv for kv in zip(keys, vals))
        prm, sc = params_from_par(text, b_brightness)
        v = types.SimpleNamespace(campos=s["campos"], lookat=s["lookat"], sky=s["sky"],
                                  center=s["center"], xres=sc["xres"], yres=sc["yres"],
                                  fov=sc["fov"])
        cam = camera_setup(v, zmin=sc["zmin"], zmax=sc["zmax"], projection=sc["projection"],
                           minrad_pix=sc["minrad_pix"])
        out.append((cam, prm, s["fidx"], s["keep_particles"]))
    return out


def tga_header(xres: int, yres: int) -> bytes:
    """The 18-byte header LS_Image::write_TGA writes (ls_image.cc:183-186)."""
    return bytes([0, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0, xres % 256, xres // 256, yres % 256,
                  yres // 256, 24, 32])


class Renderer:
    """One context on one device (cuda_rendercontext_init + cuda_renderer_init)."""

    def __init__(self, device: int = 0):
        self._c = C.c_void_p()
        rc = lib().splhip_create(C.c_int(device), C.byref(self._c))
        if rc != 0:
            raise SplhipError("splhip_create failed (%d)" % rc)
        self.device = device

    def close(self):
        if self._c:
            lib().splhip_destroy(self._c)
            self._c = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().splhip_last_error(self._c)
            raise SplhipError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def set_profiling(self, on: bool = True):
        self._check(lib().splhip_set_profiling(self._c, int(on)), "set_profiling")

    def stage_times(self):
        """Last frame's per-stage GPU ms (dict) and pass counts."""
        ms = np.zeros(len(STAGES), np.float32)
        cnt = np.zeros(4, np.int64)
        self._check(lib().splhip_stage_times(self._c, _np(ms), _np(cnt)), "stage_times")
        counts = dict(ordered_active=int(cnt[0]), ordered_entries=int(cnt[1]),
                      additive_active=int(cnt[2]), additive_entries=int(cnt[3]))
        return dict(zip(STAGES, ms.tolist())), counts

    def binning(self):
        """Region entries each pass (ordered, additive) of the last render was
        binned through; 0 where it sorted its block entries directly."""
        v = np.zeros(2, np.int64)
        self._check(lib().splhip_last_binning(self._c, _np(v)), "splhip_last_binning")
        return int(v[0]), int(v[1])

    def tuning(self):
        """The splat tuning the last render used: dict variant, split,
        heavy_max, whether it came from the autotuner, the work region
        (region, px) and particles per wave (ppw, 0 = auto)."""
        v = np.zeros(6, np.int64)
        self._check(lib().splhip_last_tuning(self._c, _np(v)), "splhip_last_tuning")
        return {"variant": int(v[0]), "split": int(v[1]), "heavy_max": int(v[2]),
                "autotuned": bool(v[3]), "region": int(v[4]), "ppw": int(v[5])}

    def set_trace(self, on: bool = True):
        self._check(lib().splhip_set_trace(self._c, C.c_int32(int(on))), "set_trace")

    def trace_items(self, ordered: bool = True):
        """Last frame's splat work items: array (n, 4) of start, end (100 MHz
        ticks), list length | channel << 32, CU id."""
        n = C.c_int64()
        L = lib()
        L.splhip_trace_items(self._c, C.c_int32(int(ordered)), None, C.c_int64(0), C.byref(n))
        out = np.zeros((max(n.value, 1), 4), np.uint64)
        L.splhip_trace_items(self._c, C.c_int32(int(ordered)), _np(out), C.c_int64(n.value),
                             C.byref(n))
        return out[:n.value]

    def set_palette(self, ptype: int, x, rgb):
        x = np.ascontiguousarray(x, np.float64)
        rgb = np.ascontiguousarray(rgb, np.float32).reshape(-1, 3)
        self._check(lib().splhip_set_palette(self._c, ptype, len(x), _np(x), _np(rgb)),
                    "set_palette")

    @staticmethod
    def particles(dev: dict) -> Particles:
        """dev: dict of contiguous cuda tensors x,y,z,hsml,inten,col[,c1,c2,type]."""
        P = Particles()
        n = int(dev["x"].numel())
        for k in ("x", "y", "z", "hsml", "inten", "col", "c1", "c2", "type"):
            t = dev.get(k)
            if t is not None:
                assert t.is_contiguous() and t.numel() == n, k
                setattr(P, k, t.data_ptr())
        P.n = n
        return P

    @staticmethod
    def _stream(stream):
        if stream is None:
            import torch
            return C.c_void_p(torch.cuda.current_stream().cuda_stream)
        return C.c_void_p(stream)

    def render(self, dev: dict, cam: Camera, prm: Params, pic, stream=None):
        """pic: cuda float32 tensor of cam.xres*cam.yres*3 elements (x-major RGB)."""
        assert pic.is_contiguous() and pic.numel() == cam.xres * cam.yres * 3
        P = self.particles(dev)
        self._check(lib().splhip_render(self._c, C.byref(P), C.byref(cam), C.byref(prm),
                                        C.c_void_p(pic.data_ptr()), self._stream(stream)),
                    "splhip_render")
        return pic

    def postprocess(self, pic, xres: int, yres: int, q: Post, out=None, rgb8=None,
                    layout: int = 1, stream=None):
        """The epilogue over a device image: out (float32 tensor, may be pic)
        and/or rgb8 (uint8 tensor of 3*xres*yres; layout 1 = TGA body,
        0 = x-major RGB).  Returns (out, rgb8)."""
        n = xres * yres * 3
        assert pic.is_contiguous() and pic.numel() == n
        for t in (out, rgb8):
            assert t is None or (t.is_contiguous() and t.numel() == n)
        self._check(lib().splhip_postprocess(
            self._c, C.c_void_p(pic.data_ptr()), C.c_int32(xres), C.c_int32(yres), C.byref(q),
            C.c_void_p(out.data_ptr() if out is not None else None),
            C.c_void_p(rgb8.data_ptr() if rgb8 is not None else None), C.c_int32(layout),
            self._stream(stream)), "splhip_postprocess")
        return out, rgb8

    def normalize(self, dev: dict, q: Norm, stream=None, ranges: bool = True,
                  noalloc: bool = False):
        """particle_normalize in place on device fields (dev: inten, col, c1,
        c2, hsml, type as cuda tensors).  Returns the per-type ranges (list of
        dicts; the call then waits for the stream), or None with
        ranges=False.  noalloc: through splhip_normalize_noalloc."""
        n = int(dev["x"].numel()) if "x" in dev else int(dev["inten"].numel())

        def ptr(k):
            t = dev.get(k)
            return C.c_void_p(t.data_ptr() if t is not None else None)
        rr = (NormRanges * MAX_TYPES)() if ranges else None
        fn = lib().splhip_normalize_noalloc if noalloc else lib().splhip_normalize
        self._check(fn(
            self._c, C.c_int64(n), ptr("inten"), ptr("col"), ptr("c1"), ptr("c2"), ptr("hsml"),
            ptr("type"), C.byref(q), rr, self._stream(stream)),
            "splhip_normalize_noalloc" if noalloc else "splhip_normalize")
        if not ranges:
            return None
        return [{k: getattr(rr[t], k) for k, _ in NormRanges._fields_} for t in range(q.ptypes)]

    def read_bin_block(self, text: str, stream=None):
        """bin_reader_block into new device tensors: dict x, y, z, hsml,
        inten, col, c1, c2 (float32) and type (uint8), the SoA the render
        takes."""
        import torch
        n = bin_block_count(text)
        d = {k: torch.empty(n, dtype=torch.float32, device="cuda")
             for k in ("x", "y", "z", "hsml", "inten", "col", "c1", "c2")}
        d["type"] = torch.empty(n, dtype=torch.uint8, device="cuda")
        self._check(lib().splhip_read_bin_block(
            self._c, text.encode(), C.c_int64(n),
            *[C.c_void_p(d[k].data_ptr()) for k in ("x", "y", "z", "hsml", "inten", "col", "c1",
                                                    "c2", "type")],
            self._stream(stream)), "splhip_read_bin_block")
        return d

    def render_scene_file(self, particles_of, par_text: str, scene_text: str, stream=None):
        """The frames of a scene file: consecutive scenes that share their
        particles (keep_particles) go through one splhip_render_batched
        call.  particles_of(fidx) returns the device fields of a fidx.
        Returns the images (cuda float32 tensors, x-major RGB)."""
        import torch
        frames = scene_views(par_text, scene_text)
        images, group = [], []

        def flush():
            if not group:
                return
            dev = particles_of(group[0][2])
            views = []
            for cam, prm, _, _ in group:
                pic = torch.empty(cam.xres * cam.yres * 3, dtype=torch.float32, device="cuda")
                views.append((dev, cam, prm, pic))
            self.render_batched(views, stream)
            images.extend(v[3] for v in views)
            group.clear()

        for f in frames:
            group.append(f)
            if not f[3]:  # the next scene loads other particles
                flush()
        flush()
        return images

    def _jobs(self, views):
        jobs = (Job * max(len(views), 1))()
        for j, (dev, cam, prm, pic) in enumerate(views):
            assert pic.is_contiguous() and pic.numel() == cam.xres * cam.yres * 3
            jobs[j].parts = self.particles(dev)
            jobs[j].cam = cam
            jobs[j].prm = prm
            jobs[j].pic = pic.data_ptr()
        return jobs

    def render_batched(self, views, stream=None):
        """views: list of (dev, cam, prm, pic) as for render(); all images in
        one pipeline (on a capturing stream: the reserved batch with device
        counts).  Returns the list of images."""
        jobs = self._jobs(views)
        self._check(lib().splhip_render_batched(self._c, C.c_int32(len(views)),
                                                C.cast(jobs, C.c_void_p),
                                                self._stream(stream)), "splhip_render_batched")
        return [v[3] for v in views]

    def capture_reserve(self, views, headroom: float = 1.25, stream=None):
        """Reserve the batch `views` for rendering under stream capture
        (splhip_capture_reserve); renders it (twice).  Returns the images."""
        jobs = self._jobs(views)
        self._check(lib().splhip_capture_reserve(self._c, C.c_int32(len(views)),
                                                 C.cast(jobs, C.c_void_p), C.c_float(headroom),
                                                 self._stream(stream)),
                    "splhip_capture_reserve")
        return [v[3] for v in views]

    def capture_status(self) -> int:
        """Overflow bits of the last device-count render (bit 0 ordered pass,
        bit 1 additive pass); waits for the device."""
        v = C.c_int32(0)
        self._check(lib().splhip_capture_status(self._c, C.byref(v)), "splhip_capture_status")
        return int(v.value)

    def backward(self, grad_pic, n: int, vector: bool = False, screen: bool = False,
                 stream=None):
        """Gradient of the last render() with respect to the particle fields,
        given grad_pic = dL/dpic (cuda float32, same layout as pic).  Returns a
        dict of cuda tensors x, y, z, hsml, inten, col [, c1, c2] [, screen
        (n, 6)]."""
        import torch
        assert grad_pic.is_contiguous() and grad_pic.dtype == torch.float32
        dev = grad_pic.device
        keys = ["x", "y", "z", "hsml", "inten", "col"] + (["c1", "c2"] if vector else [])
        out = {k: torch.empty(max(n, 1), dtype=torch.float32, device=dev) for k in keys}
        if screen:
            out["screen"] = torch.empty((max(n, 1), 6), dtype=torch.float32, device=dev)
        g = Grads()
        for k, t in out.items():
            setattr(g, k, t.data_ptr())
        self._check(lib().splhip_render_backward(self._c, C.c_void_p(grad_pic.data_ptr()),
                                                 C.byref(g), self._stream(stream)),
                    "splhip_render_backward")
        return {k: t[:n] for k, t in out.items()}

    def render_host(self, aos: np.ndarray, cam: Camera, prm: Params, stream=None):
        """cuda_rendering semantics: host particle_sim records in, host image out."""
        aos = np.ascontiguousarray(aos)
        assert aos.dtype == PARTICLE_SIM
        pic = np.zeros((cam.xres, cam.yres, 3), np.float32)
        self._check(lib().splhip_render_host(self._c, _np(aos), C.c_int64(len(aos)),
                                             C.byref(cam), C.byref(prm), _np(pic),
                                             self._stream(stream)),
                    "splhip_render_host")
        return pic

    def project(self, dev: dict, cam: Camera, prm: Params, stream=None):
        import torch
        P = self.particles(dev)
        rec = torch.empty((max(P.n, 1), 8), dtype=torch.float32, device=dev["x"].device)
        act = torch.empty((max(P.n, 1),), dtype=torch.uint8, device=dev["x"].device)
        self._check(lib().splhip_project(self._c, C.byref(P), C.byref(cam), C.byref(prm),
                                         C.c_void_p(rec.data_ptr()), C.c_void_p(act.data_ptr()),
                                         self._stream(stream)), "splhip_project")
        return rec[:P.n], act[:P.n]

    def bins(self, dev: dict, cam: Camera, prm: Params, ordered: bool, stream=None):
        """Reference-tile CSR (tile_off int64, idx = original indices) and render order."""
        P = self.particles(dev)
        T = prm.tile_size
        ntiles = ((cam.xres + T - 1) // T) * ((cam.yres + T - 1) // T)
        ne = C.c_int64()
        no = C.c_int64()
        self._check(lib().splhip_bins(self._c, C.byref(P), C.byref(cam), C.byref(prm),
                                      int(ordered), None, None, 0, C.byref(ne), None,
                                      C.byref(no), self._stream(stream)), "splhip_bins")
        off = np.zeros(ntiles + 1, np.int64)
        idx = np.zeros(max(ne.value, 1), np.uint32)
        order = np.zeros(max(no.value, 1), np.uint32)
        self._check(lib().splhip_bins(self._c, C.byref(P), C.byref(cam), C.byref(prm),
                                      int(ordered), _np(off), _np(idx), C.c_int64(len(idx)),
                                      C.byref(ne), _np(order), C.byref(no),
                                      self._stream(stream)), "splhip_bins")
        if no.value == 0:
            off[:] = 0
        return off, idx[:ne.value], order[:no.value]


def to_device(field, device="cuda"):
    """synth.Field -> dict of cuda tensors (SoA)."""
    import torch
    d = {}
    for k, v in field.arrays().items():
        if v is not None:
            d[k] = torch.from_numpy(np.ascontiguousarray(v)).to(device)
    return d


def to_particle_sim(field) -> np.ndarray:
    """synth.Field -> reference particle_sim records (e.r = colour scalar)."""
    a = np.zeros(field.n, PARTICLE_SIM)
    a["x"], a["y"], a["z"], a["r"], a["I"] = field.x, field.y, field.z, field.hsml, field.inten
    a["er"] = field.col
    if field.c1 is not None:
        a["eg"], a["eb"] = field.c1, field.c2
    if field.type is not None:
        a["type"] = field.type
    return a
