"""C-ABI library: loads without a GPU, exports every declared symbol, and its
host-only functions (camera setup, exp tables, .par parsing) agree with the
oracle bit for bit."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import pyoracle as O
import splotch_amd as S
from helpers import assert_bitexact, view_from

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            syms |= set(re.findall(r"^\s*(?:int|void|const char \*)\s*\**\s*(splhip_\w+)\s*\(",
                                   txt, re.M))
    return syms


def test_library_exports_declared_symbols():
    L = S.capi.lib()
    syms = declared_symbols()
    assert len(syms) >= 12
    for s in sorted(syms):
        assert hasattr(L, s), s
    assert set(S.capi.EXPORTS) <= syms


def test_exptable_matches_oracle():
    tab, k = S.exptable()
    t1, t2, expfac, tl = O.exptable()
    assert_bitexact(tab[:1024], t1)
    assert_bitexact(tab[1024:], t2)
    assert k[0] == np.float32(expfac) and k[1] == np.float32(tl)
    assert k[4] == np.float32(0.75)


def test_camera_setup_matches_reference_fixture(golden):
    f = golden("camera")
    for i in range(int(f["nviews"][0])):
        v = view_from(f, "v%d_" % i)
        cam = S.camera_setup(v)
        assert_bitexact(np.array(cam.mat[:], np.float32), f["v%d_mat" % i])
        assert cam.xshift == f["v%d_shift" % i][0] and cam.yshift == f["v%d_shift" % i][1]
        assert np.float32(cam.xfac) == f["v%d_f" % i][2]
        oc = O.camera(v)
        assert (cam.res2, cam.ycorr, cam.fovfct, cam.dist) == (oc.res2, oc.ycorr, oc.fovfct,
                                                               oc.dist)


def test_camera_degenerate_raises():
    from splotch_amd import synth
    with pytest.raises(S.SplhipError):
        S.camera_setup(synth.View((50, 50, -150), (50, 50, 50), (0, 0, 1), 64, 64))


def test_params_from_par_defaults_and_keys():
    p, sc = S.params_from_par("")
    assert sc == dict(xres=800, yres=800, fov=45.0, zmin=0.0, zmax=np.float32(1e23),
                      projection=True, minrad_pix=1.0)
    assert p.ptypes == 1 and p.sort_type == 1 and p.tile_size == 50
    assert abs(p.gray_absorption - 0.2) < 1e-7 and p.a_eq_e[0] == 1
    txt = """# comment
    xres = 640
    fov=30   # trailing comment
    ptypes=2
    brightness0=0.5
    brightness1=2.0
    smooth_factor1 = 0.8
    color_is_vector1 = T
    a_eq_e = false
    a_eq_e1 = .true.
    tile_size_cpu=32.0
    gray_absorption=0.3
    sort_type=2
    projection=FALSE
    """
    p, sc = S.params_from_par(txt, b_brightness=1.5)
    assert sc["xres"] == 640 and sc["yres"] == 640 and sc["fov"] == 30 and not sc["projection"]
    assert p.ptypes == 2 and p.a_eq_e[0] == 0 and p.a_eq_e[1] == 1
    assert p.brightness[0] == np.float32(np.float32(0.5) * np.float32(1.5))
    assert p.brightness[1] == np.float32(3.0)
    assert abs(p.smooth_fac[1] - 0.8) < 1e-7 and p.col_vector[1] == 1 and p.col_vector[0] == 0
    assert p.tile_size == 32 and p.sort_type == 2
    with pytest.raises(S.SplhipError):
        S.params_from_par("projection=1")        # not a bool in the reference's parser
    with pytest.raises(S.SplhipError):
        S.params_from_par("xres=12abc")


def test_post_from_par_defaults_and_keys(golden):
    """The epilogue keys (splotch.cc:186,307; splotchutils.cc:232-234) read as
    the oracle's tests read them from the fixture's .par excerpts."""
    from helpers import parse_kv, post_args
    q = S.post_from_par("")
    assert (q.a_eq_e, q.colorbar, q.ptypes, q.gamma, q.brightness, q.contrast) == (
        1, 0, 1, 1.0, 0.0, 1.0)
    f = golden("postprocess")
    for k in range(int(f["ncases"][0])):
        text = str(f["pp%d_kv" % k])
        q = S.post_from_par(text)
        want = post_args(parse_kv(text))
        assert bool(q.a_eq_e) == want["a_eq_e"] and bool(q.colorbar) == want["colorbar"]
        assert (q.gamma, q.brightness, q.contrast) == (want["gamma"], want["brightness"],
                                                       want["contrast"])
        assert list(q.col_vector[:q.ptypes]) == want["col_vector"]
    with pytest.raises(S.SplhipError):
        S.post_from_par("colorbar=2")


def test_norm_from_par_keys():
    """The particle_normalize keys (scenemaker.cc:51-55, 223-233)."""
    q = S.norm_from_par("")
    assert q.ptypes == 1 and q.norm_intensity == 1
    assert not q.intensity_log[0] and not q.color_log[0] and not q.has_color_min[0]
    q = S.norm_from_par("ptypes=3\nintensity_log1=T\ncolor_log2=true\ncolor_asinh0=.true.\n"
                        "color_is_vector2=t\ncolor_min0=0.25\nintensity_max2=5\n")
    assert q.ptypes == 3 and q.intensity_log[1] and q.color_log[2] and q.color_asinh[0]
    assert q.col_vector[2] and not q.col_vector[0]
    assert q.has_color_min[0] and q.color_min[0] == 0.25 and not q.has_color_max[0]
    assert q.has_intensity_max[2] and q.intensity_max[2] == 5.0
    with pytest.raises(S.SplhipError):
        S.norm_from_par("intensity_log0=maybe")


def test_bin_block_count(golden, tmp_path):
    """The particle count of a block file and the key checks (host only)."""
    from helpers import bb_case
    f = golden("ingest")
    kv, want = bb_case(f, 0, tmp_path / "b.bin")
    import pyoracle as O
    assert S.bin_block_count(O.kv_text(kv)) == want["x"].size
    # column numbers <= 0 name no column (bin_reader.cc:60-67), as in the
    # reference's own .par files ("C2 = -1")
    assert S.bin_block_count(O.kv_text(dict(kv, C2=-1, r=-1, I=0))) == want["x"].size
    with pytest.raises(S.SplhipError):   # a column past num_blocks
        S.bin_block_count(O.kv_text(dict(kv, C3=99)))
    with pytest.raises(S.SplhipError):   # no such file
        S.bin_block_count(O.kv_text(dict(kv, infile=str(tmp_path / "missing.bin"))))


def test_scene_file_parse_fixture(golden):
    """The scene list and per-scene cameras against the genuine sceneMaker
    (old and new scene-file formats, scene_start / scene_incr / last_scene,
    center and sky from the parameters, fidx and its 1e-5 reuse test)."""
    f = golden("scenes")
    for k in range(int(f["ncases"][0])):
        got = S.scene_file_parse(str(f["sc%d_par" % k]), str(f["sc%d_text" % k]))
        want = f["sc%d_out" % k]
        assert len(got) == len(want), k
        for g, w in zip(got, want):
            row = list(g["campos"]) + list(g["lookat"]) + list(g["sky"]) + list(g["center"]) + \
                [g["fidx"], float(g["keep_particles"])]
            assert np.array_equal(np.array(row), w), (k, row, w)
    with pytest.raises(S.SplhipError):   # a scene line with a missing value
        S.scene_file_parse(str(f["sc0_par"]), "camera_x camera_y\n1 2\n3\n")


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors in capi.py have the C header's layout: a C program
    compiled against include/splotch_hip.h prints sizeof and every field's
    offsetof, compared with ctypes.sizeof and the ctypes field offsets."""
    import shutil
    import subprocess
    from splotch_amd import capi
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {"splhip_camera": capi.Camera, "splhip_params": capi.Params,
               "splhip_particles": capi.Particles, "splhip_job": capi.Job,
               "splhip_grads": capi.Grads, "splhip_post": capi.Post,
               "splhip_norm": capi.Norm, "splhip_norm_ranges": capi.NormRanges,
               "splhip_scene": capi.Scene}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "splotch_hip.h"',
             "int main(void) {"]
    for name, cls in structs.items():
        lines.append('printf("%s sizeof %%zu\\n", sizeof(%s));' % (name, name))
        for fname, _ in cls._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));'
                         % (name, fname, name, fname))
    lines.append('printf("splhip_particle_sim sizeof %zu\\n", sizeof(splhip_particle_sim));')
    for fname in capi.PARTICLE_SIM.names:
        lines.append('printf("splhip_particle_sim %s %%zu\\n", offsetof(splhip_particle_sim, %s));'
                     % (fname, fname))
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for ln in out.splitlines():
        name, fname, v = ln.split()
        got[(name, fname)] = int(v)
    for name, cls in structs.items():
        assert got[(name, "sizeof")] == C.sizeof(cls), name
        for fname, _ in cls._fields_:
            assert got[(name, fname)] == getattr(cls, fname).offset, (name, fname)
    assert got[("splhip_particle_sim", "sizeof")] == capi.PARTICLE_SIM.itemsize
    for fname in capi.PARTICLE_SIM.names:
        assert got[("splhip_particle_sim", fname)] == capi.PARTICLE_SIM.fields[fname][1], fname
