"""The Splotch-side binding (splotch_amd/shim/hip_splotch.cc), compiled
against the reference headers and called with GENUINE Splotch objects (a
paramfile, the colour maps get_colourmaps loads from palette files, vec3
cameras): the render keys, camera and palettes it hands to the C-ABI equal
what the library derives from the same .par text and palette text, bit for
bit.  Host only (no device call)."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as O
import splotch_amd as S
from splotch_amd import synth

pytestmark = pytest.mark.skipif(not O.shim_available(),
                                reason="binding test library not built (make -C oracle shim)")


def _pal_file(tmp_path, name, rgb255):
    path = tmp_path / name
    path.write_text(synth.palette_text(rgb255, name))
    return str(path)


@pytest.mark.parametrize("case", range(3))
def test_binding_setup_matches_library(tmp_path, case):
    hot = synth.hot_palette(256)
    ramp = np.stack([np.arange(7) * 40, 255 - np.arange(7) * 40, np.full(7, 128)], 1)
    kv = [
        dict(xres=320, yres=200, fov=30, palette0=_pal_file(tmp_path, "hot.pal", hot)),
        dict(xres=256, fov=45, ptypes=2, a_eq_e="F", brightness0=0.7, brightness1=2.5,
             smooth_factor1=1.3, sort_type=2, gray_absorption=0.35, tile_size_cpu=32,
             palette0=_pal_file(tmp_path, "hot.pal", hot),
             palette1=_pal_file(tmp_path, "ramp.pal", ramp), minrad_pix=1.5, hip_ppw=4),
        dict(xres=128, yres=96, fov=20, ptypes=3, projection="F", zmin=10, zmax=400,
             color_is_vector1="T", palette0=_pal_file(tmp_path, "ramp.pal", ramp),
             palette2=_pal_file(tmp_path, "hot.pal", hot), a_eq_e1="F"),
    ][case]
    view = [synth.default_view(320, 200), synth.orbit_view(0.4, 0.3, 150.0, 256),
            synth.View((10.0, 20.0, -100.0), (50.0, 40.0, 50.0), (0.0, 1.0, 0.0), 128, 96,
                       20.0, (12.0, 21.0, -99.0))][case]
    b_bright = [1.0, 0.8, 1.7][case]
    prm, cam, pals = O.shim_setup(kv, view, b_bright)
    want_prm, keys = S.params_from_par(O.kv_text(kv), b_bright)
    for name, _ in S.capi.Params._fields_:  # field by field: padding bytes are unset
        f = getattr(S.capi.Params, name)
        assert bytes(prm)[f.offset:f.offset + f.size] == \
            bytes(want_prm)[f.offset:f.offset + f.size], name
    kview = synth.View(view.campos, view.lookat, view.sky, keys["xres"], keys["yres"],
                       keys["fov"], view.centerpos)
    want_cam = S.camera_setup(kview, projection=bool(keys["projection"]), zmin=keys["zmin"],
                              zmax=keys["zmax"], minrad_pix=keys["minrad_pix"])
    for name, _ in S.capi.Camera._fields_:
        f = getattr(S.capi.Camera, name)
        assert bytes(cam)[f.offset:f.offset + f.size] == \
            bytes(want_cam)[f.offset:f.offset + f.size], name
    for t in range(int(kv.get("ptypes", 1))):
        key = "palette%d" % t
        if key not in kv or kv.get("color_is_vector%d" % t) == "T":
            assert len(pals[t][0]) == 0
            continue
        x, rgb = synth.parse_palette(open(kv[key]).read())
        assert np.array_equal(pals[t][0], x)
        assert np.array_equal(pals[t][1].view(np.uint32), rgb.astype(np.float32).view(np.uint32))
