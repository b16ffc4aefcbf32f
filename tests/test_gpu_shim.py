"""The Splotch-side binding on the GPU: hip_rendering (splotch_amd/shim/
hip_splotch.cc, the cuda_rendering replacement) called with GENUINE Splotch
objects — a paramfile, the colour maps get_colourmaps loads from palette
files, a vector<particle_sim> — beside the genuine host_rendering on the same
objects.  Additive frames are bit-exact; ordered frames are bit-exact against
the oracle in the stable (z, index) depth order, and their difference from
the genuine std::sort order (equal depths) is stated.  The records handed to
hip_rendering come back unmodified (checked in oracle/shim_driver.cpp).
"""
import math
import sys

import numpy as np
import pytest

import pyoracle as O
import splotch_amd as S
from helpers import assert_bitexact, max_rel_err
from splotch_amd import synth

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not O.shim_available(),
                                 reason="binding test library not built (make -C oracle shim)")]


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def shim_lib():
    yield O.shim()
    O.shim().shim_release()


@pytest.mark.parametrize("aeq", [True, False])
def test_hip_rendering_with_genuine_objects(shim_lib, tmp_path, aeq):
    log("case binding hip_rendering a_eq_e", aeq)
    pal_path = tmp_path / "hot.pal"
    pal_path.write_text(synth.palette_text(synth.hot_palette(256)))
    fld = synth.plummer_field(120_000, 2900, hsml_mu=math.log(0.5))
    view = synth.default_view(384, 320)
    kv = dict(xres=384, yres=320, fov=30, a_eq_e="T" if aeq else "F", brightness0=0.8,
              palette0=str(pal_path))
    rec = S.to_particle_sim(fld)
    got = O.shim_render(kv, rec, view, genuine=False)
    gen = O.shim_render(kv, rec, view, genuine=True)
    if aeq:
        assert_bitexact(got, gen, "hip_rendering vs host_rendering (additive)")
        return
    pal = synth.parse_palette(pal_path.read_text())
    want, _ = O.host_rendering(fld, view, [pal], a_eq_e=False, brightness=[0.8], stable=True)
    assert_bitexact(got, want, "hip_rendering vs oracle (stable depth order)")
    err = max_rel_err(got, gen)
    nd = int((got.view(np.uint32) != gen.view(np.uint32)).any(axis=2).sum())
    log("ordered: vs genuine std::sort order: max rel err %.3g, %d pixels differ" % (err, nd))
