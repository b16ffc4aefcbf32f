"""The oracle (oracle/splotch_oracle.cpp) against the genuine reference's fixtures.

Every fixture under tests/golden/ was written by the genuine Splotch host code
(tests/golden/make_golden.py); the oracle must reproduce each bit for bit.
"""
import numpy as np
import pytest

import pyoracle as O
from helpers import assert_bitexact, field_from, view_from
from splotch_amd import synth


def pal(golden):
    p = golden("palette")
    return (p["x"], p["rgb"])


def test_exptable_fixture(golden):
    f = golden("exptable")
    assert_bitexact(O.exp_op(f["arg"], False), f["op"], "exptable op")
    assert_bitexact(O.exp_op(f["arg"], True), f["expm1"], "exptable expm1")
    _, _, _, tl = O.exptable()
    assert np.float32(tl) == f["taylorlimit"][0]


def test_camera_fixture(golden):
    f = golden("camera")
    for i in range(int(f["nviews"][0])):
        v = view_from(f, "v%d_" % i)
        cam = O.camera(v)
        assert_bitexact(np.array(cam.mat[:], np.float32), f["v%d_mat" % i], "matrix %d" % i)
        assert cam.xshift == f["v%d_shift" % i][0] and cam.yshift == f["v%d_shift" % i][1]
        ff = f["v%d_f" % i]
        assert (np.float32(cam.res2), np.float32(cam.ycorr), np.float32(cam.xfac),
                np.float32(cam.fovfct), np.float32(cam.dist)) == tuple(ff)


@pytest.mark.parametrize("name", ["additive_small", "ordered_small", "ortho_zclip"])
def test_project_colorize_fixture(golden, name):
    f = golden(name)
    kw = {}
    if name == "ordered_small":
        kw["brightness"] = [float(f["brightness"][0])]
    if name == "ortho_zclip":
        kw.update(projection=False, zmin=120.0, zmax=260.0)
    p = O.project_colorize(field_from(f), view_from(f), [pal(golden)], **kw)
    for k, v in p.fields().items():
        assert_bitexact(v, f["p_" + k], "%s %s" % (name, k))


def test_project_colorize_two_types(golden):
    f = golden("two_types")
    pal1 = (f["pal1_x"], f["pal1_rgb"])
    p = O.project_colorize(field_from(f), view_from(f), [pal(golden), pal1],
                           smooth_fac=[1.3, 0.8], brightness=[0.7, 2.0], col_vector=[0, 1],
                           b_brightness=float(f["b_brightness"][0]), minrad_pix=1.5)
    for k, v in p.fields().items():
        assert_bitexact(v, f["p_" + k], "two_types " + k)


def _proj_from_fixture(f):
    p = O.Projected(len(f["p_x"]))
    for k in ("x", "y", "z", "r", "I", "e0", "e1", "e2", "active"):
        getattr(p, k)[:] = f["p_" + k]
    return p


@pytest.mark.parametrize("T", [50, 16, 8, 64])
def test_render_additive_fixture(golden, T):
    f = golden("additive_small")
    p = _proj_from_fixture(f)
    v = view_from(f)
    assert_bitexact(O.render(p, v.xres, v.yres, True, 0.2, T), f["pic_T%d" % T], "T=%d" % T)


def test_render_two_types_fixture(golden):
    f = golden("two_types")
    p = _proj_from_fixture(f)
    v = view_from(f)
    for T in (50, 24):
        assert_bitexact(O.render(p, v.xres, v.yres, True, 0.2, T), f["pic_T%d" % T])


def test_sort_fixture(golden):
    f = golden("ordered_small")
    p = _proj_from_fixture(f)
    assert_bitexact(O.sort_perm(p, 1, stable=False), f["perm_std"], "std::sort order")
    assert_bitexact(O.sort_perm(p, 1, stable=True), f["perm_stable"], "stable order")
    # stable order == numpy's stable lexsort on (z, index)
    assert_bitexact(np.argsort(p.z, kind="stable").astype(np.uint32), f["perm_stable"])


def test_sort_types_fixture(golden):
    f = golden("sort_types")
    p = O.Projected(len(f["z"]))
    p.z[:] = f["z"]; p.e0[:] = f["e0"]; p.r[:] = f["r"]
    for st in (1, 2, 3, 4):
        assert_bitexact(O.sort_perm(p, st, stable=False), f["perm%d" % st], "sort_type %d" % st)
    key = {1: p.z, 2: -p.e0, 3: p.e0, 4: -p.r}
    for st, k in key.items():
        assert_bitexact(O.sort_perm(p, st, stable=True),
                        np.argsort(k, kind="stable").astype(np.uint32), "stable %d" % st)
    with pytest.raises(ValueError):
        O.sort_perm(p, 7)


def test_render_ordered_fixture(golden):
    f = golden("ordered_small")
    p = _proj_from_fixture(f)
    v = view_from(f)
    for T in (50, 16):
        ps = p.take(f["perm_std"])
        assert_bitexact(O.render(ps, v.xres, v.yres, False, 0.2, T), f["pic_T%d" % T], "T%d" % T)
    ps = p.take(f["perm_stable"])
    assert_bitexact(O.render(ps, v.xres, v.yres, False, 0.2, 50), f["pic_stable_T50"])
    assert_bitexact(O.render(ps, v.xres, v.yres, False, 0.5, 50), f["pic_stable_T50_g05"])


def test_ortho_render_fixture(golden):
    f = golden("ortho_zclip")
    p = _proj_from_fixture(f)
    v = view_from(f)
    assert_bitexact(O.render(p, v.xres, v.yres, True, 0.2, 50), f["pic_T50"])


def test_host_rendering_fixture(golden):
    f = golden("host_rendering")
    v = view_from(f)
    pic, _ = O.host_rendering(field_from(f), v, [pal(golden)], a_eq_e=False, sort_type=1,
                              grayabsorb=0.3, tile=32, stable=False)
    assert_bitexact(pic, f["pic"], "host_rendering")


def test_postprocess_fixture(golden):
    """-expm1, colour_adjust, colorbars, Colour8 quantization and the TGA file:
    the oracle against the genuine code, bit for bit (NaN, inf and values
    outside the int range included)."""
    from helpers import post_args, pp_case
    f = golden("postprocess")
    for k in range(int(f["ncases"][0])):
        pic, kv, pals, want, tga = pp_case(f, pal(golden), k)
        got, gtga = O.postprocess(pic, pals, **post_args(kv))
        assert_bitexact(got, want, "post-processed image, case %d" % k)
        assert_bitexact(gtga, tga, "TGA file, case %d" % k)


def test_postprocess_closed_forms(golden):
    """Bar rows at palette nodes are the node colours; the 8-bit code of
    k/256 is k (255 for 1.0 and above, 0 below 0, for NaN and for +-inf,
    which the x86 conversion maps to INT_MIN)."""
    x, rgb = pal(golden)
    n = len(x)
    assert np.allclose(x, np.arange(n) / (n - 1))
    img = np.zeros((n - 1, 30, 3), np.float32)
    out, tga = O.postprocess(img, [(x, rgb)], a_eq_e=False, colorbar=True)
    for ix in range(n - 1):   # ix / double(xres) == x[ix] exactly
        for iy in range(20, 30):
            assert_bitexact(out[ix, iy], rgb[ix], "bar pixel %d,%d" % (ix, iy))
    assert not np.any(out[:, :20])
    vals = np.array([k / 256 for k in range(256)] + [1.0, 2.0, -1e-7, -3.0, np.nan, np.inf,
                                                      -np.inf, 3e9], np.float32)
    img = np.repeat(vals[:, None, None], 3, axis=2)
    _, tga = O.postprocess(img, [], a_eq_e=False)
    codes = tga[18::3]
    want = list(range(256)) + [255, 255, 0, 0, 0, 0, 0, 0]
    assert codes.tolist() == want


def test_normalize_fixture(golden):
    """particle_normalize: the oracle against the genuine code, bit for bit
    (logs through glibc's log10f, asinh in double, NaN never in a range), and
    the ranges it records (the parameters hold them with 8 digits)."""
    from helpers import nm_case
    f = golden("normalize")
    nt = int(f["ntypes"][0])
    for k in range(int(f["ncases"][0])):
        kv, fields, want, ranges = nm_case(f, k)
        got, out, _ = O.normalize(fields, kv, nt)
        for q in ("I", "er", "eg", "eb"):
            assert_bitexact(got[q], want[q], "case %d field %s" % (k, q))
        assert np.allclose(out[:, :4], ranges, rtol=1e-7, atol=0), (k, out, ranges)


def test_normalize_closed_forms():
    """Per type the smallest value maps to 0 and the largest to 1; a type with
    one value maps to 1; log10 of powers of ten is exact; non-positive
    values under a log become -38 and stay out of the range."""
    I = np.array([2.0, 5.0, 3.0, 7.0, 7.0, 1e-3, 1.0, 1e3, -4.0, 0.0], np.float32)
    er = np.array([0.5, 0.25, 1.0, 3.0, 3.0, 1.0, 10.0, 100.0, 0.0, 5.0], np.float32)
    typ = np.array([0, 0, 0, 1, 1, 2, 2, 2, 2, 2], np.uint8)
    z = np.zeros_like(I)
    f = dict(I=I, r=np.ones_like(I), er=er, eg=z, eb=z, type=typ)
    kv = dict(intensity_log2="true", color_log2="true")
    got, out, _ = O.normalize(f, kv, 3)
    assert got["I"][0] == 0 and got["I"][3 - 2] == 1 and got["I"][2] == np.float32(1 / 3)
    assert got["er"][1] == 0 and got["er"][2] == 1
    assert got["I"][3] == 1 and got["I"][4] == 1 and got["er"][3] == 1
    assert out[2, 4] == -3 and out[2, 5] == 3          # log10 of 1e-3 and 1e3
    assert out[2, 0] == 0 and out[2, 1] == 2           # log10 of 1 and 100 (0 excluded)
    assert got["I"][5] == 0 and got["I"][7] == 1 and got["I"][6] == 0.5
    assert got["I"][8] == 0 and got["I"][9] == 0       # -38 clamps to the range minimum
    assert got["er"][8] == 0


def test_bin_block_fixture(golden, tmp_path):
    """bin_reader_block: the numpy restatement against the genuine reader, bit
    for bit (byte orders, defaults of missing columns, smooth_factor, the
    NaN payloads of the r column quieted, the others copied)."""
    from helpers import bb_case
    f = golden("ingest")
    for k in range(int(f["ncases"][0])):
        kv, want = bb_case(f, k, tmp_path / ("b%d.bin" % k))
        got = O.read_bin_block(kv)
        for q, w in want.items():
            assert_bitexact(got[q], w, "case %d %s" % (k, q))
