"""GPU parity of the plain render path, through the C-ABI.

Bar: bit-exact against the genuine reference fixtures and the oracle for
projected particles, tile lists and raw images (the per-pixel summation order
of render_new is reproduced).  Ordered mode uses the stable (z, index) depth
order; the genuine std::sort order differs only between equal keys, and the
error against it is stated (max per-channel relative error, floor 1e-6*max).
"""
import math
import sys

import numpy as np
import pytest

import pyoracle as O
from helpers import assert_bitexact, field_from, max_rel_err, view_from
from splotch_amd import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def R():
    import splotch_amd as S
    assert torch.cuda.is_available(), "GPU test without a GPU"
    r = S.Renderer(0)
    yield r
    torch.cuda.synchronize()
    r.close()


def palette(golden):
    p = golden("palette")
    return p["x"], p["rgb"]


def render(R, field, view, pals, prm, **camkw):
    import splotch_amd as S
    for t, (x, rgb) in enumerate(pals):
        R.set_palette(t, x, rgb)
    cam = S.camera_setup(view, **camkw)
    dev = S.to_device(field)
    pic = torch.empty(view.xres * view.yres * 3, dtype=torch.float32, device="cuda")
    R.render(dev, cam, prm, pic)
    torch.cuda.synchronize()
    return pic.cpu().numpy().reshape(view.xres, view.yres, 3)


CASES = {
    "additive_small": dict(),
    "ordered_small": dict(brightness=None),
    "ortho_zclip": dict(cam=dict(projection=False, zmin=120.0, zmax=260.0)),
}


@pytest.mark.parametrize("name", ["additive_small", "ordered_small", "ortho_zclip", "two_types"])
def test_project_colorize_bitexact(R, golden, name):
    import splotch_amd as S
    log("case project", name)
    f = golden(name)
    pals = [palette(golden)]
    camkw = {}
    kw = {}
    if name == "ordered_small":
        kw["brightness"] = [float(f["brightness"][0])]
    if name == "ortho_zclip":
        camkw = dict(projection=False, zmin=120.0, zmax=260.0)
    if name == "two_types":
        pals.append((f["pal1_x"], f["pal1_rgb"]))
        kw = dict(ptypes=2, smooth_fac=[1.3, 0.8], brightness=[0.7, 2.0], col_vector=[0, 1],
                  b_brightness=float(f["b_brightness"][0]))
        camkw = dict(minrad_pix=1.5)
    for t, (x, rgb) in enumerate(pals):
        R.set_palette(t, x, rgb)
    v = view_from(f)
    cam = S.camera_setup(v, **camkw)
    prm = S.params(**kw)
    field = field_from(f)
    rec, act = R.project(S.to_device(field), cam, prm)
    rec = rec.cpu().numpy()
    act = (act.cpu().numpy() != 0).astype(np.uint8)
    assert_bitexact(act, f["p_active"], "active")
    for k, col in (("x", 0), ("y", 1), ("r", 2), ("z", 6), ("I", 7)):
        assert_bitexact(rec[:, col], f["p_" + k], k)
    for k, col in (("e0", 3), ("e1", 4), ("e2", 5)):
        a = rec[:, col]
        b = f["p_" + k]
        m = act == 1
        assert_bitexact(a[m], b[m], k)


@pytest.mark.parametrize("T", [50, 16, 8, 64])
def test_bins_bitexact(R, golden, T):
    import splotch_amd as S
    log("case bins T", T)
    f = golden("additive_small")
    R.set_palette(0, *palette(golden))
    v = view_from(f)
    prm = S.params(tile_size=T)
    off, idx, order = R.bins(S.to_device(field_from(f)), S.camera_setup(v), prm, ordered=False)
    p = O.Projected(len(f["p_x"]))
    for k in ("x", "y", "z", "r", "I", "e0", "e1", "e2", "active"):
        getattr(p, k)[:] = f["p_" + k]
    ooff, oidx = O.bins(p, v.xres, v.yres, T)
    assert_bitexact(off, ooff, "tile offsets")
    assert_bitexact(idx, oidx, "tile lists")
    assert_bitexact(order, np.nonzero(p.active)[0].astype(np.uint32), "order")


def test_bins_ordered_mode(R, golden):
    import splotch_amd as S
    f = golden("ordered_small")
    R.set_palette(0, *palette(golden))
    v = view_from(f)
    prm = S.params(a_eq_e=False, brightness=[float(f["brightness"][0])])
    off, idx, order = R.bins(S.to_device(field_from(f)), S.camera_setup(v), prm, ordered=True)
    p = O.Projected(len(f["p_x"]))
    for k in ("x", "y", "z", "r", "I", "e0", "e1", "e2", "active"):
        getattr(p, k)[:] = f["p_" + k]
    perm = f["perm_stable"]
    act_sorted = perm[p.active[perm] == 1]
    assert_bitexact(order, act_sorted, "depth order")
    ps = p.take(perm)
    ooff, oidx = O.bins(ps, v.xres, v.yres, 50)
    assert_bitexact(off, ooff, "tile offsets")
    assert_bitexact(idx, perm[oidx], "tile lists (original indices)")


@pytest.mark.parametrize("G", [8, 16, 32, 64])
@pytest.mark.parametrize("T", [50, 16, 8, 64])
def test_render_additive_fixture(R, golden, T, G):
    import splotch_amd as S
    log("case additive T", T, "G", G)
    f = golden("additive_small")
    pic = render(R, field_from(f), view_from(f), [palette(golden)],
                 S.params(tile_size=T, hip_region=G))
    assert_bitexact(pic, f["pic_T%d" % T], "raw pic T=%d G=%d" % (T, G))


@pytest.mark.parametrize("G", [8, 16, 32, 64])
def test_render_ordered_fixture(R, golden, G):
    import splotch_amd as S
    log("case ordered G", G)
    f = golden("ordered_small")
    v = view_from(f)
    fld = field_from(f)
    b = [float(f["brightness"][0])]
    pic = render(R, fld, v, [palette(golden)], S.params(a_eq_e=False, brightness=b, hip_region=G))
    assert_bitexact(pic, f["pic_stable_T50"], "ordered stable order")
    pic = render(R, fld, v, [palette(golden)],
                 S.params(a_eq_e=False, brightness=b, gray_absorption=0.5, hip_region=G))
    assert_bitexact(pic, f["pic_stable_T50_g05"], "ordered g=0.5")
    pic = render(R, fld, v, [palette(golden)],
                 S.params(a_eq_e=False, brightness=b, tile_size=16, hip_region=G))
    err = max_rel_err(pic, f["pic_T16"])
    log("ordered T16 vs genuine std::sort order: max rel err %.3g" % err)
    assert err < 1e-3


def test_render_two_types_fixture(R, golden):
    import splotch_amd as S
    f = golden("two_types")
    pals = [palette(golden), (f["pal1_x"], f["pal1_rgb"])]
    kw = dict(ptypes=2, smooth_fac=[1.3, 0.8], brightness=[0.7, 2.0], col_vector=[0, 1],
              b_brightness=float(f["b_brightness"][0]))
    for T in (50, 24):
        pic = render(R, field_from(f), view_from(f), pals, S.params(tile_size=T, **kw),
                     minrad_pix=1.5)
        assert_bitexact(pic, f["pic_T%d" % T], "two types T=%d" % T)


def test_render_ortho_fixture(R, golden):
    import splotch_amd as S
    f = golden("ortho_zclip")
    pic = render(R, field_from(f), view_from(f), [palette(golden)], S.params(),
                 projection=False, zmin=120.0, zmax=260.0)
    assert_bitexact(pic, f["pic_T50"], "orthographic")


def test_render_host_records(R, golden):
    """cuda_rendering semantics: host particle_sim records -> host image."""
    import splotch_amd as S
    f = golden("host_rendering")
    v = view_from(f)
    fld = field_from(f)
    R.set_palette(0, *palette(golden))
    aos = S.to_particle_sim(fld)
    before = aos.copy()
    prm = S.params(a_eq_e=False, tile_size=32, gray_absorption=0.3)
    pic = R.render_host(aos, S.camera_setup(v), prm)
    assert np.array_equal(aos.view(np.uint8), before.view(np.uint8)), "host records modified"
    opic, _ = O.host_rendering(fld, v, [palette(golden)], a_eq_e=False, tile=32, grayabsorb=0.3,
                               stable=True)
    assert_bitexact(pic, opic, "host path vs oracle (stable order)")
    err = max_rel_err(pic, f["pic"])
    log("host_rendering fixture (std::sort order): max rel err %.3g" % err)
    assert err < 1e-3


def test_render_host_records_vector_colour(R, golden):
    """Host records with a vector-coloured type (all three colour components
    uploaded) and more than one upload chunk: the genuine two-type fixture;
    and a 600k-record field equal to its device-resident render."""
    import splotch_amd as S
    f = golden("two_types")
    pals = [palette(golden), (f["pal1_x"], f["pal1_rgb"])]
    for t, (x, rgb) in enumerate(pals):
        R.set_palette(t, x, rgb)
    kw = dict(ptypes=2, smooth_fac=[1.3, 0.8], brightness=[0.7, 2.0], col_vector=[0, 1],
              b_brightness=float(f["b_brightness"][0]))
    fld = field_from(f)
    pic = R.render_host(S.to_particle_sim(fld), S.camera_setup(view_from(f), minrad_pix=1.5),
                        S.params(tile_size=50, **kw))
    assert_bitexact(pic, f["pic_T50"], "host records, vector colour")
    big = synth.concat(synth.plummer_field(300_000, 1580, ptype=0),
                       synth.plummer_field(300_000, 1581, ptype=1, hsml_fixed=0.05))
    v = synth.default_view(320, 256)
    pal = synth.default_palette()
    prm = S.params(ptypes=2, a_eq_e=[0, 1], brightness=[0.5, 2.0])
    want = render(R, big, v, [pal, pal], prm)
    got = R.render_host(S.to_particle_sim(big), S.camera_setup(v), prm)
    assert_bitexact(got.reshape(want.shape), want, "host records, 600k, 5 chunks")


@pytest.mark.parametrize("seed,n,res,aeq,T,G", [
    (101, 200_000, (512, 384), True, 50, 32),
    (102, 200_000, (384, 512), False, 50, 32),
    (103, 100_000, (300, 300), True, 37, 16),
    (104, 100_000, (256, 256), False, 64, 64),
])
def test_render_random_vs_oracle(R, seed, n, res, aeq, T, G):
    import splotch_amd as S
    log("case random", seed, n, res, aeq, T, G)
    fld = synth.plummer_field(n, seed, hsml_mu=math.log(0.5))
    v = synth.default_view(*res)
    pal = synth.default_palette()
    pic = render(R, fld, v, [pal], S.params(a_eq_e=aeq, tile_size=T, hip_region=G))
    opic, _ = O.host_rendering(fld, v, [pal], a_eq_e=aeq, tile=T, stable=True)
    assert_bitexact(pic, opic, "random field vs oracle")


def test_sort_types_vs_oracle(R):
    import splotch_amd as S
    fld = synth.plummer_field(30_000, 105)
    v = synth.default_view(160, 160)
    pal = synth.default_palette()
    for st in (0, 2, 3, 4):
        log("case sort_type", st)
        pic = render(R, fld, v, [pal], S.params(a_eq_e=False, sort_type=st))
        opic, _ = O.host_rendering(fld, v, [pal], a_eq_e=False, sort_type=st, stable=True)
        assert_bitexact(pic, opic, "sort_type %d" % st)


def test_mixed_modes_combine_rule(R):
    """gas ordered + stars additive: pic = P_ordered(gas) + P_additive(stars)."""
    import splotch_amd as S
    fld = synth.gas_and_stars(40_000, 40_000, 106, 107)
    v = synth.default_view(256, 256)
    pal = synth.default_palette()
    prm = S.params(ptypes=2, a_eq_e=[0, 1], brightness=[0.5, 2.0])
    pic = render(R, fld, v, [pal, pal], prm)
    p = O.project_colorize(fld, v, [pal, pal], brightness=[0.5, 2.0])
    gas = p.take(np.arange(p.n)); gas.active = (p.active * (p.type == 0)).astype(np.uint8)
    stars = p.take(np.arange(p.n)); stars.active = (p.active * (p.type == 1)).astype(np.uint8)
    perm = O.sort_perm(gas, 1, stable=True)
    pg = O.render(gas.take(perm), v.xres, v.yres, False, 0.2, 50)
    ps = O.render(stars, v.xres, v.yres, True, 0.2, 50)
    assert_bitexact(pic, (pg + ps).astype(np.float32), "combine rule")


def test_empty_and_culled(R):
    import splotch_amd as S
    v = synth.default_view(64, 48)
    pal = synth.default_palette()
    fld = synth.plummer_field(1000, 108)
    fld.z[:] = -400.0  # all behind the camera (camera at z=-150 looking +z)
    pic = render(R, fld, v, [pal], S.params())
    assert not pic.any()
    fld0 = synth.plummer_field(0, 109)
    pic = render(R, fld0, v, [pal], S.params())
    assert not pic.any()
    one = synth.plummer_field(1, 110)
    one.x[:] = 50; one.y[:] = 50; one.z[:] = 50; one.hsml[:] = 3.0
    pic = render(R, one, v, [pal], S.params())
    opic, _ = O.host_rendering(one, v, [pal])
    assert_bitexact(pic, opic, "single particle")
    assert pic.any()


def test_deterministic(R):
    import splotch_amd as S
    fld = synth.plummer_field(150_000, 111)
    v = synth.default_view(320, 240)
    pal = synth.default_palette()
    a = render(R, fld, v, [pal], S.params(a_eq_e=False))
    b = render(R, fld, v, [pal], S.params(a_eq_e=False))
    assert_bitexact(a, b, "repeat")


@pytest.mark.parametrize("aeq", [False, True])
def test_dense_cluster_vs_oracle(R, aeq):
    """A dense cluster puts >10k particles into single 8x8 blocks, so the
    ordered pass issues them as per-channel work items."""
    import splotch_amd as S
    log("case dense cluster a_eq_e", aeq)
    fld = synth.plummer_field(60_000, 113, halo_frac=1.0, n_halos=1, hsml_mu=math.log(0.8))
    v = synth.default_view(160, 144)
    pal = synth.default_palette()
    opic, _ = O.host_rendering(fld, v, [pal], a_eq_e=aeq, stable=True)
    for G, var, split, heavy, hmax in ((8, 3, 0, 0, 0), (8, 1, 0, -1, 0), (8, 2, 0, -1, 0),
                                       (8, 3, 64, -1, 0), (8, 1, 1 << 30, -1, 0), (8, 4, 0, -1, 0),
                                       (8, 4, 64, 200, 0), (8, 4, 0, 1, 0), (8, 4, 0, 1, 1 << 30),
                                       (8, 4, 64, 5000, 3), (32, 0, 0, 0, 0)):
        pic = render(R, fld, v, [pal], S.params(a_eq_e=aeq, hip_region=G, hip_variant=var,
                                                hip_split=split, hip_heavy=heavy,
                                                hip_heavy_max=hmax))
        assert_bitexact(pic, opic, "dense cluster G=%d variant=%d split=%d heavy=%d max=%d"
                        % (G, var, split, heavy, hmax))


@pytest.mark.parametrize("var,heavy,hmax", [(4, -1, 0), (3, -1, 0), (4, 0, 0), (4, 1, 1 << 30)])
def test_zero_gray_absorption_nan_footprints(R, var, heavy, hmax):
    """gray_absorption = 0 makes q = e / (e + 0) = 0/0 for colour channels
    with e = 0 (the hot palette's green and blue below a third): the
    reference writes NaN into those channels of the particle's footprint
    pixels only.  The 8x8-block splat gives NaN in exactly the oracle's pixels
    and every other value bit for bit: its fast chunks apply a particle to the
    whole block with zero operands outside the footprint, so such a particle
    must neither take the fast table path nor leave its q where a masked slot
    reads it.  The producer/consumer kernel would spread the NaN over the
    rest of the block, so a pass with a non-finite q sends no list there, with
    the default long-list threshold (heavy 0) and with every list eligible
    (heavy 1)."""
    import splotch_amd as S
    log("case gray_absorption 0 variant", var, "heavy", heavy, hmax)
    fld = synth.plummer_field(20_000, 1620, halo_frac=1.0, n_halos=3, hsml_mu=math.log(0.6))
    v = synth.default_view(176, 144)
    pal = synth.default_palette()
    opic, _ = O.host_rendering(fld, v, [pal], a_eq_e=False, grayabsorb=0.0, stable=True)
    pic = render(R, fld, v, [pal], S.params(a_eq_e=False, gray_absorption=0.0, hip_heavy=heavy,
                                            hip_heavy_max=hmax, hip_variant=var))
    on = np.isnan(opic)
    log("NaN channels: oracle %d of %d" % (int(on.sum()), on.size))
    assert on.any() and not on.all()
    assert np.array_equal(np.isnan(pic), on), "NaN pixels differ"
    assert_bitexact(pic[~on], opic[~on], "finite pixels, gray_absorption 0")


@pytest.mark.parametrize("grp", [4, 2])
@pytest.mark.parametrize("T", [50, 37, 24, 16])
def test_region_binning_vs_oracle(R, T, grp):
    """Large footprints are binned through regions of grp x grp 8x8 blocks
    (split into the blocks' lists afterwards), point-like ones per block; both
    give the oracle's image bit for bit, with tile sizes whose 8x8 grid is not
    a multiple of the region (regions with blocks past the tile edge) or is."""
    import splotch_amd as S
    log("case region binning T", T, "group", grp)
    gas = synth.plummer_field(30_000, 1520 + T, ptype=0, halo_frac=1.0, n_halos=2,
                              hsml_mu=math.log(3.0))  # radii of ~10-30 px
    stars = synth.plummer_field(20_000, 1540 + T, ptype=1, hsml_mu=math.log(0.01))
    fld = synth.concat(gas, stars)
    v = synth.default_view(200, 168)
    pal = synth.default_palette()
    prm = S.params(ptypes=2, a_eq_e=[0, 1], brightness=[0.5, 2.0], tile_size=T,
                   hip_bin_group=grp)
    pic = render(R, fld, v, [pal, pal], prm)
    regions = R.binning()
    _, counts = R.stage_times()
    log("region entries", regions, "block entries", counts["ordered_entries"],
        counts["additive_entries"])
    assert regions[0] > 0 and 3 * regions[0] <= 2 * counts["ordered_entries"]
    assert regions[1] == 0
    p = O.project_colorize(fld, v, [pal, pal], brightness=[0.5, 2.0])
    gas = p.take(np.arange(p.n)); gas.active = (p.active * (p.type == 0)).astype(np.uint8)
    stars = p.take(np.arange(p.n)); stars.active = (p.active * (p.type == 1)).astype(np.uint8)
    pg = O.render(gas.take(O.sort_perm(gas, 1, stable=True)), v.xres, v.yres, False, 0.2, T)
    ps = O.render(stars, v.xres, v.yres, True, 0.2, T)
    assert_bitexact(pic, (pg + ps).astype(np.float32), "region binning T=%d" % T)


def test_profiling_and_trace_both_passes(R):
    """Stage timing and the work-item trace on a frame with both passes (the
    additive splat on the second stream): the image is unchanged, every
    stage of both passes has a time, and each pass's trace has one record per
    work item with end >= start."""
    import splotch_amd as S
    fld = synth.gas_and_stars(30_000, 30_000, 1590, 1591)
    v = synth.default_view(192, 160)
    pal = synth.default_palette()
    prm = S.params(ptypes=2, a_eq_e=[0, 1], brightness=[0.5, 2.0])
    want = render(R, fld, v, [pal, pal], prm)
    R.set_profiling(True)
    R.set_trace(True)
    try:
        got = render(R, fld, v, [pal, pal], prm)
        st, counts = R.stage_times()
        tr = {o: R.trace_items(o) for o in (True, False)}
    finally:
        R.set_trace(False)
        R.set_profiling(False)
    assert_bitexact(got, want, "profiled and traced frame")
    log("stages", st)
    for k in ("project", "ordered_bin", "ordered_splat", "additive_bin", "additive_splat"):
        assert st[k] > 0, k
    for o, t in tr.items():
        assert len(t) > 0, o
        assert np.all(t[:, 1] >= t[:, 0]), o


def test_chunked_region_binning(R):
    """Chunks of a pass binned through regions of 8x8 blocks (large footprints): the
    image equals the unchunked render bit for bit."""
    import splotch_amd as S
    gas = synth.plummer_field(20_000, 1560, ptype=0, halo_frac=1.0, n_halos=2,
                              hsml_mu=math.log(3.0))
    stars = synth.plummer_field(20_000, 1561, ptype=1, hsml_mu=math.log(2.0))
    fld = synth.concat(gas, stars)
    v = synth.default_view(176, 160)
    pal = synth.default_palette()
    for aeq in ([0, 1], [0, 0], [1, 1]):
        want = render(R, fld, v, [pal, pal], S.params(ptypes=2, a_eq_e=aeq,
                                                      brightness=[0.5, 2.0]))
        got = render(R, fld, v, [pal, pal], S.params(ptypes=2, a_eq_e=aeq, brightness=[0.5, 2.0],
                                                     hip_chunk_entries=60_000))
        regions = R.binning()
        _, counts = R.stage_times()
        log("chunked regions", aeq, regions, counts["ordered_entries"],
            counts["additive_entries"])
        assert max(counts["ordered_entries"], counts["additive_entries"]) > 60_000
        assert regions[0] > 0 or regions[1] > 0
        assert_bitexact(got, want, "chunked region binning %s" % (aeq,))


@pytest.mark.parametrize("aeq", [False, True])
def test_heavy_blocks_on_fixtures(R, golden, aeq):
    """Every non-empty block through the producer/consumer kernel."""
    import splotch_amd as S
    name = "additive_small" if aeq else "ordered_small"
    f = golden(name)
    kw = {} if aeq else dict(brightness=[float(f["brightness"][0])])
    pic = render(R, field_from(f), view_from(f), [palette(golden)],
                 S.params(a_eq_e=aeq, hip_heavy=1, hip_heavy_max=1 << 30, **kw))
    want = f["pic_T50"] if aeq else f["pic_stable_T50"]
    assert_bitexact(pic, want, "heavy kernel " + name)


def test_errors(R):
    import splotch_amd as S
    v = synth.default_view(64)
    fld = synth.plummer_field(100, 112)
    cam = S.camera_setup(v)
    pic = torch.empty(64 * 64 * 3, dtype=torch.float32, device="cuda")
    dev = S.to_device(fld)
    with pytest.raises(S.SplhipError):
        R.render(dev, cam, S.params(sort_type=9), pic)
    with pytest.raises(S.SplhipError):
        R.render(dev, cam, S.params(hip_region=24), pic)
    fld.c1 = None
    with pytest.raises(S.SplhipError):   # vector colour without c1/c2
        R.set_palette(0, *synth.default_palette())
        R.render(S.to_device(fld), cam, S.params(col_vector=[1]), pic)


@pytest.mark.parametrize("chunk,heavy,hmax", [(20_000, 0, 0), (3_000, 0, 0), (20_000, 1, 0),
                                              (20_000, 1, 1 << 30)])
def test_chunked_passes_bitexact(R, chunk, heavy, hmax):
    """A pass with more bin entries than hip_chunk_entries is rendered as
    consecutive ranges of its render order, each continuing the pixels'
    chains from the image: the image equals the unchunked render bit for bit
    (ordered, additive, and additive on top of ordered via a scratch image)."""
    import splotch_amd as S
    log("case chunked", chunk, heavy)
    gas = synth.plummer_field(30_000, 1500, ptype=0, hsml_mu=math.log(0.8))
    stars = synth.plummer_field(30_000, 1501, ptype=1, hsml_mu=math.log(0.4))
    fld = synth.concat(gas, stars)
    v = synth.default_view(192, 160)
    pal = synth.default_palette()
    for aeq in ([0, 1], [0, 0], [1, 1]):
        base = S.params(ptypes=2, a_eq_e=aeq, brightness=[0.5, 2.0], hip_heavy=heavy,
                        hip_heavy_max=hmax)
        want = render(R, fld, v, [pal, pal], base)
        prm = S.params(ptypes=2, a_eq_e=aeq, brightness=[0.5, 2.0], hip_heavy=heavy,
                       hip_heavy_max=hmax, hip_chunk_entries=chunk)
        got = render(R, fld, v, [pal, pal], prm)
        _, counts = R.stage_times()
        assert max(counts["ordered_entries"], counts["additive_entries"]) > chunk
        assert_bitexact(got, want, "chunked %s" % (aeq,))


def test_type_in_both_passes(R):
    """A type with a_eq_e 2 takes part in both passes (the large-share frame:
    additive over all particles plus the ordered gas pass): the image is the
    oracle's P_ordered (gas, stable depth order) + P_additive (gas and the
    rest), bit for bit, through the 8x8 splat and the producer/consumer
    kernel."""
    import splotch_amd as S
    log("case both passes")
    gas = synth.plummer_field(40_000, 1410, ptype=0, hsml_mu=math.log(0.6))
    rest = synth.plummer_field(40_000, 1411, ptype=1, hsml_mu=math.log(0.3))
    fld = synth.concat(gas, rest)
    v = synth.default_view(224, 192)
    pal = synth.default_palette()
    p = O.project_colorize(fld, v, [pal, pal], brightness=[0.5, 2.0])
    g = p.take(np.arange(p.n))
    g.active = (p.active * (p.type == 0)).astype(np.uint8)
    g = g.take(O.sort_perm(g, 1, stable=True))
    want = (O.render(g, v.xres, v.yres, False, 0.2, 50) +
            O.render(p, v.xres, v.yres, True, 0.2, 50)).astype(np.float32)
    for heavy, hmax in ((0, 0), (1, 1 << 30)):
        pic = render(R, fld, v, [pal, pal], S.params(ptypes=2, a_eq_e=[2, 1],
                                                      brightness=[0.5, 2.0], hip_heavy=heavy,
                                                      hip_heavy_max=hmax))
        assert_bitexact(pic, want, "type 0 in both passes, heavy %d" % heavy)
    _, counts = R.stage_times()
    assert counts["ordered_active"] == int(g.active.sum())
    assert counts["additive_active"] == int(p.active.sum())
