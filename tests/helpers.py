"""Shared helpers for the parity tests (fixture decoding, comparisons)."""
import numpy as np

from splotch_amd import synth


def view_from(f, prefix=""):
    res = f[prefix + "res"]
    return synth.View(tuple(f[prefix + "campos"]), tuple(f[prefix + "lookat"]),
                      tuple(f[prefix + "sky"]), int(res[0]), int(res[1]),
                      float(f[prefix + "fov"][0]), tuple(f[prefix + "centerpos"]))


def field_from(f):
    fl = synth.Field(f["x"], f["y"], f["z"], f["hsml"], f["inten"], f["col"],
                     f["type"] if "type" in f else None)
    if "c1" in f:
        fl.c1 = f["c1"]
        fl.c2 = f["c2"]
    return fl


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def assert_bitexact(a, b, what=""):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    ba, bb = bits(a), bits(b)
    bad = np.nonzero(ba.reshape(-1) != bb.reshape(-1))[0]
    if bad.size:
        i = bad[0]
        raise AssertionError("%s: %d of %d differ; first at %d: %r vs %r"
                             % (what, bad.size, ba.size, i, a.reshape(-1)[i], b.reshape(-1)[i]))


def max_rel_err(a, b, floor_frac=1e-6):
    """Max per-channel relative error |a-b| / max(|b|, floor_frac*max|b|)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.maximum(np.abs(b), floor_frac * max(np.abs(b).max(), 1e-30))
    return float((np.abs(a - b) / den).max())


def parse_kv(text):
    """'key=value' lines (a .par excerpt) -> dict of strings."""
    return dict(line.split("=", 1) for line in str(text).splitlines() if "=" in line)


def post_args(kv):
    """The post-processing switches of a .par excerpt, as the oracle takes them
    (defaults of splotch.cc:186,307 and splotchutils.cc:232-234)."""
    truth = ("t", "y", "true", ".true.")  # stringToData<bool>, string_utils.cc:131-140
    ptypes = int(kv.get("ptypes", 1))
    return dict(a_eq_e=kv.get("a_eq_e", "true").lower() in truth,
                gamma=float(kv.get("pic_gamma", 1.0)),
                brightness=float(kv.get("pic_brightness", 0.0)),
                contrast=float(kv.get("pic_contrast", 1.0)),
                colorbar=kv.get("colorbar", "false").lower() in truth,
                col_vector=[int(kv.get("color_is_vector%d" % t, "false").lower() in truth)
                            for t in range(ptypes)])


def pp_case(f, pal, k):
    """Case k of the postprocess fixture: (raw image, kv dict, palettes,
    expected float image, expected TGA bytes)."""
    kv = parse_kv(f["pp%d_kv" % k])
    pal3 = (f["pal3_x"], f["pal3_rgb"])
    pals = [pal, pal3, pal3][:int(kv.get("ptypes", 1))]
    return f["pp%d_in" % k], kv, pals, f["pp%d_out" % k], f["pp%d_tga" % k]


def nm_case(f, k):
    """Case k of the normalize fixture: (kv dict, input fields, expected
    output fields, range parameters float64[ntypes, 4])."""
    kv = parse_kv(f["nm%d_kv" % k])
    fields = {q: f["nm%d_%s" % (k, q)] for q in ("I", "r", "er", "eg", "eb", "type")}
    want = {q: f["nm%d_out_%s" % (k, q)] for q in ("I", "er", "eg", "eb")}
    return kv, fields, want, f["nm%d_ranges" % k]


def bb_case(f, k, path):
    """Case k of the ingest fixture, its file written to path: (kv dict with
    infile = path, expected fields)."""
    f["bb%d_file" % k].tofile(path)
    kv = dict(parse_kv(f["bb%d_kv" % k]), infile=str(path))
    want = {q: f["bb%d_%s" % (k, q)] for q in ("x", "y", "z", "r", "I", "er", "eg", "eb", "type")}
    return kv, want
