"""GPU parity of the post-processing and 8-bit output (splhip_postprocess),
through the C-ABI.

Bar: the genuine reference's fixtures (tests/golden/postprocess.npz, written
by the genuine -expm1 / colour_adjust / add_colorbar / LS_Image code) and the
oracle, bit for bit, for the float image and the TGA bytes.  The one
exception is colour_adjust with gamma != 1: it calls pow(double, double),
whose last bit may differ between ROCm's device library and glibc; after the
rounding to float the images agree within one float ulp (the test states the
count), and the bytes agree wherever the floats do.
"""
import sys

import numpy as np
import pytest

import pyoracle as O
from helpers import assert_bitexact, post_args, pp_case
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


def _run(R, pic, q, layout, inplace=False):
    xres, yres = pic.shape[:2]
    d = torch.from_numpy(np.ascontiguousarray(pic, np.float32)).reshape(-1).cuda()
    out = d if inplace else torch.full_like(d, float("nan"))
    rgb8 = torch.zeros(d.numel(), dtype=torch.uint8, device="cuda")
    R.postprocess(d, xres, yres, q, out=out, rgb8=rgb8, layout=layout)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(pic.shape), rgb8.cpu().numpy()


def _ulps(a, b):
    ia = np.ascontiguousarray(a, np.float32).view(np.int32).astype(np.int64)
    ib = np.ascontiguousarray(b, np.float32).view(np.int32).astype(np.int64)
    both_nan = np.isnan(a) & np.isnan(b)
    return np.where(both_nan, 0, np.abs(ia - ib))


def _check(got, gbytes, want, tga, layout, what, exact):
    xres, yres = want.shape[:2]
    body = tga[18:]
    want8 = body if layout == 1 else body.reshape(yres, xres, 3)[:, :, ::-1].transpose(
        1, 0, 2).reshape(-1)
    if exact:
        assert_bitexact(got, want, what + " float")
        assert_bitexact(gbytes, want8, what + " bytes")
        return
    u = _ulps(got, want)
    log("%s: %d of %d floats differ, at most %d ulp" % (what, int((u > 0).sum()), u.size,
                                                        int(u.max())))
    assert u.max() <= 1
    same = (u == 0).reshape(-1)
    if layout == 1:
        same = same.reshape(xres, yres, 3)[:, :, ::-1].transpose(1, 0, 2).reshape(-1)
    assert np.array_equal(gbytes[same], want8[same]), what + " bytes"


@pytest.mark.parametrize("layout", [1, 0])
def test_postprocess_fixtures(R, golden, layout):
    import splotch_amd as S
    p = golden("palette")
    f = golden("postprocess")
    for k in range(int(f["ncases"][0])):
        pic, kv, pals, want, tga = pp_case(f, (p["x"], p["rgb"]), k)
        log("case postprocess", k, kv, "layout", layout)
        for t, (x, rgb) in enumerate(pals):
            R.set_palette(t, x, rgb)
        q = S.post_from_par(O.kv_text(kv))
        got, gbytes = _run(R, pic, q, layout, inplace=(k % 2 == 1))
        _check(got, gbytes, want, tga, layout, "fixture case %d" % k, q.gamma == 1.0)


def test_postprocess_full_frame_vs_oracle(R):
    """A 2048x1536 frame: additive -expm1 with two colorbars, then an ordered
    frame with brightness / contrast (gamma 1)."""
    import splotch_amd as S
    rng = np.random.default_rng(77)
    pal = synth.default_palette()
    pal3 = (np.array([0.0, 0.3, 1.0]),
            np.array([[0.1, 0.2, 0.3], [0.9, 0.5, 0.1], [1.0, 1.0, 0.8]], np.float32))
    R.set_palette(0, *pal)
    R.set_palette(1, *pal3)
    pic = (-rng.exponential(0.8, (2048, 1536, 3))).astype(np.float32)
    kw = dict(a_eq_e=True, colorbar=True, col_vector=[0, 0])
    want, tga = O.postprocess(pic, [pal, pal3], **kw)
    got, gbytes = _run(R, pic, S.post(ptypes=2, **kw), 1)
    _check(got, gbytes, want, tga, 1, "additive frame", True)
    pic = rng.uniform(-0.2, 1.3, (2048, 1536, 3)).astype(np.float32)
    kw = dict(a_eq_e=False, brightness=-0.03, contrast=1.2)
    want, tga = O.postprocess(pic, [pal], **kw)
    got, gbytes = _run(R, pic, S.post(**kw), 0, inplace=True)
    _check(got, gbytes, want, tga, 0, "ordered frame", True)


def test_postprocess_errors(R):
    import splotch_amd as S
    d = torch.zeros(12, device="cuda")
    with pytest.raises(S.SplhipError):
        R.postprocess(d, 2, 2, S.post(), out=None, rgb8=None)
    with pytest.raises(S.SplhipError):
        R.postprocess(d, 2, 2, S.post(), out=d, layout=2)
    with pytest.raises(S.SplhipError):   # a colorbar for a type without a palette
        R.postprocess(d, 2, 2, S.post(colorbar=True, ptypes=8), out=d)
