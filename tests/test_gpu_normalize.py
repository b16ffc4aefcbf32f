"""GPU parity of particle_normalize (splhip_normalize), through the C-ABI.

Bar: against the genuine sceneMaker::particle_normalize fixtures
(tests/golden/normalize.npz) and the oracle.  Without log or asinh the
result is bit for bit.  log10 and asinh are evaluated in double and rounded
to float on the GPU, while the reference calls glibc's log10f, which is not
correctly rounded (one ulp off for about 1.4% of arguments), and its own
double asinh: logged values therefore agree within one float ulp, and the
normalized values, which also carry the logged extremes, within 1e-6
(absolute; they lie in [0, 1]).  -38 markers and NaN positions agree exactly.
"""
import sys

import numpy as np
import pytest

import pyoracle as O
from helpers import assert_bitexact, nm_case

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FIELDS = (("I", "inten"), ("er", "col"), ("eg", "c1"), ("eb", "c2"))


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


def _run(R, fields, kv, nt):
    import splotch_amd as S
    dev = {dk: torch.from_numpy(np.ascontiguousarray(fields[fk], np.float32)).cuda()
           for fk, dk in FIELDS}
    dev["hsml"] = torch.from_numpy(np.ascontiguousarray(fields["r"], np.float32)).cuda()
    dev["type"] = torch.from_numpy(np.ascontiguousarray(fields["type"], np.uint8)).cuda()
    q = S.norm_from_par(O.kv_text(dict(kv, ptypes=nt)))
    ranges = R.normalize(dev, q)
    torch.cuda.synchronize()
    return {fk: dev[dk].cpu().numpy() for fk, dk in FIELDS}, ranges


def _compare(got, want, exact, what):
    for q in ("I", "er", "eg", "eb"):
        g, w = got[q], want[q]
        assert np.array_equal(np.isnan(g), np.isnan(w)), what + " NaN positions " + q
        assert np.array_equal(g == -38, w == -38), what + " -38 markers " + q
        if exact:
            assert_bitexact(g, w, what + " " + q)
        else:
            m = ~np.isnan(w)
            d = np.abs(g[m].astype(np.float64) - w[m])
            log("%s %s: %d of %d differ, max |diff| %.3g" % (what, q, int((d > 0).sum()),
                                                          d.size, d.max() if d.size else 0))
            assert d.max() <= 1e-6, (what, q)


def test_normalize_fixtures(R, golden):
    f = golden("normalize")
    nt = int(f["ntypes"][0])
    for k in range(int(f["ncases"][0])):
        kv, fields, want, ranges = nm_case(f, k)
        log("case normalize", k, kv)
        got, rr = _run(R, fields, kv, nt)
        exact = not any(k2.startswith(("intensity_log", "color_log", "color_asinh"))
                        for k2 in kv)
        _compare(got, want, exact, "fixture case %d" % k)
        for t in range(nt):
            g = [rr[t]["range_min"], rr[t]["range_max"], rr[t]["clamp_min"], rr[t]["clamp_max"]]
            assert np.allclose(g, ranges[t], rtol=1e-6, atol=0), (k, t, g, ranges[t])


def test_normalize_large_vs_oracle(R):
    """4M particles of two types with logs, asinh and a vector colour."""
    rng = np.random.default_rng(91)
    n = 4_000_000
    fields = dict(I=rng.lognormal(0, 2, n).astype(np.float32),
                  r=rng.lognormal(-1, 0.5, n).astype(np.float32),
                  er=rng.lognormal(0, 1.5, n).astype(np.float32),
                  eg=rng.lognormal(0, 1, n).astype(np.float32),
                  eb=rng.lognormal(0, 1, n).astype(np.float32),
                  type=rng.integers(0, 2, n).astype(np.uint8))
    fields["I"][::1000] = -1.0
    for kv, exact in ((dict(), True),
                      (dict(intensity_log0="true", color_log0="true", color_asinh1="true",
                            color_is_vector1="true", color_max0=1.5), False)):
        want, out, size = O.normalize(fields, kv, 2)
        got, rr = _run(R, fields, kv, 2)
        _compare(got, want, exact, "4M %s" % kv)
        for t in range(2):
            assert rr[t]["size_min"] == size[t, 0] and rr[t]["size_max"] == size[t, 1]
            if exact:
                assert (rr[t]["range_min"], rr[t]["range_max"], rr[t]["intensity_min"],
                        rr[t]["intensity_max"]) == (out[t, 0], out[t, 1], out[t, 4], out[t, 5])


def test_normalize_empty_and_errors(R):
    import splotch_amd as S
    e = torch.zeros(0, device="cuda")
    rr = R.normalize(dict(inten=e, col=e, hsml=e), S.norm_from_par("ptypes=2"))
    assert rr[1]["range_min"] == np.float32(1e37) and rr[1]["range_max"] == np.float32(-1e37)
    one = torch.ones(4, device="cuda")
    with pytest.raises(S.SplhipError):   # vector colour without c1 / c2
        R.normalize(dict(inten=one, col=one, hsml=one), S.norm_from_par("color_is_vector0=T"))
