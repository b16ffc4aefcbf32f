"""splhip_normalize_noalloc (the entry point that allocates nothing), through
the C-ABI, against splhip_normalize.

Bar: bit-exact fields and equal per-type ranges, on two calls in a row, the
second with a smaller input (the shared scratch keeps the first call's
extremes; the call must clear them), on the genuine fixture cases (logs,
asinh, vector colours, overrides) and on a 3M-particle three-type field.
"""
import sys

import numpy as np
import pytest

import pyoracle as O
from helpers import nm_case

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


def _dev(fields):
    dev = {dk: torch.from_numpy(np.ascontiguousarray(fields[fk], np.float32)).cuda()
           for fk, dk in FIELDS}
    dev["hsml"] = torch.from_numpy(np.ascontiguousarray(fields["r"], np.float32)).cuda()
    dev["type"] = torch.from_numpy(np.ascontiguousarray(fields["type"], np.uint8)).cuda()
    return dev


def _both(R, fields, q, what):
    """Normalize copies of the fields with both entry points; compare."""
    a, b = _dev(fields), _dev(fields)
    ra = R.normalize(a, q)
    rb = R.normalize(b, q, noalloc=True)
    torch.cuda.synchronize()
    for _, dk in FIELDS:
        x, y = a[dk].cpu().numpy(), b[dk].cpu().numpy()
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), (what, dk)
    assert ra == rb, (what, ra, rb)


def _big(n, seed):
    rng = np.random.default_rng(seed)
    return {"I": rng.lognormal(0, 2, n).astype(np.float32),
            "er": rng.lognormal(0, 1, n).astype(np.float32),
            "eg": rng.uniform(0, 5, n).astype(np.float32),
            "eb": rng.uniform(-1, 1, n).astype(np.float32),
            "r": rng.lognormal(-1, 0.5, n).astype(np.float32),
            "type": rng.integers(0, 3, n).astype(np.uint8)}


def test_noalloc_equals_normalize_big_then_small(R):
    import splotch_amd as S
    q = S.norm_from_par("ptypes=3\nintensity_log0=T\ncolor_log1=T\ncolor_is_vector2=T\n")
    log("case noalloc: 3M particles, then 1000 (a smaller input after a larger one)")
    _both(R, _big(3_000_000, 41), q, "3M")
    small = _big(1000, 42)
    for k in ("I", "er"):
        small[k] = small[k] * np.float32(1e-3)  # extremes inside the first call's range
    _both(R, small, q, "1000 after 3M")


def test_noalloc_fixtures_twice(R, golden):
    import splotch_amd as S
    f = golden("normalize")
    nt = int(f["ntypes"][0])
    cases = list(range(int(f["ncases"][0])))
    log("case noalloc fixtures", len(cases))
    for rep in range(2):
        for k in cases:
            kv, fields, _, _ = nm_case(f, k)
            q = S.norm_from_par(O.kv_text(dict(kv, ptypes=nt)))
            _both(R, fields, q, "fixture %d pass %d" % (k, rep))
            part = {key: v[: max(1, len(v) // 3)] for key, v in fields.items()}
            _both(R, part, q, "fixture %d third pass %d" % (k, rep))


def test_noalloc_errors(R):
    import ctypes as C
    import splotch_amd as S
    from splotch_amd import capi
    q = S.norm_from_par("ptypes=2\n")
    q.ptypes = 9
    one = torch.zeros(4, device="cuda")
    rc = capi.lib().splhip_normalize_noalloc(R._c, C.c_int64(4),
                                             *([C.c_void_p(one.data_ptr())] * 5), None,
                                             C.byref(q), None, None)
    assert rc != 0
