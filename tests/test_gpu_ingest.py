"""GPU parity of the block-file SoA ingest (splhip_read_bin_block), through
the C-ABI.

Bar: bit for bit against the genuine bin_reader_block's fixtures
(tests/golden/ingest.npz: both byte orders, missing columns, smooth_factor,
NaN payloads) and against the oracle's numpy restatement on a 3M-particle
file that spans several staging chunks per column.
"""
import sys

import numpy as np
import pytest

import pyoracle as O
from helpers import assert_bitexact, bb_case

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

NAMES = (("x", "x"), ("y", "y"), ("z", "z"), ("r", "hsml"), ("I", "inten"), ("er", "col"),
         ("eg", "c1"), ("eb", "c2"), ("type", "type"))


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


def _check(R, kv, want, what):
    d = R.read_bin_block(O.kv_text(kv))
    torch.cuda.synchronize()
    for q, dk in NAMES:
        assert_bitexact(d[dk].cpu().numpy(), want[q], "%s %s" % (what, q))


def test_ingest_fixtures(R, golden, tmp_path):
    f = golden("ingest")
    for k in range(int(f["ncases"][0])):
        kv, want = bb_case(f, k, tmp_path / ("b%d.bin" % k))
        log("case ingest", k, kv)
        _check(R, kv, want, "fixture case %d" % k)


def test_ingest_large_vs_oracle(R, tmp_path):
    """3M particles, 8 columns, big-endian, in 16 MiB staging chunks."""
    rng = np.random.default_rng(17)
    n = 3_000_000
    vals = rng.lognormal(0, 1, (8, n)).astype(np.float32)
    path = tmp_path / "large.bin"
    vals.astype(">f4").tofile(path)
    kv = dict(infile=str(path), num_blocks=8, x=1, y=2, z=3, r=4, I=5, C1=6, C2=7, C3=8,
              smooth_factor=0.7)
    _check(R, kv, O.read_bin_block(kv), "3M")


def test_ingest_errors(R, golden, tmp_path):
    import splotch_amd as S
    f = golden("ingest")
    kv, _ = bb_case(f, 0, tmp_path / "b.bin")
    one = torch.zeros(4, device="cuda")
    import ctypes as C
    from splotch_amd import capi
    rc = capi.lib().splhip_read_bin_block(R._c, O.kv_text(kv).encode(), C.c_int64(4),
                                          *([C.c_void_p(one.data_ptr())] * 8), None, None)
    assert rc != 0   # n differs from the file's count
    with pytest.raises(S.SplhipError):
        R.read_bin_block(O.kv_text(dict(kv, infile=str(tmp_path / "missing.bin"))))
