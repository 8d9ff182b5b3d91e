"""Verify flags cleared and reported by one launch (src/core/verify.rs:14-237 semantics:
true iff every parity shard equals the recomputed parity).  The kernel clears a
caller's flag itself, so these tests start every flag dirty and check that each
call's flag ends exactly 0 for a codeword and nonzero for a corrupted stripe:
many calls over several streams, every kernel variant, parity split over two
row launches, lengths the byte kernel covers alone, and GF(2^16)."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import announce, to_dev

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
rsgpu = pytest.importorskip("rsgpu")


def _stripes(f, k, m, n, count, seed):
    _, mat = O.build_matrix(f, O.VANDERMONDE, k, m)
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        data = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k)]
        out.append([to_dev(x) for x in data + list(O.encode(f, mat, k, m, data))])
    return out


def _corrupt(stripes, k, m, n, rng, every=2):
    bad = set()
    for s in range(0, len(stripes), every):
        j = int(rng.integers(0, k + m))
        pos = int(rng.integers(0, n))
        stripes[s][j][pos] ^= 1 << int(rng.integers(0, 8))
        bad.add(s)
    return bad


def _check_async(rs, stripes, bad, streams, tag):
    flags = torch.full((len(stripes),), 7, dtype=torch.int32, device="cuda")  # dirty
    torch.cuda.synchronize()
    for s, sh in enumerate(stripes):
        st = streams[s % len(streams)]
        rs.verify_async(sh, flags[s:s + 1], stream=st)
    torch.cuda.synchronize()
    got = flags.cpu().numpy()
    for s in range(len(stripes)):
        assert (got[s] != 0) == (s in bad), (tag, s, int(got[s]))


def test_many_calls_on_several_streams_start_from_dirty_flags():
    announce("verify flags: 96 calls on 3 streams")
    k, m, n = 10, 4, 65536
    stripes = _stripes(O.GF8, k, m, n, 96, 1)
    bad = _corrupt(stripes, k, m, n, np.random.default_rng(2), every=3)
    rs = rsgpu.ReedSolomon(k, m)
    streams = [torch.cuda.Stream() for _ in range(3)]
    for rep in range(2):  # the flags' own history must not matter
        _check_async(rs, stripes, bad, streams, rep)
    # the synchronous form on the same stripes, alternating good and bad
    for s in range(8):
        assert rs.verify(stripes[s]) == (s not in bad), s


def test_every_kernel_variant_clears_and_reports():
    announce("verify flags: forced variants")
    k, m, n = 12, 4, 4096 * 3
    stripes = _stripes(O.GF8, k, m, n, 8, 3)
    bad = _corrupt(stripes, k, m, n, np.random.default_rng(4))
    rs = rsgpu.ReedSolomon(k, m)
    for v in [-1] + [v for v in range(rsgpu.variant_count()) if rs.variant_valid("verify", 0, v)]:
        rs.set_variant(v)
        _check_async(rs, stripes, bad, [torch.cuda.current_stream()], f"variant {v}")
    rs.set_variant(-1)


@pytest.mark.parametrize("field,k,m,n", [("gf8", 12, 10, 8192),     # parity rows in two launches
                                         ("gf8", 6, 3, 4096 + 3),    # vector kernel + byte tail
                                         ("gf8", 5, 2, 8),           # the byte kernel alone
                                         ("gf16", 6, 3, 4096 + 2),   # tower elements, 16-byte tile
                                         ("gf16", 4, 2, 6)])         # tower, byte kernel alone
def test_shapes_with_tails_and_row_chunks(field, k, m, n):
    announce(f"verify flags: {field} {k}+{m} n={n}")
    f = O.GF8 if field == "gf8" else O.GF16
    stripes = _stripes(f, k, m, n, 6, k * 100 + n)
    rng = np.random.default_rng(n)
    bad = set()
    for s in (1, 4):
        j = k + m - 1 if s == 1 else int(rng.integers(0, k))  # a parity row of the last launch
        stripes[s][j][n - 1] ^= 0x81
        bad.add(s)
    rs = rsgpu.ReedSolomon(k, m, field=field)
    _check_async(rs, stripes, bad, [torch.cuda.Stream(), torch.cuda.Stream()], field)
