"""GPU parity at the classic codecs' large sizes: GF(2^8) at the 256-shard limit
and GF(2^16) with more than 256 data shards, more than 256 reconstruct targets or
more than 128 missing data shards (the decode solve then runs as C^-1 by one
workgroup -- in LDS, or over a device work matrix beyond 128 rows -- plus a grid
over the coefficients), against the oracle's encode and reconstruct."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import announce, rand_shards, to_dev, to_host

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
rsgpu = pytest.importorskip("rsgpu")

F = {"gf8": O.GF8, "gf16": O.GF16}


def _check(field, k, m, n, erasure_sets, seed):
    rng = np.random.default_rng(seed)
    rc, mat = O.build_matrix(F[field], O.VANDERMONDE, k, m)
    assert rc == 0
    data = rand_shards(rng, k, n)
    full = data + list(O.encode(F[field], mat, k, m, data))
    rs = rsgpu.ReedSolomon(k, m, field=field)
    dev = torch.from_numpy(np.stack(full)).cuda()
    par = dev[k:].clone().zero_()
    rs.encode_sep(list(dev[:k]), list(par))
    assert torch.equal(par, dev[k:])
    assert rs.verify(list(dev))
    bad = dev.clone()
    bad[k + m - 1, n - 1] ^= 1  # last byte of the last parity row
    assert not rs.verify(list(bad))
    for missing in erasure_sets:
        work = [None if i in missing else dev[i] for i in range(k + m)]
        rs.reconstruct(work)
        for i in sorted(missing):
            assert (to_host(work[i]) == full[i]).all(), (len(missing), i)


@pytest.mark.parametrize("k,m", [(128, 128), (200, 56), (255, 1), (1, 255)])
def test_gf8_at_the_shard_limit(k, m):
    announce(f"gf8 limit {k}+{m}")
    rng = np.random.default_rng(k)
    sets = [set(range(min(m, k))),  # as many data shards as parity can solve
            set(rng.choice(k + m, m, replace=False).tolist()),
            {k + m - 1}]
    _check("gf8", k, m, 1024 + 7, sets, k * 7 + m)


@pytest.mark.parametrize("k,m,n_missing_data,n_missing_parity",
                         [(300, 40, 40, 0), (300, 40, 17, 23), (300, 40, 0, 40),
                          (100, 300, 30, 270), (100, 1100, 10, 1050)])
def test_gf16_beyond_256(k, m, n_missing_data, n_missing_parity):
    announce(f"gf16 large {k}+{m} missing {n_missing_data}d {n_missing_parity}p")
    rng = np.random.default_rng(k + m + n_missing_data)
    missing = set(rng.choice(k, n_missing_data, replace=False).tolist()) | \
        set((k + rng.choice(m, n_missing_parity, replace=False)).tolist())
    _check("gf16", k, m, 256 + 2, [missing], k * 3 + m)


@pytest.mark.parametrize("k,m,n_missing_data,n_missing_parity",
                         [(300, 200, 150, 0), (300, 200, 200, 0), (300, 200, 129, 60),
                          (200, 150, 140, 10), (250, 6, 6, 0)])
def test_gf16_more_missing_data_than_one_lds_solve(k, m, n_missing_data, n_missing_parity):
    """More than 128 missing data shards (the reference solves any pattern of up to
    m erasures through its full inverse, reconstruct.rs:1164-1173): C^-1 runs as a
    Gauss-Jordan over a device work matrix.  Codeword round trip, then the bytes of
    a non-codeword against the oracle."""
    announce(f"gf16 big solve {k}+{m} missing {n_missing_data}d {n_missing_parity}p")
    rng = np.random.default_rng(k + m + n_missing_data + n_missing_parity)
    missing = set(rng.choice(k, n_missing_data, replace=False).tolist()) | \
        set((k + rng.choice(m, n_missing_parity, replace=False)).tolist())
    _check("gf16", k, m, 128 + 2, [missing], k * 5 + m)
    n = 64
    rc, mat = O.build_matrix(O.GF16, O.VANDERMONDE, k, m)
    shards = rand_shards(rng, k + m, n)  # not a codeword
    present = [0 if i in missing else 1 for i in range(k + m)]
    rc, exp = O.reconstruct(O.GF16, mat, k, m, [x.copy() for x in shards], present)
    assert rc == 0
    rs = rsgpu.ReedSolomon(k, m, field="gf16")
    slots = [rsgpu.ShardSlot(to_dev(shards[i]) if i not in missing else
                             torch.full((n,), 0xA5, dtype=torch.uint8, device="cuda"),
                             present=i not in missing) for i in range(k + m)]
    rs.reconstruct(slots)
    torch.cuda.synchronize()
    for i in sorted(missing):
        assert (to_host(slots[i].data) == exp[i]).all(), i


def test_batch_wide_stripes_with_tails():
    """Batch encode of stripes wider than one 64-input byte job, at lengths with a
    len % 16 tail: the tail runs as chained byte jobs."""
    announce("batch wide stripes with tails")
    batch = rsgpu.Batch()
    cases = []
    for k, m, n in ((100, 4, 1000 + 5), (70, 8, 4096 + 9), (65, 3, 33)):
        rng = np.random.default_rng(k + n)
        rc, mat = O.build_matrix(O.GF8, O.VANDERMONDE, k, m)
        data = rand_shards(rng, k, n)
        exp = O.encode(O.GF8, mat, k, m, data)
        rs = rsgpu.ReedSolomon(k, m)
        sh = [to_dev(d) for d in data] + [torch.full((n,), 0x3C, dtype=torch.uint8, device="cuda")
                                          for _ in range(m)]
        batch.add_encode(rs, sh)
        cases.append((rs, sh, exp, k, m))
    batch.run()
    torch.cuda.synchronize()
    for rs, sh, exp, k, m in cases:
        for j in range(m):
            assert (to_host(sh[k + j]) == exp[j]).all(), (k, m, j)


def test_gf16_field_order_restriction():
    """correct_field_order_restriction (src/tests/galois_16.rs:91-102)."""
    announce("gf16 field order")
    order = 1 << 16
    for args in ((order, 1), (1, order)):
        with pytest.raises(rsgpu.RsError) as ei:
            rsgpu.ReedSolomon(*args, field="gf16")
        assert ei.value.name == "TooManyShards"
    rs = rsgpu.ReedSolomon(1, order - 1, field="gf16")
    n = 64
    data = torch.arange(n, dtype=torch.uint8, device="cuda")
    par = torch.zeros((order - 1, n), dtype=torch.uint8, device="cuda")
    rs.encode_sep([data], list(par[:: 1]))
    # one data shard: every parity row is that shard times its coefficient, and
    # the first Vandermonde parity row of a 1-data code is the identity
    rc, mat = O.build_matrix(O.GF16, O.VANDERMONDE, 1, 3)
    exp = O.encode(O.GF16, mat, 1, 3, [data.cpu().numpy()])
    for j in range(3):
        assert (to_host(par[j]) == exp[j]).all(), j
