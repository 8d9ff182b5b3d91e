"""Incremental parity on the GPU path (src/core/encode.rs:892-1132,
src/core/shard_by_shard.rs): encode_single / encode_single_sep one shard at a
time, update with changed shards, ShardByShard's call protocol, the reference's
validation errors -- bytes against the oracle's full encode."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import announce, rand_shards, to_dev, to_host

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
rsgpu = pytest.importorskip("rsgpu")

FIELDS = {"gf8": O.GF8, "gf16": O.GF16}


def expected(field, k, m, data):
    rc, mat = O.build_matrix(FIELDS[field], O.VANDERMONDE, k, m)
    assert rc == 0
    return O.encode(FIELDS[field], mat, k, m, data)


def zeros(n):
    return torch.zeros(n, dtype=torch.uint8, device="cuda")


@pytest.mark.parametrize("field,k,m,n", [("gf8", 10, 4, 65536 + 7), ("gf8", 3, 2, 4096),
                                         ("gf8", 20, 8, 1 << 20), ("gf16", 6, 3, 2048 + 2)])
def test_encode_single_in_order_equals_encode(field, k, m, n):
    announce(f"encode_single {field} {k}+{m} n={n}")
    rng = np.random.default_rng(k * m + n)
    data = rand_shards(rng, k, n)
    exp = expected(field, k, m, data)
    rs = rsgpu.ReedSolomon(k, m, field=field)
    sh = [to_dev(d) for d in data] + [to_dev(rng.integers(0, 256, n, dtype=np.uint8)) for _ in range(m)]
    for i in range(k):  # garbage parity is overwritten by shard 0, then accumulated
        rs.encode_single(i, sh)
    for j in range(m):
        assert (to_host(sh[k + j]) == exp[j]).all(), j
    par = [zeros(n) for _ in range(m)]
    for i in range(k):
        rs.encode_single_sep(i, to_dev(data[i]), par)
    for j in range(m):
        assert (to_host(par[j]) == exp[j]).all(), j


@pytest.mark.parametrize("field,k,m,n,changed", [("gf8", 10, 4, 65536, (0, 3, 9)), ("gf8", 6, 3, 1000, (2,)),
                                                 ("gf8", 40, 4, 8192, tuple(range(40))),
                                                 ("gf16", 5, 4, 4096, (1, 4))])
def test_update_equals_encode_of_new_data(field, k, m, n, changed):
    announce(f"update {field} {k}+{m} n={n} changed={len(changed)}")
    rng = np.random.default_rng(n + len(changed))
    old = rand_shards(rng, k, n)
    new = [rng.integers(0, 256, n, dtype=np.uint8) if i in changed else old[i] for i in range(k)]
    exp_old, exp_new = expected(field, k, m, old), expected(field, k, m, new)
    rs = rsgpu.ReedSolomon(k, m, field=field)
    par = [to_dev(p) for p in exp_old]
    rs.update([to_dev(d) for d in old], [to_dev(new[i]) if i in changed else None for i in range(k)], par)
    for j in range(m):
        assert (to_host(par[j]) == exp_new[j]).all(), j
    # nothing changed: parity untouched
    rs.update([to_dev(d) for d in new], [None] * k, par)
    for j in range(m):
        assert (to_host(par[j]) == exp_new[j]).all(), j


def test_update_with_fast_one_parity():
    announce("update fast one parity")
    k, n = 7, 3000
    rng = np.random.default_rng(77)
    old = rand_shards(rng, k, n)
    new = list(old)
    new[2] = rng.integers(0, 256, n, dtype=np.uint8)
    rs = rsgpu.ReedSolomon(k, 1, fast_one_parity=True)
    par = [zeros(n)]
    rs.encode([to_dev(d) for d in old] + par)
    rs.update([to_dev(d) for d in old], [None, None, to_dev(new[2])] + [None] * 4, par)
    want = np.bitwise_xor.reduce(np.stack(new), axis=0)
    assert (to_host(par[0]) == want).all()


def test_incremental_validation_matches_reference():
    announce("incremental validation")
    k, m, n = 4, 2, 256
    rs = rsgpu.ReedSolomon(k, m)
    sh = [zeros(n) for _ in range(k + m)]

    def err(fn):
        with pytest.raises(rsgpu.RsError) as ei:
            fn()
        return ei.value.name

    assert err(lambda: rs.encode_single(k, sh)) == "InvalidIndex"
    assert err(lambda: rs.encode_single(0, sh[:-1])) == "TooFewShards"
    assert err(lambda: rs.encode_single(0, sh[:-1] + [zeros(n - 1)])) == "IncorrectShardSize"
    assert err(lambda: rs.encode_single_sep(0, zeros(n), sh[:1])) == "TooFewParityShards"
    assert err(lambda: rs.encode_single_sep(0, zeros(n + 1), sh[:m])) == "IncorrectShardSize"
    assert err(lambda: rs.update(sh[:k - 1], [None] * k, sh[:m])) == "TooFewDataShards"
    assert err(lambda: rs.update(sh[:k], [None] * k, sh[:m + 1])) == "TooManyParityShards"
    assert err(lambda: rs.update(sh[:k], [zeros(n + 1)] + [None] * (k - 1), sh[:m])) == "IncorrectShardSize"
    leo = rsgpu.ReedSolomon(k, m, codec_family="leopard_gf8")
    assert err(lambda: leo.encode_single(0, sh)) == "UnsupportedCodecFamily"
    assert err(lambda: leo.update(sh[:k], [None] * k, sh[:m])) == "UnsupportedCodecFamily"


def test_shard_by_shard_protocol():
    """ShardByShard (shard_by_shard.rs:59-180) over the GPU encode_single."""
    announce("shard by shard")
    k, m, n = 5, 3, 4096
    rng = np.random.default_rng(5)
    data = rand_shards(rng, k, n)
    exp = expected("gf8", k, m, data)
    rs = rsgpu.ReedSolomon(k, m)
    sbs = rsgpu.ShardByShard(rs)
    sh = [to_dev(d) for d in data] + [zeros(n) for _ in range(m)]
    for i in range(k):
        assert not sbs.parity_ready() and sbs.cur_input_index() == i
        sbs.encode(sh)
    assert sbs.parity_ready()
    for j in range(m):
        assert (to_host(sh[k + j]) == exp[j]).all()
    with pytest.raises(rsgpu.SBSError) as ei:
        sbs.encode(sh)
    assert ei.value.kind == "TooManyCalls"
    sbs.reset()
    par = [zeros(n) for _ in range(m)]
    sbs.encode_sep([to_dev(d) for d in data], par)
    with pytest.raises(rsgpu.SBSError) as ei:
        sbs.reset()
    assert ei.value.kind == "LeftoverShards"
    with pytest.raises(rsgpu.SBSError) as ei:
        sbs.encode(sh[:-1])
    assert ei.value.kind == "RSError" and ei.value.error.name == "TooFewShards"
    sbs.reset_force()
    assert sbs.cur_input_index() == 0
