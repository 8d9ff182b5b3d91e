"""GPU parity of the classic GF(2^8) / GF(2^16) data path against the oracle.

Every byte produced on the device through the C ABI is compared with the CPU
oracle (bit-exact) on seeded inputs; the full-size cases use size-independent
properties (encode -> erase -> reconstruct round trips, verify verdicts).
"""
import itertools

import numpy as np
import pytest

import oracle_lib as O
from gpu_util import announce, rand_shards, to_dev, to_host

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
rsgpu = pytest.importorskip("rsgpu")

ORA_FIELD = {"gf8": O.GF8, "gf16": O.GF16}
ORA_MODE = {"vandermonde": O.VANDERMONDE, "cauchy": O.CAUCHY, "jerasure_like": O.JERASURE}


def _encode_dev(rs, data):
    shards = [to_dev(d) for d in data] + [torch.zeros(len(data[0]), dtype=torch.uint8, device="cuda")
                                         for _ in range(rs.parity_shard_count())]
    rs.encode(shards)
    torch.cuda.synchronize()
    return shards


def test_golden_4x2_on_device():
    announce("golden 4+2")
    rs = rsgpu.ReedSolomon(4, 2)
    data = [np.arange(16 * i, 16 * i + 16, dtype=np.uint8) for i in range(4)]
    shards = _encode_dev(rs, data)
    out = [to_host(s) for s in shards]
    assert out[4].tolist() == list(range(64, 80))
    assert out[5].tolist() == list(range(80, 96))
    assert O.checksum_shards(out) == 0x9118ae5245b6279b
    assert rs.verify(shards)
    shards[5][3] ^= 0xFF
    assert not rs.verify(shards)


@pytest.mark.parametrize("k,m", [(4, 2), (6, 3), (8, 3), (8, 4), (10, 4), (12, 4), (16, 4), (20, 4),
                                 (3, 9), (70, 5), (17, 13)])
@pytest.mark.parametrize("length", [1, 15, 16, 17, 1000, 4096 + 48])
def test_encode_matches_oracle(k, m, length):
    announce(f"encode {k}+{m} len={length}")
    rng = np.random.default_rng(k * 1000 + m * 10 + length)
    rs = rsgpu.ReedSolomon(k, m)
    mat = rs.matrix()
    rc, omat = O.build_matrix(O.GF8, O.VANDERMONDE, k, m)
    assert (mat == omat).all()
    data = rand_shards(rng, k, length)
    shards = _encode_dev(rs, data)
    expect = O.encode(O.GF8, omat, k, m, data)
    for j in range(m):
        assert (to_host(shards[k + j]) == expect[j]).all(), j
    assert rs.verify(shards)


@pytest.mark.parametrize("field", ["gf8", "gf16"])
@pytest.mark.parametrize("mode", ["vandermonde", "cauchy", "jerasure_like"])
def test_matrix_modes_and_fields(field, mode):
    announce(f"modes {field} {mode}")
    k, m, n = 7, 4, 2 * 1000 + 2
    rs = rsgpu.ReedSolomon(k, m, field=field, matrix_mode=mode)
    rc, omat = O.build_matrix(ORA_FIELD[field], ORA_MODE[mode], k, m)
    assert rc == 0
    assert (rs.matrix() == omat).all()
    rng = np.random.default_rng(11)
    data = rand_shards(rng, k, n)
    shards = _encode_dev(rs, data)
    expect = O.encode(ORA_FIELD[field], omat, k, m, data)
    for j in range(m):
        assert (to_host(shards[k + j]) == expect[j]).all()
    assert rs.verify(shards)
    full = [to_host(s) for s in shards]
    for missing in [(0,), (1, 8), (0, 2, 9), (3, 4, 5, 6)]:
        work = [None if i in missing else to_dev(full[i]) for i in range(k + m)]
        rs.reconstruct(work)
        for i in range(k + m):
            assert (to_host(work[i]) == full[i]).all(), (missing, i)


def test_every_erasure_pattern_6x3():
    announce("all patterns 6+3")
    k, m, n = 6, 3, 4096 + 16 + 3
    rs = rsgpu.ReedSolomon(k, m)
    rng = np.random.default_rng(2)
    data = rand_shards(rng, k, n)
    full = [to_host(s) for s in _encode_dev(rs, data)]
    for e in range(1, m + 1):
        for missing in itertools.combinations(range(k + m), e):
            work = [None if i in missing else to_dev(full[i]) for i in range(k + m)]
            rs.reconstruct(work)
            for i in range(k + m):
                assert (to_host(work[i]) == full[i]).all(), (missing, i)
    st = rs.cache_stats()
    assert st["misses"] > 0


def test_reconstruct_matches_oracle_on_inconsistent_input():
    """Survivor choice (first k present) matters when the shards are not a codeword:
    the device must pick the same survivors as the reference."""
    announce("inconsistent input")
    k, m, n = 5, 3, 333
    _, omat = O.build_matrix(O.GF8, O.VANDERMONDE, k, m)
    rs = rsgpu.ReedSolomon(k, m)
    rng = np.random.default_rng(9)
    shards = rand_shards(rng, k + m, n)  # random parity: not a codeword
    for missing in [(1,), (0, 6), (2, 3, 4)]:
        present = [0 if i in missing else 1 for i in range(k + m)]
        rc, expect = O.reconstruct(O.GF8, omat, k, m, [s.copy() for s in shards], present)
        assert rc == 0
        work = [None if i in missing else to_dev(shards[i]) for i in range(k + m)]
        rs.reconstruct(work)
        for i in range(k + m):
            assert (to_host(work[i]) == expect[i]).all(), (missing, i)


def test_reconstruct_data_and_some():
    announce("reconstruct_data / reconstruct_some")
    k, m, n = 10, 4, 1 << 16
    rs = rsgpu.ReedSolomon(k, m)
    rng = np.random.default_rng(3)
    full = [to_host(s) for s in _encode_dev(rs, rand_shards(rng, k, n))]
    work = [None if i in (1, 4, 11) else to_dev(full[i]) for i in range(k + m)]
    rs.reconstruct_data(work)
    assert work[11] is None
    assert (to_host(work[1]) == full[1]).all() and (to_host(work[4]) == full[4]).all()
    # reconstruct_some: only shard 12 requested
    work = [None if i in (0, 12, 13) else to_dev(full[i]) for i in range(k + m)]
    req = [0] * (k + m)
    req[12] = 1
    rs.reconstruct_some(work, req)
    assert work[0] is None and work[13] is None
    assert (to_host(work[12]) == full[12]).all()
    # ShardSlot buffers are reused and marked present
    slots = [rsgpu.ShardSlot(to_dev(full[i]), True) for i in range(k + m)]
    for i in (2, 10):
        slots[i].data.zero_()
        slots[i].mark_missing()
    rs.reconstruct(slots)
    assert all(s.present for s in slots)
    assert all((to_host(s.data) == full[i]).all() for i, s in enumerate(slots))


def test_full_size_10x4_round_trip():
    announce("10+4 x 1 MiB round trip")
    k, m, n = 10, 4, 1 << 20
    rs = rsgpu.ReedSolomon(k, m)
    g = torch.Generator(device="cuda").manual_seed(0)
    data = [torch.randint(0, 256, (n,), dtype=torch.uint8, device="cuda", generator=g) for _ in range(k)]
    parity = [torch.empty(n, dtype=torch.uint8, device="cuda") for _ in range(m)]
    rs.encode(data + parity)
    # check a 64 KiB window of every shard against the oracle (columns are independent)
    sl = slice(n - 65536, n)
    _, omat = O.build_matrix(O.GF8, O.VANDERMONDE, k, m)
    exp = O.encode(O.GF8, omat, k, m, [to_host(d[sl]) for d in data])
    for j in range(m):
        assert (to_host(parity[j][sl]) == exp[j]).all()
    full = data + parity
    assert rs.verify(full)
    ref = [t.clone() for t in full]
    rng = np.random.default_rng(4)
    for e in range(1, m + 1):
        missing = set(rng.choice(k + m, e, replace=False).tolist())
        work = [None if i in missing else full[i] for i in range(k + m)]
        rs.reconstruct(work)
        for i in missing:
            assert torch.equal(work[i], ref[i]), (e, i)
    # a single flipped byte anywhere flips the verdict
    full[k + 2][n // 3] ^= 1
    assert not rs.verify(full)


def test_unaligned_views():
    announce("unaligned views")
    k, m, n = 5, 2, 777
    rs = rsgpu.ReedSolomon(k, m)
    _, omat = O.build_matrix(O.GF8, O.VANDERMONDE, k, m)
    rng = np.random.default_rng(6)
    data = rand_shards(rng, k, n)
    big = torch.zeros((k + m) * (n + 3) + 8, dtype=torch.uint8, device="cuda")
    views = [big[i * (n + 3) + 1: i * (n + 3) + 1 + n] for i in range(k + m)]
    for i in range(k):
        views[i].copy_(to_dev(data[i]))
    rs.encode(views)
    exp = O.encode(O.GF8, omat, k, m, data)
    for j in range(m):
        assert (to_host(views[k + j]) == exp[j]).all()


def test_custom_and_fast_one_parity():
    announce("custom matrix / fast one parity")
    k, m, n = 4, 2, 64
    rows = np.array([[1, 2, 3, 4], [5, 6, 7, 8]], np.uint8)
    rs = rsgpu.ReedSolomon.with_custom_matrix(k, m, rows)
    rc, omat = O.build_matrix(O.GF8, O.CUSTOM, k, m, rows.astype(np.uint16))
    assert (rs.matrix() == omat).all()
    rng = np.random.default_rng(8)
    data = rand_shards(rng, k, n)
    shards = _encode_dev(rs, data)
    exp = O.encode(O.GF8, omat, k, m, data)
    assert all((to_host(shards[k + j]) == exp[j]).all() for j in range(m))
    # singular survivor system -> InvalidCustomMatrix before any write
    bad = rsgpu.ReedSolomon.with_custom_matrix(2, 2, np.array([[1, 1], [1, 1]], np.uint8))
    full = _encode_dev(bad, rand_shards(rng, 2, 32))
    sentinel = torch.full((32,), 0xAB, dtype=torch.uint8, device="cuda")
    slots = [rsgpu.ShardSlot(sentinel, False), rsgpu.ShardSlot(full[1].clone(), False),
             rsgpu.ShardSlot(full[2], True), rsgpu.ShardSlot(full[3], True)]
    with pytest.raises(rsgpu.RsError) as ei:
        bad.reconstruct(slots)
    assert ei.value.name == "InvalidCustomMatrix"
    assert int(sentinel[0]) == 0xAB
    # fast one parity: XOR of the data
    f1 = rsgpu.ReedSolomon(5, 1, fast_one_parity=True)
    data = rand_shards(rng, 5, 100)
    sh = _encode_dev(f1, data)
    assert (to_host(sh[5]) == np.bitwise_xor.reduce(np.stack(data))).all()


def test_validation_errors_match_reference_order():
    announce("validation errors")
    rs = rsgpu.ReedSolomon(4, 2)
    t = lambda n=16: torch.zeros(n, dtype=torch.uint8, device="cuda")
    cases = [
        (lambda: rs.encode([t() for _ in range(5)]), "TooFewShards"),
        (lambda: rs.encode([t() for _ in range(7)]), "TooManyShards"),
        (lambda: rs.encode([t(0)] + [t() for _ in range(5)]), "EmptyShard"),
        (lambda: rs.encode([t(), t(8)] + [t() for _ in range(4)]), "IncorrectShardSize"),
        (lambda: rs.encode_sep([t() for _ in range(3)], [t(), t()]), "TooFewDataShards"),
        (lambda: rs.encode_sep([t() for _ in range(4)], [t()]), "TooFewParityShards"),
        (lambda: rs.encode_sep([t() for _ in range(4)], [t(8), t(8)]), "IncorrectShardSize"),
        (lambda: rs.verify([t() for _ in range(5)]), "TooFewShards"),
        (lambda: rs.reconstruct([None, None, None, t(), t(), t()]), "TooFewShardsPresent"),
        (lambda: rs.reconstruct([t(0), None, t(), t(), t(), t()]), "EmptyShard"),
        (lambda: rs.reconstruct([t(8), None, t(), t(), t(), t()]), "IncorrectShardSize"),
        (lambda: rs.reconstruct_some([None, t(), t(), t(), t(), t()], [1, 0, 0]), "InvalidShardFlags"),
        (lambda: rs.verify_with_buffer([t() for _ in range(6)], [t()]), "TooFewBufferShards"),
    ]
    for fn, name in cases:
        with pytest.raises(rsgpu.RsError) as ei:
            fn()
        assert ei.value.name == name
    for args, name in [((0, 2), "TooFewDataShards"), ((2, 0), "TooFewParityShards"),
                       ((200, 57), "TooManyShards")]:
        with pytest.raises(rsgpu.RsError) as ei:
            rsgpu.ReedSolomon(*args)
        assert ei.value.name == name


def test_verify_with_buffer_and_async():
    announce("verify_with_buffer / async")
    k, m, n = 8, 3, 5000
    rs = rsgpu.ReedSolomon(k, m)
    rng = np.random.default_rng(12)
    shards = _encode_dev(rs, rand_shards(rng, k, n))
    buf = [torch.zeros(n, dtype=torch.uint8, device="cuda") for _ in range(m)]
    assert rs.verify_with_buffer(shards, buf)
    assert all(torch.equal(b, shards[k + j]) for j, b in enumerate(buf))
    flag = torch.ones(1, dtype=torch.int32, device="cuda")
    rs.verify_async(shards, flag)
    torch.cuda.synchronize()
    assert int(flag) == 0
    shards[k][n - 1] ^= 0x40
    rs.verify_async(shards, flag)
    torch.cuda.synchronize()
    assert int(flag) != 0


def test_apply_rows_primitive_and_host_paths():
    announce("apply_rows / host paths")
    rng = np.random.default_rng(13)
    for field in ("gf8", "gf16"):
        r, k, n = 3, 5, 130
        rows = rng.integers(0, 256 if field == "gf8" else 65536, (r, k))
        ins = rand_shards(rng, k, n)
        outs = [torch.zeros(n, dtype=torch.uint8, device="cuda") for _ in range(r)]
        rsgpu.apply_rows(rows, [to_dev(x) for x in ins], outs, field=field)
        exp = O.apply_rows(ORA_FIELD[field], rows.astype(np.uint16), ins)
        assert all((to_host(o) == e).all() for o, e in zip(outs, exp))
    rs = rsgpu.ReedSolomon(6, 3)
    _, omat = O.build_matrix(O.GF8, O.VANDERMONDE, 6, 3)
    host = rand_shards(rng, 6, 999) + [np.zeros(999, np.uint8) for _ in range(3)]
    rs.encode_host(host)
    exp = O.encode(O.GF8, omat, 6, 3, host[:6])
    assert all((host[6 + j] == exp[j]).all() for j in range(3))
    present = np.array([1, 0, 1, 1, 0, 1, 1, 1, 0], np.uint8)
    damaged = [h if p else np.zeros_like(h) for h, p in zip(host, present)]
    arrs, pres = rs.reconstruct_host(damaged, present)
    assert pres.all() and all((a == h).all() for a, h in zip(arrs, host))


def test_apply_rows_table_cache_back_to_back():
    """rsgpu_apply_rows keeps coefficient tables per row set: 70 row sets (more than
    the 64 kept, so the cache is cleared on the way) and a repeated set, launched back
    to back on one stream without synchronisation, each equal the oracle."""
    announce("apply_rows table cache")
    rng = np.random.default_rng(21)
    r, k, n = 2, 6, 4096 + 48
    ins = rand_shards(rng, k, n)
    dins = [to_dev(x) for x in ins]
    sets = [rng.integers(0, 256, (r, k)) for _ in range(70)]
    sets.append(sets[3])
    outs = []
    for rows in sets:
        o = [torch.zeros(n, dtype=torch.uint8, device="cuda") for _ in range(r)]
        rsgpu.apply_rows(rows, dins, o)
        outs.append(o)
    torch.cuda.synchronize()
    for i, (rows, o) in enumerate(zip(sets, outs)):
        exp = O.apply_rows(O.GF8, rows.astype(np.uint16), ins)
        assert all((to_host(a) == e).all() for a, e in zip(o, exp)), i
