"""Progressive recovery, decode_idx (src/galois_8/policy.rs:543-704): the
reference's own tests (src/tests/mod.rs:3970-4176, tests/compact_decode_matrix.rs:270-300)
and a per-call comparison with the oracle: each call adds sum_col row[col] * input
over its subset of the expected survivors, with the rows read off the oracle's
decode matrix (composed through the parity rows for parity destinations)."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import announce, to_dev, to_host

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
rsgpu = pytest.importorskip("rsgpu")


def _encoded(rs, k, m, n, seed):
    rng = np.random.default_rng(seed)
    sh = [to_dev(rng.integers(0, 256, n, dtype=np.uint8)) for _ in range(k)] + \
         [torch.zeros(n, dtype=torch.uint8, device="cuda") for _ in range(m)]
    rs.encode(sh)
    torch.cuda.synchronize()
    return sh


def _zeros(n):
    return torch.zeros(n, dtype=torch.uint8, device="cuda")


def test_progressive_matches_reconstruct_some():
    announce("decode_idx progressive 5+3")
    r = rsgpu.ReedSolomon(5, 3)
    n = 256 * 1024
    shards = _encoded(r, 5, 3, n, 1)
    expected = [s.clone() for s in shards]
    expected[1] = None
    expected[4] = None
    required = [False] * 8
    required[1] = required[4] = True
    r.reconstruct_some(expected, required)
    dst = [None] * 8
    dst[1], dst[4] = _zeros(n), _zeros(n)
    expect_input = [True, False, True, True, False, True, True, False]
    first = [None] * 8
    first[0], first[2] = shards[0], shards[2]
    r.decode_idx(dst, expect_input, first)
    second = [None] * 8
    second[3], second[5], second[6] = shards[3], shards[5], shards[6]
    r.decode_idx(dst, expect_input, second)
    torch.cuda.synchronize()
    assert torch.equal(expected[1], dst[1]) and torch.equal(expected[4], dst[4])
    assert torch.equal(dst[1], shards[1]) and torch.equal(dst[4], shards[4])


def test_merge_mode_accumulates_partial_results():
    announce("decode_idx merge mode")
    r = rsgpu.ReedSolomon(5, 3)
    n = 128 * 1024
    shards = _encoded(r, 5, 3, n, 2)
    expect_input = [True, False, True, True, False, True, True, False]
    pa, pb = [None] * 8, [None] * 8
    pa[1], pa[4], pb[1], pb[4] = _zeros(n), _zeros(n), _zeros(n), _zeros(n)
    ia, ib = [None] * 8, [None] * 8
    ia[0], ia[2] = shards[0], shards[2]
    ib[3], ib[5], ib[6] = shards[3], shards[5], shards[6]
    r.decode_idx(pa, expect_input, ia)
    r.decode_idx(pb, expect_input, ib)
    r.decode_idx(pa, None, pb)
    torch.cuda.synchronize()
    assert torch.equal(pa[1], shards[1]) and torch.equal(pa[4], shards[4])


def test_merge_mode_lengths_may_differ_and_unaligned_views():
    announce("decode_idx merge ragged")
    r = rsgpu.ReedSolomon(3, 2)
    rng = np.random.default_rng(3)
    lens = [17, 4096 + 3, 0, 1 << 20, 33]
    a = [to_dev(rng.integers(0, 256, x, dtype=np.uint8)) for x in lens]
    b = [to_dev(rng.integers(0, 256, x + 1, dtype=np.uint8))[1:] for x in lens]  # odd offsets
    want = [to_host(x) ^ to_host(y) for x, y in zip(a, b)]
    inp = list(b)
    inp[2] = None
    r.decode_idx(a, None, inp)
    torch.cuda.synchronize()
    for i in range(5):
        exp = want[i] if inp[i] is not None else to_host(a[i])
        assert (to_host(a[i]) == exp).all(), i


def _err(fn):
    with pytest.raises(rsgpu.RsError) as ei:
        fn()
    return ei.value.name


def test_errors_in_the_reference_order():
    announce("decode_idx errors")
    r = rsgpu.ReedSolomon(5, 3)
    assert _err(lambda: r.decode_idx([None] * 8, [True, False], [None] * 8)) == "InvalidShardFlags"
    assert _err(lambda: r.decode_idx([None] * 7, [True] * 8, [None] * 8)) == "TooFewShards"
    assert _err(lambda: r.decode_idx([None] * 8, [True] * 8, [None] * 7)) == "TooFewShards"
    assert _err(lambda: r.decode_idx([None] * 9, [True] * 8, [None] * 8)) == "TooManyShards"
    gf8 = rsgpu.ReedSolomon(32, 16, codec_family="leopard_gf8")
    assert _err(lambda: gf8.decode_idx([None] * 48, None, [None] * 48)) == "UnsupportedCodecFamily"
    gf16 = rsgpu.ReedSolomon(254, 3, leopard_mode="as_needed")
    assert gf16.codec_family() == "leopard_gf16"
    assert _err(lambda: gf16.decode_idx([None] * 257, None, [None] * 257)) == "UnsupportedCodecFamily"
    dst = [None] * 8
    dst[1] = _zeros(16)
    inp = [None] * 8
    inp[0] = torch.ones(8, dtype=torch.uint8, device="cuda")
    ex = [True, False, True, True, False, True, True, False]
    assert _err(lambda: r.decode_idx(dst, ex, inp)) == "IncorrectShardSize"
    assert int(dst[1].sum()) == 0
    m_inp = [None] * 8
    m_inp[1] = torch.ones(8, dtype=torch.uint8, device="cuda")
    assert _err(lambda: r.decode_idx([None] * 8, None, m_inp)) == "TooFewShards"
    dst = [None] * 8
    dst[1] = _zeros(8)
    few = [True, False, True, False, False, False, False, False]
    assert _err(lambda: r.decode_idx(dst, few, [None] * 8)) == "TooFewShardsPresent"
    # more than k expected survivors: get_data_decode_matrix takes exactly k
    inp = [None] * 8
    inp[0] = torch.ones(8, dtype=torch.uint8, device="cuda")
    assert _err(lambda: r.decode_idx(dst, [True] * 8, inp)) == "TooFewShardsPresent"
    # a merge batch with a bad pair after a good one changes nothing
    a, b = _zeros(8), torch.ones(8, dtype=torch.uint8, device="cuda")
    dst = [a, None, None, None, None, None, None, _zeros(9)]
    inp = [b, None, None, None, None, None, None, torch.ones(8, dtype=torch.uint8, device="cuda")]
    assert _err(lambda: r.decode_idx(dst, None, inp)) == "IncorrectShardSize"
    torch.cuda.synchronize()
    assert int(a.sum()) == 0
    # no destination, or no input among the survivors: Ok, nothing changes
    only_dst = [None] * 8
    only_dst[1] = _zeros(8)
    r.decode_idx(only_dst, ex, [None] * 8)
    off_survivor = [None] * 8
    off_survivor[1] = torch.ones(8, dtype=torch.uint8, device="cuda")  # not expected: ignored
    r.decode_idx(only_dst, ex, off_survivor)
    only_inp = [None] * 8
    only_inp[0] = torch.ones(8, dtype=torch.uint8, device="cuda")
    r.decode_idx([None] * 8, ex, only_inp)
    torch.cuda.synchronize()
    assert int(only_dst[1].sum()) == 0
    # with neither a destination nor an input there is no shard length
    assert _err(lambda: r.decode_idx([None] * 8, ex, [None] * 8)) == "TooFewShardsPresent"


def test_compact_rows_recover_missing_identity_and_parity_rows():
    """64+6 with four erasures, batches of 11 survivors; destinations include a
    surviving data shard (2) and a present parity shard outside the survivors (69)."""
    announce("decode_idx 64+6 progressive batches")
    k, m, n = 64, 6, 65
    codec = rsgpu.ReedSolomon(k, m)
    expected = _encoded(codec, k, m, n, 65)
    missing = [0, 21, 63, 64]
    survivors = [i for i in range(70) if i not in missing][:64]
    expected_inputs = [i in survivors for i in range(70)]
    targets = [0, 2, 21, 63, 64, 69]
    output = [_zeros(n) if i in targets else None for i in range(70)]
    for b0 in range(0, len(survivors), 11):
        batch = survivors[b0:b0 + 11]
        inp = [expected[i] if i in batch else None for i in range(70)]
        codec.decode_idx(output, expected_inputs, inp)
    torch.cuda.synchronize()
    for i in range(70):
        if i in targets:
            assert torch.equal(output[i], expected[i]), i
        else:
            assert output[i] is None


def _oracle_rows(field, mat, k, m, expect, targets):
    """Rows over the k expected survivors: data target t -> row t of the decode
    matrix, parity target -> the parity row times the decode matrix."""
    rc, valid, dec = O.decode_matrix(field, mat, k, m, np.array(expect, np.uint8))
    assert rc == 0
    mul = O.gf8_mul if field == O.GF8 else O.gf16_mul
    rows = []
    for t in targets:
        if t < k:
            rows.append([int(dec[t][c]) for c in range(k)])
        else:
            row = []
            for c in range(k):
                v = 0
                for i in range(k):
                    v ^= mul(int(mat[t][i]), int(dec[i][c]))
                row.append(v)
            rows.append(row)
    return valid, rows


@pytest.mark.parametrize("field,mode,k,m,n", [("gf8", O.VANDERMONDE, 10, 4, 4096 + 48),
                                              ("gf8", O.CAUCHY, 20, 4, 1000),
                                              ("gf8", O.VANDERMONDE, 10, 2, 640),
                                              ("gf8", O.JERASURE, 12, 6, 192),
                                              ("gf16", O.VANDERMONDE, 8, 4, 2048 + 6)])
def test_each_call_matches_the_oracle_on_non_codewords(field, mode, k, m, n):
    """Random inputs (not a codeword), k random expected survivors, destinations
    anywhere (missing or present, data or parity), survivors split into three
    random calls: after every call the destinations equal the oracle's sum."""
    announce(f"decode_idx oracle {field} {k}+{m} mode={mode} n={n}")
    f = O.GF8 if field == "gf8" else O.GF16
    mname = {O.VANDERMONDE: "vandermonde", O.CAUCHY: "cauchy", O.JERASURE: "jerasure_like"}[mode]
    rs = rsgpu.ReedSolomon(k, m, field=field, matrix_mode=mname)
    _, mat = O.build_matrix(f, mode, k, m)
    rng = np.random.default_rng(k * 31 + m + n)
    total = k + m
    host = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(total)]
    dev = [to_dev(h) for h in host]
    for trial in range(3):
        surv = sorted(rng.choice(total, k, replace=False).tolist())
        expect = [i in surv for i in range(total)]
        n_t = int(rng.integers(1, total + 1)) if trial else total  # trial 0: every shard
        targets = sorted(rng.choice(total, n_t, replace=False).tolist())
        valid, rows = _oracle_rows(f, mat, k, m, expect, targets)
        assert valid == surv
        dst = [_zeros(n) if i in targets else None for i in range(total)]
        want = {t: np.zeros(n, np.uint8) for t in targets}
        cols = list(range(k))
        rng.shuffle(cols)
        parts = [sorted(cols[:3]), sorted(cols[3:k // 2]), sorted(cols[k // 2:])]
        for part in parts:
            if not part:
                continue
            inp = [None] * total
            for c in part:
                inp[surv[c]] = dev[surv[c]]
            rs.decode_idx(dst, expect, inp)
            sub = [[row[c] for c in part] for row in rows]
            outs = O.apply_rows(f, np.array(sub), [host[surv[c]] for c in part])
            for t, o in zip(targets, outs):
                want[t] ^= o
            torch.cuda.synchronize()
            for t in targets:
                assert (to_host(dst[t]) == want[t]).all(), (trial, part, t)
