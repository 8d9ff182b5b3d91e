"""Every kernel variant a dispatcher may pick must produce the oracle's bytes:
the tiled matrix-apply variants on classic GF(2^8) and tower codecs, the
generator-matrix path of Leopard GF(2^8) codecs, and the Leopard FFT column
tiles.  Also checks the dispatch table and the validity rules."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import announce, rand_shards, to_dev, to_host

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
rsgpu = pytest.importorskip("rsgpu")

NV = None


def n_variants():
    global NV
    if NV is None:
        NV = rsgpu.variant_count()
    return NV


def _classic_expected(field, k, m, data):
    f = O.GF8 if field == "gf8" else O.GF16
    rc, mat = O.build_matrix(f, O.VANDERMONDE, k, m)
    assert rc == 0
    return mat, O.encode(f, mat, k, m, data)


@pytest.mark.parametrize("field,k,m,length", [("gf8", 10, 4, 65536 + 48), ("gf8", 3, 9, 4096),
                                              ("gf8", 70, 3, 1000), ("gf8", 7, 3, 4099),
                                              ("gf16", 6, 3, 8192 + 2),
                                              ("gf16", 5, 5, 333 * 2)])
def test_classic_every_variant_matches_oracle(field, k, m, length):
    announce(f"variants classic {field} {k}+{m} len={length}")
    rng = np.random.default_rng(k * 100 + m + length)
    data = rand_shards(rng, k, length)
    mat, par = _classic_expected(field, k, m, data)
    rs = rsgpu.ReedSolomon(k, m, field=field)
    erased = sorted(rng.choice(k + m, m, replace=False).tolist())
    for v in range(n_variants()):
        valid = rs.variant_valid("encode", 0, v)
        assert valid == (v < 15), v
        if not valid:
            continue
        rs.set_variant(v)
        sh = [to_dev(d) for d in data] + [torch.zeros(length, dtype=torch.uint8, device="cuda")
                                         for _ in range(m)]
        rs.encode(sh)
        for j in range(m):
            assert (to_host(sh[k + j]) == par[j]).all(), (v, j)
        assert rs.verify(sh), v
        sh[k + m - 1][length // 2] ^= 7
        assert not rs.verify(sh), v
        sh[k + m - 1][length // 2] ^= 7
        full = [to_host(x) for x in sh]
        work = [None if i in erased else to_dev(full[i]) for i in range(k + m)]
        rs.reconstruct(work)
        for i in erased:
            assert (to_host(work[i]) == full[i]).all(), (v, i)
    rs.set_variant(-1)


@pytest.mark.parametrize("bits,k,m,length", [(8, 10, 4, 4096), (8, 128, 64, 1024), (8, 33, 17, 640),
                                             (16, 10, 4, 4096), (16, 300, 45, 640)])
def test_leopard_every_variant_matches_oracle(bits, k, m, length):
    announce(f"variants leopard gf{bits} {k}+{m} len={length}")
    fam = "leopard_gf8" if bits == 8 else "leopard_gf16"
    rng = np.random.default_rng(bits * 7 + k + m)
    data = rand_shards(rng, k, length)
    rc, par = O.leo_encode(bits, k, m, data)
    assert rc == 0
    rs = rsgpu.ReedSolomon(k, m, codec_family=fam)
    full = data + list(par)
    exactly_k = sorted(rng.choice(k + m, m, replace=False).tolist())
    fewer = exactly_k[: max(1, m // 2)]
    for v in range(n_variants()):
        ok_enc = rs.variant_valid("encode", 0, v)
        assert ok_enc == (v >= 15 or bits == 8), v
        if not ok_enc:
            continue
        rs.set_variant(v)
        sh = [to_dev(d) for d in data] + [torch.zeros(length, dtype=torch.uint8, device="cuda")
                                         for _ in range(m)]
        rs.encode(sh)
        for j in range(m):
            assert (to_host(sh[k + j]) == par[j]).all(), (v, j)
        assert rs.verify(sh), v
        sh[0][3] ^= 1
        assert not rs.verify(sh), v
        sh[0][3] ^= 1
        for erased in (exactly_k, fewer):
            if not rs.variant_valid("reconstruct", len(erased), v):
                assert v < 15 and len(erased) != m
                continue
            work = [None if i in erased else to_dev(full[i]) for i in range(k + m)]
            rs.reconstruct(work)
            for i in erased:
                assert (to_host(work[i]) == full[i]).all(), (v, erased, i)
    rs.set_variant(-1)


def test_leopard_matrix_path_matches_fft_on_non_codewords():
    """With exactly k shards present the generator-matrix solve and the FFT decoder
    agree on arbitrary (non-codeword) inputs: both equal the oracle's decoder."""
    announce("variants leopard matrix path non-codeword")
    k, m, n = 12, 4, 256
    rng = np.random.default_rng(99)
    shards = rand_shards(rng, k + m, n)
    missing = (1, 5, 13, 14)
    present = [0 if i in missing else 1 for i in range(k + m)]
    rc, exp = O.leo_reconstruct(8, k, m, [s.copy() for s in shards], present)
    assert rc == 0
    rs = rsgpu.ReedSolomon(k, m, codec_family="leopard_gf8")
    for v in (0, 7, 14, -1):
        rs.set_variant(v)
        work = [None if i in missing else to_dev(shards[i]) for i in range(k + m)]
        rs.reconstruct(work)
        for i in missing:
            assert (to_host(work[i]) == exp[i]).all(), (v, i)
    rs.set_variant(-1)


def test_variant_table_routes_calls():
    announce("variant table")
    n_ops, n_buckets, b0 = rsgpu.dispatch_dims()
    k, m = 8, 4
    rs = rsgpu.ReedSolomon(k, m)
    table = np.full((n_ops, m + 1, n_buckets), -1, np.int16)
    table[0, 0, :] = 12          # encode: 64-lane, 1 row per workgroup
    table[2, 2, :] = 16          # reconstruct with 2 erasures: a Leopard tile -> not valid here
    table[1, 0, 20] = 3
    rs.set_variant_table(table)
    assert rs.variant_for("encode", 0, 4096) == 12
    assert rs.variant_for("reconstruct", 2, 4096) == -1
    assert rs.variant_for("verify", 0, 1 << 16) == 3   # round(2*16) - 12 = 20
    assert rs.variant_for("verify", 0, 1 << 17) == -1
    rng = np.random.default_rng(3)
    data = rand_shards(rng, k, 4096)
    mat, par = _classic_expected("gf8", k, m, data)
    sh = [to_dev(d) for d in data] + [torch.zeros(4096, dtype=torch.uint8, device="cuda")
                                     for _ in range(m)]
    rs.encode(sh)
    for j in range(m):
        assert (to_host(sh[k + j]) == par[j]).all()
    rs.set_variant_table(None)
    assert rs.variant_for("encode", 0, 4096) == -1
    with pytest.raises(rsgpu.RsError):
        rs.set_variant(n_variants())
