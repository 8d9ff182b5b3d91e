"""GPU parity of the Leopard GF(2^8) / GF(2^16) FFT codecs against the oracle
(bit-exact parity, rebuilt shards and verify verdicts), including the
reference's Go-derived GF16 known answers."""
import itertools

import numpy as np
import pytest

import oracle_lib as O
from gpu_util import announce, to_dev, to_host

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
rsgpu = pytest.importorskip("rsgpu")

FAMILY = {8: "leopard_gf8", 16: "leopard_gf16"}


def _encode(rs, data, m):
    n = len(data[0])
    shards = [to_dev(d) for d in data] + [torch.zeros(n, dtype=torch.uint8, device="cuda")
                                         for _ in range(m)]
    rs.encode(shards)
    torch.cuda.synchronize()
    return shards


def test_gf16_known_answer_parity():
    announce("leopard gf16 3+2 KAT")
    rs = rsgpu.ReedSolomon(3, 2, codec_family="leopard_gf16")
    data = [np.array([(i * 64 + j) & 0xFF for j in range(64)], np.uint8) for i in range(3)]
    sh = _encode(rs, data, 2)
    # src/core/leopard_gf16/tests.rs:646-697
    assert to_host(sh[3])[:16].tolist() == [59, 148, 199, 104, 147, 60, 111, 192, 227, 76, 31, 176,
                                             75, 228, 183, 24]
    assert to_host(sh[4])[:16].tolist() == [251, 85, 5, 171, 87, 249, 169, 7, 43, 133, 213, 123, 135,
                                             41, 121, 215]


@pytest.mark.parametrize("bits,k,m", [(8, 10, 4), (8, 3, 1), (8, 1, 2), (8, 12, 7), (8, 33, 17),
                                      (8, 128, 64), (8, 5, 60), (8, 192, 64),
                                      (16, 10, 4), (16, 3, 2), (16, 33, 17), (16, 100, 30),
                                      (16, 300, 45)])
@pytest.mark.parametrize("length", [64, 640])
def test_encode_matches_oracle(bits, k, m, length):
    announce(f"leopard gf{bits} encode {k}+{m} len={length}")
    rng = np.random.default_rng(bits * 1000 + k * 10 + m + length)
    data = [rng.integers(0, 256, length, dtype=np.uint8) for _ in range(k)]
    rs = rsgpu.ReedSolomon(k, m, codec_family=FAMILY[bits])
    sh = _encode(rs, data, m)
    rc, exp = O.leo_encode(bits, k, m, data)
    assert rc == 0
    for j in range(m):
        assert (to_host(sh[k + j]) == exp[j]).all(), j
    assert rs.verify(sh)
    sh[k][length - 1] ^= 1
    assert not rs.verify(sh)


@pytest.mark.parametrize("bits,k,m,length", [(8, 10, 4, (1 << 20) + 192), (8, 32, 16, (1 << 20) + 64),
                                             (8, 192, 64, (64 << 10) + 64), (16, 64, 16, (256 << 10) + 64),
                                             (16, 1024, 256, 8192)])
def test_encode_long_shards_match_oracle(bits, k, m, length):
    """Multi-workgroup launches with 16-byte accesses and a ragged last tile: the
    encode that holds every data group in LDS (small codecs) and the one that
    streams the groups through two buffers (LDS-bound codecs)."""
    announce(f"leopard gf{bits} long encode {k}+{m} len={length}")
    rng = np.random.default_rng(bits + k + m + length)
    data = [rng.integers(0, 256, length, dtype=np.uint8) for _ in range(k)]
    rs = rsgpu.ReedSolomon(k, m, codec_family=FAMILY[bits])
    sh = _encode(rs, data, m)
    rc, exp = O.leo_encode(bits, k, m, data)
    assert rc == 0
    for j in range(m):
        assert (to_host(sh[k + j]) == exp[j]).all(), j
    assert rs.verify(sh)


@pytest.mark.parametrize("bits,k,m", [(8, 1, 1), (16, 1, 1), (8, 2, 1), (8, 10, 4), (8, 12, 7),
                                      (8, 128, 64), (16, 10, 4), (16, 33, 17), (16, 300, 45)])
def test_reconstruct_random_patterns_match_oracle(bits, k, m):
    announce(f"leopard gf{bits} reconstruct {k}+{m}")
    n = 256
    rng = np.random.default_rng(k + m + bits)
    data = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k)]
    rs = rsgpu.ReedSolomon(k, m, codec_family=FAMILY[bits])
    full = [to_host(s) for s in _encode(rs, data, m)]
    for trial in range(6):
        e = int(rng.integers(1, m + 1))
        missing = set(rng.choice(k + m, e, replace=False).tolist())
        work = [None if i in missing else to_dev(full[i]) for i in range(k + m)]
        rs.reconstruct(work)
        for i in range(k + m):
            assert (to_host(work[i]) == full[i]).all(), (trial, i, sorted(missing))


@pytest.mark.parametrize("k,m,length", [(4096, 1024, 128), (10000, 2000, 64)])
def test_gf16_wide_codecs_match_oracle(k, m, length):
    """GF16 shapes whose decode transform (8192 and 16384 lanes) leaves room in
    LDS for only 2 words and 1 word per lane: the narrow-tile kernel paths."""
    announce(f"leopard gf16 wide {k}+{m} len={length}")
    rng = np.random.default_rng(k + m)
    data = [rng.integers(0, 256, length, dtype=np.uint8) for _ in range(k)]
    rc, par = O.leo_encode(16, k, m, data)
    assert rc == 0
    full = np.stack(data + list(par))
    rs = rsgpu.ReedSolomon(k, m, codec_family="leopard_gf16")
    dev = torch.from_numpy(full).cuda()
    dev[k:] = 0
    rs.encode(list(dev))
    assert (dev[k:].cpu().numpy() == full[k:]).all()
    assert rs.verify(list(dev))
    for e in (m, m // 3, 1):
        missing = set(rng.choice(k + m, e, replace=False).tolist())
        work = [None if i in missing else dev[i] for i in range(k + m)]
        rs.reconstruct(work)
        for i in sorted(missing):
            assert (to_host(work[i]) == full[i]).all(), (e, i)


def test_every_pattern_small_gf8_and_gf16():
    announce("leopard all patterns 5+3")
    for bits in (8, 16):
        k, m, n = 5, 3, 128
        rng = np.random.default_rng(bits)
        data = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k)]
        rs = rsgpu.ReedSolomon(k, m, codec_family=FAMILY[bits])
        full = [to_host(s) for s in _encode(rs, data, m)]
        for e in range(1, m + 1):
            for missing in itertools.combinations(range(k + m), e):
                work = [None if i in missing else to_dev(full[i]) for i in range(k + m)]
                rs.reconstruct(work)
                for i in range(k + m):
                    assert (to_host(work[i]) == full[i]).all(), (bits, missing, i)


def test_inconsistent_input_matches_oracle():
    """Non-codeword inputs: the device decoder must produce the oracle's bytes."""
    announce("leopard inconsistent input")
    for bits in (8, 16):
        k, m, n = 6, 4, 192
        rng = np.random.default_rng(40 + bits)
        shards = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k + m)]
        rs = rsgpu.ReedSolomon(k, m, codec_family=FAMILY[bits])
        for missing in [(0,), (2, 7), (1, 5, 8, 9)]:
            present = [0 if i in missing else 1 for i in range(k + m)]
            rc, exp = O.leo_reconstruct(bits, k, m, [s.copy() for s in shards], present)
            assert rc == 0
            work = [None if i in missing else to_dev(shards[i]) for i in range(k + m)]
            rs.reconstruct(work)
            for i in missing:
                assert (to_host(work[i]) == exp[i]).all(), (bits, missing, i)


@pytest.mark.parametrize("k,m,n", [(6, 4, 192), (10, 4, (64 << 10) + 64), (12, 7, 4096), (32, 16, 640),
                                   (128, 64, 1024)])
def test_decoder_matrix_reconstruct_matches_oracle(k, m, n):
    """Matrix variants with more than k shards present apply the FFT decoder's own
    matrix over all of them: on non-codeword inputs the bytes equal the oracle's
    decoder, for every erasure count below m, for reconstruct_data, and again
    from the cached matrix."""
    announce(f"leopard decoder matrix {k}+{m} len={n}")
    rng = np.random.default_rng(k * 100 + m)
    shards = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k + m)]
    rs = rsgpu.ReedSolomon(k, m, codec_family="leopard_gf8")
    rs.set_variant(5)  # a matrix-apply variant
    for e in (range(1, m) if m <= 16 else (1, 2, 17, m - 1)):  # > 64 present: split over inputs
        missing = sorted(rng.choice(k + m, e, replace=False).tolist())
        present = [0 if i in missing else 1 for i in range(k + m)]
        rc, exp = O.leo_reconstruct(8, k, m, [s.copy() for s in shards], present)
        assert rc == 0
        for rep in range(2):  # the second call uses the cached matrix
            work = [None if i in missing else to_dev(shards[i]) for i in range(k + m)]
            rs.reconstruct(work)
            for i in missing:
                assert (to_host(work[i]) == exp[i]).all(), (e, missing, i, rep)
        work = [None if i in missing else to_dev(shards[i]) for i in range(k + m)]
        rs.reconstruct_data(work)
        for i in missing:
            if i < k:
                assert (to_host(work[i]) == exp[i]).all(), (e, missing, i)
            else:
                assert work[i] is None


@pytest.mark.parametrize("k,m,n,counts", [(10, 4, 192, None), (64, 16, 4096, None),
                                           (300, 45, 640, (1, 2, 9, 44, 45)),
                                           (1024, 256, 4096, (1, 3, 256))])
def test_gf16_decoder_matrix_reconstruct_matches_oracle(k, m, n, counts):
    """Leopard GF(2^16) matrix variants apply the FFT decoder's own matrix over
    every present shard on the split-layout vector tiles: on non-codeword inputs
    the bytes equal the oracle's decoder for each erasure count (up to m, i.e.
    exactly k present), for reconstruct_data, and again from the cached matrix."""
    announce(f"leopard gf16 decoder matrix {k}+{m} len={n}")
    rng = np.random.default_rng(k * 100 + m + 16)
    shards = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k + m)]
    rs = rsgpu.ReedSolomon(k, m, codec_family="leopard_gf16")
    assert rs.variant_valid("reconstruct", 1, 5) and not rs.variant_valid("encode", 0, 5)
    for e in (counts or range(1, m + 1)):
        rs.set_variant(5 if e % 2 else 0)  # two matrix-apply tiles
        missing = sorted(rng.choice(k + m, e, replace=False).tolist())
        present = [0 if i in missing else 1 for i in range(k + m)]
        rc, exp = O.leo_reconstruct(16, k, m, [s.copy() for s in shards], present)
        assert rc == 0
        for rep in range(2):  # the second call uses the cached matrix
            work = [None if i in missing else to_dev(shards[i]) for i in range(k + m)]
            rs.reconstruct(work)
            for i in missing:
                assert (to_host(work[i]) == exp[i]).all(), (e, missing, i, rep)
        work = [None if i in missing else to_dev(shards[i]) for i in range(k + m)]
        rs.reconstruct_data(work)
        for i in missing:
            if i < k:
                assert (to_host(work[i]) == exp[i]).all(), (e, missing, i)
            else:
                assert work[i] is None
    rs.set_variant(-1)


def test_reconstruct_data_some_and_errors():
    announce("leopard reconstruct_data / errors")
    k, m, n = 8, 4, 128
    rs = rsgpu.ReedSolomon(k, m, codec_family="leopard_gf8")
    rng = np.random.default_rng(5)
    full = [to_host(s) for s in _encode(rs, [rng.integers(0, 256, n, dtype=np.uint8)
                                             for _ in range(k)], m)]
    work = [None if i in (1, 9) else to_dev(full[i]) for i in range(k + m)]
    rs.reconstruct_data(work)
    assert work[9] is None and (to_host(work[1]) == full[1]).all()
    with pytest.raises(rsgpu.RsError) as ei:
        rs.reconstruct([None] * 5 + [to_dev(x) for x in full[5:]])
    assert ei.value.name == "TooFewShardsPresent"
    bad = [to_dev(np.zeros(100, np.uint8)) for _ in range(k + m)]
    with pytest.raises(rsgpu.RsError) as ei:
        rs.encode(bad)
    assert ei.value.name == "IncorrectShardSize"
    with pytest.raises(rsgpu.RsError) as ei:
        rsgpu.ReedSolomon(200, 100, codec_family="leopard_gf8")
    assert ei.value.name == "TooManyShards"


def test_prefer_leopard_mode_resolves_family():
    # resolve_codec_family (src/core/leopard.rs:191-223)
    assert rsgpu.ReedSolomon(10, 4, leopard_mode="prefer_leopard").codec_family() == "leopard_gf8"
    assert rsgpu.ReedSolomon(10, 4, leopard_mode="prefer_gf16").codec_family() == "leopard_gf16"
    assert rsgpu.ReedSolomon(10, 4, leopard_mode="as_needed").codec_family() == "classic"
    assert rsgpu.ReedSolomon(300, 100, leopard_mode="as_needed").codec_family() == "leopard_gf16"


@pytest.mark.parametrize("bits", [8, 16])
def test_plug_points_match_oracle(bits):
    """leopard_gf8_encode / leopard_gf16_encode and reconstruct_with_tables by shape,
    without a codec object (src/core/leopard.rs:277-317): parity equals the oracle,
    requested outputs are rebuilt, an output of length 0 is not requested, and
    validation follows validate_leopard_decode_buffers (reconstruct.rs:140-185)."""
    announce(f"leopard gf{bits} plug points")
    k, m, n = 20, 6, 4096
    rng = np.random.default_rng(70 + bits)
    data = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k)]
    rc, par = O.leo_encode(bits, k, m, data)
    assert rc == 0
    dd = [to_dev(d) for d in data]
    pp = [torch.zeros(n, dtype=torch.uint8, device="cuda") for _ in range(m)]
    rsgpu.leopard_encode(bits, k, m, dd, pp)
    for j in range(m):
        assert (to_host(pp[j]) == par[j]).all(), j
    full = data + list(par)
    missing = [0, 3, k + 1, k + 4]
    present = [i not in missing for i in range(k + m)]
    inputs = [to_dev(full[i]) if present[i] else None for i in range(k + m)]
    outputs = [None] * (k + m)
    for i in missing[:3]:
        outputs[i] = torch.zeros(n, dtype=torch.uint8, device="cuda")
    rsgpu.leopard_reconstruct(bits, present, outputs, inputs, k, m)
    for i in missing[:3]:
        assert (to_host(outputs[i]) == full[i]).all(), i
    assert outputs[missing[3]] is None
    # errors, in the reference's order
    with pytest.raises(rsgpu.RsError) as ei:
        rsgpu.leopard_encode(bits, k, m, dd[:-1], pp)
    assert ei.value.name == "TooFewShards"
    short = [d[:100] for d in dd]
    with pytest.raises(rsgpu.RsError) as ei:
        rsgpu.leopard_encode(bits, k, m, short, [p[:100] for p in pp])
    assert ei.value.name == "IncorrectShardSize"
    too_few = [i not in range(m + 1) for i in range(k + m)]
    with pytest.raises(rsgpu.RsError) as ei:
        rsgpu.leopard_reconstruct(bits, too_few, outputs, [inputs[i] if too_few[i] else None
                                                            for i in range(k + m)], k, m)
    assert ei.value.name == "TooFewShardsPresent"
    bad = list(outputs)
    bad[missing[0]] = torch.zeros(n - 64, dtype=torch.uint8, device="cuda")
    with pytest.raises(rsgpu.RsError) as ei:
        rsgpu.leopard_reconstruct(bits, present, bad, inputs, k, m)
    assert ei.value.name == "IncorrectShardSize"
