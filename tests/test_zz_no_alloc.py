"""rsgpu_reconstruct_noalloc (the entry point that allocates nothing) against
rsgpu_reconstruct / rsgpu_reconstruct_data on the same inputs: bit-exact on two
calls in a row, the second with a smaller input, then the size limits."""
import numpy as np
import pytest

from gpu_util import announce

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
rsgpu = pytest.importorskip("rsgpu")

POISON = 0xA5


def _stripe(rs, k, m, length, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    full = torch.empty((k + m, length), dtype=torch.uint8, device="cuda")
    full[:k].random_(0, 256, generator=g)
    rs.encode(list(full))
    torch.cuda.synchronize()
    return full


def _run(rs, full, missing, noalloc, data_only):
    slots = []
    for i in range(full.shape[0]):
        if i in missing:
            slots.append(rsgpu.ShardSlot(torch.full((full.shape[1],), POISON, dtype=torch.uint8,
                                                    device="cuda"), present=False))
        else:
            slots.append(rsgpu.ShardSlot(full[i], present=True))
    if noalloc:
        rs.reconstruct_noalloc(slots, data_only=data_only)
    elif data_only:
        rs.reconstruct_data(slots)
    else:
        rs.reconstruct(slots)
    torch.cuda.synchronize()
    return slots


def _compare(rs, full, missing, data_only):
    a = _run(rs, full, missing, True, data_only)
    b = _run(rs, full, missing, False, data_only)
    for i in range(full.shape[0]):
        assert a[i].present == b[i].present, i
        assert torch.equal(a[i].data, b[i].data), (i, missing)
        if i in missing and a[i].present:
            assert torch.equal(a[i].data, full[i]), i


@pytest.mark.parametrize("field", ["gf8", "gf16"])
def test_noalloc_matches_reconstruct_two_calls_second_smaller(field):
    rng = np.random.default_rng(11)
    calls = [(10, 4, (1 << 20) + 48), (10, 4, 4096 + 16), (6, 3, 640)]
    if field == "gf16":
        calls = [(20, 8, (256 << 10) + 2), (20, 8, 4096), (5, 2, 66)]
    for k, m, length in calls:
        announce(f"noalloc {field} {k}+{m} len={length}")
        rs = rsgpu.ReedSolomon(k, m, field=field)
        full = _stripe(rs, k, m, length, k * 1000 + length)
        for data_only in (False, True):
            e = int(rng.integers(1, m + 1))
            missing = set(rng.choice(k + m, e, replace=False).tolist())
            _compare(rs, full, missing, data_only)


def test_noalloc_widest_gf8_codec_and_parity_only():
    announce("noalloc gf8 128+128 and parity-only patterns")
    k, m = 128, 128
    rs = rsgpu.ReedSolomon(k, m)
    full = _stripe(rs, k, m, 4096, 5)
    rng = np.random.default_rng(3)
    _compare(rs, full, set(rng.choice(k + m, m, replace=False).tolist()), False)
    _compare(rs, full, set(range(k, k + m, 3)), False)
    rs2 = rsgpu.ReedSolomon(10, 4)
    full2 = _stripe(rs2, 10, 4, 1024, 6)
    _compare(rs2, full2, {0, 11}, False)


def test_noalloc_limits_and_errors():
    announce("noalloc limits")
    # GF(2^16) 128+128 with 128 targets needs 2 MiB of tables: beyond the scratch
    rs = rsgpu.ReedSolomon(128, 128, field="gf16")
    full = torch.zeros((256, 64), dtype=torch.uint8, device="cuda")
    slots = [rsgpu.ShardSlot(full[i], present=i >= 128) for i in range(256)]
    with pytest.raises(rsgpu.RsError) as ei:
        rs.reconstruct_noalloc(slots)
    assert ei.value.name == "Unsupported"
    assert all(not s.present for s in slots[:128])
    big = rsgpu.ReedSolomon(300, 20, field="gf16")
    with pytest.raises(rsgpu.RsError) as ei:
        big.reconstruct_noalloc([rsgpu.ShardSlot(torch.zeros(64, dtype=torch.uint8, device="cuda"))
                                 for _ in range(320)])
    assert ei.value.name == "Unsupported"
    leo = rsgpu.ReedSolomon(10, 4, codec_family="leopard_gf8")
    with pytest.raises(rsgpu.RsError) as ei:
        leo.reconstruct_noalloc([rsgpu.ShardSlot(torch.zeros(64, dtype=torch.uint8, device="cuda"))
                                 for _ in range(14)])
    assert ei.value.name == "UnsupportedCodecFamily"
    # validation as the allocating entry point
    rs = rsgpu.ReedSolomon(4, 2)
    t = [torch.zeros(64, dtype=torch.uint8, device="cuda") for _ in range(6)]
    with pytest.raises(rsgpu.RsError) as ei:
        rs.reconstruct_noalloc([rsgpu.ShardSlot(x, present=i > 2) for i, x in enumerate(t)])
    assert ei.value.name == "TooFewShardsPresent"
    with pytest.raises(rsgpu.RsError) as ei:
        rs.reconstruct_noalloc([rsgpu.ShardSlot(x) for x in t[:5]])
    assert ei.value.name == "TooFewShards"
