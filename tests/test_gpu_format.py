"""Format identity and split / join on the GPU path: the reference's literal
version-1 identifier fixtures (tests/encoding_format_identity.rs:8-27), identity
rules between matrix modes and families, and split / join round trips with the
classic and Leopard padding rules (src/core/codec.rs:476-538, leopard.rs:75-98)."""
import numpy as np
import pytest

from fixtures import golden
from gpu_util import announce, to_dev, to_host

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
rsgpu = pytest.importorskip("rsgpu")

# version-1 identifiers of the 1+1 custom matrices [0x87] (GF(2^8)) and [[0x12, 0x34]]
# (tower GF(2^16)), tests/encoding_format_identity.rs:10-27
_IDS = golden("format_identity_v1")
GF8_ID_V1 = bytes.fromhex(_IDS["gf8_custom_1x1_0x87"])
TOWER_GF16_ID_V1 = bytes.fromhex(_IDS["tower_gf16_custom_1x1_0x1234"])


def test_stable_ids_match_reference_fixtures():
    announce("format id fixtures")
    gf8 = rsgpu.ReedSolomon.with_custom_matrix(1, 1, [[0x87]])
    tower = rsgpu.ReedSolomon.with_custom_matrix(1, 1, [[0x12, 0x34]], field="gf16")
    assert gf8.stable_id_bytes() == GF8_ID_V1
    assert tower.stable_id_bytes() == TOWER_GF16_ID_V1


def test_identity_follows_effective_coefficients_and_family():
    announce("format id rules")
    ids = set()
    for mode in ("vandermonde", "cauchy", "jerasure_like"):
        rs = rsgpu.ReedSolomon(4, 2, matrix_mode=mode)
        rows = rs.matrix()[4:]
        same = rsgpu.ReedSolomon.with_custom_matrix(4, 2, rows)
        assert rs.stable_id_bytes() == same.stable_id_bytes()
        ids.add(rs.stable_id_bytes())
    assert len(ids) == 3
    leo8 = rsgpu.ReedSolomon(4, 2, codec_family="leopard_gf8").stable_id_bytes()
    leo16 = rsgpu.ReedSolomon(4, 2, codec_family="leopard_gf16").stable_id_bytes()
    assert leo8[10] == 1 and leo16[10] == 2 and leo8 != leo16
    assert leo8[-16:] == bytes(16)  # 0 x 0 matrix
    assert rsgpu.ReedSolomon(4, 2, leopard_mode="prefer_leopard").stable_id_bytes() == leo8


@pytest.mark.parametrize("family,k,n", [("classic", 4, 61), ("classic", 10, 1 << 20), ("classic", 3, 0),
                                        ("leopard_gf8", 4, 61), ("leopard_gf16", 10, 5000),
                                        ("classic", 7, 7 * 64)])
def test_split_join_round_trip(family, k, n):
    announce(f"split/join {family} k={k} n={n}")
    rs = rsgpu.ReedSolomon(k, 2, codec_family=family)
    rng = np.random.default_rng(n + k)
    data = rng.integers(0, 256, n, dtype=np.uint8)
    L = rs.split_len(n)
    if family == "classic":
        assert L == -(-n // k)
    else:
        assert L % 64 == 0 and L >= 64 and L * k >= n
    shards = rs.split(to_dev(data) if n else torch.empty(0, dtype=torch.uint8, device="cuda"))
    flat = np.concatenate([to_host(s) for s in shards]) if L else np.zeros(0, np.uint8)
    assert (flat[:n] == data).all() and not flat[n:].any()
    if L:
        par = [torch.zeros(L, dtype=torch.uint8, device="cuda") for _ in range(2)]
        rs.encode(shards + par)
        back = rs.join(shards, n)
        assert (to_host(back) == data).all()
        assert rs.join(shards, 10 * n + 5).numel() == L * k  # truncated to what the shards hold


def test_join_validation():
    announce("join validation")
    rs = rsgpu.ReedSolomon(3, 2)
    z = [torch.zeros(8, dtype=torch.uint8, device="cuda") for _ in range(3)]
    with pytest.raises(rsgpu.RsError) as ei:
        rs.join(z[:2], 10)
    assert ei.value.name == "TooFewDataShards"
    with pytest.raises(rsgpu.RsError) as ei:
        rs.join(z[:2] + [torch.zeros(7, dtype=torch.uint8, device="cuda")], 10)
    assert ei.value.name == "IncorrectShardSize"
