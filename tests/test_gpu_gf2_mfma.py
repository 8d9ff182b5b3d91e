"""GPU parity of the int8 matrix-core encode (GF(2) bit-matrix product, exact via
the parity of int32 sums) against the oracle's encode (src/core/encode.rs:935-998
semantics), and its argument limits."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import announce, rand_shards, to_dev, to_host

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
rsgpu = pytest.importorskip("rsgpu")


@pytest.mark.parametrize("k,m,n", [(10, 4, 1 << 20), (10, 4, 4), (10, 4, 132), (1, 1, 512),
                                   (3, 2, 4096 + 4), (8, 3, 65536), (17, 5, 12345 * 4),
                                   (24, 8, 1 << 16), (31, 8, 260), (31, 1, 1024), (4, 7, 2048)])
@pytest.mark.parametrize("matrix_mode", ["vandermonde", "cauchy"])
def test_matches_oracle(k, m, n, matrix_mode):
    announce(f"gf2 mfma encode {k}+{m} n={n} {matrix_mode}")
    rng = np.random.default_rng(k * 131 + m * 7 + n)
    data = rand_shards(rng, k, n)
    mode = O.VANDERMONDE if matrix_mode == "vandermonde" else O.CAUCHY
    rc, mat = O.build_matrix(O.GF8, mode, k, m)
    assert rc == 0
    exp = O.encode(O.GF8, mat, k, m, data)
    rs = rsgpu.ReedSolomon(k, m, matrix_mode=matrix_mode)
    par = [torch.full((n,), 0x5A, dtype=torch.uint8, device="cuda") for _ in range(m)]
    rs.encode_sep_gf2mfma([to_dev(d) for d in data], par)
    for j in range(m):
        assert (to_host(par[j]) == exp[j]).all(), j


def test_fast_one_parity_is_xor():
    announce("gf2 mfma fast one parity")
    k, n = 7, 4096
    rng = np.random.default_rng(9)
    data = rand_shards(rng, k, n)
    rs = rsgpu.ReedSolomon(k, 1, fast_one_parity=True)
    par = [torch.zeros(n, dtype=torch.uint8, device="cuda")]
    rs.encode_sep_gf2mfma([to_dev(d) for d in data], par)
    assert (to_host(par[0]) == np.bitwise_xor.reduce(np.stack(data), axis=0)).all()


def test_equals_vector_path_on_large_stripe():
    announce("gf2 mfma vs vector path 12+4 x 8 MiB")
    k, m, n = 12, 4, 8 << 20
    g = torch.Generator(device="cuda").manual_seed(5)
    data = [torch.randint(0, 256, (n,), dtype=torch.uint8, device="cuda", generator=g) for _ in range(k)]
    rs = rsgpu.ReedSolomon(k, m)
    a = [torch.zeros(n, dtype=torch.uint8, device="cuda") for _ in range(m)]
    b = [torch.ones(n, dtype=torch.uint8, device="cuda") for _ in range(m)]
    rs.encode_sep(data, a)
    rs.encode_sep_gf2mfma(data, b)
    torch.cuda.synchronize()
    for j in range(m):
        assert torch.equal(a[j], b[j]), j


def test_limits():
    announce("gf2 mfma limits")

    def err(rs, k, m, n, off=0):
        buf = torch.zeros(n + off, dtype=torch.uint8, device="cuda")
        data = [buf[off:off + n]] * k
        par = [torch.zeros(n, dtype=torch.uint8, device="cuda") for _ in range(m)]
        with pytest.raises(rsgpu.RsError) as ei:
            rs.encode_sep_gf2mfma(data, par)
        return ei.value.name

    assert err(rsgpu.ReedSolomon(32, 4), 32, 4, 64) == "InvalidArgument"
    assert err(rsgpu.ReedSolomon(10, 9), 10, 9, 64) == "InvalidArgument"
    assert err(rsgpu.ReedSolomon(10, 4), 10, 4, 66) == "InvalidArgument"
    assert err(rsgpu.ReedSolomon(10, 4), 10, 4, 64, off=2) == "InvalidArgument"
    assert err(rsgpu.ReedSolomon(10, 4, field="gf16"), 10, 4, 64) == "UnsupportedCodecFamily"
    assert err(rsgpu.ReedSolomon(10, 4, codec_family="leopard_gf8"), 10, 4, 64) == "UnsupportedCodecFamily"
    assert err(rsgpu.ReedSolomon(10, 4), 9, 4, 64) == "TooFewDataShards"
    assert err(rsgpu.ReedSolomon(10, 4), 10, 5, 64) == "TooManyParityShards"
