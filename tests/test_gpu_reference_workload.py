"""The reference's CPU-runnable workload on the GPU path, with the reference's
own input generators: 64 stripes of RS 10+4 x 64 KiB, data from SmallRng seeded
with derived_seed(op, case) ^ stripe (benches/common/mod.rs:10,224-250, the
stream the reference's seeded golden checksum pins), erasures from the SplitMix
partial Fisher-Yates (src/tests/mod.rs:67-78).  Encode (plain and batched),
verify and reconstruct with 4 erasures, against the oracle."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import announce, to_host

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
rsgpu = pytest.importorskip("rsgpu")

BASE_SEED = 0x00EC_5EED_2026_0524
OP_TAG = {"encode": 0x11, "verify": 0x22, "reconstruct": 0x33}
K, M, S, STRIPES = 10, 4, 64 << 10, 64


def derived_seed(op, k, m, size):
    return BASE_SEED ^ OP_TAG[op] ^ (k << 48) ^ (m << 32) ^ size


def data_shards(seed):
    raw = np.frombuffer(bytes(O.smallrng_bytes(seed, K * S)), np.uint8)
    return [raw[i * S:(i + 1) * S].copy() for i in range(K)]


def test_reference_workload_10x4_64k():
    announce("reference workload 64 x 10+4 x 64 KiB")
    rc, mat = O.build_matrix(O.GF8, O.VANDERMONDE, K, M)
    assert rc == 0
    rs = rsgpu.ReedSolomon(K, M)
    stripes = []
    for s in range(STRIPES):
        data = data_shards(derived_seed("encode", K, M, S) ^ s)
        stripes.append((data, O.encode(O.GF8, mat, K, M, data)))
    # plain encode, one launch per stripe
    dev = torch.zeros((STRIPES, K + M, S), dtype=torch.uint8, device="cuda")
    for s, (data, _) in enumerate(stripes):
        dev[s, :K] = torch.from_numpy(np.stack(data)).cuda()
        rs.encode(list(dev[s]))
    torch.cuda.synchronize()
    for s, (_, par) in enumerate(stripes):
        assert (dev[s, K:].cpu().numpy() == np.stack(par)).all(), s
    # batched encode of all 64 stripes in one launch
    dev[:, K:] = 0
    batch = rsgpu.Batch()
    for s in range(STRIPES):
        batch.add_encode(rs, list(dev[s]))
    batch.run()
    torch.cuda.synchronize()
    for s, (_, par) in enumerate(stripes):
        assert (dev[s, K:].cpu().numpy() == np.stack(par)).all(), s
    # verify: every stripe, then one corrupted byte
    assert all(rs.verify(list(dev[s])) for s in range(STRIPES))
    dev[7, K + 2, S // 3] ^= 0x40
    assert not rs.verify(list(dev[7]))
    dev[7, K + 2, S // 3] ^= 0x40
    # reconstruct with 4 erasures per stripe (data and parity mixed)
    for s in range(STRIPES):
        data = data_shards(derived_seed("reconstruct", K, M, S) ^ s)
        full = data + list(O.encode(O.GF8, mat, K, M, data))
        lost = O.corrupt_positions(4, K + M, O.qc_seed([K, M, s]))
        work = [None if i in lost else torch.from_numpy(full[i]).cuda() for i in range(K + M)]
        rs.reconstruct(work)
        for i in lost:
            assert (to_host(work[i]) == full[i]).all(), (s, i, lost)
