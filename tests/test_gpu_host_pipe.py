"""GPU parity of the host-pointer entry points (shards in caller memory, as the
reference takes them): the chunked pipeline over shard columns must give the
same bytes and verdicts as the oracle at every length -- one chunk, an exact
chunk, ragged tails, many chunks -- and reject bad lengths before writing."""
import threading

import numpy as np
import pytest

import oracle_lib as O
from gpu_util import announce

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
rsgpu = pytest.importorskip("rsgpu")

CHUNK = 512 << 10


def _encoded(k, m, n, seed):
    rng = np.random.default_rng(seed)
    data = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k)]
    rc, mat = O.build_matrix(O.GF8, O.VANDERMONDE, k, m)
    assert rc == 0
    return data + list(O.encode(O.GF8, mat, k, m, data))


@pytest.mark.parametrize("n", [64, 1000, CHUNK, CHUNK + 1, 3 * CHUNK + 17, 5 << 20])
def test_encode_verify_reconstruct(n):
    announce(f"host pipe 10+4 n={n}")
    k, m = 10, 4
    full = _encoded(k, m, n, n)
    rs = rsgpu.ReedSolomon(k, m)
    host = [d.copy() for d in full[:k]] + [np.full(n, 0xA5, np.uint8) for _ in range(m)]
    rs.encode_host(host)
    for j in range(m):
        assert (host[k + j] == full[k + j]).all(), j
    assert rs.verify_host(host)
    for pos in (0, n - 1):  # first and last chunk
        bad = [h.copy() for h in host]
        bad[k + 1][pos] ^= 0x10
        assert not rs.verify_host(bad), pos
    present = np.ones(k + m, np.uint8)
    present[[1, 7, 11, 13]] = 0
    damaged = [h.copy() if p else np.zeros(n, np.uint8) for h, p in zip(full, present)]
    arrs, pres = rs.reconstruct_host(damaged, present)
    assert pres.all()
    assert all((a == f).all() for a, f in zip(arrs, full))
    damaged = [h.copy() if p else np.zeros(n, np.uint8) for h, p in zip(full, present)]
    arrs, pres = rs.reconstruct_host(damaged, present, data_only=True)
    assert pres[:k].all() and not pres[11] and not pres[13]
    assert all((arrs[i] == full[i]).all() for i in range(k))


def test_leopard_and_tower_lengths():
    announce("host pipe leopard / gf16")
    k, m, n = 12, 4, CHUNK * 2 + 640
    rng = np.random.default_rng(3)
    data = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k)]
    leo = rsgpu.ReedSolomon(k, m, codec_family="leopard_gf8")
    host = [d.copy() for d in data] + [np.zeros(n, np.uint8) for _ in range(m)]
    leo.encode_host(host)
    rc, exp = O.leo_encode(8, k, m, data)
    assert rc == 0
    assert all((host[k + j] == exp[j]).all() for j in range(m))
    assert leo.verify_host(host)
    bad = [d.copy() for d in data] + [np.full(n + 0, 7, np.uint8) for _ in range(m)]
    bad = [b[:n - 32].copy() for b in bad]  # not a multiple of 64
    with pytest.raises(rsgpu.RsError) as ei:
        leo.encode_host(bad)
    assert ei.value.name == "IncorrectShardSize"
    assert all((b == 7).all() for b in bad[k:])  # nothing written
    tower = rsgpu.ReedSolomon(5, 3, field="gf16")
    odd = [np.zeros(CHUNK + 3, np.uint8) for _ in range(8)]
    with pytest.raises(rsgpu.RsError) as ei:
        tower.encode_host(odd)
    assert ei.value.name == "IncorrectShardSize"


def test_concurrent_callers():
    """Two threads (each with its own staging) share the copy workers."""
    announce("host pipe two threads")
    results = {}

    def worker(tid, k, m, n):
        full = _encoded(k, m, n, 100 + tid)
        rs = rsgpu.ReedSolomon(k, m)
        ok = True
        for _ in range(3):
            host = [d.copy() for d in full[:k]] + [np.zeros(n, np.uint8) for _ in range(m)]
            rs.encode_host(host)
            ok &= all((host[k + j] == full[k + j]).all() for j in range(m))
        results[tid] = ok

    ts = [threading.Thread(target=worker, args=(0, 10, 4, 3 * CHUNK + 5)),
          threading.Thread(target=worker, args=(1, 6, 3, 2 * CHUNK + 64))]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert results == {0: True, 1: True}
