"""Decode-row cache on the GPU (src/core/inversion_cache.rs semantics, device
resident): entries solved on one stream and read on another, and entry slots
recycled after eviction, all bit-exact against the oracle."""
import itertools

import numpy as np
import pytest

import oracle_lib as O
from gpu_util import announce

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
rsgpu = pytest.importorskip("rsgpu")


def _stripe(k, m, n, seed):
    rng = np.random.default_rng(seed)
    data = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k)]
    rc, mat = O.build_matrix(O.GF8, O.VANDERMONDE, k, m)
    assert rc == 0
    return np.stack(data + list(O.encode(O.GF8, mat, k, m, data)))


def test_hit_on_other_stream_waits_for_the_solve():
    """The solve of a new pattern is queued on stream A behind a long encode; the
    same pattern is then reconstructed on stream B at once."""
    announce("decode cache cross-stream hit")
    k, m, n = 10, 4, 1 << 16
    full = _stripe(k, m, n, 1)
    rs = rsgpu.ReedSolomon(k, m)
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    big = [torch.zeros(64 << 20, dtype=torch.uint8, device="cuda") for _ in range(k + m)]
    dev = torch.from_numpy(full).cuda()
    torch.cuda.synchronize()
    for trial, missing in enumerate([(0, 3, 11), (1, 2, 12, 13), (4,)]):
        work_a = [None if i in missing else dev[i] for i in range(k + m)]
        work_b = [None if i in missing else dev[i] for i in range(k + m)]
        for _ in range(4):
            rs.encode(big, stream=sa)  # keeps stream A busy for milliseconds
        rs.reconstruct(work_a, stream=sa)  # miss: solve queued behind the encodes
        rs.reconstruct(work_b, stream=sb)  # hit on another stream
        torch.cuda.synchronize()
        for i in missing:
            assert (work_a[i].cpu().numpy() == full[i]).all(), (trial, i)
            assert (work_b[i].cpu().numpy() == full[i]).all(), (trial, i)
    st = rs.cache_stats()
    assert st["misses"] == 3 and st["hits"] == 3


def test_slots_recycle_after_eviction():
    """A 4-entry cache over every 1- and 2-erasure pattern of 8+3 (several times
    the recycle batch): evicted slots retire and come back after a sync."""
    announce("decode cache slot recycling")
    k, m, n = 8, 3, 4096
    full = _stripe(k, m, n, 2)
    rs = rsgpu.ReedSolomon(k, m, inversion_cache_capacity=4)
    dev = torch.from_numpy(full).cuda()
    patterns = [p for e in (1, 2) for p in itertools.combinations(range(k + m), e)]
    batch_outs = []
    for rnd in range(3):
        for missing in patterns:
            work = [None if i in missing else dev[i] for i in range(k + m)]
            rs.reconstruct(work)
            for i in missing:
                batch_outs.append((work[i], i))
        torch.cuda.synchronize()
        for t, i in batch_outs:
            assert (t.cpu().numpy() == full[i]).all(), (rnd, i)
        batch_outs.clear()
    st = rs.cache_stats()
    assert st["entries"] <= 4
    assert st["misses"] >= 3 * len(patterns) - 4
