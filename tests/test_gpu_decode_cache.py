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


def test_host_threads_share_codecs():
    """Four host threads call one classic and one Leopard GF(2^8) codec at once, each
    on its own stream: decode-cache misses, hits and evictions (capacity 16), slot
    recycling, and Leopard decoder matrices and plans built concurrently.  Every
    rebuilt shard equals the original."""
    import threading

    announce("decode caches under concurrent host threads")
    k, m, n = 10, 4, 1 << 16
    full = _stripe(k, m, n, 7)
    rs = rsgpu.ReedSolomon(k, m, inversion_cache_capacity=16)
    rng0 = np.random.default_rng(8)
    ldata = [rng0.integers(0, 256, n, dtype=np.uint8) for _ in range(k)]
    rc, lpar = O.leo_encode(8, k, m, ldata)
    assert rc == 0
    lfull = np.stack(ldata + list(lpar))
    leo = rsgpu.ReedSolomon(k, m, codec_family="leopard_gf8")
    leo.set_variant(5)  # matrix variant: decoder matrices for fewer than m erasures
    dev, ldev = torch.from_numpy(full).cuda(), torch.from_numpy(lfull).cuda()
    torch.cuda.synchronize()
    patterns = [p for e in (1, 2) for p in itertools.combinations(range(k + m), e)]
    errors = []

    def worker(tid):
        try:
            s = torch.cuda.Stream()
            rng = np.random.default_rng(100 + tid)
            for it in range(40):
                missing = patterns[int(rng.integers(len(patterns)))]
                codec, src, ref = (rs, dev, full) if (it + tid) % 2 == 0 else (leo, ldev, lfull)
                work = [None if i in missing else src[i] for i in range(k + m)]
                codec.reconstruct(work, stream=s)
                s.synchronize()
                for i in missing:
                    if not (work[i].cpu().numpy() == ref[i]).all():
                        errors.append((tid, it, missing, i))
        except Exception as e:  # reported below
            errors.append((tid, repr(e)))

    ts = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors[:5]
    assert rs.cache_stats()["entries"] <= 16


def test_leopard_gf16_plans_under_concurrent_host_threads():
    """Four host threads share one Leopard GF(2^16) codec, each on its own stream.
    They ask for 96 erasure patterns (past the 64-plan cache, so least recently
    used plans are evicted while others are in use), with every thread starting
    on the same pattern (one build in flight, the others wait for it).  Every
    rebuilt shard equals the original."""
    import threading

    announce("leopard gf16 decode plans under concurrent host threads")
    k, m, n = 20, 6, 4096
    rng0 = np.random.default_rng(16)
    data = [rng0.integers(0, 256, n, dtype=np.uint8) for _ in range(k)]
    rc, par = O.leo_encode(16, k, m, data)
    assert rc == 0
    full = np.stack(data + list(par))
    leo = rsgpu.ReedSolomon(k, m, codec_family="leopard_gf16")
    dev = torch.from_numpy(full).cuda()
    torch.cuda.synchronize()
    prng = np.random.default_rng(17)
    patterns = []
    while len(patterns) < 96:
        p = tuple(sorted(prng.choice(k + m, int(prng.integers(1, m + 1)), replace=False).tolist()))
        if p not in patterns:
            patterns.append(p)
    errors = []
    start = threading.Barrier(4)

    def worker(tid):
        try:
            s = torch.cuda.Stream()
            rng = np.random.default_rng(200 + tid)
            start.wait()
            for it in range(60):
                missing = patterns[0] if it == 0 else patterns[int(rng.integers(len(patterns)))]
                slots = [rsgpu.ShardSlot(torch.full((n,), 0xA5, dtype=torch.uint8, device="cuda")
                                         if i in missing else dev[i], present=i not in missing)
                         for i in range(k + m)]
                s.wait_stream(torch.cuda.current_stream())
                leo.reconstruct(slots, stream=s)
                s.synchronize()
                for i in missing:
                    if not (slots[i].data.cpu().numpy() == full[i]).all():
                        errors.append((tid, it, missing, i))
        except Exception as e:  # reported below
            errors.append((tid, repr(e)))

    ts = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors[:5]


@pytest.mark.parametrize("field", ["gf8", "gf16"])
def test_host_solved_entry_serves_device_table_callers(field):
    """A pattern first reconstructed by an aligned call (decode rows solved on the
    host, tables passed in the kernel arguments) is then used by callers that need
    the tables in device memory: a batch, a call with an unaligned length (byte
    tail), and a prepared workspace on unaligned views.  Every result equals the
    oracle, and the cache counts one miss for the pattern."""
    announce(f"host-solved entry reused by device-table callers {field}")
    f = O.GF8 if field == "gf8" else O.GF16
    k, m = (10, 4) if field == "gf8" else (4, 2)
    n = 1 << 16
    rng = np.random.default_rng(17)
    data = [rng.integers(0, 256, n + 64, dtype=np.uint8) for _ in range(k)]
    _, mat = O.build_matrix(f, O.VANDERMONDE, k, m)
    full = np.stack(data + list(O.encode(f, mat, k, m, data)))
    rs = rsgpu.ReedSolomon(k, m, field=field)
    missing = (1, k + 1)
    dev = torch.from_numpy(full).cuda()

    # 1. aligned plain call: a miss solved on the host
    work = [None if i in missing else dev[i][:n] for i in range(k + m)]
    rs.reconstruct(work)
    torch.cuda.synchronize()
    for i in missing:
        assert (work[i].cpu().numpy() == full[i][:n]).all(), ("plain", i)
    assert rs.cache_stats()["misses"] == 1
    # 2. the same pattern in a batch: device copy of the host tables
    slots = [rsgpu.ShardSlot(dev[i][:n].clone() if i not in missing else
                             torch.full((n,), 0xA5, dtype=torch.uint8, device="cuda"),
                             present=i not in missing) for i in range(k + m)]
    b = rsgpu.Batch()
    b.add_reconstruct(rs, slots)
    b.run()
    torch.cuda.synchronize()
    for i in missing:
        assert (slots[i].data.cpu().numpy() == full[i][:n]).all(), ("batch", i)
    # 3. an unaligned length: the byte-tail kernel reads device tables
    odd = n + (2 if field == "gf16" else 3)
    work = [None if i in missing else dev[i][:odd] for i in range(k + m)]
    rs.reconstruct(work)
    torch.cuda.synchronize()
    for i in missing:
        assert (work[i].cpu().numpy() == full[i][:odd]).all(), ("unaligned", i)
    st = rs.cache_stats()
    assert st["misses"] == 1 and st["hits"] == 2
    # 4. a workspace prepared from the host-solved pattern, used on views at an
    #    odd offset (not 16-byte aligned): device tables on demand
    layout = [None if i in missing else dev[i][:n] for i in range(k + m)]
    ws = rs.prepare_reconstruct_workspace(layout)
    off = 2 if field == "gf16" else 1
    views = [rsgpu.ShardSlot(dev[i][off:off + n].clone() if i not in missing else
                             torch.zeros(n + 16, dtype=torch.uint8, device="cuda")[off:off + n],
                             present=i not in missing) for i in range(k + m)]
    shifted = np.stack([fr[off:off + n] for fr in full])
    # the shifted columns are not a codeword: compare with the oracle's reconstruct
    present = [0 if i in missing else 1 for i in range(k + m)]
    rc, exp = O.reconstruct(f, mat, k, m, [x.copy() for x in shifted], present)
    assert rc == 0
    rs.reconstruct_with_workspace(views, ws)
    torch.cuda.synchronize()
    for i in missing:
        assert (views[i].data.cpu().numpy() == exp[i]).all(), ("workspace unaligned", i)
