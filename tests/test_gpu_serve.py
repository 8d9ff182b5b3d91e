"""The request server: every result of a request served in an overlapping mix
equals the result of the same request run alone through the plain entry points
(bit-exact bytes and verify verdicts); admission keeps the in-flight footprint
under the HBM budget; validation errors stay with their own request."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import announce

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
rsgpu = pytest.importorskip("rsgpu")
from rsgpu import serve as SV  # noqa: E402


class _View:
    """A device byte range as a torch tensor (no copy)."""

    def __init__(self, ptr, n):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": "|u1", "data": (ptr, False),
                                         "version": 2}


def view(ptr, n):
    return torch.as_tensor(_View(ptr, n), device="cuda")


def _alone(store, meta_i):
    """The request run alone through the plain API: (outputs, verify verdict)."""
    (k, m), S, op, erased, stripe = meta_i
    rs = store.codecs[(k, m)]
    ptrs = store.shard_ptrs((k, m), stripe)
    if op == 0:
        sh = [view(p, S).clone() for p in ptrs[:k]] + [torch.zeros(S, dtype=torch.uint8, device="cuda")
                                                       for _ in range(m)]
        rs.encode(sh)
        return [x.cpu().numpy() for x in sh[k:]], None
    if op == 1:
        return [], rs.verify([view(p, S).clone() for p in ptrs])
    work = [None if i in erased else view(p, S).clone() for i, p in enumerate(ptrs)]
    rs.reconstruct(work)
    return [work[i].cpu().numpy() for i in erased], None


def _check_results(server, store, comps, meta):
    for c, mi in zip(comps, meta):
        assert c.status == 0, (c.tag, c.status)
        (k, m), S, op, erased, stripe = mi
        exp, ok = _alone(store, mi)
        if op == 1:
            assert bool(c.verify_ok) == ok and ok
            continue
        outs = server.outputs(c.tag, len(exp))
        for j, (p, e) in enumerate(zip(outs, exp)):
            got = view(p, S).cpu().numpy()
            assert np.array_equal(got, e), (c.tag, op, j)
        # the store holds codewords: rebuilt shards equal the stored ones
        if op == 2:
            ptrs = store.shard_ptrs((k, m), stripe)
            for j, i in enumerate(erased):
                assert np.array_equal(exp[j], view(ptrs[i], S).cpu().numpy())


def test_results_identical_alone_and_in_mix():
    announce("serve mix vs alone")
    store = SV.Store(torch, per_shape=3, max_shard=256 << 10)
    reqs, keep, meta = SV.poisson_requests(store, 240, rate=400000.0, seed=3, lo=4 << 10,
                                           hi=2 << 20)
    srv = SV.Server(hbm_budget=8 << 30, arena_bytes=1 << 30, max_inflight_batches=4, max_batch_ops=16)
    comps = srv.replay(reqs, keep_outputs=True)
    st = srv.stats()
    print("stats", st, SV.latency_summary(comps, meta), flush=True)
    assert st["completed"] == 240 and st["batches"] >= 15
    assert st["max_inflight_batches"] >= 2  # batches overlapped on the device
    _check_results(srv, store, comps, meta)
    for c in comps:
        srv.release(c.tag)
    assert srv.stats()["footprint"] == 0
    srv.close()


def test_admission_keeps_footprint_under_budget():
    announce("serve admission")
    store = SV.Store(torch, per_shape=2, max_shard=512 << 10)
    reqs, keep, meta = SV.poisson_requests(store, 120, rate=1e7, seed=9, lo=1 << 20, hi=8 << 20)
    biggest = max(c.shard_len * 20 for c in reqs)
    budget = 3 * biggest
    srv = SV.Server(hbm_budget=budget, arena_bytes=1 << 30, max_inflight_batches=3, max_batch_ops=64)
    comps = srv.replay(reqs, keep_outputs=False)
    st = srv.stats()
    print("stats", st, flush=True)
    assert st["completed"] == 120
    assert st["peak_footprint"] <= budget
    assert st["footprint"] == 0
    # requests waited for memory: admissions are spread out, in arrival order
    admits = [c.admit_ns for c in comps]
    assert admits == sorted(admits)
    assert max(c.admit_ns - c.arrival_ns for c in comps) > 0
    srv.close()


def test_validation_error_stays_with_its_request():
    announce("serve validation error")
    store = SV.Store(torch, per_shape=1, max_shard=64 << 10)
    reqs, keep, meta = SV.poisson_requests(store, 12, rate=1e6, seed=1, lo=16 << 10, hi=256 << 10)
    bad = 5
    reqs[bad].op = 0
    reqs[bad].erased = None
    reqs[bad].shard_len = 0  # encode of empty shards: EmptyShard (macros.rs check_slices!)
    srv = SV.Server(hbm_budget=1 << 30, arena_bytes=256 << 20)
    comps = srv.replay(reqs, keep_outputs=True)
    for i, c in enumerate(comps):
        if i == bad:
            assert rsgpu.ERROR_NAMES[c.status] == "EmptyShard"
    good = [i for i in range(12) if i != bad]
    _check_results(srv, store, [comps[i] for i in good], [meta[i] for i in good])
    srv.close()


def test_replay_returns_when_kept_outputs_fill_the_arena():
    """Kept outputs are never released during a replay: once they fill the arena, a
    request that cannot be admitted while nothing is in flight completes with
    OutOfMemory instead of blocking the replay forever.  Admitted requests keep
    their exact results."""
    announce("serve kept outputs exceed the arena")
    store = SV.Store(torch, per_shape=1, max_shard=64 << 10)
    reqs, keep, meta = SV.poisson_requests(store, 16, rate=1e6, seed=4, lo=512 << 10, hi=640 << 10,
                                           mix=(1.0, 0.0, 0.0))
    srv = SV.Server(hbm_budget=1 << 30, arena_bytes=1 << 20, max_inflight_batches=2, max_batch_ops=4)
    comps = srv.replay(reqs, keep_outputs=True)
    names = [rsgpu.ERROR_NAMES[c.status] if c.status < 100 else c.status for c in comps]
    print("statuses", names, flush=True)
    good = [i for i, c in enumerate(comps) if c.status == 0]
    oom = [i for i, c in enumerate(comps) if c.status == 104]  # RSGPU_ERR_OUT_OF_MEMORY
    assert len(good) + len(oom) == 16 and oom and good
    _check_results(srv, store, [comps[i] for i in good], [meta[i] for i in good])
    for c in comps:
        srv.release(c.tag)
    assert srv.stats()["footprint"] == 0
    srv.close()
