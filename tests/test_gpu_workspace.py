"""Prepared reconstruction (ReconstructWorkspace, src/core/reconstruct.rs:539-700):
the reference's doc examples, oracle parity over repeated use (codewords and
not), independence from the decode cache, and the validation order of
validate_reconstruct_workspace (reconstruct.rs:669-700)."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import announce, to_dev, to_host

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
rsgpu = pytest.importorskip("rsgpu")

POISON = 0xA5


def _err(fn):
    with pytest.raises(rsgpu.RsError) as ei:
        fn()
    return ei.value.name


def test_some_workspace_doc_example():
    """reconstruct.rs:566-580: a plan for shard 4 only; the mask can change afterwards."""
    announce("workspace doc example (some)")
    codec = rsgpu.ReedSolomon(4, 3)
    shards = [torch.full((65,), 2, dtype=torch.uint8, device="cuda") for _ in range(7)]
    codec.encode(shards)
    full = [s.clone() for s in shards]
    shards = list(shards)
    for i in (0, 1, 4):
        shards[i] = None
    required = [False] * 7
    required[4] = True
    plan = codec.prepare_reconstruct_some_workspace(shards, required)
    required = [False] * 7
    codec.reconstruct_with_workspace(shards, plan)
    torch.cuda.synchronize()
    assert shards[0] is None and shards[1] is None
    assert shards[4] is not None and torch.equal(shards[4], full[4])
    assert plan.parity_targets() == [4] and plan.data_targets() == []


def test_gf16_workspace_reused_doc_example():
    """reconstruct.rs:620-633: galois_16 4+2, shard 0 rebuilt three times from one plan."""
    announce("workspace doc example (gf16)")
    codec = rsgpu.ReedSolomon(4, 2, field="gf16")
    one = torch.tensor([3, 7] * 65, dtype=torch.uint8, device="cuda")
    shards = [one.clone() for _ in range(6)]
    codec.encode(shards)
    full = [s.clone() for s in shards]
    slots = [rsgpu.ShardSlot.new_present(s) for s in shards]
    slots[0].mark_missing()
    ws = codec.prepare_reconstruct_workspace(slots)
    for _ in range(3):
        slots[0].data.fill_(POISON)
        slots[0].mark_missing()
        codec.reconstruct_with_workspace_par(slots, ws)
        torch.cuda.synchronize()
        assert slots[0].present and torch.equal(slots[0].data, full[0])


@pytest.mark.parametrize("field,k,m,n", [("gf8", 10, 4, 1 << 20), ("gf8", 8, 3, 4096 + 16),
                                         ("gf8", 20, 4, 333), ("gf16", 12, 6, 2048 + 2)])
@pytest.mark.parametrize("goal", ["all", "data", "some"])
def test_workspace_matches_oracle_on_repeated_inputs(field, k, m, n, goal):
    """One plan per erasure pattern, used on three different inputs (two
    non-codewords): every requested target equals the oracle's reconstruct."""
    announce(f"workspace oracle {field} {k}+{m} n={n} goal={goal}")
    f = O.GF8 if field == "gf8" else O.GF16
    codec = rsgpu.ReedSolomon(k, m, field=field)
    _, mat = O.build_matrix(f, O.VANDERMONDE, k, m)
    rng = np.random.default_rng(k + m + n)
    total = k + m
    for trial in range(2):
        e = int(rng.integers(1, m + 1))
        missing = set(rng.choice(total, e, replace=False).tolist())
        required = [bool(rng.integers(0, 2)) for _ in range(total)] if goal == "some" else None
        layout = [None if i in missing else torch.empty(n, dtype=torch.uint8, device="cuda")
                  for i in range(total)]
        if goal == "all":
            ws = codec.prepare_reconstruct_workspace(layout)
        elif goal == "data":
            ws = codec.prepare_reconstruct_data_workspace(layout)
        else:
            ws = codec.prepare_reconstruct_some_workspace(layout, required)
        for use in range(3):
            host = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(total)]
            if use == 0:  # a codeword
                host[k:] = O.encode(f, mat, k, m, host[:k])
            present = [0 if i in missing else 1 for i in range(total)]
            rc, exp = O.reconstruct(f, mat, k, m, [h.copy() for h in host], present,
                                    data_only=(goal == "data"),
                                    required=None if required is None else [int(r) for r in required])
            assert rc == 0
            slots = [rsgpu.ShardSlot(to_dev(host[i]) if i not in missing else
                                     torch.full((n,), POISON, dtype=torch.uint8, device="cuda"),
                                     present=i not in missing) for i in range(total)]
            codec.reconstruct_with_workspace(slots, ws)
            torch.cuda.synchronize()
            targets = set(ws.data_targets()) | set(ws.parity_targets())
            for i in missing:
                if i in targets:
                    assert slots[i].present and (to_host(slots[i].data) == exp[i]).all(), (trial, use, i)
                else:
                    assert not slots[i].present and int(slots[i].data[0]) == POISON


def test_workspace_outlives_the_decode_cache():
    """With the inversion cache off, and after enough other patterns to evict an
    entry, a prepared plan still rebuilds its pattern."""
    announce("workspace without cache")
    for opts in ({"inversion_cache": False}, {"inversion_cache_capacity": 4}):
        k, m, n = 6, 3, 4096
        codec = rsgpu.ReedSolomon(k, m, **opts)
        rng = np.random.default_rng(9)
        data = [to_dev(rng.integers(0, 256, n, dtype=np.uint8)) for _ in range(k)]
        full = data + [torch.zeros(n, dtype=torch.uint8, device="cuda") for _ in range(m)]
        codec.encode(full)
        layout = [None if i in (1, 7) else full[i] for i in range(k + m)]
        ws = codec.prepare_reconstruct_workspace(layout)
        for a in range(k + m):  # other single-erasure patterns churn the cache
            work = [None if i == a else full[i].clone() for i in range(k + m)]
            codec.reconstruct(work)
        torch.cuda.synchronize()
        slots = [rsgpu.ShardSlot(full[i] if i not in (1, 7) else
                                 torch.full((n,), POISON, dtype=torch.uint8, device="cuda"),
                                 present=i not in (1, 7)) for i in range(k + m)]
        codec.reconstruct_with_workspace(slots, ws)
        torch.cuda.synchronize()
        assert torch.equal(slots[1].data, full[1]) and torch.equal(slots[7].data, full[7])


def test_validation_order_and_equivalent_codecs():
    announce("workspace validation")
    k, m, n = 5, 3, 256
    codec = rsgpu.ReedSolomon(k, m)
    full = [torch.randint(0, 256, (n,), dtype=torch.uint8, device="cuda") for _ in range(k)] + \
           [torch.zeros(n, dtype=torch.uint8, device="cuda") for _ in range(m)]
    codec.encode(full)
    layout = [None if i in (0, 6) else full[i] for i in range(k + m)]
    ws = codec.prepare_reconstruct_workspace(layout)
    assert ws.data_targets() == [0] and ws.parity_targets() == [6]

    def slots(missing, length=n):
        return [rsgpu.ShardSlot(torch.zeros(length, dtype=torch.uint8, device="cuda") if i in missing
                                else full[i][:length].clone(), present=i not in missing)
                for i in range(k + m)]

    assert _err(lambda: codec.reconstruct_with_workspace(slots({0, 6})[:-1], ws)) == "TooFewShards"
    assert _err(lambda: codec.reconstruct_with_workspace(slots({0, 6}) + slots({})[:1], ws)) == "TooManyShards"
    assert _err(lambda: codec.reconstruct_with_workspace(slots({0, 5}), ws)) == "InvalidShardFlags"
    assert _err(lambda: codec.reconstruct_with_workspace(slots({0, 6}, 128), ws)) == "IncorrectShardSize"
    wider = rsgpu.ReedSolomon(6, 3)
    assert _err(lambda: wider.reconstruct_with_workspace(slots({0, 6}) + slots({})[:1], ws)) == \
        "TooFewDataShards"
    narrower = rsgpu.ReedSolomon(4, 3)
    assert _err(lambda: narrower.reconstruct_with_workspace(slots({0, 6})[1:], ws)) == "TooManyDataShards"
    more_parity = rsgpu.ReedSolomon(5, 4)
    assert _err(lambda: more_parity.reconstruct_with_workspace(slots({0, 6}) + slots({})[:1], ws)) == \
        "TooFewParityShards"
    cauchy = rsgpu.ReedSolomon(k, m, matrix_mode="cauchy")
    assert _err(lambda: cauchy.reconstruct_with_workspace(slots({0, 6}), ws)) == "InvalidCustomMatrix"
    leo = rsgpu.ReedSolomon(k, m, codec_family="leopard_gf8")
    assert _err(lambda: leo.prepare_reconstruct_workspace(layout)) == "UnsupportedCodecFamily"
    assert _err(lambda: leo.reconstruct_with_workspace(slots({0, 6}), ws)) == "UnsupportedCodecFamily"
    assert _err(lambda: codec.prepare_reconstruct_some_workspace(layout, [True] * 3)) == "InvalidShardFlags"
    assert _err(lambda: codec.prepare_reconstruct_workspace([None] * 4 + full[4:])) == "TooFewShardsPresent"
    empty = list(layout)
    empty[2] = full[2][:0]
    assert _err(lambda: codec.prepare_reconstruct_workspace(empty)) == "EmptyShard"
    # a validation error leaves every destination untouched
    bad = slots({0, 6})
    bad[0].data.fill_(POISON)
    bad[3].mark_missing()
    assert _err(lambda: codec.reconstruct_with_workspace(bad, ws)) == "InvalidShardFlags"
    assert int(bad[0].data[0]) == POISON
    # an equivalent codec (same format) reuses the plan
    twin = rsgpu.ReedSolomon(k, m)
    ok = slots({0, 6})
    twin.reconstruct_with_workspace(ok, ws)
    torch.cuda.synchronize()
    assert torch.equal(ok[0].data, full[0]) and torch.equal(ok[6].data, full[6])
    # an all-present plan accepts only the all-present layout
    ws_all = codec.prepare_reconstruct_workspace(full)
    assert ws_all.data_targets() == [] and ws_all.parity_targets() == []
    codec.reconstruct_with_workspace([s.clone() for s in full], ws_all)
    assert _err(lambda: codec.reconstruct_with_workspace(slots({0}), ws_all)) == "InvalidShardFlags"
