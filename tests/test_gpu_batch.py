"""GPU parity of the ragged-batch kernel: every stripe of a mixed batch
(shapes, shard sizes, encode/verify/reconstruct, erasures) equals the oracle,
also when the batch is replayed from a hipGraph and after autotuning."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import announce, to_host

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
rsgpu = pytest.importorskip("rsgpu")

SHAPES = [(4, 2), (6, 3), (8, 3), (8, 4), (10, 4), (12, 4), (16, 4), (20, 4)]


def _mixed_batch(seed, n_stripes, max_len=1 << 16, tails=True):
    rng = np.random.default_rng(seed)
    codecs = {s: rsgpu.ReedSolomon(*s) for s in SHAPES}
    mats = {s: O.build_matrix(O.GF8, O.VANDERMONDE, *s)[1] for s in SHAPES}
    items = []
    for i in range(n_stripes):
        k, m = SHAPES[rng.integers(len(SHAPES))]
        length = int(rng.integers(1, max_len // 64 + 1)) * 64
        if tails and i % 7 == 3:
            length += int(rng.integers(1, 16))
        op = ["encode", "encode", "encode", "verify", "reconstruct"][rng.integers(5)]
        data = [rng.integers(0, 256, length, dtype=np.uint8) for _ in range(k)]
        parity = O.encode(O.GF8, mats[(k, m)], k, m, data)
        items.append(dict(k=k, m=m, length=length, op=op, data=data, parity=parity))
    return rng, codecs, mats, items


def test_mixed_batch_matches_oracle_and_graph_replays():
    announce("mixed batch 120 stripes")
    rng, codecs, mats, items = _mixed_batch(1, 120)
    batch = rsgpu.Batch()
    flags = torch.zeros(len(items), dtype=torch.int32, device="cuda")
    views = []
    for idx, it in enumerate(items):
        k, m, n = it["k"], it["m"], it["length"]
        rs = codecs[(k, m)]
        full = it["data"] + it["parity"]
        if it["op"] == "encode":
            sh = [torch.from_numpy(d).cuda() for d in it["data"]] + \
                 [torch.zeros(n, dtype=torch.uint8, device="cuda") for _ in range(m)]
            batch.add_encode(rs, sh)
            views.append(("encode", sh, full, k))
        elif it["op"] == "verify":
            sh = [torch.from_numpy(x).cuda() for x in full]
            corrupt = idx % 2 == 0
            if corrupt:
                sh[k + int(rng.integers(m))][int(rng.integers(n))] ^= 0x5A
            batch.add_verify(rs, sh, flags[idx:idx + 1])
            views.append(("verify", idx, corrupt))
        else:
            e = int(rng.integers(1, m + 1))
            missing = set(rng.choice(k + m, e, replace=False).tolist())
            slots = [rsgpu.ShardSlot(torch.from_numpy(x).cuda() if i not in missing
                                     else torch.zeros(n, dtype=torch.uint8, device="cuda"),
                                     i not in missing) for i, x in enumerate(full)]
            batch.add_reconstruct(rs, slots)
            views.append(("reconstruct", slots, full, missing))

    def check():
        torch.cuda.synchronize()
        for v in views:
            if v[0] in ("encode", "reconstruct"):
                tensors = v[1] if v[0] == "encode" else [s.data for s in v[1]]
                for i, t in enumerate(tensors):
                    assert (to_host(t) == v[2][i]).all(), (v[0], i)
            else:
                assert (int(flags[v[1]]) != 0) == v[2], v

    batch.run()
    check()
    # graph replays, with eager work and host syncs between them; every replay re-checked
    for rep in range(3):
        for v in views:  # clear every output so that the replay must rewrite it
            if v[0] == "encode":
                for t in v[1][v[3]:]:
                    t.zero_()
            elif v[0] == "reconstruct":
                for i in v[3]:
                    v[1][i].data.zero_()
        flags.zero_()
        batch.graph_launch()
        check()
        torch.cuda.synchronize()


def test_autotune_keeps_outputs_bit_exact():
    announce("batch autotune")
    _, codecs, mats, items = _mixed_batch(2, 40, max_len=1 << 15, tails=False)
    batch = rsgpu.Batch()
    outs = []
    for it in items:
        k, m, n = it["k"], it["m"], it["length"]
        sh = [torch.from_numpy(d).cuda() for d in it["data"]] + \
             [torch.zeros(n, dtype=torch.uint8, device="cuda") for _ in range(m)]
        batch.add_encode(codecs[(k, m)], sh)
        outs.append((sh, it))
    best = batch.autotune(reps=2)
    assert best["inputs_per_group"] in (4, 6, 8, 12)
    for sh, it in outs:
        for t in sh[it["k"]:]:
            t.zero_()
    batch.run()
    torch.cuda.synchronize()
    for sh, it in outs:
        for j in range(it["m"]):
            assert (to_host(sh[it["k"] + j]) == it["parity"][j]).all()
    st = batch.stats()
    assert st["ops"] == len(items)


@pytest.mark.parametrize("kb,nv,bpc", [(8, 1, 0), (4, 2, 0), (12, 1, 0), (6, 1, 1), (6, 2, 32)])
def test_every_launch_shape_is_bit_exact(kb, nv, bpc):
    """Persistent grids of any size and the one-workgroup-per-tile grid (0)."""
    announce(f"batch tune kb={kb} nv={nv} blocks_per_cu={bpc}")
    _, codecs, mats, items = _mixed_batch(11, 30, max_len=1 << 15, tails=True)
    batch = rsgpu.Batch()
    outs = []
    for it in items:
        k, m, n = it["k"], it["m"], it["length"]
        sh = [torch.from_numpy(d).cuda() for d in it["data"]] + \
             [torch.full((n,), 0xA5, dtype=torch.uint8, device="cuda") for _ in range(m)]
        batch.add_encode(codecs[(k, m)], sh)
        outs.append((sh, it))
    batch.set_tune(kb, nv, bpc)
    batch.run()
    torch.cuda.synchronize()
    for sh, it in outs:
        for j in range(it["m"]):
            assert (to_host(sh[it["k"] + j]) == it["parity"][j]).all(), (it["k"], it["m"], it["length"], j)


def test_batch_tower_encode_small():
    announce("batch tower encode 5+3")
    k, m, n = 5, 3, 256
    rs16 = rsgpu.ReedSolomon(k, m, field="gf16")
    _, omat = O.build_matrix(O.GF16, O.VANDERMONDE, k, m)
    rng = np.random.default_rng(3)
    data = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k)]
    sh = [torch.from_numpy(d).cuda() for d in data] + [torch.zeros(n, dtype=torch.uint8, device="cuda")
                                                       for _ in range(m)]
    b = rsgpu.Batch()
    b.add_encode(rs16, sh)
    b.run()
    torch.cuda.synchronize()
    exp = O.encode(O.GF16, omat, k, m, data)
    assert all((to_host(sh[k + j]) == exp[j]).all() for j in range(m))


def test_leopard_gf8_ops_in_the_batch_kernel_match_the_fft():
    """Leopard GF(2^8) codecs with k <= 32 and no dispatch state put their matrix forms
    into the batch kernel (encode and verify over [I; G], reconstruct with exactly k
    present as the [I; G] solve, with more as the decoder's own matrix); larger ones
    run their FFT as deferred launches.  Every output equals the oracle's FFT codec,
    on non-codeword inputs for the decoder-matrix reconstruct, and a corrupted stripe
    fails its verify."""
    announce("batch leopard gf8 matrix forms")
    rng = np.random.default_rng(88)
    b = rsgpu.Batch()
    checks = []
    for k, m, n in [(10, 4, 4096), (32, 16, 1024), (128, 64, 640)]:
        # rs is rebound on each pass: the batch holds every op's codec until it runs
        rs = rsgpu.ReedSolomon(k, m, codec_family="leopard_gf8")
        data = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k)]
        rc, par = O.leo_encode(8, k, m, data)
        assert rc == 0
        # encode
        sh = [torch.from_numpy(d).cuda() for d in data] + [torch.zeros(n, dtype=torch.uint8, device="cuda")
                                                           for _ in range(m)]
        b.add_encode(rs, sh)
        checks.append(("encode", sh, [(k + j, par[j]) for j in range(m)]))
        # verify: a codeword and a corrupted copy
        full = data + list(par)
        good = [torch.from_numpy(x.copy()).cuda() for x in full]
        bad = [torch.from_numpy(x.copy()).cuda() for x in full]
        bad[k + m - 1][5] ^= 1
        flags = torch.zeros(2, dtype=torch.int32, device="cuda")
        b.add_verify(rs, good, flags[0:1])
        b.add_verify(rs, bad, flags[1:2])
        checks.append(("verify", flags, None))
        # reconstruct, exactly k present (a codeword)
        miss = sorted(rng.choice(k + m, m, replace=False).tolist())
        slots = [rsgpu.ShardSlot(torch.zeros(n, dtype=torch.uint8, device="cuda"), False) if i in miss
                 else rsgpu.ShardSlot(torch.from_numpy(full[i].copy()).cuda()) for i in range(k + m)]
        b.add_reconstruct(rs, slots)
        checks.append(("rec", slots, [(i, full[i]) for i in miss]))
        # reconstruct, more than k present, non-codeword inputs (the decoder's bytes)
        junk = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k + m)]
        miss = sorted(rng.choice(k + m, 2, replace=False).tolist())
        present = [0 if i in miss else 1 for i in range(k + m)]
        rc, exp = O.leo_reconstruct(8, k, m, [x.copy() for x in junk], present)
        assert rc == 0
        slots = [rsgpu.ShardSlot(torch.zeros(n, dtype=torch.uint8, device="cuda"), False) if i in miss
                 else rsgpu.ShardSlot(torch.from_numpy(junk[i].copy()).cuda()) for i in range(k + m)]
        b.add_reconstruct(rs, slots)
        checks.append(("rec", slots, [(i, exp[i]) for i in miss]))
    b.run()
    torch.cuda.synchronize()
    for kind, obj, want in checks:
        if kind == "encode":
            for i, w in want:
                assert (to_host(obj[i]) == w).all(), (kind, i)
        elif kind == "verify":
            assert int(obj[0]) == 0 and int(obj[1]) != 0
        else:
            for i, w in want:
                assert obj[i].present and (to_host(obj[i].data) == w).all(), (kind, i)


def _dev_view(host, off):
    """A device copy of `host` starting `off` bytes into its allocation."""
    buf = torch.zeros(host.size + off, dtype=torch.uint8, device="cuda")
    v = buf[off:off + host.size]
    v.copy_(torch.from_numpy(host))
    return v


def test_classic_gf16_ops_in_the_tower_batch_kernel_match_the_oracle():
    """Classic GF(2^16) encode, verify and reconstruct (all and data only) run in
    the tower batch kernel beside GF(2^8) stripes: more outputs than one
    descriptor holds, more inputs than one apply launch takes, 16-byte tails and
    unaligned views through the byte kernel.  Reconstructs take non-codeword
    inputs, where the survivors chosen decide the bytes.  Every output equals the
    oracle's, and a corrupted stripe fails its verify."""
    announce("batch classic gf16 tower kernel")
    rng = np.random.default_rng(16)
    b = rsgpu.Batch()
    checks = []
    for k, m, n, off in [(10, 4, 4096, 0), (5, 3, 2006, 0), (12, 6, 2048, 2), (20, 8, 65538, 0),
                         (70, 6, 4096, 0), (3, 5, 256, 0), (8, 3, 1 << 20, 0)]:
        for field in ("gf16", "gf8") if k == 10 else ("gf16",):
            F = O.GF16 if field == "gf16" else O.GF8
            rs = rsgpu.ReedSolomon(k, m, field=field)
            mat = O.build_matrix(F, O.VANDERMONDE, k, m)[1]
            data = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k)]
            par = O.encode(F, mat, k, m, data)
            sh = [_dev_view(d, off) for d in data] + [_dev_view(np.zeros(n, np.uint8), off) for _ in range(m)]
            b.add_encode(rs, sh)
            checks.append(("encode", sh, [(k + j, par[j]) for j in range(m)]))
            full = data + list(par)
            flags = torch.zeros(2, dtype=torch.int32, device="cuda")
            good = [_dev_view(x, off) for x in full]
            bad = [_dev_view(x, off) for x in full]
            bad[rng.integers(k + m)][int(rng.integers(n))] ^= 0x40
            b.add_verify(rs, good, flags[0:1])
            b.add_verify(rs, bad, flags[1:2])
            checks.append(("verify", flags, None))
            for data_only in (False, True):
                junk = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k + m)]
                e = int(rng.integers(1, m + 1))
                miss = sorted(rng.choice(k + m, e, replace=False).tolist())
                present = [0 if i in miss else 1 for i in range(k + m)]
                rc, exp = O.reconstruct(F, mat, k, m, [x.copy() for x in junk], present, data_only)
                assert rc == 0
                slots = [rsgpu.ShardSlot(_dev_view(np.zeros(n, np.uint8), off), False) if i in miss
                         else rsgpu.ShardSlot(_dev_view(junk[i], off)) for i in range(k + m)]
                b.add_reconstruct(rs, slots, data_only=data_only)
                want = [i for i in miss if i < k or not data_only]
                checks.append(("rec", slots, [(i, exp[i]) for i in want]))
    assert b.stats()["tiles"] > 0
    b.run()
    torch.cuda.synchronize()
    for kind, obj, want in checks:
        if kind == "encode":
            for i, w in want:
                assert (to_host(obj[i]) == w).all(), (kind, i)
        elif kind == "verify":
            assert int(obj[0]) == 0 and int(obj[1]) != 0
        else:
            for i, w in want:
                assert obj[i].present and (to_host(obj[i].data) == w).all(), (kind, i)


def test_leopard_gf16_reconstruct_in_the_split_batch_kernel_matches_the_oracle():
    """A Leopard GF(2^16) reconstruct whose dispatch picks a matrix variant runs
    as the decoder's own matrix in the split-layout batch kernel (more targets than
    one descriptor holds, thousands of inputs); one with more erasures than the
    matrix form takes runs its FFT after the kernels.  Inputs are not codewords;
    every output equals the oracle's FFT decoder."""
    announce("batch leopard gf16 split kernel")
    rng = np.random.default_rng(61)
    b = rsgpu.Batch()
    checks = []
    for k, m, n, e in [(10, 4, 4096, 1), (10, 4, 640, 3), (64, 16, 4096, 16), (300, 45, 640, 9),
                       (1024, 256, 4096, 2), (300, 45, 640, 40)]:
        rs = rsgpu.ReedSolomon(k, m, codec_family="leopard_gf16")
        rs.set_variant(0)  # a matrix variant: taken by reconstructs of at most 32 erasures
        junk = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k + m)]
        miss = sorted(rng.choice(k + m, e, replace=False).tolist())
        present = [0 if i in miss else 1 for i in range(k + m)]
        rc, exp = O.leo_reconstruct(16, k, m, [x.copy() for x in junk], present)
        assert rc == 0
        slots = [rsgpu.ShardSlot(torch.zeros(n, dtype=torch.uint8, device="cuda"), False) if i in miss
                 else rsgpu.ShardSlot(torch.from_numpy(junk[i].copy()).cuda()) for i in range(k + m)]
        b.add_reconstruct(rs, slots)
        assert all(slots[i].present for i in miss)
        checks.append((slots, [(i, exp[i]) for i in miss]))
    assert b.stats()["tiles"] > 0
    b.run()
    torch.cuda.synchronize()
    for slots, want in checks:
        for i, w in want:
            assert (to_host(slots[i].data) == w).all(), i


def test_leopard_gf16_few_erasures_enter_the_batch_without_dispatch_state():
    """With no dispatch state, a Leopard GF(2^16) reconstruct of at most 4 erasures
    runs as the decoder's own matrix in the split-layout kernel (the batch gains
    tiles); 5 or more run the FFT after the kernels.  Bytes equal the oracle's."""
    announce("batch leopard gf16 default matrix rule")
    rng = np.random.default_rng(62)
    k, m, n = 100, 20, 1024
    rs = rsgpu.ReedSolomon(k, m, codec_family="leopard_gf16")
    for e, in_kernel in ((1, True), (4, True), (5, False)):
        b = rsgpu.Batch()
        junk = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k + m)]
        miss = sorted(rng.choice(k + m, e, replace=False).tolist())
        rc, exp = O.leo_reconstruct(16, k, m, [x.copy() for x in junk],
                                    [0 if i in miss else 1 for i in range(k + m)])
        assert rc == 0
        slots = [rsgpu.ShardSlot(torch.zeros(n, dtype=torch.uint8, device="cuda"), False) if i in miss
                 else rsgpu.ShardSlot(torch.from_numpy(junk[i].copy()).cuda()) for i in range(k + m)]
        b.add_reconstruct(rs, slots)
        assert (b.stats()["tiles"] > 0) == in_kernel, e
        b.run()
        torch.cuda.synchronize()
        for i in miss:
            assert slots[i].present and (to_host(slots[i].data) == exp[i]).all(), (e, i)


def test_graph_replays_every_element_type():
    """One batch holding GF(2^8), tower GF(2^16) and split-layout Leopard GF(2^16)
    stripes is three kernel launches, so graph_launch captures and replays them.
    The outputs are overwritten between replays (eager work on the same stream)
    and every replay must rebuild the oracle's bytes."""
    announce("batch graph replay, three element types")
    rng = np.random.default_rng(77)
    b = rsgpu.Batch()
    outs = []
    for field, k, m, n in [("gf8", 10, 4, 8192), ("gf16", 6, 3, 4096 + 2), ("gf16", 12, 4, 16384)]:
        F = O.GF16 if field == "gf16" else O.GF8
        rs = rsgpu.ReedSolomon(k, m, field=field)
        mat = O.build_matrix(F, O.VANDERMONDE, k, m)[1]
        data = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k)]
        par = O.encode(F, mat, k, m, data)
        sh = [torch.from_numpy(d).cuda() for d in data] + [torch.zeros(n, dtype=torch.uint8, device="cuda")
                                                           for _ in range(m)]
        b.add_encode(rs, sh)
        outs += [(sh[k + j], par[j]) for j in range(m)]
    k, m, n = 64, 16, 4096
    rs = rsgpu.ReedSolomon(k, m, codec_family="leopard_gf16")
    rs.set_variant(0)
    junk = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k + m)]
    miss = [3, 40, 70]
    rc, exp = O.leo_reconstruct(16, k, m, [x.copy() for x in junk], [0 if i in miss else 1 for i in range(k + m)])
    assert rc == 0
    slots = [rsgpu.ShardSlot(torch.zeros(n, dtype=torch.uint8, device="cuda"), False) if i in miss
             else rsgpu.ShardSlot(torch.from_numpy(junk[i].copy()).cuda()) for i in range(k + m)]
    b.add_reconstruct(rs, slots)
    outs += [(slots[i].data, exp[i]) for i in miss]
    for rep in range(3):
        for t, _ in outs:
            t.fill_(0xA5 ^ rep)
        torch.cuda.synchronize()
        b.graph_launch()
        torch.cuda.synchronize()
        for t, w in outs:
            assert (to_host(t) == w).all(), rep


def test_ragged_workload_plain_equals_batched():
    """The ragged mix (mixed shapes, 4 KiB-256 KiB shards, encode/verify/reconstruct with
    0..m erasures): the batch kernel writes the same bytes as per-stripe launches."""
    announce("ragged workload plain == batched")
    from rsgpu import workload

    spec = workload.ragged_spec(96, seed=11, hi=256 << 10)
    rg = workload.RaggedSet(rsgpu, torch, spec, seed=3)
    sp = rsgpu._stream(None)
    rg.make_consistent(sp)
    torch.cuda.synchronize()
    golden = rg.buf.clone()
    # wipe all parity of encode stripes and all erased shards, then run per stripe
    base = rg.buf.data_ptr()

    def wipe():
        for (k, m, size, op, er), off in zip(rg.spec, rg.offs):
            targets = range(k, k + m) if op == "encode" else (er if op == "reconstruct" else [])
            for i in targets:
                o = off + i * (size + 256)
                rg.buf[o:o + size].zero_()

    wipe()
    rg.run_plain(sp)
    torch.cuda.synchronize()
    assert torch.equal(rg.buf, golden)
    assert int(rg.flags.sum()) == 0
    wipe()
    b = rg.make_batch(sp)
    b.run()
    torch.cuda.synchronize()
    assert torch.equal(rg.buf, golden)
    assert int(rg.flags.sum()) == 0
    # spot-check one stripe of every shape against the oracle
    seen = set()
    for (k, m, size, op, er), off in zip(rg.spec, rg.offs):
        if (k, m) in seen:
            continue
        seen.add((k, m))
        sh = [to_host(rg.buf[off + i * (size + 256): off + i * (size + 256) + size]) for i in range(k + m)]
        _, omat = O.build_matrix(O.GF8, O.VANDERMONDE, k, m)
        exp = O.encode(O.GF8, omat, k, m, sh[:k])
        assert all((sh[k + j] == exp[j]).all() for j in range(m)), (k, m)


def test_growth_and_streams():
    """A batch grown while an earlier run is still queued, rerun on another
    stream without a new upload, then rebuilt: every run equals the oracle."""
    announce("batch growth across streams")
    _, codecs, mats, items = _mixed_batch(21, 120, max_len=1 << 14, tails=False)
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    batch = rsgpu.Batch()
    outs = []

    def add(it):
        k, m, n = it["k"], it["m"], it["length"]
        sh = [torch.from_numpy(d).cuda() for d in it["data"]] + \
             [torch.zeros(n, dtype=torch.uint8, device="cuda") for _ in range(m)]
        batch.add_encode(codecs[(k, m)], sh)
        outs.append((sh, it))

    def check():
        for sh, it in outs:
            for j in range(it["m"]):
                assert (to_host(sh[it["k"] + j]) == it["parity"][j]).all()

    torch.cuda.synchronize()
    for it in items[:4]:
        add(it)
    batch.run(stream=sa)          # small upload on stream A
    for it in items[4:]:
        add(it)                   # the tables outgrow their buffers
    batch.run(stream=sa)
    torch.cuda.synchronize()
    check()
    for sh, it in outs:
        for t in sh[it["k"]:]:
            t.zero_()
    torch.cuda.synchronize()
    batch.run(stream=sb)          # no new upload: stream B waits for A's copies
    torch.cuda.synchronize()
    check()
    batch.clear()
    outs.clear()
    for it in items[:50]:
        add(it)
    batch.run(stream=sb)
    torch.cuda.synchronize()
    check()
