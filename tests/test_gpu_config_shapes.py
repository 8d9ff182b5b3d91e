"""GPU parity at the benchmark configurations' own shapes: Leopard GF(2^16)
1024+256 and 4096+1024, classic 8+3 and 10+4 at 1 MiB, and the full 8192-stripe
ragged mix.  Every rebuilt shard is written into a buffer poisoned beforehand,
so a reconstruct that leaves its targets untouched fails.

The Leopard and classic codes are column-independent in 64-byte blocks
(leopard_gf16/encode.rs:262-371 and decode.rs:918-995 process chunk columns;
classic rows are bytewise), so the oracle checks column windows of the long
shards; round trips and the device's own cross-checks cover the full length."""
import numpy as np
import pytest

import oracle_lib as O
from gpu_util import announce, to_dev, to_host

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
rsgpu = pytest.importorskip("rsgpu")

POISON = 0xA5


def _windows(length, width):
    """Column windows [a, b) (multiples of 64) at the start, middle and end."""
    mid = (length // 2) // 64 * 64
    return [(0, width), (mid, mid + width), (length - width, length)]


def _slots(dev_shards, missing, length):
    slots = []
    for i, t in enumerate(dev_shards):
        if i in missing:
            slots.append(rsgpu.ShardSlot(torch.full((length,), POISON, dtype=torch.uint8, device="cuda"),
                                         present=False))
        else:
            slots.append(rsgpu.ShardSlot(t, present=True))
    return slots


def _leo_encode_and_check(bits, k, m, length, seed, win):
    rng = np.random.default_rng(seed)
    full = torch.empty((k + m, length), dtype=torch.uint8, device="cuda")
    full[:k].random_(0, 256, generator=torch.Generator(device="cuda").manual_seed(seed))
    full[k:].fill_(POISON)
    rs = rsgpu.ReedSolomon(k, m, codec_family="leopard_gf16" if bits == 16 else "leopard_gf8")
    rs.encode(list(full))
    torch.cuda.synchronize()
    host = full.cpu().numpy()
    for a, b in _windows(length, win):
        rc, par = O.leo_encode(bits, k, m, [host[i, a:b] for i in range(k)])
        assert rc == 0
        for j in range(m):
            assert (host[k + j, a:b] == par[j]).all(), (a, j)
    assert rs.verify(list(full))
    return rs, full, host, rng


def _leo_reconstruct_check(bits, rs, k, m, full, host, missing, win, expect_host=None):
    """Device reconstruct of `missing` into poisoned slots; every rebuilt byte equals
    the original codeword (expect_host None) or the oracle decoder on windows."""
    length = full.shape[1]
    slots = _slots(list(full), set(missing), length)
    rs.reconstruct(slots)
    torch.cuda.synchronize()
    for i in missing:
        assert slots[i].present
    if expect_host is None:
        for i in missing:
            assert torch.equal(slots[i].data, full[i]), i
        return
    present = [0 if i in missing else 1 for i in range(k + m)]
    for a, b in _windows(length, win):
        rc, out = O.leo_reconstruct(bits, k, m, [expect_host[i, a:b].copy() for i in range(k + m)], present)
        assert rc == 0
        for i in missing:
            got = to_host(slots[i].data[a:b])
            assert (got == out[i]).all(), (a, i)


@pytest.mark.parametrize("length", [64 << 10])
def test_leopard_gf16_1024_256_reconstruct_matches_oracle(length):
    """1024+256 (n = 2048 work lanes: the 8-word, 1024-thread decode tile) at 64 KiB:
    m erasures and random counts, codewords and non-codewords, against the oracle."""
    k, m = 1024, 256
    announce(f"leopard gf16 {k}+{m} len={length}")
    rs, full, host, rng = _leo_encode_and_check(16, k, m, length, 1024256, 1024)
    for e in (m, int(rng.integers(1, m)), 1):
        missing = sorted(rng.choice(k + m, e, replace=False).tolist())
        announce(f"  erasures={e}")
        _leo_reconstruct_check(16, rs, k, m, full, host, missing, 1024)
    # non-codeword input: the device decoder gives the oracle decoder's bytes
    noise = torch.empty_like(full).random_(0, 256, generator=torch.Generator(device="cuda").manual_seed(9))
    nh = noise.cpu().numpy()
    for e in (m, 37):
        missing = sorted(rng.choice(k + m, e, replace=False).tolist())
        announce(f"  non-codeword erasures={e}")
        _leo_reconstruct_check(16, rs, k, m, noise, nh, missing, 512, expect_host=nh)


def test_leopard_gf16_1024_256_roundtrip_256k():
    k, m, length = 1024, 256, 256 << 10
    announce(f"leopard gf16 {k}+{m} len={length} round trip")
    rs, full, host, rng = _leo_encode_and_check(16, k, m, length, 77, 256)
    for e in (m, 128):
        missing = sorted(rng.choice(k + m, e, replace=False).tolist())
        _leo_reconstruct_check(16, rs, k, m, full, host, missing, 256)


def test_leopard_gf16_4096_1024_at_4k_matches_oracle():
    """4096+1024 (n = 8192 lanes, the two-word tile) at 4 KiB with m erasures."""
    k, m, length = 4096, 1024, 4 << 10
    announce(f"leopard gf16 {k}+{m} len={length}")
    rs, full, host, rng = _leo_encode_and_check(16, k, m, length, 40961024, 256)
    for e in (m, 300):
        missing = sorted(rng.choice(k + m, e, replace=False).tolist())
        announce(f"  erasures={e}")
        _leo_reconstruct_check(16, rs, k, m, full, host, missing, 256)
    noise = torch.empty_like(full).random_(0, 256, generator=torch.Generator(device="cuda").manual_seed(5))
    nh = noise.cpu().numpy()
    missing = sorted(rng.choice(k + m, m, replace=False).tolist())
    _leo_reconstruct_check(16, rs, k, m, noise, nh, missing, 128, expect_host=nh)


def test_leopard_gf8_128_64_reconstruct_256k():
    k, m, length = 128, 64, 256 << 10
    announce(f"leopard gf8 {k}+{m} len={length}")
    rs, full, host, rng = _leo_encode_and_check(8, k, m, length, 12864, 4096)
    for e in (m, 17, 1):
        missing = sorted(rng.choice(k + m, e, replace=False).tolist())
        _leo_reconstruct_check(8, rs, k, m, full, host, missing, 4096)
    noise = torch.empty_like(full).random_(0, 256, generator=torch.Generator(device="cuda").manual_seed(3))
    nh = noise.cpu().numpy()
    missing = sorted(rng.choice(k + m, 40, replace=False).tolist())
    _leo_reconstruct_check(8, rs, k, m, noise, nh, missing, 4096, expect_host=nh)


@pytest.mark.parametrize("k,m", [(8, 3), (10, 4)])
def test_classic_1mib_windows_and_every_erasure_count(k, m):
    """Classic 8+3 and 10+4 at 1 MiB: parity against the oracle on windows, then
    reconstruct, reconstruct_data and verify for 1..m erasures at random positions."""
    length = 1 << 20
    announce(f"classic {k}+{m} len={length}")
    rs, full, host, rng = None, None, None, np.random.default_rng(k * 10 + m)
    full = torch.empty((k + m, length), dtype=torch.uint8, device="cuda")
    full[:k].random_(0, 256, generator=torch.Generator(device="cuda").manual_seed(k + m))
    full[k:].fill_(POISON)
    rs = rsgpu.ReedSolomon(k, m)
    rs.encode(list(full))
    torch.cuda.synchronize()
    host = full.cpu().numpy()
    _, omat = O.build_matrix(O.GF8, O.VANDERMONDE, k, m)
    for a, b in _windows(length, 64 << 10):
        exp = O.encode(O.GF8, omat, k, m, [host[i, a:b] for i in range(k)])
        for j in range(m):
            assert (host[k + j, a:b] == exp[j]).all(), (a, j)
    assert rs.verify(list(full))
    for e in range(1, m + 1):
        for trial in range(2):
            missing = sorted(rng.choice(k + m, e, replace=False).tolist())
            slots = _slots(list(full), set(missing), length)
            rs.reconstruct(slots)
            torch.cuda.synchronize()
            for i in missing:
                assert slots[i].present and torch.equal(slots[i].data, full[i]), (e, missing, i)
            slots = _slots(list(full), set(missing), length)
            rs.reconstruct_data(slots)
            torch.cuda.synchronize()
            for i in missing:
                if i < k:
                    assert torch.equal(slots[i].data, full[i]), (e, missing, i)
                else:
                    assert not slots[i].present
                    assert int(slots[i].data[0]) == POISON
    bad = full.clone()
    bad[k - 1, length - 1] ^= 1
    assert not rs.verify(list(bad))


def test_full_ragged_spec_plain_equals_batched():
    """The benchmark's whole ragged mix (8192 stripes, 8 shapes, 4 KiB-4 MiB shards,
    60/20/20 encode/verify/reconstruct with 0..m erasures): every output wiped, then
    written by per-stripe launches and again by one batch launch; both equal the
    consistent bytes, verify flags stay clear, and one stripe per shape matches
    the oracle on a column window."""
    from rsgpu import workload

    announce("full ragged spec plain == batched")
    spec = workload.ragged_spec(8192)
    rg = workload.RaggedSet(rsgpu, torch, spec, seed=5)
    sp = rsgpu._stream(None)
    rg.make_consistent(sp)
    torch.cuda.synchronize()
    golden = rg.buf.clone()
    print(f"ragged bytes {rg.buf.numel() / 1e9:.1f} GB", flush=True)

    def wipe():
        for (k, m, size, op, er), off in zip(rg.spec, rg.offs):
            targets = range(k, k + m) if op == "encode" else (er if op == "reconstruct" else [])
            for i in targets:
                o = off + i * (size + 256)
                rg.buf[o:o + size].fill_(POISON)

    wipe()
    rg.run_plain(sp)
    torch.cuda.synchronize()
    announce("  plain done")
    assert torch.equal(rg.buf, golden)
    assert int(rg.flags.sum()) == 0
    wipe()
    b = rg.make_batch(sp)
    b.run()
    torch.cuda.synchronize()
    announce("  batched done")
    assert torch.equal(rg.buf, golden)
    assert int(rg.flags.sum()) == 0
    seen = set()
    for (k, m, size, op, er), off in zip(rg.spec, rg.offs):
        if (k, m) in seen:
            continue
        seen.add((k, m))
        w = min(size, 64 << 10)
        sh = [to_host(rg.buf[off + i * (size + 256) + size - w: off + i * (size + 256) + size])
              for i in range(k + m)]
        _, omat = O.build_matrix(O.GF8, O.VANDERMONDE, k, m)
        exp = O.encode(O.GF8, omat, k, m, sh[:k])
        assert all((sh[k + j] == exp[j]).all() for j in range(m)), (k, m)
    del golden
