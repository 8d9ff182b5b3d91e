"""Block-stream encode / verify / reconstruct (src/core/stream.rs:677-1031) on the
GPU path: bytes against the oracle, the zero-padding and length rules, the
error kinds and indices of StreamError, and the reference's own stream tests
(stream.rs tests: basic, multi-block, partial last block, verify, reconstruct)."""
import io

import numpy as np
import pytest

import oracle_lib as O
from gpu_util import announce

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
rsgpu = pytest.importorskip("rsgpu")


def ref_data(n):
    # random_data of the reference's stream tests (stream.rs tests)
    return bytes(((i * 73 + 17) & 0xFF) for i in range(n))


def oracle_parity(k, m, data):
    rc, mat = O.build_matrix(O.GF8, O.VANDERMONDE, k, m)
    arrs = [np.frombuffer(d, np.uint8) for d in data]
    return [p.tobytes() for p in O.encode(O.GF8, mat, k, m, arrs)]


@pytest.mark.parametrize("k,m,total,block", [(3, 2, 4096, 4 << 20), (3, 2, 10 * 1024, 4096),
                                             (10, 4, (1 << 20) + 333, 65536), (6, 3, 5000, 0),
                                             (20, 4, 3 << 20, 1 << 20)])
def test_encode_stream_matches_oracle(k, m, total, block):
    announce(f"stream encode {k}+{m} total={total} block={block}")
    rng = np.random.default_rng(total + k)
    data = [rng.integers(0, 256, total, dtype=np.uint8).tobytes() for _ in range(k)]
    rs = rsgpu.ReedSolomon(k, m)
    writers = [io.BytesIO() for _ in range(m)]
    rs.encode_stream([io.BytesIO(d) for d in data], writers, block_size=block)
    exp = oracle_parity(k, m, data)
    for j in range(m):
        assert writers[j].getvalue() == exp[j], j
    assert rs.verify_stream([io.BytesIO(d) for d in data] + [io.BytesIO(w.getvalue()) for w in writers],
                            block_size=block)


def test_encode_stream_zero_pads_short_streams():
    """Each block is as long as the longest read; shorter data streams are padded
    with zeros (stream.rs:259-283)."""
    announce("stream encode zero padding")
    k, m = 4, 2
    lens = [3000, 2500, 3000, 100]
    rng = np.random.default_rng(1)
    data = [rng.integers(0, 256, n, dtype=np.uint8).tobytes() for n in lens]
    rs = rsgpu.ReedSolomon(k, m)
    writers = [io.BytesIO() for _ in range(m)]
    rs.encode_stream([io.BytesIO(d) for d in data], writers, block_size=1024)
    padded = [d + bytes(3000 - len(d)) for d in data]
    exp = oracle_parity(k, m, padded)
    assert [w.getvalue() for w in writers] == exp


def test_verify_stream_detects_corruption_and_length_mismatch():
    announce("stream verify")
    k, m, n = 3, 2, 9000
    data = [ref_data(n) for _ in range(k)]
    par = oracle_parity(k, m, data)
    rs = rsgpu.ReedSolomon(k, m)
    good = data + par
    assert rs.verify_stream([io.BytesIO(x) for x in good], block_size=2048)
    bad = list(good)
    b = bytearray(bad[4])
    b[8500] ^= 0x40  # in the last (partial) block
    bad[4] = bytes(b)
    assert not rs.verify_stream([io.BytesIO(x) for x in bad], block_size=2048)
    short = list(good)
    short[1] = short[1][:5000]
    with pytest.raises(rsgpu.StreamError) as ei:
        rs.verify_stream([io.BytesIO(x) for x in short], block_size=2048)
    assert ei.value.kind == "codec" and ei.value.shard_index == 1
    assert ei.value.error.name == "IncorrectShardSize"
    # empty streams verify as true (no blocks)
    assert rs.verify_stream([io.BytesIO(b"") for _ in range(k + m)])


@pytest.mark.parametrize("missing", [(0,), (1, 4), (2, 3), (0, 1)])
def test_reconstruct_stream_from_rebuilds_missing(missing):
    announce(f"stream reconstruct {missing}")
    k, m, n = 3, 2, 10 * 1024 + 7
    rng = np.random.default_rng(sum(missing) + 3)
    data = [rng.integers(0, 256, n, dtype=np.uint8).tobytes() for _ in range(k)]
    full = data + oracle_parity(k, m, data)
    rs = rsgpu.ReedSolomon(k, m)
    readers = [None if i in missing else io.BytesIO(full[i]) for i in range(k + m)]
    writers = [io.BytesIO() if i in missing else None for i in range(k + m)]
    rs.reconstruct_stream_from(readers, writers, block_size=4096)
    for i in missing:
        assert writers[i].getvalue() == full[i], i


def test_reconstruct_stream_cursor_form():
    announce("stream reconstruct cursors")
    k, m, n = 4, 2, 6000
    data = [ref_data(n) for _ in range(k)]
    full = data + oracle_parity(k, m, data)
    rs = rsgpu.ReedSolomon(k, m)
    shards = [io.BytesIO(x) for x in full]
    for s in shards:
        s.seek(0, 2)  # present cursors are read from position 0 regardless
    shards[1] = io.BytesIO()
    shards[5] = io.BytesIO()
    rs.reconstruct_stream(shards, block_size=1024)
    assert shards[1].getvalue() == full[1] and shards[5].getvalue() == full[5]
    # nothing present: nothing to do
    rs.reconstruct_stream([io.BytesIO() for _ in range(k + m)])
    with pytest.raises(rsgpu.StreamError) as ei:
        rs.reconstruct_stream([io.BytesIO(full[0])] + [io.BytesIO() for _ in range(k + m - 1)])
    assert ei.value.error.name == "TooFewShardsPresent"


def test_stream_errors_match_reference():
    announce("stream errors")
    k, m = 3, 2
    rs = rsgpu.ReedSolomon(k, m)
    with pytest.raises(rsgpu.StreamError) as ei:
        rs.encode_stream([io.BytesIO(b"x")] * 2, [io.BytesIO(), io.BytesIO()])
    assert ei.value.error.name == "TooFewShards" and ei.value.shard_index == 0
    with pytest.raises(rsgpu.StreamError) as ei:
        rs.verify_stream([io.BytesIO(b"x")] * 6)
    assert ei.value.error.name == "TooManyShards"
    leo = rsgpu.ReedSolomon(k, m, codec_family="leopard_gf8")
    with pytest.raises(rsgpu.StreamError) as ei:
        leo.encode_stream([io.BytesIO(b"x")] * 3, [io.BytesIO(), io.BytesIO()])
    assert ei.value.error.name == "UnsupportedCodecFamily"
    # layout: a shard with both a reader and a writer
    readers = [io.BytesIO(b"abc")] * 5
    writers = [None, io.BytesIO(), None, None, None]
    with pytest.raises(rsgpu.StreamError) as ei:
        rs.reconstruct_stream_from(readers, writers)
    assert ei.value.error.name == "InvalidShardFlags" and ei.value.shard_index == 1
    with pytest.raises(rsgpu.StreamError) as ei:
        rs.reconstruct_stream_from([None, None, None, io.BytesIO(b"a"), io.BytesIO(b"b")],
                                   [io.BytesIO(), io.BytesIO(), io.BytesIO(), None, None])
    assert ei.value.error.name == "TooFewShardsPresent"

    class Boom:
        def __init__(self, after):
            self.left = after

        def readinto(self, buf):
            if self.left <= 0:
                raise IOError("disk gone")
            n = min(len(buf), self.left, 1024)
            buf[:n] = b"\x01" * n
            self.left -= n
            return n

    # a read error on the second block: the first block's parity is already written
    writers = [io.BytesIO(), io.BytesIO()]
    with pytest.raises(rsgpu.StreamError) as ei:
        rs.encode_stream([io.BytesIO(ref_data(8192)), Boom(4096), io.BytesIO(ref_data(8192))], writers,
                         block_size=4096)
    assert ei.value.kind == "read" and ei.value.shard_index == 1
    assert isinstance(ei.value.error, IOError)
    first = oracle_parity(k, m, [ref_data(4096), b"\x01" * 4096, ref_data(4096)])
    assert [w.getvalue() for w in writers] == first

    class Full:
        def write(self, b):
            raise IOError("no space")

    with pytest.raises(rsgpu.StreamError) as ei:
        rs.encode_stream([io.BytesIO(ref_data(100))] * 3, [io.BytesIO(), Full()])
    assert ei.value.kind == "write" and ei.value.shard_index == k + 1


def _host_streams(arrs):
    return [rsgpu.HostStream(a) for a in arrs]


@pytest.mark.parametrize("io_mode", ["auto", "serial", "parallel"])
@pytest.mark.parametrize("k,m,total,block", [(10, 4, (8 << 20) + 4097, 4 << 20), (3, 2, 10 * 1024 + 5, 4096),
                                             (12, 4, 3 << 20, 1 << 20)])
def test_native_memory_streams_match_oracle(io_mode, k, m, total, block):
    """HostStream readers/writers run the library's C callbacks: parity, verify and
    a 2-erasure rebuild equal the oracle's bytes in every I/O mode."""
    announce(f"stream native {io_mode} {k}+{m} total={total} block={block}")
    rng = np.random.default_rng(total + k + len(io_mode))
    data = [rng.integers(0, 256, total, dtype=np.uint8) for _ in range(k)]
    parity = [np.zeros(total, np.uint8) for _ in range(m)]
    rs = rsgpu.ReedSolomon(k, m)
    readers, writers = _host_streams(data), _host_streams(parity)
    rs.encode_stream(readers, writers, block_size=block, io_mode=io_mode)
    exp = oracle_parity(k, m, [d.tobytes() for d in data])
    assert all(r.pos == total for r in readers) and all(w.pos == total for w in writers)
    for j in range(m):
        assert parity[j].tobytes() == exp[j], j
    full = data + parity
    assert rs.verify_stream(_host_streams(full), block_size=block, io_mode=io_mode)
    lost = (1, k + 1)
    out = {i: np.zeros(total, np.uint8) for i in lost}
    rd = [None if i in lost else rsgpu.HostStream(full[i]) for i in range(k + m)]
    wr = [rsgpu.HostStream(out[i]) if i in lost else None for i in range(k + m)]
    rs.reconstruct_stream_from(rd, wr, block_size=block, io_mode=io_mode)
    for i in lost:
        assert (out[i] == full[i]).all(), i
    bad = [f.copy() for f in full]
    bad[k][total - 1] ^= 1
    assert not rs.verify_stream(_host_streams(bad), block_size=block, io_mode=io_mode)


@pytest.mark.parametrize("io_mode", ["serial", "parallel"])
def test_python_streams_in_every_io_mode(io_mode):
    """Python callbacks run on the library's worker threads in parallel mode."""
    announce(f"stream python {io_mode}")
    k, m, n = 10, 4, (4 << 20) * 2 + 100
    rng = np.random.default_rng(5)
    data = [rng.integers(0, 256, n, dtype=np.uint8).tobytes() for _ in range(k)]
    rs = rsgpu.ReedSolomon(k, m)
    writers = [io.BytesIO() for _ in range(m)]
    rs.encode_stream([io.BytesIO(d) for d in data], writers, io_mode=io_mode)
    assert [w.getvalue() for w in writers] == oracle_parity(k, m, data)


def test_parallel_errors_report_the_lowest_shard():
    announce("stream parallel errors")
    k, m = 10, 4
    rs = rsgpu.ReedSolomon(k, m)

    class Fail:
        def readinto(self, b):
            raise IOError("bad sector")

    readers = [io.BytesIO(ref_data(5000)) for _ in range(k)]
    readers[3], readers[7] = Fail(), Fail()
    with pytest.raises(rsgpu.StreamError) as ei:
        rs.encode_stream(readers, [io.BytesIO() for _ in range(m)], block_size=1024, io_mode="parallel")
    assert ei.value.kind == "read" and ei.value.shard_index == 3
    # fixed-size writers: the third (1000 bytes) is full within block 0 -> a write error at k + 2
    data = [np.frombuffer(ref_data(3000), np.uint8).copy() for _ in range(k)]
    outs = [np.zeros(3000, np.uint8), np.zeros(2000, np.uint8), np.zeros(1000, np.uint8),
            np.zeros(3000, np.uint8)]
    with pytest.raises(rsgpu.StreamError) as ei:
        rs.encode_stream(_host_streams(data), _host_streams(outs), block_size=1024, io_mode="parallel")
    assert ei.value.kind == "write" and ei.value.shard_index == k + 2
    with pytest.raises(rsgpu.RsError):
        rs.encode_stream(_host_streams(data), _host_streams(outs), io_mode=3)


def test_reader_positions_after_errors():
    """The documented positions: after a write error in block i or a failing verify
    of block i, readers have consumed block i+1 (the read-ahead); on success they
    are at the end; the outputs of earlier blocks are always written."""
    announce("stream reader positions")
    k, m, blk = 3, 2, 1024
    rs = rsgpu.ReedSolomon(k, m)
    data = [np.frombuffer(ref_data(4 * blk), np.uint8).copy() for _ in range(k)]
    outs = [np.zeros(4 * blk, np.uint8), np.zeros(blk + 10, np.uint8)]
    readers = _host_streams(data)
    with pytest.raises(rsgpu.StreamError) as ei:
        rs.encode_stream(readers, _host_streams(outs), block_size=blk, io_mode="serial")
    assert ei.value.kind == "write" and ei.value.shard_index == k + 1
    assert [r.pos for r in readers] == [3 * blk] * k  # blocks 0..2: the write of block 1 failed
    exp = oracle_parity(k, m, [d.tobytes() for d in data])
    assert outs[0][:blk].tobytes() == exp[0][:blk]  # block 0 is written everywhere
    par = oracle_parity(k, m, [d.tobytes() for d in data])
    full = data + [np.frombuffer(p, np.uint8).copy() for p in par]
    full[0][10] ^= 1  # block 0 is inconsistent
    readers = _host_streams(full)
    assert not rs.verify_stream(readers, block_size=blk, io_mode="serial")
    assert [r.pos for r in readers] == [2 * blk] * (k + m)
    full[0][10] ^= 1
    readers = _host_streams(full)
    assert rs.verify_stream(readers, block_size=blk)
    assert [r.pos for r in readers] == [4 * blk] * (k + m)


def test_staging_is_reused_across_sizes_and_nested_calls():
    """Calls on one thread reuse (and grow) its staging; a reader that itself runs
    a stream call gets staging of its own."""
    announce("stream staging reuse")
    k, m = 4, 2
    rs = rsgpu.ReedSolomon(k, m)
    rng = np.random.default_rng(11)
    for total, blk in ((5000, 1024), ((16 << 20) + 3, 16 << 20), (70000, 4096)):
        data = [rng.integers(0, 256, total, dtype=np.uint8) for _ in range(k)]
        parity = [np.zeros(total, np.uint8) for _ in range(m)]
        rs.encode_stream(_host_streams(data), _host_streams(parity), block_size=blk)
        assert [p.tobytes() for p in parity] == oracle_parity(k, m, [d.tobytes() for d in data])

    inner_data = [np.frombuffer(ref_data(3000), np.uint8).copy() for _ in range(k)]
    inner_out = [np.zeros(3000, np.uint8) for _ in range(m)]

    class Nested(io.BytesIO):
        def readinto(self, b):
            rs.encode_stream(_host_streams(inner_data), _host_streams(inner_out), block_size=1024)
            return super().readinto(b)

    data = [rng.integers(0, 256, 9000, dtype=np.uint8).tobytes() for _ in range(k)]
    writers = [io.BytesIO() for _ in range(m)]
    rs.encode_stream([Nested(data[0])] + [io.BytesIO(d) for d in data[1:]], writers, block_size=2048)
    assert [w.getvalue() for w in writers] == oracle_parity(k, m, data)
    assert [o.tobytes() for o in inner_out] == oracle_parity(k, m, [d.tobytes() for d in inner_data])


@pytest.mark.parametrize("io_mode", ["serial", "parallel"])
def test_reconstruct_stream_c_entry_presized_and_full(io_mode):
    """rsgpu_reconstruct_stream called through the C ABI with grow == NULL: missing
    vectors sized by the caller receive the rebuilt bytes; present vectors are read
    from position 0 and end at their length; a missing vector without room is a
    write error of that shard (stream.rs:839-900)."""
    import ctypes as C
    announce(f"stream reconstruct C cursors {io_mode}")
    k, m, n = 5, 3, 9000
    data = [ref_data(n)[i:] + ref_data(i) for i in range(k)]
    full = data + oracle_parity(k, m, data)
    rs = rsgpu.ReedSolomon(k, m)
    missing = {0, 3, 6}
    bufs = [bytearray(n) if i in missing else bytearray(full[i]) for i in range(k + m)]
    arrs = [(C.c_uint8 * n).from_buffer(b) for b in bufs]
    vecs = (rsgpu._Vec * (k + m))()
    for i in range(k + m):
        vecs[i].data, vecs[i].cap = C.addressof(arrs[i]), n
        vecs[i].len = 0 if i in missing else n
        vecs[i].pos = 7  # reset to 0 for present shards, kept for missing ones
    opts = rsgpu._stream_options(2048, io_mode)
    err = rsgpu._StreamErr()
    null_grow = rsgpu._GROW_CB()  # a NULL function pointer
    rc = rsgpu._lib.rsgpu_reconstruct_stream(rs._h, vecs, k + m, null_grow, None, C.byref(opts), C.byref(err))
    assert rc == 0, (rc, err.kind, err.shard_index)
    for i in range(k + m):
        assert bytes(bufs[i]) == full[i], i
        assert vecs[i].len == n
        assert vecs[i].pos == (7 if i in missing else n), i
    # a missing vector one block short of room: write error of that shard, the
    # earlier shards' blocks already appended
    small = bytearray(1024)
    arr_small = (C.c_uint8 * 1024).from_buffer(small)
    for i in range(k + m):
        vecs[i].len = 0 if i in missing else n
        vecs[i].pos = 0
    vecs[3].data, vecs[3].cap = C.addressof(arr_small), 1024
    rc = rsgpu._lib.rsgpu_reconstruct_stream(rs._h, vecs, k + m, null_grow, None, C.byref(opts), C.byref(err))
    assert rc == 105 and err.kind == 2 and err.shard_index == 3
    assert vecs[3].len == 0  # not even the first block fits
    assert vecs[0].len >= 2048 and bytes(bufs[0][:2048]) == full[0][:2048]


def test_reconstruct_stream_c_entry_errors():
    """Count errors, the Leopard refusal, the empty layout and unequal present
    lengths of rsgpu_reconstruct_stream, in the reference's order."""
    import ctypes as C
    announce("stream reconstruct C cursor errors")
    k, m = 3, 2
    rs = rsgpu.ReedSolomon(k, m)
    err = rsgpu._StreamErr()
    grow = rsgpu._GROW_CB(lambda *a: -1)

    def call(codec, vecs, n):
        return rsgpu._lib.rsgpu_reconstruct_stream(codec._h, vecs, n, grow, None, None, C.byref(err))

    vecs = (rsgpu._Vec * 6)()
    assert call(rs, vecs, 4) == 1 and err.kind == 3 and err.shard_index == 0  # TooFewShards
    assert call(rs, vecs, 6) == 2  # TooManyShards
    assert call(rs, vecs, 5) == 0  # nothing present: nothing to do
    leo = rsgpu.ReedSolomon(k, m, codec_family="leopard_gf8")
    assert call(leo, vecs, 5) == 15  # UnsupportedCodecFamily, before presence is looked at
    b = [bytearray(b"abcd"), bytearray(b"abcd"), bytearray(b"abc")]
    arrs = [(C.c_uint8 * len(x)).from_buffer(x) for x in b]
    for i in range(3):
        vecs[i].data, vecs[i].len, vecs[i].cap, vecs[i].pos = C.addressof(arrs[i]), len(b[i]), len(b[i]), 0
    vecs[3] = rsgpu._Vec()
    vecs[4] = rsgpu._Vec()
    assert call(rs, vecs, 5) == 9 and err.shard_index == 2  # IncorrectShardSize at the short shard
    vecs[2] = rsgpu._Vec()
    assert call(rs, vecs, 5) == 10  # TooFewShardsPresent
