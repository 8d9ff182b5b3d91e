"""CPU checks of argument validation at the C ABI: calls whose arguments fail the
reference's checks return its Error variant before any device work, so they run
without a GPU (shard pointers are never dereferenced here).  Follows
validate_leopard_decode_buffers (src/core/reconstruct.rs:140-185) and
encode_with_tables (src/core/leopard_gf8/encode.rs:201-210)."""
import ctypes as C
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "rustfs-erasure-codec_amd", "rsgpu", "librsgpu.so")


class Slice(C.Structure):
    _fields_ = [("ptr", C.c_void_p), ("len", C.c_size_t)]


@pytest.fixture(scope="module")
def lib():
    lib = C.CDLL(LIB)
    lib.rsgpu_status_name.restype = C.c_char_p
    lib.rsgpu_leopard_encode.argtypes = [C.c_int32, C.c_size_t, C.c_size_t, C.POINTER(Slice), C.c_size_t,
                                         C.POINTER(Slice), C.c_size_t, C.c_void_p]
    lib.rsgpu_leopard_reconstruct.argtypes = [C.c_int32, C.c_size_t, C.c_size_t, C.POINTER(C.c_uint8),
                                              C.POINTER(Slice), C.POINTER(Slice), C.c_void_p]
    return lib


def _name(lib, rc):
    return lib.rsgpu_status_name(rc).decode()


def _slices(n, length, base=0x1000):
    return (Slice * max(1, n))(*[Slice(base + 0x100000 * i, length) for i in range(n)])


def test_leopard_encode_counts_and_length(lib):
    k, m = 10, 4
    assert _name(lib, lib.rsgpu_leopard_encode(8, k, m, _slices(k - 1, 64), k - 1, _slices(m, 64), m,
                                               None)) == "TooFewShards"
    assert _name(lib, lib.rsgpu_leopard_encode(8, k, m, _slices(k, 64), k, _slices(m + 1, 64), m + 1,
                                               None)) == "TooFewShards"
    assert _name(lib, lib.rsgpu_leopard_encode(8, k, m, _slices(k, 100), k, _slices(m, 100), m,
                                               None)) == "IncorrectShardSize"
    assert _name(lib, lib.rsgpu_leopard_encode(16, k, m, _slices(k, 0), k, _slices(m, 0), m,
                                               None)) == "IncorrectShardSize"


def _reconstruct(lib, k, m, present, inputs, outputs):
    pres = (C.c_uint8 * (k + m))(*[1 if p else 0 for p in present])
    ins = (Slice * (k + m))(*inputs)
    outs = (Slice * (k + m))(*outputs)
    return _name(lib, lib.rsgpu_leopard_reconstruct(8, k, m, pres, ins, outs, None))


def test_leopard_reconstruct_validation_order(lib):
    k, m, n = 6, 3, 128
    ok_in = [Slice(0x1000 + 0x10000 * i, n) for i in range(k + m)]
    none = [Slice(0, 0)] * (k + m)
    present = [True] * k + [False] * m
    assert _name(lib, lib.rsgpu_leopard_reconstruct(8, 0, m, (C.c_uint8 * 1)(1), (Slice * 1)(),
                                                    (Slice * 1)(), None)) == "TooFewDataShards"
    assert _name(lib, lib.rsgpu_leopard_reconstruct(8, k, 0, (C.c_uint8 * 1)(1), (Slice * 1)(),
                                                    (Slice * 1)(), None)) == "TooFewParityShards"
    # a present shard without data, an empty one, a length off the 64-byte grid
    missing_ptr = list(ok_in)
    missing_ptr[0] = Slice(0, n)
    assert _reconstruct(lib, k, m, present, missing_ptr, none) == "TooFewShardsPresent"
    empty = list(ok_in)
    empty[1] = Slice(0x1000, 0)
    assert _reconstruct(lib, k, m, present, empty, none) == "EmptyShard"
    odd = list(ok_in)
    odd[2] = Slice(0x1000, n + 1)
    assert _reconstruct(lib, k, m, present, odd, none) == "IncorrectShardSize"
    # fewer than k present
    assert _reconstruct(lib, k, m, [True] * (k - 1) + [False] * (m + 1), ok_in, none) == \
        "TooFewShardsPresent"
    # an output whose length differs from the shards'
    outs = list(none)
    outs[k] = Slice(0x9000, n - 64)
    assert _reconstruct(lib, k, m, present, ok_in, outs) == "IncorrectShardSize"
    # nothing requested: Ok without touching the device (decode.rs:305-309)
    assert _reconstruct(lib, k, m, present, ok_in, none) == "Ok"
    assert _reconstruct(lib, k, m, [True] * (k + m), ok_in, none) == "Ok"
