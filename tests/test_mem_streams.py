"""CPU checks of the library's in-memory stream callbacks (rsgpu_mem_stream_read /
rsgpu_mem_stream_write), the C form of the io::Cursor streams of the reference's
stream tests (src/core/stream.rs:839-846), and of StreamOptions::default()
(stream.rs:73-80).  Host code only: no device calls."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "rustfs-erasure-codec_amd", "rsgpu", "librsgpu.so")


class MemStream(C.Structure):
    _fields_ = [("data", C.c_void_p), ("len", C.c_size_t), ("pos", C.c_size_t)]


class MemStreams(C.Structure):
    _fields_ = [("readers", C.POINTER(MemStream)), ("writers", C.POINTER(MemStream))]


class StreamOptions(C.Structure):
    _fields_ = [("block_size", C.c_size_t), ("io_mode", C.c_int32), ("reserved", C.c_int32)]


def _lib():
    lib = C.CDLL(LIB)
    for f in (lib.rsgpu_mem_stream_read, lib.rsgpu_mem_stream_write):
        f.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]
        f.restype = C.c_int64
    lib.rsgpu_stream_options_default.argtypes = [C.POINTER(StreamOptions)]
    lib.rsgpu_stream_options_default.restype = None
    return lib


def test_stream_options_default_is_the_reference_default():
    o = StreamOptions(0, 7, 7)
    _lib().rsgpu_stream_options_default(C.byref(o))
    assert (o.block_size, o.io_mode, o.reserved) == (4 << 20, 0, 0)  # 4 MiB, Auto


def test_reads_advance_to_the_end_then_return_zero():
    lib = _lib()
    a = np.arange(1000, dtype=np.uint8)
    b = np.arange(10, dtype=np.uint8)[::-1].copy()
    readers = (MemStream * 2)(MemStream(a.ctypes.data, a.size, 0), MemStream(b.ctypes.data, b.size, 3))
    ms = MemStreams(readers, None)
    out = np.zeros(600, np.uint8)
    assert lib.rsgpu_mem_stream_read(C.byref(ms), 0, out.ctypes.data, 600) == 600
    assert (out == a[:600]).all() and readers[0].pos == 600
    assert lib.rsgpu_mem_stream_read(C.byref(ms), 0, out.ctypes.data, 600) == 400
    assert (out[:400] == a[600:]).all() and readers[0].pos == 1000
    assert lib.rsgpu_mem_stream_read(C.byref(ms), 0, out.ctypes.data, 600) == 0
    # a reader that starts mid-buffer (a cursor not at 0) reads from its position
    assert lib.rsgpu_mem_stream_read(C.byref(ms), 1, out.ctypes.data, 600) == 7
    assert (out[:7] == b[3:]).all()
    assert lib.rsgpu_mem_stream_read(C.byref(ms), 1, out.ctypes.data, 0) == 0


def test_writes_fill_the_buffer_then_accept_nothing():
    lib = _lib()
    dst = np.zeros(100, np.uint8)
    writers = (MemStream * 3)()
    writers[2] = MemStream(dst.ctypes.data, dst.size, 0)
    ms = MemStreams(None, writers)
    src = np.arange(70, dtype=np.uint8)
    assert lib.rsgpu_mem_stream_write(C.byref(ms), 2, src.ctypes.data, 70) == 70
    assert lib.rsgpu_mem_stream_write(C.byref(ms), 2, src.ctypes.data, 70) == 30
    assert writers[2].pos == 100
    assert (dst[:70] == src).all() and (dst[70:] == src[:30]).all()
    # full: 0 bytes, which the stream calls turn into a write error (WriteZero)
    assert lib.rsgpu_mem_stream_write(C.byref(ms), 2, src.ctypes.data, 70) == 0
