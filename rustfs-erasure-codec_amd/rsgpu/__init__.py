"""rsgpu -- Python mirror of the rustfs-erasure-codec ``ReedSolomon`` API over
the MI355X C ABI (include/rsgpu.h, librsgpu.so).

Shards are device tensors (``torch.uint8`` on ``cuda``), one per shard, as the
reference takes one slice per shard.  Every call validates its arguments in
the reference's order before any device work is enqueued and raises
``RsError`` carrying the reference's ``Error`` variant name.

There is no CPU fallback: importing this module fails if librsgpu.so has not
been built (``python rustfs-erasure-codec_amd/build.py``).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librsgpu.so")
if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: build it with rustfs-erasure-codec_amd/build.py")

# torch ships its own HIP runtime; loading it first makes librsgpu.so bind to the
# same runtime, so device pointers and streams are shared with torch tensors.
try:
    import torch  # noqa: F401
except ImportError:
    pass

_lib = C.CDLL(LIB_PATH)

ERROR_NAMES = [
    "Ok", "TooFewShards", "TooManyShards", "TooFewDataShards", "TooManyDataShards",
    "TooFewParityShards", "TooManyParityShards", "TooFewBufferShards", "TooManyBufferShards",
    "IncorrectShardSize", "TooFewShardsPresent", "EmptyShard", "InvalidShardFlags", "InvalidIndex",
    "InvalidCustomMatrix", "UnsupportedCodecFamily", "UnsupportedLeopardPrototype",
]

FIELDS = {"gf8": 0, "galois_8": 0, "gf16": 1, "galois_16": 1}
FAMILIES = {"classic": 0, "leopard_gf8": 1, "leopard_gf16": 2}
MATRIX_MODES = {"vandermonde": 0, "cauchy": 1, "jerasure_like": 2, "custom": 3}
LEOPARD_MODES = {"disabled": 0, "as_needed": 1, "prefer_gf16": 2, "prefer_leopard": 3}


class RsError(Exception):
    """A validation or device error; ``name`` is the reference Error variant."""

    def __init__(self, code):
        self.code = code
        self.name = _lib.rsgpu_status_name(code).decode()
        detail = _lib.rsgpu_last_hip_error().decode() if code >= 100 else ""
        super().__init__(f"{self.name}" + (f" ({detail})" if detail else ""))


class StreamError(Exception):
    """StreamError of the block-stream calls (src/core/stream.rs:107-171): `kind` is
    "read", "write" or "codec", `shard_index` the shard, `error` the exception the
    reader/writer raised (or its negative return value), or the codec RsError."""

    def __init__(self, kind, shard_index, error):
        self.kind, self.shard_index, self.error = kind, shard_index, error
        super().__init__(f"{kind} error on shard {shard_index}: {error}")


class SBSError(Exception):
    """SBSError (src/errors.rs): "TooManyCalls", "LeftoverShards", or "RSError" with
    the codec's RsError in `error`."""

    def __init__(self, kind, error=None):
        self.kind, self.error = kind, error
        super().__init__(kind if error is None else f"{kind}({error})")


class Options(C.Structure):
    _fields_ = [
        ("field", C.c_int32), ("codec_family", C.c_int32), ("matrix_mode", C.c_int32),
        ("leopard_mode", C.c_int32), ("fast_one_parity", C.c_int32), ("inversion_cache", C.c_int32),
        ("inversion_cache_capacity", C.c_uint64), ("inversion_cache_max_bytes", C.c_uint64),
    ]


class CSlice(C.Structure):
    _fields_ = [("ptr", C.c_void_p), ("len", C.c_size_t)]


class Slice(C.Structure):
    _fields_ = [("ptr", C.c_void_p), ("len", C.c_size_t)]


class Slot(C.Structure):
    _fields_ = [("ptr", C.c_void_p), ("len", C.c_size_t), ("present", C.c_int32)]


_vp = C.c_void_p
_sz = C.c_size_t
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)
_lib.rsgpu_options_default.argtypes = [C.POINTER(Options)]
_lib.rsgpu_codec_new.argtypes = [_sz, _sz, C.POINTER(Options), C.POINTER(_vp)]
_lib.rsgpu_codec_with_custom_matrix.argtypes = [_sz, _sz, _u8p, _sz, _sz, C.POINTER(Options), C.POINTER(_vp)]
_lib.rsgpu_codec_free.argtypes = [_vp]
_lib.rsgpu_codec_info.argtypes = [_vp, C.POINTER(_sz), C.POINTER(_sz), _i32p, _i32p]
_lib.rsgpu_codec_matrix.argtypes = [_vp, _u8p, _sz]
_lib.rsgpu_encode.argtypes = [_vp, C.POINTER(Slice), _sz, _vp]
_lib.rsgpu_encode_sep.argtypes = [_vp, C.POINTER(CSlice), _sz, C.POINTER(Slice), _sz, _vp]
_lib.rsgpu_encode_sep_gf2mfma.argtypes = [_vp, C.POINTER(CSlice), _sz, C.POINTER(Slice), _sz, _vp]
_lib.rsgpu_verify.argtypes = [_vp, C.POINTER(CSlice), _sz, _i32p, _vp]
_lib.rsgpu_verify_async.argtypes = [_vp, C.POINTER(CSlice), _sz, _vp, _vp]
_lib.rsgpu_verify_with_buffer.argtypes = [_vp, C.POINTER(CSlice), _sz, C.POINTER(Slice), _sz, _i32p, _vp]
_lib.rsgpu_reconstruct.argtypes = [_vp, C.POINTER(Slot), _sz, _vp]
_lib.rsgpu_reconstruct_data.argtypes = [_vp, C.POINTER(Slot), _sz, _vp]
_lib.rsgpu_reconstruct_some.argtypes = [_vp, C.POINTER(Slot), _sz, _u8p, _sz, _vp]
_lib.rsgpu_prepare_reconstruct_workspace.argtypes = [_vp, C.POINTER(Slot), _sz, _vp, C.POINTER(_vp)]
_lib.rsgpu_prepare_reconstruct_data_workspace.argtypes = [_vp, C.POINTER(Slot), _sz, _vp, C.POINTER(_vp)]
_lib.rsgpu_prepare_reconstruct_some_workspace.argtypes = [_vp, C.POINTER(Slot), _sz, _u8p, _sz, _vp,
                                                          C.POINTER(_vp)]
_lib.rsgpu_reconstruct_with_workspace.argtypes = [_vp, C.POINTER(Slot), _sz, _vp, _vp]
_lib.rsgpu_reconstruct_workspace_targets.argtypes = [_vp, C.POINTER(_sz), _sz, C.POINTER(_sz), C.POINTER(_sz)]
_lib.rsgpu_reconstruct_workspace_free.argtypes = [_vp]
_lib.rsgpu_decode_idx.argtypes = [_vp, C.POINTER(Slice), _sz, _u8p, _sz, C.POINTER(CSlice), _sz, _vp]
_lib.rsgpu_reconstruct_noalloc.argtypes = [_vp, C.POINTER(Slot), _sz, C.c_int32, _vp]
_lib.rsgpu_encode_host.argtypes = [_vp, C.POINTER(_vp), _sz, _sz]
_lib.rsgpu_verify_host.argtypes = [_vp, C.POINTER(_vp), _sz, _sz, _i32p]
_lib.rsgpu_reconstruct_host.argtypes = [_vp, C.POINTER(_vp), _u8p, _sz, _sz, C.c_int32]
_lib.rsgpu_apply_rows.argtypes = [C.c_int32, _u8p, _sz, _sz, C.POINTER(CSlice), C.POINTER(Slice), _vp]
_lib.rsgpu_leopard_encode.argtypes = [C.c_int32, _sz, _sz, C.POINTER(CSlice), _sz, C.POINTER(Slice), _sz, _vp]
_lib.rsgpu_leopard_reconstruct.argtypes = [C.c_int32, _sz, _sz, _u8p, C.POINTER(CSlice), C.POINTER(Slice), _vp]
_lib.rsgpu_cache_stats.argtypes = [_vp] + [C.POINTER(C.c_uint64)] * 4
_lib.rsgpu_codec_format_id.argtypes = [_vp, _u8p, _sz, C.POINTER(_sz)]
_lib.rsgpu_split_len.argtypes = [_vp, _sz, C.POINTER(_sz)]
_lib.rsgpu_split.argtypes = [_vp, _vp, _sz, C.POINTER(Slice), _sz, _vp]
_lib.rsgpu_join.argtypes = [_vp, C.POINTER(CSlice), _sz, _sz, _vp, C.POINTER(_sz), _vp]
_lib.rsgpu_encode_single.argtypes = [_vp, _sz, C.POINTER(Slice), _sz, _vp]
_lib.rsgpu_encode_single_sep.argtypes = [_vp, _sz, CSlice, C.POINTER(Slice), _sz, _vp]
_lib.rsgpu_update.argtypes = [_vp, C.POINTER(CSlice), _sz, C.POINTER(CSlice), _sz, C.POINTER(Slice), _sz, _vp]
_lib.rsgpu_status_name.restype = C.c_char_p
_lib.rsgpu_status_name.argtypes = [C.c_int]
_lib.rsgpu_last_hip_error.restype = C.c_char_p
_lib.rsgpu_init.argtypes = [C.c_int]


def _check(rc):
    if rc != 0:
        raise RsError(rc)


def lib():
    return _lib


def _stream(stream):
    if stream is not None:
        return C.c_void_p(stream if isinstance(stream, int) else stream.cuda_stream)
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _torch_uint8():
    import torch

    return torch.uint8


def _ptr_len(t):
    """(device pointer, byte length) of a uint8 device tensor view."""
    if t is None:
        return 0, 0
    if not t.is_cuda:
        raise ValueError("shards must be device (cuda) tensors; host pointers would reach a kernel")
    if t.dtype != _torch_uint8():
        raise ValueError(f"shards must be uint8 tensors, not {t.dtype}")
    if not t.is_contiguous():
        raise ValueError("shard views must be contiguous")
    return t.data_ptr(), t.numel()


class ShardSlot:
    """A shard buffer kept across calls with a presence flag (src/lib.rs:250-304)."""

    def __init__(self, data, present=True):
        self.data = data
        self.present = bool(present)

    @classmethod
    def new_present(cls, data):
        return cls(data, True)

    @classmethod
    def new_missing(cls, data):
        return cls(data, False)

    def is_present(self):
        return self.present

    def mark_missing(self):
        self.present = False

    def mark_present(self):
        self.present = True


class ReconstructWorkspace:
    """ReconstructWorkspace (src/core/workspace.rs:85-140): a prepared Classic plan for
    one erasure layout, with its device decode rows."""

    def __init__(self, handle):
        self._h = handle
        cap = 65536
        buf = (_sz * cap)()
        nd, np_ = _sz(), _sz()
        _check(_lib.rsgpu_reconstruct_workspace_targets(handle, buf, cap, C.byref(nd), C.byref(np_)))
        self._data = [buf[i] for i in range(nd.value)]
        self._parity = [buf[nd.value + i] for i in range(np_.value)]

    def data_targets(self):
        return list(self._data)

    def parity_targets(self):
        return list(self._parity)

    def __del__(self, _free=_lib.rsgpu_reconstruct_workspace_free):  # bound now: globals go at exit
        if getattr(self, "_h", None):
            _free(self._h)
            self._h = None


class ReedSolomon:
    """ReedSolomon<F> with CodecOptions (src/core/mod.rs:89, src/core/options.rs:99-147)."""

    def __init__(self, data_shards, parity_shards, *, field="gf8", codec_family="classic",
                 matrix_mode="vandermonde", leopard_mode="disabled", fast_one_parity=False,
                 inversion_cache=True, inversion_cache_capacity=0, inversion_cache_max_bytes=0,
                 _custom=None):
        opts = Options()
        _lib.rsgpu_options_default(C.byref(opts))
        opts.field = FIELDS[field]
        opts.codec_family = FAMILIES[codec_family]
        opts.matrix_mode = MATRIX_MODES[matrix_mode]
        opts.leopard_mode = LEOPARD_MODES[leopard_mode]
        opts.fast_one_parity = int(bool(fast_one_parity))
        opts.inversion_cache = int(bool(inversion_cache))
        opts.inversion_cache_capacity = inversion_cache_capacity
        opts.inversion_cache_max_bytes = inversion_cache_max_bytes
        h = C.c_void_p()
        self._h = None
        if _custom is not None:
            rows = np.ascontiguousarray(_custom, np.uint8)
            eb = 2 if opts.field == 1 else 1
            n_rows = rows.shape[0]
            row_len = rows.shape[1] // eb if rows.ndim == 2 else 0
            rc = _lib.rsgpu_codec_with_custom_matrix(data_shards, parity_shards,
                                                     rows.ctypes.data_as(_u8p), n_rows, row_len,
                                                     C.byref(opts), C.byref(h))
        else:
            rc = _lib.rsgpu_codec_new(data_shards, parity_shards, C.byref(opts), C.byref(h))
        _check(rc)
        self._h = h
        k, m, f, fam = _sz(), _sz(), C.c_int32(), C.c_int32()
        _lib.rsgpu_codec_info(h, C.byref(k), C.byref(m), C.byref(f), C.byref(fam))
        self._k, self._m, self._field, self._family = k.value, m.value, f.value, fam.value

    @classmethod
    def new(cls, data_shards, parity_shards):
        return cls(data_shards, parity_shards)

    @classmethod
    def with_custom_matrix(cls, data_shards, parity_shards, rows, **opts):
        """rows: (n_rows, row_len * elem_bytes) uint8 array, GF16 elements as [x, c] pairs."""
        return cls(data_shards, parity_shards, matrix_mode="vandermonde", _custom=rows, **opts)

    def __del__(self, _free=_lib.rsgpu_codec_free):  # bound now: module globals are gone at exit
        if getattr(self, "_h", None):
            _free(self._h)
            self._h = None

    # ---- shape
    def data_shard_count(self):
        return self._k

    def parity_shard_count(self):
        return self._m

    def total_shard_count(self):
        return self._k + self._m

    def codec_family(self):
        return {0: "classic", 1: "leopard_gf8", 2: "leopard_gf16"}[self._family]

    def field(self):
        return "gf16" if self._field == 1 else "gf8"

    def matrix(self):
        eb = 2 if self._field == 1 else 1
        out = np.zeros(((self._k + self._m) * self._k * eb,), np.uint8)
        _check(_lib.rsgpu_codec_matrix(self._h, out.ctypes.data_as(_u8p), out.size))
        if eb == 2:
            out = (out[0::2].astype(np.uint16) << 8) | out[1::2]
        return out.reshape(self._k + self._m, self._k)

    # ---- encode
    def encode(self, shards, stream=None):
        arr = (Slice * len(shards))(*[Slice(*_ptr_len(s)) for s in shards])
        _check(_lib.rsgpu_encode(self._h, arr, len(shards), _stream(stream)))

    def encode_sep(self, data, parity, stream=None):
        d = (CSlice * len(data))(*[CSlice(*_ptr_len(s)) for s in data])
        p = (Slice * len(parity))(*[Slice(*_ptr_len(s)) for s in parity])
        _check(_lib.rsgpu_encode_sep(self._h, d, len(data), p, len(parity), _stream(stream)))

    def encode_sep_gf2mfma(self, data, parity, stream=None):
        """encode_sep through the int8 matrix-core formulation (bit-expanded GF(2)
        product); classic GF(2^8), k <= 31, m <= 8, lengths a multiple of 4."""
        d = (CSlice * len(data))(*[CSlice(*_ptr_len(s)) for s in data])
        p = (Slice * len(parity))(*[Slice(*_ptr_len(s)) for s in parity])
        _check(_lib.rsgpu_encode_sep_gf2mfma(self._h, d, len(data), p, len(parity), _stream(stream)))

    # ---- verify
    def verify(self, shards, stream=None):
        arr = (CSlice * len(shards))(*[CSlice(*_ptr_len(s)) for s in shards])
        ok = C.c_int32(0)
        _check(_lib.rsgpu_verify(self._h, arr, len(shards), C.byref(ok), _stream(stream)))
        return bool(ok.value)

    def verify_with_workspace(self, shards, workspace=None, stream=None):
        """The classic GPU verify compares in registers and needs no workspace."""
        return self.verify(shards, stream)

    def verify_async(self, shards, device_flag, stream=None):
        arr = (CSlice * len(shards))(*[CSlice(*_ptr_len(s)) for s in shards])
        _check(_lib.rsgpu_verify_async(self._h, arr, len(shards), device_flag.data_ptr(), _stream(stream)))

    def verify_with_buffer(self, shards, buffer, stream=None):
        arr = (CSlice * len(shards))(*[CSlice(*_ptr_len(s)) for s in shards])
        buf = (Slice * len(buffer))(*[Slice(*_ptr_len(s)) for s in buffer])
        ok = C.c_int32(0)
        _check(_lib.rsgpu_verify_with_buffer(self._h, arr, len(shards), buf, len(buffer),
                                             C.byref(ok), _stream(stream)))
        return bool(ok.value)

    # ---- reconstruct
    def _slots(self, shards):
        """Accepts a list of Optional[tensor] or of ShardSlot; returns (ctypes slots, fixups)."""
        import torch

        shard_len, device = None, None
        for s in shards:
            t = s.data if isinstance(s, ShardSlot) else s
            present = s.present if isinstance(s, ShardSlot) else s is not None
            if present and t is not None:
                shard_len = t.numel()
                device = t.device
                break
        slots = (Slot * len(shards))()
        allocated = {}
        for i, s in enumerate(shards):
            if isinstance(s, ShardSlot):
                p, n = _ptr_len(s.data)
                slots[i] = Slot(p, n if s.present else n, int(s.present))
            elif s is not None:
                p, n = _ptr_len(s)
                slots[i] = Slot(p, n, 1)
            else:
                if shard_len is not None:
                    t = torch.empty(shard_len, dtype=torch.uint8, device=device)
                    allocated[i] = t
                    slots[i] = Slot(t.data_ptr(), shard_len, 0)
                else:
                    slots[i] = Slot(None, 0, 0)
        return slots, allocated

    def _finish(self, shards, slots, allocated):
        for i, s in enumerate(shards):
            if isinstance(s, ShardSlot):
                s.present = bool(slots[i].present)
            elif s is None and i in allocated and slots[i].present:
                shards[i] = allocated[i]

    def reconstruct(self, shards, stream=None):
        slots, alloc = self._slots(shards)
        _check(_lib.rsgpu_reconstruct(self._h, slots, len(shards), _stream(stream)))
        self._finish(shards, slots, alloc)

    def reconstruct_data(self, shards, stream=None):
        slots, alloc = self._slots(shards)
        _check(_lib.rsgpu_reconstruct_data(self._h, slots, len(shards), _stream(stream)))
        self._finish(shards, slots, alloc)

    def reconstruct_some(self, shards, required, stream=None):
        slots, alloc = self._slots(shards)
        req = np.ascontiguousarray(required, np.uint8)
        _check(_lib.rsgpu_reconstruct_some(self._h, slots, len(shards), req.ctypes.data_as(_u8p),
                                           req.size, _stream(stream)))
        self._finish(shards, slots, alloc)

    # ---- prepared reconstruction (src/core/reconstruct.rs:539-700)
    @staticmethod
    def _layout(shards):
        """Slots describing the erasure layout only (missing entries need no buffer)."""
        slots = (Slot * len(shards))()
        for i, s in enumerate(shards):
            if isinstance(s, ShardSlot):
                p, n = _ptr_len(s.data)
                slots[i] = Slot(p, n, int(s.present))
            elif s is not None:
                p, n = _ptr_len(s)
                slots[i] = Slot(p, n, 1)
            else:
                slots[i] = Slot(None, 0, 0)
        return slots

    def _prepare(self, fn, shards, *extra, stream=None):
        h = _vp()
        _check(fn(self._h, self._layout(shards), len(shards), *extra, _stream(stream), C.byref(h)))
        return ReconstructWorkspace(h)

    def prepare_reconstruct_workspace(self, shards, stream=None):
        return self._prepare(_lib.rsgpu_prepare_reconstruct_workspace, shards, stream=stream)

    def prepare_reconstruct_data_workspace(self, shards, stream=None):
        return self._prepare(_lib.rsgpu_prepare_reconstruct_data_workspace, shards, stream=stream)

    def prepare_reconstruct_some_workspace(self, shards, required, stream=None):
        req = np.ascontiguousarray([1 if r else 0 for r in required], np.uint8)
        return self._prepare(_lib.rsgpu_prepare_reconstruct_some_workspace, shards,
                             req.ctypes.data_as(_u8p), req.size, stream=stream)

    def reconstruct_with_workspace(self, shards, workspace, stream=None):
        slots, alloc = self._slots(shards)
        _check(_lib.rsgpu_reconstruct_with_workspace(self._h, slots, len(shards), workspace._h,
                                                     _stream(stream)))
        self._finish(shards, slots, alloc)

    reconstruct_with_workspace_par = reconstruct_with_workspace

    def decode_idx(self, dst, expect_input, input, stream=None):
        """Progressive recovery (src/galois_8/policy.rs:543-704).  dst / input: lists of
        k+m Optional device tensors; expect_input: list of k+m bools, or None to merge
        partial results by XOR.  Destinations accumulate in place."""
        d = (Slice * max(1, len(dst)))(*[Slice(*_ptr_len(t)) if t is not None else Slice(None, 0)
                                         for t in dst])
        i = (CSlice * max(1, len(input)))(*[CSlice(*_ptr_len(t)) if t is not None else CSlice(None, 0)
                                            for t in input])
        for arr, lst in ((d, dst), (i, input)):  # a Some of length 0 still needs a non-NULL ptr
            for j, t in enumerate(lst):
                if t is not None and t.numel() == 0:
                    arr[j].ptr = 1
        if expect_input is None:
            ep, ne = None, 0
        else:
            ex = np.ascontiguousarray([1 if x else 0 for x in expect_input], np.uint8)
            ep, ne = ex.ctypes.data_as(_u8p), ex.size
        _check(_lib.rsgpu_decode_idx(self._h, d, len(dst), ep, ne, i, len(input), _stream(stream)))

    def reconstruct_noalloc(self, shards, data_only=False, stream=None):
        """reconstruct / reconstruct_data through the entry point that allocates
        nothing (decode solve in the library's per-device scratch).  Pass ShardSlots
        or tensors: a None entry would need an allocation here, so it is refused."""
        if any(s is None for s in shards):
            raise ValueError("reconstruct_noalloc needs a buffer for every shard (use ShardSlot)")
        slots, alloc = self._slots(shards)
        _check(_lib.rsgpu_reconstruct_noalloc(self._h, slots, len(shards), int(bool(data_only)),
                                              _stream(stream)))
        self._finish(shards, slots, alloc)

    # ---- host buffers (PCIe-inclusive path)
    def encode_host(self, shards):
        arrs = [np.ascontiguousarray(s, np.uint8) for s in shards]
        ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        _check(_lib.rsgpu_encode_host(self._h, ptrs, len(arrs), arrs[0].size))
        for s, a in zip(shards, arrs):
            if s is not a:
                s[...] = a

    def verify_host(self, shards):
        arrs = [np.ascontiguousarray(s, np.uint8) for s in shards]
        ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        ok = C.c_int32(0)
        _check(_lib.rsgpu_verify_host(self._h, ptrs, len(arrs), arrs[0].size, C.byref(ok)))
        return bool(ok.value)

    def reconstruct_host(self, shards, present, data_only=False):
        arrs = [np.ascontiguousarray(s, np.uint8) for s in shards]
        ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        pres = np.ascontiguousarray(present, np.uint8).copy()
        _check(_lib.rsgpu_reconstruct_host(self._h, ptrs, pres.ctypes.data_as(_u8p), len(arrs),
                                           arrs[0].size, int(data_only)))
        return arrs, pres

    def cache_stats(self):
        vals = [C.c_uint64() for _ in range(4)]
        _check(_lib.rsgpu_cache_stats(self._h, *[C.byref(v) for v in vals]))
        return dict(zip(["requests", "hits", "misses", "entries"], [v.value for v in vals]))

    # ---- format identity, split, join (src/core/format.rs, src/core/codec.rs:476-538)
    def stable_id_bytes(self):
        n = _sz()
        _check(_lib.rsgpu_codec_format_id(self._h, None, 0, C.byref(n)))
        buf = (C.c_uint8 * n.value)()
        _check(_lib.rsgpu_codec_format_id(self._h, buf, n.value, C.byref(n)))
        return bytes(buf)

    def split_len(self, data_len):
        n = _sz()
        _check(_lib.rsgpu_split_len(self._h, data_len, C.byref(n)))
        return n.value

    def split(self, data, stream=None):
        """k zero-padded device shards holding the bytes of the uint8 device tensor `data`."""
        import torch

        n = data.numel()
        L = self.split_len(n)
        shards = [torch.empty(L, dtype=torch.uint8, device=data.device) for _ in range(self._k)]
        arr = (Slice * self._k)(*[Slice(t.data_ptr(), L) for t in shards])
        _check(_lib.rsgpu_split(self._h, data.data_ptr() if n else None, n, arr, self._k, _stream(stream)))
        return shards

    def join(self, shards, out_len, stream=None):
        import torch

        arr = (CSlice * len(shards))(*[CSlice(*_ptr_len(t)) for t in shards])
        cap = min(out_len, sum(t.numel() for t in shards))
        out = torch.empty(max(cap, 1), dtype=torch.uint8, device=shards[0].device if shards else "cuda")
        w = _sz()
        _check(_lib.rsgpu_join(self._h, arr, len(shards), out_len, out.data_ptr(), C.byref(w),
                               _stream(stream)))
        return out[:w.value]

    # ---- incremental encode (src/core/encode.rs:892-1132)
    def encode_single(self, i_data, shards, stream=None):
        arr = (Slice * len(shards))(*[Slice(*_ptr_len(t)) for t in shards])
        _check(_lib.rsgpu_encode_single(self._h, i_data, arr, len(shards), _stream(stream)))

    def encode_single_sep(self, i_data, single, parity, stream=None):
        par = (Slice * len(parity))(*[Slice(*_ptr_len(t)) for t in parity])
        _check(_lib.rsgpu_encode_single_sep(self._h, i_data, CSlice(*_ptr_len(single)), par,
                                            len(parity), _stream(stream)))

    def update(self, old_data, new_data, parity, stream=None):
        """new_data[i] is None for an unchanged shard."""
        old = (CSlice * len(old_data))(*[CSlice(*_ptr_len(t)) for t in old_data])
        new = (CSlice * len(new_data))(*[CSlice(*_ptr_len(t)) for t in new_data])
        par = (Slice * len(parity))(*[Slice(*_ptr_len(t)) for t in parity])
        _check(_lib.rsgpu_update(self._h, old, len(old_data), new, len(new_data), par, len(parity),
                                 _stream(stream)))

    # ---- block streams (src/core/stream.rs:677-1031)
    # block_size and io_mode are the reference's StreamOptions (stream.rs:49-104);
    # io_mode "auto" | "serial" | "parallel".  With HostStream readers and writers
    # the library's native in-memory callbacks run (no Python per block);
    # other streams go through Python callbacks (readinto/read, write).
    def encode_stream(self, data, parity, block_size=4 << 20, io_mode="auto"):
        """Reads the k data streams block by block and writes the m parity streams."""
        io = _stream_io(list(data), list(parity))
        err = _StreamErr()
        rc = _lib.rsgpu_encode_stream(self._h, len(io.readers), len(io.writers), io.rd, io.wr, io.user,
                                      C.byref(_stream_options(block_size, io_mode)), C.byref(err))
        io.finish()
        io.raise_for(rc, err, writer_base=self._k)

    def verify_stream(self, shards, block_size=4 << 20, io_mode="auto"):
        io = _stream_io(list(shards), [])
        err = _StreamErr()
        ok = C.c_int32()
        rc = _lib.rsgpu_verify_stream(self._h, len(io.readers), io.rd, io.user,
                                      C.byref(_stream_options(block_size, io_mode)), C.byref(ok),
                                      C.byref(err))
        io.finish()
        io.raise_for(rc, err)
        return bool(ok.value)

    def reconstruct_stream_from(self, readers, writers, block_size=4 << 20, io_mode="auto"):
        """readers[i] / writers[i]: a stream or None (the reference's Option layout)."""
        readers, writers = list(readers), list(writers)
        io = _stream_io(readers, writers)
        hr = (C.c_uint8 * max(1, len(readers)))(*[r is not None for r in readers])
        hw = (C.c_uint8 * max(1, len(writers)))(*[w is not None for w in writers])
        err = _StreamErr()
        rc = _lib.rsgpu_reconstruct_stream_from(self._h, len(readers), len(writers), hr, hw, io.rd,
                                                io.wr, io.user,
                                                C.byref(_stream_options(block_size, io_mode)),
                                                C.byref(err))
        io.finish()
        io.raise_for(rc, err)

    def reconstruct_stream(self, shards, block_size=4 << 20, io_mode="auto"):
        """In-memory form (stream.rs:839-900) through rsgpu_reconstruct_stream:
        `shards` are io.BytesIO; non-empty ones are present (read from position 0,
        left where the reading stopped), empty ones receive the rebuilt bytes
        (appended, their position unmoved)."""
        n = len(shards)
        vecs = (_Vec * max(1, n))()
        views, arrays, grown = [], [], {}
        for i, s in enumerate(shards):
            mv = s.getbuffer()
            if len(mv):
                arr = (C.c_char * len(mv)).from_buffer(mv)
                views.append(mv)
                arrays.append(arr)
                vecs[i].data, vecs[i].len, vecs[i].cap = C.addressof(arr), len(mv), len(mv)
                vecs[i].pos = s.tell()
            else:
                mv.release()

        def grow(_user, shard, vec, need):
            try:
                v = vec.contents
                buf = C.create_string_buffer(need)
                if v.len:
                    C.memmove(buf, v.data, v.len)
                grown[shard] = buf
                v.data, v.cap = C.addressof(buf), need
                return 0
            except Exception:  # reported as a write error of this shard
                return -1

        cb = _GROW_CB(grow)
        err = _StreamErr()
        try:
            rc = _lib.rsgpu_reconstruct_stream(self._h, vecs, n, cb, None,
                                               C.byref(_stream_options(block_size, io_mode)),
                                               C.byref(err))
        finally:
            del arrays[:]
            for mv in views:
                mv.release()
            for i, s in enumerate(shards):
                if vecs[i].data is None:
                    continue
                if i in grown:
                    pos = s.tell()
                    s.seek(0, 2)
                    s.write(C.string_at(vecs[i].data, vecs[i].len))
                    s.seek(pos)
                else:
                    s.seek(vecs[i].pos)
        _raise_stream(rc, err, {}, 0)

    # ---- kernel-variant dispatch (every variant produces the same bytes)
    def set_variant(self, variant):
        """Force one kernel variant for every call (-1: back to table/default)."""
        _check(_lib.rsgpu_codec_set_variant(self._h, int(variant)))

    def set_variant_table(self, table):
        """Install an int16 [n_ops][n_erasures][n_buckets] table (None clears)."""
        if table is None:
            _check(_lib.rsgpu_codec_set_variant_table(self._h, None, 0))
            return
        t = np.ascontiguousarray(table, np.int16)
        n_ops, n_buckets, _ = dispatch_dims()
        if t.ndim != 3 or t.shape[0] != n_ops or t.shape[2] != n_buckets:
            raise ValueError(f"table must be [{n_ops}][n_erasures][{n_buckets}]")
        self._vtab = t  # the library copies it; kept only for inspection
        _check(_lib.rsgpu_codec_set_variant_table(self._h, t.ctypes.data_as(C.POINTER(C.c_int16)),
                                                  t.shape[1]))

    def variant_valid(self, op, erasures, variant):
        return bool(_lib.rsgpu_variant_valid(self._h, OPS[op] if isinstance(op, str) else op,
                                             erasures, int(variant)))

    def variant_for(self, op, erasures, shard_len):
        v = C.c_int32()
        _check(_lib.rsgpu_codec_variant_for(self._h, OPS[op] if isinstance(op, str) else op,
                                            erasures, shard_len, C.byref(v)))
        return v.value


class _StreamErr(C.Structure):
    _fields_ = [("kind", C.c_int32), ("shard_index", C.c_int32), ("io_error", C.c_int64)]


_READ_CB = C.CFUNCTYPE(C.c_int64, _vp, _sz, C.POINTER(C.c_uint8), _sz)
_WRITE_CB = C.CFUNCTYPE(C.c_int64, _vp, _sz, C.POINTER(C.c_uint8), _sz)
IO_MODES = {"auto": 0, "serial": 1, "parallel": 2}


class StreamOptions(C.Structure):
    _fields_ = [("block_size", _sz), ("io_mode", C.c_int32), ("reserved", C.c_int32)]


def _stream_options(block_size, io_mode):
    mode = IO_MODES[io_mode] if isinstance(io_mode, str) else int(io_mode)
    return StreamOptions(int(block_size), mode, 0)


class _MemStream(C.Structure):
    _fields_ = [("data", _vp), ("len", _sz), ("pos", _sz)]


class _MemStreams(C.Structure):
    _fields_ = [("readers", C.POINTER(_MemStream)), ("writers", C.POINTER(_MemStream))]


_SOPT = C.POINTER(StreamOptions)
_lib.rsgpu_stream_options_default.argtypes = [_SOPT]
_lib.rsgpu_stream_options_default.restype = None
_lib.rsgpu_mem_stream_read.argtypes = [_vp, _sz, C.POINTER(C.c_uint8), _sz]
_lib.rsgpu_mem_stream_read.restype = C.c_int64
_lib.rsgpu_mem_stream_write.argtypes = [_vp, _sz, C.POINTER(C.c_uint8), _sz]
_lib.rsgpu_mem_stream_write.restype = C.c_int64
_lib.rsgpu_encode_stream.argtypes = [_vp, _sz, _sz, _READ_CB, _WRITE_CB, _vp, _SOPT, C.POINTER(_StreamErr)]
_lib.rsgpu_verify_stream.argtypes = [_vp, _sz, _READ_CB, _vp, _SOPT, _i32p, C.POINTER(_StreamErr)]
_lib.rsgpu_reconstruct_stream_from.argtypes = [_vp, _sz, _sz, _u8p, _u8p, _READ_CB, _WRITE_CB, _vp, _SOPT,
                                               C.POINTER(_StreamErr)]
class _Vec(C.Structure):
    _fields_ = [("data", _vp), ("len", _sz), ("cap", _sz), ("pos", _sz)]


_GROW_CB = C.CFUNCTYPE(C.c_int, _vp, _sz, C.POINTER(_Vec), _sz)
_lib.rsgpu_reconstruct_stream.argtypes = [_vp, C.POINTER(_Vec), _sz, _GROW_CB, _vp, _SOPT,
                                          C.POINTER(_StreamErr)]
# the library's own in-memory callbacks, as C function pointers
_NATIVE_RD = _READ_CB(C.cast(_lib.rsgpu_mem_stream_read, _vp).value)
_NATIVE_WR = _WRITE_CB(C.cast(_lib.rsgpu_mem_stream_write, _vp).value)


def stream_options_default():
    o = StreamOptions()
    _lib.rsgpu_stream_options_default(C.byref(o))
    return o


class HostStream:
    """A shard stream over host memory, the reference's io::Cursor over a fixed
    slice: `buf` is a contiguous uint8 numpy array (or bytes / bytearray, read-only
    for readers).  Reads take buf[pos:]; writes fill buf[pos:] and fail once it is
    full.  `pos` advances as bytes move; value() is buf[:pos]."""

    def __init__(self, buf, pos=0):
        if not isinstance(buf, np.ndarray):
            buf = np.frombuffer(buf, np.uint8)
        if buf.dtype != np.uint8 or buf.ndim != 1 or not buf.flags.c_contiguous:
            raise ValueError("HostStream needs a contiguous 1-D uint8 buffer")
        self.buf, self.pos = buf, int(pos)

    def value(self):
        return self.buf[:self.pos]

    # Python-callback form, for calls that mix HostStream with other streams
    def readinto(self, b):
        n = max(0, min(len(b), self.buf.size - self.pos))
        b[:n] = self.buf[self.pos:self.pos + n].tobytes()
        self.pos += n
        return n

    def write(self, b):
        n = max(0, min(len(b), self.buf.size - self.pos))
        if n == 0 and len(b):
            raise OSError("failed to write whole buffer")
        self.buf[self.pos:self.pos + n] = np.frombuffer(b, np.uint8, n)
        self.pos += n
        return n


def _raise_stream(rc, err, exc, writer_base):
    """StreamError (or RsError) for a stream call's status; exc maps (kind, local
    index) to the exception a Python callback raised."""
    if rc == 0:
        return
    if rc == 105:  # RSGPU_ERR_IO
        kind = "read" if err.kind == 1 else "write"
        idx = err.shard_index
        local = idx - writer_base if kind == "write" else idx
        e = exc.get((kind, local))
        if e is None:
            e = OSError(f"{kind} returned {err.io_error}")
        raise StreamError(kind, idx, e)
    if 1 <= rc <= 16:
        raise StreamError("codec", err.shard_index, RsError(rc))
    raise RsError(rc)


class _NativeIO:
    """The library's in-memory callbacks over HostStream buffers."""

    def __init__(self, readers, writers):
        self.readers, self.writers = readers, writers

        def table(streams):
            t = (_MemStream * max(1, len(streams)))()
            for i, s in enumerate(streams):
                if s is not None:
                    t[i].data, t[i].len, t[i].pos = s.buf.ctypes.data, s.buf.size, s.pos
            return t

        self._r, self._w = table(readers), table(writers)
        self._ms = _MemStreams(self._r, self._w)
        self.user = C.cast(C.pointer(self._ms), _vp)
        self.rd, self.wr = _NATIVE_RD, _NATIVE_WR

    def finish(self):
        for t, streams in ((self._r, self.readers), (self._w, self.writers)):
            for i, s in enumerate(streams):
                if s is not None:
                    s.pos = t[i].pos

    def raise_for(self, rc, err, writer_base=0):
        _raise_stream(rc, err, {}, writer_base)


def _stream_io(readers, writers):
    live = [s for s in readers + writers if s is not None]
    if live and all(isinstance(s, HostStream) for s in live):
        return _NativeIO(readers, writers)
    return _StreamIO(readers, writers)


class _StreamIO:
    """ctypes callbacks over Python readers (readinto/read) and writers (write).
    In parallel I/O mode the library calls them from its worker threads; ctypes
    takes the GIL for each call."""

    def __init__(self, readers, writers):
        self.readers, self.writers = readers, writers
        self.exc = {}
        self.user = None

        def rd(_user, shard, buf, n):
            try:
                r = self.readers[shard]
                view = (C.c_uint8 * n).from_address(C.addressof(buf.contents))
                if hasattr(r, "readinto"):
                    got = r.readinto(memoryview(view).cast("B"))
                    return int(got or 0)
                data = r.read(n)
                C.memmove(view, data, len(data))
                return len(data)
            except Exception as e:  # reported as StreamError("read", shard, e)
                self.exc[("read", shard)] = e
                return -1

        def wr(_user, shard, buf, n):
            try:
                # io's write() reads the buffer only during the call, so the staging
                # memory is handed over without a copy
                view = (C.c_uint8 * n).from_address(C.addressof(buf.contents))
                got = self.writers[shard].write(memoryview(view).cast("B"))
                return n if got is None else int(got)
            except Exception as e:
                self.exc[("write", shard)] = e
                return -1

        self.rd, self.wr = _READ_CB(rd), _WRITE_CB(wr)

    def finish(self):
        pass

    def raise_for(self, rc, err, writer_base=0):
        _raise_stream(rc, err, self.exc, writer_base)


_lib.rsgpu_variant_count.argtypes = []
_lib.rsgpu_dispatch_dims.argtypes = [_i32p, _i32p, _i32p]
_lib.rsgpu_variant_info.argtypes = [C.c_int32] + [_i32p] * 5
_lib.rsgpu_variant_valid.argtypes = [_vp, C.c_int32, _sz, C.c_int32]
_lib.rsgpu_codec_set_variant.argtypes = [_vp, C.c_int32]
_lib.rsgpu_codec_set_variant_table.argtypes = [_vp, C.POINTER(C.c_int16), _sz]
_lib.rsgpu_codec_variant_for.argtypes = [_vp, C.c_int32, _sz, _sz, _i32p]

OPS = {"encode": 0, "verify": 1, "reconstruct": 2}


def variant_count():
    return _lib.rsgpu_variant_count()


def dispatch_dims():
    """(n_ops, n_buckets, bucket0): bucket = clamp(round(2*log2(len)) - bucket0)."""
    v = [C.c_int32() for _ in range(3)]
    _check(_lib.rsgpu_dispatch_dims(*[C.byref(x) for x in v]))
    return tuple(x.value for x in v)


def variant_info(variant):
    v = [C.c_int32() for _ in range(5)]
    _check(_lib.rsgpu_variant_info(int(variant), *[C.byref(x) for x in v]))
    return dict(zip(["kind", "lanes", "vectors_per_lane", "rows_per_group", "leopard_tile_words"],
                    [x.value for x in v]))


_f32p = C.POINTER(C.c_float)
_lib.rsgpu_mlp_new.argtypes = [C.c_int32, C.c_int32, C.c_uint64, C.POINTER(_vp)]
_lib.rsgpu_mlp_free.argtypes = [_vp]
_lib.rsgpu_mlp_dims.argtypes = [_vp, _i32p, _i32p, _i32p, C.POINTER(_sz)]
_lib.rsgpu_mlp_get_params.argtypes = [_vp, _f32p, _sz]
_lib.rsgpu_mlp_set_params.argtypes = [_vp, _f32p, _sz]
_lib.rsgpu_mlp_forward.argtypes = [_vp, _vp, _sz, _vp, _vp]
_lib.rsgpu_mlp_loss_grad.argtypes = [_vp, _vp, _vp, _vp, _sz, _f32p, _f32p, _sz, _vp]
_lib.rsgpu_mlp_train.argtypes = [_vp, _vp, _vp, _vp, _sz, C.c_int32, C.c_float, _f32p, _vp]
_lib.rsgpu_codec_install_dispatch.argtypes = [_vp, _vp, C.POINTER(C.c_int64), _vp]


class Mlp:
    """The dispatcher network (fp32, n_in -> 64 -> 64 -> n_out) on the device.
    Array arguments are float32 device tensors [batch][n_in] / [batch][n_out]."""

    def __init__(self, n_in, n_out, seed=0):
        self._h = _vp()
        _check(_lib.rsgpu_mlp_new(n_in, n_out, seed, C.byref(self._h)))
        a, hdim, b, n = C.c_int32(), C.c_int32(), C.c_int32(), _sz()
        _check(_lib.rsgpu_mlp_dims(self._h, C.byref(a), C.byref(hdim), C.byref(b), C.byref(n)))
        self.n_in, self.n_hidden, self.n_out, self.n_params = a.value, hdim.value, b.value, n.value

    def __del__(self, _free=_lib.rsgpu_mlp_free):  # bound now: module globals are gone at exit
        if getattr(self, "_h", None):
            _free(self._h)
            self._h = None

    def get_params(self):
        p = np.zeros(self.n_params, np.float32)
        _check(_lib.rsgpu_mlp_get_params(self._h, p.ctypes.data_as(_f32p), p.size))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters")
        _check(_lib.rsgpu_mlp_set_params(self._h, p.ctypes.data_as(_f32p), p.size))

    @staticmethod
    def _dev(t, cols):
        if t.dtype != _torch().float32 or not t.is_cuda or not t.is_contiguous() or t.shape[-1] != cols:
            raise ValueError("expected a contiguous float32 cuda tensor with %d columns" % cols)
        return t.data_ptr()

    def forward(self, x, stream=None):
        torch = _torch()
        y = torch.empty((x.shape[0], self.n_out), dtype=torch.float32, device=x.device)
        _check(_lib.rsgpu_mlp_forward(self._h, self._dev(x, self.n_in), x.shape[0], y.data_ptr(),
                                      _stream(stream)))
        return y

    def loss_grad(self, x, t, w, stream=None):
        loss = C.c_float()
        g = np.zeros(self.n_params, np.float32)
        _check(_lib.rsgpu_mlp_loss_grad(self._h, self._dev(x, self.n_in), self._dev(t, self.n_out),
                                        self._dev(w, self.n_out), x.shape[0], C.byref(loss),
                                        g.ctypes.data_as(_f32p), g.size, _stream(stream)))
        return loss.value, g

    def train(self, x, t, w, steps, lr, stream=None):
        loss = C.c_float()
        _check(_lib.rsgpu_mlp_train(self._h, self._dev(x, self.n_in), self._dev(t, self.n_out),
                                    self._dev(w, self.n_out), x.shape[0], int(steps), float(lr),
                                    C.byref(loss), _stream(stream)))
        return loss.value


def _torch():
    import torch
    return torch


_lib.rsgpu_batch_new.argtypes = [C.POINTER(_vp)]
_lib.rsgpu_batch_free.argtypes = [_vp]
_lib.rsgpu_batch_clear.argtypes = [_vp]
_lib.rsgpu_batch_add_encode.argtypes = [_vp, _vp, C.POINTER(Slice), _sz]
_lib.rsgpu_batch_add_verify.argtypes = [_vp, _vp, C.POINTER(CSlice), _sz, _vp]
_lib.rsgpu_batch_add_reconstruct.argtypes = [_vp, _vp, C.POINTER(Slot), _sz, C.c_int32, _vp]
_lib.rsgpu_batch_set_tune.argtypes = [_vp, C.c_int32, C.c_int32, C.c_int32]
_lib.rsgpu_batch_autotune.argtypes = [_vp, _vp, C.c_int32, _i32p, _i32p, _i32p, C.POINTER(C.c_float)]
_lib.rsgpu_batch_run.argtypes = [_vp, _vp]
_lib.rsgpu_batch_graph_launch.argtypes = [_vp, _vp]
_lib.rsgpu_batch_stats.argtypes = [_vp] + [C.POINTER(C.c_uint64)] * 4


class Batch:
    """Many stripe operations executed by one batch-kernel launch.  The batch holds
    each op's codec and shards until it is cleared or dropped."""

    def __init__(self):
        h = C.c_void_p()
        _check(_lib.rsgpu_batch_new(C.byref(h)))
        self._h = h
        self._keep = []

    def __del__(self, _free=_lib.rsgpu_batch_free):  # bound now: module globals are gone at exit
        if getattr(self, "_h", None):
            _free(self._h)
            self._h = None

    def clear(self):
        self._keep = []
        _check(_lib.rsgpu_batch_clear(self._h))

    def add_encode(self, rs, shards):
        arr = (Slice * len(shards))(*[Slice(*_ptr_len(s)) for s in shards])
        _check(_lib.rsgpu_batch_add_encode(self._h, rs._h, arr, len(shards)))
        self._keep.append((rs, shards))

    def add_verify(self, rs, shards, device_flag):
        arr = (CSlice * len(shards))(*[CSlice(*_ptr_len(s)) for s in shards])
        _check(_lib.rsgpu_batch_add_verify(self._h, rs._h, arr, len(shards), device_flag.data_ptr()))
        self._keep.append((rs, shards, device_flag))

    def add_reconstruct(self, rs, slots, data_only=False, stream=None):
        """slots: list of ShardSlot (buffers for missing shards must be allocated)."""
        arr = (Slot * len(slots))(*[Slot(*_ptr_len(s.data), int(s.present)) for s in slots])
        _check(_lib.rsgpu_batch_add_reconstruct(self._h, rs._h, arr, len(slots), int(data_only),
                                                _stream(stream)))
        for s, a in zip(slots, arr):
            s.present = bool(a.present)
        self._keep.append((rs, slots))

    def set_tune(self, inputs_per_group, vectors_per_lane, blocks_per_cu):
        _check(_lib.rsgpu_batch_set_tune(self._h, inputs_per_group, vectors_per_lane, blocks_per_cu))

    def autotune(self, reps=3, stream=None):
        kb, nv, bpc, ms = C.c_int32(), C.c_int32(), C.c_int32(), C.c_float()
        _check(_lib.rsgpu_batch_autotune(self._h, _stream(stream), reps, C.byref(kb), C.byref(nv),
                                         C.byref(bpc), C.byref(ms)))
        return {"inputs_per_group": kb.value, "vectors_per_lane": nv.value,
                "blocks_per_cu": bpc.value, "ms": ms.value}

    def run(self, stream=None):
        _check(_lib.rsgpu_batch_run(self._h, _stream(stream)))

    def graph_launch(self, stream=None):
        _check(_lib.rsgpu_batch_graph_launch(self._h, _stream(stream)))

    def stats(self):
        vals = [C.c_uint64() for _ in range(4)]
        _check(_lib.rsgpu_batch_stats(self._h, *[C.byref(v) for v in vals]))
        return dict(zip(["ops", "bytes_moved", "data_bytes", "tiles"], [v.value for v in vals]))


def apply_rows(rows, inputs, outputs, field="gf8", stream=None):
    """out_j = XOR_i rows[j][i] * in_i (SelectedEncodeKernel::execute semantics)."""
    rows = np.asarray(rows)
    f = FIELDS[field]
    r, k = rows.shape
    if f == 1:
        rb = np.zeros((r, k, 2), np.uint8)
        rb[..., 0] = (rows.astype(np.uint32) >> 8) & 0xFF
        rb[..., 1] = rows.astype(np.uint32) & 0xFF
        rows_b = np.ascontiguousarray(rb.reshape(-1))
    else:
        rows_b = np.ascontiguousarray(rows.astype(np.uint8).reshape(-1))
    ins = (CSlice * k)(*[CSlice(*_ptr_len(t)) for t in inputs])
    outs = (Slice * r)(*[Slice(*_ptr_len(t)) for t in outputs])
    _check(_lib.rsgpu_apply_rows(f, rows_b.ctypes.data_as(_u8p), r, k, ins, outs, _stream(stream)))


def leopard_encode(bits, data_shards, parity_shards, data, parity, stream=None):
    """leopard_gf8_encode / leopard_gf16_encode (src/core/leopard.rs:277-296): the FFT
    encoder by shape, without a codec object.  bits: 8 or 16."""
    ins = (CSlice * max(1, len(data)))(*[CSlice(*_ptr_len(t)) for t in data])
    outs = (Slice * max(1, len(parity)))(*[Slice(*_ptr_len(t)) for t in parity])
    _check(_lib.rsgpu_leopard_encode(bits, data_shards, parity_shards, ins, len(data), outs, len(parity),
                                     _stream(stream)))


def leopard_reconstruct(bits, present, outputs, inputs, data_shards, parity_shards, stream=None):
    """reconstruct_with_tables (leopard_gf8/mod.rs:546-562, leopard.rs:299-317): present[i]
    marks shard i present (its input is read); outputs[i] receives missing shard i, None or
    an empty tensor = not requested; inputs[i] is the present shard (None when missing)."""
    total = data_shards + parity_shards
    # the C side reads data_shards + parity_shards entries of each list
    # (IncorrectShardSize as validate_leopard_decode_buffers, reconstruct.rs:155-157)
    if not (len(present) == len(inputs) == len(outputs) == total):
        raise RsError(9)
    pres = np.ascontiguousarray([1 if p else 0 for p in present], np.uint8)
    ins = (CSlice * total)(*[CSlice(*_ptr_len(t)) for t in inputs])
    outs = (Slice * total)(*[Slice(*_ptr_len(t)) for t in outputs])
    _check(_lib.rsgpu_leopard_reconstruct(bits, data_shards, parity_shards, pres.ctypes.data_as(_u8p),
                                          ins, outs, _stream(stream)))


class ShardByShard:
    """Encodes the data shards of one stripe one call at a time
    (src/core/shard_by_shard.rs:59-180) through ReedSolomon.encode_single."""

    def __init__(self, codec):
        self.codec = codec
        self.cur_input = 0

    def parity_ready(self):
        return self.cur_input == self.codec.data_shard_count()

    def reset(self):
        if self.cur_input > 0 and not self.parity_ready():
            raise SBSError("LeftoverShards")
        self.cur_input = 0

    def reset_force(self):
        self.cur_input = 0

    def cur_input_index(self):
        return self.cur_input

    def _checks(self, fn):
        if self.parity_ready():
            raise SBSError("TooManyCalls")
        try:
            fn()
        except RsError as e:
            raise SBSError("RSError", e) from None

    def encode(self, shards, stream=None):
        shards = list(shards)
        c = self.codec

        def checks():
            n = c.total_shard_count()
            if len(shards) != n:
                raise RsError(1 if len(shards) < n else 2)
            lens = [0 if t is None else t.numel() for t in shards]
            if lens[0] == 0:
                raise RsError(11)
            if any(x != lens[0] for x in lens):
                raise RsError(9)

        self._checks(checks)
        try:
            c.encode_single(self.cur_input, shards, stream)
        except RsError as e:
            raise SBSError("RSError", e) from None
        self.cur_input += 1

    def encode_sep(self, data, parity, stream=None):
        data, parity = list(data), list(parity)
        c = self.codec

        def checks():
            if len(data) != c.data_shard_count():
                raise RsError(3 if len(data) < c.data_shard_count() else 4)
            if len(parity) != c.parity_shard_count():
                raise RsError(5 if len(parity) < c.parity_shard_count() else 6)
            for group in (data, parity):
                if group[0].numel() == 0:
                    raise RsError(11)
                if any(t.numel() != group[0].numel() for t in group):
                    raise RsError(9)
            if data[0].numel() != parity[0].numel():
                raise RsError(9)

        self._checks(checks)
        try:
            c.encode_single_sep(self.cur_input, data[self.cur_input], parity, stream)
        except RsError as e:
            raise SBSError("RSError", e) from None
        self.cur_input += 1
