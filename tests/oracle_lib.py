"""ctypes wrapper around the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure only: the parity tests use it as the checker.  The oracle
is built from oracle/*.c by oracle/Makefile (``__graft_entry__.build()`` runs it).
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "_build", "liboracle.so")

GF8, GF16 = 8, 16
VANDERMONDE, CAUCHY, JERASURE, CUSTOM, FAST_ONE_PARITY = 0, 1, 2, 3, 4

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
        _lib = C.CDLL(ORACLE_SO)
        _declare(_lib)
    return _lib


u8p = C.POINTER(C.c_uint8)
u16p = C.POINTER(C.c_uint16)
ptrs = C.POINTER(C.c_void_p)


def _declare(L):
    L.ora_gf8_mul.restype = C.c_uint8
    L.ora_gf8_mul.argtypes = [C.c_uint8, C.c_uint8]
    L.ora_gf8_div.restype = C.c_uint8
    L.ora_gf8_div.argtypes = [C.c_uint8, C.c_uint8]
    L.ora_gf8_exp.restype = C.c_uint8
    L.ora_gf8_exp.argtypes = [C.c_uint8, C.c_size_t]
    L.ora_gf8_tables.argtypes = [u8p, u8p]
    L.ora_gf8_mul_slice.argtypes = [C.c_uint8, u8p, u8p, C.c_size_t]
    L.ora_gf8_mul_slice_xor.argtypes = [C.c_uint8, u8p, u8p, C.c_size_t]
    for n in ("ora_gf16_mul", "ora_gf16_div"):
        getattr(L, n).restype = C.c_uint16
        getattr(L, n).argtypes = [C.c_uint16, C.c_uint16]
    L.ora_gf16_exp.restype = C.c_uint16
    L.ora_gf16_exp.argtypes = [C.c_uint16, C.c_size_t]
    L.ora_gf16_inverse.restype = C.c_uint16
    L.ora_gf16_inverse.argtypes = [C.c_uint16]
    L.ora_matrix_invert.argtypes = [C.c_int, u16p, C.c_int, u16p]
    L.ora_matrix_multiply.argtypes = [C.c_int, u16p, C.c_int, C.c_int, u16p, C.c_int, u16p]
    L.ora_matrix_vandermonde.argtypes = [C.c_int, C.c_int, C.c_int, u16p]
    L.ora_build_matrix.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, u16p, u16p]
    L.ora_apply_rows.argtypes = [C.c_int, u16p, C.c_int, C.c_int, ptrs, ptrs, C.c_size_t]
    L.ora_encode.argtypes = [C.c_int, u16p, C.c_int, C.c_int, ptrs, ptrs, C.c_size_t]
    L.ora_verify.argtypes = [C.c_int, u16p, C.c_int, C.c_int, ptrs, C.c_size_t, C.POINTER(C.c_int)]
    L.ora_reconstruct.argtypes = [C.c_int, u16p, C.c_int, C.c_int, ptrs, u8p, C.c_size_t, C.c_int, u8p]
    L.ora_decode_matrix.argtypes = [C.c_int, u16p, C.c_int, C.c_int, u8p, C.POINTER(C.c_int), u16p]
    L.ora_leo8_tables.argtypes = [u8p, u8p, u8p, u8p]
    L.ora_leo8_encode.argtypes = [C.c_int, C.c_int, ptrs, ptrs, C.c_size_t]
    L.ora_leo8_reconstruct.argtypes = [C.c_int, C.c_int, ptrs, u8p, u8p, C.c_size_t]
    L.ora_leo8_error_locs.argtypes = [C.c_int, C.c_int, u8p, u8p]
    L.ora_leo8_mul_log.restype = C.c_uint8
    L.ora_leo8_mul_log.argtypes = [C.c_uint8, C.c_uint8]
    L.ora_leo16_tables.argtypes = [u16p, u16p, u16p, u16p]
    L.ora_leo16_encode.argtypes = [C.c_int, C.c_int, ptrs, ptrs, C.c_size_t]
    L.ora_leo16_reconstruct.argtypes = [C.c_int, C.c_int, ptrs, u8p, u8p, C.c_size_t]
    L.ora_leo16_error_locs.argtypes = [C.c_int, C.c_int, u8p, u16p]
    L.ora_leo16_mul_log.restype = C.c_uint16
    L.ora_leo16_mul_log.argtypes = [C.c_uint16, C.c_uint16]
    L.ora_qc_mix.restype = C.c_uint64
    L.ora_qc_mix.argtypes = [C.POINTER(C.c_uint64)]
    L.ora_qc_seed.restype = C.c_uint64
    L.ora_qc_seed.argtypes = [C.POINTER(C.c_uint64), C.c_size_t]
    L.ora_qc_fill.argtypes = [u8p, C.c_size_t, C.c_uint64]
    L.ora_corrupt_positions.argtypes = [C.c_size_t, C.c_size_t, C.c_uint64, C.POINTER(C.c_size_t)]
    L.ora_smallrng_fill.argtypes = [C.c_uint64, u8p, C.c_size_t]
    L.ora_derived_seed.restype = C.c_uint64
    L.ora_derived_seed.argtypes = [C.c_int, C.c_size_t, C.c_size_t, C.c_size_t]
    L.ora_cpu_isa.restype = C.c_int
    L.ora_cpu_encode.argtypes = [u16p, C.c_int, C.c_int, ptrs, ptrs, C.c_size_t, C.c_int, C.c_int]
    L.ora_cpu_encode.restype = C.c_int
    L.ora_cpu_verify.argtypes = [u16p, C.c_int, C.c_int, ptrs, C.c_size_t, C.c_int, C.c_int,
                                 C.POINTER(C.c_int)]
    L.ora_cpu_verify.restype = C.c_int
    L.ora_cpu_reconstruct.argtypes = [u16p, C.c_int, C.c_int, ptrs, u8p, u8p, C.c_size_t, C.c_int, C.c_int]
    L.ora_cpu_reconstruct.restype = C.c_int
    L.ora_checksum_shards.restype = C.c_uint64
    L.ora_checksum_shards.argtypes = [ptrs, C.POINTER(C.c_size_t), C.c_size_t]


def _u8(a):
    return a.ctypes.data_as(u8p)


def _u16(a):
    return a.ctypes.data_as(u16p)


def _ptrs(arrs):
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


# ---------------------------------------------------------------- helpers
def gf8_mul(a, b):
    return lib().ora_gf8_mul(a, b)


def gf8_div(a, b):
    return lib().ora_gf8_div(a, b)


def gf8_exp(a, n):
    return lib().ora_gf8_exp(a, n)


def gf8_tables():
    log = np.zeros(256, np.uint8)
    exp = np.zeros(510, np.uint8)
    lib().ora_gf8_tables(_u8(log), _u8(exp))
    return log, exp


def gf8_mul_slice(c, inp, out=None, xor=False):
    inp = np.ascontiguousarray(inp, np.uint8)
    out = np.zeros_like(inp) if out is None else out
    f = lib().ora_gf8_mul_slice_xor if xor else lib().ora_gf8_mul_slice
    f(c, _u8(inp), _u8(out), inp.size)
    return out


def gf16_mul(a, b):
    return lib().ora_gf16_mul(a, b)


def gf16_div(a, b):
    return lib().ora_gf16_div(a, b)


def gf16_exp(a, n):
    return lib().ora_gf16_exp(a, n)


def matrix_invert(field, a):
    a = np.ascontiguousarray(a, np.uint16)
    n = a.shape[0]
    out = np.zeros((n, n), np.uint16)
    rc = lib().ora_matrix_invert(field, _u16(a), n, _u16(out))
    return rc, out


def _field(f):
    """The oracle's field ids are 8 (GF(2^8)) and 16 (the GF(2^16) tower); the C
    side would read any other value as the tower field."""
    if f not in (GF8, GF16):
        raise ValueError(f"oracle field must be GF8 (8) or GF16 (16), not {f!r}")
    return f


def build_matrix(field, mode, k, m, custom=None):
    _field(field)
    out = np.zeros((k + m, k), np.uint16)
    cp = None
    if custom is not None:
        custom = np.ascontiguousarray(custom, np.uint16)
        cp = _u16(custom)
    rc = lib().ora_build_matrix(field, mode, k, m, cp, _u16(out))
    return rc, out


def apply_rows(field, rows, inputs):
    _field(field)
    rows = np.ascontiguousarray(rows, np.uint16)
    r, k = rows.shape
    inputs = [np.ascontiguousarray(x, np.uint8) for x in inputs]
    outs = [np.zeros_like(inputs[0]) for _ in range(r)]
    lib().ora_apply_rows(field, _u16(rows), r, k, _ptrs(inputs), _ptrs(outs), inputs[0].size)
    return outs


def encode(field, matrix, k, m, data):
    _field(field)
    data = [np.ascontiguousarray(x, np.uint8) for x in data]
    parity = [np.zeros_like(data[0]) for _ in range(m)]
    matrix = np.ascontiguousarray(matrix, np.uint16)
    lib().ora_encode(field, _u16(matrix), k, m, _ptrs(data), _ptrs(parity), data[0].size)
    return parity


def verify(field, matrix, k, m, shards):
    _field(field)
    shards = [np.ascontiguousarray(x, np.uint8) for x in shards]
    ok = C.c_int(0)
    matrix = np.ascontiguousarray(matrix, np.uint16)
    lib().ora_verify(field, _u16(matrix), k, m, _ptrs(shards), shards[0].size, C.byref(ok))
    return bool(ok.value)


def reconstruct(field, matrix, k, m, shards, present, data_only=False, required=None):
    _field(field)
    """shards: list of uint8 arrays (missing ones are overwritten in place)."""
    shards = [np.ascontiguousarray(x, np.uint8) for x in shards]
    pres = np.ascontiguousarray(present, np.uint8)
    req = None if required is None else np.ascontiguousarray(required, np.uint8)
    matrix = np.ascontiguousarray(matrix, np.uint16)
    rc = lib().ora_reconstruct(field, _u16(matrix), k, m, _ptrs(shards), _u8(pres),
                               shards[0].size, int(data_only), None if req is None else _u8(req))
    return rc, shards


def decode_matrix(field, matrix, k, m, present):
    _field(field)
    pres = np.ascontiguousarray(present, np.uint8)
    valid = (C.c_int * k)()
    dec = np.zeros((k, k), np.uint16)
    matrix = np.ascontiguousarray(matrix, np.uint16)
    rc = lib().ora_decode_matrix(field, _u16(matrix), k, m, _u8(pres), valid, _u16(dec))
    return rc, list(valid), dec


def leo8_tables():
    skew = np.zeros(255, np.uint8)
    walsh = np.zeros(256, np.uint8)
    log = np.zeros(256, np.uint8)
    exp = np.zeros(256, np.uint8)
    lib().ora_leo8_tables(_u8(skew), _u8(walsh), _u8(log), _u8(exp))
    return skew, walsh, log, exp


def leo16_tables():
    skew = np.zeros(65535, np.uint16)
    walsh = np.zeros(65536, np.uint16)
    log = np.zeros(65536, np.uint16)
    exp = np.zeros(131072, np.uint16)
    lib().ora_leo16_tables(_u16(skew), _u16(walsh), _u16(log), _u16(exp))
    return skew, walsh, log, exp


def leo_encode(bits, k, m, data):
    data = [np.ascontiguousarray(x, np.uint8) for x in data]
    parity = [np.zeros_like(data[0]) for _ in range(m)]
    f = lib().ora_leo8_encode if bits == 8 else lib().ora_leo16_encode
    rc = f(k, m, _ptrs(data), _ptrs(parity), data[0].size)
    return rc, parity


def leo_reconstruct(bits, k, m, shards, present, required=None):
    shards = [np.ascontiguousarray(x, np.uint8) for x in shards]
    pres = np.ascontiguousarray(present, np.uint8)
    req = np.ones(k + m, np.uint8) if required is None else np.ascontiguousarray(required, np.uint8)
    f = lib().ora_leo8_reconstruct if bits == 8 else lib().ora_leo16_reconstruct
    rc = f(k, m, _ptrs(shards), _u8(pres), _u8(req), shards[0].size)
    return rc, shards


def leo_error_locs(bits, k, m, present):
    pres = np.ascontiguousarray(present, np.uint8)
    if bits == 8:
        out = np.zeros(256, np.uint8)
        lib().ora_leo8_error_locs(k, m, _u8(pres), _u8(out))
    else:
        out = np.zeros(65536, np.uint16)
        lib().ora_leo16_error_locs(k, m, _u8(pres), _u16(out))
    return out


def qc_seed(parts):
    arr = (C.c_uint64 * len(parts))(*parts)
    return lib().ora_qc_seed(arr, len(parts))


def qc_fill(n, seed):
    out = np.zeros(n, np.uint8)
    lib().ora_qc_fill(_u8(out), n, seed & 0xFFFFFFFFFFFFFFFF)
    return out


def corrupt_positions(n, total, seed):
    out = (C.c_size_t * max(n, 1))()
    lib().ora_corrupt_positions(n, total, seed & 0xFFFFFFFFFFFFFFFF, out)
    return list(out)[: min(n, total)]


def smallrng_bytes(seed, n):
    out = np.zeros(n, np.uint8)
    lib().ora_smallrng_fill(seed & 0xFFFFFFFFFFFFFFFF, _u8(out), n)
    return out


def derived_seed(op_tag, k, m, shard_size):
    return lib().ora_derived_seed(op_tag, k, m, shard_size)


def checksum_shards(shards):
    shards = [np.ascontiguousarray(x, np.uint8) for x in shards]
    lens = (C.c_size_t * len(shards))(*[s.size for s in shards])
    return lib().ora_checksum_shards(_ptrs(shards), lens, len(shards))


CPU_ISA = {0: "auto", 1: "scalar", 2: "avx2", 3: "gfni-avx512"}


def cpu_encode(matrix, k, m, data, threads=1, isa=0, parity=None):
    """The benchmark's CPU baseline (oracle/cpu_simd.c); returns (isa used, parity)."""
    data = [np.ascontiguousarray(x, np.uint8) for x in data]
    if parity is None:
        parity = [np.zeros_like(data[0]) for _ in range(m)]
    matrix = np.ascontiguousarray(matrix, np.uint16)
    used = lib().ora_cpu_encode(_u16(matrix), k, m, _ptrs(data), _ptrs(parity), data[0].size,
                                threads, isa)
    return used, parity


def cpu_verify(matrix, k, m, shards, threads=1, isa=0):
    """The CPU baseline's verify; returns (isa used, ok)."""
    shards = [np.ascontiguousarray(x, np.uint8) for x in shards]
    matrix = np.ascontiguousarray(matrix, np.uint16)
    ok = C.c_int(0)
    used = lib().ora_cpu_verify(_u16(matrix), k, m, _ptrs(shards), shards[0].size, threads, isa, C.byref(ok))
    return used, bool(ok.value)


def cpu_data_rows(matrix, k, m, present):
    """Decode rows (over the first k present shards) of the missing data shards,
    as one uint8 array of n_missing_data x k: the baseline's cached inverse."""
    rc, valid, dec = decode_matrix(GF8, matrix, k, m, present)
    assert rc == 0
    rows = [dec[d] for d in range(k) if not present[d]]
    return np.ascontiguousarray(np.array(rows, np.uint8).reshape(-1) if rows else np.zeros(1, np.uint8))


def cpu_reconstruct(matrix, k, m, shards, present, data_rows, threads=1, isa=0):
    """The CPU baseline's reconstruct into the missing entries of `shards` (in place)."""
    matrix = np.ascontiguousarray(matrix, np.uint16)
    pres = np.ascontiguousarray(present, np.uint8)
    return lib().ora_cpu_reconstruct(_u16(matrix), k, m, _ptrs(shards), _u8(pres), _u8(data_rows),
                                     shards[0].size, threads, isa)
