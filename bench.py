"""Benchmark of the Reed-Solomon data path on one MI355X.

Workload: 512 stripes of RS 10+4 (GF(2^8), Vandermonde) with 1 MiB shards of uniform random
bytes, device-resident.  One step encodes all 512 stripes.

  value               : batched encode -- the 512 stripes in ONE batch-kernel launch
                        replayed from a hipGraph (data bytes k*S per second)
  encode_plain_gbs    : the same stripes launched one stripe at a time
  verify_*/reconstruct_*: verify (register compare) and reconstruct with 1..4 random erasures
                        per stripe (data and parity mixed), plain and batched
  encode_8x3_*        : the 8+3 x 1 MiB shape
  stripe_latency_ms   : p50/p99 device time of one plain stripe encode
  *_dispatched_*      : per-stripe calls with the learned dispatcher's variant tables
  dispatch_live       : the dispatcher judged live on unseen call shapes (regret vs exhaustive
                        timing, gain vs the default tiling)
  leopard             : Leopard GF8 128+64 / GF16 1024+256 / 4096+1024, 4-256 KiB shards
  serving             : open-loop Poisson request streams through the request server
                        (continuous batching, 200 GB admission budget): p50/p99 latency, GB/s
  host_boundary       : PCIe-inclusive rates of the host-pointer and block-stream entry points
  roofline            : the batch kernel (dominant kernel) vs the 8 TB/s HBM peak
  cpu_baseline        : C restatement of the reference's SIMD CPU encode (GFNI / AVX2 row
                        kernels, chunked over a thread pool) on distinct host-resident stripes

Prints ONE JSON line.
"""
import argparse
import ctypes as C
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rustfs-erasure-codec_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--stripes", type=int, default=512)
    ap.add_argument("--shard", type=int, default=1 << 20)
    ap.add_argument("--cpu-stripes", type=int, default=96)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="host threads of the CPU baseline (the box's CPU share is 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ragged", type=int, default=8192,
                    help="stripes in the ragged mix (0 skips it)")
    ap.add_argument("--no-leopard", dest="leopard", action="store_false",
                    help="skip the Leopard GF8 128+64 / GF16 1024+256 / 4096+1024 section")
    ap.add_argument("--dispatch-live", type=int, default=32,
                    help="call shapes on which the dispatcher is judged live (0 skips)")
    ap.add_argument("--rates", default="2000,10000,50000",
                    help="request rates (req/s) for the serving section ('' skips)")
    ap.add_argument("--leopard-only", action="store_true",
                    help="only the Leopard section (for profiling its kernels)")
    ap.add_argument("--leopard-case", nargs=4, metavar=("FAMILY", "K", "M", "SHARD"),
                    help="one Leopard shape (e.g. leopard_gf8 10 4 1048576): encode and "
                         "reconstruct with m erasures, default kernel choice")
    ap.add_argument("--leopard-erasures", type=int, default=None, metavar="E",
                    help="with --leopard-case: reconstruct with E erasures instead of m, "
                         "also with the matrix-apply variant (--leopard-variant)")
    ap.add_argument("--leopard-variant", type=int, default=-1,
                    help="with --leopard-case: the forced kernel variant of the timed calls")
    ap.add_argument("--mfma-only", action="store_true",
                    help="only the int8 matrix-core vs VALU encode comparison")
    ap.add_argument("--gf16-batch-only", action="store_true",
                    help="only the GF(2^16) batched-vs-plain section")
    ap.add_argument("--serving-only", action="store_true",
                    help="only the serving section (request server latency and capacity)")
    ap.add_argument("--host-only", action="store_true",
                    help="only the host-boundary section (host-pointer and block-stream rates)")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the tuned batch encode launches plus a 1 GiB device copy "
                         "(the reference kernel for scaling rocprofv3 byte counters)")
    return ap.parse_args()


def roofline_only(args):
    """Short run for counter collection: a 1 GiB device copy of known size, then the tuned
    10+4 batch encode launched `steps` times."""
    import torch

    import rsgpu

    torch.cuda.init()
    src = torch.ones(1 << 30, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    sp = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    k, m, S, N = 10, 4, args.shard, args.stripes
    st = Stripes(rsgpu, torch, k, m, N, S, seed=0x00EC5EED)
    b = st.batch("encode", sp)
    tune = b.autotune(reps=2)
    for _ in range(args.steps):
        b.graph_launch()
    torch.cuda.synchronize()
    print(json.dumps({"roofline_only": True, "tune": tune, "bytes_per_launch": N * (k + m) * S,
                      "copy_bytes": 1 << 30}))


class Stripes:
    """n stripes of k+m shards in one device allocation, with prebuilt C-ABI views.

    Shards sit at a pitch of shard + `pad` bytes.  An exact power-of-two pitch puts the
    k+m streams of a column on the same HBM channel bits; independent allocations
    would not, so the default pitch adds 64 KiB + 256 B (both pitches are reported)."""

    def __init__(self, rsgpu, torch, k, m, n, shard, seed, pad=65536 + 256, codec_kw=None):
        self.k, self.m, self.n, self.shard = k, m, n, shard
        self.pitch = shard + pad
        g = torch.Generator(device="cuda").manual_seed(seed)
        self.store = torch.empty((n, k + m, self.pitch), dtype=torch.uint8, device="cuda")
        self.buf = self.store[:, :, :shard]
        self.buf[:, :k].random_(0, 256, generator=g)
        self.rs = rsgpu.ReedSolomon(k, m, **(codec_kw or {}))
        self.lib = rsgpu.lib()
        self.rsgpu = rsgpu
        self.enc, self.ver = [], []
        for s in range(n):
            ptrs = [self.ptr(s, i) for i in range(k + m)]
            self.enc.append((rsgpu.Slice * (k + m))(*[rsgpu.Slice(p, shard) for p in ptrs]))
            self.ver.append((rsgpu.CSlice * (k + m))(*[rsgpu.CSlice(p, shard) for p in ptrs]))

    def ptr(self, s, i):
        return self.store.data_ptr() + (s * (self.k + self.m) + i) * self.pitch

    def _ok(self, rc):
        if rc:
            raise self.rsgpu.RsError(rc)

    def encode(self, s, stream):
        self._ok(self.lib.rsgpu_encode(self.rs._h, self.enc[s], self.k + self.m, stream))

    def verify_async(self, s, flag_ptr, stream):
        self._ok(self.lib.rsgpu_verify_async(self.rs._h, self.ver[s], self.k + self.m, flag_ptr, stream))

    def erasures(self, seed):
        """Per stripe: 1..m erased positions at random (data and parity mixed)."""
        import numpy as np

        rng = np.random.default_rng(seed)
        out = []
        for _ in range(self.n):
            e = int(rng.integers(1, self.m + 1))
            out.append(set(rng.choice(self.k + self.m, e, replace=False).tolist()))
        return out

    def slots(self, s, missing):
        arr = (self.rsgpu.Slot * (self.k + self.m))()
        for i in range(self.k + self.m):
            arr[i] = self.rsgpu.Slot(self.ptr(s, i), self.shard, 0 if i in missing else 1)
        return arr

    def reconstruct(self, arr, stream):
        self._ok(self.lib.rsgpu_reconstruct(self.rs._h, arr, self.k + self.m, stream))

    def batch(self, op, stream, erasures=None, flags=None):
        b = self.rsgpu.Batch()
        lib = self.lib
        for s in range(self.n):
            if op == "encode":
                self._ok(lib.rsgpu_batch_add_encode(b._h, self.rs._h, self.enc[s], self.k + self.m))
            elif op == "verify":
                self._ok(lib.rsgpu_batch_add_verify(b._h, self.rs._h, self.ver[s], self.k + self.m,
                                                    flags.data_ptr() + 4 * s))
            else:
                self._ok(lib.rsgpu_batch_add_reconstruct(b._h, self.rs._h, self.slots(s, erasures[s]),
                                                         self.k + self.m, 0, stream))
        return b


def gf16_batch_section(torch, rsgpu, steps):
    """GF(2^16) stripes in the batch kernels against one call per stripe.  Classic
    (tower) 10+4 x 256 KiB over 256 stripes: encode, and reconstruct of 1..4 erasures;
    Leopard GF(2^16) 64+16 x 64 KiB over 64 stripes: reconstruct of 1..4 erasures as
    the decoder's own matrix (matrix variant forced; the batch puts it in the
    split-layout kernel).  The stripes are codewords (encoded first), rebuilt shards
    are overwritten with a pattern before each timed run and compared with the
    codeword after it."""
    import ctypes as C

    import numpy as np

    sp = torch.cuda.current_stream().cuda_stream
    rate = lambda nbytes, t: round(nbytes / t / 1e9, 2)
    out = {}
    cases = [("classic_gf16_10x4_256KiB", {"field": "gf16"}, 10, 4, 256, 256 << 10),
             ("leopard_gf16_64x16_64KiB", {"codec_family": "leopard_gf16"}, 64, 16, 64, 64 << 10)]
    for name, kw, k, m, n, S in cases:
        st = Stripes(rsgpu, torch, k, m, n, S, seed=7, codec_kw=kw)
        row = {}
        enc_plain = lambda: [st.encode(s, sp) for s in range(n)]
        t_plain = timed(torch, enc_plain, steps, 1)
        ref = st.buf.clone()
        if name.startswith("classic"):
            b = st.batch("encode", sp)
            st.buf[:, k:].fill_(0xA5)
            t_batch = timed(torch, b.graph_launch, steps, 1)
            assert torch.equal(st.buf, ref), name + " encode: batched differs from plain"
            row["encode"] = {"plain_gbs": rate(n * k * S, t_plain), "batched_gbs": rate(n * k * S, t_batch),
                             "batched_vs_plain": round(t_plain / t_batch, 2)}
            del b
        else:
            st.rs.set_variant(0)
        rng = np.random.default_rng(5)
        er = [set(rng.choice(k + m, int(rng.integers(1, 5)), replace=False).tolist()) for _ in range(n)]
        layouts = [st.slots(s, er[s]) for s in range(n)]
        work = [st.slots(s, er[s]) for s in range(n)]
        nbytes = C.sizeof(layouts[0])

        def rec_plain():
            for s in range(n):
                C.memmove(work[s], layouts[s], nbytes)  # presence as laid out (a call marks targets)
                st.reconstruct(work[s], sp)

        def poison():
            for s in range(n):
                for i in er[s]:
                    st.buf[s, i].fill_(0x5A)

        poison()
        t_plain = timed(torch, rec_plain, steps, 1)
        assert torch.equal(st.buf, ref), name + " reconstruct: plain differs from the codeword"
        b = st.batch("reconstruct", sp, erasures=er)
        poison()
        t_batch = timed(torch, b.graph_launch, steps, 1)
        assert torch.equal(st.buf, ref), name + " reconstruct: batched differs from the codeword"
        row["reconstruct_1_4_erasures"] = {"plain_gbs": rate(n * k * S, t_plain),
                                           "batched_gbs": rate(n * k * S, t_batch),
                                           "batched_vs_plain": round(t_plain / t_batch, 2)}
        row["batch_tiles"] = b.stats()["tiles"]
        out[name] = row
        del b, ref, st
    return out


def config1_section(torch, rsgpu, steps):
    """64 stripes of RS 10+4 x 64 KiB: encode, verify, and reconstruct of 4 erasures
    per stripe (positions uniform over the 14 shards), batched (one batch-kernel launch
    per step, hipGraph replay) and plain (one call per stripe from Python).  Encoded
    shards and rebuilt shards are overwritten with a pattern before each timed run and
    compared with the codeword after it; verify runs on codewords and must pass."""
    import ctypes as C

    import numpy as np

    sp = torch.cuda.current_stream().cuda_stream
    k, m, n, S = 10, 4, 64, 64 << 10
    data = n * k * S
    rate = lambda t: round(data / t / 1e9, 2)
    st = Stripes(rsgpu, torch, k, m, n, S, seed=0x00C0F161)
    out = {"stripes": n, "shard_bytes": S}
    st.buf[:, k:].fill_(0xA5)
    t_plain = timed(torch, lambda: [st.encode(s, sp) for s in range(n)], steps, 2)
    ref = st.buf.clone()
    b = st.batch("encode", sp)
    b.autotune(reps=3)
    st.buf[:, k:].fill_(0xA5)
    t_batch = timed(torch, b.graph_launch, steps, 2)
    assert torch.equal(st.buf, ref), "config 1 encode: batched differs from plain"
    out["encode"] = {"plain_gbs": rate(t_plain), "batched_gbs": rate(t_batch)}
    del b
    flags = torch.zeros(n, dtype=torch.int32, device="cuda")
    t_plain = timed(torch, lambda: [st.verify_async(s, flags.data_ptr() + 4 * s, sp) for s in range(n)],
                    steps, 2)
    assert int(flags.sum()) == 0, "config 1 verify (plain) reported a mismatch"
    b = st.batch("verify", sp, flags=flags)
    b.autotune(reps=3)

    def ver_batch():
        flags.zero_()
        b.graph_launch()

    t_batch = timed(torch, ver_batch, steps, 2)
    assert int(flags.sum()) == 0, "config 1 verify (batched) reported a mismatch"
    out["verify"] = {"plain_gbs": rate(t_plain), "batched_gbs": rate(t_batch)}
    del b
    rng = np.random.default_rng(0xC1)
    er = [set(rng.choice(k + m, 4, replace=False).tolist()) for _ in range(n)]
    layouts = [st.slots(s, er[s]) for s in range(n)]
    work = [st.slots(s, er[s]) for s in range(n)]
    nbytes = C.sizeof(layouts[0])

    def rec_plain():
        for s in range(n):
            C.memmove(work[s], layouts[s], nbytes)  # presence as laid out (a call marks targets)
            st.reconstruct(work[s], sp)

    def poison():
        for s in range(n):
            for i in er[s]:
                st.buf[s, i].fill_(0x5A)

    poison()
    t_plain = timed(torch, rec_plain, steps, 2)
    assert torch.equal(st.buf, ref), "config 1 reconstruct: plain differs from the codeword"
    b = st.batch("reconstruct", sp, erasures=er)
    b.autotune(reps=3)
    poison()
    t_batch = timed(torch, b.graph_launch, steps, 2)
    assert torch.equal(st.buf, ref), "config 1 reconstruct: batched differs from the codeword"
    out["reconstruct_4_erasures"] = {"plain_gbs": rate(t_plain), "batched_gbs": rate(t_batch)}
    del b, ref, st
    return out


def timed(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def event_durations(torch, fn, count):
    """Device durations (ms) of `count` calls of fn(i), events recorded on torch's stream
    (the stream every launch here is enqueued on)."""
    import numpy as np

    stream = torch.cuda.current_stream()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(count)]
    for i in range(count):
        evs[i][0].record(stream)
        fn(i)
        evs[i][1].record(stream)
    torch.cuda.synchronize()
    return np.array([a.elapsed_time(b) for a, b in evs])


def measured_traffic(bytes_per_launch):
    """HBM bytes per launch of the encode batch kernel, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (scaled by a 1 GiB copy of known size); None when the
    workload differs from the profiled one."""
    path = os.path.join(ROOT, "profiles", "batch_kernel_encode_10x4_1MiB_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except OSError:
        return None
    if d.get("bytes_per_launch_algorithmic") != bytes_per_launch:
        return None
    return d["traffic_bytes_per_launch"]


def cpu_baseline(k, m, shard, stripes, threads, erasures):
    """The reference's SIMD CPU path restated in C (oracle/cpu_simd.c: GFNI affine
    rows on AVX-512, AVX2 nibble tables otherwise; 16/64/256 KiB chunks spread over
    a thread pool), on `stripes` distinct host-resident stripes so the data is not
    cache-resident: encode, verify (recompute and compare) and reconstruct (the
    reference's data-then-parity stages over the GPU run's erasure patterns, decode
    rows computed once per pattern as its inversion cache holds them).  Plus the
    64 x 10+4 x 64 KiB set with 4 random erasures per stripe."""
    import numpy as np

    import oracle_lib as O

    _, mat = O.build_matrix(O.GF8, O.VANDERMONDE, k, m)

    def run(S, n, erased, min_s=1.0):
        rng = np.random.default_rng(S + n)
        data = [[np.frombuffer(rng.bytes(S), np.uint8).copy() for _ in range(k)] for _ in range(n)]
        par = [[np.zeros(S, np.uint8) for _ in range(m)] for _ in range(n)]
        isa, _ = O.cpu_encode(mat, k, m, data[0], threads=threads, parity=par[0])  # warm
        exp = O.encode(O.GF8, mat, k, m, data[0])
        assert all(np.array_equal(a, b) for a, b in zip(par[0], exp)), "CPU baseline differs from oracle"
        out = {}

        def timeit(fn):
            fn()
            reps, t0 = 0, time.perf_counter()
            while True:
                fn()
                reps += 1
                dt = time.perf_counter() - t0
                if dt >= min_s:
                    return reps * n * k * S / dt / 1e9, reps

        out["encode"], reps_e = timeit(lambda: [O.cpu_encode(mat, k, m, data[s], threads=threads, parity=par[s])
                                                for s in range(n)])
        full = [data[s] + par[s] for s in range(n)]
        oks = []
        out["verify"], _ = timeit(lambda: oks.extend(O.cpu_verify(mat, k, m, full[s], threads=threads)[1]
                                                     for s in range(n)))
        assert all(oks), "CPU verify reported a mismatch on encoded stripes"
        present = [[0 if i in erased[s] else 1 for i in range(k + m)] for s in range(n)]
        rows = [O.cpu_data_rows(mat, k, m, present[s]) for s in range(n)]
        ref = [[x.copy() for x in full[s]] for s in range(n)]
        for s in range(n):
            for i in erased[s]:
                full[s][i][:] = 0xA5
        out["reconstruct"], _ = timeit(lambda: [O.cpu_reconstruct(mat, k, m, full[s], present[s], rows[s],
                                                                  threads=threads) for s in range(n)])
        assert all(np.array_equal(a, b) for s in range(n) for a, b in zip(full[s], ref[s])), \
            "CPU reconstruct differs from the encoded stripes"
        return {kk: round(v, 3) for kk, v in out.items()}, isa

    main, isa = run(shard, stripes, erasures[:stripes])
    rng = np.random.default_rng(4)
    er1 = [set(rng.choice(k + m, 4, replace=False).tolist()) for _ in range(64)]
    small, _ = run(64 << 10, 64, er1)
    return {"value": main["encode"], "unit": "GB/s", "cores": threads, "kind": "port",
            "encode": main["encode"], "verify": main["verify"], "reconstruct": main["reconstruct"],
            "config1_64x10+4x64KiB": small,
            "sample": f"{stripes} distinct host-resident stripes of RS {k}+{m} x {shard // 1024} KiB, repeated "
                      f"for >= 1 s per op (reconstruct: the GPU run's 1..{m} erasures per stripe), and "
                      f"64 x {k}+{m} x 64 KiB with 4 erasures; C restatement of the reference's SIMD row kernels "
                      f"({O.CPU_ISA[isa]}), {threads} threads; data GB/s"}


def leopard_section(torch, rsgpu, DSP, disp):
    """Leopard codecs at 4-256 KiB shards: device time per call with the default
    kernel choice and with the dispatcher's variant tables (data GB/s = k*S/t)."""
    out = {}
    for name, fam, k, m in [("gf8_128+64", "leopard_gf8", 128, 64),
                            ("gf16_1024+256", "leopard_gf16", 1024, 256),
                            ("gf16_4096+1024", "leopard_gf16", 4096, 1024)]:
        rs = rsgpu.ReedSolomon(k, m, codec_family=fam)
        rows = {}
        for S in (4096, 65536, 262144):
            # m erasures, and on GF(2^16) at 256 KiB also 1 and 8 (few lost shards,
            # where the decoder-matrix variants compete with the FFT)
            extra = ((1, "reconstruct_1_erasure"), (8, "reconstruct_8_erasures")) \
                if fam == "leopard_gf16" and S == 262144 else ()
            for op, e, label in (("encode", 0, "encode"), ("reconstruct", m, "reconstruct"),
                                 *[("reconstruct", x, lab) for x, lab in extra]):
                case = DSP.OpCase(torch, rs, fam, k, m, S, op, e, seed=S + e)
                rs.set_variant_table(None)
                t_def = case.time([-1])[-1]
                disp.install(rs, fam)
                v = rs.variant_for(op, e, S)
                t_dsp = case.time([-1])[-1]
                rs.set_variant_table(None)
                rows[f"{label}_{S // 1024}KiB"] = {
                    "default_gbs": round(k * S / t_def / 1e3, 2),
                    "dispatched_gbs": round(k * S / t_dsp / 1e3, 2),
                    "dispatched_variant": v, "dispatched_us": round(t_dsp, 2)}
                del case
        out[name] = rows
        del rs
    # the two Leopard GF8 shapes the reference publishes (1 MiB shards, data GiB/s on
    # Apple M-series NEON: docs/benchmark-results-2026-06-04.md:53-54)
    for name, k, m, pub_gib in (("gf8_10+4_1MiB", 10, 4, 10.47), ("gf8_32+16_1MiB", 32, 16, 7.04)):
        rs = rsgpu.ReedSolomon(k, m, codec_family="leopard_gf8")
        S = 1 << 20
        rows = {"reference_published_encode_gbs": round(pub_gib * 2 ** 30 / 1e9, 2)}
        # reconstruct with m erasures (exactly k present) and with one
        for name_op, op, e in (("encode", "encode", 0), ("reconstruct", "reconstruct", m),
                               ("reconstruct_1_erasure", "reconstruct", 1)):
            case = DSP.OpCase(torch, rs, "leopard_gf8", k, m, S, op, e, seed=k + e)
            rs.set_variant_table(None)
            t_def = case.time([-1])[-1]
            disp.install(rs, "leopard_gf8")
            v = rs.variant_for(op, e, S)
            t_dsp = case.time([-1])[-1]
            rs.set_variant_table(None)
            rows[name_op] = {"default_gbs": round(k * S / t_def / 1e3, 2),
                             "dispatched_gbs": round(k * S / t_dsp / 1e3, 2), "dispatched_variant": v}
            del case
        out[name] = rows
        del rs
    return out


def gf2_mfma_section(torch, rsgpu):
    """The int8 matrix-core formulation (bit-expanded GF(2) product) against the VALU
    apply kernels on single large stripes: device time per encode (HIP events), the
    best of all 15 apply variants and the default one, outputs compared byte for byte."""
    out = {}
    for k, m, S in ((10, 4, 64 << 20), (24, 8, 32 << 20), (31, 8, 32 << 20)):
        rs = rsgpu.ReedSolomon(k, m)
        g = torch.Generator(device="cuda").manual_seed(k * 100 + m)
        data = [torch.randint(0, 256, (S,), dtype=torch.uint8, device="cuda", generator=g)
                for _ in range(k)]
        par_v = [torch.zeros(S, dtype=torch.uint8, device="cuda") for _ in range(m)]
        par_x = [torch.zeros(S, dtype=torch.uint8, device="cuda") for _ in range(m)]

        def dev_ms(fn):
            fn()
            return float(event_durations(torch, lambda i: fn(), 5).mean())

        rs.set_variant(-1)
        t_default = dev_ms(lambda: rs.encode_sep(data, par_v))
        best_v, t_best = -1, t_default
        for v in range(rsgpu.variant_count()):
            if not rs.variant_valid("encode", 0, v):
                continue
            rs.set_variant(v)
            t = dev_ms(lambda: rs.encode_sep(data, par_v))
            if t < t_best:
                best_v, t_best = v, t
        rs.set_variant(-1)
        t_mfma = dev_ms(lambda: rs.encode_sep_gf2mfma(data, par_x))
        torch.cuda.synchronize()
        same = all(torch.equal(a, b) for a, b in zip(par_v, par_x))
        moved = (k + m) * S
        out[f"{k}+{m}_{S >> 20}MiB"] = {
            "vector_best_variant": best_v, "vector_best_ms": round(t_best, 4),
            "vector_default_ms": round(t_default, 4), "mfma_ms": round(t_mfma, 4),
            "vector_best_moved_gbs": round(moved / t_best / 1e6, 1),
            "mfma_moved_gbs": round(moved / t_mfma / 1e6, 1),
            "mfma_vs_vector_best": round(t_best / t_mfma, 3), "identical": same}
        del data, par_v, par_x, rs
        torch.cuda.empty_cache()
    return out


def host_section(rsgpu):
    """PCIe-inclusive rates (never `value`) on 10+4 with host-resident shards: the
    host-pointer entry points (a two-slot pipeline over 512 KiB column chunks:
    threaded copies into pinned staging, both PCIe directions and the kernels
    overlap) at 4 MiB and 64 MiB shards, and the block streams (4 MiB blocks
    through pinned double-buffered staging; the library's C callbacks over host
    memory, so no Python runs per block)."""
    import numpy as np

    k, m = 10, 4
    rs = rsgpu.ReedSolomon(k, m)
    rng = np.random.default_rng(3)
    out = {}

    def per_call(fn, reps):
        fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        return (time.perf_counter() - t0) / reps

    for S, reps in ((4 << 20, 16), (64 << 20, 4)):
        shards = [np.frombuffer(rng.bytes(S), np.uint8).copy() for _ in range(k)] + \
                 [np.zeros(S, np.uint8) for _ in range(m)]
        tag = f"{S >> 20}MiB"
        out[f"encode_host_{tag}_gbs"] = round(k * S / per_call(lambda: rs.encode_host(shards), reps) / 1e9, 2)
        if S == 64 << 20:
            out[f"verify_host_{tag}_gbs"] = round(k * S / per_call(lambda: rs.verify_host(shards), reps) / 1e9, 2)
            present = np.ones(k + m, np.uint8)
            present[[0, 5, 10, 13]] = 0
            out[f"reconstruct_host_{tag}_4_erasures_gbs"] = round(
                k * S / per_call(lambda: rs.reconstruct_host(shards, present), reps) / 1e9, 2)
        del shards
    # block streams over host memory through the library's native C callbacks
    # (HostStream): 64 MiB per shard in 4 MiB blocks, each I/O mode
    S2, blk = 64 << 20, 4 << 20
    data = [np.frombuffer(rng.bytes(S2), np.uint8).copy() for _ in range(k)]
    parity = [np.zeros(S2, np.uint8) for _ in range(m)]
    HS = rsgpu.HostStream
    for mode in ("auto", "serial", "parallel"):
        t = per_call(lambda: rs.encode_stream([HS(d) for d in data], [HS(p) for p in parity],
                                              block_size=blk, io_mode=mode), 2)
        out[f"encode_stream_{mode}_gbs"] = round(k * S2 / t / 1e9, 2)
    full = data + parity
    t = per_call(lambda: rs.verify_stream([HS(x) for x in full], block_size=blk), 2)
    out["verify_stream_auto_gbs"] = round(k * S2 / t / 1e9, 2)
    lost = (0, 5, 10, 13)
    rebuilt = {i: np.zeros(S2, np.uint8) for i in lost}

    def rebuild():
        rs.reconstruct_stream_from([None if i in lost else HS(full[i]) for i in range(k + m)],
                                   [HS(rebuilt[i]) if i in lost else None for i in range(k + m)],
                                   block_size=blk)

    t = per_call(rebuild, 2)
    out["reconstruct_stream_4_erasures_auto_gbs"] = round(k * S2 / t / 1e9, 2)
    out["stream_outputs_match"] = bool(all((rebuilt[i] == full[i]).all() for i in lost))
    out.update({"stream_bytes_per_shard": S2, "stream_block_bytes": blk,
                "stream_callbacks": "native C (rsgpu_mem_stream_read / _write)"})
    return out


def serving_section(torch, rates, budget=200e9):
    """Open-loop Poisson streams through the request server (continuous batching,
    HBM admission); objects log-uniform 16 KiB-64 MiB on 10+4 / 12+4 / 16+4."""
    from rsgpu import serve as SV

    store = SV.Store(torch, per_shape=8)
    srv = SV.Server(hbm_budget=int(budget), arena_bytes=16 << 30, max_inflight_batches=4,
                    max_batch_ops=4096)
    reqs, keep, meta = SV.poisson_requests(store, 400, 20000.0, seed=1)
    srv.replay(reqs)  # warm-up: decode rows of the erasure patterns, batch buffers
    out = {"budget_bytes": int(budget), "store_bytes": int(store.buf.numel()),
           "batch_slots": 4, "rates": {}}
    for rate in rates:
        n = int(min(rate * 1.0, 25000))
        reqs, keep, meta = SV.poisson_requests(store, n, float(rate), seed=rate)
        comps = srv.replay(reqs)
        summ = SV.latency_summary(comps, meta)
        summ = {k: (round(float(v), 4) if isinstance(v, float) or hasattr(v, "dtype") else v)
                for k, v in summ.items()}
        out["rates"][str(rate)] = summ
    # capacity: 4000 requests arriving at once; the first such burst also grows
    # the batch staging and decode-slot pools, the second (different requests,
    # new erasure patterns) runs on the grown server
    for name, seed in (("saturated_first_burst", 77), ("saturated_4000_requests", 78)):
        reqs, keep, meta = SV.poisson_requests(store, 4000, 1e12, seed=seed)
        comps = srv.replay(reqs)
        cap = SV.latency_summary(comps, meta)
        out[name] = {"data_gbs": round(float(cap["data_gbs"]), 1),
                     "moved_gbs": round(float(cap["moved_gbs"]), 1),
                     "seconds": round(float(cap["seconds"]), 4)}
    out["stats"] = srv.stats()
    srv.close()
    del store
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(stripes are independent: more GPUs would each take their own "
                                   "stripes, with no exchange between them)"}))
        sys.exit(2)
    if args.roofline_only:
        roofline_only(args)
        return
    if args.leopard_case:
        import torch

        import rsgpu
        from rsgpu import dispatch as DSP

        torch.cuda.init()
        fam, k, m, S = args.leopard_case[0], *map(int, args.leopard_case[1:])
        rs = rsgpu.ReedSolomon(k, m, codec_family=fam)
        e_rec = m if args.leopard_erasures is None else args.leopard_erasures
        v = args.leopard_variant
        out = {}
        for op, e in (("encode", 0), ("reconstruct", e_rec)):
            vv = v if rs.variant_valid(op, e, v) else -1
            t = DSP.OpCase(torch, rs, fam, k, m, S, op, e, seed=1).time([vv])[vv]
            out[op] = {"us": round(t, 1), "data_gbs": round(k * S / t / 1e3, 1), "erasures": e,
                       "variant": vv}
        print(json.dumps({"leopard_case": {"family": fam, "k": k, "m": m, "shard": S, **out}}))
        return
    if args.serving_only:
        import torch

        torch.cuda.init()
        print(json.dumps({"serving": serving_section(torch, [int(r) for r in args.rates.split(",")])}))
        return
    if args.host_only:
        import torch

        import rsgpu

        torch.cuda.init()
        print(json.dumps({"host_boundary": host_section(rsgpu)}))
        return
    if args.gf16_batch_only:
        import torch

        import rsgpu

        torch.cuda.init()
        print(json.dumps({"gf16_batch": gf16_batch_section(torch, rsgpu, 10),
                          "config1_64x10+4x64KiB": config1_section(torch, rsgpu, args.steps)}))
        return
    if args.mfma_only:
        import torch

        import rsgpu

        torch.cuda.init()
        print(json.dumps({"gf2_mfma": gf2_mfma_section(torch, rsgpu)}))
        return
    if args.leopard_only:
        import torch

        import rsgpu
        from rsgpu import dispatch as DSP

        torch.cuda.init()
        print(json.dumps({"leopard": leopard_section(torch, rsgpu, DSP, DSP.Dispatcher.load())}))
        return
    import numpy as np
    import torch

    import rsgpu

    torch.cuda.init()
    stream = torch.cuda.current_stream()
    sp = C.c_void_p(stream.cuda_stream)
    k, m, S, N = 10, 4, args.shard, args.stripes
    st = Stripes(rsgpu, torch, k, m, N, S, seed=0x00EC5EED)
    data_bytes = N * k * S
    pitch = st.pitch

    # ---- plain: one launch per stripe
    t_enc_plain = timed(torch, lambda: [st.encode(s, sp) for s in range(N)], args.steps, args.warmup)
    lat = event_durations(torch, lambda i: st.encode(i % N, sp), min(N, 256))
    ref_parity = st.buf[:, k:].clone()
    from rsgpu import dispatch as DSP

    disp = DSP.Dispatcher.load()
    disp.install(st.rs, "gf8")
    enc_variant = st.rs.variant_for("encode", 0, S)
    t_enc_plain_d = timed(torch, lambda: [st.encode(s, sp) for s in range(N)], args.steps, args.warmup)
    st.rs.set_variant_table(None)
    assert torch.equal(st.buf[:, k:], ref_parity), "dispatched encode differs"

    # ---- batched encode: one batch-kernel launch, replayed from a hipGraph
    b_enc = st.batch("encode", sp)
    tune = b_enc.autotune(reps=3)
    st.buf[:, k:].zero_()
    b_enc.graph_launch()
    torch.cuda.synchronize()
    assert torch.equal(st.buf[:, k:], ref_parity), "batched encode differs from plain encode"
    t_enc_batch = timed(torch, b_enc.graph_launch, args.steps, args.warmup)
    # roofline: the kernel's own duration, HIP events around 10 direct launches of the
    # same tuned batch, on the stream it is launched on
    d_batch = event_durations(torch, lambda i: b_enc.run(), 10)
    bytes_per_launch = N * (k + m) * S
    achieved = bytes_per_launch / (float(d_batch.mean()) * 1e-3) / 1e9

    # ---- verify
    flags = torch.zeros(N, dtype=torch.int32, device="cuda")
    t_ver_plain = timed(torch, lambda: [st.verify_async(s, flags.data_ptr() + 4 * s, sp)
                                        for s in range(N)], max(1, args.steps // 2), 1)
    assert int(flags.sum()) == 0
    b_ver = st.batch("verify", sp, flags=flags)
    b_ver.autotune(reps=2)

    def ver_batch():
        flags.zero_()
        b_ver.graph_launch()

    t_ver_batch = timed(torch, ver_batch, args.steps, 1)
    assert int(flags.sum()) == 0, "verify reported a mismatch on freshly encoded stripes"

    # ---- reconstruct 1..4 erasures per stripe, in place (survivors read, erased rewritten)
    er = st.erasures(7)
    n_erased = sum(len(e) for e in er)
    slot_arrays = [st.slots(s, er[s]) for s in range(N)]
    # pristine erasure layouts: a call marks its rebuilt slots present, so each
    # timed call first restores its stripe's layout with one memmove
    layouts = [st.slots(s, er[s]) for s in range(N)]
    slot_bytes = C.sizeof(layouts[0])

    def rec_plain():
        for s in range(N):
            arr = slot_arrays[s]
            C.memmove(arr, layouts[s], slot_bytes)
            st.reconstruct(arr, sp)

    ref = st.buf.clone()

    def poison_erased():
        # the erased shards are overwritten with a pattern first, so a reconstruct
        # that leaves its targets alone fails the comparison below
        for s in range(N):
            for i in er[s]:
                st.buf[s, i].fill_(0xA5)

    poison_erased()
    t_rec_plain = timed(torch, rec_plain, max(1, args.steps // 2), 1)
    assert torch.equal(ref, st.buf), "plain reconstruct: rebuilt shards differ"
    # plain, through prepared workspaces (prepare_reconstruct_workspace once per
    # stripe's pattern; each timed call only validates and enqueues)
    workspaces = []
    for s in range(N):
        h = C.c_void_p()
        st._ok(st.lib.rsgpu_prepare_reconstruct_workspace(st.rs._h, layouts[s], k + m, sp, C.byref(h)))
        workspaces.append(rsgpu.ReconstructWorkspace(h))

    def rec_plain_ws():
        for s in range(N):
            arr = slot_arrays[s]
            C.memmove(arr, layouts[s], slot_bytes)
            st._ok(st.lib.rsgpu_reconstruct_with_workspace(st.rs._h, arr, k + m, workspaces[s]._h, sp))

    poison_erased()
    t_rec_plain_ws = timed(torch, rec_plain_ws, max(1, args.steps // 2), 1)
    assert torch.equal(ref, st.buf), "plain reconstruct with workspaces: rebuilt shards differ"
    del workspaces
    b_rec = st.batch("reconstruct", sp, erasures=er)
    b_rec.autotune(reps=2)
    poison_erased()
    t_rec_batch = timed(torch, b_rec.graph_launch, args.steps, 1)
    assert torch.equal(ref, st.buf), "batched reconstruct: rebuilt shards differ"
    del ref

    # ---- the same encode batch with shards at an exact 1 MiB pitch
    del b_rec
    st_p2 = Stripes(rsgpu, torch, k, m, N, S, seed=5, pad=0)
    b_p2 = st_p2.batch("encode", sp)
    b_p2.autotune(reps=2)
    t_enc_p2 = timed(torch, b_p2.graph_launch, args.steps, 1)
    del b_p2, st_p2

    # ---- 8+3
    del b_enc, b_ver
    st83 = Stripes(rsgpu, torch, 8, 3, N, S, seed=83)
    t83_plain = timed(torch, lambda: [st83.encode(s, sp) for s in range(N)], max(1, args.steps // 2), 1)
    b83 = st83.batch("encode", sp)
    b83.autotune(reps=2)
    t83_batch = timed(torch, b83.graph_launch, args.steps, 1)
    del b83, st83

    gbs = lambda nbytes, t: round(nbytes / t / 1e9, 2)

    # ---- ragged mix: mixed (k, m), 4 KiB-4 MiB shards, 60/20/20 encode/verify/reconstruct
    ragged = None
    if args.ragged:
        from rsgpu import workload

        del st
        spec = workload.ragged_spec(args.ragged)
        rg = workload.RaggedSet(rsgpu, torch, spec)
        rg.make_consistent(sp)
        torch.cuda.synchronize()
        # every output region (encode parity, erased shards) is poisoned before each
        # timed run and compared with the consistent bytes after it
        rg_golden = rg.output_copies()
        rg.poison_outputs()
        t_rg_plain = timed(torch, lambda: rg.run_plain(sp), 1, 1)
        assert rg.outputs_equal(rg_golden), "ragged plain: outputs differ"
        for (rk, rm), rcodec in rg.codecs.items():
            disp.install(rcodec, "gf8")
        rg.poison_outputs()
        t_rg_plain_d = timed(torch, lambda: rg.run_plain(sp), 1, 1)
        assert rg.outputs_equal(rg_golden), "ragged plain (dispatched): outputs differ"
        for rcodec in rg.codecs.values():
            rcodec.set_variant_table(None)
        b_rg = rg.make_batch(sp)
        rg_tune = b_rg.autotune(reps=2)

        def rg_batch():
            rg.flags.zero_()
            b_rg.graph_launch()

        rg.poison_outputs()
        t_rg_batch = timed(torch, rg_batch, args.steps, 1)
        assert int(rg.flags.sum()) == 0, "ragged verify reported a mismatch"
        assert rg.outputs_equal(rg_golden), "ragged batched: outputs differ"
        del rg_golden
        ragged = {"stripes": len(spec), "data_bytes": rg.data_bytes, "moved_bytes": rg.moved_bytes,
                  "plain_gbs": gbs(rg.data_bytes, t_rg_plain),
                  "batched_gbs": gbs(rg.data_bytes, t_rg_batch),
                  "batched_bytes_moved_gbs": gbs(rg.moved_bytes, t_rg_batch),
                  "plain_dispatched_gbs": gbs(rg.data_bytes, t_rg_plain_d),
                  "batched_vs_plain": round(t_rg_plain / t_rg_batch, 2),
                  "batched_vs_plain_dispatched": round(t_rg_plain_d / t_rg_batch, 2),
                  "batched_ms": round(t_rg_batch * 1e3, 3), "plain_ms": round(t_rg_plain * 1e3, 3),
                  "tune": rg_tune}
        del b_rg, rg

    # the per-call sections below are host-bound at a few microseconds a call: start
    # each without the ragged section's Python objects pending collection
    gc.collect()
    leopard = leopard_section(torch, rsgpu, DSP, disp) if args.leopard else None
    gc.collect()
    gf16_batch = gf16_batch_section(torch, rsgpu, 10)
    gc.collect()
    config1 = config1_section(torch, rsgpu, args.steps)
    gc.collect()
    gf2_mfma = gf2_mfma_section(torch, rsgpu)
    dispatch_live = None
    if args.dispatch_live:
        dispatch_live = DSP.evaluate_live(torch, disp, DSP.sample_specs(args.dispatch_live,
                                                                        seed=20261204))
    serving = serving_section(torch, [int(r) for r in args.rates.split(",")]) if args.rates else None
    host = host_section(rsgpu)

    cpu = None if args.no_cpu else cpu_baseline(k, m, S, args.cpu_stripes, args.cpu_threads, er)
    rec_moved = N * k * S + n_erased * S
    out = {
        "metric": "data GB/s RS 10+4 encode/reconstruct, batched vs plain, bit-exact; p99 ms",
        "value": gbs(data_bytes, t_enc_batch),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_enc_batch * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": f"RS 10+4 GF(2^8) encode of {N} stripes x {S // 1024} KiB shards, "
                               "device-resident, batched: one batch-kernel launch per step "
                               "(rsgpu_batch_graph_launch, which launches a one-kernel batch "
                               "directly); plain per-stripe launches reported alongside",
                   "stripes": N, "shard_bytes": S, "data_shards": k, "parity_shards": m,
                   "shard_pitch_bytes": pitch},
        "encode_batched_gbs": gbs(data_bytes, t_enc_batch),
        "encode_plain_gbs": gbs(data_bytes, t_enc_plain),
        "encode_plain_dispatched_gbs": gbs(data_bytes, t_enc_plain_d),
        "encode_plain_dispatched_variant": enc_variant,
        "encode_batched_bytes_moved_gbs": gbs(N * (k + m) * S, t_enc_batch),
        "encode_batched_gbs_pow2_pitch": gbs(data_bytes, t_enc_p2),
        "encode_batched_vs_plain": round(t_enc_plain / t_enc_batch, 2),
        "verify_batched_gbs": gbs(data_bytes, t_ver_batch),
        "verify_plain_gbs": gbs(data_bytes, t_ver_plain),
        "reconstruct_batched_gbs": gbs(data_bytes, t_rec_batch),
        "reconstruct_plain_gbs": gbs(data_bytes, t_rec_plain),
        "reconstruct_plain_workspace_gbs": gbs(data_bytes, t_rec_plain_ws),
        "reconstruct_batched_bytes_moved_gbs": gbs(rec_moved, t_rec_batch),
        "encode_8x3_batched_gbs": gbs(N * 8 * S, t83_batch),
        "encode_8x3_plain_gbs": gbs(N * 8 * S, t83_plain),
        "config1_64x10+4x64KiB": config1,
        "ragged": ragged,
        "dispatch_live": dispatch_live,
        "leopard": leopard,
        "gf16_batch": gf16_batch,
        "gf2_mfma": gf2_mfma,
        "serving": serving,
        "host_boundary": host,
        "stripe_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                              "p99": round(float(np.percentile(lat, 99)), 4)},
        "batch_tune": tune,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(bytes_per_launch),
                     "traffic_source": "rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE runs of this kernel and "
                                       "workload, scaled by a 1 GiB copy (profiles/"
                                       "batch_kernel_encode_10x4_1MiB_traffic.json); not re-measured here",
                     "kernel": "batch_kernel (512 x RS 10+4 x 1 MiB encode, one launch)",
                     "launches_timed": int(len(d_batch)),
                     "bytes_per_launch": bytes_per_launch,
                     "avg_launch_ms": round(float(d_batch.mean()), 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
