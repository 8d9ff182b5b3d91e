"""Benchmark of the Reed-Solomon data path on one MI355X.

Workload (one step): 512 stripes of RS 10+4 with 1 MiB shards, device-resident uniform
random bytes, each stripe encoded by its own launch ("plain" per-stripe launches).
The same run also measures verify and reconstruct (1..4 random erasures per stripe),
the 8+3 shape, per-stripe latency (p50/p99), the roofline of the encode kernel and a
CPU baseline (the oracle's restatement of the reference algorithm) on a bounded sample.

Prints ONE JSON line.  `value` = data bytes (k * shard) encoded per second.
"""
import argparse
import ctypes as C
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
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--stripes", type=int, default=512)
    ap.add_argument("--shard", type=int, default=1 << 20)
    ap.add_argument("--cpu-stripes", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


class Stripes:
    """n stripes of k+m shards in one device allocation, with prebuilt C-ABI views."""

    def __init__(self, rsgpu, torch, k, m, n, shard, seed):
        self.k, self.m, self.n, self.shard = k, m, n, shard
        g = torch.Generator(device="cuda").manual_seed(seed)
        self.buf = torch.empty((n, k + m, shard), dtype=torch.uint8, device="cuda")
        self.buf[:, :k].random_(0, 256, generator=g)
        self.rs = rsgpu.ReedSolomon(k, m)
        self.lib = rsgpu.lib()
        base, stride_s, stride_i = self.buf.data_ptr(), (k + m) * shard, shard
        self.enc = []
        self.ver = []
        for s in range(n):
            arr = (rsgpu.Slice * (k + m))(*[rsgpu.Slice(base + s * stride_s + i * stride_i, shard)
                                            for i in range(k + m)])
            self.enc.append(arr)
            carr = (rsgpu.CSlice * (k + m))(*[rsgpu.CSlice(base + s * stride_s + i * stride_i, shard)
                                              for i in range(k + m)])
            self.ver.append(carr)
        self.rsgpu = rsgpu

    def encode(self, s, stream):
        rc = self.lib.rsgpu_encode(self.rs._h, self.enc[s], self.k + self.m, stream)
        if rc:
            raise self.rsgpu.RsError(rc)

    def verify_async(self, s, flag_ptr, stream):
        rc = self.lib.rsgpu_verify_async(self.rs._h, self.ver[s], self.k + self.m, flag_ptr, stream)
        if rc:
            raise self.rsgpu.RsError(rc)

    def erasure_slots(self, seed):
        """Per-stripe slot arrays with 1..m erasures at random positions (data and parity
        mixed); the erased shards are rebuilt into a separate output buffer."""
        import numpy as np

        rng = np.random.default_rng(seed)
        self.rbuf = None
        slots = []
        base = self.buf.data_ptr()
        n_erased = 0
        for s in range(self.n):
            e = int(rng.integers(1, self.m + 1))
            missing = set(rng.choice(self.k + self.m, e, replace=False).tolist())
            n_erased += e
            arr = (self.rsgpu.Slot * (self.k + self.m))()
            for i in range(self.k + self.m):
                ptr = base + (s * (self.k + self.m) + i) * self.shard
                arr[i] = self.rsgpu.Slot(ptr, self.shard, 0 if i in missing else 1)
            slots.append((arr, missing))
        return slots, n_erased

    def reconstruct(self, arr, stream):
        for i in range(self.k + self.m):
            pass
        rc = self.lib.rsgpu_reconstruct(self.rs._h, arr, self.k + self.m, stream)
        if rc:
            raise self.rsgpu.RsError(rc)


def timed(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def per_launch_events(torch, fn, count):
    """Average / p50 / p99 device duration of single launches via events on torch's stream."""
    import numpy as np

    stream = torch.cuda.current_stream()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(count)]
    for i in range(count):
        evs[i][0].record(stream)
        fn(i)
        evs[i][1].record(stream)
    torch.cuda.synchronize()
    d = np.array([a.elapsed_time(b) for a, b in evs])
    return float(d.mean()), float(np.percentile(d, 50)), float(np.percentile(d, 99))


def cpu_baseline(k, m, shard, stripes):
    """Oracle (scalar C restatement of the reference's code_some_slices) on a bounded sample."""
    import numpy as np

    import oracle_lib as O

    _, mat = O.build_matrix(O.GF8, O.VANDERMONDE, k, m)
    rng = np.random.default_rng(1)
    data = [rng.integers(0, 256, shard, dtype=np.uint8) for _ in range(k)]
    O.encode(O.GF8, mat, k, m, data)  # warm the tables
    t0 = time.perf_counter()
    for _ in range(stripes):
        O.encode(O.GF8, mat, k, m, data)
    dt = time.perf_counter() - t0
    return {"value": round(stripes * k * shard / dt / 1e9, 3), "unit": "GB/s", "cores": 1,
            "kind": "port",
            "sample": f"{stripes} stripes of RS {k}+{m} x {shard // 1024} KiB encode, scalar oracle, 1 thread"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(stripes shard across GPUs with no exchange; see DESIGN.md)"}))
        sys.exit(2)
    import numpy as np
    import torch

    import rsgpu

    torch.cuda.init()
    stream = torch.cuda.current_stream()
    sp = C.c_void_p(stream.cuda_stream)
    k, m, S, N = 10, 4, args.shard, args.stripes

    st = Stripes(rsgpu, torch, k, m, N, S, seed=0x00EC5EED)

    def enc_all():
        for s in range(N):
            st.encode(s, sp)

    t_enc = timed(torch, enc_all, args.steps, args.warmup)
    enc_gbs = N * k * S / t_enc / 1e9

    # roofline + latency of single encode launches (events on the launch stream)
    avg_ms, p50_ms, p99_ms = per_launch_events(torch, lambda i: st.encode(i % N, sp), min(N, 256))
    bytes_per_launch = (k + m) * S
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9

    # verify (register compare, no parity writes)
    flags = torch.zeros(N, dtype=torch.int32, device="cuda")

    def ver_all():
        for s in range(N):
            st.verify_async(s, flags.data_ptr() + 4 * s, sp)

    t_ver = timed(torch, ver_all, max(1, args.steps // 2), 1)
    assert int(flags.sum()) == 0, "verify reported a mismatch on freshly encoded stripes"
    ver_gbs = N * k * S / t_ver / 1e9

    # reconstruct 1..m random erasures per stripe (in place: survivors read, erased rewritten)
    slots, n_erased = st.erasure_slots(7)
    ref = st.buf[: min(N, 8)].clone()

    def rec_all():
        for arr, _ in slots:
            for i in range(k + m):
                arr[i].present = 0 if i in _ else 1
            st.reconstruct(arr, sp)

    t_rec = timed(torch, rec_all, max(1, args.steps // 2), 1)
    assert torch.equal(ref, st.buf[: ref.shape[0]]), "reconstruct changed stripe bytes"
    rec_gbs = N * k * S / t_rec / 1e9
    rec_moved = (N * k * S + n_erased * S) / t_rec / 1e9

    # 8+3 plain encode
    del slots
    st83 = Stripes(rsgpu, torch, 8, 3, N, S, seed=83)

    def enc83():
        for s in range(N):
            st83.encode(s, sp)

    t83 = timed(torch, enc83, max(1, args.steps // 2), 1)
    enc83_gbs = N * 8 * S / t83 / 1e9
    del st83

    cpu = None if args.no_cpu else cpu_baseline(k, m, S, args.cpu_stripes)

    out = {
        "metric": "data GB/s RS 10+4 encode/reconstruct, batched vs plain, bit-exact; p99 ms",
        "value": round(enc_gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_enc * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": f"RS 10+4 GF(2^8) encode, {N} stripes x {S // 1024} KiB shards, "
                               "device-resident, one launch per stripe",
                   "stripes": N, "shard_bytes": S, "data_shards": k, "parity_shards": m},
        "encode_plain_gbs": round(enc_gbs, 2),
        "encode_plain_bytes_moved_gbs": round(enc_gbs * (k + m) / k, 2),
        "verify_plain_gbs": round(ver_gbs, 2),
        "reconstruct_plain_gbs": round(rec_gbs, 2),
        "reconstruct_plain_bytes_moved_gbs": round(rec_moved, 2),
        "encode_8x3_plain_gbs": round(enc83_gbs, 2),
        "stripe_latency_ms": {"p50": round(p50_ms, 4), "p99": round(p99_ms, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "apply_kernel (encode 10+4 x 1 MiB, one stripe per launch)",
                     "bytes_per_launch": bytes_per_launch, "avg_launch_ms": round(avg_ms, 5)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
