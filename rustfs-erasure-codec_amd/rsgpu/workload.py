"""Synthetic Reed-Solomon workloads of uniform random bytes (device-resident).

ragged(): thousands of stripes with mixed (k, m), log-uniform shard sizes and a mix of
encode / verify / reconstruct, laid out in one device allocation, with helpers to run
them one stripe per launch ("plain") or as one batch-kernel launch.
"""
import ctypes as C
import math

import numpy as np

RAGGED_SHAPES = [(4, 2), (6, 3), (8, 3), (8, 4), (10, 4), (12, 4), (16, 4), (20, 4)]


def log_uniform_size(rng, lo, hi, align=64):
    s = math.exp(rng.uniform(math.log(lo), math.log(hi)))
    return max(align, int(round(s / align)) * align)


def ragged_spec(n_stripes, seed=0x3A66ED, shapes=RAGGED_SHAPES, lo=4 << 10, hi=4 << 20,
                mix=(0.6, 0.2, 0.2)):
    """[(k, m, shard_bytes, op, erased positions)] with op in encode/verify/reconstruct."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_stripes):
        k, m = shapes[int(rng.integers(len(shapes)))]
        size = log_uniform_size(rng, lo, hi)
        u = rng.random()
        op = "encode" if u < mix[0] else ("verify" if u < mix[0] + mix[1] else "reconstruct")
        erased = []
        if op == "reconstruct":
            e = int(rng.integers(0, m + 1))
            erased = sorted(rng.choice(k + m, e, replace=False).tolist())
        out.append((k, m, size, op, erased))
    return out


class RaggedSet:
    """The stripes of a spec in one device buffer (pitch = shard + 256 B) with prebuilt
    C-ABI argument arrays for the plain and the batched paths."""

    def __init__(self, rsgpu, torch, spec, seed=7):
        self.rsgpu, self.torch, self.spec = rsgpu, torch, spec
        self.lib = rsgpu.lib()
        self.codecs = {}
        offs, total = [], 0
        for k, m, size, _, _ in spec:
            if (k, m) not in self.codecs:
                self.codecs[(k, m)] = rsgpu.ReedSolomon(k, m)
            pitch = size + 256
            offs.append(total)
            total += (k + m) * pitch
        self.buf = torch.empty(total, dtype=torch.uint8, device="cuda")
        g = torch.Generator(device="cuda").manual_seed(seed)
        self.buf.random_(0, 256, generator=g)
        self.offs = offs
        self.flags = torch.zeros(max(1, len(spec)), dtype=torch.int32, device="cuda")
        base = self.buf.data_ptr()
        self.args = []
        for idx, ((k, m, size, op, erased), off) in enumerate(zip(spec, offs)):
            ptrs = [base + off + i * (size + 256) for i in range(k + m)]
            if op == "encode":
                a = (rsgpu.Slice * (k + m))(*[rsgpu.Slice(p, size) for p in ptrs])
            elif op == "verify":
                a = (rsgpu.CSlice * (k + m))(*[rsgpu.CSlice(p, size) for p in ptrs])
            else:
                a = (rsgpu.Slot * (k + m))(*[rsgpu.Slot(p, size, 0 if i in erased else 1)
                                             for i, p in enumerate(ptrs)])
            self.args.append(a)
        self.data_bytes = sum(k * size for k, m, size, _, _ in spec)
        self.moved_bytes = sum((k + (m if op != "reconstruct" else len(er))) * size
                               for k, m, size, op, er in spec)

    def _ok(self, rc):
        if rc:
            raise self.rsgpu.RsError(rc)

    def make_consistent(self, stream):
        """Encode every stripe once so that verify passes and reconstruct rebuilds the
        bytes already in place (repeat runs then rewrite identical bytes)."""
        b = self.rsgpu.Batch()
        base = self.buf.data_ptr()
        for (k, m, size, _, _), off in zip(self.spec, self.offs):
            ptrs = [base + off + i * (size + 256) for i in range(k + m)]
            arr = (self.rsgpu.Slice * (k + m))(*[self.rsgpu.Slice(p, size) for p in ptrs])
            self._ok(self.lib.rsgpu_batch_add_encode(b._h, self.codecs[(k, m)]._h, arr, k + m))
        self._ok(self.lib.rsgpu_batch_run(b._h, stream))

    def output_regions(self):
        """(offset, length) of every byte range an op writes: encode parity shards and
        the erased shards of reconstruct ops."""
        out = []
        for (k, m, size, op, er), off in zip(self.spec, self.offs):
            targets = range(k, k + m) if op == "encode" else (er if op == "reconstruct" else [])
            out += [(off + i * (size + 256), size) for i in targets]
        return out

    def output_copies(self):
        return [self.buf[o:o + n].clone() for o, n in self.output_regions()]

    def poison_outputs(self, value=0xA5):
        for o, n in self.output_regions():
            self.buf[o:o + n].fill_(value)

    def outputs_equal(self, copies):
        return all(self.torch.equal(self.buf[o:o + n], c)
                   for (o, n), c in zip(self.output_regions(), copies))

    def reset_slots(self):
        for (k, m, size, op, erased), a in zip(self.spec, self.args):
            if op == "reconstruct":
                for i in range(k + m):
                    a[i].present = 0 if i in erased else 1

    def run_plain(self, stream):
        """One C-ABI call (one launch) per stripe."""
        self.reset_slots()
        lib = self.lib
        fptr = self.flags.data_ptr()
        for idx, ((k, m, size, op, erased), a) in enumerate(zip(self.spec, self.args)):
            h = self.codecs[(k, m)]._h
            if op == "encode":
                rc = lib.rsgpu_encode(h, a, k + m, stream)
            elif op == "verify":
                rc = lib.rsgpu_verify_async(h, a, k + m, fptr + 4 * idx, stream)
            else:
                rc = lib.rsgpu_reconstruct(h, a, k + m, stream)
            self._ok(rc)

    def make_batch(self, stream):
        self.reset_slots()
        b = self.rsgpu.Batch()
        lib = self.lib
        fptr = self.flags.data_ptr()
        for idx, ((k, m, size, op, erased), a) in enumerate(zip(self.spec, self.args)):
            h = self.codecs[(k, m)]._h
            if op == "encode":
                rc = lib.rsgpu_batch_add_encode(b._h, h, a, k + m)
            elif op == "verify":
                rc = lib.rsgpu_batch_add_verify(b._h, h, a, k + m, fptr + 4 * idx)
            else:
                rc = lib.rsgpu_batch_add_reconstruct(b._h, h, a, k + m, 0, stream)
            self._ok(rc)
        return b
