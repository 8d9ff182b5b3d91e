"""Shared helpers for the GPU parity tests (seeded inputs, torch <-> numpy)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


def dev():
    return torch.device("cuda:0")


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.uint8)).to(dev())


def to_host(t):
    return t.detach().cpu().numpy()


def rand_shards(rng, n, length):
    return [rng.integers(0, 256, length, dtype=np.uint8) for _ in range(n)]


def announce(msg):
    print(f"[gpu-case] {msg}", flush=True)
