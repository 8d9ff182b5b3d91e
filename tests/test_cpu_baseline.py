"""The benchmark's CPU baseline (oracle/cpu_simd.c: GFNI / AVX2 / scalar row kernels,
multithreaded) must produce the scalar oracle's parity, verify verdicts and rebuilt
shards for every instruction set this host has, every thread count and tail length."""
import numpy as np
import pytest

import oracle_lib as O


@pytest.mark.parametrize("k,m,n", [(10, 4, 1 << 20), (4, 2, 16384 + 33), (20, 4, 70000), (3, 9, 100),
                                   (12, 7, 65536 * 5 + 64)])
@pytest.mark.parametrize("threads", [1, 4])
def test_cpu_baseline_matches_oracle(k, m, n, threads):
    rng = np.random.default_rng(k * 1000 + n)
    data = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k)]
    rc, mat = O.build_matrix(O.GF8, O.VANDERMONDE, k, m)
    exp = O.encode(O.GF8, mat, k, m, data)
    best = O.lib().ora_cpu_isa()
    for isa in range(1, best + 1):
        used, par = O.cpu_encode(mat, k, m, data, threads=threads, isa=isa)
        assert used == isa
        for j in range(m):
            assert np.array_equal(par[j], exp[j]), (O.CPU_ISA[isa], j)


@pytest.mark.parametrize("k,m,n", [(10, 4, 1 << 20), (4, 2, 16384 + 33), (20, 4, 70000), (6, 3, 100)])
@pytest.mark.parametrize("threads", [1, 4])
def test_cpu_baseline_verify_and_reconstruct_match_oracle(k, m, n, threads):
    rng = np.random.default_rng(k * 100 + n + threads)
    data = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k)]
    rc, mat = O.build_matrix(O.GF8, O.VANDERMONDE, k, m)
    full = data + O.encode(O.GF8, mat, k, m, data)
    best = O.lib().ora_cpu_isa()
    for isa in range(1, best + 1):
        used, ok = O.cpu_verify(mat, k, m, full, threads=threads, isa=isa)
        assert used == isa and ok
        bad = [x.copy() for x in full]
        bad[k + m - 1][n - 1] ^= 1
        assert not O.cpu_verify(mat, k, m, bad, threads=threads, isa=isa)[1]
        for missing in ({0}, {k}, {0, k + m - 1}, set(range(min(m, k))), set(range(k, k + m))):
            present = [0 if i in missing else 1 for i in range(k + m)]
            rows = O.cpu_data_rows(mat, k, m, present)
            work = [x.copy() if present[i] else np.full(n, 0xA5, np.uint8) for i, x in enumerate(full)]
            assert O.cpu_reconstruct(mat, k, m, work, present, rows, threads=threads, isa=isa) == isa
            for i in missing:
                assert np.array_equal(work[i], full[i]), (O.CPU_ISA[isa], sorted(missing), i)
        # non-codeword input: the same bytes as the oracle's reconstruct
        noise = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k + m)]
        present = [0 if i in (1, k + 1) else 1 for i in range(k + m)]
        rc, exp = O.reconstruct(O.GF8, mat, k, m, [x.copy() for x in noise], present)
        work = [x.copy() for x in noise]
        O.cpu_reconstruct(mat, k, m, work, present, O.cpu_data_rows(mat, k, m, present),
                          threads=threads, isa=isa)
        for i in (1, k + 1):
            assert np.array_equal(work[i], exp[i]), (O.CPU_ISA[isa], i)
