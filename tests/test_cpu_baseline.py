"""The benchmark's CPU baseline (oracle/cpu_simd.c: GFNI / AVX2 / scalar row kernels,
multithreaded) must produce the scalar oracle's parity for every instruction set
this host has, every thread count and tail length."""
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
