"""Pins the CPU oracle against the reference's own known answers.

The known answers come from the tests of rustfs-erasure-codec 9.1.0, kept as data in
tests/golden/reference_vectors.json with the file:line of each; the other checks follow
from closed forms of the mathematics.  The oracle is
the checker for every GPU parity test, so these run on the CPU only.
"""
import itertools

import numpy as np
import pytest

import oracle_lib as O
from fixtures import golden, hexint

# ------------------------------------------------------------------ GF(2^8)
def test_gf8_log_table_same_as_backblaze():
    log, _ = O.gf8_tables()
    assert list(log) == golden("gf8_log_table")["log"]


def test_gf8_known_answers():
    g = golden("gf8_arithmetic")
    for x, y, want in g["mul"]:
        assert O.gf8_mul(x, y) == want
    for x, n, want in g["exp"]:
        assert O.gf8_exp(x, n) == want
    inp = np.array(g["slice_input"], np.uint8)
    out = None
    for step in g["slice_steps"]:
        if step["xor"]:
            O.gf8_mul_slice(step["c"], inp, out, xor=True)
        else:
            out = O.gf8_mul_slice(step["c"], inp)
        assert list(out) == step["out"], step["c"]


def _carryless_mod(a, b, poly=0x11D):
    """Textbook GF(2)[x] multiply modulo the field polynomial."""
    r = 0
    for i in range(8):
        if (b >> i) & 1:
            r ^= a << i
    for i in range(15, 7, -1):
        if (r >> i) & 1:
            r ^= poly << (i - 8)
    return r


def test_gf8_mul_matches_carryless_polynomial_product():
    # closed form: the table multiply is polynomial multiplication mod 0x11d
    for a in range(256):
        for b in range(0, 256, 7):
            assert O.gf8_mul(a, b) == _carryless_mod(a, b)


def test_gf8_div_exp_identities():
    for a in range(1, 256):
        assert O.gf8_mul(a, O.gf8_div(1, a)) == 1
        assert O.gf8_div(a, 0) == 0
        assert O.gf8_exp(a, 255) == 1
    assert O.gf8_exp(0, 0) == 1 and O.gf8_exp(0, 5) == 0


# ------------------------------------------------------------------ matrices
def test_matrix_inverse_known_answers():
    for case in golden("matrix_inverse")["cases"]:
        rc, inv = O.matrix_invert(O.GF8, case["matrix"])
        if case["inverse"] is None:
            assert rc != 0
        else:
            assert rc == 0
            assert inv.tolist() == case["inverse"]


def test_systematic_matrix_rows():
    # parity rows that reproduce the 4+2 and smoke golden checksums below
    rc, m42 = O.build_matrix(O.GF8, O.VANDERMONDE, 4, 2)
    assert rc == 0
    assert m42[:4].tolist() == np.eye(4, dtype=int).tolist()
    assert m42[4:].tolist() == [[27, 28, 18, 20], [28, 27, 20, 18]]
    rc, m104 = O.build_matrix(O.GF8, O.VANDERMONDE, 10, 4)
    assert m104[10:].tolist() == [[129, 150, 175, 184, 210, 196, 254, 232, 3, 2],
                                  [150, 129, 184, 175, 196, 210, 232, 254, 2, 3],
                                  [191, 214, 98, 10, 6, 111, 223, 183, 5, 4],
                                  [214, 191, 10, 98, 111, 6, 183, 223, 4, 5]]


# ----------------------------------------------------- classic golden vectors
def fixed_4x2():
    return [np.arange(16 * i, 16 * i + 16, dtype=np.uint8) for i in range(4)]


def test_golden_encode_4x2_incrementing_input():
    g = golden("classic_4x2")
    _, mat = O.build_matrix(O.GF8, O.VANDERMONDE, 4, 2)
    data = fixed_4x2()
    parity = O.encode(O.GF8, mat, 4, 2, data)
    assert [p.tolist() for p in parity] == g["parity"]
    assert O.checksum_shards(data + parity) == hexint(g["codeword_checksum"])


def test_golden_reconstruct_4x2():
    g = golden("classic_4x2")
    _, mat = O.build_matrix(O.GF8, O.VANDERMONDE, 4, 2)
    data = fixed_4x2()
    full = data + O.encode(O.GF8, mat, 4, 2, data)
    # reconstruct_data with data shard 1 and parity shard 0 missing
    missing = g["reconstruct_data_missing"]
    shards = [s.copy() for s in full]
    for i in missing:
        shards[i][:] = 0
    present = [0 if i in missing else 1 for i in range(6)]
    rc, out = O.reconstruct(O.GF8, mat, 4, 2, shards, present, data_only=True)
    assert rc == 0
    assert out[1].tolist() == list(range(16, 32))
    recovered = [out[i] if i < 4 or i not in missing else np.zeros(0, np.uint8) for i in range(6)]
    assert O.checksum_shards(recovered) == hexint(g["reconstruct_data_checksum"])
    # full reconstruct
    missing = g["reconstruct_missing"]
    shards = [s.copy() for s in full]
    for i in missing:
        shards[i][:] = 0
    rc, out = O.reconstruct(O.GF8, mat, 4, 2, shards, [0 if i in missing else 1 for i in range(6)])
    assert rc == 0 and all((a == b).all() for a, b in zip(out, full))
    assert O.checksum_shards(out) == hexint(g["codeword_checksum"])


def test_golden_verify_detects_corruption():
    c = golden("classic_4x2")["verify_corrupt"]
    _, mat = O.build_matrix(O.GF8, O.VANDERMONDE, 4, 2)
    data = fixed_4x2()
    full = data + O.encode(O.GF8, mat, 4, 2, data)
    assert O.verify(O.GF8, mat, 4, 2, full)
    full[c["shard"]][c["byte"]] ^= c["xor"]
    assert not O.verify(O.GF8, mat, 4, 2, full)


def _split(data, k):
    # ReedSolomon::split for Classic (src/core/codec.rs:486-508)
    n = -(-len(data) // k)
    out = []
    for i in range(k):
        s = np.zeros(n, np.uint8)
        chunk = data[i * n:(i + 1) * n]
        s[:len(chunk)] = chunk
        out.append(s)
    return out


@pytest.mark.parametrize("kind", ["seeded", "zeros", "pattern"])
def test_golden_split_encode(kind):
    # 61-byte payloads through split + encode
    g = golden("split_encode_4x2")
    n = g["payload_bytes"]
    if kind == "seeded":
        payload = O.smallrng_bytes(hexint(g["seeded"]["seed"]), n)
    elif kind == "zeros":
        payload = np.zeros(n, np.uint8)
    else:
        payload = np.resize(np.array(g["pattern"]["repeat"], np.uint8), n)
    _, mat = O.build_matrix(O.GF8, O.VANDERMONDE, 4, 2)
    data = _split(payload, 4)
    parity = O.encode(O.GF8, mat, 4, 2, data)
    assert O.verify(O.GF8, mat, 4, 2, data + parity)
    assert O.checksum_shards(data + parity) == hexint(g[kind]["checksum"])


def test_qc_mix_known_answer():
    g = golden("qc_mix")
    seed = O.qc_seed(g["parts"])
    assert O.qc_fill(len(g["fill"]), seed ^ hexint(g["xor"])).tolist() == g["fill"]


def test_corrupt_positions_is_partial_permutation():
    for seed in range(20):
        pos = O.corrupt_positions(4, 14, seed)
        assert len(set(pos)) == 4 and all(0 <= p < 14 for p in pos)


def test_decode_matrix_is_inverse_of_survivor_rows():
    # mirrors src/core/reconstruct/tests.rs:175-360: every small survivor subset
    k, m = 4, 3
    _, mat = O.build_matrix(O.GF8, O.VANDERMONDE, k, m)
    for missing in itertools.chain.from_iterable(itertools.combinations(range(k + m), e)
                                                 for e in range(1, m + 1)):
        present = [0 if i in missing else 1 for i in range(k + m)]
        rc, valid, dec = O.decode_matrix(O.GF8, mat, k, m, present)
        assert rc == 0
        assert valid == [i for i in range(k + m) if present[i]][:k]
        sel = mat[valid]
        prod = np.zeros((k, k), np.uint16)
        for r in range(k):
            for c in range(k):
                acc = 0
                for i in range(k):
                    acc ^= O.gf8_mul(int(dec[r, i]), int(sel[i, c]))
                prod[r, c] = acc
        assert (prod == np.eye(k, dtype=np.uint16)).all()


@pytest.mark.parametrize("mode", [O.VANDERMONDE, O.CAUCHY, O.JERASURE])
@pytest.mark.parametrize("field", [O.GF8, O.GF16])
def test_classic_roundtrip_every_erasure_pattern(mode, field):
    # MDS property: any <= m erasures reconstruct the codeword (src/tests/mod.rs round trips)
    k, m, n = 5, 3, 64
    rc, mat = O.build_matrix(field, mode, k, m)
    assert rc == 0
    rng = np.random.default_rng(1 + mode + field)
    data = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k)]
    full = data + O.encode(field, mat, k, m, data)
    assert O.verify(field, mat, k, m, full)
    for e in range(1, m + 1):
        for missing in itertools.combinations(range(k + m), e):
            shards = [s.copy() if i not in missing else np.zeros(n, np.uint8)
                      for i, s in enumerate(full)]
            present = [0 if i in missing else 1 for i in range(k + m)]
            rc, out = O.reconstruct(field, mat, k, m, shards, present)
            assert rc == 0
            assert all((a == b).all() for a, b in zip(out, full)), missing
    # too many erasures
    shards = [s.copy() for s in full]
    rc, _ = O.reconstruct(field, mat, k, m, shards, [0] * (m + 1) + [1] * (k - 1))
    assert rc == 10  # TooFewShardsPresent


def test_classic_linearity():
    k, m, n = 10, 4, 256
    _, mat = O.build_matrix(O.GF8, O.VANDERMONDE, k, m)
    rng = np.random.default_rng(5)
    a = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k)]
    b = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k)]
    pa = O.encode(O.GF8, mat, k, m, a)
    pb = O.encode(O.GF8, mat, k, m, b)
    pab = O.encode(O.GF8, mat, k, m, [x ^ y for x, y in zip(a, b)])
    assert all(((x ^ y) == z).all() for x, y, z in zip(pa, pb, pab))


# ------------------------------------- closed forms of the matrix constructions
_FIELD_OPS = {O.GF8: (O.gf8_mul, O.gf8_div), O.GF16: (O.gf16_mul, O.gf16_div)}


def _prod(mul, vals):
    acc = 1
    for v in vals:
        acc = mul(acc, v)
    return acc


def _det(mul, div, a):
    """Determinant by elimination (characteristic 2: row swaps keep the sign)."""
    a = [list(map(int, row)) for row in a]
    n, det = len(a), 1
    for c in range(n):
        p = next((r for r in range(c, n) if a[r][c]), None)
        if p is None:
            return 0
        a[c], a[p] = a[p], a[c]
        det = mul(det, a[c][c])
        for r in range(c + 1, n):
            f = div(a[r][c], a[c][c])
            a[r] = [x ^ mul(f, y) for x, y in zip(a[r], a[c])]
    return det


@pytest.mark.parametrize("field", [O.GF8, O.GF16])
@pytest.mark.parametrize("k,m", [(5, 3), (10, 4), (3, 1)])
def test_vandermonde_parity_is_lagrange_basis(field, k, m):
    """systematic rows of build_matrix (src/core/codec.rs:66-88): V * Vtop^-1 takes
    data row i to the Lagrange basis polynomial L_i over the points nth(0..k-1), so
    parity row r, column i is L_i(nth(r)) = prod_{j != i} (r + j) / (i + j)."""
    mul, div = _FIELD_OPS[field]
    rc, mat = O.build_matrix(field, O.VANDERMONDE, k, m)
    assert rc == 0
    assert mat[:k].tolist() == np.eye(k, dtype=int).tolist()
    for r in range(k, k + m):
        for i in range(k):
            others = [j for j in range(k) if j != i]
            want = div(_prod(mul, [r ^ j for j in others]), _prod(mul, [i ^ j for j in others]))
            assert int(mat[r, i]) == want, (r, i)


@pytest.mark.parametrize("field", [O.GF8, O.GF16])
@pytest.mark.parametrize("k,m", [(5, 3), (10, 4), (4, 1)])
def test_cauchy_parity_obeys_cauchy_determinant(field, k, m):
    """build_cauchy_matrix (src/core/codec.rs:90-112): entry (r, c) is the inverse of
    nth(r) + nth(c), and every square block of the parity rows has the Cauchy
    determinant prod_{i<j}(x_i + x_j) prod_{i<j}(y_i + y_j) / prod_{i,j}(x_i + y_j)."""
    mul, div = _FIELD_OPS[field]
    rc, mat = O.build_matrix(field, O.CAUCHY, k, m)
    assert rc == 0
    assert mat[:k].tolist() == np.eye(k, dtype=int).tolist()
    for r in range(k, k + m):
        for c in range(k):
            assert mul(int(mat[r, c]), r ^ c) == 1, (r, c)
    for size in range(1, m + 1):
        for rows in itertools.combinations(range(k, k + m), size):
            for cols in itertools.combinations(range(k), size):
                num = _prod(mul, [a ^ b for a, b in itertools.combinations(rows, 2)])
                num = mul(num, _prod(mul, [a ^ b for a, b in itertools.combinations(cols, 2)]))
                den = _prod(mul, [x ^ y for x in rows for y in cols])
                block = [[mat[x, y] for y in cols] for x in rows]
                assert _det(mul, div, block) == div(num, den), (rows, cols)


@pytest.mark.parametrize("field", [O.GF8, O.GF16])
@pytest.mark.parametrize("k,m", [(5, 3), (10, 4), (4, 1), (6, 2), (12, 4)])
def test_jerasure_like_parity_closed_form(field, k, m):
    """build_jerasure_like_matrix (src/core/codec.rs:114-197).  Its top k rows are the
    untouched Vandermonde rows over nth(0..k-1), so the column reduction is a right
    multiplication by Vtop^-1 and never swaps rows.  Before the scalings, parity row r
    (r < k+m-1) is [L_i(nth(r))] and the last row, e_{k-1} * Vtop^-1, is the leading
    coefficients 1 / prod_{j != i} (i + j).  Columns are then scaled so row k is all
    ones, and rows after k so column 0 is one."""
    mul, div = _FIELD_OPS[field]
    rc, mat = O.build_matrix(field, O.JERASURE, k, m)
    assert rc == 0
    assert mat[:k].tolist() == np.eye(k, dtype=int).tolist()

    def lagrange(r, i):
        others = [j for j in range(k) if j != i]
        return div(_prod(mul, [r ^ j for j in others]), _prod(mul, [i ^ j for j in others]))

    raw = [[lagrange(r, i) for i in range(k)] for r in range(k, k + m - 1)]
    raw.append([div(1, _prod(mul, [i ^ j for j in range(k) if j != i])) for i in range(k)])
    col = [div(1, v) for v in raw[0]]
    scaled = [[mul(v, s) for v, s in zip(row, col)] for row in raw]
    want = [scaled[0]] + [[div(v, row[0]) for v in row] for row in scaled[1:]]
    assert mat[k:].tolist() == want
    assert mat[k].tolist() == [1] * k
    assert all(int(mat[r, 0]) == 1 for r in range(k, k + m))


def test_gf16_classic_encode_element_layout():
    """A GF(2^16) element is two bytes, high byte first: one() = [0, 1] and
    nth(n) = [n >> 8, n & 255] (src/galois_16.rs:55-65).  A unit element at
    position e of data shard i therefore encodes to the matrix coefficient
    (r, i) at position e of parity row r, high byte first; a general element
    encodes to the scalar tower product, element by element."""
    k, m, n = 6, 3, 16
    rc, mat = O.build_matrix(O.GF16, O.VANDERMONDE, k, m)
    assert rc == 0
    for i in range(k):
        data = [np.zeros(2 * n, np.uint8) for _ in range(k)]
        data[i][2 * 5:2 * 5 + 2] = [0, 1]
        parity = O.encode(O.GF16, mat, k, m, data)
        for r in range(m):
            coef = int(mat[k + r, i])
            want = np.zeros(2 * n, np.uint8)
            want[10:12] = [coef >> 8, coef & 255]
            assert (parity[r] == want).all(), (i, r)
    rng = np.random.default_rng(11)
    data = [rng.integers(0, 256, 2 * n, dtype=np.uint8) for _ in range(k)]
    parity = O.encode(O.GF16, mat, k, m, data)
    for r in range(m):
        for e in range(n):
            acc = 0
            for i in range(k):
                acc ^= O.gf16_mul(int(mat[k + r, i]), (int(data[i][2 * e]) << 8) | int(data[i][2 * e + 1]))
            assert (int(parity[r][2 * e]) << 8) | int(parity[r][2 * e + 1]) == acc, (r, e)


# ---------------------------------------------------------------- GF(2^16) tower
def test_gf16_tower_field_axioms():
    rng = np.random.default_rng(7)
    vals = [0, 1, 2, 0x100, 0xFFFF] + [int(x) for x in rng.integers(0, 65536, 60)]
    for a in vals:
        if a:
            assert O.gf16_mul(a, O.gf16_div(1, a)) == 1
        for b in vals[:20]:
            assert O.gf16_mul(a, b) == O.gf16_mul(b, a)
            for c in vals[:6]:
                assert O.gf16_mul(a, b ^ c) == O.gf16_mul(a, b) ^ O.gf16_mul(a, c)
                assert O.gf16_mul(O.gf16_mul(a, b), c) == O.gf16_mul(a, O.gf16_mul(b, c))
    # X^2 = 2X + 128 (src/galois_16.rs:9-11)
    assert O.gf16_mul(0x100, 0x100) == (2 << 8) | 128
    # multiplicative group of order 65535
    for a in vals[1:8]:
        if a:
            assert O.gf16_exp(a, 65535) == 1


# ---------------------------------------------------------------- Leopard
def test_leopard_gf8_table_shapes():
    # src/core/leopard_gf8/tests.rs:5-14
    skew, walsh, log, exp = O.leo8_tables()
    assert log[0] == 255 and exp[0] == 1
    # log/exp are inverse bijections on nonzero elements
    assert sorted(log[1:].tolist()) == list(range(255))
    for a in range(1, 256):
        assert exp[log[a]] == a


def test_leopard_gf16_fft_skew_known_answer():
    want = golden("leopard_gf16")["fft_skew"]["first8"]
    skew, _, log, exp = O.leo16_tables()
    assert skew[:8].tolist() == want


def test_leopard_gf16_parity_known_answer():
    # Go-derived parity bytes
    want = golden("leopard_gf16")["parity_3x2"]["parity_first16"]
    data = [np.array([(i * 64 + j) & 0xFF for j in range(64)], np.uint8) for i in range(3)]
    rc, parity = O.leo_encode(16, 3, 2, data)
    assert rc == 0
    assert [p[:16].tolist() for p in parity] == want


def test_leopard_gf16_coefficients_known_answer():
    expected = golden("leopard_gf16")["coefficients_3x2"]["rows"]
    for d in range(3):
        data = [np.zeros(64, np.uint8) for _ in range(3)]
        data[d][0] = 1
        _, parity = O.leo_encode(16, 3, 2, data)
        for p in range(2):
            got = int(parity[p][0]) | (int(parity[p][32]) << 8)  # element 0: bytes 0 and 32
            assert got == expected[p][d]


def _leo_mul8(a, log_b, log, exp):
    if a == 0:
        return 0
    s = int(log[a]) + int(log_b)
    return int(exp[(s + (s >> 8)) & 0xFF])


# Published Cantor bases of the Leopard fields (src/core/leopard_gf8/tables.rs:25,
# src/core/leopard_gf16/tables.rs:24-27): field polynomials 0x11D and 0x1002D.
LEO_CANTOR = {
    8: (0x11D, [1, 214, 152, 146, 86, 200, 88, 230]),
    16: (0x1002D, [0x0001, 0xACCA, 0x3C0E, 0x163E, 0xC582, 0xED2E, 0x914C, 0x4012,
                   0x6C98, 0x10D8, 0x6A72, 0xB900, 0xFDB8, 0xFB34, 0xFF38, 0x991E]),
}


def _poly_mul(a, b, bits, poly):
    """Carry-less product reduced by the field polynomial (textbook GF(2^n))."""
    r = 0
    while b:
        if b & 1:
            r ^= a
        b >>= 1
        a <<= 1
        if a >> bits:
            a ^= poly
    return r


def _leo_field(bits):
    """Multiply and divide of a Leopard field through its log / exp tables."""
    _, _, log, exp = O.leo8_tables() if bits == 8 else O.leo16_tables()
    mod = (1 << bits) - 1
    log = [int(v) for v in log]
    exp = [int(v) for v in exp[:mod + 1]]

    def mul(a, b):
        return 0 if a == 0 or b == 0 else exp[(log[a] + log[b]) % mod]

    def div(a, b):
        assert b
        return 0 if a == 0 else exp[(log[a] - log[b]) % mod]

    return mul, div


@pytest.mark.parametrize("bits", [8, 16])
def test_leopard_cantor_basis_recurrence(bits):
    """A Cantor basis has b_0 = 1 and b_i^2 + b_i = b_{i-1} in the polynomial field."""
    poly, basis = LEO_CANTOR[bits]
    assert basis[0] == 1
    for i in range(1, bits):
        assert _poly_mul(basis[i], basis[i], bits, poly) ^ basis[i] == basis[i - 1], i


@pytest.mark.parametrize("bits", [8, 16])
def test_leopard_tables_are_the_polynomial_field_in_cantor_coordinates(bits):
    """Leopard stores element x as its coordinates in the Cantor basis: the map
    phi(x) = XOR of b_i over the set bits i of x is a field isomorphism from the
    table arithmetic onto GF(2)[X]/poly (init_luts, leopard_gf8/tables.rs:25-60)."""
    poly, basis = LEO_CANTOR[bits]
    size = 1 << bits
    phi = [0] * size
    for x in range(1, size):
        low = x & -x
        phi[x] = phi[x ^ low] ^ basis[low.bit_length() - 1]
    assert sorted(phi) == list(range(size))
    inv = [0] * size
    for x, y in enumerate(phi):
        inv[y] = x
    mul, _ = _leo_field(bits)
    rng = np.random.default_rng(bits)
    if bits == 8:
        pairs = itertools.product(range(256), repeat=2)
    else:
        pairs = [(int(a), int(b)) for a, b in rng.integers(0, size, (20000, 2))]
        pairs += [(0, 5), (5, 0), (1, 0xFFFF), (0xFFFF, 0xFFFF)]
    for a, b in pairs:
        assert mul(a, b) == inv[_poly_mul(phi[a], phi[b], bits, poly)], (a, b)


def _leo_elements(bits, shard):
    if bits == 8:
        return [int(v) for v in shard]
    out = []
    for base in range(0, shard.size, 64):
        out += [int(shard[base + i]) | int(shard[base + 32 + i]) << 8 for i in range(32)]
    return out


@pytest.mark.parametrize("bits,k,m", [(16, 3, 2), (16, 10, 4), (16, 40, 16), (16, 33, 5),
                                      (8, 3, 2), (8, 10, 4), (8, 17, 8), (8, 100, 28),
                                      (8, 5, 3), (8, 128, 64)])
def test_leopard_parity_is_coset_interpolation(bits, k, m):
    """Closed form of the Leopard encoder (leopard_gf8/encode.rs, leopard_gf16/encode.rs).
    With M = next power of two >= m and element x standing for the point w_x, data
    shard c*M + t sits at the point M*(c+1) + t.  The IFFT over the coset
    M*(c+1) + {0..M-1} and the FFT over {0..M-1} evaluate the degree < M interpolant
    of chunk c (absent data shards count as zeros) at the parity points, so
    parity i = sum_{c,t} d_{cM+t} prod_{s != t} (i + M(c+1) + s) / (t + s).
    The GF(2^16) cases, whose parity bytes are pinned by the Go known answers,
    confirm the form; the GF(2^8) cases then pin the GF(2^8) encoder."""
    mul, div = _leo_field(bits)
    M = 1 << (m - 1).bit_length()
    rng = np.random.default_rng(bits * 1000 + k * 10 + m)
    data = [rng.integers(0, 256, 64, dtype=np.uint8) for _ in range(k)]
    rc, parity = O.leo_encode(bits, k, m, data)
    assert rc == 0
    elems = [_leo_elements(bits, d) for d in data]
    got = [_leo_elements(bits, p) for p in parity]
    for i in range(m):
        want = [0] * len(elems[0])
        for d in range(k):
            c, t = divmod(d, M)
            base = M * (c + 1)
            coef = 1
            for s in range(M):
                if s != t:
                    coef = mul(coef, div(i ^ (base + s), t ^ s))
            for e, v in enumerate(elems[d]):
                want[e] ^= mul(coef, v)
        assert got[i] == want, i


@pytest.mark.parametrize("bits", [8, 16])
def test_leopard_encode_is_linear_systematic_code(bits):
    """Leopard parity is a fixed linear map of the data: encode(a^b) = encode(a)^encode(b),
    and a single nonzero symbol at data shard d produces parity = coeff * symbol."""
    k, m, n = 7, 5, 128
    rng = np.random.default_rng(bits)
    a = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k)]
    b = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k)]
    _, pa = O.leo_encode(bits, k, m, a)
    _, pb = O.leo_encode(bits, k, m, b)
    _, pab = O.leo_encode(bits, k, m, [x ^ y for x, y in zip(a, b)])
    assert all(((x ^ y) == z).all() for x, y, z in zip(pa, pb, pab))


@pytest.mark.parametrize("bits,k,m", [(8, 10, 4), (8, 12, 7), (8, 3, 1), (8, 1, 2), (8, 33, 17),
                                      (16, 10, 4), (16, 12, 7), (16, 3, 1), (16, 33, 17)])
def test_leopard_roundtrip_random_erasures(bits, k, m):
    n = 192
    rng = np.random.default_rng(k * 100 + m + bits)
    data = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k)]
    rc, parity = O.leo_encode(bits, k, m, data)
    assert rc == 0
    full = data + parity
    for trial in range(12):
        e = int(rng.integers(1, m + 1))
        missing = rng.choice(k + m, e, replace=False)
        present = np.ones(k + m, np.uint8)
        present[missing] = 0
        shards = [s.copy() if present[i] else np.zeros(n, np.uint8) for i, s in enumerate(full)]
        rc, out = O.leo_reconstruct(bits, k, m, shards, present)
        assert rc == 0
        for i in range(k + m):
            assert (out[i] == full[i]).all(), (trial, i, missing)


@pytest.mark.parametrize("bits", [8, 16])
def test_leopard_error_locator_equals_log_convolution(bits):
    """Closed form of the FWHT locator: err_loc[i] == sum_{j erased} log[i ^ j] mod (2^n - 1)
    (the identity the reference's sparse GF16 locator uses, leopard_gf16/decode.rs:412-436)."""
    k, m = (20, 6) if bits == 8 else (40, 9)
    mm = 1 << (m - 1).bit_length()
    mod = (1 << bits) - 1
    log = O.leo8_tables()[2] if bits == 8 else O.leo16_tables()[2]
    rng = np.random.default_rng(3)
    for _ in range(4):
        present = np.ones(k + m, np.uint8)
        present[rng.choice(k + m, int(rng.integers(1, m + 1)), replace=False)] = 0
        pos = [i - k for i in range(k, k + m) if not present[i]] + list(range(m, mm)) + \
              [mm + i for i in range(k) if not present[i]]
        el = O.leo_error_locs(bits, k, m, present)
        for i in range(mm + k):
            expect = sum(int(log[i ^ j]) for j in pos) % mod
            assert int(el[i]) % mod == expect


@pytest.mark.parametrize("bits,k,m,erased", [(8, 6, 4, 1), (8, 6, 4, 3), (8, 10, 4, 2), (8, 5, 3, 2),
                                             (8, 20, 11, 5), (8, 12, 7, 7), (16, 6, 4, 1), (16, 10, 4, 3),
                                             (16, 9, 7, 4), (16, 3, 2, 1)])
def test_leopard_decoder_on_non_codewords_is_forney_form(bits, k, m, erased):
    """Closed form of the Leopard decoder for ANY input, codeword or not
    (leopard_gf8/decode.rs:290-551, leopard_gf16/decode.rs:507-1100).  Parity j
    sits at the point w_j, data i at w_{M+i} (M = next power of two >= m); the
    erased set R holds the missing shards' points and the padding points
    w_m..w_{M-1}.  With E(x) = prod_{j in R} (x + w_j), the decoder multiplies
    present values by E, interpolates over all n points (IFFT), takes the formal
    derivative and evaluates (FFT).  Because the subspace polynomial of the n
    points has a constant derivative, the derivative of the interpolant at an
    erased point w_r is sum_i v_i E(w_i) / (w_r + w_i), so
        out_r = sum_{i present} v_i * E(w_i) / (E'(w_r) * (w_r + w_i)),
        E'(w_r) = prod_{j in R, j != r} (w_r + w_j).
    On non-codewords this fixes every byte of the decoder's output, which round
    trips (codewords only) leave open."""
    mul, div = _leo_field(bits)
    M = 1 << (m - 1).bit_length()
    rng = np.random.default_rng(bits * 100 + k * 7 + m + erased)
    n = 64
    shards = [rng.integers(0, 256, n, dtype=np.uint8) for _ in range(k + m)]  # not a codeword
    missing = sorted(rng.choice(k + m, erased, replace=False).tolist())
    present = np.ones(k + m, np.uint8)
    present[missing] = 0
    rc, out = O.leo_reconstruct(bits, k, m, [s.copy() for s in shards], present)
    assert rc == 0

    def point(idx):
        return idx - k if idx >= k else M + idx

    R = [point(i) for i in missing] + list(range(m, M))
    pres = [i for i in range(k + m) if present[i]]
    elems = {i: _leo_elements(bits, shards[i]) for i in pres}
    for r in missing:
        wr = point(r)
        dr = 1
        for j in R:
            if j != wr:
                dr = mul(dr, wr ^ j)
        want = [0] * len(elems[pres[0]])
        for i in pres:
            wi = point(i)
            ei = 1
            for j in R:
                ei = mul(ei, wi ^ j)
            coef = div(ei, mul(dr, wr ^ wi))
            for e, v in enumerate(elems[i]):
                want[e] ^= mul(coef, v)
        assert _leo_elements(bits, out[r]) == want, (r, missing)


def test_smallrng_seeded_payload_matches_golden_checksum_path():
    # seeded_data(61, 0xA5A5_1024) feeds test_golden_split_encode; check the stream is stable
    a = O.smallrng_bytes(0xA5A51024, 61)
    b = O.smallrng_bytes(0xA5A51024, 128)
    assert (a == b[:61]).all()
