"""Pins the oracle to the reference's own known-answer tests and test streams.

Each test names the reference test it reproduces. The FFT product is also
checked against an exact 128-bit schoolbook product (orc_inner_product_exact)
on the std::mt19937_64(12345) streams of crypto/polynomial_test.cc, which the
reference asserts equal to its exact NTT product.
"""
import numpy as np
import pytest

import oracle as O

M32 = (1 << 32) - 1
M64 = (1 << 64) - 1


def u32(a):
    return [int(x) & M32 for x in a]


# ---- crypto/matrix_test.cc ----

def test_multiply_valid_kat():
    # matrix_test.cc:82-88: {{1,2},{3,4}} * {5,6} = {17, 39} (generic Multiply).
    out = O.multiply_i16(np.array([[1, 2], [3, 4]], np.int16), [5, 6])
    assert out.tolist() == [17, 39]


def test_condensed_u16_kat():
    # matrix_test.cc:90-111.
    m = np.tile(np.array([1, 2, 3, 4], np.uint16), (2, 8))
    out = O.condensed_multiply_u16(O.condense_u16(m), 32, np.ones(32, np.uint64))
    assert out.tolist() == [80, 80]


def test_condensed_i32_kat():
    # matrix_test.cc:113-134.
    m = np.tile(np.array([-1, 2], np.int32), (2, 16))
    out = O.condensed_multiply_i32(O.condense_i32(m), 32, np.ones(32, np.uint64))
    assert out.tolist() == [16, 16]


@pytest.mark.parametrize("cols,expect", [(64, 288), (1024, 4608)])
def test_condensed_u8_kat(cols, expect):
    # matrix_test.cc:136-183.
    m = np.tile(np.arange(1, 9, dtype=np.uint8), (2, cols // 8))
    out = O.condensed_multiply_u8(O.condense_u8(m), cols, np.ones(cols, np.uint64))
    assert out.tolist() == [expect, expect]


def test_condensed_layout():
    # matrix.cc:107-139: lanes least-significant first; int32 low half first.
    c = O.condense_u16(np.array([[1, 2, 3, 4]], np.uint16))
    assert int(c[0, 0]) == 1 | (2 << 16) | (3 << 32) | (4 << 48)
    c = O.condense_i32(np.array([[-1, 5]], np.int32))
    assert int(c[0, 0]) == 0xFFFFFFFF | (5 << 32)


def test_condensed_lane_check():
    # matrix.cc:269-272: columns must be a multiple of 4 * Lanes.
    m = np.ones((1, 16), np.uint16)
    with pytest.raises(O.OracleError, match="multiple of 4 \\* Lanes"):
        O.condensed_multiply_u16(O.condense_u16(m), 16, np.ones(16, np.uint64), lanes=8)


@pytest.mark.parametrize("kind", ["u16", "u8", "i32", "i16"])
def test_matrix_products_closed_form(kind):
    rng = np.random.default_rng(1)
    rows, cols = 5, 128
    vec = rng.integers(0, 1 << 64, size=cols, dtype=np.uint64)
    if kind == "u16":
        m = rng.integers(0, 1 << 16, size=(rows, cols), dtype=np.uint16)
        got = O.condensed_multiply_u16(O.condense_u16(m), cols, vec)
    elif kind == "u8":
        m = rng.integers(0, 1 << 8, size=(rows, cols), dtype=np.uint8)
        got = O.condensed_multiply_u8(O.condense_u8(m), cols, vec)
    elif kind == "i32":
        m = rng.integers(-(1 << 31), 1 << 31, size=(rows, cols)).astype(np.int32)
        got = O.condensed_multiply_i32(O.condense_i32(m), cols, vec)
    else:
        m = rng.integers(-(1 << 15), 1 << 15, size=(rows, cols)).astype(np.int16)
        got = O.multiply_i16(m, vec)
    want = [sum(int(a) * int(b) for a, b in zip(row, vec)) & M64 for row in m.tolist()]
    assert [int(x) for x in got] == want


# ---- crypto/polynomial_test.cc coefficient KATs ----

def test_right_shift_kat():
    # polynomial_test.cc:140-151.
    p = [0x12345678, 0xFFFFFFFF, 0, 1]
    assert O.right_shift(p, 4).tolist() == [0x01234567, 0x0FFFFFFF, 0, 0]
    assert O.right_shift(p, 64).tolist() == [0, 0, 0, 0]


def test_low_bits_kat():
    # polynomial_test.cc:161-173.
    p = [0x12345678, 0xFFFFFFFF, 0, 1]
    assert O.low_bits(p, 4).tolist() == [0x8, 0xF, 0, 1]
    assert O.low_bits(p, 64).tolist() == p


def test_gadget_inv_kat():
    # polynomial_test.cc:182-198 (Polynomial<uint32_t>, log_modulus 0 -> 32).
    p = [0x12345678, 0xFFFFFFFF, 0, 1]
    g = O.gadget_inv(p, 32, 8, 4)
    assert g.tolist() == [[0x78, 0xFF, 0, 1], [0x56, 0xFF, 0, 0], [0x34, 0xFF, 0, 0],
                          [0x12, 0xFF, 0, 0]]
    g = O.gadget_inv(p, 32, 16, 3)
    assert g.tolist() == [[0x5678, 0xFFFF, 0, 1], [0x1234, 0xFFFF, 0, 0], [0, 0, 0, 0]]


def test_approximate_gadget_inv_kat():
    # polynomial_test.cc:200-210.
    p = [0x12345678, 0xFFFFFFFF, 0, 1, 0x80000000, 0x7FFFFFFF, 0, 0]
    g = O.gadget_inv(p, 32, 8, 2)
    assert [u32(x) for x in g] == [[0x34, 0, 0, 0, 0, 0, 0, 0],
                                   [0x12, 0, 0, 0, 0x80, 0x80, 0, 0]]


def test_signed_gadget_inv_kat():
    # polynomial_test.cc:212-232, including the reconstruction property.
    p = [0x12345698, 0xFFFFFFFF, 0, 1]
    g = O.signed_gadget_inv(p, 32, 8, 4)
    assert [u32(x) for x in g] == [[0xFFFFFF98, 0xFFFFFFFF, 0, 1], [0x57, 0, 0, 0],
                                   [0x34, 0, 0, 0], [0x12, 0, 0, 0]]
    rec = [sum(int(g[i][j]) << (8 * i) for i in range(4)) & M32 for j in range(4)]
    assert rec == p


def test_approximate_signed_gadget_inv_kat():
    # polynomial_test.cc:234-242.
    g = O.signed_gadget_inv([0x12345678, 0x129B9856, 0, 1], 32, 8, 2)
    assert [u32(x) for x in g] == [[0x34, 0xFFFFFF9C, 0, 0], [0x12, 0x13, 0, 0]]


def test_gadget_inv_param_errors():
    # polynomial_test.cc:244-276.
    for ld, nd in [(0, 4), (-1, 4), (7, 0), (7, -1)]:
        with pytest.raises(O.OracleError, match="log_digit and num_digits must be positive"):
            O.gadget_inv([1, 2, 3, 4], 32, ld, nd)
        with pytest.raises(O.OracleError, match="log_digit and num_digits must be positive"):
            O.signed_gadget_inv([1, 2, 3, 4], 32, ld, nd)


def test_automorph_kat():
    # polynomial_test.cc:103-114.
    assert u32(O.automorph([1, 2, 3, 4], 3)) == [1, 4, M32 - 2, 2]
    for bad in (-1, 2, 8, 9):
        with pytest.raises(O.OracleError, match="non-negative odd integer"):
            O.automorph([1, 2, 3, 4], bad)


def test_rescale_closed_form():
    # polynomial.cc:297-302: (c * 2^new + 2^(cur-1)) / 2^cur, truncated.
    rng = np.random.default_rng(2)
    c = rng.integers(0, 1 << 64, size=64, dtype=np.uint64)
    for cur, new in [(52, 29), (52, 18), (64, 61), (62, 20), (56, 56)]:
        want = [((int(x) << new) + (1 << (cur - 1))) // (1 << cur) & M64 for x in c]
        assert [int(x) for x in O.rescale(c, cur, new)] == want


# ---- FFT products (crypto/polynomial_test.cc:297-612) ----

def test_mult_kat_degree4_u32():
    # polynomial_test.cc:297-309 and :399-408: {1,2,3,4}*{5,6,7,8} = {-56,-36,2,60}.
    out, _ = O.inner_product_fft_w([1, 2, 3, 4], [5, 6, 7, 8], 32, 32, 20, width=32)
    assert out.tolist() == [(1 << 32) - 56, (1 << 32) - 36, 2, 60]


def test_inner_product_kat_u32():
    # polynomial_test.cc:530-549.
    u = [[1, 2, 3, 4], [2, 3, 4, 5]]
    v = [[5, 6, 7, 8], [1, 1, 1, 1]]
    out, _ = O.inner_product_fft_w(u, v, 32, 32, 20, width=32)
    assert out.tolist() == O.inner_product_exact_w(u, v, 32, 32, 20, width=32).tolist()


def test_mult_fft_degree16_u64():
    # polynomial_test.cc:410-424.
    c1 = [(0x123456789ABCDEF * (i + 1)) & M64 for i in range(16)]
    c2 = [(0xFEDCBA987654321 * (16 - i)) & M64 for i in range(16)]
    out, _ = O.inner_product_fft(c1, c2, 64, 64, 20)
    exact = [0] * 16
    for i in range(16):
        for j in range(16):
            k = i + j
            if k < 16:
                exact[k] += c1[i] * c2[j]
            else:
                exact[k - 16] -= c1[i] * c2[j]
    assert out.tolist() == [x & M64 for x in exact]


def test_mult_fft_degree1024_u64():
    # polynomial_test.cc:426-440.
    c1 = np.array([i * 1000000000 for i in range(1024)], np.uint64)
    c2 = np.array([(1024 - i) * 1000000000 for i in range(1024)], np.uint64)
    out, _ = O.inner_product_fft(c1, c2, 64, 64, 20)
    assert np.array_equal(out, O.inner_product_exact(c1, c2, 64, 64, 20))


def test_mult_fft_random_stream():
    # polynomial_test.cc:442-489 with std::mt19937_64(12345). The test draws
    # through uniform_int_distribution<uint64_t>(0, UINT64_MAX), which for the
    # generator's full range returns the raw draw (libstdc++).
    draws = iter(O.mt19937_64(12345, 200000).tolist())
    for degree in (128, 256, 512, 1024, 2048):
        for width in (32, 64):
            c1, c2 = [], []
            for _ in range(degree):
                c1.append(next(draws))
                c2.append(next(draws))
            if width == 32:
                c1 = u32(c1)
                c2 = u32(c2)
            got, err = O.inner_product_fft_w(c1, c2, width, width, 20, width=width)
            want = O.inner_product_exact_w(c1, c2, width, width, 20, width=width)
            assert np.array_equal(got, want), (degree, width, err)


def test_mult_fft_explicit_bits_stream():
    # polynomial_test.cc:491-528 (draw order reproduced exactly).
    draws = iter(O.mt19937_64(12345, 100000).tolist())
    for degree in (128, 256, 512, 1024, 2048):
        this_bits = next(draws) % 31 + 1
        that_bits = next(draws) % 31 + 1
        m1, m2 = (1 << this_bits) - 1, (1 << that_bits) - 1
        c1, c2 = [], []
        for _ in range(degree):
            c1.append(next(draws) & M32 & m1)
            c2.append(next(draws) & M32 & m2)
        want = O.inner_product_exact_w(c1, c2, this_bits, that_bits, 12, width=32)
        for chunk in (12, 16, 20):
            got, _ = O.inner_product_fft_w(c1, c2, this_bits, that_bits, chunk, width=32)
            assert np.array_equal(got, want), (degree, chunk)


def test_inner_product_fft_random_stream():
    # polynomial_test.cc:551-612 (draw order reproduced exactly).
    draws = iter(O.mt19937_64(12345, 400000).tolist())
    for degree in (128, 512, 2048):
        for k in (1, 5, 10):
            this_bits = next(draws) % 31 + 1
            that_bits = next(draws) % 31 + 1
            m1, m2 = (1 << this_bits) - 1, (1 << that_bits) - 1
            u = np.zeros((k, degree), np.uint64)
            v = np.zeros((k, degree), np.uint64)
            for i in range(k):
                for j in range(degree):
                    u[i, j] = next(draws) & M32 & m1
                    v[i, j] = next(draws) & M32 & m2
            want = O.inner_product_exact_w(u, v, this_bits, that_bits, 12, width=32)
            chunks = [12, 16, 20]
            boundary = int((53.0 - np.log2(k) - np.log2(degree)) / 2.0)
            if boundary > 0:
                chunks.append(boundary)
            for chunk in chunks:
                if np.log2(k) + np.log2(degree) + 2.0 * chunk <= 53.0:
                    got, _ = O.inner_product_fft_w(u, v, this_bits, that_bits, chunk, width=32)
                    assert np.array_equal(got, want), (degree, k, chunk)


def test_inner_product_precision_guard():
    # polynomial_fft.cc:90-96 error messages.
    u = np.ones((2, 2048), np.uint64)
    with pytest.raises(O.OracleError, match="Potential precision overflow"):
        O.inner_product_fft(u, u, 64, 64, 21)
    with pytest.raises(O.OracleError, match="Invalid bits or chunk_bits"):
        O.inner_product_fft(u, u, 0, 64, 20)


def test_external_product_rounding_at_d4096_is_not_exact():
    """Characterises the reference arithmetic: with ExternalProduct's chunk of
    19 bits at d = 4096 (external_product.cc:100-102) the double-precision
    products reach 2^52.6 and the rounded FFT result differs from the exact
    product. The GPU path therefore reproduces the FFT arithmetic, not the
    exact product."""
    rng = np.random.default_rng(5)
    d = 4096
    u = rng.integers(0, 1 << 19, size=(6, d), dtype=np.uint64)
    v = rng.integers(0, 1 << 52, size=(6, d), dtype=np.uint64)
    got, err = O.inner_product_fft(u, v, 19, 52, 19)
    exact = O.inner_product_exact(u, v, 19, 52, 19)
    assert err >= 0.25
    assert not np.array_equal(got, exact)
