"""What the mathematics fixes of InnerProductFft's roundings
(crypto/polynomial_fft.cc:68-197), at the shapes of the answer path.

InnerProductFft rounds, for every accumulated chunk pair (a, b), the value
pre[j] = IFFT(sum_i U_i,a * V_i,b)[j] / 2d, whose exact value is the integer
E[j] = sum_i (u_i,a * v_i,b)[j] of the linear convolution of the chunk
polynomials. The tests compute E independently of the oracle's C++ (numpy
integer convolutions, exact below 2^63) and check:

  (i)   |round(pre) - E| <= 1 at every position of every pair, in the
        reference's uncontracted arithmetic and in a contracted build;
  (ii)  rounding is fixed by worst-case error analysis only where Percival's
        bound (C. Percival, "Rapid multiplication modulo the sum and
        difference of highly composite numbers", Math. Comp. 72 (2003)
        387-395, Theorem 5.1; restated in Brent and Zimmermann, "Modern
        Computer Arithmetic", Cambridge 2010, section 3.3) is below 1/2: there
        round(pre) == E everywhere. At the answer path's shapes the bound is
        far above 1/2 (tens of units at d = 4096), so no rounding there is
        fixed by it;
  (iii) the oracle's result is exactly the shifted, folded sum of its rounded
        pairs, and the independent 128-bit closed form is the same sum of the
        exact pairs, so the result differs from the exact product only at
        coefficients that a pair rounded away from its exact value reaches.

What fixes the rest is the arithmetic itself: the oracle's pre-rounding
doubles equal, bit for bit, an independent numpy restatement of crypto/fft.h
(explicit real arithmetic, the C library's twiddles as a g++ build computes
them, no fused multiply-adds). The rounding at d = 4096 depends on how the reference is
compiled: a build whose compiler fuses the complex products into FMAs rounds
the ExternalProduct (crypto/external_product.cc:100-102) differently, and
then every coefficient of an n = 2^15, d = 4096, t = 8 response differs;
at d = 2048 the builds agree.
"""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O

EPS = 2.0 ** -53
# Twiddle error bound: the angle -2*pi*i/N in double precision is within
# 2|fl(pi) - pi| + 2 ulp(pi) < 1.1e-15 of the exact angle, and the C
# library's cos/sin are within one ulp (< 1.2e-16) of the rounded angle's.
BETA = 2.0 ** -49

# (name, terms, d, u_bits, v_bits, chunk, u_signed): ExternalProduct's
# product (external_product.cc:100-102) and FinalizeMatrixPack's
# (lwes_to_rlwe.cc:60-74) at the path's ring degrees.
SHAPES = [
    ("external_product_d4096", 6, 4096, 19, 52, 19, False),
    ("finalize_matrix_pack_d4096", 2, 4096, 19, 52, 20, True),
    ("external_product_d2048", 6, 2048, 19, 52, 19, False),
    ("finalize_matrix_pack_d2048", 2, 2048, 19, 52, 20, True),
]


def _operands(shape, seed):
    _, n, d, ub, vb, _, signed = shape
    rng = np.random.default_rng(seed)
    u = rng.integers(0, 1 << 64 if signed else 1 << ub, size=(n, d), dtype=np.uint64)
    v = rng.integers(0, 1 << vb, size=(n, d), dtype=np.uint64)
    return u, v


def _num_chunks(bits, chunk):
    return min((bits + chunk - 1) // chunk, (64 + chunk - 1) // chunk)


def _chunk(x, chunk, c, bits, signed, nchunks):
    """Chunk c of every coefficient as polynomial_fft.cc:121-135 reads it."""
    ch = ((x >> np.uint64(chunk * c)) & np.uint64((1 << chunk) - 1)).astype(np.int64)
    if signed and c == nchunks - 1:
        b = bits - chunk * c
        ch &= (1 << b) - 1
        ch = np.where(ch >= 1 << (b - 1), ch - (1 << b), ch)
    return ch


def _exact_pairs(u, v, ub, vb, chunk, signed, pairs):
    n, d = u.shape
    nu, nv = _num_chunks(ub, chunk), _num_chunks(vb, chunk)
    E = np.zeros((len(pairs), 2 * d), dtype=np.int64)
    for p, (a, b) in enumerate(pairs):
        for i in range(n):
            E[p, :2 * d - 1] += np.convolve(_chunk(u[i], chunk, a, ub, signed, nu),
                                            _chunk(v[i], chunk, b, vb, False, nv))
    return E


def _round_half_away(x):
    return np.where(x >= 0, np.floor(x + 0.5), -np.floor(-x + 0.5)).astype(np.int64)


def _fold(vals, d, chunk, pairs):
    """sum_p (vals_p << chunk*(a+b)) with X^d = -1, mod 2^64
    (polynomial_fft.cc:186-193)."""
    out = np.zeros(d, dtype=np.uint64)
    with np.errstate(over="ignore"):
        for p, (a, b) in enumerate(pairs):
            w = vals[p].astype(np.uint64) << np.uint64(chunk * (a + b))
            out += w[:d]
            out -= w[d:]
    return out


def _percival_bound(x, y, n_log):
    return (np.linalg.norm(x) * np.linalg.norm(y) *
            ((1 + EPS) ** (3 * n_log) * (1 + EPS * math.sqrt(5)) ** (3 * n_log + 1) *
             (1 + BETA) ** (3 * n_log) - 1))


@pytest.mark.parametrize("shape", SHAPES, ids=[s[0] for s in SHAPES])
@pytest.mark.parametrize("fused", [False, True], ids=["uncontracted", "contracted"])
def test_rounded_pairs_within_one_of_exact(shape, fused):
    name, n, d, ub, vb, chunk, signed = shape
    u, v = _operands(shape, 17)
    pairs, pre = O.inner_product_fft_pairs(u, v, ub, vb, chunk, signed, fused=fused)
    assert pairs == [(a, b) for a in range(_num_chunks(ub, chunk))
                     for b in range(_num_chunks(vb, chunk)) if chunk * (a + b) < 64]
    E = _exact_pairs(u, v, ub, vb, chunk, signed, pairs)
    R = _round_half_away(pre)
    assert np.abs(R - E).max() <= 1                                 # (i)
    assert not R[:, 2 * d - 1].any()                                 # the padding's top is 0
    if name == "external_product_d4096":
        # The exact values reach 2^50.6, where a double resolves quarters:
        # the FFT rounds a few percent of them to a neighbour.
        wrong = int((R != E).sum())
        assert 0 < wrong < 0.15 * R.size
    else:
        assert np.array_equal(R, E)


@pytest.mark.parametrize("shape", SHAPES, ids=[s[0] for s in SHAPES])
def test_result_is_folded_sum_of_rounded_pairs(shape):
    name, n, d, ub, vb, chunk, signed = shape
    u, v = _operands(shape, 23)
    pairs, pre = O.inner_product_fft_pairs(u, v, ub, vb, chunk, signed)
    E = _exact_pairs(u, v, ub, vb, chunk, signed, pairs)
    R = _round_half_away(pre)
    got, _ = O.inner_product_fft(u, v, ub, vb, chunk, signed)
    exact = O.inner_product_exact(u, v, ub, vb, chunk, signed)
    assert np.array_equal(got, _fold(R, d, chunk, pairs))          # (iii)
    assert np.array_equal(exact, _fold(E, d, chunk, pairs))
    off = np.zeros(d, dtype=bool)
    for p in range(len(pairs)):
        wrong = np.nonzero(R[p] != E[p])[0]
        off[wrong % d] = True
    assert not np.any((got != exact) & ~off)


def test_worst_case_bound_fixes_small_chunk_products_only():
    # MultFft with 12-bit chunks at degree 2048 (polynomial_test.cc:491-528
    # runs chunk 12 over 1..31-bit operands): the bound is far below 1/2 for
    # every chunk pair, so each rounding is fixed, and the oracle's is exact.
    rng = np.random.default_rng(29)
    d, chunk, bits = 2048, 12, 31
    u = rng.integers(0, 1 << bits, size=(1, d), dtype=np.uint64)
    v = rng.integers(0, 1 << bits, size=(1, d), dtype=np.uint64)
    pairs, pre = O.inner_product_fft_pairs(u, v, bits, bits, chunk)
    nu = _num_chunks(bits, chunk)
    for a, b in pairs:
        x = _chunk(u[0], chunk, a, bits, False, nu).astype(np.float64)
        y = _chunk(v[0], chunk, b, bits, False, nu).astype(np.float64)
        assert _percival_bound(x, y, 12) < 0.01                      # (ii)
    E = _exact_pairs(u, v, bits, bits, chunk, False, pairs)
    assert np.array_equal(_round_half_away(pre), E)
    # At the answer path's shapes one product term alone has a bound of
    # several units: worst-case analysis fixes none of those roundings.
    for shape in SHAPES:
        name, n, d, ub, vb, chunk, signed = shape
        u, v = _operands(shape, 31)
        x = _chunk(u[0], chunk, 0, ub, signed, _num_chunks(ub, chunk)).astype(np.float64)
        y = _chunk(v[0], chunk, 0, vb, False, _num_chunks(vb, chunk)).astype(np.float64)
        assert _percival_bound(x, y, int(math.log2(2 * d))) > 0.5, name


# ---- crypto/fft.h restated with numpy, one IEEE operation at a time ----

def _sincos():
    import ctypes
    import ctypes.util
    libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
    f = libm.sincos
    f.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    f.restype = None

    def sincos(x):
        s, c = ctypes.c_double(), ctypes.c_double()
        f(x, ctypes.byref(s), ctypes.byref(c))
        return s.value, c.value
    return sincos


def _twiddles(n, sign):
    """fft.h:71-80 as g++ compiles it (GCC's sincos pass and the fold of
    cos(-x) to cos(x)): the forward pair is one sincos(angle) call of the C
    library, the backward real part reuses that cosine, and the backward
    imaginary part is a separate sin(-angle) call. sincos and sin differ in
    the last bit of a few results (2 of 4096 at n = 8192), so this is the
    table of a g++ build; a compiler that calls sin and cos separately gets
    a slightly different forward table."""
    sincos = _sincos()
    pi = math.acos(-1.0)
    re = np.empty(n // 2)
    im = np.empty(n // 2)
    for i in range(n // 2):
        angle = -2.0 * pi * i / n          # fft.h:75, same operation order
        s, c = sincos(angle)
        re[i] = c
        im[i] = s if sign < 0 else math.sin(-angle)
    return re, im


def _fft(re, im, tw):
    """ComputeFftInternal (fft.h:106-129): bit-reversed copy, then radix-2
    stages; v = out[hi] * w as (xr*wr - xi*wi, xr*wi + xi*wr); each numpy
    operation rounds once, as an uncontracted build does."""
    n = re.size
    log_n = n.bit_length() - 1
    idx = np.arange(n)
    rev = np.zeros(n, dtype=np.int64)
    for b in range(log_n):
        rev |= ((idx >> b) & 1) << (log_n - 1 - b)
    xr = np.empty(n)
    xi = np.empty(n)
    xr[rev] = re
    xi[rev] = im
    twr, twi = tw
    length = 2
    while length <= n:
        half = length // 2
        step = n // length
        xr = xr.reshape(-1, length)
        xi = xi.reshape(-1, length)
        wr = twr[np.arange(half) * step]
        wi = twi[np.arange(half) * step]
        ur, ui = xr[:, :half].copy(), xi[:, :half].copy()
        hr, hi = xr[:, half:], xi[:, half:]
        vr = hr * wr - hi * wi
        vi = hr * wi + hi * wr
        xr = np.concatenate([ur + vr, ur - vr], axis=1).reshape(-1)
        xi = np.concatenate([ui + vi, ui - vi], axis=1).reshape(-1)
        length *= 2
    return xr, xi


def _restated_pairs(u, v, ub, vb, chunk, signed):
    """polynomial_fft.cc:109-184 around _fft."""
    n, d = u.shape
    N = 2 * d
    fwd, bwd = _twiddles(N, -1), _twiddles(N, +1)
    nu, nv = _num_chunks(ub, chunk), _num_chunks(vb, chunk)
    pairs = [(a, b) for a in range(nu) for b in range(nv) if chunk * (a + b) < 64]
    acc = {p: (np.zeros(N), np.zeros(N)) for p in pairs}
    zeros = np.zeros(N)
    for i in range(n):
        us, vs = [], []
        for c in range(nu):
            x = np.zeros(N)
            x[:d] = _chunk(u[i], chunk, c, ub, signed, nu).astype(np.float64)
            us.append(_fft(x, zeros, fwd))
        for c in range(nv):
            x = np.zeros(N)
            x[:d] = _chunk(v[i], chunk, c, vb, False, nv).astype(np.float64)
            vs.append(_fft(x, zeros, fwd))
        for (a, b) in pairs:
            (ar, ai), (br, bi) = us[a], vs[b]
            sr, si = acc[(a, b)]
            acc[(a, b)] = (sr + (ar * br - ai * bi), si + (ar * bi + ai * br))
    return pairs, np.stack([_fft(*acc[p], bwd)[0] / N for p in pairs])


@pytest.mark.parametrize("shape", SHAPES[:2], ids=[s[0] for s in SHAPES[:2]])
def test_oracle_doubles_equal_a_restatement_of_fft_h(shape):
    name, n, d, ub, vb, chunk, signed = shape
    u, v = _operands(shape, 37)
    pairs, pre = O.inner_product_fft_pairs(u, v, ub, vb, chunk, signed)
    rpairs, rpre = _restated_pairs(u, v, ub, vb, chunk, signed)
    assert pairs == rpairs
    assert np.array_equal(pre.view(np.uint64), rpre.view(np.uint64))


def _cpu_has(*flags):
    try:
        with open("/proc/cpuinfo") as f:
            words = set(f.read().split())
    except OSError:
        return False
    return all(fl in words for fl in flags)


@pytest.mark.skipif(shutil.which("g++") is None or not _cpu_has("avx512f", "fma"),
                    reason="needs g++ and an AVX-512 host to run a contracted build")
def test_gcc_avx512_build_is_the_contracted_arithmetic(tmp_path):
    """The oracle's source compiled as the reference's README builds it for
    AVX-512 (-O2 -march=skylake-avx512, GCC's default contraction) rounds
    exactly as the contracted arithmetic (fma(xr, wr, -(xi*wi)),
    fma(xr, wi, xi*wr)) does, and differently from the uncontracted one."""
    import ctypes
    src = os.path.join(os.path.dirname(O.__file__), "pir_oracle.cpp")
    so = str(tmp_path / "liboracle_avx512.so")
    subprocess.run(["g++", "-std=c++17", "-O2", "-march=skylake-avx512", "-fPIC", "-shared",
                    "-pthread", "-o", so, src], check=True, capture_output=True)
    lib = ctypes.CDLL(so)
    shape = SHAPES[0]
    name, n, d, ub, vb, chunk, signed = shape
    u, v = _operands(shape, 41)
    nu, nv = _num_chunks(ub, chunk), _num_chunks(vb, chunk)
    pre = np.empty((nu * nv, 2 * d))
    pa = np.empty(nu * nv, np.int32)
    pb = np.empty(nu * nv, np.int32)
    k = ctypes.c_int()
    err = ctypes.create_string_buffer(256)

    def ptr(a, t):
        return a.ctypes.data_as(ctypes.POINTER(t))

    rc = lib.orc_inner_product_fft_pairs(ptr(u, ctypes.c_uint64), ptr(v, ctypes.c_uint64), n, d,
                                         ub, vb, chunk, int(signed), 0, ptr(pre, ctypes.c_double),
                                         ptr(pa, ctypes.c_int), ptr(pb, ctypes.c_int),
                                         ctypes.byref(k), err, 256)
    assert rc == 0, err.value
    built = pre[:k.value]
    _, fused = O.inner_product_fft_pairs(u, v, ub, vb, chunk, signed, fused=True)
    _, plain = O.inner_product_fft_pairs(u, v, ub, vb, chunk, signed)
    assert np.array_equal(built.view(np.uint64), fused.view(np.uint64))
    assert not np.array_equal(built, plain)


def test_response_depends_on_the_build_at_d4096_only():
    """ProcessResponse (pir_server.cc:152-306) under both arithmetics: at
    d = 4096 the Horner steps carry the ExternalProduct's differing roundings
    into every coefficient; at d = 2048 the two builds answer identically."""
    try:
        for (n, log_d, t, all_differ) in [(32768, 12, 8, True), (8192, 11, 4, False)]:
            P = O.metrics_tool_params(n=n, log_d=log_d, t=t)
            mat = O.synthetic_material(P, 5)
            req = O.synthetic_request(P, 101)
            O.set_fft_contraction(False)
            plain, _, _ = O.process_response(P, mat, req)
            O.set_fft_contraction(True)
            fused, _, _ = O.process_response(P, mat, req)
            O.set_fft_contraction(False)
            if all_differ:
                assert (plain != fused).mean() > 0.99
            else:
                assert np.array_equal(plain, fused)
    finally:
        O.set_fft_contraction(False)
