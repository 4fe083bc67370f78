"""CPU oracle for the ReinsPIRe online answer path -- TEST INFRASTRUCTURE ONLY.

Python (ctypes + numpy) view of ``oracle/_build/liboracle.so``, the C++
restatement of google/InsPIRe's ``PirServer::ProcessResponse`` path
(crypto/pir/server/pir_server.cc:152-306) and of the preprocessing / client
code that makes decryptable inputs for it. Each wrapped function names the
reference file:line it restates; the C++ source (``pir_oracle.cpp``) holds the
line-level citations and the list of checks that pin each function.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package, and only as the checker or the timed CPU
baseline. The product library (``inspire_amd``) never imports it.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u16p = ctypes.POINTER(ctypes.c_uint16)
_i16p = ctypes.POINTER(ctypes.c_int16)
_dblp = ctypes.POINTER(ctypes.c_double)


class OracleError(RuntimeError):
    """A non-OK status from the oracle: ``code`` is the absl::StatusCode."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


class _CParams(ctypes.Structure):
    _fields_ = [
        ("log_degree", ctypes.c_int),
        ("log_q", ctypes.c_int),
        ("log_q1", ctypes.c_int),
        ("log_q2", ctypes.c_int),
        ("variance", ctypes.c_int),
        ("plaintext_modulus", ctypes.c_uint64),
        ("pack_log_digit", ctypes.c_int),
        ("pack_num_digits", ctypes.c_int),
        ("eval_log_digit", ctypes.c_int),
        ("eval_num_digits", ctypes.c_int),
        ("num_entries", ctypes.c_int),
        ("interp_degree", ctypes.c_int),
        ("entry_size_multiple", ctypes.c_int),
        ("db_bytes", ctypes.c_int),
        ("mat_bytes", ctypes.c_int),
        ("coeff_bits", ctypes.c_int),
    ]


class _CSizes(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "d", "t", "ell", "k", "e", "r", "S", "K", "M", "db_words_per_row",
        "pack_cols", "pack_words_per_row")]


@dataclasses.dataclass(frozen=True)
class Params:
    """PirParams<uint64_t> fields (crypto/pir/pir_params.h:36-105)."""

    log_degree: int = 11
    log_q: int = 52
    log_q1: int = 29
    log_q2: int = 18
    variance: int = 40
    plaintext_modulus: int = 65535
    pack_log_digit: int = 19
    pack_num_digits: int = 2
    eval_log_digit: int = 19
    eval_num_digits: int = 3
    num_entries: int = 4096
    interp_degree: int = 2
    entry_size_multiple: int = 1
    db_bytes: int = 2
    mat_bytes: int = 4
    coeff_bits: int = 64   # CoeffType: 64 uint64_t, 32 uint32_t

    @property
    def d(self) -> int:
        return 1 << self.log_degree

    def c(self) -> _CParams:
        return _CParams(*[getattr(self, f.name) for f in dataclasses.fields(self)])


def metrics_tool_params(n: int, log_d: int, t: int, ell: int = 1) -> Params:
    """The defaults of pir_metrics_tool (crypto/pir/benchmarking/pir_metrics_tool.cc:79-87,111-114)."""
    log_p = 16  # bit_width(65535)
    return Params(log_degree=log_d, log_q=52, log_q1=log_p + log_d + 2,
                  log_q2=log_p + 2, variance=40, plaintext_modulus=65535,
                  pack_log_digit=19, pack_num_digits=2, eval_log_digit=19,
                  eval_num_digits=3, num_entries=n, interp_degree=t,
                  entry_size_multiple=ell)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (oracle/Makefile)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def _p(a: np.ndarray, ctype=_u64p):
    return a.ctypes.data_as(ctype)


def _call(fn, *args):
    err = ctypes.create_string_buffer(512)
    rc = fn(*args, err, ctypes.c_size_t(512))
    if rc != 0:
        raise OracleError(rc, err.value.decode())


def _u64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


def _u64_2d(a) -> np.ndarray:
    # Convert before reshaping: numpy infers float64 for Python ints >= 2^63.
    a = _u64(a)
    return a.reshape(1, -1) if a.ndim == 1 else a


@dataclasses.dataclass
class Sizes:
    d: int
    t: int
    ell: int
    k: int
    e: int
    r: int
    S: int
    K: int
    M: int
    db_words_per_row: int
    pack_cols: int
    pack_words_per_row: int


def sizes(P: Params) -> Sizes:
    s = _CSizes()
    cp = P.c()
    _call(lib().orc_get_sizes, ctypes.byref(cp), ctypes.byref(s))
    return Sizes(*[getattr(s, f[0]) for f in _CSizes._fields_])


# ---- polynomial coefficient ops (crypto/polynomial.cc) ----

def low_bits(x, amount: int) -> np.ndarray:
    """Polynomial::LowBits (crypto/polynomial.cc:323-338)."""
    x = _u64(x)
    out = np.empty_like(x)
    lib().orc_low_bits(_p(x), len(x), amount, _p(out))
    return out


def right_shift(x, amount: int) -> np.ndarray:
    """Polynomial::RightShift (crypto/polynomial.cc:306-321)."""
    x = _u64(x)
    out = np.empty_like(x)
    lib().orc_right_shift(_p(x), len(x), amount, _p(out))
    return out


def rescale(x, log_cur: int, log_new: int) -> np.ndarray:
    """Polynomial::Rescale (crypto/polynomial.cc:287-304)."""
    x = _u64(x)
    out = np.empty_like(x)
    lib().orc_rescale(_p(x), len(x), log_cur, log_new, _p(out))
    return out


def gadget_inv(x, log_modulus: int, log_digit: int, num_digits: int) -> np.ndarray:
    """Polynomial::GadgetInv (crypto/polynomial.cc:340-374); returns num_digits x len."""
    x = _u64(x)
    out = np.empty((max(num_digits, 1), len(x)), dtype=np.uint64)
    _call(lib().orc_gadget_inv, _p(x), len(x), log_modulus, log_digit, num_digits, _p(out))
    return out


def signed_gadget_inv(x, log_modulus: int, log_digit: int, num_digits: int) -> np.ndarray:
    """Polynomial::SignedGadgetInv (crypto/polynomial.cc:376-427)."""
    x = _u64(x)
    out = np.empty((max(num_digits, 1), len(x)), dtype=np.uint64)
    _call(lib().orc_signed_gadget_inv, _p(x), len(x), log_modulus, log_digit,
          num_digits, _p(out))
    return out


def automorph(x, power: int) -> np.ndarray:
    """Polynomial::Automorph (crypto/polynomial.cc:429-456)."""
    x = _u64(x)
    out = np.empty_like(x)
    _call(lib().orc_automorph, _p(x), len(x), power, _p(out))
    return out


def inner_product_fft(u, v, u_bits=64, v_bits=64, chunk_bits=20, u_signed=False):
    """Polynomial::InnerProductFft (crypto/polynomial_fft.cc:68-197).

    Returns (result, max |x - round(x)| over every rounded value)."""
    u = _u64_2d(u)
    v = _u64_2d(v)
    n, d = u.shape
    out = np.empty(d, dtype=np.uint64)
    me = ctypes.c_double(0.0)
    _call(lib().orc_inner_product_fft, _p(u), _p(v), n, d, u_bits, v_bits,
          chunk_bits, int(u_signed), _p(out), ctypes.byref(me))
    return out, me.value


def inner_product_fft_pairs(u, v, u_bits, v_bits, chunk_bits, u_signed=False, fused=False):
    """The values InnerProductFft rounds (polynomial_fft.cc:176-184), per
    accumulated chunk pair: returns (pairs, pre) with pairs a list of (a, b)
    in the reference's order and pre[p, j] = backward_out[j].real() / 2d for
    j < 2d. fused evaluates every complex product with FMA contraction."""
    u = _u64_2d(u)
    v = _u64_2d(v)
    n, d = u.shape
    max_chunks = (64 + chunk_bits - 1) // chunk_bits
    nu = min((u_bits + chunk_bits - 1) // chunk_bits, max_chunks)
    nv = min((v_bits + chunk_bits - 1) // chunk_bits, max_chunks)
    pre = np.empty((nu * nv, 2 * d), dtype=np.float64)
    pa = np.empty(nu * nv, dtype=np.int32)
    pb = np.empty(nu * nv, dtype=np.int32)
    npairs = ctypes.c_int(0)
    ip = ctypes.POINTER(ctypes.c_int)
    _call(lib().orc_inner_product_fft_pairs, _p(u), _p(v), n, d, u_bits, v_bits, chunk_bits,
          int(u_signed), int(fused), pre.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
          pa.ctypes.data_as(ip), pb.ctypes.data_as(ip), ctypes.byref(npairs))
    k = npairs.value
    return [(int(pa[i]), int(pb[i])) for i in range(k)], pre[:k].copy()


def set_fft_contraction(fused: bool) -> None:
    """Process-wide: InnerProductFft and what uses it evaluate complex products
    with FMA contraction (True) or as the reference's uncontracted build (False,
    the default). For rounding analysis only."""
    lib().orc_set_fft_contraction(int(bool(fused)))


def inner_product_exact(u, v, u_bits=64, v_bits=64, chunk_bits=20, u_signed=False):
    """Exact 128-bit schoolbook counterpart of inner_product_fft (closed form)."""
    u = _u64_2d(u)
    v = _u64_2d(v)
    n, d = u.shape
    out = np.empty(d, dtype=np.uint64)
    _call(lib().orc_inner_product_exact, _p(u), _p(v), n, d, u_bits, v_bits,
          chunk_bits, int(u_signed), _p(out))
    return out


def inner_product_fft_w(u, v, u_bits, v_bits, chunk_bits, u_signed=False, width=64):
    """InnerProductFft for Polynomial<uint32_t> (width 32) or <uint64_t> (64)."""
    u = _u64_2d(u)
    v = _u64_2d(v)
    n, d = u.shape
    out = np.empty(d, dtype=np.uint64)
    me = ctypes.c_double(0.0)
    _call(lib().orc_inner_product_fft_w, _p(u), _p(v), n, d, u_bits, v_bits,
          chunk_bits, int(u_signed), width, _p(out), ctypes.byref(me))
    return out, me.value


def inner_product_exact_w(u, v, u_bits, v_bits, chunk_bits, u_signed=False, width=64):
    u = _u64_2d(u)
    v = _u64_2d(v)
    n, d = u.shape
    out = np.empty(d, dtype=np.uint64)
    _call(lib().orc_inner_product_exact_w, _p(u), _p(v), n, d, u_bits, v_bits,
          chunk_bits, int(u_signed), width, _p(out))
    return out


def mt19937_64(seed: int, n: int) -> np.ndarray:
    """Raw std::mt19937_64 draws (the generator of crypto/polynomial_test.cc:443)."""
    out = np.empty(n, dtype=np.uint64)
    lib().orc_mt19937_64(ctypes.c_uint64(seed), _p(out), ctypes.c_size_t(n))
    return out


# ---- matrices (crypto/matrix.cc) ----

def condense_u16(m) -> np.ndarray:
    """Matrix<uint16_t>::CreateCondensed (crypto/matrix.cc:107-123)."""
    m = np.ascontiguousarray(m, dtype=np.uint16)
    rows, cols = m.shape
    out = np.empty((rows, cols // 4), dtype=np.uint64)
    _call(lib().orc_condense_u16, m.ctypes.data_as(_u16p), rows, cols, _p(out))
    return out


def condense_u8(m) -> np.ndarray:
    """Matrix<uint8_t>::CreateCondensed (crypto/matrix.cc:90-106)."""
    m = np.ascontiguousarray(m, dtype=np.uint8)
    rows, cols = m.shape
    out = np.empty((rows, cols // 8), dtype=np.uint64)
    _call(lib().orc_condense_u8, m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
          rows, cols, _p(out))
    return out


def condense_i32(m) -> np.ndarray:
    """Matrix<int32_t>::CreateCondensed (crypto/matrix.cc:124-140)."""
    m = np.ascontiguousarray(m, dtype=np.int32)
    rows, cols = m.shape
    out = np.empty((rows, cols // 2), dtype=np.uint64)
    _call(lib().orc_condense_i32, m.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
          rows, cols, _p(out))
    return out


def condensed_multiply_u16(cdata, cols: int, vec, lanes: int = 8) -> np.ndarray:
    """CondensedMultiply<uint16_t> (crypto/matrix.cc:261-318)."""
    cdata = _u64(cdata)
    vec = _u64(vec)
    rows = cdata.size // (cols // 4)
    out = np.empty(rows, dtype=np.uint64)
    _call(lib().orc_condensed_multiply_u16, _p(cdata), rows, cols, _p(vec), lanes, _p(out))
    return out


def condensed_multiply_u8(cdata, cols: int, vec, lanes: int = 8) -> np.ndarray:
    """CondensedMultiply<uint8_t> (crypto/matrix.cc:376-462)."""
    cdata = _u64(cdata)
    vec = _u64(vec)
    rows = cdata.size // (cols // 8)
    out = np.empty(rows, dtype=np.uint64)
    _call(lib().orc_condensed_multiply_u8, _p(cdata), rows, cols, _p(vec), lanes, _p(out))
    return out


def condensed_multiply_i32(cdata, cols: int, vec, lanes: int = 8) -> np.ndarray:
    """CondensedMultiply<int32_t> (crypto/matrix.cc:320-374)."""
    cdata = _u64(cdata)
    vec = _u64(vec)
    rows = cdata.size // (cols // 2)
    out = np.empty(rows, dtype=np.uint64)
    _call(lib().orc_condensed_multiply_i32, _p(cdata), rows, cols, _p(vec), lanes, _p(out))
    return out


def multiply_i16(m, vec) -> np.ndarray:
    """Matrix<int16_t>::Multiply<uint64_t> (crypto/matrix.h:183-194)."""
    m = np.ascontiguousarray(m, dtype=np.int16)
    vec = _u64(vec)
    rows, cols = m.shape
    out = np.empty(rows, dtype=np.uint64)
    lib().orc_multiply_i16(m.ctypes.data_as(_i16p), rows, cols, _p(vec), _p(out))
    return out


def condensed_multiply_u16_rows(cdata, cols, row0, nrows, vec) -> np.ndarray:
    cdata = _u64(cdata)
    vec = _u64(vec)
    out = np.empty(nrows, dtype=np.uint64)
    lib().orc_condensed_multiply_u16_rows(_p(cdata), cols, row0, nrows, _p(vec), _p(out))
    return out


def condensed_multiply_i32_rows(cdata, cols, row0, nrows, vec) -> np.ndarray:
    cdata = _u64(cdata)
    vec = _u64(vec)
    out = np.empty(nrows, dtype=np.uint64)
    lib().orc_condensed_multiply_i32_rows(_p(cdata), cols, row0, nrows, _p(vec), _p(out))
    return out


# ---- packing / external product (crypto/lwes_to_rlwe.cc, external_product.cc) ----

def finalize_matrix_pack(P: Params, b, b_part, y_vec_h, t_vec_h, a_tilde):
    """FinalizeMatrixPack (crypto/lwes_to_rlwe.cc:448-479); returns (a, b, max_err)."""
    d = P.d
    b, b_part, y_vec_h, t_vec_h, a_tilde = map(_u64, (b, b_part, y_vec_h, t_vec_h, a_tilde))
    oa = np.empty(d, dtype=np.uint64)
    ob = np.empty(d, dtype=np.uint64)
    me = ctypes.c_double(0.0)
    cp = P.c()
    _call(lib().orc_finalize_matrix_pack, ctypes.byref(cp), _p(b), _p(b_part),
          _p(y_vec_h), _p(t_vec_h), _p(a_tilde), _p(oa), _p(ob), ctypes.byref(me))
    return oa, ob, me.value


def matrix_pack(P: Params, b, y_vec_g, y_vec_h, pack, t_vec_h, a_tilde):
    """MatrixPack (crypto/lwes_to_rlwe.cc:481-508)."""
    d = P.d
    b, y_vec_g, y_vec_h, t_vec_h, a_tilde = map(_u64, (b, y_vec_g, y_vec_h, t_vec_h, a_tilde))
    pack = np.ascontiguousarray(pack)
    oa = np.empty(d, dtype=np.uint64)
    ob = np.empty(d, dtype=np.uint64)
    cp = P.c()
    _call(lib().orc_matrix_pack, ctypes.byref(cp), _p(b), _p(y_vec_g), _p(y_vec_h),
          pack.ctypes.data_as(_u64p), _p(t_vec_h), _p(a_tilde), _p(oa), _p(ob))
    return oa, ob


def external_product(P: Params, ct_a, ct_b, rgsw_a, rgsw_b):
    """ExternalProduct (crypto/external_product.cc:35-114); returns (a, b, max_err)."""
    d = P.d
    ct_a, ct_b, rgsw_a, rgsw_b = map(_u64, (ct_a, ct_b, rgsw_a, rgsw_b))
    oa = np.empty(d, dtype=np.uint64)
    ob = np.empty(d, dtype=np.uint64)
    me = ctypes.c_double(0.0)
    cp = P.c()
    _call(lib().orc_external_product, ctypes.byref(cp), _p(ct_a), _p(ct_b),
          _p(rgsw_a), _p(rgsw_b), _p(oa), _p(ob), ctypes.byref(me))
    return oa, ob, me.value


def preprocess_matrix_pack(P: Params, a_is, w):
    """PreprocessMatrixPack (crypto/lwes_to_rlwe.cc:421-446) on explicit LWE a-vectors.

    Returns (pack, a_tilde, t_vec_h); pack is d x (k*d/2) condensed words
    (mat_bytes 4) or d x k*d int16 (mat_bytes 2)."""
    s = sizes(P)
    d, k = s.d, s.k
    a_is, w = _u64(a_is), _u64(w)
    pack = _alloc_pack(P, d)
    a_tilde = np.empty(d, dtype=np.uint64)
    t_vec_h = np.empty((k, d), dtype=np.uint64)
    cp = P.c()
    _call(lib().orc_preprocess_matrix_pack, ctypes.byref(cp), _p(a_is), _p(w),
          pack.ctypes.data_as(ctypes.c_void_p), _p(a_tilde), _p(t_vec_h))
    return pack, a_tilde, t_vec_h


def pack_reference(P: Params, a_is, w, b, y_vec_g, y_vec_h):
    """Pack (crypto/lwes_to_rlwe.cc:370-419): the automorphism-based packing."""
    d = P.d
    a_is, w, b, y_vec_g, y_vec_h = map(_u64, (a_is, w, b, y_vec_g, y_vec_h))
    oa = np.empty(d, dtype=np.uint64)
    ob = np.empty(d, dtype=np.uint64)
    cp = P.c()
    _call(lib().orc_pack_reference, ctypes.byref(cp), _p(a_is), _p(w), _p(b),
          _p(y_vec_g), _p(y_vec_h), _p(oa), _p(ob))
    return oa, ob


# ---- server material / process (crypto/pir/server) ----

def _alloc_pack(P: Params, rows: int) -> np.ndarray:
    s = sizes(P)
    if P.mat_bytes == 4:
        return np.zeros((rows, s.pack_words_per_row), dtype=np.uint64)
    if P.mat_bytes == 8:
        return np.zeros((rows, s.pack_cols), dtype=np.uint64)
    return np.zeros((rows, s.pack_cols), dtype=np.int16)


@dataclasses.dataclass
class Material:
    """The resident server material, laid out exactly as PirServer::Create
    concatenates it (crypto/pir/server/pir_server.cc:100-118)."""

    db: np.ndarray          # M x db_words_per_row condensed uint64 words
    pack: np.ndarray        # M x pack_words_per_row uint64 (int32 condensed) or M x k*d int16
    a_tilde: np.ndarray     # (ell*t) x d
    t_vec_h: np.ndarray     # (ell*t) x k x d
    second_dim_a: np.ndarray  # 2e x d (zeros when t == 1)
    db_cols: int            # original (padded) DB columns K


def preprocess(P: Params, db_raw: np.ndarray, public_seed: int) -> Material:
    """PreprocessingServer::Preprocess (crypto/pir/server/preprocessing_server.h:68-160)."""
    s = sizes(P)
    db_raw = np.ascontiguousarray(db_raw, dtype=np.uint16)
    assert db_raw.size == P.num_entries * s.d * s.ell
    db = np.zeros((s.M, s.db_words_per_row), dtype=np.uint64)
    pack = _alloc_pack(P, s.M)
    a_tilde = np.zeros((s.ell * s.t, s.d), dtype=np.uint64)
    t_vec_h = np.zeros((s.ell * s.t, s.k, s.d), dtype=np.uint64)
    sda = np.zeros((2 * s.e, s.d), dtype=np.uint64)
    cp = P.c()
    _call(lib().orc_preprocess, ctypes.byref(cp), db_raw.ctypes.data_as(_u16p),
          ctypes.c_uint64(public_seed), _p(db), pack.ctypes.data_as(ctypes.c_void_p),
          _p(a_tilde), _p(t_vec_h), _p(sda))
    return Material(db, pack, a_tilde, t_vec_h, sda, s.K)


def synthetic_material(P: Params, seed: int) -> Material:
    """Range-faithful random material (no preprocessing): DB entries uniform in
    [0, p), pack entries uniform in [-2^(pack_log_digit+log d-1), 2^(...)) (the
    range of BuildPreprocessMatrix's output), t_vec_h signed pack_log_digit-bit
    digits, a_tilde and second_dim_a uniform mod q."""
    s = sizes(P)
    rng = np.random.default_rng(seed)
    p = P.plaintext_modulus
    dbv = rng.integers(0, p, size=(s.M, s.K), dtype=np.int64)
    dbv[:, s.r:] = 0
    if P.db_bytes == 2:
        db = condense_u16(dbv.astype(np.uint16))
    else:
        db = condense_u8(dbv.astype(np.uint8))
    bound = 1 << min(30, P.pack_log_digit + P.log_degree - 1)
    if P.mat_bytes == 2:
        bound = min(bound, 1 << 15)
    pv = rng.integers(-bound, bound, size=(s.M, s.pack_cols), dtype=np.int64)
    if P.mat_bytes == 4:
        pack = condense_i32(pv.astype(np.int32))
    elif P.mat_bytes == 8:
        pack = pv.astype(np.int64).view(np.uint64)
    else:
        pack = pv.astype(np.int16)
    qmask = (1 << P.log_q) - 1 if P.log_q < 64 else (1 << 64) - 1
    a_tilde = (rng.integers(0, 1 << 63, size=(s.ell * s.t, s.d), dtype=np.uint64) * 2 +
               rng.integers(0, 2, size=(s.ell * s.t, s.d), dtype=np.uint64)) & np.uint64(qmask)
    h = 1 << (P.pack_log_digit - 1)
    t_vec_h = rng.integers(-h, h, size=(s.ell * s.t, s.k, s.d), dtype=np.int64).astype(np.uint64)
    sda = (rng.integers(0, 1 << 63, size=(2 * s.e, s.d), dtype=np.uint64) * 2) & np.uint64(qmask)
    return Material(db, pack, a_tilde, t_vec_h, sda, s.K)


@dataclasses.dataclass
class Request:
    """PirRequest<uint64_t> (crypto/pir/pir_params.h:196-208)."""

    first_dim: np.ndarray      # r values
    second_dim: np.ndarray     # 2e x d (empty when t == 1)
    packing_key: np.ndarray    # 2k x d


def synthetic_request(P: Params, seed: int) -> Request:
    s = sizes(P)
    rng = np.random.default_rng(seed)
    qmask = np.uint64((1 << P.log_q) - 1 if P.log_q < 64 else (1 << 64) - 1)
    def u(shape):
        return rng.integers(0, 1 << 64, size=shape, dtype=np.uint64) & qmask
    sd = u((2 * s.e, s.d)) if s.t > 1 else np.zeros((0, s.d), dtype=np.uint64)
    return Request(u(s.r), sd, u((2 * s.k, s.d)))


def client_request(P: Params, public_seed: int, secret_seed: int, index: int) -> Request:
    """PirClient::CreateRequest (crypto/pir/client/pir_client.cc:61-80)."""
    s = sizes(P)
    fd = np.zeros(s.r, dtype=np.uint64)
    sd = np.zeros((2 * s.e if s.t > 1 else 0, s.d), dtype=np.uint64)
    pk = np.zeros((2 * s.k, s.d), dtype=np.uint64)
    cp = P.c()
    sd_buf = sd if sd.size else np.zeros(1, dtype=np.uint64)
    _call(lib().orc_client_request, ctypes.byref(cp), ctypes.c_uint64(public_seed),
          ctypes.c_uint64(secret_seed), index, _p(fd), _p(sd_buf), _p(pk))
    return Request(fd, sd, pk)


def client_decrypt(P: Params, secret_seed: int, response: np.ndarray) -> np.ndarray:
    """PirClient::ProcessResponse (crypto/pir/client/pir_client.cc:260-281)."""
    s = sizes(P)
    response = _u64(response)
    out = np.zeros(s.ell * s.d, dtype=np.uint64)
    cp = P.c()
    _call(lib().orc_client_decrypt, ctypes.byref(cp), ctypes.c_uint64(secret_seed),
          _p(response), _p(out))
    return out


def set_matrix_threads(n: int) -> None:
    """Threads for ProcessResponse's DB and pack products (1 = the reference's
    single-threaded loop); results are unchanged."""
    lib().orc_set_matrix_threads.restype = None
    lib().orc_set_matrix_threads(ctypes.c_int(int(n)))


def process_response(P: Params, mat: Material, req: Request, first_dim_len=None):
    """PirServer::ProcessResponse (crypto/pir/server/pir_server.cc:152-306).

    Returns (response ell x 2 x d, stage_ms[3], max rounding distance)."""
    s = sizes(P)
    fd = _u64(req.first_dim)
    sd = _u64(req.second_dim) if req.second_dim.size else np.zeros(1, dtype=np.uint64)
    pk = _u64(req.packing_key)
    out = np.zeros((s.ell, 2, s.d), dtype=np.uint64)
    stage = (ctypes.c_double * 3)()
    me = ctypes.c_double(0.0)
    cp = P.c()
    sd_polys = req.second_dim.shape[0] if req.second_dim.ndim == 2 else 0
    _call(lib().orc_process_response, ctypes.byref(cp), _p(_u64(mat.db)), mat.db_cols,
          np.ascontiguousarray(mat.pack).ctypes.data_as(ctypes.c_void_p),
          _p(_u64(mat.a_tilde)), _p(_u64(mat.t_vec_h)), _p(_u64(mat.second_dim_a)),
          _p(fd), len(fd) if first_dim_len is None else first_dim_len, _p(sd), sd_polys,
          _p(pk), pk.shape[0], _p(out), stage, ctypes.byref(me))
    return out, list(stage), me.value


def create_check(P: Params, num_db_matrices, db_rows, db_cols, num_outputs,
                 outputs_len, second_dim_a_len):
    """PirServer::Create validation (crypto/pir/server/pir_server.cc:52-95)."""
    cp = P.c()
    _call(lib().orc_create_check, ctypes.byref(cp), num_db_matrices, db_rows, db_cols,
          num_outputs, outputs_len, second_dim_a_len)


def interpolate(row, d: int, t: int, p: int) -> np.ndarray:
    """Interpolate (crypto/db_row_interpolation.cc:115-228)."""
    row = _u64(row)
    out = np.empty(d * t, dtype=np.uint64)
    _call(lib().orc_interpolate, _p(row), d, t, ctypes.c_uint64(p), _p(out))
    return out


def compute_hint_matrix(db, db_bits: int, a) -> np.ndarray:
    """ComputeHintMatrix (crypto/hint_computation.cc:78-160): db is rows x cols
    (values below 2^db_bits), a is n x d; returns rows x d uint64."""
    db = np.ascontiguousarray(db, dtype=np.uint16)
    a = _u64(a)
    rows, cols = db.shape
    n, d = a.shape
    out = np.empty((rows, d), dtype=np.uint64)
    _call(lib().orc_compute_hint_matrix, db.ctypes.data_as(_u16p), rows, cols, db_bits,
          _p(a), n, d, _p(out))
    return out


def sample_a_components(P: Params, public_seed: int):
    """The a-components preprocess() draws from public_seed:
    (first_dim_a, second_dim_a or None, pack_a)."""
    s = sizes(P)
    first = np.empty((s.S, s.d), dtype=np.uint64)
    second = np.empty((2 * s.e, s.d), dtype=np.uint64) if s.t > 1 else None
    pack_a = np.empty((2 * s.k, s.d), dtype=np.uint64)
    cp = P.c()
    _call(lib().orc_sample_a_components, ctypes.byref(cp), ctypes.c_uint64(public_seed),
          _p(first), _p(second) if second is not None else None, _p(pack_a))
    return first, second, pack_a


def fill_db_u16(seed: int, n: int, p: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint16)
    lib().orc_fill_db_u16(ctypes.c_uint64(seed), out.ctypes.data_as(_u16p),
                          ctypes.c_size_t(n), ctypes.c_uint64(p))
    return out


def fill_u64(seed: int, n: int, mask: int = (1 << 64) - 1) -> np.ndarray:
    out = np.empty(n, dtype=np.uint64)
    lib().orc_fill_u64(ctypes.c_uint64(seed), _p(out), ctypes.c_size_t(n),
                       ctypes.c_uint64(mask))
    return out



# ---- bit-packed formats (crypto/polynomial.cc, crypto/pir/server/pir_server.cc) ----

_u8p = ctypes.POINTER(ctypes.c_uint8)


def poly_to_proto(coeffs, log_modulus: int) -> bytes:
    """Polynomial::ToProto(log_modulus).encoded_coeffs (polynomial.cc:562-590)."""
    c = _u64(coeffs)
    cap = (c.size * 64 + 7) // 8 + 1
    out = np.zeros(cap, dtype=np.uint8)
    n = ctypes.c_size_t(0)
    _call(lib().orc_poly_to_proto, _p(c), ctypes.c_int(c.size), ctypes.c_int(log_modulus),
          _p(out, _u8p), ctypes.c_size_t(cap), ctypes.byref(n))
    return out[:n.value].tobytes()


def poly_from_proto(data: bytes, length: int, log_modulus: int) -> np.ndarray:
    """Polynomial::CreateFromProto (polynomial.cc:208-252)."""
    b = np.frombuffer(data, dtype=np.uint8) if len(data) else np.zeros(1, np.uint8)
    out = np.zeros(max(length, 1), dtype=np.uint64)
    _call(lib().orc_poly_from_proto, _p(np.ascontiguousarray(b), _u8p),
          ctypes.c_size_t(len(data)), ctypes.c_int(length), ctypes.c_int(log_modulus), _p(out))
    return out[:length]


def bit_pack(values, bits: int, is_signed: bool = False, value_bits: int = 64) -> bytes:
    """BitPackSigned / BitPackCoefficients (pir_server.cc:309-399)."""
    v = _u64(np.asarray(values).astype(np.int64).view(np.uint64) if is_signed else values)
    lib().orc_bit_pack.restype = ctypes.c_size_t
    n = lib().orc_bit_pack(_p(v), ctypes.c_size_t(v.size), ctypes.c_int(bits),
                           ctypes.c_int(int(is_signed)), ctypes.c_int(value_bits), None)
    out = np.zeros(max(n, 1), dtype=np.uint8)
    lib().orc_bit_pack(_p(v), ctypes.c_size_t(v.size), ctypes.c_int(bits),
                       ctypes.c_int(int(is_signed)), ctypes.c_int(value_bits), _p(out, _u8p))
    return out[:n].tobytes()


def bit_unpack(data: bytes, n: int, bits: int, is_signed: bool = False,
               value_bits: int = 64) -> np.ndarray:
    """BitUnpackSigned / BitUnpackCoefficients (pir_server.cc:338-429); signed
    values come back as int64."""
    b = np.frombuffer(data, dtype=np.uint8).copy() if len(data) else np.zeros(1, np.uint8)
    out = np.zeros(max(n, 1), dtype=np.uint64)
    lib().orc_bit_unpack.restype = None
    lib().orc_bit_unpack(_p(b, _u8p), ctypes.c_size_t(n), ctypes.c_int(bits),
                         ctypes.c_int(int(is_signed)), ctypes.c_int(value_bits), _p(out))
    return out[:n].view(np.int64) if is_signed else out[:n]


def save_server(P: Params, mat: Material) -> bytes:
    """PirServer::SaveToFile's byte layout (pir_server.cc:433-592) for the
    combined material."""
    cp = P.c()
    f = lib().orc_save_server
    f.restype = ctypes.c_size_t
    pack = np.ascontiguousarray(mat.pack)
    args = (ctypes.byref(cp), _p(_u64(mat.db)), ctypes.c_int(mat.db_cols),
            pack.ctypes.data_as(ctypes.c_void_p), _p(_u64(mat.a_tilde)),
            _p(_u64(mat.t_vec_h)), _p(_u64(mat.second_dim_a)))
    n = f(*args, None)
    out = np.zeros(n, dtype=np.uint8)
    f(*args, _p(out, _u8p))
    return out.tobytes()


def load_server(P: Params, data: bytes) -> Material:
    """PirServer::LoadFromBuffer (pir_server.cc:594-793) back into the combined
    material."""
    s = sizes(P)
    db = np.zeros((s.M, s.db_words_per_row), dtype=np.uint64)
    pack = _alloc_pack(P, s.M)
    a_tilde = np.zeros((s.ell * s.t, s.d), dtype=np.uint64)
    t_vec_h = np.zeros((s.ell * s.t, s.k, s.d), dtype=np.uint64)
    sda = np.zeros((2 * s.e, s.d), dtype=np.uint64)
    b = np.frombuffer(data, dtype=np.uint8)
    cp = P.c()
    _call(lib().orc_load_server, ctypes.byref(cp), _p(np.ascontiguousarray(b), _u8p),
          ctypes.c_size_t(len(data)), _p(db), pack.ctypes.data_as(ctypes.c_void_p),
          _p(a_tilde), _p(t_vec_h), _p(sda))
    return Material(db, pack, a_tilde, t_vec_h, sda, s.K)


# ---- vPIR online path (crypto/vpir/server/vpir_server.cc) ----

def vpir_process(P: Params, local_finalize: bool, db, db_cols: int, pack, a_tilde, t_vec_h,
                 first_dim, packing_key, matmul: bool = False) -> np.ndarray:
    """VpirServer::ProcessResponse (vpir_server.cc:101-195), or with matmul
    ProcessResponseWithMatMul (:197-287); returns ell x d b polynomials."""
    s = sizes(P)
    d, ell = s.d, s.ell
    fd = _u64(first_dim).reshape(-1)
    pk = _u64(packing_key).reshape(-1)
    out = np.zeros((ell, d), dtype=np.uint64)
    cp = P.c()
    pack = np.ascontiguousarray(pack)
    if matmul:
        _call(lib().orc_vpir_process_matmul, ctypes.byref(cp), ctypes.c_int(int(local_finalize)),
              _p(_u64(db)), ctypes.c_int(db_cols), pack.ctypes.data_as(ctypes.c_void_p),
              _p(_u64(t_vec_h)), _p(fd), ctypes.c_int(fd.size // d), _p(pk),
              ctypes.c_int(pk.size // d), _p(out))
    else:
        _call(lib().orc_vpir_process, ctypes.byref(cp), ctypes.c_int(int(local_finalize)),
              _p(_u64(db)), ctypes.c_int(db_cols), pack.ctypes.data_as(ctypes.c_void_p),
              _p(_u64(a_tilde)), _p(_u64(t_vec_h)), _p(fd), ctypes.c_int(fd.size // d), _p(pk),
              ctypes.c_int(pk.size // d), _p(out))
    return out
