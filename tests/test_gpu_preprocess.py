"""GPU offline preprocessing (PreprocessingServer::Preprocess,
crypto/pir/server/preprocessing_server.h:68-273) against the oracle, stage by
stage and end to end, bit for bit; and servers built from GPU-preprocessed
material answer requests that decrypt to the DB records."""
import numpy as np
import pytest

import inspire_amd as ia
import oracle as O
from helpers import (PIR_SERVER_TEST_TYPES, PIR_SERVER_TEST_TYPES_U32, gpu_request,
                     pir_server_test_params, to_gpu_params)

pytestmark = pytest.mark.gpu


# ---- Interpolate (crypto/db_row_interpolation.cc:115-228) ----

@pytest.mark.parametrize("d,t,p,db_bytes", [(32, 2, 257, 2), (32, 8, 65535, 2), (16, 32, 257, 2),
                                            (64, 4, 65535, 2), (4096, 8, 65535, 2),
                                            (64, 16, 251, 1), (8, 1, 257, 2)])
def test_interpolate_matches_oracle(d, t, p, db_bytes):
    rng = np.random.default_rng(d * t + p)
    nrows = 5
    rows = rng.integers(0, p, size=(nrows, t * d), dtype=np.uint64)
    got = ia.interpolate(rows, d, t, p, db_bytes)
    for i in range(nrows):
        want = O.interpolate(rows[i], d, t, p)
        np.testing.assert_array_equal(got[i].astype(np.uint64), want)


def test_interpolate_errors():
    rows = np.zeros(64, dtype=np.uint16)
    with pytest.raises(ia.PirError, match="ring_degree must be a power of two."):
        ia.interpolate(rows, 12, 2, 257)
    with pytest.raises(ia.PirError, match="interpolation_degree must be a power of two."):
        ia.interpolate(rows, 16, 3, 257)
    with pytest.raises(ia.PirError, match="less than or equal to 2 \\* ring_degree"):
        ia.interpolate(rows, 2, 8, 257)
    with pytest.raises(ia.PirError, match="plaintext_modulus must be odd."):
        ia.interpolate(rows, 16, 4, 256)


# ---- ComputeHintMatrix (crypto/hint_computation.cc:78-160) ----

@pytest.mark.parametrize("d,n,rows,db_bytes", [(2, 4, 7, 2), (16, 4, 21, 2), (16, 2, 16, 1),
                                               (256, 40, 5, 2), (2048, 5, 7, 2),
                                               (4096, 32, 3, 2), (4096, 3, 2, 1)])
def test_hint_matrix_matches_oracle(d, n, rows, db_bytes):
    # d = 2048: batches of safe_n = 4 with a short last batch; d = 4096: 2.
    rng = np.random.default_rng(d + n + rows)
    bits = 8 * db_bytes
    db = rng.integers(0, 1 << bits, size=(rows, d * n), dtype=np.uint64)
    a = rng.integers(0, 1 << 64, size=(n, d), dtype=np.uint64)
    got = ia.compute_hint_matrix(db, a, db_bytes)
    want = O.compute_hint_matrix(db.astype(np.uint16), bits, a)
    np.testing.assert_array_equal(got, want)


def test_hint_matrix_many_rows_in_blocks(monkeypatch):
    # A small scratch budget forces many row blocks and a short last one.
    monkeypatch.setenv("PIRG_HINT_BUDGET_BYTES", str(3 * 64 * 1024))
    rng = np.random.default_rng(5)
    d, n, rows = 64, 3, 37
    db = rng.integers(0, 1 << 16, size=(rows, d * n), dtype=np.uint64)
    a = rng.integers(0, 1 << 64, size=(n, d), dtype=np.uint64)
    got = ia.compute_hint_matrix(db, a)
    np.testing.assert_array_equal(got, O.compute_hint_matrix(db.astype(np.uint16), 16, a))


def test_hint_matrix_errors():
    a = np.zeros((1, 4), dtype=np.uint64)
    with pytest.raises(ia.PirError, match="Database matrix cannot be empty."):
        ia.compute_hint_matrix(np.zeros((0, 4), dtype=np.uint16), a)
    with pytest.raises(ia.PirError, match="a_components cannot be empty."):
        ia.compute_hint_matrix(np.zeros((2, 4), dtype=np.uint16), np.zeros((0, 4), np.uint64))
    with pytest.raises(ia.PirError, match="Database columns must match d"):
        ia.compute_hint_matrix(np.zeros((2, 3), dtype=np.uint16), a)


# ---- PreprocessMatrixPack (crypto/lwes_to_rlwe.cc:163-446) ----

def _pack_params(log_d, log_q, ld, nd, mat_bytes):
    return O.Params(log_degree=log_d, log_q=log_q, log_q1=log_q, log_q2=log_q, variance=4,
                    plaintext_modulus=257, pack_log_digit=ld, pack_num_digits=nd,
                    eval_log_digit=ld, eval_num_digits=nd, num_entries=1 << log_d,
                    interp_degree=1, entry_size_multiple=1, mat_bytes=mat_bytes)


@pytest.mark.parametrize("log_d,log_q,ld,nd,mat_bytes,nblocks", [
    (1, 62, 2, 30, 4, 1), (4, 64, 2, 31, 2, 2), (5, 62, 2, 30, 4, 3), (8, 52, 19, 2, 4, 2),
    (11, 52, 19, 2, 4, 1), (6, 32, 8, 4, 2, 2)])
def test_preprocess_matrix_pack_matches_oracle(log_d, log_q, ld, nd, mat_bytes, nblocks):
    P = _pack_params(log_d, log_q, ld, nd, mat_bytes)
    d = 1 << log_d
    rng = np.random.default_rng(log_d * 1000 + log_q)
    a_is = rng.integers(0, 1 << 64, size=(nblocks * d, d), dtype=np.uint64)
    w = rng.integers(0, 1 << 64, size=(2 * nd, d), dtype=np.uint64) & np.uint64(
        (1 << log_q) - 1 if log_q < 64 else (1 << 64) - 1)
    pack, a_tilde, t_vec_h = ia.preprocess_matrix_pack(to_gpu_params(P), a_is, w)
    for b in range(nblocks):
        wp, wa, wt = O.preprocess_matrix_pack(P, a_is[b * d:(b + 1) * d], w)
        np.testing.assert_array_equal(pack[b * d:(b + 1) * d], wp)
        np.testing.assert_array_equal(a_tilde[b], wa)
        np.testing.assert_array_equal(t_vec_h[b], wt)


def test_preprocess_matrix_pack_d4096_matches_oracle():
    # The ring degree of the benchmark configuration: 2047 sequential steps
    # per reduction chain, 2-digit signed gadget, q = 2^52.
    P = _pack_params(12, 52, 19, 2, 4)
    d = 4096
    rng = np.random.default_rng(4096)
    a_is = rng.integers(0, 1 << 64, size=(d, d), dtype=np.uint64)
    w = rng.integers(0, 1 << 52, size=(4, d), dtype=np.uint64)
    pack, a_tilde, t_vec_h = ia.preprocess_matrix_pack(to_gpu_params(P), a_is, w)
    wp, wa, wt = O.preprocess_matrix_pack(P, a_is, w)
    np.testing.assert_array_equal(a_tilde[0], wa)
    np.testing.assert_array_equal(t_vec_h[0], wt)
    np.testing.assert_array_equal(pack, wp)


# ---- Preprocess, end to end ----

def _gpu_preprocess(P, db, seed):
    first, _, pack_a = O.sample_a_components(P, seed)
    return ia.preprocess(to_gpu_params(P), db, first, pack_a)


def _check_material(got, want):
    assert got["db_cols"] == want.db_cols
    np.testing.assert_array_equal(got["db"], want.db)
    np.testing.assert_array_equal(got["pack"], want.pack)
    np.testing.assert_array_equal(got["a_tilde"], want.a_tilde)
    np.testing.assert_array_equal(got["t_vec_h"], want.t_vec_h)


@pytest.mark.parametrize("name,ell,n", [("Uint64Mod62", 1, 1024), ("Uint64Mod0", 1, 1024),
                                        ("Uint64Mod0Interp1", 1, 1024),
                                        ("Uint64Mod62", 3, 1024), ("Uint64Mod62Interp1", 1, 50),
                                        ("Uint64Mod62ModSwitch", 1, 50 * 32)])
def test_preprocess_matches_oracle(name, ell, n):
    P = pir_server_test_params(*PIR_SERVER_TEST_TYPES[name], ell=ell, n=n)
    s = O.sizes(P)
    db = O.fill_db_u16(17, P.num_entries * s.d * s.ell, P.plaintext_modulus)
    _check_material(_gpu_preprocess(P, db, 9), O.preprocess(P, db, public_seed=9))


@pytest.mark.parametrize("name", sorted(PIR_SERVER_TEST_TYPES_U32))
def test_preprocess_uint32_types_match_oracle(name):
    # PreprocessingServer<uint16_t, uint32_t, int16_t> of pir_server_test.cc's
    # Uint32 types: the material Polynomial<uint32_t> holds.
    P = pir_server_test_params(*PIR_SERVER_TEST_TYPES_U32[name], coeff_bits=32)
    s = O.sizes(P)
    db = O.fill_db_u16(19, P.num_entries * s.d * s.ell, P.plaintext_modulus)
    got = _gpu_preprocess(P, db, 9)
    assert int(got["t_vec_h"].max()) < 1 << 32
    _check_material(got, O.preprocess(P, db, public_seed=9))


@pytest.mark.parametrize("t,ell,mat_bytes", [(4, 1, 4), (1, 2, 2), (8, 2, 4)])
def test_preprocess_u8_db_matches_oracle(t, ell, mat_bytes):
    # Db = uint8_t: 64-column SIMD padding of the condensed DB, 8-bit hint
    # chunks, the plaintext modulus as a uint8_t (preprocessing_server.h:196).
    P = O.Params(log_degree=6, log_q=56, log_q1=56, log_q2=56, variance=4,
                 plaintext_modulus=251, pack_log_digit=8, pack_num_digits=7,
                 eval_log_digit=8, eval_num_digits=7, num_entries=40 * t,
                 interp_degree=t, entry_size_multiple=ell, db_bytes=1, mat_bytes=mat_bytes)
    s = O.sizes(P)
    rng = np.random.default_rng(t * 10 + ell)
    db = rng.integers(0, 251, size=P.num_entries * s.d * s.ell, dtype=np.uint16)
    first, _, pack_a = O.sample_a_components(P, 13)
    got = ia.preprocess(to_gpu_params(P), db.astype(np.uint8), first, pack_a)
    _check_material(got, O.preprocess(P, db, public_seed=13))


@pytest.mark.parametrize("log_d,n,t,ell", [(8, 512, 2, 1), (8, 256, 4, 2), (6, 2048, 8, 1)])
def test_preprocess_metrics_params_matches_oracle(log_d, n, t, ell):
    P = O.metrics_tool_params(n=n, log_d=log_d, t=t, ell=ell)
    s = O.sizes(P)
    db = O.fill_db_u16(log_d + t, P.num_entries * s.d * s.ell, P.plaintext_modulus)
    got = _gpu_preprocess(P, db, 21)
    _check_material(got, O.preprocess(P, db, public_seed=21))
    st = got["stage_ms"]
    assert st["total"] >= st["hints"] >= 0 and st["preprocess_pack"] >= 0


def _decrypt_check(P, srv, db, idxs, pub):
    s = O.sizes(P)
    for idx in idxs:
        req = O.client_request(P, pub, 500 + idx, idx)
        got = srv.process_response(gpu_request(req)).ciphertexts
        msg = O.client_decrypt(P, 500 + idx, got)
        np.testing.assert_array_equal(msg.astype(np.uint16),
                                      db[idx * s.d * s.ell:(idx + 1) * s.d * s.ell])


@pytest.mark.parametrize("case", ["pir_server_test", "metrics"])
def test_server_from_gpu_preprocessing_decrypts(case):
    if case == "pir_server_test":
        P = pir_server_test_params(*PIR_SERVER_TEST_TYPES["Uint64Mod62"], ell=2)
        idxs = [0, 1, 10, P.num_entries - 1]
    else:
        P = O.metrics_tool_params(n=1024, log_d=8, t=4, ell=2)
        idxs = [0, 333, 1023]
    s = O.sizes(P)
    db = O.fill_db_u16(23, P.num_entries * s.d * s.ell, P.plaintext_modulus)
    first, second, pack_a = O.sample_a_components(P, 31)
    with ia.PirServer.create_preprocessed(to_gpu_params(P), db, first, second, pack_a) as srv:
        _decrypt_check(P, srv, db, idxs, 31)
        # Same answers as a server built from the oracle's material.
        mat = O.preprocess(P, db, public_seed=31)
        req = O.client_request(P, 31, 7, idxs[-1])
        want, _, _ = O.process_response(P, mat, req)
        np.testing.assert_array_equal(srv.process_response(gpu_request(req)).ciphertexts, want)


def test_preprocess_d4096_t8_matches_oracle_stagewise():
    # Ring degree 4096, t = 8 with the pir_metrics_tool parameters (the
    # benchmark's shape at a smaller n). The oracle is too slow for the whole
    # pipeline here, so each stage is checked on the GPU's own intermediate:
    # interpolated columns of sampled row groups, then the hint and
    # PreprocessMatrixPack of one block recomputed by the oracle from the
    # GPU's DB matrix.
    P = O.metrics_tool_params(n=8 * 4096, log_d=12, t=8)
    s = O.sizes(P)
    d, t = s.d, s.t
    rng = np.random.default_rng(12)
    db = rng.integers(0, P.plaintext_modulus, size=P.num_entries * d, dtype=np.uint16)
    first, _, pack_a = O.sample_a_components(P, 41)
    got = ia.preprocess(to_gpu_params(P), db, first, pack_a)
    T = got["db"].view(np.uint16).reshape(s.M, s.K)
    for g in (0, 1, 2047, s.r - 1):
        want = O.interpolate(db[g * t * d:(g + 1) * t * d], d, t, P.plaintext_modulus)
        np.testing.assert_array_equal(T[:, g].astype(np.uint64), want)
    b = 3
    hint = O.compute_hint_matrix(T[b * d:(b + 1) * d, :s.S * d], 16, first)
    wp, wa, wt = O.preprocess_matrix_pack(P, hint, pack_a)
    np.testing.assert_array_equal(got["a_tilde"][b], wa)
    np.testing.assert_array_equal(got["t_vec_h"][b], wt)
    np.testing.assert_array_equal(got["pack"][b * d:(b + 1) * d], wp)


def test_server_from_gpu_preprocessing_d4096_decrypts():
    # Ring degree 4096, t = 8, q = 2^52 -> 2^30 / 2^18 as the metrics tool,
    # with 13-bit gadgets: the tool's 19-bit gadgets leave the noise at the
    # decryption boundary at this degree (a few percent of coefficients off
    # by one), which would make an exact-decrypt check flaky.
    P = O.Params(log_degree=12, log_q=52, log_q1=30, log_q2=18, variance=40,
                 plaintext_modulus=65535, pack_log_digit=13, pack_num_digits=3,
                 eval_log_digit=13, eval_num_digits=4, num_entries=8 * 4096, interp_degree=8,
                 entry_size_multiple=1)
    s = O.sizes(P)
    rng = np.random.default_rng(12)
    db = rng.integers(0, P.plaintext_modulus, size=P.num_entries * s.d * s.ell, dtype=np.uint16)
    first, second, pack_a = O.sample_a_components(P, 41)
    with ia.PirServer.create_preprocessed(to_gpu_params(P), db, first, second, pack_a) as srv:
        _decrypt_check(P, srv, db, [0, 4097, P.num_entries - 1], 41)
