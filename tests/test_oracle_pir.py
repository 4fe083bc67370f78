"""Pins the oracle's ProcessResponse, packing and interpolation by the
properties the reference's own tests check: decrypt(response) == DB record
(crypto/pir/server/pir_server_test.cc:246-412), MatrixPack == Pack and
decrypt(pack) == message (crypto/lwes_to_rlwe_test.cc:150-468), the packing
matrix norm bound (:470-526), Interpolate(Evaluate(P)) == P
(crypto/db_row_interpolation_test.cc:91-142) and the reference's error
messages (pir_server_test.cc:343-525)."""
import numpy as np
import pytest

import oracle as O
from tests.helpers import PIR_SERVER_TEST_TYPES, PIR_SERVER_TEST_TYPES_U32, pir_server_test_params


def _e2e(P, idxs, db_seed=1, pub=77):
    s = O.sizes(P)
    db = O.fill_db_u16(db_seed, P.num_entries * s.d * s.ell, P.plaintext_modulus)
    mat = O.preprocess(P, db, public_seed=pub)
    for idx in idxs:
        req = O.client_request(P, pub, 1000 + idx, idx)
        assert req.first_dim.size == s.r
        resp, _, _ = O.process_response(P, mat, req)
        msg = O.client_decrypt(P, 1000 + idx, resp)
        np.testing.assert_array_equal(msg.astype(np.uint16),
                                      db[idx * s.d * s.ell:(idx + 1) * s.d * s.ell])


@pytest.mark.parametrize("name", sorted(PIR_SERVER_TEST_TYPES))
def test_end_to_end(name):
    # pir_server_test.cc:246-275.
    P = pir_server_test_params(*PIR_SERVER_TEST_TYPES[name])
    idxs = [0, 1, 10, P.num_entries - 2, P.num_entries - 1] + list(range(0, P.num_entries, 211))
    _e2e(P, idxs)


@pytest.mark.parametrize("name", sorted(PIR_SERVER_TEST_TYPES_U32))
def test_end_to_end_uint32(name):
    # pir_server_test.cc:246-275 for the CoeffType = uint32_t types (:55-93):
    # every coefficient wraps at 2^32 and the answers decrypt.
    P = pir_server_test_params(*PIR_SERVER_TEST_TYPES_U32[name], coeff_bits=32)
    idxs = [0, 1, 10, P.num_entries - 2, P.num_entries - 1] + list(range(0, P.num_entries, 211))
    _e2e(P, idxs)


@pytest.mark.parametrize("name", sorted(PIR_SERVER_TEST_TYPES_U32))
def test_uint32_answers_are_32_bit_and_differ_from_uint64_only_through_wraps(name):
    # With t > 1 the Horner state and Rescale see the 2^32 wrap, so the
    # uint64_t arithmetic on the same inputs answers differently; with t = 1
    # every value before Rescale is already reduced mod q <= 2^32.
    import dataclasses
    P = pir_server_test_params(*PIR_SERVER_TEST_TYPES_U32[name], coeff_bits=32)
    s = O.sizes(P)
    db = O.fill_db_u16(2, P.num_entries * s.d * s.ell, P.plaintext_modulus)
    mat = O.preprocess(P, db, public_seed=3)
    assert int(mat.t_vec_h.max()) < 1 << 32 and int(mat.a_tilde.max()) < 1 << 32
    req = O.client_request(P, 3, 44, 7)
    got, _, _ = O.process_response(P, mat, req)
    assert int(got.max()) < 1 << 32
    wide, _, _ = O.process_response(dataclasses.replace(P, coeff_bits=64), mat, req)
    if P.interp_degree > 1:
        assert not np.array_equal(got, wide)
    else:
        np.testing.assert_array_equal(got, wide)


@pytest.mark.parametrize("name", ["Uint64Mod62", "Uint64Mod0"])
def test_uint64_pack_matrix_answers_as_the_signed_one(name):
    # PirServer<uint16_t, uint64_t, uint64_t> (pir_server.cc:806-818): the
    # pack matrix holds BuildPreprocessMatrix's centered values in two's
    # complement (lwes_to_rlwe.cc:210-222), so its products equal the int32_t
    # matrix's mod 2^64 and the answers are the same, and decrypt.
    import dataclasses
    lq, q1, q2, t, _ = PIR_SERVER_TEST_TYPES[name]
    P4 = pir_server_test_params(lq, q1, q2, t, 4)
    P8 = dataclasses.replace(P4, mat_bytes=8)
    s = O.sizes(P4)
    db = O.fill_db_u16(4, P4.num_entries * s.d * s.ell, P4.plaintext_modulus)
    m4 = O.preprocess(P4, db, public_seed=9)
    m8 = O.preprocess(P8, db, public_seed=9)
    # The entries are NTT outputs over Z_2^64 (not reduced mod q): equal to
    # the int32 matrix's values mod 2^log_q, whose condensed words are the
    # row-major int32 values.
    assert m8.pack.shape == (s.M, s.pack_cols)
    q = 1 << P4.log_q if P4.log_q < 64 else 1 << 64
    centered = [((int(x) % q) + q // 2) % q - q // 2 for x in m8.pack.reshape(-1)[:4096]]
    np.testing.assert_array_equal(np.array(centered, dtype=np.int64),
                                  m4.pack.view(np.int32).reshape(-1)[:4096].astype(np.int64))
    for idx in (0, 5, P4.num_entries - 1):
        req = O.client_request(P4, 9, 70 + idx, idx)
        a4, _, _ = O.process_response(P4, m4, req)
        a8, _, _ = O.process_response(P8, m8, req)
        np.testing.assert_array_equal(a8, a4)
        msg = O.client_decrypt(P8, 70 + idx, a8)
        np.testing.assert_array_equal(msg.astype(np.uint16), db[idx * s.d:(idx + 1) * s.d])


def test_uint32_rejects_log_q_above_32():
    import dataclasses
    P = pir_server_test_params(*PIR_SERVER_TEST_TYPES["Uint64Mod62"], coeff_bits=32)
    with pytest.raises(O.OracleError, match="coefficient width"):
        O.sizes(dataclasses.replace(P))


@pytest.mark.parametrize("name", ["Uint64Mod62", "Uint64Mod0Interp1"])
def test_end_to_end_large_entry_size_multiple(name):
    # pir_server_test.cc:277-341 (ell = 3).
    P = pir_server_test_params(*PIR_SERVER_TEST_TYPES[name], ell=3)
    _e2e(P, [0, 1, 10, P.num_entries - 1])


@pytest.mark.parametrize("name", ["Uint64Mod62", "Uint64Mod62Interp1"])
def test_end_to_end_non_ring_degree_multiple(name):
    # pir_server_test.cc:354-412: num_entries = 50 t, so n/t is not a multiple of d.
    lq, q1, q2, t, mb = PIR_SERVER_TEST_TYPES[name]
    P = pir_server_test_params(lq, q1, q2, t, mb, n=50 * t)
    _e2e(P, [0, 1, 10, 50 * t - 2, 50 * t - 1])


def test_end_to_end_metrics_params():
    # The pir_metrics_tool parameters (q = 2^52, p = 65535, gadgets (19,2)/(19,3)).
    _e2e(O.metrics_tool_params(n=512, log_d=8, t=2), [0, 123, 511])
    _e2e(O.metrics_tool_params(n=256, log_d=8, t=4, ell=2), [0, 77, 255])


def test_process_response_errors():
    # pir_server_test.cc:343-449.
    P = pir_server_test_params(*PIR_SERVER_TEST_TYPES["Uint64Mod62"])
    s = O.sizes(P)
    mat = O.synthetic_material(P, seed=3)
    req = O.synthetic_request(P, seed=4)
    bad = O.Request(req.first_dim, req.second_dim, req.packing_key[:-1])
    with pytest.raises(O.OracleError, match="Packing key must contain 2 \\* num_digits"):
        O.process_response(P, mat, bad)
    bad = O.Request(req.first_dim[:-1], req.second_dim, req.packing_key)
    with pytest.raises(O.OracleError, match="First dimension query size \\(31\\) is invalid "
                                            "for DB columns \\(32\\)"):
        O.process_response(P, mat, bad)
    big = np.zeros(100000, np.uint64)
    with pytest.raises(O.OracleError, match="First dimension query size"):
        O.process_response(P, mat, O.Request(big, req.second_dim, req.packing_key),
                           first_dim_len=s.K + 1)
    bad = O.Request(req.first_dim, req.second_dim[:-1], req.packing_key)
    with pytest.raises(O.OracleError, match="second_dimension_query size mismatch"):
        O.process_response(P, mat, bad)


def test_create_errors():
    # pir_server_test.cc:451-525 (PirServer::Create validation, pir_server.cc:62-95).
    P = pir_server_test_params(*PIR_SERVER_TEST_TYPES["Uint64Mod62"])
    s = O.sizes(P)
    ok = (s.ell, s.t * s.d, s.K, s.ell, s.t, 2 * s.e)
    O.create_check(P, *ok)
    cases = [
        ((2,) + ok[1:], "db_matrices size mismatch."),
        ((s.ell, 1) + ok[2:], "DB matrix must have td rows."),
        ((s.ell, s.t * s.d, 1) + ok[3:], "DB matrix must have at least N/t columns."),
        (ok[:3] + (2,) + ok[4:], "preprocessed_outputs size mismatch."),
        (ok[:4] + (s.t - 1, ok[5]), "Preprocessed outputs must be of length t."),
        (ok[:5] + (2 * s.e - 1,), "second_dim_query_a must have length 2"),
    ]
    for args, msg in cases:
        with pytest.raises(O.OracleError, match=msg.replace(".", "\\.").replace("*", "\\*")):
            O.create_check(P, *args)


def test_params_errors():
    # pir_params_test.cc / pir_params.cc:35-62.
    base = dict(log_degree=5, num_entries=1024, interp_degree=32)
    for over, msg in [(dict(num_entries=0), "num_entries must be strictly positive"),
                      (dict(interp_degree=3), "interpolation_degree must be a power of two"),
                      (dict(num_entries=1000), "must divide num_entries"),
                      (dict(interp_degree=128), "must be less than or equal to 2 \\* ring degree"),
                      (dict(entry_size_multiple=0), "entry_size_multiple must be strictly positive")]:
        kw = dict(base)
        kw.update(over)
        with pytest.raises(O.OracleError, match=msg):
            O.sizes(O.Params(**kw))


def _lwe_setup(P, seed):
    """Random LWE a-vectors and packing-key a-parts for the packing tests."""
    rng = np.random.default_rng(seed)
    d, k = P.d, P.pack_num_digits
    qmask = np.uint64((1 << P.log_q) - 1 if P.log_q < 64 else (1 << 64) - 1)
    a_is = rng.integers(0, 1 << 64, size=(d, d), dtype=np.uint64) & qmask
    w = rng.integers(0, 1 << 64, size=(2 * k, d), dtype=np.uint64) & qmask
    b = rng.integers(0, 1 << 64, size=d, dtype=np.uint64) & qmask
    y = rng.integers(0, 1 << 64, size=(2 * k, d), dtype=np.uint64) & qmask
    return a_is, w, b, y[:k], y[k:]


@pytest.mark.parametrize("log_q,ld,mat_bytes", [(45, 4, 4), (54, 4, 4), (64, 4, 4), (52, 19, 4),
                                                 (52, 19, 2)])
def test_matrix_pack_equals_pack(log_q, ld, mat_bytes):
    # lwes_to_rlwe_test.cc:364-468: the matrix of BuildPreprocessMatrix applies the
    # same linear map as Pack's automorphism sum (lwes_to_rlwe.cc:128-162).
    nd = (log_q + ld - 1) // ld - 2 if ld == 4 else 2
    P = O.Params(log_degree=6, log_q=log_q, log_q1=log_q, log_q2=log_q, pack_log_digit=ld,
                 pack_num_digits=nd, num_entries=64, interp_degree=1, mat_bytes=mat_bytes)
    a_is, w, b, yg, yh = _lwe_setup(P, log_q)
    pack, a_tilde, t_vec_h = O.preprocess_matrix_pack(P, a_is, w)
    ma, mb = O.matrix_pack(P, b, yg, yh, pack, t_vec_h, a_tilde)
    ra, rb = O.pack_reference(P, a_is, w, b, yg, yh)
    np.testing.assert_array_equal(ma, ra)
    if mat_bytes == 4 or ld == 4:
        np.testing.assert_array_equal(mb, rb)


@pytest.mark.parametrize("log_q", [45, 54, 64])
def test_pack_matrix_norm_bound(log_q):
    # lwes_to_rlwe_test.cc:470-526: |entry| < 4 sqrt(d) 2^log_digit.
    ld = 4
    nd = (log_q + ld - 1) // ld
    P = O.Params(log_degree=6, log_q=log_q, log_q1=log_q, log_q2=log_q, pack_log_digit=ld,
                 pack_num_digits=nd, num_entries=64, interp_degree=1, mat_bytes=4)
    a_is, w, _, _, _ = _lwe_setup(P, 9)
    pack, _, _ = O.preprocess_matrix_pack(P, a_is, w)
    vals = pack.view(np.int32)
    assert np.abs(vals.astype(np.int64)).max() < 4.0 * np.sqrt(P.d) * (1 << ld)


def _evaluate(coeffs, d, t, p):
    """P(omega^j) with omega = X^(2d/t) in Z_p[X]/(X^d+1) (the inverse of
    db_row_interpolation.cc:115-228)."""
    polys = coeffs.reshape(t, d).astype(np.int64)
    out = np.zeros((t, d), np.int64)
    for j in range(t):
        acc = np.zeros(d, np.int64)
        for i in range(t):
            power = (2 * d // t) * i * j % (2 * d)
            sh = np.zeros(d, np.int64)
            neg = power >= d
            pw = power - d if neg else power
            sh[pw:] = polys[i][:d - pw]
            sh[:pw] = -polys[i][d - pw:]
            if neg:
                sh = -sh
            acc = (acc + sh) % p
        out[j] = acc
    return out.reshape(-1).astype(np.uint64)


@pytest.mark.parametrize("d,t,p", [(32, 2, 257), (32, 8, 65535), (16, 32, 257), (64, 4, 65535)])
def test_interpolate_round_trip(d, t, p):
    rng = np.random.default_rng(d * t)
    coeffs = rng.integers(0, p, size=d * t, dtype=np.uint64)
    evals = _evaluate(coeffs, d, t, p)
    np.testing.assert_array_equal(O.interpolate(evals, d, t, p), coeffs)


def test_interpolate_errors():
    with pytest.raises(O.OracleError, match="plaintext_modulus must be odd"):
        O.interpolate(np.zeros(64, np.uint64), 32, 2, 256)


def test_row_parallel_products_match_single_thread():
    # The oracle's DB and pack products split over threads (the bench's
    # parallel CPU baseline) answer exactly as the single-threaded loop.
    P = O.metrics_tool_params(n=4096, log_d=8, t=4, ell=2)
    mat = O.synthetic_material(P, seed=31)
    req = O.synthetic_request(P, seed=32)
    want, _, _ = O.process_response(P, mat, req)
    try:
        for n in (2, 3, 7):
            O.set_matrix_threads(n)
            got, _, _ = O.process_response(P, mat, req)
            np.testing.assert_array_equal(got, want)
    finally:
        O.set_matrix_threads(1)
