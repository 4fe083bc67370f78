"""GPU parity: every HIP stage against the oracle on identical seeded inputs,
bit-exact (the path is integer-exact; its FFT rounds like the reference's)."""
import numpy as np
import pytest

import inspire_amd as ia
import oracle as O
from tests.helpers import (PIR_SERVER_TEST_TYPES, PIR_SERVER_TEST_TYPES_U32, gpu_request,
                           pir_server_test_params, server_from_material)

pytestmark = pytest.mark.gpu


def test_matrix_multiply_u16_kat():
    # crypto/matrix_test.cc:90-111: rows [1,2,3,4]*8, vec of ones -> 80, 80.
    m = np.tile(np.array([1, 2, 3, 4], np.uint16), (2, 8))
    out = ia.matrix_multiply_u16(O.condense_u16(m), 32, np.ones(32, np.uint64))
    assert out.tolist() == [80, 80]


def test_matrix_multiply_i32_kat():
    # crypto/matrix_test.cc:113-134: rows [-1, 2]*16, vec of ones -> 16, 16.
    m = np.tile(np.array([-1, 2], np.int32), (2, 16))
    out = ia.matrix_multiply_i32(O.condense_i32(m), 32, np.ones(32, np.uint64))
    assert out.tolist() == [16, 16]


@pytest.mark.parametrize("rows,cols", [(16, 64), (37, 96), (300, 4096), (64, 70000)])
def test_matrix_multiply_u16_random(rows, cols):
    rng = np.random.default_rng(rows * 7 + cols)
    m = rng.integers(0, 1 << 16, size=(rows, cols), dtype=np.uint16)
    vec = rng.integers(0, 1 << 64, size=cols, dtype=np.uint64)
    c = O.condense_u16(m)
    want = O.condensed_multiply_u16(c, cols, vec, lanes=1)
    got = ia.matrix_multiply_u16(c, cols, vec)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("rows,cols", [(16, 64), (33, 128), (200, 8192)])
def test_matrix_multiply_i32_random(rows, cols):
    rng = np.random.default_rng(rows + cols)
    m = rng.integers(-(1 << 31), 1 << 31, size=(rows, cols), dtype=np.int64).astype(np.int32)
    m[0, :4] = [-(1 << 31), (1 << 31) - 1, -1, 0]
    vec = rng.integers(0, 1 << 64, size=cols, dtype=np.uint64)
    c = O.condense_i32(m)
    want = O.condensed_multiply_i32(c, cols, vec, lanes=1)
    got = ia.matrix_multiply_i32(c, cols, vec)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("n,d,ub,vb,cb,signed", [
    (1, 2, 20, 20, 20, False),
    (1, 4, 32, 32, 20, False),
    (2, 8, 19, 52, 20, True),
    (2, 16, 19, 52, 19, False),
    (2, 64, 19, 52, 20, True),
    (6, 2048, 19, 52, 19, False),
    (6, 4096, 19, 52, 19, False),
    (2, 4096, 19, 52, 20, True),
    (3, 256, 64, 64, 16, False),
])
def test_inner_product_fft(n, d, ub, vb, cb, signed):
    rng = np.random.default_rng(d + n)
    u = rng.integers(0, 1 << 64, size=(n, d), dtype=np.uint64)
    v = rng.integers(0, 1 << 64, size=(n, d), dtype=np.uint64)
    if ub < 64:
        u &= np.uint64((1 << ub) - 1)
    if vb < 64:
        v &= np.uint64((1 << vb) - 1)
    want, _ = O.inner_product_fft(u, v, ub, vb, cb, signed)
    got = ia.inner_product_fft(u, v, ub, vb, cb, signed)
    np.testing.assert_array_equal(got, want)


def _compare(P, mat, reqs):
    with server_from_material(P, mat) as srv:
        for req in reqs:
            want, _, _ = O.process_response(P, mat, req)
            got = srv.process_response(gpu_request(req)).ciphertexts
            np.testing.assert_array_equal(got, want)


ALL_TYPES = {**{n: (v, 64) for n, v in PIR_SERVER_TEST_TYPES.items()},
             **{n: (v, 32) for n, v in PIR_SERVER_TEST_TYPES_U32.items()}}


@pytest.mark.parametrize("name", sorted(ALL_TYPES))
def test_process_pir_server_test_types(name):
    # crypto/pir/server/pir_server_test.cc:246-275 with real preprocessing,
    # for all nine typed configurations (:55-143), uint32_t and uint64_t.
    vals, coeff_bits = ALL_TYPES[name]
    P = pir_server_test_params(*vals, coeff_bits=coeff_bits)
    s = O.sizes(P)
    db = O.fill_db_u16(11, P.num_entries * s.d * s.ell, P.plaintext_modulus)
    mat = O.preprocess(P, db, public_seed=5)
    idxs = [0, 1, 10, P.num_entries - 2, P.num_entries - 1]
    with server_from_material(P, mat) as srv:
        for idx in idxs:
            req = O.client_request(P, 5, 900 + idx, idx)
            want, _, _ = O.process_response(P, mat, req)
            got = srv.process_response(gpu_request(req)).ciphertexts
            np.testing.assert_array_equal(got, want)
            msg = O.client_decrypt(P, 900 + idx, got)
            np.testing.assert_array_equal(msg.astype(np.uint16),
                                          db[idx * s.d * s.ell:(idx + 1) * s.d * s.ell])


@pytest.mark.parametrize("name,coeff_bits,mat_bytes", [
    ("Uint64Mod62", 64, 8), ("Uint64Mod0", 64, 8),                 # Mat = uint64_t
    ("Uint32Mod30", 32, 4), ("Uint32Mod0Interp1", 32, 4),          # Mat = uint32_t / int32_t
])
def test_unsigned_and_int32_pack_matrices(name, coeff_bits, mat_bytes):
    # PirServer<uint16_t, Coeff, Mat> with Mat = CoeffType (pir_server.cc:806,
    # :807) and <uint16_t, uint32_t, int32_t>: a uint32_t matrix has the bytes
    # of the condensed int32_t one; a uint64_t matrix holds the NTT outputs
    # over Z_2^64 and is reduced mod q on the way in.
    types = PIR_SERVER_TEST_TYPES if coeff_bits == 64 else PIR_SERVER_TEST_TYPES_U32
    lq, q1, q2, t, _ = types[name]
    P = pir_server_test_params(lq, q1, q2, t, mat_bytes, coeff_bits=coeff_bits)
    s = O.sizes(P)
    db = O.fill_db_u16(13, P.num_entries * s.d * s.ell, P.plaintext_modulus)
    mat = O.preprocess(P, db, public_seed=6)
    with server_from_material(P, mat) as srv:
        for idx in (0, 17, P.num_entries - 1):
            req = O.client_request(P, 6, 500 + idx, idx)
            want, _, _ = O.process_response(P, mat, req)
            got = srv.process_response(gpu_request(req)).ciphertexts
            np.testing.assert_array_equal(got, want)
            msg = O.client_decrypt(P, 500 + idx, got)
            np.testing.assert_array_equal(msg.astype(np.uint16), db[idx * s.d:(idx + 1) * s.d])


@pytest.mark.parametrize("log_d,n,t,ell", [(8, 512, 2, 1), (8, 512, 4, 2), (11, 4096, 2, 1),
                                            (12, 1024, 2, 1)])
def test_process_metrics_params_synthetic(log_d, n, t, ell):
    P = O.metrics_tool_params(n=n, log_d=log_d, t=t, ell=ell)
    mat = O.synthetic_material(P, seed=log_d * 100 + t)
    reqs = [O.synthetic_request(P, seed=i) for i in range(2)]
    _compare(P, mat, reqs)


def test_inner_product_fft_d4096_many_trials():
    # At d = 4096 the ExternalProduct chunks sit at the rounding limit, so any
    # difference in FFT arithmetic (e.g. a one-ulp twiddle) shows up here.
    rng = np.random.default_rng(4096)
    for _ in range(8):
        u = rng.integers(0, 1 << 19, size=(6, 4096), dtype=np.uint64)
        v = rng.integers(0, 1 << 52, size=(6, 4096), dtype=np.uint64)
        want, err = O.inner_product_fft(u, v, 19, 52, 19)
        got = ia.inner_product_fft(u, v, 19, 52, 19)
        np.testing.assert_array_equal(got, want)


def test_batch_equals_sequential_with_invalid_request():
    P = O.metrics_tool_params(n=2048, log_d=8, t=4, ell=2)
    mat = O.synthetic_material(P, seed=21)
    reqs = [O.synthetic_request(P, seed=100 + i) for i in range(37)]
    bad = O.Request(reqs[5].first_dim, reqs[5].second_dim, reqs[5].packing_key[:-1])
    greqs = [gpu_request(r) for r in reqs]
    greqs[5] = gpu_request(bad)
    with server_from_material(P, mat) as srv:
        batch, statuses, msg = srv.process_batch(greqs)
        assert statuses[5] == ia.INVALID_ARGUMENT
        assert "Packing key must contain 2 * num_digits polynomials." in msg
        assert batch[5] is None
        for i, r in enumerate(reqs):
            if i == 5:
                continue
            single = srv.process_response(gpu_request(r)).ciphertexts
            np.testing.assert_array_equal(batch[i].ciphertexts, single)
            if i % 6 == 0:
                want, _, _ = O.process_response(P, mat, r)
                np.testing.assert_array_equal(single, want)


@pytest.mark.parametrize("nq,wide_min,digit_major,tall,fuse,pack_stream", [
    (1, 10, 1, 1, 1, 0), (2, 10, 1, 1, 1, 0), (3, 10, 1, 1, 1, 0), (9, 10, 1, 1, 1, 0),
    (1, 1, 1, 1, 1, 0), (3, 1, 1, 1, 1, 0), (16, 10, 1, 1, 1, 0), (33, 10, 0, 1, 1, 0),
    (33, 10, 1, 1, 1, 0), (33, 10, 1, 0, 1, 0), (33, 10, 1, 1, 1, 1), (33, 10, 1, 1, 1, 2),
    (64, 10, 0, 1, 1, 0), (64, 10, 1, 1, 1, 0), (64, 10, 1, 1, 0, 0), (64, 10, 1, 0, 1, 0),
    (64, 10, 1, 1, 1, 1), (64, 10, 1, 1, 1, 2), (70, 10, 0, 1, 1, 0), (70, 10, 1, 1, 1, 0),
    (70, 10, 1, 0, 1, 0), (70, 10, 1, 1, 1, 1), (70, 10, 1, 1, 1, 2), (129, 10, 1, 1, 1, 0),
    (129, 10, 1, 1, 0, 0), (129, 10, 1, 0, 1, 0), (129, 10, 1, 1, 1, 1), (129, 10, 1, 1, 1, 2)])
def test_wide_batches_match_oracle(nq, wide_min, digit_major, tall, fuse, pack_stream,
                                   monkeypatch):
    # Every scan configuration: NarrowScan with 1, 2 and 4 tiles per group,
    # WideScan with 2, 4 and 7 tiles per wave in the query-major B layout and
    # with 7 digit tiles in the digit-major one (8 strips per workgroup, or
    # the tall kernel's 12, which here leaves the last row block a third
    # full: 1024 rows = 5 1/3 blocks), several workgroup column blocks (70 and
    # 129 queries) and partly filled groups. With the tall kernel the DB and
    # pack scans run as one launch (fuse = 1) or as two; from 22 requests on,
    # EvalPoly runs in two halves. The four-plane pack product runs the plain
    # WideScan (pack_stream = 0) or the streaming schedule with 10 strips per
    # workgroup (1; the pack matrix's 1024 rows leave the last block 2/5
    # full) or 8 (2).
    monkeypatch.setenv("PIRG_WIDE_MIN", str(wide_min))
    monkeypatch.setenv("PIRG_DIGIT_MAJOR", str(digit_major))
    monkeypatch.setenv("PIRG_TALL", str(tall))
    monkeypatch.setenv("PIRG_FUSE_SCANS", str(fuse))
    monkeypatch.setenv("PIRG_PACK_STREAM", str(pack_stream))
    P = O.metrics_tool_params(n=4096, log_d=9, t=2)
    mat = O.synthetic_material(P, seed=nq)
    reqs = [O.synthetic_request(P, seed=1000 * nq + i) for i in range(nq)]
    with server_from_material(P, mat) as srv:
        batch, statuses, _ = srv.process_batch([gpu_request(r) for r in reqs])
        if nq >= 33:
            assert srv.scans_fused(nq) == bool(fuse and digit_major and tall and not pack_stream)
    assert statuses == [0] * nq
    for i in range(0, nq, max(1, nq // 5)):
        want, _, _ = O.process_response(P, mat, reqs[i])
        np.testing.assert_array_equal(batch[i].ciphertexts, want)


@pytest.mark.parametrize("nq,ell", [(11, 5), (2, 4), (1, 5), (9, 1)])
def test_finalize_tiles_match_oracle(nq, ell):
    # FinalizeMatrixPack's inner products in tiles of 8 requests x 8 (shard,
    # i) blocks when a request has 32 or more blocks (l * t = 40 or 32 here:
    # ragged request and block tiles), and one request's group otherwise.
    # Every request is checked: a job the tiling skipped or repeated would
    # show in its own request's answer.
    P = O.metrics_tool_params(n=8192, log_d=8, t=8, ell=ell)
    mat = O.synthetic_material(P, seed=40 + nq)
    reqs = [O.synthetic_request(P, seed=4000 + 50 * ell + i) for i in range(nq)]
    with server_from_material(P, mat) as srv:
        batch, statuses, _ = srv.process_batch([gpu_request(r) for r in reqs])
    assert statuses == [0] * nq
    for i in range(nq):
        want, _, _ = O.process_response(P, mat, reqs[i])
        np.testing.assert_array_equal(batch[i].ciphertexts, want, err_msg=f"request {i}")


def test_multi_chunk_upload(monkeypatch):
    # The staged host -> HBM conversion in many small chunks (chunk boundaries
    # that split 16-row strips).
    monkeypatch.setenv("PIRG_UPLOAD_CHUNK_BYTES", "3000")
    P = O.metrics_tool_params(n=8192, log_d=8, t=4, ell=2)
    mat = O.synthetic_material(P, seed=8)
    req = O.synthetic_request(P, seed=9)
    want, _, _ = O.process_response(P, mat, req)
    with server_from_material(P, mat) as srv:
        got = srv.process_response(gpu_request(req)).ciphertexts
    np.testing.assert_array_equal(got, want)


def test_process_errors_match_reference_messages():
    # pir_server_test.cc:343-449.
    P = O.metrics_tool_params(n=2048, log_d=8, t=2)
    s = O.sizes(P)
    mat = O.synthetic_material(P, seed=1)
    req = O.synthetic_request(P, seed=2)
    cases = [
        (O.Request(req.first_dim, req.second_dim, req.packing_key[:-1]),
         "Packing key must contain 2 * num_digits polynomials."),
        (O.Request(req.first_dim[:-1], req.second_dim, req.packing_key),
         f"First dimension query size ({s.r - 1}) is invalid for DB columns ({s.K})."),
        (O.Request(np.zeros(100000, np.uint64), req.second_dim, req.packing_key),
         f"First dimension query size (100000) is invalid for DB columns ({s.K})."),
        (O.Request(req.first_dim, req.second_dim[:-1], req.packing_key),
         "second_dimension_query size mismatch. Must be 2 * num_digits."),
    ]
    with server_from_material(P, mat) as srv:
        for bad, msg in cases:
            with pytest.raises(ia.PirError) as ei:
                srv.process_response(gpu_request(bad))
            assert ei.value.code == ia.INVALID_ARGUMENT
            assert ei.value.message == msg
        # The server keeps answering after rejected requests.
        want, _, _ = O.process_response(P, mat, req)
        np.testing.assert_array_equal(srv.process_response(gpu_request(req)).ciphertexts, want)


@pytest.fixture(scope="module")
def full_size():
    """The headline shape (n = 2^20, d = 4096, t = 8) on range-faithful
    synthetic material: (params, material, server)."""
    import bench
    P = ia.metrics_tool_params(n=1 << 20, log_d=12, t=8)
    mat = bench.synthetic_material(P, seed=2026)
    srv = bench.server_from(P, mat)
    yield P, mat, srv
    srv.close()


def test_full_size_response_matches_oracle(full_size):
    """One request at the headline shape, GPU vs the oracle, bit for bit."""
    import bench
    P, mat, srv = full_size
    req = bench.synthetic_requests(P, 1, seed=7)[0]
    OP = O.Params(**{f: getattr(P, f) for f in O.Params.__dataclass_fields__})
    omat = O.Material(mat["db"], mat["pack"], mat["a_tilde"], mat["t_vec_h"],
                      mat["second_dim_a"], mat["db_cols"])
    want, _, _ = O.process_response(OP, omat, O.Request(req.first_dimension_query,
                                                        req.second_dimension_query,
                                                        req.packing_key))
    np.testing.assert_array_equal(srv.process_response(req).ciphertexts, want)


def test_full_size_batches_match_single_answers(full_size):
    """Batches of 256 (the bench's width, after autotuning) and of a ragged 200
    requests at the headline shape answer each request exactly as the
    single-request path does (which the test above pins to the oracle); the
    graph-replayed batch equals the eager one."""
    import bench
    P, _, srv = full_size
    reqs = bench.synthetic_requests(P, 256, seed=77)
    srv.autotune([256])
    singles = {i: srv.process_response(reqs[i]).ciphertexts for i in (0, 1, 131, 199, 255)}
    for n in (256, 200):
        got, st, _ = srv.process_batch(reqs[:n])
        assert st == [0] * n
        for i, want in singles.items():
            if i < n:
                np.testing.assert_array_equal(got[i].ciphertexts, want, err_msg=f"n={n} i={i}")
    srv.batch_stage(reqs)
    srv.use_graphs(True)
    try:
        for _ in range(2):
            srv.batch_run(256)
            out = srv.batch_fetch(256)
            for i, want in singles.items():
                np.testing.assert_array_equal(out[i].ciphertexts, want, err_msg=f"graph i={i}")
    finally:
        srv.use_graphs(False)


def test_autotuned_plans_match_oracle():
    # After the autotuner picks scan configurations per width, batches of
    # those widths (and of others) still answer bit for bit.
    P = O.metrics_tool_params(n=8192, log_d=9, t=4)
    mat = O.synthetic_material(P, seed=404)
    reqs = [O.synthetic_request(P, seed=4040 + i) for i in range(20)]
    with server_from_material(P, mat) as srv:
        report = srv.autotune([1, 5, 20])
        lines = report.splitlines()
        assert sum(1 for ln in lines if ln.startswith(("db ", "pack "))) == 6, report
        # Widths whose plans could share one scan launch get a timed pair line.
        assert all(ln.endswith(("-> one launch", "-> two launches"))
                   for ln in lines if ln.startswith("pair ")), report
        for n in (1, 5, 7, 20):
            got, st, _ = srv.process_batch([gpu_request(r) for r in reqs[:n]])
            assert st == [0] * n
            for i in (0, n - 1):
                want, _, _ = O.process_response(P, mat, reqs[i])
                np.testing.assert_array_equal(got[i].ciphertexts, want)
