"""The plain (n = 4096, d = 2048, t = 2) and batched (n = 65536, d = 2048,
t = 4, 64 queries in one launch) workloads end to end with real material:
seeded DB, GPU preprocessing (checked against the oracle's), client queries
from the oracle's restatement of the reference client, answers bit-identical
to the oracle's and decrypting to the records."""
import numpy as np
import pytest

import inspire_amd as ia
import oracle as O
from helpers import gpu_request, to_gpu_params

pytestmark = pytest.mark.gpu


def _material(P, seed):
    s = O.sizes(P)
    db = O.fill_db_u16(seed, P.num_entries * s.d * s.ell, P.plaintext_modulus)
    first, second, pack_a = O.sample_a_components(P, seed)
    got = ia.preprocess(to_gpu_params(P), db, first, pack_a)
    mat = O.Material(got["db"], got["pack"], got["a_tilde"], got["t_vec_h"], second,
                     got["db_cols"])
    return db, mat


def _record(P, db, idx):
    s = O.sizes(P)
    return db[idx * s.d * s.ell:(idx + 1) * s.d * s.ell]


def test_plain_query_config():
    P = O.metrics_tool_params(n=4096, log_d=11, t=2)
    db, mat = _material(P, 101)
    want_mat = O.preprocess(P, db, public_seed=101)
    np.testing.assert_array_equal(mat.db, want_mat.db)
    np.testing.assert_array_equal(mat.pack, want_mat.pack)
    np.testing.assert_array_equal(mat.a_tilde, want_mat.a_tilde)
    np.testing.assert_array_equal(mat.t_vec_h, want_mat.t_vec_h)
    s = O.sizes(P)
    with ia.PirServer.create_flat(to_gpu_params(P), mat.db, mat.db_cols, mat.pack, mat.a_tilde,
                                  mat.t_vec_h, mat.second_dim_a) as srv:
        for idx in (0, 1234, P.num_entries - 1):
            req = O.client_request(P, 101, 7000 + idx, idx)
            assert req.first_dim.size == s.r
            got = srv.process_response(gpu_request(req)).ciphertexts
            want, _, _ = O.process_response(P, mat, req)
            np.testing.assert_array_equal(got, want)
            np.testing.assert_array_equal(O.client_decrypt(P, 7000 + idx, got).astype(np.uint16),
                                          _record(P, db, idx))


def test_batched_64_queries_config():
    # 64 client queries in one launch, bit-identical to 64 sequential
    # answers of the oracle. Decryption: with the metrics tool's parameters
    # (q2 = 2^18 over p = 65535, 19-bit gadgets) the noise occasionally
    # crosses the rounding boundary at d = 2048 (here 1 coefficient in
    # 131,072 decrypts off by one), so the check bounds those events instead
    # of requiring none.
    P = O.metrics_tool_params(n=65536, log_d=11, t=4)
    db, mat = _material(P, 202)
    idxs = [int(i) for i in np.random.default_rng(202).integers(0, P.num_entries, size=64)]
    reqs = [O.client_request(P, 202, 9000 + j, idx) for j, idx in enumerate(idxs)]
    with ia.PirServer.create_flat(to_gpu_params(P), mat.db, mat.db_cols, mat.pack, mat.a_tilde,
                                  mat.t_vec_h, mat.second_dim_a) as srv:
        batch, st, _ = srv.process_batch([gpu_request(r) for r in reqs])
        assert st == [0] * 64
        for j in (0, 17, 63):
            alone = srv.process_response(gpu_request(reqs[j])).ciphertexts
            np.testing.assert_array_equal(batch[j].ciphertexts, alone)
    off, total = 0, 0
    for j, idx in enumerate(idxs):
        want, _, _ = O.process_response(P, mat, reqs[j])
        np.testing.assert_array_equal(batch[j].ciphertexts, want)
        msg = O.client_decrypt(P, 9000 + j, batch[j].ciphertexts).astype(np.int64)
        diff = np.abs(msg - _record(P, db, idx).astype(np.int64))
        diff = np.minimum(diff, P.plaintext_modulus - diff)
        assert int(diff.max()) <= 1
        off += int(np.count_nonzero(diff))
        total += diff.size
    assert off * 10000 <= total, (off, total)
