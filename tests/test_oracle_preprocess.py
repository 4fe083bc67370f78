"""Pins the oracle's offline preprocessing beyond the end-to-end decrypt tests:
ComputeHintMatrix against the reference test's naive negacyclic product
(crypto/hint_computation_test.cc:38-64, 150-231) and its error messages
(:233-268), and the a-component sampler the GPU tests share with
oracle.preprocess."""
import numpy as np
import pytest

import oracle as O


def naive_hint(db: np.ndarray, a: np.ndarray) -> np.ndarray:
    """NaiveComputeHintMatrix: sum over k of db[:, k*d:(k+1)*d] @ M_k, where
    M_k[i][j] = a_k[i - j] for i >= j and -a_k[d + i - j] otherwise, mod 2^64."""
    n, d = a.shape
    rows = db.shape[0]
    out = np.zeros((rows, d), dtype=np.uint64)
    idx = np.arange(d)
    diff = idx[:, None] - idx[None, :]
    for k in range(n):
        ak = a[k]
        m = np.where(diff >= 0, ak[diff % d], np.uint64(0) - ak[(d + diff) % d])
        sub = db[:, k * d:(k + 1) * d].astype(np.uint64)
        for i in range(rows):
            # uint64 products and sums wrap mod 2^64.
            out[i] += (sub[i][:, None] * m).sum(axis=0, dtype=np.uint64)
    return out


@pytest.mark.parametrize("db_bits", [16, 8])
def test_hint_matrix_matches_naive(db_bits):
    rng = np.random.default_rng(db_bits)
    for log_d in range(1, 5):
        d = 1 << log_d
        for n in (1, 2, 4):
            for rows in (1, 2, d, d + 5):
                db = rng.integers(0, 1 << db_bits, size=(rows, d * n), dtype=np.uint64)
                a = rng.integers(0, 1 << 64, size=(n, d), dtype=np.uint64)
                got = O.compute_hint_matrix(db.astype(np.uint16), db_bits, a)
                np.testing.assert_array_equal(got, naive_hint(db, a))


def test_hint_matrix_larger_degree_matches_naive():
    # d = 256: several InnerProductFft batches per row (safe_n = 2^(53-8-40)).
    rng = np.random.default_rng(256)
    d, n, rows = 256, 40, 3
    db = rng.integers(0, 1 << 16, size=(rows, d * n), dtype=np.uint64)
    a = rng.integers(0, 1 << 64, size=(n, d), dtype=np.uint64)
    got = O.compute_hint_matrix(db.astype(np.uint16), 16, a)
    np.testing.assert_array_equal(got, naive_hint(db, a))


def test_hint_matrix_errors():
    a = np.zeros((1, 4), dtype=np.uint64)
    with pytest.raises(O.OracleError, match="Database matrix cannot be empty."):
        O.compute_hint_matrix(np.zeros((0, 4), dtype=np.uint16), 16, a)
    with pytest.raises(O.OracleError, match="a_components cannot be empty."):
        O.compute_hint_matrix(np.zeros((2, 4), dtype=np.uint16), 16,
                              np.zeros((0, 4), dtype=np.uint64))
    with pytest.raises(O.OracleError, match="Database columns must match d"):
        O.compute_hint_matrix(np.zeros((2, 3), dtype=np.uint16), 16, a)


def test_sampled_components_are_the_preprocess_ones():
    P = O.metrics_tool_params(n=256, log_d=6, t=4)
    s = O.sizes(P)
    db = O.fill_db_u16(3, P.num_entries * s.d * s.ell, P.plaintext_modulus)
    mat = O.preprocess(P, db, public_seed=77)
    first, second, pack_a = O.sample_a_components(P, 77)
    assert first.shape == (s.S, s.d) and pack_a.shape == (2 * s.k, s.d)
    np.testing.assert_array_equal(second, mat.second_dim_a)
    assert int(first.max()) < (1 << P.log_q)
