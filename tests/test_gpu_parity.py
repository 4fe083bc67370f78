"""GPU parity: every HIP stage against the oracle on identical seeded inputs,
bit-exact (the path is integer-exact; its FFT rounds like the reference's)."""
import numpy as np
import pytest

import inspire_amd as ia
import oracle as O
from tests.helpers import (PIR_SERVER_TEST_TYPES, gpu_request, pir_server_test_params,
                           server_from_material)

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
    (1, 4, 32, 32, 20, False),
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


@pytest.mark.parametrize("name", sorted(PIR_SERVER_TEST_TYPES))
def test_process_pir_server_test_types(name):
    # crypto/pir/server/pir_server_test.cc:246-275 with real preprocessing.
    P = pir_server_test_params(*PIR_SERVER_TEST_TYPES[name])
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


@pytest.mark.parametrize("log_d,n,t,ell", [(8, 512, 2, 1), (8, 512, 4, 2), (11, 4096, 2, 1),
                                            (12, 1024, 2, 1)])
def test_process_metrics_params_synthetic(log_d, n, t, ell):
    P = O.metrics_tool_params(n=n, log_d=log_d, t=t, ell=ell)
    mat = O.synthetic_material(P, seed=log_d * 100 + t)
    reqs = [O.synthetic_request(P, seed=i) for i in range(2)]
    _compare(P, mat, reqs)
