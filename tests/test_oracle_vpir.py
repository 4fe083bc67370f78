"""The oracle's vPIR online path (crypto/vpir/server/vpir_server.cc:101-287),
pinned by the reference's own invariant -- ProcessResponse equals
ProcessResponseWithMatMul (vpir_server_test.cc:326-345), i.e. the FFT
FinalizeMatrixPack equals the exact negacyclic H' product -- and, for local
finalize, by LowBits(DB q + pack y_g) computed here with Python integers."""
import numpy as np
import pytest

import oracle as O
from vpir_helpers import VPIR_TEST_TYPES, request, vpir_params


@pytest.mark.parametrize("name", sorted(VPIR_TEST_TYPES))
@pytest.mark.parametrize("ell", [1, 2])
@pytest.mark.parametrize("local", [False, True])
def test_fft_finalize_equals_h_prime_product(name, ell, local):
    P = vpir_params(ell=ell, **VPIR_TEST_TYPES[name])
    mat = O.synthetic_material(P, seed=3 + ell)
    for seed in (1, 2, 3):
        fd, pk = request(P, seed, local)
        a = O.vpir_process(P, local, mat.db, mat.db_cols, mat.pack, mat.a_tilde, mat.t_vec_h,
                           fd, pk)
        b = O.vpir_process(P, local, mat.db, mat.db_cols, mat.pack, mat.a_tilde, mat.t_vec_h,
                           fd, pk, matmul=True)
        np.testing.assert_array_equal(a, b)


def test_local_finalize_closed_form():
    P = vpir_params(log_q=40, mat_bytes=4, ell=2, log_d=3, n=32, pack_num_digits=3)
    s = O.sizes(P)
    rng = np.random.default_rng(0)
    db = rng.integers(0, 257, size=(s.M, P.num_entries)).astype(np.uint16)
    pack = rng.integers(-(1 << 20), 1 << 20, size=(s.M, s.k * s.d)).astype(np.int32)
    fd, pk = request(P, 9, True)
    got = O.vpir_process(P, True, O.condense_u16(db), P.num_entries, O.condense_i32(pack),
                         np.zeros((s.ell, s.d), np.uint64),
                         np.zeros((s.ell, s.k, s.d), np.uint64), fd, pk)
    y = [int(v) for v in pk.reshape(-1)]
    q = [int(v) for v in fd]
    for row in range(s.M):
        want = (sum(int(a) * b for a, b in zip(db[row], q)) +
                sum(int(a) * b for a, b in zip(pack[row], y))) % (1 << P.log_q)
        assert int(got.reshape(-1)[row]) == want


@pytest.mark.parametrize("local,nk,nfd,msg", [
    (False, 3, 32, "Packing key must contain {} polynomials."),
    (True, 2, 31, "Packing key must contain {} polynomials."),
    (False, None, 31, "Flattened first dimension query must match DB matrix columns (N)."),
])
def test_request_errors(local, nk, nfd, msg):
    P = vpir_params(log_q=62, mat_bytes=4)
    s = O.sizes(P)
    mat = O.synthetic_material(P, seed=1)
    fd, pk = request(P, 1, local)
    if nk is not None:
        pk = np.zeros((nk, s.d), np.uint64)
    fd = np.zeros(nfd * s.d, np.uint64)
    with pytest.raises(O.OracleError) as ei:
        O.vpir_process(P, local, mat.db, mat.db_cols, mat.pack, mat.a_tilde, mat.t_vec_h, fd, pk)
    want = msg.format(s.k if local else 2 * s.k)
    assert ei.value.code == 3 and ei.value.message == want
