"""vPIR online path on the GPU (crypto/vpir/server/vpir_server.cc:45-195)
against the oracle, bit for bit: the reference's test types (d = 32), the
metrics tool's online parameters (q = 2^32, uint8 DB, gadget (8, 4)) at
d = 2048, local and FFT finalize, batches, the device-resident batch path
and the reference's error messages."""
import numpy as np
import pytest

import inspire_amd as ia
import oracle as O
from helpers import to_gpu_params
from vpir_helpers import VPIR_TEST_TYPES, metrics_like, request, vpir_params

pytestmark = pytest.mark.gpu


def make(P, local, seed):
    mat = O.synthetic_material(P, seed=seed)
    s = O.sizes(P)
    srv = ia.VpirServer.create(to_gpu_params(P), local, mat.db, s.M, mat.db_cols, mat.pack,
                               mat.a_tilde, mat.t_vec_h)
    return srv, mat


def oracle_answer(P, local, mat, fd, pk):
    return O.vpir_process(P, local, mat.db, mat.db_cols, mat.pack, mat.a_tilde, mat.t_vec_h,
                          fd, pk)


CONFIGS = {
    "Uint64Mod0_ell1": lambda: vpir_params(ell=1, **VPIR_TEST_TYPES["Uint64Mod0"]),
    "Uint64Mod62_ell2": lambda: vpir_params(ell=2, **VPIR_TEST_TYPES["Uint64Mod62"]),
    "q52_d256": lambda: vpir_params(52, 4, ell=2, log_d=8, n=2048, pack_log_digit=19,
                                    pack_num_digits=2),
    "metrics_online_d2048": lambda: metrics_like(ell=4),
}


@pytest.mark.parametrize("local", [False, True])
@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_vpir_answers_match_oracle(name, local):
    P = CONFIGS[name]()
    srv, mat = make(P, local, seed=21)
    with srv:
        reqs = [request(P, 100 + i, local) for i in range(5)]
        got, st, _ = srv.process_batch(reqs)
        one = srv.process_response(*reqs[0])
    assert st == [0] * len(reqs)
    np.testing.assert_array_equal(one, got[0])
    for i, (fd, pk) in enumerate(reqs):
        np.testing.assert_array_equal(got[i], oracle_answer(P, local, mat, fd, pk),
                                      err_msg=f"request {i}")


@pytest.mark.parametrize("cfg,nq,pack_stream", [
    ("u8", 1, -1), ("u8", 20, -1), ("u8", 70, 0), ("u8", 70, 1), ("u8", 70, 2), ("u16", 40, -1),
    ("u16", 129, 0), ("u16", 129, 1), ("u16", 129, 2), ("u16", 200, -1)])
def test_vpir_device_batches_and_replays(cfg, nq, pack_stream, monkeypatch):
    # log q = 32: 4-digit queries, digit-major wide scans above 32 queries
    # (uint8 DB: 1 plane, uint16: 2 planes, int32 pack: 4 planes). The pack
    # product runs the plain digit-major WideScan (pack_stream = 0) or the
    # streaming scans with two query groups per wave and 6 (1) or 8 (2)
    # strips per workgroup (70 queries fill 5 of a workgroup's 8 groups, 129
    # queries 9 of two workgroups' 16); -1 keeps the default plan (from 97
    # queries on the 8-strip streaming scan).
    if pack_stream >= 0:
        monkeypatch.setenv("PIRG_PACK_STREAM", str(pack_stream))
    if cfg == "u8":
        P = metrics_like(ell=2)
    else:
        P = vpir_params(32, 4, ell=2, log_d=8, n=2048, pack_log_digit=8, pack_num_digits=4)
    srv, mat = make(P, False, seed=5)
    reqs = [request(P, 500 + i, False) for i in range(nq)]
    with srv:
        srv.batch_stage(reqs)
        srv.batch_run(nq)
        first = srv.batch_fetch(nq)
        srv.batch_run(nq)  # graph replay
        again = srv.batch_fetch(nq)
    for i in range(nq):
        np.testing.assert_array_equal(again[i], first[i])
    for i in (0, nq - 1):
        np.testing.assert_array_equal(first[i], oracle_answer(P, False, mat, *reqs[i]))


@pytest.mark.parametrize("local", [False, True])
def test_vpir_many_shards_match_oracle(local):
    # 37 shards: FinalizeMatrixPack's inner products run in tiles of 8
    # requests x 8 shards (ragged in both), 11 requests.
    P = vpir_params(32, 4, ell=37, log_d=8, n=2048, pack_log_digit=8, pack_num_digits=4,
                    db_bytes=1, p=256)
    srv, mat = make(P, local, seed=31)
    reqs = [request(P, 700 + i, local) for i in range(11)]
    with srv:
        got, st, _ = srv.process_batch(reqs)
    assert st == [0] * len(reqs)
    for i in range(len(reqs)):
        np.testing.assert_array_equal(got[i], oracle_answer(P, local, mat, *reqs[i]),
                                      err_msg=f"request {i}")


def test_vpir_errors_match_oracle():
    P = vpir_params(ell=2, **VPIR_TEST_TYPES["Uint64Mod62"])
    s = O.sizes(P)
    srv, mat = make(P, False, seed=2)
    good = request(P, 1, False)
    bad = [(good[0], good[1][:3]),                                  # packing key count
           (good[0][:-s.d], good[1])]                               # first dimension size
    with srv:
        got, st, msg = srv.process_batch([bad[0], good, bad[1]])
        assert st[1] == 0 and st[0] == ia.INVALID_ARGUMENT and st[2] == ia.INVALID_ARGUMENT
        np.testing.assert_array_equal(got[1], oracle_answer(P, False, mat, *good))
        for fd, pk in bad:
            with pytest.raises(ia.PirError) as eg:
                srv.process_response(fd, pk)
            with pytest.raises(O.OracleError) as eo:
                oracle_answer(P, False, mat, fd, pk)
            assert eg.value.message == eo.value.message


def test_vpir_create_errors():
    P = vpir_params(ell=2, **VPIR_TEST_TYPES["Uint64Mod62"])
    s = O.sizes(P)
    mat = O.synthetic_material(P, seed=3)
    gp = to_gpu_params(P)
    for args, msg in [((s.M - 32, mat.db_cols), "DB matrix must have ell * d rows."),
                      ((s.M, mat.db_cols - 4), "DB matrix must have N columns.")]:
        with pytest.raises(ia.PirError) as ei:
            ia.VpirServer.create(gp, False, mat.db, args[0], args[1], mat.pack, mat.a_tilde,
                                 mat.t_vec_h)
        assert ei.value.message == msg
