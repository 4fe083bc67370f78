"""Servers built from the reference's protobuf forms of the preprocessed data
(crypto/pir/server/pir_preprocessed_data.h:52-320): the single-message file
and the metadata + shard directory, written here with the protobuf runtime
from the reference's schema, answer exactly as the oracle does."""
import numpy as np
import pytest

import inspire_amd as ia
import oracle as O
import pir_proto
from helpers import (PIR_SERVER_TEST_TYPES, gpu_request, pir_server_test_params,
                     to_gpu_params)

pytestmark = pytest.mark.gpu

CONFIGS = {
    "metrics_t2": lambda: O.metrics_tool_params(n=4096, log_d=9, t=2),
    "metrics_ell2": lambda: O.metrics_tool_params(n=2048, log_d=8, t=2, ell=2),
    "metrics_t1": lambda: O.metrics_tool_params(n=1024, log_d=9, t=1),
    "Uint64Mod0": lambda: pir_server_test_params(*PIR_SERVER_TEST_TYPES["Uint64Mod0"][:3], 4,
                                                 PIR_SERVER_TEST_TYPES["Uint64Mod0"][4]),
}


def check(P, mat, srv):
    reqs = [O.synthetic_request(P, seed=70 + i) for i in range(3)]
    got, st, _ = srv.process_batch([gpu_request(r) for r in reqs])
    assert st == [0] * len(reqs)
    for i, r in enumerate(reqs):
        want, _, _ = O.process_response(P, mat, r)
        np.testing.assert_array_equal(got[i].ciphertexts, want)


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_server_from_proto_file_matches_oracle(name):
    P = CONFIGS[name]()
    mat = O.synthetic_material(P, seed=17)
    with ia.PirServer.create_from_proto(to_gpu_params(P), pir_proto.serialize_data(P, mat)) as s:
        check(P, mat, s)


@pytest.mark.parametrize("name", ["metrics_ell2", "Uint64Mod0"])
def test_server_from_directory_matches_oracle(name, tmp_path):
    P = CONFIGS[name]()
    mat = O.synthetic_material(P, seed=18)
    pir_proto.save_directory(P, mat, str(tmp_path))
    with ia.PirServer.create_from_directory(to_gpu_params(P), str(tmp_path)) as s:
        check(P, mat, s)


def test_decryptable_answer_from_proto():
    P = pir_server_test_params(*PIR_SERVER_TEST_TYPES["Uint64Mod62"])
    s = O.sizes(P)
    db = O.fill_db_u16(19, P.num_entries * s.d * s.ell, P.plaintext_modulus)
    mat = O.preprocess(P, db, public_seed=4)
    idx = 9
    with ia.PirServer.create_from_proto(to_gpu_params(P), pir_proto.serialize_data(P, mat)) as sv:
        req = O.client_request(P, 4, 400 + idx, idx)
        got = sv.process_response(gpu_request(req)).ciphertexts
    msg = O.client_decrypt(P, 400 + idx, got)
    np.testing.assert_array_equal(msg.astype(np.uint16),
                                  db[idx * s.d * s.ell:(idx + 1) * s.d * s.ell])
