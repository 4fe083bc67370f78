"""PirServer::SaveToFile / LoadFromFile (crypto/pir/server/pir_server.cc:433-592,
:794-804) on the GPU path, against the oracle's restatement of the writer.

The GPU writes the image from the resident server (DB words and pack entries
rebuilt from their HBM tiles, everything bit-packed on the GPU); it must equal
the oracle's SaveToFile bytes for the same material, for every typed
configuration of pir_server_test.cc, ell = 2 and a uint64_t pack matrix. A
file round trip (save, LoadFromFile) must then answer as the oracle does.
"""
import numpy as np
import pytest

import inspire_amd as ia
import oracle as O
from helpers import (PIR_SERVER_TEST_TYPES, PIR_SERVER_TEST_TYPES_U32, gpu_request,
                     pir_server_test_params, server_from_material, to_gpu_params)

pytestmark = pytest.mark.gpu

CASES = ([(n, v, 64, 1, None) for n, v in sorted(PIR_SERVER_TEST_TYPES.items())] +
         [(n, v, 32, 1, None) for n, v in sorted(PIR_SERVER_TEST_TYPES_U32.items())] +
         [("Uint64Mod62", PIR_SERVER_TEST_TYPES["Uint64Mod62"], 64, 2, None),
          ("Uint64Mod0Interp1", PIR_SERVER_TEST_TYPES["Uint64Mod0Interp1"], 64, 2, None),
          ("Uint64Mod62", PIR_SERVER_TEST_TYPES["Uint64Mod62"], 64, 1, 8)])


def _params(vals, coeff_bits, ell, mat_bytes):
    lq, q1, q2, t, mb = vals
    return pir_server_test_params(lq, q1, q2, t, mat_bytes or mb, ell=ell,
                                  coeff_bits=coeff_bits)


@pytest.mark.parametrize("name,vals,coeff_bits,ell,mat_bytes", CASES,
                         ids=[f"{c[0]}-c{c[2]}-l{c[3]}-m{c[4] or c[1][4]}" for c in CASES])
def test_image_equals_oracle_and_round_trips(tmp_path, name, vals, coeff_bits, ell, mat_bytes):
    P = _params(vals, coeff_bits, ell, mat_bytes)
    s = O.sizes(P)
    db = O.fill_db_u16(23, P.num_entries * s.d * s.ell, P.plaintext_modulus)
    mat = O.preprocess(P, db, public_seed=4)
    want_image = O.save_server(P, mat)
    path = str(tmp_path / "server.bin")
    with server_from_material(P, mat) as srv:
        assert srv.image_size() == len(want_image)
        got_image = srv.save_to_buffer()
        assert got_image.tobytes() == want_image
        srv.save_to_file(path)
    with open(path, "rb") as f:
        assert f.read() == want_image
    with ia.PirServer.load_from_file(to_gpu_params(P), path) as srv2:
        for idx in (0, P.num_entries - 1):
            req = O.client_request(P, 4, 60 + idx, idx)
            want, _, _ = O.process_response(P, mat, req)
            got = srv2.process_response(gpu_request(req)).ciphertexts
            np.testing.assert_array_equal(got, want)
        # A loaded server writes its image back unchanged.
        assert srv2.save_to_buffer().tobytes() == want_image


def test_image_of_a_loaded_metrics_tool_server(tmp_path):
    # The pir_metrics_tool parameters (d = 2048, q = 2^52, t = 4, ell = 2):
    # 31-bit pack entries straddle bytes and words.
    P = O.metrics_tool_params(n=2048, log_d=11, t=4, ell=2)
    mat = O.synthetic_material(P, 8)
    image = O.save_server(P, mat)
    with ia.PirServer.load_from_buffer(to_gpu_params(P), image) as srv:
        assert srv.save_to_buffer().tobytes() == image
        req = O.synthetic_request(P, 3)
        want, _, _ = O.process_response(P, mat, req)
        np.testing.assert_array_equal(srv.process_response(gpu_request(req)).ciphertexts, want)
