"""GPU bit-packed formats against the oracle, byte for byte:
Polynomial::ToProto / CreateFromProto (crypto/polynomial.cc:562-590,
208-252) on the host ABI and on device memory, and LoadFromBuffer
(crypto/pir/server/pir_server.cc:594-793): a server loaded from the
SaveToFile image answers exactly as the oracle does on the original
material."""
import numpy as np
import pytest

import inspire_amd as ia
import oracle as O
from helpers import (PIR_SERVER_TEST_TYPES, gpu_request, pir_server_test_params,
                     server_from_material, to_gpu_params)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("bits", [1, 2, 7, 8, 13, 19, 30, 31, 32, 33, 52, 56, 63, 64])
@pytest.mark.parametrize("length,npolys", [(1, 3), (2, 5), (4, 1), (16, 7), (4096, 3)])
def test_to_and_from_proto_match_oracle(bits, length, npolys):
    x = O.fill_u64(1000 * bits + length, npolys * length).reshape(npolys, length)
    got = ia.poly_to_proto(x, bits)
    for p in range(npolys):
        assert got[p] == O.poly_to_proto(x[p], bits), (p, bits, length)
    back = ia.poly_from_proto(got, length, bits)
    for p in range(npolys):
        np.testing.assert_array_equal(back[p], O.poly_from_proto(got[p], length, bits))


@pytest.mark.parametrize("length,logm,nbytes", [(3, 8, 3), (4, 0, 4), (4, 8, 0), (4, 8, 5)])
def test_from_proto_errors_equal_oracle(length, logm, nbytes):
    with pytest.raises(O.OracleError) as eo:
        O.poly_from_proto(bytes(nbytes), length, logm)
    with pytest.raises(ia.PirError) as eg:
        ia.poly_from_proto([bytes(nbytes)], length, logm)
    assert (eg.value.code, eg.value.message) == (eo.value.code, eo.value.message)


def test_device_forms_match_host_forms():
    from inspire_amd.hiprt import DeviceBuffer, synchronize
    length, npolys, bits = 4096, 6, 52
    x = O.fill_u64(77, npolys * length).reshape(npolys, length)
    pitch = (length * bits + 63) // 64 + 3  # a pitch above the stream length
    with DeviceBuffer(x.nbytes) as dx, DeviceBuffer(npolys * pitch * 8) as dw, \
            DeviceBuffer(x.nbytes) as dy:
        dx.upload(x)
        dw.fill(0xFF)
        ia.poly_to_proto_device(dx.addr, npolys, length, bits, dw.addr, pitch)
        ia.poly_from_proto_device(dw.addr, pitch, npolys, length, bits, dy.addr)
        synchronize()
        words = dw.download(np.uint64, npolys * pitch).reshape(npolys, pitch)
        back = dy.download(np.uint64, npolys * length).reshape(npolys, length)
    nb = (length * bits + 7) // 8
    for p in range(npolys):
        assert words[p].tobytes()[:nb] == O.poly_to_proto(x[p], bits)
        assert not words[p, (length * bits + 63) // 64:].any()  # padding written as zero
    np.testing.assert_array_equal(back, x & np.uint64((1 << bits) - 1))


def test_full_size_query_round_trip():
    # The n = 2^20, d = 4096, t = 8 first-dimension query: 131072 values mod 2^52.
    x = O.fill_u64(5, 131072, (1 << 52) - 1).reshape(1, -1)
    enc = ia.poly_to_proto(x, 52)
    assert enc[0] == O.poly_to_proto(x[0], 52)
    np.testing.assert_array_equal(ia.poly_from_proto(enc, 131072, 52), x)


def _configs():
    return {
        "metrics_t2": O.metrics_tool_params(n=4096, log_d=9, t=2),
        "metrics_ell2": O.metrics_tool_params(n=2048, log_d=8, t=2, ell=2),
        "metrics_t1": O.metrics_tool_params(n=1024, log_d=9, t=1),
        "Uint64Mod0": pir_server_test_params(*PIR_SERVER_TEST_TYPES["Uint64Mod0"][:3], 4,
                                             PIR_SERVER_TEST_TYPES["Uint64Mod0"][4]),
        "Uint64Mod62": pir_server_test_params(*PIR_SERVER_TEST_TYPES["Uint64Mod62"][:3], 4,
                                              PIR_SERVER_TEST_TYPES["Uint64Mod62"][4]),
    }


@pytest.mark.parametrize("name", sorted(_configs()))
def test_server_loaded_from_image_matches_oracle(name, monkeypatch):
    P = _configs()[name]
    mat = O.synthetic_material(P, seed=31)
    img = O.save_server(P, mat)
    # Small staging chunks: several row blocks of the bit-packed pack matrix.
    monkeypatch.setenv("PIRG_UPLOAD_CHUNK_BYTES", str(1 << 16))
    reqs = [O.synthetic_request(P, seed=310 + i) for i in range(3)]
    with ia.PirServer.load_from_buffer(to_gpu_params(P), img) as srv, \
            server_from_material(P, mat) as ref:
        got, st, _ = srv.process_batch([gpu_request(r) for r in reqs])
        want_gpu, _, _ = ref.process_batch([gpu_request(r) for r in reqs])
    assert st == [0] * len(reqs)
    for i, r in enumerate(reqs):
        want, _, _ = O.process_response(P, mat, r)
        np.testing.assert_array_equal(got[i].ciphertexts, want)
        np.testing.assert_array_equal(want_gpu[i].ciphertexts, want)


@pytest.mark.parametrize("name", ["Uint64Mod0", "Uint64Mod62"])
def test_decryptable_answer_from_loaded_image(name):
    # pir_server_test.cc:623-668: preprocess, save, load, answer, decrypt.
    P = pir_server_test_params(*PIR_SERVER_TEST_TYPES[name])
    s = O.sizes(P)
    db = O.fill_db_u16(13, P.num_entries * s.d * s.ell, P.plaintext_modulus)
    mat = O.preprocess(P, db, public_seed=7)
    img = O.save_server(P, mat)
    idx = 5
    with ia.PirServer.load_from_buffer(to_gpu_params(P), img) as srv:
        req = O.client_request(P, 7, 700 + idx, idx)
        got = srv.process_response(gpu_request(req)).ciphertexts
    want, _, _ = O.process_response(P, mat, req)
    np.testing.assert_array_equal(got, want)
    msg = O.client_decrypt(P, 700 + idx, got)
    np.testing.assert_array_equal(msg.astype(np.uint16),
                                  db[idx * s.d * s.ell:(idx + 1) * s.d * s.ell])


def test_bench_image_writer_matches_oracle_savetofile():
    # The bench writes the full-size image with the GPU encoder; at a small
    # size its bytes must equal the oracle's SaveToFile restatement.
    import bench
    P = ia.metrics_tool_params(n=4096, log_d=9, t=2)
    mat = bench.synthetic_material(P, seed=3)
    img = bench.server_image(P, mat)
    OP = O.Params(**{f: getattr(P, f) for f in O.Params.__dataclass_fields__})
    om = O.Material(mat["db"], mat["pack"], mat["a_tilde"], mat["t_vec_h"],
                    mat["second_dim_a"], mat["db_cols"])
    assert img == O.save_server(OP, om)
