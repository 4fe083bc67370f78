"""The oracle's bit-packed formats, pinned without the reference:
Polynomial::ToProto / CreateFromProto (crypto/polynomial.cc:562-590,
208-252), BitPack*/BitUnpack* (crypto/pir/server/pir_server.cc:309-429) and
the SaveToFile / LoadFromBuffer image (pir_server.cc:433-793).

The reference's tests hold no serialized fixtures for these (pir_server_test
writes and reads a temporary file), so the checks are: hand-computed bytes,
the cases that reduce to plain byte copies (8, 16 and 64 bits per value), an
independent bit-string construction on small inputs, round trips, the
reference's error messages, and -- for the image -- that a server rebuilt
from it answers exactly as the original material does."""
import numpy as np
import pytest

import oracle as O


def bitstring_pack(values, bits):
    # Independent formulation: the LSB-first bit string, then bytes LSB-first.
    s = "".join(format(int(v) & ((1 << bits) - 1), f"0{bits}b")[::-1] for v in values)
    s += "0" * (-len(s) % 8)
    return bytes(int(s[i:i + 8][::-1], 2) for i in range(0, len(s), 8))


def test_hand_computed_bytes():
    # 1 | 2 << 2 | 3 << 4 | 0 << 6 = 0x39
    assert O.poly_to_proto([1, 2, 3, 0], 2) == bytes([0x39])
    assert O.poly_to_proto([0xFF, 0x01], 8) == bytes([0xFF, 0x01])
    # 3-bit values 5, 6, 7: bits 101 011 111 (LSB first) -> 0b11110101, 0b1
    assert O.poly_to_proto([5, 6, 7, 0], 3) == bytes([0xF5, 0x01])


@pytest.mark.parametrize("bits,dtype", [(8, np.uint8), (16, np.uint16), (32, np.uint32),
                                        (64, np.uint64)])
def test_whole_byte_widths_are_little_endian_copies(bits, dtype):
    x = O.fill_u64(7, 64)
    assert O.poly_to_proto(x, bits) == x.astype(dtype).tobytes()


@pytest.mark.parametrize("bits", [1, 2, 3, 5, 7, 13, 19, 30, 31, 33, 52, 56, 63])
def test_matches_bit_string_and_round_trips(bits):
    x = O.fill_u64(bits, 32)
    enc = O.poly_to_proto(x, bits)
    assert enc == bitstring_pack(x, bits)
    assert len(enc) == (32 * bits + 7) // 8
    mask = np.uint64((1 << bits) - 1)
    np.testing.assert_array_equal(O.poly_from_proto(enc, 32, bits), x & mask)


@pytest.mark.parametrize("length,logm,nbytes,msg", [
    (3, 8, 3, "Invalid length."),
    (0, 8, 1, "Invalid length."),
    (4, 0, 4, "Invalid log_modulus."),
    (4, 65, 4, "Invalid log_modulus."),
    (4, 8, 0, "Empty serialized polynomial."),
    (4, 8, 5, "Unexpected serialized size."),
])
def test_from_proto_errors_match_reference(length, logm, nbytes, msg):
    with pytest.raises(O.OracleError) as ei:
        O.poly_from_proto(bytes(nbytes), length, logm)
    assert ei.value.code == 3 and ei.value.message == msg


def test_signed_pack_sign_extends():
    v = np.array([-1, 5, -(1 << 30), (1 << 30) - 1, 0, -7], dtype=np.int64)
    enc = O.bit_pack(v, 31, is_signed=True, value_bits=32)
    # The field holds the low 31 bits of the int32 pattern.
    assert enc == bitstring_pack(v.astype(np.int32).view(np.uint32), 31)
    np.testing.assert_array_equal(O.bit_unpack(enc, v.size, 31, True, 32), v)
    # int16 entries at 12 bits.
    w = np.array([-2048, 2047, -1, 0, 17], dtype=np.int64)
    enc = O.bit_pack(w, 12, is_signed=True, value_bits=16)
    np.testing.assert_array_equal(O.bit_unpack(enc, w.size, 12, True, 16), w)


@pytest.mark.parametrize("cfg", ["metrics", "int16_mod0", "ell2"])
def test_server_image_round_trip_answers_identically(cfg):
    from tests.helpers import pir_server_test_params
    if cfg == "metrics":
        P = O.metrics_tool_params(n=4096, log_d=9, t=2)
    elif cfg == "int16_mod0":
        P = pir_server_test_params(64, 61, 60, 4, 2)
    else:
        P = O.metrics_tool_params(n=2048, log_d=8, t=2, ell=2)
    mat = O.synthetic_material(P, seed=5)
    img = O.save_server(P, mat)
    back = O.load_server(P, img)
    np.testing.assert_array_equal(back.db, mat.db)
    np.testing.assert_array_equal(back.pack, mat.pack)
    np.testing.assert_array_equal(back.a_tilde, mat.a_tilde)
    logm = min(P.log_q, 64)
    mask = np.uint64((1 << logm) - 1) if logm < 64 else np.uint64((1 << 64) - 1)
    np.testing.assert_array_equal(back.t_vec_h, mat.t_vec_h & mask)
    req = O.synthetic_request(P, seed=9)
    a, _, _ = O.process_response(P, mat, req)
    b, _, _ = O.process_response(P, back, req)
    np.testing.assert_array_equal(a, b)
