"""Shared builders for the parity tests: the same inputs go to the oracle and
to the GPU library."""
import numpy as np

import inspire_amd as ia
import oracle as O


def to_gpu_params(P: O.Params) -> ia.PirParams:
    return ia.PirParams(**{f: getattr(P, f) for f in ia.PirParams.__dataclass_fields__})


def server_from_material(P: O.Params, mat: O.Material) -> ia.PirServer:
    s = O.sizes(P)
    sda = mat.second_dim_a if s.t > 1 else None
    return ia.PirServer.create_flat(to_gpu_params(P), mat.db, mat.db_cols, mat.pack,
                                    mat.a_tilde, mat.t_vec_h, sda)


def gpu_request(req: O.Request) -> ia.PirRequest:
    return ia.PirRequest(req.first_dim, req.second_dim, req.packing_key)


def pir_server_test_params(log_q, log_q1, log_q2, t, mat_bytes, ell=1, n=1024, coeff_bits=64):
    """The typed configurations of crypto/pir/server/pir_server_test.cc:55-143
    (d = 32, p = 257, log_digit 2, num_digits = ceil(log_q/2) - 1)."""
    nd = (log_q + 1) // 2 - 1
    return O.Params(log_degree=5, log_q=log_q, log_q1=log_q1, log_q2=log_q2, variance=4,
                    plaintext_modulus=257, pack_log_digit=2, pack_num_digits=nd,
                    eval_log_digit=2, eval_num_digits=nd, num_entries=n, interp_degree=t,
                    entry_size_multiple=ell, mat_bytes=mat_bytes, coeff_bits=coeff_bits)


# (log_q, log_q1, log_q2, t, mat_bytes) of pir_server_test.cc Uint64* types.
PIR_SERVER_TEST_TYPES = {
    "Uint64Mod0": (64, 61, 60, 32, 2),
    "Uint64Mod0Interp1": (64, 61, 60, 1, 2),
    "Uint64Mod62": (62, 61, 60, 32, 4),
    "Uint64Mod62ModSwitch": (62, 28, 20, 32, 4),
    "Uint64Mod62Interp1": (62, 61, 60, 1, 4),
}

# (log_q, log_q1, log_q2, t, mat_bytes) of the pir_server_test.cc Uint32*
# types (:55-93): CoeffType uint32_t, modulus 0 (log q = 32) or 2^30.
PIR_SERVER_TEST_TYPES_U32 = {
    "Uint32Mod0": (32, 31, 30, 32, 2),
    "Uint32Mod0Interp1": (32, 31, 30, 1, 2),
    "Uint32Mod30": (30, 29, 28, 32, 2),
    "Uint32Mod30Interp1": (30, 29, 28, 1, 2),
}
