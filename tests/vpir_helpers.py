"""vPIR parameter sets and material for the tests (the typed configurations
of crypto/vpir/server/vpir_server_test.cc:52-105 and the online parameters of
crypto/vpir/benchmarking/vpir_metrics_tool.cc:64-89)."""
import numpy as np

import oracle as O


def vpir_params(log_q, mat_bytes, ell=1, log_d=5, n=1024, pack_log_digit=2,
                pack_num_digits=None, db_bytes=2, p=257):
    if pack_num_digits is None:
        pack_num_digits = (log_q + 1) // 2 - 1
    return O.Params(log_degree=log_d, log_q=log_q, log_q1=log_q, log_q2=log_q, variance=4,
                    plaintext_modulus=p, pack_log_digit=pack_log_digit,
                    pack_num_digits=pack_num_digits, eval_log_digit=2, eval_num_digits=1,
                    num_entries=n, interp_degree=1, entry_size_multiple=ell,
                    db_bytes=db_bytes, mat_bytes=mat_bytes)


# vpir_server_test.cc Uint64 types (d = 32, n = 1024, log_digit 2).
VPIR_TEST_TYPES = {
    "Uint64Mod0": dict(log_q=64, mat_bytes=2),
    "Uint64Mod62": dict(log_q=62, mat_bytes=4),
}


def metrics_like(ell=4, n=2048, log_d=11):
    # vpir_metrics_tool's online phase: q = 2^32 (uint32 modulus 0), uint8 DB,
    # int32 pack with gadget (8, 4), p = 256.
    return vpir_params(32, 4, ell=ell, log_d=log_d, n=n, pack_log_digit=8, pack_num_digits=4,
                       db_bytes=1, p=256)


def request(P, seed, local_finalize):
    s = O.sizes(P)
    rng = np.random.default_rng(seed)
    qm = np.uint64((1 << P.log_q) - 1) if P.log_q < 64 else np.uint64((1 << 64) - 1)
    fd = rng.integers(0, 1 << 63, size=P.num_entries, dtype=np.uint64) & qm
    nk = s.k if local_finalize else 2 * s.k
    pk = rng.integers(0, 1 << 63, size=(nk, s.d), dtype=np.uint64) & qm
    return fd, pk
