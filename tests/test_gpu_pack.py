"""Batched MatrixPack on the GPU against the oracle's MatrixPack
(crypto/lwes_to_rlwe.cc:481-508), job by job, bit for bit.

The oracle's MatrixPack is pinned by Pack == MatrixPack and by decryption
(test_oracle_pir.py); here every GPU job must equal it on the same material,
for ragged LWE counts (zero-padded b), shared and per-job keys, int32 and
int16 pack matrices, graph replays and eager launches.
"""
import numpy as np
import pytest

import inspire_amd as ia
import oracle as O
from helpers import to_gpu_params

pytestmark = pytest.mark.gpu


def pack_params(log_d, mat_bytes=4, log_q=56):
    return O.Params(log_degree=log_d, log_q=log_q, log_q1=log_q, log_q2=log_q, variance=11,
                    plaintext_modulus=1 << 16, pack_log_digit=19, pack_num_digits=2,
                    eval_log_digit=19, eval_num_digits=3, num_entries=1 << log_d,
                    interp_degree=1, entry_size_multiple=1, db_bytes=2, mat_bytes=mat_bytes)


def random_job(P, rng):
    d, k = 1 << P.log_degree, P.pack_num_digits
    if P.mat_bytes == 4:
        vals = rng.integers(-(1 << 31), 1 << 31, size=(d, k * d), dtype=np.int64)
        pack = (vals.astype(np.int32).view(np.uint32).astype(np.uint64))
        pack = (pack[:, 0::2] | (pack[:, 1::2] << np.uint64(32))).astype(np.uint64)
    else:
        pack = rng.integers(-(1 << 15), 1 << 15, size=(d, k * d)).astype(np.int16)
    qmask = np.uint64((1 << P.log_q) - 1) if P.log_q < 64 else np.uint64(~0 & ((1 << 64) - 1))
    a_tilde = rng.integers(0, 1 << 63, size=d, dtype=np.uint64) & qmask
    h = 1 << (P.pack_log_digit - 1)
    t_vec_h = rng.integers(-h, h, size=(k, d)).astype(np.uint64)
    return pack, a_tilde, t_vec_h


def random_key(P, rng):
    d, k = 1 << P.log_degree, P.pack_num_digits
    qmask = np.uint64((1 << P.log_q) - 1) if P.log_q < 64 else np.uint64((1 << 64) - 1)
    return rng.integers(0, 1 << 63, size=(2 * k, d), dtype=np.uint64) & qmask


def oracle_pack(P, pack, a_tilde, t_vec_h, b, key):
    d, k = 1 << P.log_degree, P.pack_num_digits
    bb = np.zeros(d, dtype=np.uint64)
    bb[:len(b)] = b
    a, bo = O.matrix_pack(P, bb, key[:k], key[k:], pack, t_vec_h, a_tilde)
    return np.stack([a, bo])


@pytest.mark.parametrize("log_d,mat_bytes,njobs", [(8, 4, 5), (6, 4, 3), (8, 2, 4), (11, 4, 2)])
def test_packer_matches_oracle(log_d, mat_bytes, njobs):
    P = pack_params(log_d, mat_bytes)
    d = 1 << log_d
    rng = np.random.default_rng(100 * log_d + njobs)
    jobs = [random_job(P, rng) for _ in range(njobs)]
    keys = [random_key(P, rng) for _ in range(njobs)]
    keys[-1] = keys[0]  # two jobs share a key
    counts = [d, 16, 1, 0, d - 3][:njobs]
    qmask = (1 << P.log_q) - 1
    bs = [rng.integers(0, 1 << 62, size=m, dtype=np.uint64) & np.uint64(qmask) for m in counts]
    with ia.PirPacker.create(to_gpu_params(P), [j[0] for j in jobs], [j[1] for j in jobs],
                             [j[2] for j in jobs]) as pk:
        got = pk.run(bs, keys)
        again = pk.run(bs, keys)  # graph replay
        pk.use_graph(False)
        eager = pk.run(bs, keys)
    np.testing.assert_array_equal(again, got)
    np.testing.assert_array_equal(eager, got)
    for j in range(njobs):
        want = oracle_pack(P, jobs[j][0], jobs[j][1], jobs[j][2], bs[j], keys[j])
        np.testing.assert_array_equal(got[j], want, err_msg=f"job {j}")


def test_packer_on_preprocessed_material():
    # Material from the oracle's PreprocessMatrixPack on explicit LWE vectors.
    P = pack_params(7)
    d = 1 << 7
    rng = np.random.default_rng(5)
    qmask = np.uint64((1 << P.log_q) - 1)
    mats, keys, bs = [], [], []
    for m in (d, 40):
        a_is = rng.integers(0, 1 << 63, size=(d, d), dtype=np.uint64) & qmask
        a_is[m:] = 0
        w = rng.integers(0, 1 << 63, size=(2 * P.pack_num_digits, d), dtype=np.uint64) & qmask
        mats.append(O.preprocess_matrix_pack(P, a_is, w))
        keys.append(random_key(P, rng))
        bs.append(rng.integers(0, 1 << 62, size=m, dtype=np.uint64) & qmask)
    with ia.PirPacker.create(to_gpu_params(P), [m[0] for m in mats], [m[1] for m in mats],
                             [m[2] for m in mats]) as pk:
        got = pk.run(bs, keys)
    for j, (pack, a_tilde, t_vec_h) in enumerate(mats):
        want = oracle_pack(P, pack, a_tilde, t_vec_h, bs[j], keys[j])
        np.testing.assert_array_equal(got[j], want)


def test_packer_rejects_oversized_jobs():
    P = pack_params(6)
    rng = np.random.default_rng(1)
    job = random_job(P, rng)
    with ia.PirPacker.create(to_gpu_params(P), [job[0]], [job[1]], [job[2]]) as pk:
        with pytest.raises(ia.PirError) as ei:
            pk.run([np.zeros(65, dtype=np.uint64)], [random_key(P, rng)])
        assert ei.value.code == ia.INVALID_ARGUMENT
        assert "exceeds the ring degree" in ei.value.message
