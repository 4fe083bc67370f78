"""The packing batch and the serving loop at their full workload shapes, on
the GPU, against the oracle.

- 512 MatrixPack jobs (crypto/lwes_to_rlwe.cc:481-508) at d = 2048,
  q = 2^56, pack_log_digit 19, with LWE counts drawn uniformly from 16..2048
  and distinct material per job (16 GiB of pack tiles), answered by one
  graph launch: jobs at the start, middle and end of the tile buffer must
  equal the oracle's MatrixPack.
- The serving loop on the n = 262144, d = 2048, t = 4, ell = 2 server (2 GiB
  DB, 256 MiB pack matrix): a 64-request burst from two threads over three
  slots; sampled answers must equal the oracle's ProcessResponse
  (crypto/pir/server/pir_server.cc:152-306) and the same request answered
  alone.
- The vPIR online answer at vpir_metrics_tool's parameters (d = 2048, 256
  shards, q = 2^32, uint8 DB, int32 pack matrix with gadget (8, 4): 16 GiB):
  a batch of 256 requests (the two-groups-per-wave pack scan, the tiled
  FinalizeMatrixPack inner products) and one request alone, sampled against
  the oracle's VpirServer::ProcessResponse (crypto/vpir/server/
  vpir_server.cc:101-195).
"""
import threading

import numpy as np
import pytest

import inspire_amd as ia
import oracle as O
from helpers import gpu_request, server_from_material, to_gpu_params
from vpir_helpers import metrics_like, request

pytestmark = pytest.mark.gpu


def test_packing_512_jobs_one_graph_launch():
    log_d, k, njobs = 11, 2, 512
    d = 1 << log_d
    P = O.Params(log_degree=log_d, log_q=56, log_q1=56, log_q2=56, variance=11,
                 plaintext_modulus=1 << 16, pack_log_digit=19, pack_num_digits=k,
                 eval_log_digit=19, eval_num_digits=3, num_entries=d, interp_degree=1,
                 entry_size_multiple=1, db_bytes=2, mat_bytes=4)
    GP = ia.PirParams(**{f: getattr(P, f) for f in ia.PirParams.__dataclass_fields__})
    rng = np.random.default_rng(512)
    words = d * k * d // 2                  # one job's condensed int32 matrix
    stride = 1 << 18                        # job j is the window at j * stride words
    pool = np.frombuffer(rng.bytes((words + stride * njobs) * 8), dtype=np.uint64)
    packs = [pool[j * stride:j * stride + words].reshape(d, k * d // 2) for j in range(njobs)]
    qmask = np.uint64((1 << 56) - 1)
    a_tilde = [rng.integers(0, 1 << 63, size=d, dtype=np.uint64) & qmask for _ in range(njobs)]
    h = 1 << 18
    t_vec_h = [rng.integers(-h, h, size=(k, d)).astype(np.uint64) for _ in range(njobs)]
    counts = rng.integers(16, d + 1, size=njobs)
    bs = [rng.integers(0, 1 << 63, size=int(m), dtype=np.uint64) & qmask for m in counts]
    keys = [rng.integers(0, 1 << 63, size=(2 * k, d), dtype=np.uint64) & qmask
            for _ in range(njobs)]
    print("packer: 512 jobs, 16 GiB of tiles", flush=True)
    with ia.PirPacker.create(GP, packs, a_tilde, t_vec_h) as pk:
        got = pk.run(bs, keys)              # captured and launched as one graph
        again = pk.run(bs, keys)            # replayed
    np.testing.assert_array_equal(again, got)
    for j in (0, 1, 255, 256, 510, 511):
        bb = np.zeros(d, dtype=np.uint64)
        bb[:counts[j]] = bs[j]
        a, b = O.matrix_pack(P, bb, keys[j][:k], keys[j][k:], packs[j], t_vec_h[j], a_tilde[j])
        np.testing.assert_array_equal(got[j], np.stack([a, b]), err_msg=f"job {j}")


def test_serving_at_full_shape_matches_oracle_and_alone():
    P = O.metrics_tool_params(n=262144, log_d=11, t=4, ell=2)
    mat = O.synthetic_material(P, seed=262)
    reqs = [O.synthetic_request(P, seed=5000 + i) for i in range(64)]
    sampled = (0, 21, 42, 63)
    print("serving: n=262144 d=2048 t=4 ell=2, 64 requests, 3 slots", flush=True)
    with server_from_material(P, mat) as srv:
        alone = {i: srv.process_response(gpu_request(reqs[i])).ciphertexts for i in sampled}
        srv.serve_start(max_batch=32, slots=3, max_outstanding=128)
        try:
            def submitter(ids):
                for i in ids:
                    srv.submit(gpu_request(reqs[i]), tag=i)

            threads = [threading.Thread(target=submitter, args=(list(range(h, 64, 2)),))
                       for h in (0, 1)]
            for th in threads:
                th.start()
            for th in threads:
                th.join()
            served = {}
            while len(served) < 64:
                r = srv.poll(60000)
                assert r is not None, f"no result after 60 s ({len(served)} of 64)"
                tag, resp, status, lat, msg = r
                assert status == 0, msg
                served[tag] = resp.ciphertexts
            stats = srv.serve_stats()
        finally:
            srv.serve_stop()
    assert stats["answered"] == 64 and stats["slots"] >= 1
    for i in sampled:
        np.testing.assert_array_equal(served[i], alone[i], err_msg=f"request {i}")
        want, _, _ = O.process_response(P, mat, reqs[i])
        np.testing.assert_array_equal(served[i], want, err_msg=f"request {i}")


def test_vpir_metrics_shape_256_requests_match_oracle():
    P = metrics_like(ell=256, n=2048, log_d=11)
    s = O.sizes(P)
    d, ell, k, n = s.d, s.ell, s.k, P.num_entries
    M = ell * d
    rng = np.random.default_rng(91)
    db = np.frombuffer(rng.bytes(M * n), dtype=np.uint64).reshape(M, n // 8)
    # int32 entries inside the preprocessing bound |x| < 2^(log_digit + log d - 1).
    pack32 = (np.frombuffer(rng.bytes(M * k * d * 4), dtype=np.int32) >> 13).reshape(M, k * d)
    pack = pack32.view(np.uint64).reshape(M, k * d // 2)
    a_tilde = rng.integers(0, 1 << 32, size=(ell, d), dtype=np.uint64)
    t_vec_h = rng.integers(-128, 128, size=(ell, k, d)).astype(np.uint64)
    reqs = [request(P, 9000 + i, False) for i in range(256)]
    srv = ia.VpirServer.create(to_gpu_params(P), False, db, M, n, pack, a_tilde, t_vec_h)
    with srv:
        got, st, _ = srv.process_batch(reqs)
        one = srv.process_response(*reqs[200])
    assert st == [0] * len(reqs)
    np.testing.assert_array_equal(one, got[200])
    for i in (0, 131, 255):
        want = O.vpir_process(P, False, db, n, pack, a_tilde, t_vec_h, *reqs[i])
        np.testing.assert_array_equal(got[i], want, err_msg=f"request {i}")
