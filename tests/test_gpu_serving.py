"""hipGraph replays and the serving loop (continuous batching), on the GPU.

Parity here is "batched == sequential == oracle": the reference answers one
request per call, so every answer of a graph replay or of a served batch
must equal, bit for bit, the same request answered alone, and the oracle's
ProcessResponse on the same inputs.
"""
import threading
import time

import numpy as np
import pytest

import inspire_amd as ia
import oracle as O
from helpers import gpu_request, server_from_material

pytestmark = pytest.mark.gpu

P = O.metrics_tool_params(n=4096, log_d=8, t=4, ell=2)


@pytest.fixture(scope="module")
def material():
    return O.synthetic_material(P, seed=77)


@pytest.fixture(scope="module")
def requests():
    return [O.synthetic_request(P, seed=300 + i) for i in range(40)]


def test_graph_replays_match_eager(material, requests):
    reqs = requests[:12]
    with server_from_material(P, material) as srv:
        eager, st, _ = srv.process_batch([gpu_request(r) for r in reqs])
        assert st == [0] * len(reqs)
        srv.use_graphs(True)
        for rep in range(4):
            got, st2, _ = srv.process_batch([gpu_request(r) for r in reqs])
            assert st2 == [0] * len(reqs)
            for a, b in zip(eager, got):
                np.testing.assert_array_equal(b.ciphertexts, a.ciphertexts)
            # Eager work between the replays.
            srv.use_graphs(False)
            one = srv.process_response(gpu_request(reqs[rep]))
            np.testing.assert_array_equal(one.ciphertexts, eager[rep].ciphertexts)
            srv.use_graphs(True)
        # Another batch size gets its own graph.
        got5, _, _ = srv.process_batch([gpu_request(r) for r in reqs[:5]])
        for a, b in zip(eager[:5], got5):
            np.testing.assert_array_equal(b.ciphertexts, a.ciphertexts)
    for i in (0, 7):
        want, _, _ = O.process_response(P, material, reqs[i])
        np.testing.assert_array_equal(eager[i].ciphertexts, want)


def test_graph_scan_timing_nodes(material, requests):
    # Graphs captured with scan event nodes answer exactly as eager runs, and
    # their replays (not only eager runs) add to the scan times; switching the
    # mode back recaptures graphs without the nodes.
    reqs = [gpu_request(r) for r in requests[:12]]
    with server_from_material(P, material) as srv:
        eager, _, _ = srv.process_batch(reqs)
        srv.batch_stage(reqs)
        srv.use_graphs(True)
        srv.batch_run(12)
        srv.scan_times_ms(reset=True)
        srv.batch_run(12)
        assert srv.scan_times_ms(reset=True) == (0.0, 0.0, 0)
        srv.time_graph_scans(True)
        for rep in range(3):
            srv.batch_run(12)
            for a, b in zip(eager, srv.batch_fetch(12)):
                np.testing.assert_array_equal(b.ciphertexts, a.ciphertexts)
        db_ms, pk_ms, n = srv.scan_times_ms(reset=True)
        assert n == 3 and db_ms > 0 and pk_ms > 0
        got, _, _ = srv.process_batch(reqs)
        for a, b in zip(eager, got):
            np.testing.assert_array_equal(b.ciphertexts, a.ciphertexts)
        srv.time_graph_scans(False)
        srv.scan_times_ms(reset=True)
        srv.batch_run(12)
        assert srv.scan_times_ms(reset=True)[2] == 0
        for a, b in zip(eager, srv.batch_fetch(12)):
            np.testing.assert_array_equal(b.ciphertexts, a.ciphertexts)


def _drain(srv, n, timeout_ms=60000):
    got = {}
    while len(got) < n:
        r = srv.poll(timeout_ms)
        assert r is not None, f"no result after {timeout_ms} ms ({len(got)} of {n})"
        tag, resp, status, lat, msg = r
        assert status == 0, msg
        assert lat > 0
        got[tag] = resp.ciphertexts
    return got


def test_serving_alone_equals_in_the_mix(material, requests):
    with server_from_material(P, material) as srv:
        srv.serve_start(max_batch=8, slots=2, max_outstanding=64)
        try:
            srv.submit(gpu_request(requests[0]), tag=1000)
            alone = _drain(srv, 1)[1000]
            # The same request again, in the middle of a burst submitted from
            # two threads.
            order = list(range(1, len(requests)))
            order.insert(17, 0)
            halves = [order[0::2], order[1::2]]

            def submitter(ids):
                for i in ids:
                    srv.submit(gpu_request(requests[i]), tag=i)

            threads = [threading.Thread(target=submitter, args=(h,)) for h in halves]
            for th in threads:
                th.start()
            for th in threads:
                th.join()
            mixed = _drain(srv, len(order))
            stats = srv.serve_stats()
        finally:
            srv.serve_stop()
    np.testing.assert_array_equal(mixed[0], alone)
    want, _, _ = O.process_response(P, material, requests[0])
    np.testing.assert_array_equal(alone, want)
    for i in (5, 23, 39):
        want, _, _ = O.process_response(P, material, requests[i])
        np.testing.assert_array_equal(mixed[i], want)
    assert stats["answered"] == len(order) + 1
    assert stats["slots"] >= 1 and stats["max_batch"] >= 1


def test_serving_batches_a_burst(material, requests):
    # One slot: while it answers the first arrivals the rest of the burst
    # queues and joins the next batches.
    with server_from_material(P, material) as srv:
        srv.serve_start(max_batch=16, slots=1, max_outstanding=64)
        try:
            for i, r in enumerate(requests[:32]):
                srv.submit(gpu_request(r), tag=i)
            got = _drain(srv, 32)
            stats = srv.serve_stats()
        finally:
            srv.serve_stop()
    assert stats["batches"] < 32 and stats["mean_batch"] > 1.0, stats
    solo, _, _ = O.process_response(P, material, requests[31])
    np.testing.assert_array_equal(got[31], solo)


def test_serving_rejects_invalid_and_full_queue(material, requests):
    with server_from_material(P, material) as srv:
        with pytest.raises(ia.PirError) as ei:
            srv.submit(gpu_request(requests[0]), tag=1)
        assert ei.value.code == ia.FAILED_PRECONDITION
        srv.serve_start(max_batch=4, slots=1, max_outstanding=2)
        try:
            bad = O.Request(requests[0].first_dim, requests[0].second_dim,
                            requests[0].packing_key[:-1])
            with pytest.raises(ia.PirError) as ei:
                srv.submit(gpu_request(bad), tag=2)
            assert ei.value.code == ia.INVALID_ARGUMENT
            assert ei.value.message == "Packing key must contain 2 * num_digits polynomials."
            srv.submit(gpu_request(requests[1]), tag=3)
            srv.submit(gpu_request(requests[2]), tag=4)
            # Two submitted and not yet polled: the third is refused.
            with pytest.raises(ia.PirError) as ei:
                srv.submit(gpu_request(requests[3]), tag=5)
            assert ei.value.code == ia.RESOURCE_EXHAUSTED
            first = _drain(srv, 1)
            srv.submit(gpu_request(requests[3]), tag=5)
            rest = _drain(srv, 2)
        finally:
            srv.serve_stop()
    got = {**first, **rest}
    assert sorted(got) == [3, 4, 5]
    want, _, _ = O.process_response(P, material, requests[3])
    np.testing.assert_array_equal(got[5], want)


def test_direct_calls_while_serving(material, requests):
    with server_from_material(P, material) as srv:
        srv.serve_start(max_batch=8, slots=2, max_outstanding=64)
        try:
            for i in range(8):
                srv.submit(gpu_request(requests[i]), tag=i)
            direct = srv.process_response(gpu_request(requests[9])).ciphertexts
            served = _drain(srv, 8)
        finally:
            srv.serve_stop()
    want, _, _ = O.process_response(P, material, requests[9])
    np.testing.assert_array_equal(direct, want)
    want, _, _ = O.process_response(P, material, requests[4])
    np.testing.assert_array_equal(served[4], want)


def test_library_calls_while_serving(material, requests):
    # While the serving loop captures and replays graphs on its own thread,
    # the caller's thread grows workspaces, captures its own batch graphs and
    # runs the standalone entry points. None of them may touch the legacy
    # default stream (HIP rejects that during another thread's capture).
    rng = np.random.default_rng(5)
    u = rng.integers(0, 1 << 64, size=(3, 256), dtype=np.uint64)
    v = rng.integers(0, 1 << 64, size=(3, 256), dtype=np.uint64)
    want_ip, _ = O.inner_product_fft(u, v, 64, 64, 16)
    polys = rng.integers(0, 1 << 40, size=(4, 256), dtype=np.uint64)
    with server_from_material(P, material) as srv, \
            server_from_material(P, material) as other:
        eager = [other.process_response(gpu_request(r)) for r in requests[:6]]
        other.use_graphs(True)
        srv.serve_start(max_batch=8, slots=2, max_outstanding=64)
        try:
            tag = 0
            for wave in range(5):
                for _ in range(wave + 3):
                    srv.submit(gpu_request(requests[tag % len(requests)]), tag=tag)
                    tag += 1
                # A new batch size: a larger workspace and a new graph.
                n = wave + 2
                got, st, _ = other.process_batch([gpu_request(r) for r in requests[:n]])
                assert st == [0] * n
                for a, b in zip(eager[:n], got):
                    np.testing.assert_array_equal(b.ciphertexts, a.ciphertexts)
                np.testing.assert_array_equal(ia.inner_product_fft(u, v, 64, 64, 16), want_ip)
                enc = ia.poly_to_proto(polys, 40)
                np.testing.assert_array_equal(ia.poly_from_proto(enc, 256, 40), polys)
                assert other.fft_margin() >= 0.0
            served = _drain(srv, tag)
        finally:
            srv.serve_stop()
    for t in (0, tag - 1):
        want, _, _ = O.process_response(P, material, requests[t % len(requests)])
        np.testing.assert_array_equal(served[t], want)


def test_stop_wakes_a_blocked_poller(material, requests):
    # pirg_serve_stop while another thread waits in pirg_serve_poll: the
    # poller returns at once (FAILED_PRECONDITION, nothing left to deliver)
    # instead of using a freed loop; a result still undelivered at the stop
    # is handed to the poller first.
    with server_from_material(P, material) as srv:
        srv.serve_start(max_batch=4, slots=1, max_outstanding=8)
        srv.submit(gpu_request(requests[0]), tag=7)
        first = srv.poll(30000)
        assert first is not None and first[0] == 7
        srv.submit(gpu_request(requests[1]), tag=8)
        got = {}

        def poller():
            t0 = time.perf_counter()
            outcomes = []
            for _ in range(2):
                try:
                    outcomes.append(srv.poll(30000))
                except ia.PirError as e:
                    outcomes.append(e)
            got["outcomes"] = outcomes
            got["s"] = time.perf_counter() - t0

        th = threading.Thread(target=poller)
        th.start()
        time.sleep(0.5)
        srv.serve_stop()
        th.join(20)
        assert not th.is_alive()
        assert got["s"] < 10
        r, e = got["outcomes"]
        assert r is not None and r[0] == 8 and r[2] == 0
        assert isinstance(e, ia.PirError) and e.code == ia.FAILED_PRECONDITION
        with pytest.raises(ia.PirError) as ei:
            srv.submit(gpu_request(requests[2]), tag=9)
        assert ei.value.code == ia.FAILED_PRECONDITION
        srv.serve_stop()  # stopping twice is harmless
    want, _, _ = O.process_response(P, material, requests[1])
    np.testing.assert_array_equal(r[1].ciphertexts, want)


def test_batched_submit_and_poll(material, requests):
    # pirg_serve_submit_batch / pirg_serve_poll_batch: per-request validation
    # statuses, all-or-nothing admission, and answers equal to the oracle's.
    with server_from_material(P, material) as srv:
        srv.serve_start(max_batch=8, slots=2, max_outstanding=24)
        try:
            bad = O.Request(requests[0].first_dim, requests[0].second_dim,
                            requests[0].packing_key[:-1])
            batch = [gpu_request(r) for r in requests[:10]] + [gpu_request(bad)]
            st = srv.submit_batch(batch, list(range(11)))
            assert st[:10] == [0] * 10 and st[10] == ia.INVALID_ARGUMENT
            with pytest.raises(ia.PirError) as ei:   # 10 + 15 > 24: none queued
                srv.submit_batch([gpu_request(r) for r in requests[10:25]],
                                 list(range(100, 115)))
            assert ei.value.code == ia.RESOURCE_EXHAUSTED
            got = {}
            while len(got) < 10:
                r = srv.poll_batch(16, 30000)
                assert r is not None
                tags, outs, sts, lats = r
                for tg, o, s_, lt in zip(tags, outs, sts, lats):
                    assert s_ == 0 and lt > 0
                    got[tg] = o.copy()
            assert sorted(got) == list(range(10))
            assert srv.poll_batch(4, 50) is None
        finally:
            srv.serve_stop()
    for i in (0, 9):
        want, _, _ = O.process_response(P, material, requests[i])
        np.testing.assert_array_equal(got[i], want)



def test_zero_copy_records(material, requests):
    # pirg_serve_submit_records: records written by pirg_request_record into
    # pinned memory from pirg_host_alloc are answered as the same requests
    # submitted by copy (and as the oracle), mixed in one loop; records
    # outside that memory, or running past its end, are refused whole.
    reqs = [gpu_request(r) for r in requests[:12]]
    with server_from_material(P, material) as srv:
        alone = [srv.process_response(r).ciphertexts for r in reqs]
        words = srv.record_words()
        with ia.HostRecords(srv, 12) as hr:
            assert hr.array.shape == (12, words)
            for i, r in enumerate(reqs):
                srv.write_record(r, hr.array[i])
            bad = O.Request(requests[0].first_dim, requests[0].second_dim,
                            requests[0].packing_key[:-1])
            with pytest.raises(ia.PirError) as ei:
                srv.write_record(gpu_request(bad), hr.array[0].copy())
            assert ei.value.code == ia.INVALID_ARGUMENT
            srv.serve_start(max_batch=8, slots=2, max_outstanding=64)
            try:
                outside = np.zeros(words, dtype=np.uint64)
                with pytest.raises(ia.PirError) as ei:
                    srv.submit_records([hr.array[0], outside], [0, 1])
                assert ei.value.code == ia.INVALID_ARGUMENT
                # A record starting inside the block but ending past it.
                flat = hr.array.reshape(-1)
                with pytest.raises(ia.PirError) as ei:
                    srv.submit_records([flat[11 * words + 1:]], [0])
                assert ei.value.code == ia.INVALID_ARGUMENT
                # Records (adjacent and not, each twice) and copied requests.
                order = [0, 1, 2, 3, 7, 5, 11, 10, 9, 4, 6, 8]
                srv.submit_records([hr.array[i] for i in order], order)
                srv.submit_batch(reqs[:4], [100, 101, 102, 103])
                srv.submit_records([hr.array[i] for i in order[::-1]],
                                   [200 + i for i in order[::-1]])
                got = {}
                while len(got) < 28:
                    r = srv.poll_batch(16, 30000)
                    assert r is not None
                    for tg, o, s_ in zip(r[0], r[1], r[2]):
                        assert s_ == 0
                        got[tg] = o.copy()
            finally:
                srv.serve_stop()
    for i in range(12):
        np.testing.assert_array_equal(got[i], alone[i])
        np.testing.assert_array_equal(got[200 + i], alone[i])
    for i in range(4):
        np.testing.assert_array_equal(got[100 + i], alone[i])
    want, _, _ = O.process_response(P, material, requests[5])
    np.testing.assert_array_equal(got[5], want)


def test_admission_by_device_memory_limit(material, requests):
    # Admission by device memory, forced with a serving memory limit: slots
    # are dropped first (down to two), then the batch capacity halves; a
    # limit below two single-request slots refuses the start; the loop then
    # starts and answers under the default budget again.
    reqs = [gpu_request(r) for r in requests[:6]]
    with server_from_material(P, material) as srv:
        want = [srv.process_response(r).ciphertexts for r in reqs]
        w8, w4, w1 = (srv.serve_workspace_bytes(b) for b in (8, 4, 1))
        assert w1 < w4 < w8
        srv.serve_set_memory_limit(2 * w4 + w4 // 2)   # two slots of 4 fit, not of 8
        srv.serve_start(max_batch=8, slots=3, max_outstanding=64)
        try:
            st = srv.serve_stats()
            assert (st["slots"], st["max_batch"]) == (2, 4)
            assert srv.submit_batch(reqs, list(range(6))) == [0] * 6
            got = {}
            while len(got) < 6:
                r = srv.poll_batch(8, 30000)
                assert r is not None
                for tg, o, s_ in zip(r[0], r[1], r[2]):
                    assert s_ == 0
                    got[tg] = o.copy()
        finally:
            srv.serve_stop()
        for i in range(6):
            np.testing.assert_array_equal(got[i], want[i])
        srv.serve_set_memory_limit(2 * w1 - 1)
        with pytest.raises(ia.PirError) as ei:
            srv.serve_start(max_batch=8, slots=2, max_outstanding=64)
        assert ei.value.code == ia.RESOURCE_EXHAUSTED
        assert "not enough free device memory" in ei.value.message
        srv.serve_set_memory_limit(0)
        srv.serve_start(max_batch=8, slots=2, max_outstanding=64)
        try:
            assert srv.serve_stats()["max_batch"] == 8
            srv.submit(reqs[0], 99)
            r = srv.poll(30000)
            assert r is not None and r[0] == 99
            np.testing.assert_array_equal(r[1].ciphertexts, want[0])
        finally:
            srv.serve_stop()
