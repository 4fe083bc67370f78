"""Headline benchmark: PIR answers per second on one MI355X at n = 2^20, d = 4096.

Workload (the large resident-DB configuration): n = 1048576 records, ring
degree d = 4096, interpolation degree t = 8, ell = 1, the pir_metrics_tool
parameters (q = 2^52, p = 65535, pack gadget (19, 2), eval gadget (19, 3),
q1 = 2^30, q2 = 2^18). The 8 GiB uint16 DB and the 1 GiB int32 pack matrix
are synthetic but range-faithful (DB uniform in [0, p), pack entries uniform
in [-2^30, 2^30), t_vec_h signed 19-bit digits, a_tilde / second_dim_a /
requests uniform mod 2^52), all seeded and resident in HBM before timing.

A step answers one batch of B requests (PirServer::ProcessResponse for each,
batched); `value` is requests/s of the batched path (B = --batch); the
single-request path (B = 1) is timed beside it with its p50 / p99 latency.

The CPU baseline is the oracle's ProcessResponse restatement (single thread,
matrix loops auto-vectorised) on the same material, timed on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import inspire_amd as ia  # noqa: E402

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
INT8_PEAK_TOPS = 5000.0        # MI355X dense INT8 MFMA


def synthetic_material(P: ia.PirParams, seed: int):
    d, t, ell = P.degree, P.interp_degree, P.entry_size_multiple
    k, e = P.pack_num_digits, P.eval_num_digits
    r = P.num_entries // t
    K = (r + 31) // 32 * 32
    M = ell * t * d
    rng = np.random.default_rng(seed)
    p = P.plaintext_modulus
    db = np.frombuffer(rng.bytes(M * K * 2), dtype=np.uint16).copy().reshape(M, K)
    bad = db >= p
    while bad.any():
        db[bad] = np.frombuffer(rng.bytes(int(bad.sum()) * 2), dtype=np.uint16)
        bad = db >= p
    db[:, r:] = 0
    db_words = db.view(np.uint64).reshape(M, K // 4)
    pack32 = np.frombuffer(rng.bytes(M * k * d * 4), dtype=np.int32).reshape(M, k * d) >> 1
    pack_words = pack32.view(np.uint64).reshape(M, k * d // 2)
    qmask = np.uint64((1 << P.log_q) - 1)
    a_tilde = rng.integers(0, 1 << 63, size=(ell * t, d), dtype=np.uint64) & qmask
    h = 1 << (P.pack_log_digit - 1)
    t_vec_h = rng.integers(-h, h, size=(ell * t, k, d)).astype(np.uint64)
    sda = rng.integers(0, 1 << 63, size=(2 * e, d), dtype=np.uint64) & qmask
    return dict(db=db_words, db_cols=K, pack=pack_words, a_tilde=a_tilde, t_vec_h=t_vec_h,
                second_dim_a=sda)


def synthetic_requests(P: ia.PirParams, n: int, seed: int):
    d = P.degree
    r = P.num_entries // P.interp_degree
    rng = np.random.default_rng(seed)
    qmask = np.uint64((1 << P.log_q) - 1)
    out = []
    for _ in range(n):
        fd = rng.integers(0, 1 << 63, size=r, dtype=np.uint64) & qmask
        sd = rng.integers(0, 1 << 63, size=(2 * P.eval_num_digits, d), dtype=np.uint64) & qmask
        pk = rng.integers(0, 1 << 63, size=(2 * P.pack_num_digits, d), dtype=np.uint64) & qmask
        out.append(ia.PirRequest(fd, sd, pk))
    return out


def time_batches(srv: ia.PirServer, B: int, steps: int, warmup: int):
    for _ in range(warmup):
        srv.batch_run(B)
    srv.scan_times_ms(reset=True)
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        s0 = time.perf_counter()
        srv.batch_run(B)
        lat.append((time.perf_counter() - s0) * 1e3)
    total = time.perf_counter() - t0
    db_ms, pk_ms, n = srv.scan_times_ms(reset=True)
    return total, np.array(lat), db_ms / max(n, 1), pk_ms / max(n, 1)


def cpu_baseline(P, mat, req, max_queries=2, budget_s=30.0):
    """The oracle's ProcessResponse (one thread) on full requests of the same
    workload; stops after `max_queries` or the time budget."""
    import oracle as O
    OP = O.Params(**{f: getattr(P, f) for f in O.Params.__dataclass_fields__})
    omat = O.Material(mat["db"], mat["pack"], mat["a_tilde"], mat["t_vec_h"],
                      mat["second_dim_a"], mat["db_cols"])
    oreq = O.Request(req.first_dimension_query, req.second_dimension_query, req.packing_key)
    times, out = [], None
    start = time.perf_counter()
    for _ in range(max_queries):
        t0 = time.perf_counter()
        out, stages, _ = O.process_response(OP, omat, oreq)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - start > budget_s:
            break
    return len(times) / sum(times), len(times), out, stages


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--log-degree", type=int, default=12)
    ap.add_argument("--num-entries", type=int, default=1 << 20)
    ap.add_argument("--interp-degree", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-queries", type=int, default=1)
    ap.add_argument("--profile", action="store_true",
                    help="short run for rocprofv3: 2 single + 2 batched answers and a "
                         "2 GiB copy-kernel calibration, no CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the answer path runs on a single MI355X in this "
              "build (multi-GPU sharding is described in DESIGN.md, not implemented).",
              file=sys.stderr)
        sys.exit(2)

    P = ia.metrics_tool_params(n=args.num_entries, log_d=args.log_degree, t=args.interp_degree)
    d, t = P.degree, P.interp_degree
    r = P.num_entries // t
    M = t * d
    K = (r + 31) // 32 * 32
    print(f"[bench] generating material n={P.num_entries} d={d} t={t} "
          f"(DB {M * K * 2 / 2**30:.2f} GiB, pack {M * 2 * d * 4 / 2**30:.2f} GiB)",
          file=sys.stderr, flush=True)
    mat = synthetic_material(P, seed=2026)
    t0 = time.perf_counter()
    srv = ia.PirServer.create_flat(P, mat["db"], mat["db_cols"], mat["pack"], mat["a_tilde"],
                                   mat["t_vec_h"], mat["second_dim_a"])
    print(f"[bench] server resident in HBM ({time.perf_counter() - t0:.1f}s)",
          file=sys.stderr, flush=True)
    B = args.batch
    reqs = synthetic_requests(P, B, seed=7)
    if args.profile:
        srv.batch_stage(reqs[:1])
        for _ in range(2):
            srv.batch_run(1)
        srv.batch_stage(reqs)
        for _ in range(2):
            srv.batch_run(B)
        copy_ms = ia.copy_probe(2 << 30, repeats=3)
        read_ms = ia.read_probe(4 << 30, repeats=3)
        print(json.dumps({"profile": True, "copy_bytes": 2 << 30, "copy_ms": copy_ms,
                          "read_bytes": 4 << 30, "read_ms": read_ms,
                          "scan_ms": srv.scan_times_ms()}))
        srv.close()
        return

    # Single-request path.
    srv.batch_stage(reqs[:1])
    s_total, s_lat, s_db_ms, s_pk_ms = time_batches(srv, 1, max(args.steps, 5), args.warmup)
    single_qps = max(args.steps, 5) / s_total
    print(f"[bench] single: {np.mean(s_lat):.3f} ms/query, DB scan {s_db_ms:.3f} ms",
          file=sys.stderr, flush=True)
    # Batched path.
    srv.batch_stage(reqs)
    b_total, b_lat, b_db_ms, b_pk_ms = time_batches(srv, B, args.steps, args.warmup)
    batch_qps = B * args.steps / b_total
    print(f"[bench] batch {B}: {np.mean(b_lat):.3f} ms/batch, {batch_qps:.1f} q/s, "
          f"DB scan {b_db_ms:.3f} ms", file=sys.stderr, flush=True)

    db_bytes = M * K * 2
    nl = 7
    pairs = 13  # (plane, digit) pairs with 8 (p + c) < 52
    mfma_ops_batch = 2.0 * M * K * B * pairs
    roofline_single = {
        "bound": "hbm", "achieved": db_bytes / (s_db_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": db_bytes / (s_db_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "traffic": None}
    batch_tops = mfma_ops_batch / (b_db_ms * 1e-3) / 1e12
    batch_gbs = db_bytes / (b_db_ms * 1e-3) / 1e9
    if batch_tops / INT8_PEAK_TOPS >= batch_gbs / HBM_PEAK_GBS:
        roofline = {"bound": "mfma", "achieved": batch_tops, "peak": INT8_PEAK_TOPS,
                    "unit": "TOP/s", "frac": batch_tops / INT8_PEAK_TOPS, "traffic": None}
    else:
        roofline = {"bound": "hbm", "achieved": batch_gbs, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": batch_gbs / HBM_PEAK_GBS, "traffic": None}

    cpu = None
    if not args.no_cpu:
        qps, nq, cpu_out, stages = cpu_baseline(P, mat, reqs[0], max_queries=args.cpu_queries)
        gpu_out = srv.process_response(reqs[0]).ciphertexts
        cpu = {"value": qps, "unit": "queries/s", "cores": 1, "kind": "port",
               "sample": f"{nq} full request(s) of the same n=2^20 d=4096 t=8 workload "
                         f"through the oracle ProcessResponse (stage ms: db "
                         f"{stages[0]:.0f}, pack {stages[1]:.0f}, eval {stages[2]:.0f})",
               "bit_identical": bool(np.array_equal(cpu_out, gpu_out))}

    line = {
        "metric": "queries/sec on 1 GPU at n=2^20,d=4096 (batched vs single); p50/p99 answer ms",
        "value": batch_qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": float(np.mean(b_lat)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded, range-faithful material and requests)",
        "config": {"workload": "resident-DB PIR answer, n=2^20 d=4096 t=8 ell=1",
                   "n": P.num_entries, "d": d, "t": t, "ell": 1, "batch": B,
                   "log_q": P.log_q, "p": P.plaintext_modulus},
        "single_qps": single_qps,
        "batched_over_single": batch_qps / single_qps,
        "p50_ms_single": float(np.percentile(s_lat, 50)),
        "p99_ms_single": float(np.percentile(s_lat, 99)),
        "p50_ms_batch": float(np.percentile(b_lat, 50)),
        "p99_ms_batch": float(np.percentile(b_lat, 99)),
        "roofline": roofline,
        "roofline_single_scan": roofline_single,
        "db_scan_ms_single": s_db_ms,
        "db_scan_ms_batch": b_db_ms,
        "pack_scan_ms_single": s_pk_ms,
        "cpu_baseline": cpu,
    }
    srv.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
