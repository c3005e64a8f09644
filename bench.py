"""Benchmark: RS-encode + Merkle commit + first FRI fold of one B128 job on one GPU.

Workload (one "step"): a 2^log_n B128 message resident in HBM is Reed-Solomon
encoded at rate 1/2 with the interleaved batch 2^4 chosen by Verifier::setup
(codeword 2^(log_n+1) elements), committed with the SHA-256 Merkle tree (16-element
leaves), and folded once by the interleaved FRI fold with 4 pre-seeded
challenges.  Default log_n = 24 (256 MiB message, 512 MiB codeword).

Prints one JSON line.  `value` is message GB/s (16 * 2^log_n bytes / step time),
the convention of the reference's pcs/ntt benches.  `roofline` covers the
dominant kernel (the NTT pass), measured live with HIP events on the stream the
kernels run on.  `cpu_baseline` times the reference's x86-64 commit path
(restated in oracle/b64_cpu_baseline.c) on the host threads, on the same message.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# From rocprofv3 --pmc runs of `bench.py --counters`
# (profiles/commit_step_log_n24_counters.md), keyed by log2 codeword length and
# the schedule (variant, max_k, tile_log, tab_min, tab_group); the counters cannot
# be read inside the timed run, so other schedules report null.
#  - per launch of the forward-NTT pass kernels: HBM bytes (FETCH_SIZE + WRITE_SIZE,
#    each scaled by the same counter on a 512 MiB device copy) and VALU issue rate
#    (SQ_INSTS_VALU per CU per clock), averaged over the passes of one forward NTT;
#  - per commit step: the VALU wave-instructions of all its kernels.
TRAFFIC = {(25, 3, 4, 9, 16, 1): {"bytes": 1082729497, "valu_issue_per_cu_clk": 0.726,
                                  "step_valu_wave_insts": 1051501888}}
# MI355X VALU issue peak: one wave64 instruction per CU per clock, 256 CUs, 2.4 GHz
VALU_PEAK_WAVE_INSTS_PER_S = 256 * 2.4e9
# LDS: a wave-wide ds_read_b128 occupies a CU's LDS for 4 clocks (256 B/clock;
# SQ_LDS_IDX_ACTIVE / SQ_INSTS_LDS = 4.0 in profiles/commit_step_log_n24_counters.md)
LDS_CLK_PER_LOOKUP = 4
LDS_CU_CLK_PER_S = 256 * 2.4e9


def traffic_entry(log_cw, tuning):
    key = (log_cw, tuning["variant"], tuning["max_k"], tuning["tile_log"], tuning["tab_min"], tuning["tab_group"])
    return TRAFFIC.get(key) or {}


def ntt_pass_kernels(b64, log_d, skip_early, skip_late):
    """Kernels of a forward NTT's passes as the library plans them (b64_ntt_plan)."""
    names = [p["kernel"] for p in b64.ntt_plan(log_d, skip_early, skip_late)]
    return " + ".join(f"{nm} x{names.count(nm)}" for nm in dict.fromkeys(names))


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--log-n", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-jobs leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-proofs", action="store_true", help="skip the batched and served FRI proofs leg")
    ap.add_argument("--no-ring-switch", action="store_true", help="skip the ring-switch folds leg")
    ap.add_argument("--no-pcs", action="store_true", help="skip the PCS evaluation-proof leg")
    ap.add_argument("--no-vision", action="store_true", help="skip the Vision Merkle-commit leg")
    ap.add_argument("--ntt", default=None, metavar="V,K,T,M,G[,W]",
                    help="NTT schedule for every size in this run (replaces the autotuned table): "
                         "variant, max_k, tile_log, tab_min, tab_group[, waves]")
    ap.add_argument("--counters", action="store_true",
                    help="hardware-counter workload: one encode, one forward NTT of the codeword and a "
                         "512 MiB device copy (the byte-counter calibration), then exit")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is a design note, not a measured path)"}))
        sys.exit(2)
    import numpy as np
    import torch

    import binius64_amd as b64

    L = b64.lib()
    if b64.device_count() == 0:
        raise SystemExit("no HIP device")
    if args.ntt:
        b64.config_set("ntt.clear_sizes", 0)
        for key, v in zip(("ntt.variant", "ntt.max_k", "ntt.tile_log", "ntt.tab_min", "ntt.tab_group", "ntt.waves"),
                          args.ntt.split(",")):
            b64.config_set(key, int(v))
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    sp = C.c_void_p(stream.cuda_stream)

    n = args.log_n
    params, dom = b64.setup_commit_params(n, 1, 96)
    lbs, log_dim, lir = params.log_batch_size, params.log_dim, params.log_inv_rate
    rs_log_len = params.rs_log_len()
    log_cw = rs_log_len + lbs
    g = torch.Generator(device=dev)
    g.manual_seed(24)
    msg = torch.randint(-(2 ** 63), 2 ** 63 - 1, (2 << n,), dtype=torch.int64, device=dev, generator=g)
    cw = torch.empty((2 << log_cw,), dtype=torch.int64, device=dev)
    tree = torch.empty((4 * ((2 << rs_log_len) - 1),), dtype=torch.int64, device=dev)
    folded = torch.empty((2 << rs_log_len,), dtype=torch.int64, device=dev)
    chal = np.random.default_rng(4).integers(0, 2 ** 63, size=(lbs, 2), dtype=np.uint64)
    chal_p = chal.ctypes.data_as(C.POINTER(C.c_uint64))

    pc = params.to_c()

    if args.counters:
        # one whole commit step (its kernels in launch order), one forward NTT of
        # the codeword, then the byte-counter calibration copy
        b64.check(L.b64_fri_commit_fold(C.byref(pc), dom.ptr, msg.data_ptr(), cw.data_ptr(), tree.data_ptr(),
                                        chal_p, folded.data_ptr(), None, sp), "commit_fold")
        torch.cuda.synchronize()
        b64.check(L.b64_ntt_forward(dom.ptr, cw.data_ptr(), log_cw, lir, lbs, sp), "ntt_forward")
        torch.cuda.synchronize()
        cal = torch.empty((1 << 26,), dtype=torch.int64, device=dev)
        cal.copy_(cw[: 1 << 26])
        torch.cuda.synchronize()
        print(json.dumps({"counters_workload": {"log_n": n, "launches": "commit step, then ntt_forward",
                                                "ntt_passes": [p["kernel"] for p in b64.ntt_plan(log_cw, lir, lbs)],
                                                "copy_bytes": 8 << 26}}), flush=True)
        return

    def step():
        # encode, then one pass over the codeword for leaf digests + the folded
        # oracle, then the inner tree levels (b64_fri_commit_fold)
        b64.check(L.b64_fri_commit_fold(C.byref(pc), dom.ptr, msg.data_ptr(), cw.data_ptr(), tree.data_ptr(),
                                        chal_p, folded.data_ptr(), None, sp), "commit_fold")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    t0.record(stream)
    for _ in range(args.steps):
        step()
    t1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms = t0.elapsed_time(t1) / args.steps
    msg_bytes = 16 << n
    value = msg_bytes / (ms * 1e-3) / 1e9

    # ---- per-stage timing (HIP events on the same stream) -------------------
    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    scratch = cw.clone()
    n_layers = rs_log_len - lir  # layers lir .. log_cw - lbs
    tuning = {k: b64.config_get(f"ntt.{k}:{log_cw}")
              for k in ("variant", "max_k", "tile_log", "tab_min", "tab_group", "waves")}
    n_passes = len(b64.ntt_plan(log_cw, lir, lbs))
    kernel = ntt_pass_kernels(b64, log_cw, lir, lbs)
    stages = {
        "encode_ms": timed(lambda: b64.check(L.b64_rs_encode_batch(dom.ptr, log_dim, lir, msg.data_ptr(), lbs,
                                                                   cw.data_ptr(), sp))),
        "ntt_ms": timed(lambda: b64.check(L.b64_ntt_forward(dom.ptr, scratch.data_ptr(), log_cw, lir, lbs, sp))),
        "merkle_ms": timed(lambda: b64.check(L.b64_merkle_commit(cw.data_ptr(), 1 << log_cw, 1 << lbs,
                                                                 tree.data_ptr(), None, sp))),
        "fold_ms": timed(lambda: b64.check(L.b64_fri_fold_interleaved(cw.data_ptr(), rs_log_len, lbs, chal_p,
                                                                      folded.data_ptr(), sp))),
    }
    ntt_pass_ms = stages["ntt_ms"] / n_passes
    pass_bytes = 2 * (16 << log_cw)  # each pass reads and writes the codeword once
    achieved = pass_bytes / (ntt_pass_ms * 1e-3) / 1e9
    muls = n_layers * (1 << (log_cw - 1))  # one multiply per butterfly
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": traffic_entry(log_cw, tuning).get("bytes"),
                "valu_issue_per_cu_clk": traffic_entry(log_cw, tuning).get("valu_issue_per_cu_clk"),
                "kernel": kernel, "ntt_tuning": tuning, "launches_per_ntt": n_passes,
                "avg_launch_ms": round(ntt_pass_ms, 4),
                "algorithmic_bytes_per_launch": pass_bytes,
                "gf_mul_per_s": float(f"{muls / (stages['ntt_ms'] * 1e-3):.4g}")}
    # LDS bound of the NTT passes: with schedule variants 3-4 every product goes
    # through a 4-bit window table, 32 lookups per product, each a wave-wide
    # ds_read_b128 that holds a CU's LDS for 4 clocks (256 bytes per clock)
    lds_roofline = None
    if tuning["variant"] >= 3:
        lookups = muls * 32 // 64 // n_passes  # wave-wide lookups per launch
        floor_ms = lookups * LDS_CLK_PER_LOOKUP / LDS_CU_CLK_PER_S * 1e3
        lds_roofline = {"bound": "lds", "wave_lookups_per_launch": lookups, "clk_per_lookup": LDS_CLK_PER_LOOKUP,
                        "floor_ms": round(floor_ms, 4), "frac": round(floor_ms / ntt_pass_ms, 4)}
    step_bytes = 58 * (1 << n)  # algorithmic HBM bytes of the whole step
    path_achieved = step_bytes / (ms * 1e-3) / 1e9
    step_valu = traffic_entry(log_cw, tuning).get("step_valu_wave_insts")
    valu_roofline = {"bound": "valu", "wave_insts_per_step": step_valu,
                     "peak_wave_insts_per_s": VALU_PEAK_WAVE_INSTS_PER_S,
                     "floor_ms": round(step_valu / VALU_PEAK_WAVE_INSTS_PER_S * 1e3, 4) if step_valu else None,
                     "frac": round(step_valu / VALU_PEAK_WAVE_INSTS_PER_S / (ms * 1e-3), 4) if step_valu else None}

    batch = None if args.no_batch else batch_bench(b64, torch, stream, sp)
    serve = None if args.no_serve else serve_bench(b64, torch, sp)
    proofs = None if args.no_proofs else proofs_bench(b64, torch, sp)
    ring = None if args.no_ring_switch else ring_switch_bench(b64, torch, stream, sp, with_cpu=not args.no_cpu)
    pcs = None if args.no_pcs else pcs_bench(b64, torch, sp, log_n=n, with_cpu=not args.no_cpu)
    vision = None if args.no_vision else vision_bench(b64, torch, stream, sp, with_cpu=not args.no_cpu)

    cpu = None
    if not args.no_cpu:
        import numpy as np
        step()
        torch.cuda.synchronize()
        msg_host = msg.cpu().numpy().view(np.uint64).reshape(-1, 2)
        gpu_root = tree.cpu().numpy().view(np.uint8)[-32:].tobytes()
        gpu_folded = folded.cpu().numpy().view(np.uint64).reshape(-1, 2)
        cpu = cpu_baseline(dom.subspace(rs_log_len), log_dim, lir, lbs, msg_host, chal, gpu_root, gpu_folded)

    out = {
        "metric": "RS-encode+commit GB/s (B128, log_n=24, rate 1/2)",
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "b128",
        "data": "synthetic (uniform random B128 message, seeded)",
        "config": {"workload": f"encode+merkle+interleaved-fold, log_n={n}, rate 1/2, batch 2^{lbs}",
                   "log_n": n, "log_inv_rate": lir, "log_batch_size": lbs, "rs_log_len": rs_log_len,
                   "fold_arities": params.fold_arities, "n_test_queries": params.n_test_queries},
        "roofline": roofline,
        "path_hbm": {"algorithmic_bytes_per_step": step_bytes, "achieved_GBs": round(path_achieved, 2),
                     "frac": round(path_achieved / HBM_PEAK_GBS, 5)},
        "valu_roofline": valu_roofline,
        "lds_roofline": lds_roofline,
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "wall_s_timed_region": round(wall, 4),
        "cpu_baseline": cpu,
        "batch": batch,
        "serve": serve,
        "proofs": proofs,
        "ring_switch": ring,
        "pcs_prove": pcs,
        "vision_merkle": vision,
    }
    print(json.dumps(out), flush=True)


def batch_bench(b64, torch, stream, sp, n_jobs: int = 64, reps: int = 5):
    """64 independent commit jobs, log_n ~ U{12..20} (seed 4), challenges seeded
    by job id: per-job plain launches vs one batched hipGraph replay."""
    import numpy as np

    rng = np.random.default_rng(4)
    sizes = [int(x) for x in rng.integers(12, 21, size=n_jobs)]
    jobs = []
    for jid, n in enumerate(sizes):
        r = np.random.default_rng(1000 + jid)
        msg = r.integers(0, 2 ** 63, size=(1 << n, 2), dtype=np.uint64)
        chal = np.random.default_rng(jid).integers(0, 2 ** 63, size=(n, 2), dtype=np.uint64)
        jobs.append(b64.CommitJob(n, msg, chal))

    def plain():
        for j in jobs:
            j.run(sp.value)

    plain()  # warm: domain setup and twiddle uploads happen here
    torch.cuda.synchronize()
    want = [j.roots() for j in jobs]
    t = time.perf_counter()
    for _ in range(reps):
        plain()
    t_plain = (time.perf_counter() - t) / reps
    plan = b64.BatchPlan(jobs, use_graph=True)
    plan.run(sp.value)
    same = [j.roots() for j in jobs] == want
    t = time.perf_counter()
    for _ in range(reps):
        plan.run(sp.value)
    t_batch = (time.perf_counter() - t) / reps
    same = same and [j.roots() for j in jobs] == want
    from binius64_amd._abi import TUNING_PATH
    return {"jobs": n_jobs, "log_n_range": [12, 20], "sizes_seed": 4, "n_launches_batched": plan.n_launches(),
            "ntt_schedule": "per-size autotuned table" if TUNING_PATH.exists() else "library defaults",
            "plain_jobs_per_s": round(n_jobs / t_plain, 1), "batched_jobs_per_s": round(n_jobs / t_batch, 1),
            "speedup": round(t_plain / t_batch, 3), "plain_ms": round(1e3 * t_plain, 3),
            "batched_ms": round(1e3 * t_batch, 3), "roots_identical": bool(same)}


def serve_bench(b64, torch, sp, n_jobs: int = 400, rate: float = 200.0, detail: bool = False):
    """Commit requests arriving as a Poisson stream (mean `rate`/s, seed 5) with
    witness sizes 2^14..2^22 64-bit words, log-uniform (n = 13..21 B128
    elements after packing two words per element), uniform random contents.

    Served: the continuous-batching server (b64_server_*), messages uploaded from
    pinned host memory, p50/p99 submit-to-roots latency.  Plain: the same jobs one
    at a time (upload + b64_job_run) under the same arrival times, run by one
    worker in arrival order, latency measured from arrival to roots.  Saturated
    throughput: every job submitted at once, vs the plain loop back to back."""
    import threading

    import numpy as np

    from binius64_amd._abi import JobDescC

    L = b64.lib()
    rng = np.random.default_rng(5)
    sizes = [int(w) - 1 for w in rng.integers(14, 23, size=n_jobs)]
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_jobs))
    pool_elems = 1 << 22
    g = torch.Generator().manual_seed(5)
    pool_t = torch.randint(-(2 ** 63), 2 ** 63 - 1, (2 * pool_elems,), dtype=torch.int64, generator=g)
    pool_t = pool_t.pin_memory()
    pool = pool_t.numpy().view(np.uint64).reshape(-1, 2)
    offs = [int(rng.integers(0, pool_elems - (1 << n) + 1)) for n in sizes]
    msgs = [pool[o:o + (1 << n)] for o, n in zip(offs, sizes)]
    chals = [np.random.default_rng(10_000 + i).integers(0, 2 ** 63, size=(n, 2), dtype=np.uint64)
             for i, n in enumerate(sizes)]
    n_roots = [b64.CommitJob.n_roots_for(n) for n in sizes]

    # ---- plain: one job at a time ------------------------------------------
    nmax = max(sizes)
    d_msg = b64.DeviceBuffer(16 << nmax)
    d_ws = b64.DeviceBuffer(b64.CommitJob.workspace_size(nmax))
    roots_plain = [np.zeros(32 * r, dtype=np.uint8) for r in n_roots]
    descs = [JobDescC(n, d_msg.ptr, chals[i].ctypes.data_as(C.POINTER(C.c_uint64)), d_ws.ptr,
                      roots_plain[i].ctypes.data_as(C.POINTER(C.c_uint8))) for i, n in enumerate(sizes)]

    def plain_one(i):
        b64.check(L.b64_memcpy_htod(d_msg.ptr, msgs[i].ctypes.data, msgs[i].nbytes, sp), "htod")
        b64.check(L.b64_job_run(C.byref(descs[i]), sp), "job_run")

    for n in sorted(set(sizes)):  # warm every size (domain and plan tables)
        plain_one(sizes.index(n))
    t_all = time.perf_counter()
    for i in range(n_jobs):
        plain_one(i)
    t_plain = time.perf_counter() - t_all
    want = [r.tobytes() for r in roots_plain]

    # plain under the same arrivals: a submitter thread queues each job at its
    # arrival time, one worker runs them one at a time in arrival order
    import queue
    fifo, lat_plain = queue.Queue(), np.zeros(n_jobs)

    def plain_worker():
        for _ in range(n_jobs):
            i, t_arr = fifo.get()
            plain_one(i)
            lat_plain[i] = time.perf_counter() - t_arr

    worker = threading.Thread(target=plain_worker)
    worker.start()
    t0 = time.perf_counter()
    for i in range(n_jobs):
        dt = t0 + arrivals[i] - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        fifo.put((i, time.perf_counter()))
    worker.join()
    same_plain = [r.tobytes() for r in roots_plain] == want

    def joined(rs):
        return b"".join(rs)

    # ---- served: Poisson arrivals --------------------------------------------
    srv = b64.Server(hbm_budget_bytes=8 << 30, max_batch=64, n_streams=3)
    try:
        # warm-up (untimed): three jobs of each size, then every size at once (the
        # multi-job kernels of mixed batches), so first-launch costs (code-object
        # loads, first use of a stream slot) stay out of the latencies
        for n in sorted(set(sizes)):
            i = sizes.index(n)
            for t in [srv.submit(n, msgs[i], chals[i]) for _ in range(3)]:
                srv.wait(t, n_roots[i])
        for _ in range(2):
            burst = [(sizes.index(n), srv.submit(n, msgs[sizes.index(n)], chals[sizes.index(n)]))
                     for n in sorted(set(sizes)) for _ in range(3)]
            for i, t in burst:
                srv.wait(t, n_roots[i])
        tickets = [0] * n_jobs

        def submitter():
            t0 = time.perf_counter()
            for i in range(n_jobs):
                dt = t0 + arrivals[i] - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                tickets[i] = srv.submit(sizes[i], msgs[i], chals[i])
                ready[i].set()

        ready = [threading.Event() for _ in range(n_jobs)]
        t0 = time.perf_counter()
        th = threading.Thread(target=submitter)
        th.start()
        lat, same = np.zeros(n_jobs), True
        for i in range(n_jobs):
            ready[i].wait()
            roots, lat[i] = srv.wait(tickets[i], n_roots[i])
            same = same and joined(roots) == want[i]
        t_poisson = time.perf_counter() - t0
        th.join()
        st_poisson = srv.stats()

        # ---- served: saturated (all jobs queued at once) ---------------------
        t0 = time.perf_counter()
        tk = [srv.submit(sizes[i], msgs[i], chals[i]) for i in range(n_jobs)]
        for i in range(n_jobs):
            roots, _ = srv.wait(tk[i], n_roots[i])
            same = same and joined(roots) == want[i]
        t_sat = time.perf_counter() - t0
        st = srv.stats()
    finally:
        srv.close()

    def pct(a, q):
        return round(1e3 * float(np.percentile(a, q)), 3)

    extra = {}
    if detail:
        worst = np.argsort(lat)[-8:]
        extra = {"worst_served": [(sizes[i], round(1e3 * lat[i], 3), round(1e3 * lat_plain[i], 3),
                                   round(float(arrivals[i]), 4)) for i in worst]}
    return {**extra, "jobs": n_jobs, "arrival_rate_per_s": rate, "seed": 5, "log_n_range": [13, 21],
            "hbm_budget_bytes": 8 << 30, "max_batch": 64, "n_streams": 3,
            "served_p50_ms": pct(lat, 50), "served_p99_ms": pct(lat, 99),
            "served_jobs_per_s_at_rate": round(n_jobs / t_poisson, 1),
            "plain_p50_ms": pct(lat_plain, 50), "plain_p99_ms": pct(lat_plain, 99),
            "saturated_served_jobs_per_s": round(n_jobs / t_sat, 1),
            "plain_jobs_per_s": round(n_jobs / t_plain, 1),
            "saturated_speedup": round(t_plain / t_sat, 3),
            "batches_at_rate": st_poisson["batches"], "batches_total": st["batches"],
            "peak_pool_bytes": st["peak_pool_bytes"], "roots_identical": bool(same and same_plain),
            "plain_latency": "measured: the same arrivals, one job at a time in arrival order"}


def proofs_bench(b64, torch, sp, n_batch: int = 64, reps: int = 3, n_serve: int = 400, rate: float = 200.0):
    """Whole FRI proofs (commitment, fold rounds with challenges sampled from each
    job's transcript, query phase) of the batching and serving workloads.

    Batch: the 64-job mix (log_n ~ U{12..20}, seed 4), b64_batch_prove over all
    jobs vs the same prover one job at a time.  Serve: the seed-5 Poisson stream
    (n = 13..21, mean `rate`/s) as proof jobs through the server vs one job at a
    time under the same arrivals; p50/p99 submit-to-proof latency, and saturated
    throughput.  Every tape is compared with the one-job tape of the same job."""
    import threading

    import numpy as np

    from binius64_amd._abi import ProveJobC

    L = b64.lib()
    new_t = lambda: b64.ProverTranscript()

    def prove(jobs_in, transcripts):
        descs = (ProveJobC * len(jobs_in))(*[ProveJobC(n, m.ptr, ws.ptr, t.ptr, None)
                                             for (n, m, ws), t in zip(jobs_in, transcripts)])
        b64.check(L.b64_batch_prove(descs, len(jobs_in), sp), "batch_prove")

    # the one-job prover of the reference's call sequence (fri/tests.rs:57-88):
    # b64_fri_commit_interleaved, the root written, FRIFoldProver rounds with
    # sampled challenges, finish_proof -- into preallocated codeword and tree
    setups = {}

    def one_job(n, d_msg, t, cw, tree):
        if n not in setups:
            setups[n] = b64.setup_commit_params(n, 1, 96)
        params, dom = setups[n]
        pc = params.to_c()
        root = np.zeros(32, dtype=np.uint8)
        rp = root.ctypes.data_as(C.POINTER(C.c_uint8))
        b64.check(L.b64_fri_commit_interleaved(C.byref(pc), dom.ptr, d_msg.ptr, cw.ptr, tree.ptr, rp, sp), "commit")
        fp = C.c_void_p()
        b64.check(L.b64_fri_prover_create(C.byref(pc), dom.ptr, cw.ptr, tree.ptr, sp, C.byref(fp)), "prover")
        L.b64_transcript_message(t.ptr, rp, 32)
        ch = np.zeros(2, dtype=np.uint64)
        chp = ch.ctypes.data_as(C.POINTER(C.c_uint64))
        chb = ch.view(np.uint8).ctypes.data_as(C.POINTER(C.c_uint8))
        committed = C.c_int(0)
        for _ in range(params.n_fold_rounds()):
            L.b64_transcript_sample(t.ptr, chb, 16)
            b64.check(L.b64_fri_prover_receive_challenge(fp, chp), "receive")
            b64.check(L.b64_fri_prover_execute_fold_round(fp, rp, C.byref(committed)), "fold_round")
            if committed.value:
                L.b64_transcript_message(t.ptr, rp, 32)
        b64.check(L.b64_fri_prover_finish_proof(fp, t.ptr), "finish_proof")
        L.b64_fri_prover_destroy(fp)

    # ---- batch ------------------------------------------------------------------
    rng = np.random.default_rng(4)
    sizes = [int(x) for x in rng.integers(12, 21, size=n_batch)]
    jobs = []
    for jid, n in enumerate(sizes):
        msg = np.random.default_rng(1000 + jid).integers(0, 2 ** 63, size=(1 << n, 2), dtype=np.uint64)
        jobs.append((n, b64.DeviceBuffer.from_host(msg), b64.DeviceBuffer(b64.CommitJob.workspace_size(n))))
    ts = [new_t() for _ in jobs]
    for j, t in zip(jobs, ts):  # warm every size; these tapes are the reference
        prove([j], [t])
    want = [t.finalize() for t in ts]
    t0 = time.perf_counter()
    for _ in range(reps):
        ts = [new_t() for _ in jobs]
        for j, t in zip(jobs, ts):
            prove([j], [t])
    t_plain = (time.perf_counter() - t0) / reps
    same = [t.finalize() for t in ts] == want
    nb = max(sizes)  # codeword 2^(nb+1) elements, tree over 2^(nb-3) leaves
    cw1 = b64.DeviceBuffer(16 << (nb + 1))
    tree1 = b64.DeviceBuffer(32 * ((2 << (nb - 3)) - 1))
    for (n, m, _), t in zip(jobs, [new_t() for _ in jobs]):  # warm
        one_job(n, m, t, cw1, tree1)
    t0 = time.perf_counter()
    for _ in range(reps):
        ts = [new_t() for _ in jobs]
        for (n, m, _), t in zip(jobs, ts):
            one_job(n, m, t, cw1, tree1)
    t_one = (time.perf_counter() - t0) / reps
    same = same and [t.finalize() for t in ts] == want
    prove(jobs, [new_t() for _ in jobs])
    t0 = time.perf_counter()
    for _ in range(reps):
        ts = [new_t() for _ in jobs]
        prove(jobs, ts)
    t_batch = (time.perf_counter() - t0) / reps
    same = same and [t.finalize() for t in ts] == want
    t_best = min(t_plain, t_one)
    batch = {"jobs": n_batch, "log_n_range": [12, 20], "sizes_seed": 4,
             "plain_proofs_per_s": round(n_batch / t_best, 1), "batched_proofs_per_s": round(n_batch / t_batch, 1),
             "speedup": round(t_best / t_batch, 3), "plain_ms": round(1e3 * t_best, 3),
             "batched_ms": round(1e3 * t_batch, 3),
             "plain_paths_ms": {"one_job_prover": round(1e3 * t_one, 3),
                                "batch_prove_one_job_at_a_time": round(1e3 * t_plain, 3)},
             "tapes_identical": bool(same)}

    # ---- serve ------------------------------------------------------------------
    rng = np.random.default_rng(5)
    ssizes = [int(w) - 1 for w in rng.integers(14, 23, size=n_serve)]
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_serve))
    pool_elems = 1 << 22
    g = torch.Generator().manual_seed(5)
    pool_t = torch.randint(-(2 ** 63), 2 ** 63 - 1, (2 * pool_elems,), dtype=torch.int64, generator=g).pin_memory()
    pool = pool_t.numpy().view(np.uint64).reshape(-1, 2)
    offs = [int(rng.integers(0, pool_elems - (1 << n) + 1)) for n in ssizes]
    msgs = [pool[o:o + (1 << n)] for o, n in zip(offs, ssizes)]
    n_roots = [b64.CommitJob.n_roots_for(n) for n in ssizes]
    nmax = max(ssizes)
    d_msg = b64.DeviceBuffer(16 << nmax)
    d_ws = b64.DeviceBuffer(b64.CommitJob.workspace_size(nmax))

    use_one_job = t_one < t_plain  # the faster one-at-a-time path of the batch figures
    if use_one_job:
        cw1 = b64.DeviceBuffer(16 << (nmax + 1))
        tree1 = b64.DeviceBuffer(32 * ((2 << (nmax - 3)) - 1))

    def plain_one(i, t):
        b64.check(L.b64_memcpy_htod(d_msg.ptr, msgs[i].ctypes.data, msgs[i].nbytes, sp), "htod")
        if use_one_job:
            one_job(ssizes[i], d_msg, t, cw1, tree1)
        else:
            prove([(ssizes[i], d_msg, d_ws)], [t])

    for n in sorted(set(ssizes)):  # warm every size
        plain_one(ssizes.index(n), new_t())
    t0 = time.perf_counter()
    tp = [new_t() for _ in range(n_serve)]
    for i in range(n_serve):
        plain_one(i, tp[i])
    t_plain_s = time.perf_counter() - t0
    want_s = [t.finalize() for t in tp]

    import queue
    fifo, lat_plain = queue.Queue(), np.zeros(n_serve)

    def plain_worker():
        for _ in range(n_serve):
            i, t_arr = fifo.get()
            plain_one(i, new_t())
            lat_plain[i] = time.perf_counter() - t_arr

    worker = threading.Thread(target=plain_worker)
    worker.start()
    t0 = time.perf_counter()
    for i in range(n_serve):
        dt = t0 + arrivals[i] - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        fifo.put((i, time.perf_counter()))
    worker.join()

    srv = b64.Server(hbm_budget_bytes=8 << 30, max_batch=64, n_streams=3)
    same_s = True
    try:
        # warm-up (untimed): three proofs of each size, then every size at once
        # twice (bursts make the slots' buffers reach their working size)
        for n in sorted(set(ssizes)):
            i = ssizes.index(n)
            for tk in [srv.submit_proof(n, msgs[i], new_t()) for _ in range(3)]:
                srv.wait(tk, n_roots[i])
        for _ in range(2):
            burst = [(ssizes.index(n), srv.submit_proof(n, msgs[ssizes.index(n)], new_t()))
                     for n in sorted(set(ssizes)) for _ in range(4)]
            for i, tk in burst:
                srv.wait(tk, n_roots[i])
        tickets, tr = [0] * n_serve, [new_t() for _ in range(n_serve)]
        ready = [threading.Event() for _ in range(n_serve)]

        def submitter():
            t0 = time.perf_counter()
            for i in range(n_serve):
                dt = t0 + arrivals[i] - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                tickets[i] = srv.submit_proof(ssizes[i], msgs[i], tr[i])
                ready[i].set()

        t0 = time.perf_counter()
        th = threading.Thread(target=submitter)
        th.start()
        lat = np.zeros(n_serve)
        for i in range(n_serve):
            ready[i].wait()
            _, lat[i] = srv.wait(tickets[i], n_roots[i])
        t_poisson = time.perf_counter() - t0
        th.join()
        same_s = [t.finalize() for t in tr] == want_s
        t0 = time.perf_counter()
        tr = [new_t() for _ in range(n_serve)]
        tk = [srv.submit_proof(ssizes[i], msgs[i], tr[i]) for i in range(n_serve)]
        for i in range(n_serve):
            srv.wait(tk[i], n_roots[i])
        t_sat = time.perf_counter() - t0
        same_s = same_s and [t.finalize() for t in tr] == want_s
    finally:
        srv.close()

    def pct(a, q):
        return round(1e3 * float(np.percentile(a, q)), 3)

    serve = {"jobs": n_serve, "arrival_rate_per_s": rate, "seed": 5, "log_n_range": [13, 21],
             "served_p50_ms": pct(lat, 50), "served_p99_ms": pct(lat, 99),
             "served_proofs_per_s_at_rate": round(n_serve / t_poisson, 1),
             "plain_p50_ms": pct(lat_plain, 50), "plain_p99_ms": pct(lat_plain, 99),
             "saturated_served_proofs_per_s": round(n_serve / t_sat, 1),
             "plain_proofs_per_s": round(n_serve / t_plain_s, 1),
             "saturated_speedup": round(t_plain_s / t_sat, 3), "tapes_identical": bool(same_s),
             "plain_path": "one-job prover" if use_one_job else "batch_prove, one job at a time",
             "plain_latency": "measured: the same arrivals, one proof at a time in arrival order"}
    return {"batch": batch, "serve": serve}


def ring_switch_bench(b64, torch, stream, sp, log_len: int = 24, reps: int = 10, with_cpu: bool = True):
    """The ring-switching step of the PCS opening (prover/prover/src/pcs.rs:127-153) at
    2^24 rows: fold_1b_rows_for_b128(packed witness, suffix tensor) and
    fold_b128_elems_inplace(suffix tensor, 128-element eq tensor).  Throughput is
    the reference bench's convention (benches/ring_switch.rs): 16 * 2^log_len bytes."""
    import numpy as np

    L = b64.lib()
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    n = 1 << log_len
    mat = torch.randint(-(2 ** 63), 2 ** 63 - 1, (2 * n,), dtype=torch.int64, device=dev, generator=g)
    vec = torch.randint(-(2 ** 63), 2 ** 63 - 1, (2 * n,), dtype=torch.int64, device=dev, generator=g)
    out = torch.empty((256,), dtype=torch.int64, device=dev)
    w = np.random.default_rng(8).integers(0, 2 ** 63, size=(128, 2), dtype=np.uint64)
    wp = w.ctypes.data_as(C.POINTER(C.c_uint64))

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    t1 = timed(lambda: b64.check(L.b64_ring_switch_fold_1b_rows(mat.data_ptr(), vec.data_ptr(), log_len,
                                                                out.data_ptr(), sp), "fold_1b_rows"))
    t2 = timed(lambda: b64.check(L.b64_ring_switch_fold_elems(vec.data_ptr(), log_len, wp, sp), "fold_elems"))
    res = {"log_len": log_len,
           "fold_1b_rows_ms": round(t1, 4), "fold_1b_rows_GBs": round((16 << log_len) / (t1 * 1e-3) / 1e9, 2),
           "fold_1b_rows_hbm_frac": round((32 << log_len) / (t1 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "fold_elems_ms": round(t2, 4), "fold_elems_GBs": round((16 << log_len) / (t2 * 1e-3) / 1e9, 2),
           "fold_elems_hbm_frac": round((32 << log_len) / (t2 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    if with_cpu:
        from oracle import oracle as orc
        cores = cpu_share()
        m_h = mat.cpu().numpy().view(np.uint64).reshape(-1, 2)
        v_h = vec.cpu().numpy().view(np.uint64).reshape(-1, 2)
        ts = []
        for _ in range(3):
            t = time.perf_counter()
            orc.rs_fold_1b_rows(m_h, v_h, 0, threads=cores)
            ts.append(time.perf_counter() - t)
        ts2 = []
        for _ in range(3):
            t = time.perf_counter()
            orc.rs_fold_elems(v_h, w, 0, threads=cores)
            ts2.append(time.perf_counter() - t)
        res["cpu_baseline"] = {"cores": cores, "kind": "port", "sample": f"same 2^{log_len} inputs, median of 3",
                               "fold_1b_rows_GBs": round((16 << log_len) / sorted(ts)[1] / 1e9, 3),
                               "fold_elems_GBs": round((16 << log_len) / sorted(ts2)[1] / 1e9, 3)}
    return res


def pcs_bench(b64, torch, sp, log_n: int = 24, reps: int = 3, with_cpu: bool = True):
    """The evaluation proof of the polynomial commitment (OneBitPCSProver::prove,
    prover/prover/src/pcs.rs:107-170) for a committed multilinear of 2^log_n B128
    elements (2^(log_n+7) bits) with the Verifier::setup parameters: ring-switch
    partial evaluations, log_n BaseFold rounds (sumcheck + FRI fold and
    commitment), 232 FRI queries.  Wall time per proof, transcript included; the
    commitment is made once beforehand."""
    import numpy as np

    L = b64.lib()
    params, dom = b64.setup_commit_params(log_n, 1, 96)
    pc = params.to_c()
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(11)
    msg = torch.randint(-(2 ** 63), 2 ** 63 - 1, (2 << log_n,), dtype=torch.int64, device=dev, generator=g)
    packed = torch.empty_like(msg)
    cw = torch.empty((2 << params.log_len(),), dtype=torch.int64, device=dev)
    n_nodes = (2 << params.rs_log_len()) - 1
    tree = torch.empty((4 * n_nodes,), dtype=torch.int64, device=dev)
    root = np.zeros(32, dtype=np.uint8)
    point = np.random.default_rng(12).integers(0, 2 ** 63, size=(log_n + 7, 2), dtype=np.uint64)
    pp = point.ctypes.data_as(C.POINTER(C.c_uint64))

    def commit():
        b64.check(L.b64_fri_commit_interleaved(C.byref(pc), dom.ptr, msg.data_ptr(), cw.data_ptr(), tree.data_ptr(),
                                               b64._p8(root), sp), "commit")

    commit()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        commit()
    torch.cuda.synchronize()
    commit_ms = 1e3 * (time.perf_counter() - t) / reps

    def prove():
        packed.copy_(msg)  # the proof consumes its multilinear
        torch.cuda.synchronize()
        tr = b64.ProverTranscript()
        tr.message(root.tobytes())
        t0 = time.perf_counter()
        b64.check(L.b64_pcs_prove(C.byref(pc), dom.ptr, cw.data_ptr(), tree.data_ptr(), packed.data_ptr(), pp,
                                  log_n + 7, tr.ptr, sp), "pcs_prove")
        torch.cuda.synchronize()
        return time.perf_counter() - t0, tr.finalize()

    _, tape = prove()
    runs = [prove() for _ in range(reps)]
    same = all(r[1] == tape for r in runs)
    med = sorted(r[0] for r in runs)[reps // 2]
    res = {"log_n": log_n, "proof_ms": round(1e3 * med, 3), "proofs_per_s": round(1 / med, 2),
           "commit_ms": round(commit_ms, 3), "tape_bytes": len(tape), "deterministic": bool(same)}
    if with_cpu:
        from oracle import oracle as orc
        cores = cpu_share()
        m_h = msg.cpu().numpy().view(np.uint64).reshape(-1, 2)
        cw_h = cw.cpu().numpy().view(np.uint64).reshape(-1, 2)
        tree_h = tree.cpu().numpy().view(np.uint8).reshape(n_nodes, 32)
        ref = orc.Domain(dom.subspace(params.rs_log_len()))
        ot = orc.Transcript()
        ot.message(root.tobytes())
        t = time.perf_counter()
        orc.pcs_prove(ref, m_h, point, cw_h, tree_h, params.log_dim, 1, params.log_batch_size, params.fold_arities,
                      params.n_test_queries, ot, threads=cores)
        cpu_s = time.perf_counter() - t
        res["cpu_baseline"] = {"proof_s": round(cpu_s, 3), "cores": cores, "kind": "port",
                               "sample": f"same 2^{log_n}-element multilinear and point, one proof"}
        res["tape_identical_to_cpu"] = ot.tape() == tape
    return res


def vision_bench(b64, torch, stream, sp, log_leaves: int = 17, reps: int = 3, with_cpu: bool = True,
                 cpu_log_leaves: int = 13):
    """Vision Merkle commits at the shape of the reference's Merkle bench
    (prover/prover/benches/binary_merkle_tree.rs: 2^17 leaves of 16 B128 elements):
    Vision-6 leaves with Vision-4 compression (that bench's hasher), and Vision-4
    throughout (setup_vision4, prover/examples/src/lib.rs).  Throughput is the
    bench's convention: committed bytes per second."""
    import numpy as np

    L = b64.lib()
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(13)
    n_elems = 16 << log_leaves
    data = torch.randint(-(2 ** 63), 2 ** 63 - 1, (2 * n_elems,), dtype=torch.int64, device=dev, generator=g)
    tree = torch.empty((4 * ((2 << log_leaves) - 1),), dtype=torch.int64, device=dev)
    res = {"log_leaves": log_leaves, "leaf_elems": 16}
    for leaf_m in (6, 4):
        def run():
            b64.check(L.b64_merkle_commit_vision(leaf_m, data.data_ptr(), n_elems, 16, tree.data_ptr(), None, sp),
                      "merkle_commit_vision")
        run()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            run()
        b.record(stream)
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / reps
        res[f"vision{leaf_m}_leaves_ms"] = round(ms, 3)
        res[f"vision{leaf_m}_leaves_MBs"] = round((16 * n_elems) / (ms * 1e-3) / 1e6, 2)
    if with_cpu:
        import json as _json

        from oracle import oracle as orc
        gold = _json.loads((ROOT / "tests" / "golden" / "reference_vectors.json").read_text())
        v4, v6 = orc.Vision(4, gold["vision4"]), orc.Vision(6, gold["vision6"])
        cores = cpu_share()
        sample = data[: 2 * (16 << cpu_log_leaves)].cpu().numpy().view(np.uint64).reshape(-1, 2)
        t = time.perf_counter()
        orc.merkle_build_vision(v6, v4, sample, 16, threads=cores)
        cpu_s = time.perf_counter() - t
        res["cpu_baseline"] = {"vision6_leaves_MBs": round(16 * sample.shape[0] / cpu_s / 1e6, 3), "cores": cores,
                               "kind": "port", "sample": f"2^{cpu_log_leaves} leaves of the same data, one commit",
                               "seconds": round(cpu_s, 3)}
    return res


def cpu_share() -> int:
    """Host threads for the CPU legs: the CPUs this process may run on, capped at
    the 16 a one-GPU box allots to its GPU's jobs."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def cpu_baseline(basis, log_dim: int, lir: int, lbs: int, msg_host, chal, gpu_tree_root: bytes, gpu_folded,
                 reps: int = 5):
    """The reference's x86-64 commit path (oracle/libb64cpu.so: NeighborsLastMultiThread
    with 4-wide VPCLMULQDQ products over GenericPreExpanded twiddles, SHA-256 leaves
    and Sha256Compression nodes with the SHA extensions, packed fold_interleaved, a
    persistent pool) on the same message as the GPU step, with every host thread of
    the box's share and with one thread; its root and folded oracle must equal the
    GPU's."""
    import numpy as np

    from oracle import cpu_baseline as cbl

    cores = cpu_share()
    n_cw = 1 << (log_dim + lir + lbs)
    cw = np.zeros((n_cw, 2), dtype=np.uint64)
    tree = np.zeros((2 * (n_cw >> lbs) - 1, 32), dtype=np.uint8)
    folded = np.zeros((n_cw >> lbs, 2), dtype=np.uint64)
    msg_bytes = 16 * msg_host.shape[0]

    def timed(threads, n_reps):
        base = cbl.Baseline(basis, threads)
        base.commit_step(msg_host, log_dim, lir, lbs, chal, cw, tree, folded)  # warm-up (page faults, pool)
        ts = []
        for _ in range(n_reps):
            t = time.perf_counter()
            base.commit_step(msg_host, log_dim, lir, lbs, chal, cw, tree, folded)
            ts.append(time.perf_counter() - t)
        base.close()
        return sorted(ts)[len(ts) // 2]

    med = timed(cores, reps)
    same = tree[-1].tobytes() == gpu_tree_root and np.array_equal(folded, gpu_folded)
    single = timed(1, 1)
    return {"value": round(msg_bytes / med / 1e9, 4), "unit": "GB/s", "cores": cores, "kind": "port",
            "sample": f"the full step (2^{log_dim + lbs} B128 message, encode+merkle+fold) on the GPU step's "
                      f"message, median of {reps} after a warm-up",
            "seconds": round(med, 4), "host_cpus": os.cpu_count(),
            "single_thread": {"value": round(msg_bytes / single / 1e9, 4), "unit": "GB/s", "seconds": round(single, 3)},
            "isa": cbl.features(), "identical_to_gpu": bool(same)}


if __name__ == "__main__":
    main()
