"""GPU tests of the serving loop: every result equals the job run alone, in any
mix, under a tight HBM budget (jobs wait for admission) and with several streams."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def b64():
    import binius64_amd as b
    if b.device_count() == 0:
        pytest.fail("no HIP device visible")
    return b


def job_inputs(orc, n, seed):
    return orc.random_elems(5000 + seed, 1 << n), orc.random_elems(6000 + seed, n)


def alone(b64, n, msg, chal):
    job = b64.CommitJob(n, msg, chal)
    job.run()
    return job.roots()


def test_mixed_submissions_equal_alone(b64, orc):
    rng = np.random.default_rng(5)
    sizes = [int(x) for x in rng.integers(9, 17, size=40)]
    inputs = [job_inputs(orc, n, i) for i, n in enumerate(sizes)]
    want = [alone(b64, n, m, c) for n, (m, c) in zip(sizes, inputs)]
    srv = b64.Server(hbm_budget_bytes=256 << 20, max_batch=16, n_streams=3)
    try:
        tickets = [srv.submit(n, m, c) for n, (m, c) in zip(sizes, inputs)]
        got = [srv.wait(t, b64.CommitJob.n_roots_for(n))[0] for t, n in zip(tickets, sizes)]
        st = srv.stats()
    finally:
        srv.close()
    assert got == want
    assert st["jobs_done"] == len(sizes) and st["batches"] >= 1


def test_tight_budget_admits_as_memory_frees(b64, orc):
    sizes = [16, 14, 16, 12, 15, 16, 13]
    ws = max(b64.CommitJob.workspace_size(n) + (16 << n) for n in sizes)
    srv = b64.Server(hbm_budget_bytes=int(ws * 1.5) + (1 << 20), max_batch=8, n_streams=2)
    inputs = [job_inputs(orc, n, 100 + i) for i, n in enumerate(sizes)]
    try:
        tickets = [srv.submit(n, m, c) for n, (m, c) in zip(sizes, inputs)]
        got = [srv.wait(t, b64.CommitJob.n_roots_for(n))[0] for t, n in zip(tickets, sizes)]
        st = srv.stats()
    finally:
        srv.close()
    assert got == [alone(b64, n, m, c) for n, (m, c) in zip(sizes, inputs)]
    assert st["peak_pool_bytes"] <= int(ws * 1.5) + (1 << 20)
    assert st["batches"] >= 3  # two 2^16 jobs never fit together


def test_oversized_job_rejected(b64, orc):
    srv = b64.Server(hbm_budget_bytes=1 << 20, max_batch=4, n_streams=1)
    try:
        m, c = job_inputs(orc, 16, 1)
        with pytest.raises(b64.B64Error) as e:
            srv.submit(16, m, c)
        assert e.value.code == 23
        # a small job still runs
        m, c = job_inputs(orc, 9, 2)
        t = srv.submit(9, m, c)
        assert srv.wait(t, b64.CommitJob.n_roots_for(9))[0] == alone(b64, 9, m, c)
    finally:
        srv.close()


def test_concurrent_submitters(b64, orc):
    sizes = [10, 12, 14, 11, 13, 12, 10, 15]
    inputs = [job_inputs(orc, n, 200 + i) for i, n in enumerate(sizes)]
    srv = b64.Server(hbm_budget_bytes=128 << 20, max_batch=4, n_streams=2)
    results = {}

    def client(i):
        n = sizes[i]
        t = srv.submit(n, *inputs[i])
        results[i] = srv.wait(t, b64.CommitJob.n_roots_for(n))[0]

    try:
        th = [threading.Thread(target=client, args=(i,)) for i in range(len(sizes))]
        for t in th:
            t.start()
        for t in th:
            t.join()
    finally:
        srv.close()
    for i, n in enumerate(sizes):
        assert results[i] == alone(b64, n, *inputs[i])


def test_config5_stream_matches_oracle(b64, orc):
    """The serving workload itself, first 48 requests of the seed-5 stream:
    witnesses of 2^14..2^22 64-bit words (log-uniform), packed two words per
    B128 element by pack_witness on the device (prover/prover/src/prove.rs:286-327,
    n = 13..21), submitted at their Poisson arrival times (mean 200/s).  Every
    job's roots equal the oracle's FRIFoldProver roots with the job's challenges
    (prover/prover/src/fri/fold.rs:117-175) and the roots of the job run alone."""
    import time

    n_stream, n_jobs = 400, 48
    rng = np.random.default_rng(5)
    log_words = [int(w) for w in rng.integers(14, 23, size=n_stream)][:n_jobs]
    arrivals = np.cumsum(rng.exponential(1.0 / 200.0, size=n_stream))[:n_jobs]
    sizes = [w - 1 for w in log_words]
    print("config-5 sizes", sizes, flush=True)
    msgs, chals = [], []
    for i, lw in enumerate(log_words):
        words = np.random.default_rng(20_000 + i).integers(0, 2 ** 64, size=1 << lw, dtype=np.uint64)
        packed = b64.pack_witness(b64.DeviceBuffer.from_host(words), 1 << lw, lw - 1).to_host(1 << (lw - 1))
        assert np.array_equal(packed, orc.pack_witness(words)), f"pack_witness job {i}"
        msgs.append(packed)
        chals.append(np.random.default_rng(10_000 + i).integers(0, 2 ** 63, size=(lw - 1, 2), dtype=np.uint64))
    srv = b64.Server(hbm_budget_bytes=8 << 30, max_batch=64, n_streams=3)
    try:
        tickets, t0 = [], time.perf_counter()
        for i, n in enumerate(sizes):
            dt = t0 + arrivals[i] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            tickets.append(srv.submit(n, msgs[i], chals[i]))
        got = [srv.wait(t, b64.CommitJob.n_roots_for(n))[0] for t, n in zip(tickets, sizes)]
        st = srv.stats()
    finally:
        srv.close()
    assert st["jobs_done"] == n_jobs
    for i, n in enumerate(sizes):
        p = orc.fri_params_setup(n, 1, 96)
        dom = orc.Domain(orc.Domain(dim=n + 1).subspace(p["rs_log_len"]))
        want = orc.fri_prove(dom, msgs[i], p["log_dim"], 1, p["log_batch_size"], p["fold_arities"], 1,
                             challenges=chals[i], threads=8)["roots"]
        assert got[i] == want, f"job {i} (n={n}) != oracle"
        assert got[i] == alone(b64, n, msgs[i], chals[i]), f"job {i} (n={n}) != alone"


def test_served_proofs_match_oracle(b64, orc):
    """Proof jobs served from the seed-5 stream (n = 13..21 from pack_witness
    sizes), mixed with commit jobs and submitted at their Poisson arrival times:
    every proof job's transcript tape equals the oracle's FRI prover tape for its
    message (prover/prover/src/fri/fold.rs:117-175, 221-244: challenges sampled
    from the transcript) and the tape of the same job proved alone; the commit
    jobs in between still give their own roots."""
    import time

    rng = np.random.default_rng(5)
    sizes = [int(w) - 1 for w in rng.integers(14, 23, size=400)][:24]
    arrivals = np.cumsum(rng.exponential(1.0 / 200.0, size=400))[:24]
    print("served proof sizes", sizes, flush=True)
    msgs = [np.random.default_rng(30_000 + i).integers(0, 2 ** 63, size=(1 << n, 2), dtype=np.uint64)
            for i, n in enumerate(sizes)]
    srv = b64.Server(hbm_budget_bytes=8 << 30, max_batch=16, n_streams=3)
    transcripts = [b64.ProverTranscript() for _ in sizes]
    try:
        tickets, commits, t0 = [], [], time.perf_counter()
        for i, n in enumerate(sizes):
            dt = t0 + arrivals[i] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            tickets.append(srv.submit_proof(n, msgs[i], transcripts[i]))
            if i % 3 == 0:  # a commit job between the proofs
                m, ch = job_inputs(orc, 10 + i % 5, 700 + i)
                commits.append((10 + i % 5, m, ch, srv.submit(10 + i % 5, m, ch)))
        got = [srv.wait(t, b64.CommitJob.n_roots_for(n))[0] for t, n in zip(tickets, sizes)]
        got_commits = [srv.wait(t, b64.CommitJob.n_roots_for(n))[0] for n, _, _, t in commits]
    finally:
        srv.close()
    for i, n in enumerate(sizes):
        p = orc.fri_params_setup(n, 1, 96)
        dom = orc.Domain(orc.Domain(dim=n + 1).subspace(p["rs_log_len"]))
        want = orc.fri_prove(dom, msgs[i], p["log_dim"], 1, p["log_batch_size"], p["fold_arities"],
                             p["n_test_queries"], threads=8)
        assert transcripts[i].finalize() == want["tape"], f"proof job {i} (n={n}) tape != oracle"
        assert got[i] == want["roots"], f"proof job {i} (n={n}) roots != oracle"
    for i in (0, 11, 23):
        alone_job = b64.ProveJob(sizes[i], msgs[i])
        b64.batch_prove([alone_job])
        assert alone_job.tape() == transcripts[i].finalize(), f"proof job {i} != alone"
    for (n, m, ch, _), r in zip(commits, got_commits):
        assert r == alone(b64, n, m, ch)


def test_tight_budget_proofs_and_commits(b64, orc):
    """Proof jobs and commit jobs under a budget that holds about one and a half
    of the largest job: admission waits for memory freed by proofs delivered one
    by one; every tape equals the same job proved alone, every commit its roots."""
    sizes = [16, 14, 16, 12, 15, 16, 13, 16]
    ws = max(b64.CommitJob.workspace_size(n) + (16 << n) for n in sizes)
    budget = int(ws * 1.5) + (1 << 20)
    srv = b64.Server(hbm_budget_bytes=budget, max_batch=8, n_streams=2)
    msgs = [orc.random_elems(800 + i, 1 << n) for i, n in enumerate(sizes)]
    ts = [b64.ProverTranscript() for _ in sizes]
    try:
        tickets, commits = [], []
        for i, n in enumerate(sizes):
            tickets.append(srv.submit_proof(n, msgs[i], ts[i]))
            if i % 2:
                m, c = job_inputs(orc, 11, 900 + i)
                commits.append((m, c, srv.submit(11, m, c)))
        got = [srv.wait(t, b64.CommitJob.n_roots_for(n))[0] for t, n in zip(tickets, sizes)]
        got_c = [srv.wait(t, b64.CommitJob.n_roots_for(11))[0] for _, _, t in commits]
        st = srv.stats()
    finally:
        srv.close()
    assert st["peak_pool_bytes"] <= budget
    for i, n in enumerate(sizes):
        job = b64.ProveJob(n, msgs[i])
        b64.batch_prove([job])
        assert ts[i].finalize() == job.tape(), f"proof job {i} (n={n})"
        assert got[i] == job.roots()
    for (m, c, _), r in zip(commits, got_c):
        assert r == alone(b64, 11, m, c)
