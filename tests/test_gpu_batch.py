"""GPU parity of commit jobs: the plain per-job path and the batched (hipGraph)
path give the oracle's roots and oracles, for uneven job sizes."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def b64():
    import binius64_amd as b
    if b.device_count() == 0:
        pytest.fail("no HIP device visible")
    return b


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_job(orc, n, msg, chal):
    p = orc.fri_params_setup(n, 1, 96)
    std = orc.Domain(dim=n + 1)
    dom = orc.Domain(std.subspace(p["rs_log_len"]))
    res = orc.fri_prove(dom, msg, p["log_dim"], 1, p["log_batch_size"], p["fold_arities"], 1,
                        challenges=chal, threads=8)
    return res


def make_job(b64, orc, n, seed):
    msg = orc.random_elems(1000 + seed, 1 << n)
    chal = orc.random_elems(2000 + seed, n)
    return b64.CommitJob(n, msg, chal), msg, chal


@pytest.mark.parametrize("n", [8, 9, 12, 13, 14, 16, 17])
def test_plain_job_matches_oracle(b64, orc, n):
    say("plain job", n)
    job, msg, chal = make_job(b64, orc, n, n)
    job.run()
    want = oracle_job(orc, n, msg, chal)
    assert job.roots() == want["roots"]
    # the terminal oracle equals the oracle's terminal codeword
    assert np.array_equal(job.oracle(job.n_roots - 1), want["terminal"])


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_equals_plain_uneven_sizes(b64, orc, use_graph):
    rng = np.random.default_rng(4)
    sizes = [int(x) for x in rng.integers(12, 17, size=12)] + [8, 20]
    say("batch sizes", sizes, "graph", use_graph)
    jobs = [make_job(b64, orc, n, i)[0] for i, n in enumerate(sizes)]
    plain = []
    for j in jobs:
        j.run()
        plain.append(j.roots())
    # clear the outputs, then run the same jobs batched
    for j in jobs:
        j.roots_buf[:] = 0
    plan = b64.BatchPlan(jobs, use_graph=use_graph)
    assert plan.n_launches() < 40
    for rep in range(3):  # replays give identical results
        for j in jobs:
            j.roots_buf[:] = 0
        plan.run()
        assert [j.roots() for j in jobs] == plain, rep


def test_batch_matches_oracle_small(b64, orc):
    sizes = [12, 14, 12, 15, 13]
    made = [make_job(b64, orc, n, 50 + i) for i, n in enumerate(sizes)]
    plan = b64.BatchPlan([m[0] for m in made], use_graph=True)
    plan.run()
    for (job, msg, chal), n in zip(made, sizes):
        assert job.roots() == oracle_job(orc, n, msg, chal)["roots"]


@pytest.mark.parametrize("sizes", [[13, 13], [14, 14, 14], [16, 17, 16, 17], [19, 19]])
def test_batch_uniform_arity_matches_oracle(b64, orc, sizes):
    """Batched coset folds whose jobs all fold the same number of challenges
    (arity 2, 3 and 4 at the first folded oracle, 4 at both for n = 19) run the
    per-arity multi-job kernel; the mixed slots the general one."""
    say("uniform-arity batch", sizes)
    made = [make_job(b64, orc, n, 90 + i) for i, n in enumerate(sizes)]
    plan = b64.BatchPlan([m[0] for m in made], use_graph=False)
    plan.run()
    for (job, msg, chal), n in zip(made, sizes):
        assert job.roots() == oracle_job(orc, n, msg, chal)["roots"]


@pytest.mark.parametrize("n,use_graph", [(13, False), (20, False), (13, True), (20, True)])
def test_single_job_plan(b64, orc, n, use_graph):
    """A plan of one job: without a graph it runs the job's single-job launchers
    (the per-size schedule, register-resident NTT passes at n = 20), with one
    the multi-job kernels; both give the oracle's roots, on every replay."""
    say("single-job plan", n, "graph", use_graph)
    job, msg, chal = make_job(b64, orc, n, 70 + n)
    want = oracle_job(orc, n, msg, chal)["roots"]
    plan = b64.BatchPlan([job], use_graph=use_graph)
    for rep in range(2):
        job.roots_buf[:] = 0
        plan.run()
        assert job.roots() == want, rep


def test_config4_mix_matches_oracle(b64, orc):
    """The batching workload itself: 64 jobs, log_n ~ U{12..20} drawn with seed 4,
    uniform random messages, FRI challenges seeded by job id -- one hipGraph
    plan; every job's roots equal the oracle's FRIFoldProver roots
    (prover/prover/src/fri/fold.rs:117-175 with those challenges) and the plain
    per-job launches'."""
    rng = np.random.default_rng(4)
    sizes = [int(x) for x in rng.integers(12, 21, size=64)]
    say("config-4 sizes", sizes)
    jobs, inputs = [], []
    for jid, n in enumerate(sizes):
        msg = np.random.default_rng(1000 + jid).integers(0, 2 ** 63, size=(1 << n, 2), dtype=np.uint64)
        chal = np.random.default_rng(jid).integers(0, 2 ** 63, size=(n, 2), dtype=np.uint64)
        jobs.append(b64.CommitJob(n, msg, chal))
        inputs.append((msg, chal))
    plan = b64.BatchPlan(jobs, use_graph=True)
    plan.run()
    batched = [j.roots() for j in jobs]
    for j in jobs:
        j.roots_buf[:] = 0
        j.run()
    assert [j.roots() for j in jobs] == batched
    for jid, (n, (msg, chal)) in enumerate(zip(sizes, inputs)):
        assert batched[jid] == oracle_job(orc, n, msg, chal)["roots"], f"job {jid} (n={n})"


@pytest.mark.parametrize("cand", [(1, 4, 11, 128, 2), (3, 4, 9, 16, 1), (3, 4, 10, 16, 2), (4, 4, 9, 16, 1),
                                  (3, 3, 9, 8, 1), (3, 5, 10, 16, 1), (0, 7, 11, 1, 2)])
def test_batch_schedules_match_oracle(b64, orc, cand):
    """Batch-plan NTT schedules: variants 3-4 run each job's passes in the
    multi-job register / small-table kernels (per-job kernel choice, several
    launches per NTT slot), variants 0-1 in the multi-job table kernel; every
    schedule gives the oracle's roots for uneven sizes, with and without a graph."""
    from binius64_amd._abi import _apply_tuning
    sizes = [8, 12, 20, 13, 17, 14, 19, 16, 9, 18]
    say("batch schedule", cand, sizes)
    made = [make_job(b64, orc, n, 300 + i) for i, n in enumerate(sizes)]
    want = [oracle_job(orc, n, msg, chal)["roots"] for (job, msg, chal), n in zip(made, sizes)]
    try:
        b64.config_set("batch.clear_sizes", 0)
        for key, v in zip(("variant", "max_k", "tile_log", "tab_min", "tab_group"), cand):
            b64.config_set(f"batch.{key}", v)
        for use_graph in (False, True):
            for job, _, _ in made:
                job.roots_buf[:] = 0
            plan = b64.BatchPlan([m[0] for m in made], use_graph=use_graph)
            plan.run()
            assert [m[0].roots() for m in made] == want, use_graph
    finally:
        for key, v in zip(("variant", "max_k", "tile_log", "tab_min", "tab_group"), (1, 4, 11, 128, 2)):
            b64.config_set(f"batch.{key}", v)
        _apply_tuning(b64.lib())
