"""GPU parity of batched proofs (b64_batch_prove): every job's transcript tape --
its codeword root, the FRI rounds with challenges sampled from that transcript,
and the query phase -- equals the oracle's FRI prover tape for the same message
(prover/prover/src/fri/fold.rs:117-175, 221-244), and the one-job GPU prover's."""
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


def oracle_proof(orc, n, msg):
    """The oracle's FRI prover under a fresh transcript, challenges sampled from it,
    with the Verifier::setup parameters of a 2^n-element message."""
    p = orc.fri_params_setup(n, 1, 96)
    std = orc.Domain(dim=n + 1)
    dom = orc.Domain(std.subspace(p["rs_log_len"]))
    return orc.fri_prove(dom, msg, p["log_dim"], 1, p["log_batch_size"], p["fold_arities"],
                         p["n_test_queries"], threads=8)


def one_job_proof(b64, n, msg, transcript):
    """The one-job GPU flow on a given transcript: commit_interleaved, root written,
    FRIFoldProver rounds with sampled challenges, finish_proof."""
    params, dom = b64.setup_commit_params(n, 1, 96)
    co = b64.commit_interleaved(params, dom, b64.DeviceBuffer.from_host(msg))
    prover = b64.FRIFoldProver(params, dom, co.codeword, co.tree)
    transcript.message(co.commitment)
    roots = [co.commitment]
    for _ in range(params.n_fold_rounds()):
        prover.receive_challenge(transcript.sample())
        r = prover.execute_fold_round()
        if r is not None:
            transcript.message(r)
            roots.append(r)
    prover.finish_proof(transcript)
    return roots


def rand_msg(seed, n):
    return np.random.default_rng(seed).integers(0, 2 ** 63, size=(1 << n, 2), dtype=np.uint64)


# one job (one stream), two jobs (one per part), odd and even counts split over the two parts
@pytest.mark.parametrize("sizes", [[8], [12], [12, 13], [10, 10], [8, 9, 10, 11, 12, 13, 14, 15, 16],
                                   [16, 9, 13, 9, 16]])
def test_batch_prove_matches_oracle(b64, orc, sizes):
    say("batch_prove sizes", sizes)
    msgs = [rand_msg(300 + i, n) for i, n in enumerate(sizes)]
    jobs = [b64.ProveJob(n, m) for n, m in zip(sizes, msgs)]
    b64.batch_prove(jobs)
    for i, (n, m, j) in enumerate(zip(sizes, msgs, jobs)):
        want = oracle_proof(orc, n, m)
        assert j.roots() == want["roots"], f"job {i} (n={n}) roots"
        assert j.tape() == want["tape"], f"job {i} (n={n}) tape"


def test_batch_prove_continues_transcripts(b64):
    """Transcripts that already hold data are continued: the tape equals the
    one-job prover's on a transcript with the same prefix."""
    sizes = [10, 14, 12]
    msgs = [rand_msg(400 + i, n) for i, n in enumerate(sizes)]
    jobs = []
    for i, (n, m) in enumerate(zip(sizes, msgs)):
        t = b64.ProverTranscript()
        t.message(bytes([i]) * (7 + i))
        t.observe(b"prefix")
        jobs.append(b64.ProveJob(n, m, t))
    b64.batch_prove(jobs)
    for i, (n, m, j) in enumerate(zip(sizes, msgs, jobs)):
        t = b64.ProverTranscript()
        t.message(bytes([i]) * (7 + i))
        t.observe(b"prefix")
        roots = one_job_proof(b64, n, m, t)
        assert j.roots() == roots, f"job {i}"
        assert j.tape() == t.finalize(), f"job {i}"


def test_batch_prove_config4_mix(b64, orc):
    """The batching workload's job mix (64 jobs, log_n ~ U{12..20} drawn with seed
    4, uniform random messages), proved with transcript challenges: every tape
    equals the oracle prover's; a sample also equals the one-job GPU prover's."""
    rng = np.random.default_rng(4)
    sizes = [int(x) for x in rng.integers(12, 21, size=64)]
    say("config-4 sizes", sizes)
    msgs = [np.random.default_rng(1000 + jid).integers(0, 2 ** 63, size=(1 << n, 2), dtype=np.uint64)
            for jid, n in enumerate(sizes)]
    jobs = [b64.ProveJob(n, m) for n, m in zip(sizes, msgs)]
    b64.batch_prove(jobs)
    for jid, (n, m, j) in enumerate(zip(sizes, msgs, jobs)):
        want = oracle_proof(orc, n, m)
        assert j.tape() == want["tape"], f"job {jid} (n={n})"
    for jid in (0, 17, 63):
        t = b64.ProverTranscript()
        one_job_proof(b64, sizes[jid], msgs[jid], t)
        assert jobs[jid].tape() == t.finalize(), f"job {jid}"


def test_batch_prove_rejects_bad_jobs(b64):
    m = rand_msg(1, 8)
    with pytest.raises(b64.B64Error):
        b64.ProveJob(7, rand_msg(1, 7))  # below the smallest size
    job = b64.ProveJob(8, m)
    job.desc.transcript = None
    with pytest.raises(b64.B64Error):
        b64.batch_prove([job])
