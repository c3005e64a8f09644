"""CPU-side checks of libb64hip.so: the library loads, exports every symbol the
header declares, and its host logic (domain twiddles, FRI parameters, the
Fiat-Shamir transcript, host field arithmetic) agrees with the oracle.  No
device call is made here."""
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def b64():
    import binius64_amd as b
    b.lib()
    return b


def declared_functions():
    src = (ROOT / "include" / "b64hip.h").read_text()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(b64_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol(b64):
    names = declared_functions()
    assert len(names) > 40
    L = b64.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    from binius64_amd._abi import SIGNATURES
    assert set(SIGNATURES) == set(names), set(SIGNATURES) ^ set(names)


def test_version_and_status_strings(b64):
    assert b64.lib().b64_version().decode().count(".") == 2
    assert b"subspace" in b64.lib().b64_status_string(2)


@pytest.mark.parametrize("pclmul", [1, 0])
def test_host_field_matches_oracle(b64, orc, pclmul):
    """The host's products (domain set-up, ring-switch claim, sumcheck Horner) on
    PCLMULQDQ and on the portable loop ("host.pclmul" 0) both equal the oracle's."""
    was = b64.config_get("host.pclmul")
    b64.config_set("host.pclmul", pclmul)
    try:
        if pclmul == 0:
            assert b64.config_get("host.pclmul") == 0
        a = orc.random_elems(1, 64)
        b = orc.random_elems(2, 64)
        edge = [0, 1, 2, 0x87, 1 << 127, (1 << 128) - 1]
        xs = orc.to_ints(a) + edge + edge
        ys = orc.to_ints(b) + edge + edge[::-1]
        want = orc.to_ints(orc.gf_mul(orc.to_elems(xs), orc.to_elems(ys)))
        got = [b64.gf_mul(x, y) for x, y in zip(xs, ys)]
        assert got == want
        inv = [b64.gf_inv(x) for x in orc.to_ints(a)]
        assert inv == orc.to_ints(orc.gf_inv(a))
    finally:
        b64.config_set("host.pclmul", was)


@pytest.mark.parametrize("basis", [None, [5, 7, 22, 95, 0x1234, 0xABCDEF0123]])
def test_domain_matches_oracle(b64, orc, basis):
    if basis is None:
        dom, ref = b64.DomainContext.standard(12), orc.Domain(dim=12)
    else:
        dom, ref = b64.DomainContext(orc.to_elems(basis)), orc.Domain(orc.to_elems(basis))
    d = dom.log_domain_size()
    assert d == ref.dim
    for i in range(d + 1):
        assert np.array_equal(dom.subspace(i), ref.subspace(i))
    for layer in range(d):
        for block in list(range(min(1 << layer, 64))) + [(1 << layer) - 1]:
            assert dom.twiddle(layer, block) == ref.twiddle(layer, block)


@pytest.mark.parametrize("n", [8, 12, 16, 20, 24, 26])
def test_setup_commit_params_match_oracle(b64, orc, n):
    p, dom = b64.setup_commit_params(n, 1, 96)
    ref = orc.fri_params_setup(n, 1, 96)
    assert p.log_batch_size == ref["log_batch_size"]
    assert p.log_dim == ref["log_dim"]
    assert p.rs_log_len() == ref["rs_log_len"]
    assert p.fold_arities == ref["fold_arities"]
    assert p.n_test_queries == 232
    # the prover domain is GenericPreExpanded over the Verifier::setup RS subspace
    std = orc.Domain(dim=n + 1)
    rs_basis = std.subspace(p.rs_log_len())
    assert np.array_equal(dom.subspace(p.rs_log_len()), rs_basis)
    ref_dom = orc.Domain(rs_basis)
    for layer in range(dom.log_domain_size()):
        assert dom.twiddle(layer, (1 << layer) - 1) == ref_dom.twiddle(layer, (1 << layer) - 1)


def test_fri_param_kats(b64):
    assert b64.calculate_n_test_queries(96, 1) == 232
    assert b64.calculate_n_test_queries(96, 2) == 142
    for L in range(22, 35):
        assert b64.estimate_optimal_arity(L, 256, 128) == 4
    with pytest.raises(b64.B64Error) as e:
        b64.FRIParams.new(4, 1, 0, [3, 2], 10)
    assert e.value.code == 9  # InvalidFoldAritySequence


def test_transcript_matches_oracle(b64, orc):
    rng = np.random.default_rng(5)
    t, r = b64.ProverTranscript(), orc.Transcript()
    for step in range(60):
        op = step % 5
        data = rng.integers(0, 256, int(rng.integers(0, 200)), dtype=np.uint8).tobytes()
        if op == 0:
            t.message(data), r.message(data)
        elif op == 1:
            t.observe(data), r.observe(data)
        elif op == 2:
            t.decommit(data), r.decommit(data)
        elif op == 3:
            n = int(rng.integers(1, 70))
            assert t.sample_bytes(n) == r.sample(n)
        else:
            bits = int(rng.integers(0, 40))
            assert t.sample_bits(bits) == r.sample_bits(bits)
    assert t.finalize() == r.tape()


def test_device_calls_report_no_device_or_run(b64):
    """Without a GPU, device entry points must fail loudly with NO_DEVICE."""
    if b64.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(b64.B64Error) as e:
        b64.DeviceBuffer(64)
    assert e.value.code == 20


KNOBS = ("variant", "max_k", "tile_log", "tab_min", "tab_group", "waves")


@pytest.fixture
def four_layer_register_schedule(b64):
    """Per-size schedule 'variant 2, 4 layers per pass' for log_d 20..40 (the
    register pass's schedule), restored afterwards."""
    saved = {d: {k: b64.config_get(f"ntt.{k}:{d}") for k in KNOBS} for d in range(20, 41)}
    for d in range(20, 41):
        b64.config_set(f"ntt.variant:{d}", 2)
        b64.config_set(f"ntt.max_k:{d}", 4)
        b64.config_set(f"ntt.tile_log:{d}", 11)
    yield
    for d, knobs in saved.items():
        for k, v in knobs.items():
            b64.config_set(f"ntt.{k}:{d}", v)


@pytest.mark.parametrize("log_d", list(range(20, 41)))
def test_ntt_plan_register_pass_bounds(b64, four_layer_register_schedule, log_d):
    """The register-resident pass addresses a hi group with 32-bit lane offsets,
    so the planner gives it only passes whose flipped bits start at index bit
    10..32 (checked on a 256-CU device)."""
    for encode in (False, True):
        plan = b64.ntt_plan(log_d, 1, 4, encode=encode, n_cus=256)
        assert sum(p["n_layers"] for p in plan) == log_d - 5
        at = 1
        for p in plan:
            assert p["layer0"] == at
            at += p["n_layers"]
            lb = log_d - p["layer0"] - p["n_layers"]
            if p["kernel"] == "k_ntt_pass_r16":
                assert p["n_layers"] == 4 and p["extra_bits"] == 0 and 10 <= lb <= 32
        kinds = [p["kernel"] for p in plan]
        lbs = [log_d - p["layer0"] - p["n_layers"] for p in plan]
        assert all(k == "k_ntt_pass_r16" for k, lb, p in zip(kinds, lbs, plan)
                   if 10 <= lb <= 32 and p["n_layers"] == 4 and p["extra_bits"] == 0)
        if log_d >= 38:
            # the first pass flips bits above index bit 32: not the register pass
            assert lbs[0] > 32 and kinds[0] != "k_ntt_pass_r16"


def test_ntt_plan_large_transform_refused_without_launch(b64):
    """A transform whose pass grids exceed what HIP launches is an argument
    error, reported before anything runs (on a device-less host the call stops
    at the device check, so only the planner is exercised here)."""
    plan = b64.ntt_plan(40, 0, 0, n_cus=256)
    assert max(p["workitems"] for p in plan) >= 1 << 32
    small = b64.ntt_plan(24, 1, 4, n_cus=256)
    assert max(p["workitems"] for p in small) < 1 << 32


def test_fri_params_choose_refuses_too_many_arities(b64):
    dom = b64.DomainContext.standard(48)
    # arity 1, rate 1/2: rs_log_len = log_msg_len, terminal dim 8 (232 queries),
    # so log_msg_len - 8 fold rounds: 32 fit, 33 do not
    p = b64.FRIParams.choose_with_constant_fold_arity(dom, 40, 96, 1, 1)
    assert p.fold_arities == [1] * 32
    with pytest.raises(b64.B64Error) as e:
        b64.FRIParams.choose_with_constant_fold_arity(dom, 41, 96, 1, 1)
    assert e.value.code == 1


def test_batch_schedule_keys_are_separate(b64):
    """Batch plans keep their own NTT schedule table ("batch.<knob>"), applied
    from the tuning file's batch section; setting it leaves the single-job
    schedule alone."""
    from binius64_amd._abi import TUNING_PATH
    from binius64_amd.autotune import KNOBS, load_table
    single, batch = load_table(TUNING_PATH), load_table(TUNING_PATH, "batch_sizes")
    assert single and batch
    for d, cand in batch.items():
        assert tuple(b64.config_get(f"batch.{k.split('.')[1]}:{d}") for k in KNOBS) == cand
        assert cand[0] in (0, 1, 3, 4)
    d = 20
    before = b64.config_get(f"ntt.tab_group:{d}")
    old = b64.config_get(f"batch.tab_group:{d}")
    b64.config_set(f"batch.tab_group:{d}", 8 if old != 8 else 4)
    try:
        assert b64.config_get(f"ntt.tab_group:{d}") == before
    finally:
        b64.config_set(f"batch.tab_group:{d}", old)
    with pytest.raises(b64.B64Error):
        b64.config_set(f"batch.variant:{d}", 2)  # batch plans have variants 0, 1, 3, 4


def test_batched_proof_entry_points_reject_bad_arguments(b64):
    """b64_batch_prove / b64_server_submit_proof check their arguments before any
    device call (no GPU here)."""
    import ctypes as C
    from binius64_amd._abi import ProveJobC
    L = b64.lib()
    assert L.b64_batch_prove(None, 0, None) == 1
    t = b64.ProverTranscript()
    bad = (ProveJobC * 1)(ProveJobC(12, None, None, t.ptr, None))  # no message, no workspace
    assert L.b64_batch_prove(bad, 1, None) == 1
    no_t = (ProveJobC * 1)(ProveJobC(12, 16, 16, None, None))  # no transcript
    assert L.b64_batch_prove(no_t, 1, None) == 1
    ticket = C.c_uint64()
    msg = np.zeros((1 << 12, 2), dtype=np.uint64)
    assert L.b64_server_submit_proof(None, 12, msg.ctypes.data_as(C.POINTER(C.c_uint64)), t.ptr,
                                     C.byref(ticket)) == 1
