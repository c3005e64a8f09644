"""ctypes binding of libb64hip.so (the C ABI declared in include/b64hip.h).

The library is the product: importing this module never falls back to any CPU
implementation.  If libb64hip.so is missing the import fails with a clear error
(build it with ``python -m binius64_amd.build``).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

# B64_SANITIZED_LIBS names a folder of host-sanitizer builds (tests/sanitize/run.sh)
LIB_PATH = Path(os.environ["B64_SANITIZED_LIBS"]) / "libb64hip.so" if os.environ.get("B64_SANITIZED_LIBS") \
    else Path(__file__).resolve().parent / "libb64hip.so"

u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
vp = C.c_void_p

MAX_FOLD_ARITIES = 32


class FriParamsC(C.Structure):
    _fields_ = [
        ("log_dim", C.c_uint32),
        ("log_inv_rate", C.c_uint32),
        ("log_batch_size", C.c_uint32),
        ("n_test_queries", C.c_uint32),
        ("n_fold_arities", C.c_uint32),
        ("fold_arities", C.c_uint32 * MAX_FOLD_ARITIES),
        ("log_terminal_dim", C.c_uint32),
    ]


class NttPassInfoC(C.Structure):
    _fields_ = [
        ("layer0", C.c_uint32),
        ("n_layers", C.c_uint32),
        ("low_bits", C.c_uint32),
        ("extra_bits", C.c_uint32),
        ("kernel", C.c_uint32),
        ("workitems", C.c_uint64),
    ]


NTT_KERNEL_NAMES = {0: "k_ntt_pass", 1: "k_ntt_pass_tab", 2: "k_ntt_pass_tab", 3: "k_ntt_pass_tab_w4",
                    4: "k_ntt_pass_r16", 5: "k_ntt_pass_t8", 6: "k_ntt_pass_r16w8", 7: "k_ntt_pass_t4u"}


class ServerConfigC(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_size_t),
        ("max_batch", C.c_uint32),
        ("n_streams", C.c_uint32),
    ]


class JobDescC(C.Structure):
    _fields_ = [
        ("log_msg_len", C.c_uint32),
        ("d_msg", C.c_void_p),
        ("challenges", C.POINTER(C.c_uint64)),
        ("d_workspace", C.c_void_p),
        ("roots_out", C.POINTER(C.c_uint8)),
    ]


class ProveJobC(C.Structure):
    _fields_ = [
        ("log_msg_len", C.c_uint32),
        ("d_msg", C.c_void_p),
        ("d_workspace", C.c_void_p),
        ("transcript", C.c_void_p),
        ("roots_out", C.POINTER(C.c_uint8)),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "b64_batch_prove": (C.c_int, [C.POINTER(ProveJobC), C.c_uint32, vp]),
    "b64_job_workspace_size": (C.c_size_t, [C.c_uint32]),
    "b64_job_n_roots": (C.c_uint32, [C.c_uint32]),
    "b64_job_oracle": (C.c_int, [C.POINTER(JobDescC), C.c_uint32, C.POINTER(vp), u32p]),
    "b64_job_run": (C.c_int, [C.POINTER(JobDescC), vp]),
    "b64_batch_plan_create": (C.c_int, [C.POINTER(JobDescC), C.c_uint32, C.c_int, vp, C.POINTER(vp)]),
    "b64_batch_plan_run": (C.c_int, [vp, vp]),
    "b64_batch_plan_enqueue": (C.c_int, [vp, vp]),
    "b64_batch_plan_collect": (C.c_int, [vp]),
    "b64_server_create": (C.c_int, [C.POINTER(ServerConfigC), C.POINTER(vp)]),
    "b64_server_submit": (C.c_int, [vp, C.c_uint32, u64p, u64p, C.POINTER(C.c_uint64)]),
    "b64_server_submit_proof": (C.c_int, [vp, C.c_uint32, u64p, vp, C.POINTER(C.c_uint64)]),
    "b64_server_poll": (C.c_int, [vp, C.c_uint64, u8p, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "b64_server_wait": (C.c_int, [vp, C.c_uint64, u8p, C.POINTER(C.c_double)]),
    "b64_server_stats": (C.c_int, [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_size_t)]),
    "b64_server_destroy": (None, [vp]),
    "b64_batch_plan_n_launches": (C.c_uint32, [vp]),
    "b64_batch_plan_destroy": (None, [vp]),
    "b64_version": (C.c_char_p, []),
    "b64_status_string": (C.c_char_p, [C.c_int]),
    "b64_device_count": (C.c_int, []),
    "b64_device_alloc": (C.c_int, [C.c_size_t, C.POINTER(vp)]),
    "b64_device_free": (C.c_int, [vp]),
    "b64_memcpy_htod": (C.c_int, [vp, vp, C.c_size_t, vp]),
    "b64_memcpy_dtoh": (C.c_int, [vp, vp, C.c_size_t, vp]),
    "b64_memcpy_dtod": (C.c_int, [vp, vp, C.c_size_t, vp]),
    "b64_stream_synchronize": (C.c_int, [vp]),
    "b64_config_set": (C.c_int, [C.c_char_p, C.c_int64]),
    "b64_config_get": (C.c_int, [C.c_char_p, C.POINTER(C.c_int64)]),
    "b64_gf_mul": (None, [u64p, u64p, u64p]),
    "b64_gf_inv": (None, [u64p, u64p]),
    "b64_gf_mul_device": (C.c_int, [vp, vp, vp, C.c_uint64, vp]),
    "b64_gf_mul_const_device": (C.c_int, [vp, u64p, vp, C.c_uint64, vp]),
    "b64_domain_create": (C.c_int, [u64p, C.c_uint32, C.POINTER(vp)]),
    "b64_domain_create_standard": (C.c_int, [C.c_uint32, C.POINTER(vp)]),
    "b64_domain_destroy": (None, [vp]),
    "b64_domain_log_size": (C.c_uint32, [vp]),
    "b64_domain_subspace": (C.c_int, [vp, C.c_uint32, u64p]),
    "b64_domain_twiddle": (C.c_int, [vp, C.c_uint32, C.c_uint64, u64p]),
    "b64_ntt_forward": (C.c_int, [vp, vp, C.c_uint32, C.c_uint32, C.c_uint32, vp]),
    "b64_ntt_plan": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_uint32,
                               C.POINTER(NttPassInfoC), C.c_uint32, C.POINTER(C.c_uint32)]),
    "b64_ntt_inverse": (C.c_int, [vp, vp, C.c_uint32, C.c_uint32, C.c_uint32, vp]),
    "b64_rs_encode_batch": (C.c_int, [vp, C.c_uint32, C.c_uint32, vp, C.c_uint32, vp, vp]),
    "b64_merkle_commit": (C.c_int, [vp, C.c_uint64, C.c_uint64, vp, u8p, vp]),
    "b64_merkle_layer": (C.c_int, [vp, C.c_uint32, C.c_uint32, u8p, vp]),
    "b64_merkle_branch": (C.c_int, [vp, C.c_uint32, C.c_uint64, C.c_uint32, u8p, vp]),
    "b64_fri_fold_interleaved": (C.c_int, [vp, C.c_uint32, C.c_uint32, u64p, vp, vp]),
    "b64_fri_fold_codeword": (C.c_int, [vp, vp, C.c_uint32, u64p, C.c_uint32, vp, vp]),
    "b64_calculate_n_test_queries": (C.c_uint32, [C.c_uint32, C.c_uint32]),
    "b64_estimate_optimal_arity": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32]),
    "b64_fri_params_new": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, u32p, C.c_uint32, C.c_uint32,
                                     C.POINTER(FriParamsC)]),
    "b64_fri_params_choose": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                        C.POINTER(FriParamsC)]),
    "b64_setup_commit_params": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(FriParamsC),
                                          C.POINTER(vp)]),
    "b64_fri_commit_interleaved": (C.c_int, [C.POINTER(FriParamsC), vp, vp, vp, vp, u8p, vp]),
    "b64_fri_commit_fold": (C.c_int, [C.POINTER(FriParamsC), vp, vp, vp, vp, u64p, vp, u8p, vp]),
    "b64_transcript_create": (C.c_int, [C.POINTER(vp)]),
    "b64_transcript_destroy": (None, [vp]),
    "b64_transcript_message": (None, [vp, u8p, C.c_size_t]),
    "b64_transcript_observe": (None, [vp, u8p, C.c_size_t]),
    "b64_transcript_decommit": (None, [vp, u8p, C.c_size_t]),
    "b64_transcript_sample": (None, [vp, u8p, C.c_size_t]),
    "b64_transcript_sample_bits": (C.c_uint32, [vp, C.c_uint32]),
    "b64_transcript_tape": (C.c_size_t, [vp, u8p, C.c_size_t]),
    "b64_proof_serialize": (C.c_int, [u8p, C.c_uint64, C.c_char_p, u8p, C.c_uint64, u64p]),
    "b64_proof_deserialize": (C.c_int, [u8p, C.c_uint64, u64p, u64p, u64p, u64p]),
    "b64_fri_prover_create": (C.c_int, [C.POINTER(FriParamsC), vp, vp, vp, vp, C.POINTER(vp)]),
    "b64_fri_prover_destroy": (None, [vp]),
    "b64_fri_prover_n_rounds": (C.c_uint32, [vp]),
    "b64_fri_prover_curr_round": (C.c_uint32, [vp]),
    "b64_fri_prover_receive_challenge": (C.c_int, [vp, u64p]),
    "b64_fri_prover_execute_fold_round": (C.c_int, [vp, u8p, C.POINTER(C.c_int)]),
    "b64_fri_prover_finish_proof": (C.c_int, [vp, vp]),
    "b64_fri_prover_oracle": (C.c_int, [vp, C.c_uint32, C.POINTER(vp), u32p, C.POINTER(vp)]),
    "b64_pack_witness": (C.c_int, [vp, C.c_uint64, C.c_uint32, vp, vp]),
    "b64_ring_switch_fold_1b_rows": (C.c_int, [vp, vp, C.c_uint32, vp, vp]),
    "b64_ring_switch_fold_1b_rows_noalloc": (C.c_int, [vp, vp, C.c_uint32, vp, vp]),
    "b64_ring_switch_fold_elems": (C.c_int, [vp, C.c_uint32, u64p, vp]),
    "b64_eq_ind_partial_eval": (C.c_int, [u64p, C.c_uint32, vp, vp]),
    "b64_sumcheck_create": (C.c_int, [vp, vp, C.c_uint32, u64p, vp, C.POINTER(vp)]),
    "b64_sumcheck_destroy": (None, [vp]),
    "b64_sumcheck_n_vars": (C.c_uint32, [vp]),
    "b64_sumcheck_execute": (C.c_int, [vp, u64p]),
    "b64_sumcheck_fold": (C.c_int, [vp, u64p]),
    "b64_sumcheck_finish": (C.c_int, [vp, u64p]),
    "b64_basefold_prove": (C.c_int, [C.POINTER(FriParamsC), vp, vp, vp, u64p, vp, vp, vp, vp]),
    "b64_vision_permute": (C.c_int, [C.c_uint32, vp, C.c_uint64, vp]),
    "b64_vision_digest": (C.c_int, [C.c_uint32, vp, C.c_uint64, C.c_uint64, vp, vp]),
    "b64_merkle_commit_vision": (C.c_int, [C.c_uint32, vp, C.c_uint64, C.c_uint64, vp, u8p, vp]),
    "b64_fri_commit_interleaved_ex": (C.c_int, [C.POINTER(FriParamsC), vp, C.c_uint32, vp, vp, vp, u8p, vp]),
    "b64_fri_prover_create_ex": (C.c_int, [C.POINTER(FriParamsC), vp, C.c_uint32, vp, vp, vp, C.POINTER(vp)]),
    "b64_basefold_prove_ex": (C.c_int, [C.POINTER(FriParamsC), vp, C.c_uint32, vp, vp, u64p, vp, vp, vp, vp]),
    "b64_pcs_prove_ex": (C.c_int, [C.POINTER(FriParamsC), vp, C.c_uint32, vp, vp, vp, u64p, C.c_uint32, vp, vp]),
    "b64_pcs_prove": (C.c_int, [C.POINTER(FriParamsC), vp, vp, vp, vp, u64p, C.c_uint32, vp, vp]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build the HIP library with `python -m binius64_amd.build` "
                "(there is no CPU fallback)")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
        _apply_tuning(L)
    return _lib


TUNING_PATH = Path(__file__).resolve().parent / "ntt_tuning.json"


def _apply_tuning(L) -> None:
    """Per-size NTT schedules measured by binius64_amd.autotune, if present:
    single-job transforms and batch plans."""
    if TUNING_PATH.exists():
        from .autotune import apply_table, load_table
        apply_table(load_table(TUNING_PATH), L)
        apply_table(load_table(TUNING_PATH, "batch_sizes"), L, prefix="batch")


class B64Error(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        self.code = code
        msg = lib().b64_status_string(code).decode()
        super().__init__(f"{what}: status {code} ({msg})" if what else f"status {code} ({msg})")


def check(code: int, what: str = "") -> None:
    if code != 0:
        raise B64Error(code, what)
