"""ctypes view of the CPU oracle (libb64oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker or the CPU baseline, never
as the product path.  Field elements are numpy uint64 arrays of shape (n, 2)
holding (lo, hi) words, i.e. the little-endian u128 byte order of the
reference's serialization (verifier/utils/src/serialization.rs:66-72).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
# B64_SANITIZED_LIBS names a folder of sanitizer builds (tests/sanitize/run.sh)
_LIB_PATH = Path(os.environ["B64_SANITIZED_LIBS"]) / "libb64oracle.so" if os.environ.get("B64_SANITIZED_LIBS") \
    else _HERE / "libb64oracle.so"
_lib = None

u64p = C.POINTER(C.c_uint64)
u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
intp = C.POINTER(C.c_int)


def build(force: bool = False) -> Path:
    if force or not _LIB_PATH.exists():
        subprocess.run(["make", "-s", "-C", str(_HERE), "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        L = C.CDLL(str(_LIB_PATH))
        L.bo_domain_new.restype = C.c_void_p
        L.bo_domain_new.argtypes = [u64p, C.c_int]
        L.bo_domain_new_standard.restype = C.c_void_p
        L.bo_domain_new_standard.argtypes = [C.c_int]
        L.bo_domain_free.argtypes = [C.c_void_p]
        L.bo_domain_dim.argtypes = [C.c_void_p]
        L.bo_domain_evals.argtypes = [C.c_void_p, u64p]
        L.bo_domain_twiddle.argtypes = [C.c_void_p, C.c_int, C.c_uint64, u64p]
        L.bo_domain_subspace.argtypes = [C.c_void_p, C.c_int, u64p]
        for name in ("bo_ntt_forward_ref", "bo_ntt_inverse_ref"):
            getattr(L, name).argtypes = [C.c_void_p, u64p, C.c_int, C.c_int, C.c_int]
        L.bo_ntt_forward_mt.argtypes = [C.c_void_p, u64p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.bo_encode_batch.argtypes = [C.c_void_p, u64p, C.c_int, C.c_int, C.c_int, u64p, C.c_int]
        L.bo_gf_mul_many.argtypes = [u64p, u64p, u64p, C.c_size_t, C.c_int]
        for name in ("bo_gf_square_many", "bo_gf_inv_many", "bo_gf_mul_x_many", "bo_gf_mul_inv_x_many"):
            getattr(L, name).argtypes = [u64p, u64p, C.c_size_t]
        L.bo_gf_pow.argtypes = [u64p, u64p, u64p]
        L.bo_sha256.argtypes = [u8p, C.c_size_t, u8p]
        L.bo_sha256_compress_block.argtypes = [u32p, u8p, u32p]
        L.bo_binius_compress.argtypes = [u8p, u8p, u8p]
        L.bo_merkle_build.argtypes = [u64p, C.c_size_t, C.c_size_t, u8p, C.c_int]
        L.bo_merkle_branch.argtypes = [u8p, C.c_int, C.c_size_t, C.c_int, u8p]
        L.bo_n_test_queries.restype = C.c_size_t
        L.bo_n_test_queries.argtypes = [C.c_int, C.c_int]
        L.bo_estimate_optimal_arity.argtypes = [C.c_int, C.c_int, C.c_int]
        L.bo_fri_params_setup.argtypes = [C.c_int, C.c_int, C.c_int, intp]
        L.bo_eq_ind.argtypes = [u64p, C.c_int, u64p]
        L.bo_fold_interleaved.argtypes = [u64p, C.c_int, C.c_int, u64p, u64p, C.c_int]
        L.bo_fold_codeword.argtypes = [C.c_void_p, u64p, C.c_int, u64p, C.c_int, u64p, C.c_int]
        L.bo_fold_chunk.argtypes = [C.c_void_p, C.c_int, C.c_uint64, u64p, u64p, C.c_int, u64p]
        L.bo_evaluate_mle.argtypes = [u64p, C.c_int, u64p, u64p]
        L.bo_pack_witness.argtypes = [u64p, C.c_size_t, u64p]
        L.bo_rs_fold_1b_rows.argtypes = [u64p, u64p, C.c_int, u64p, C.c_int, C.c_int]
        L.bo_rs_fold_elems.argtypes = [u64p, C.c_size_t, u64p, C.c_int, C.c_int]
        L.bo_fri_prove.argtypes = [C.c_void_p, u64p, C.c_int, C.c_int, C.c_int, intp, C.c_int, C.c_int,
                                   u64p, u8p, C.c_size_t, C.POINTER(C.c_size_t), u8p, u64p, u64p, C.c_int]
        L.bo_fri_prove_ex.argtypes = L.bo_fri_prove.argtypes + [C.c_void_p, C.c_void_p]
        L.bo_sumcheck_prove.argtypes = [u64p, u64p, C.c_int, u64p, C.c_void_p, u64p, u64p, C.c_int]
        L.bo_basefold_prove.argtypes = [C.c_void_p, u64p, u64p, u64p, u64p, u8p, C.c_int, C.c_int, C.c_int,
                                        intp, C.c_int, C.c_int, C.c_void_p, u64p, C.c_int]
        L.bo_tensor_transpose.argtypes = [u64p, u64p]
        L.bo_pcs_prove.argtypes = [C.c_void_p, u64p, u64p, u64p, u8p, C.c_int, C.c_int, C.c_int, intp, C.c_int,
                                   C.c_int, C.c_void_p, C.c_int]
        L.bo_basefold_prove_ex.argtypes = L.bo_basefold_prove.argtypes + [C.c_void_p, C.c_void_p]
        L.bo_pcs_prove_ex.argtypes = L.bo_pcs_prove.argtypes + [C.c_void_p, C.c_void_p]
        L.bo_vision_new.restype = C.c_void_p
        L.bo_vision_new.argtypes = [C.c_int, u64p, u64p, u64p, u64p]
        L.bo_vision_free.argtypes = [C.c_void_p]
        L.bo_vision_round_constants.argtypes = [C.c_void_p, u64p]
        L.bo_vision_b_maps.argtypes = [C.c_void_p, u64p, u64p, u64p, C.c_size_t]
        L.bo_vision_batch_invert.argtypes = [u64p, C.c_int]
        L.bo_vision_permutation.argtypes = [C.c_void_p, u64p, C.c_size_t]
        L.bo_vision_digest.argtypes = [C.c_void_p, u8p, C.c_size_t, u8p]
        L.bo_vision_compress.argtypes = [C.c_void_p, u8p, u8p, u8p]
        L.bo_merkle_build_vision.argtypes = [C.c_void_p, C.c_void_p, u64p, C.c_size_t, C.c_size_t, u8p, C.c_int]
        L.bo_transcript_new.restype = C.c_void_p
        L.bo_transcript_free.argtypes = [C.c_void_p]
        for name in ("bo_transcript_message", "bo_transcript_observe", "bo_transcript_decommit",
                     "bo_transcript_sample"):
            getattr(L, name).argtypes = [C.c_void_p, u8p, C.c_size_t]
        L.bo_transcript_sample_bits.restype = C.c_uint32
        L.bo_transcript_sample_bits.argtypes = [C.c_void_p, C.c_int]
        L.bo_transcript_tape.restype = C.c_size_t
        L.bo_transcript_tape.argtypes = [C.c_void_p, u8p, C.c_size_t]
        L.bo_splitmix_fill.argtypes = [C.c_uint64, u64p, C.c_size_t]
        _lib = L
    return _lib


def _p64(a: np.ndarray):
    assert a.dtype == np.uint64 and a.flags.c_contiguous
    return a.ctypes.data_as(u64p)


def _p8(a: np.ndarray):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(u8p)


def nthreads_default() -> int:
    return int(os.environ.get("B64_ORACLE_THREADS", os.cpu_count() or 1))


# ---- element helpers ---------------------------------------------------------
def to_elems(ints) -> np.ndarray:
    ints = list(ints)
    out = np.zeros((len(ints), 2), dtype=np.uint64)
    for i, v in enumerate(ints):
        out[i, 0] = v & 0xFFFFFFFFFFFFFFFF
        out[i, 1] = v >> 64
    return out


def to_ints(arr: np.ndarray):
    arr = np.asarray(arr, dtype=np.uint64).reshape(-1, 2)
    return [int(lo) | (int(hi) << 64) for lo, hi in arr]


def random_elems(seed: int, n: int) -> np.ndarray:
    out = np.zeros((n, 2), dtype=np.uint64)
    lib().bo_splitmix_fill(seed, _p64(out), 2 * n)
    return out


def random_words(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint64)
    lib().bo_splitmix_fill(seed, _p64(out), n)
    return out


# ---- field -------------------------------------------------------------------
def gf_mul(a: np.ndarray, b: np.ndarray, variant: int = 0) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 2)
    b = np.ascontiguousarray(b, dtype=np.uint64).reshape(-1, 2)
    out = np.zeros_like(a)
    lib().bo_gf_mul_many(_p64(a), _p64(b), _p64(out), a.shape[0], variant)
    return out


def _unary(name, a):
    a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 2)
    out = np.zeros_like(a)
    getattr(lib(), name)(_p64(a), _p64(out), a.shape[0])
    return out


def gf_square(a):
    return _unary("bo_gf_square_many", a)


def gf_inv(a):
    return _unary("bo_gf_inv_many", a)


def gf_mul_x(a):
    return _unary("bo_gf_mul_x_many", a)


def gf_mul_inv_x(a):
    return _unary("bo_gf_mul_inv_x_many", a)


def gf_pow(a: int, e: int) -> int:
    A, E, O = to_elems([a]), to_elems([e]), np.zeros((1, 2), dtype=np.uint64)
    lib().bo_gf_pow(_p64(A), _p64(E), _p64(O))
    return to_ints(O)[0]


# ---- domain ------------------------------------------------------------------
class Domain:
    """GenericPreExpanded over a subspace basis (ntt/domain_context.rs:99-159)."""

    def __init__(self, basis=None, dim: int | None = None):
        if basis is None:
            self.ptr = lib().bo_domain_new_standard(dim)
        else:
            b = np.ascontiguousarray(basis, dtype=np.uint64).reshape(-1, 2)
            self.ptr = lib().bo_domain_new(_p64(b), b.shape[0])
        if not self.ptr:
            raise ValueError("bad domain")
        self.dim = lib().bo_domain_dim(self.ptr)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().bo_domain_free(self.ptr)
            self.ptr = None

    def evals(self) -> np.ndarray:
        out = np.zeros((self.dim * self.dim, 2), dtype=np.uint64)
        lib().bo_domain_evals(self.ptr, _p64(out))
        return out.reshape(self.dim, self.dim, 2)

    def twiddle(self, layer: int, block: int) -> int:
        out = np.zeros(2, dtype=np.uint64)
        lib().bo_domain_twiddle(self.ptr, layer, block, _p64(out))
        return int(out[0]) | (int(out[1]) << 64)

    def subspace(self, i: int) -> np.ndarray:
        out = np.zeros((max(i, 1), 2), dtype=np.uint64)
        lib().bo_domain_subspace(self.ptr, i, _p64(out))
        return out[:i]


def ntt_forward(dom: Domain, data: np.ndarray, skip_early: int, skip_late: int,
                threads: int = 1, log_num_shares: int | None = None) -> np.ndarray:
    d = np.ascontiguousarray(data, dtype=np.uint64).reshape(-1, 2).copy()
    log_d = d.shape[0].bit_length() - 1
    if threads <= 1:
        rc = lib().bo_ntt_forward_ref(dom.ptr, _p64(d), log_d, skip_early, skip_late)
    else:
        lns = log_num_shares if log_num_shares is not None else threads.bit_length() - 1
        rc = lib().bo_ntt_forward_mt(dom.ptr, _p64(d), log_d, skip_early, skip_late, lns, threads)
    if rc:
        raise ValueError(f"ntt precondition failed ({rc})")
    return d


def ntt_inverse(dom: Domain, data: np.ndarray, skip_early: int, skip_late: int) -> np.ndarray:
    d = np.ascontiguousarray(data, dtype=np.uint64).reshape(-1, 2).copy()
    log_d = d.shape[0].bit_length() - 1
    if lib().bo_ntt_inverse_ref(dom.ptr, _p64(d), log_d, skip_early, skip_late):
        raise ValueError("ntt precondition failed")
    return d


def encode_batch(dom: Domain, msg: np.ndarray, log_dim: int, log_inv_rate: int, log_batch: int,
                 threads: int = 1) -> np.ndarray:
    m = np.ascontiguousarray(msg, dtype=np.uint64).reshape(-1, 2)
    assert m.shape[0] == 1 << (log_dim + log_batch)
    out = np.zeros((1 << (log_dim + log_batch + log_inv_rate), 2), dtype=np.uint64)
    rc = lib().bo_encode_batch(dom.ptr, _p64(m), log_dim, log_inv_rate, log_batch, _p64(out), threads)
    if rc:
        raise ValueError(f"encode failed ({rc})")
    return out


# ---- hashing -----------------------------------------------------------------
def sha256(data: bytes) -> bytes:
    buf = np.frombuffer(data, dtype=np.uint8).copy() if data else np.zeros(1, dtype=np.uint8)
    out = np.zeros(32, dtype=np.uint8)
    lib().bo_sha256(_p8(buf), len(data), _p8(out))
    return out.tobytes()


def sha256_compress_block(iv, block: bytes):
    ivv = np.array(iv, dtype=np.uint32)
    blk = np.frombuffer(block, dtype=np.uint8).copy()
    out = np.zeros(8, dtype=np.uint32)
    lib().bo_sha256_compress_block(ivv.ctypes.data_as(u32p), _p8(blk), out.ctypes.data_as(u32p))
    return [int(x) for x in out]


def binius_compress(left: bytes, right: bytes) -> bytes:
    l = np.frombuffer(left, dtype=np.uint8).copy()
    r = np.frombuffer(right, dtype=np.uint8).copy()
    out = np.zeros(32, dtype=np.uint8)
    lib().bo_binius_compress(_p8(l), _p8(r), _p8(out))
    return out.tobytes()


def merkle_build(elems: np.ndarray, batch: int, threads: int = 1) -> np.ndarray:
    e = np.ascontiguousarray(elems, dtype=np.uint64).reshape(-1, 2)
    n_leaves = e.shape[0] // batch if batch else 0
    tree = np.zeros((max(2 * n_leaves - 1, 1), 32), dtype=np.uint8)
    rc = lib().bo_merkle_build(_p64(e), e.shape[0], batch, _p8(tree), threads)
    if rc:
        raise ValueError(f"merkle error {rc}")
    return tree


def merkle_branch(tree: np.ndarray, log_len: int, index: int, layer_depth: int) -> np.ndarray:
    out = np.zeros((max(log_len - layer_depth, 1), 32), dtype=np.uint8)
    rc = lib().bo_merkle_branch(_p8(np.ascontiguousarray(tree)), log_len, index, layer_depth, _p8(out))
    if rc:
        raise ValueError(f"branch error {rc}")
    return out[: log_len - layer_depth]


# ---- FRI ---------------------------------------------------------------------
def n_test_queries(security_bits: int, log_inv_rate: int) -> int:
    return lib().bo_n_test_queries(security_bits, log_inv_rate)


def estimate_optimal_arity(log_block_length: int, digest_size: int, field_size: int) -> int:
    return lib().bo_estimate_optimal_arity(log_block_length, digest_size, field_size)


def fri_params_setup(log_msg_len: int, log_inv_rate: int = 1, security_bits: int = 96) -> dict:
    out = (C.c_int * 80)()
    rc = lib().bo_fri_params_setup(log_msg_len, log_inv_rate, security_bits, out)
    if rc:
        raise ValueError(f"fri params error {rc}")
    n = out[4]
    return {"log_batch_size": out[0], "log_dim": out[1], "rs_log_len": out[2], "n_test_queries": out[3],
            "fold_arities": [out[5 + i] for i in range(n)], "log_inv_rate": log_inv_rate,
            "log_msg_len": log_msg_len}


def eq_ind(point: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(point, dtype=np.uint64).reshape(-1, 2)
    out = np.zeros((1 << p.shape[0], 2), dtype=np.uint64)
    lib().bo_eq_ind(_p64(p), p.shape[0], _p64(out))
    return out


def fold_interleaved(cw: np.ndarray, log_len: int, log_batch: int, chal: np.ndarray,
                     threads: int = 1) -> np.ndarray:
    c = np.ascontiguousarray(cw, dtype=np.uint64).reshape(-1, 2)
    ch = np.ascontiguousarray(chal, dtype=np.uint64).reshape(-1, 2)
    out = np.zeros((1 << log_len, 2), dtype=np.uint64)
    if lib().bo_fold_interleaved(_p64(c), log_len, log_batch, _p64(ch), _p64(out), threads):
        raise ValueError("fold error")
    return out


def fold_codeword(dom: Domain, cw: np.ndarray, log_len: int, chal: np.ndarray, threads: int = 1) -> np.ndarray:
    c = np.ascontiguousarray(cw, dtype=np.uint64).reshape(-1, 2)
    ch = np.ascontiguousarray(chal, dtype=np.uint64).reshape(-1, 2)
    out = np.zeros((1 << (log_len - ch.shape[0]), 2), dtype=np.uint64)
    if lib().bo_fold_codeword(dom.ptr, _p64(c), log_len, _p64(ch), ch.shape[0], _p64(out), threads):
        raise ValueError("fold error")
    return out


def fold_chunk(dom: Domain, log_len: int, chunk_index: int, vals: np.ndarray, chal: np.ndarray) -> int:
    v = np.ascontiguousarray(vals, dtype=np.uint64).reshape(-1, 2)
    ch = np.ascontiguousarray(chal, dtype=np.uint64).reshape(-1, 2)
    assert v.shape[0] == 1 << ch.shape[0]
    out = np.zeros(2, dtype=np.uint64)
    lib().bo_fold_chunk(dom.ptr, log_len, chunk_index, _p64(v), _p64(ch), ch.shape[0], _p64(out))
    return int(out[0]) | (int(out[1]) << 64)


def evaluate_mle(evals: np.ndarray, point: np.ndarray) -> int:
    e = np.ascontiguousarray(evals, dtype=np.uint64).reshape(-1, 2)
    p = np.ascontiguousarray(point, dtype=np.uint64).reshape(-1, 2)
    out = np.zeros(2, dtype=np.uint64)
    lib().bo_evaluate_mle(_p64(e), p.shape[0], _p64(p), _p64(out))
    return int(out[0]) | (int(out[1]) << 64)


def pack_witness(words: np.ndarray) -> np.ndarray:
    w = np.ascontiguousarray(words, dtype=np.uint64)
    out = np.zeros((w.shape[0] // 2, 2), dtype=np.uint64)
    lib().bo_pack_witness(_p64(w), w.shape[0] // 2, _p64(out))
    return out


def rs_fold_1b_rows(mat: np.ndarray, vec: np.ndarray, variant: int = 0, threads: int | None = None) -> np.ndarray:
    """ring_switch::fold_1b_rows_for_b128: 128 elements."""
    m = np.ascontiguousarray(mat, dtype=np.uint64).reshape(-1, 2)
    v = np.ascontiguousarray(vec, dtype=np.uint64).reshape(-1, 2)
    assert m.shape == v.shape and (m.shape[0] & (m.shape[0] - 1)) == 0
    out = np.zeros((128, 2), dtype=np.uint64)
    lib().bo_rs_fold_1b_rows(_p64(m), _p64(v), m.shape[0].bit_length() - 1, _p64(out), variant,
                             threads or nthreads_default())
    return out


def rs_fold_elems(elems: np.ndarray, vec: np.ndarray, variant: int = 0, threads: int | None = None) -> np.ndarray:
    """ring_switch::fold_b128_elems_inplace (variant 0) / fold_elems_inplace (variant 1)."""
    e = np.array(elems, dtype=np.uint64).reshape(-1, 2)
    v = np.ascontiguousarray(vec, dtype=np.uint64).reshape(-1, 2)
    assert v.shape[0] == 128
    lib().bo_rs_fold_elems(_p64(e), e.shape[0], _p64(v), variant, threads or nthreads_default())
    return e


def _scheme(scheme):
    """(leaf Vision, node Vision) pointers for the Vision Merkle scheme, (None, None) for SHA-256."""
    return (None, None) if scheme is None else (scheme[0].ptr, scheme[1].ptr)


def fri_prove(dom: Domain, msg: np.ndarray, log_dim: int, log_inv_rate: int, log_batch: int,
              arities, n_queries: int, challenges: np.ndarray | None = None, threads: int = 1,
              scheme=None) -> dict:
    m = np.ascontiguousarray(msg, dtype=np.uint64).reshape(-1, 2)
    ar = (C.c_int * max(len(arities), 1))(*arities)
    n_rounds = log_dim + log_batch
    ch_out = np.zeros((max(n_rounds, 1), 2), dtype=np.uint64)
    roots = np.zeros((len(arities) + 2, 32), dtype=np.uint8)
    n_final = log_dim - sum(arities)
    term = np.zeros((1 << (n_final + log_inv_rate), 2), dtype=np.uint64)
    tape_len = C.c_size_t(0)
    ch_in = None
    if challenges is not None:
        ch_in = np.ascontiguousarray(challenges, dtype=np.uint64).reshape(-1, 2)
    # first call sizes the tape
    args = [dom.ptr, _p64(m), log_dim, log_inv_rate, log_batch, ar, len(arities), n_queries,
            _p64(ch_in) if ch_in is not None else None]
    rc = lib().bo_fri_prove_ex(*args, None, 0, C.byref(tape_len), _p8(roots), _p64(ch_out), _p64(term), threads,
                               *_scheme(scheme))
    if rc:
        raise ValueError(f"fri_prove error {rc}")
    tape = np.zeros(tape_len.value, dtype=np.uint8)
    rc = lib().bo_fri_prove_ex(*args, _p8(tape), tape.size, C.byref(tape_len), _p8(roots), _p64(ch_out),
                               _p64(term), threads, *_scheme(scheme))
    if rc:
        raise ValueError(f"fri_prove error {rc}")
    return {"tape": tape.tobytes(), "roots": [roots[i].tobytes() for i in range(len(arities) + 2)],
            "challenges": ch_out[:n_rounds], "terminal": term}


class Transcript:
    """ProverTranscript<HasherChallenger<Sha256>> (transcript.rs:205-420)."""

    def __init__(self):
        self.ptr = lib().bo_transcript_new()

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().bo_transcript_free(self.ptr)
            self.ptr = None

    def _call(self, name, data: bytes):
        buf = np.frombuffer(data, dtype=np.uint8).copy() if data else np.zeros(1, dtype=np.uint8)
        getattr(lib(), name)(self.ptr, _p8(buf), len(data))

    def message(self, data: bytes):
        self._call("bo_transcript_message", data)

    def observe(self, data: bytes):
        self._call("bo_transcript_observe", data)

    def decommit(self, data: bytes):
        self._call("bo_transcript_decommit", data)

    def sample(self, n: int) -> bytes:
        out = np.zeros(max(n, 1), dtype=np.uint8)
        lib().bo_transcript_sample(self.ptr, _p8(out), n)
        return out[:n].tobytes()

    def sample_bits(self, bits: int) -> int:
        return lib().bo_transcript_sample_bits(self.ptr, bits)

    def tape(self) -> bytes:
        n = lib().bo_transcript_tape(self.ptr, None, 0)
        out = np.zeros(max(n, 1), dtype=np.uint8)
        lib().bo_transcript_tape(self.ptr, _p8(out), n)
        return out[:n].tobytes()


# ---- sumcheck / BaseFold / one-bit PCS ----------------------------------------
def sumcheck_prove(a: np.ndarray, b: np.ndarray, claim: int, transcript: "Transcript",
                   threads: int = 1) -> dict:
    """prove_single over BivariateProductSumcheckProver: challenges and (a(r), b(r))."""
    wa = np.array(a, dtype=np.uint64).reshape(-1, 2)
    wb = np.array(b, dtype=np.uint64).reshape(-1, 2)
    assert wa.shape == wb.shape
    n_vars = wa.shape[0].bit_length() - 1
    ch = np.zeros((max(n_vars, 1), 2), dtype=np.uint64)
    ev = np.zeros(4, dtype=np.uint64)
    s = to_elems([claim])
    if lib().bo_sumcheck_prove(_p64(wa), _p64(wb), n_vars, _p64(s), transcript.ptr, _p64(ch), _p64(ev), threads):
        raise ValueError("sumcheck error")
    return {"challenges": ch[:n_vars], "evals": to_ints(ev)}


def _arities(arities):
    return (C.c_int * max(len(arities), 1))(*arities)


def basefold_prove(dom: Domain, multilinear: np.ndarray, transparent: np.ndarray, claim: int,
                   codeword: np.ndarray, tree: np.ndarray, log_dim: int, log_inv_rate: int, log_batch: int,
                   arities, n_queries: int, transcript: "Transcript", threads: int = 1, scheme=None) -> np.ndarray:
    """BaseFoldProver::prove after the caller wrote the codeword root; returns the challenges."""
    wa = np.array(multilinear, dtype=np.uint64).reshape(-1, 2)
    wb = np.array(transparent, dtype=np.uint64).reshape(-1, 2)
    cw = np.ascontiguousarray(codeword, dtype=np.uint64).reshape(-1, 2)
    tr = np.ascontiguousarray(tree, dtype=np.uint8).reshape(-1)
    n_vars = log_dim + log_batch
    assert wa.shape[0] == wb.shape[0] == 1 << n_vars
    ch = np.zeros((max(n_vars, 1), 2), dtype=np.uint64)
    rc = lib().bo_basefold_prove_ex(dom.ptr, _p64(wa), _p64(wb), _p64(to_elems([claim])), _p64(cw), _p8(tr),
                                    log_dim, log_inv_rate, log_batch, _arities(arities), len(arities), n_queries,
                                    transcript.ptr, _p64(ch), threads, *_scheme(scheme))
    if rc:
        raise ValueError(f"basefold_prove error {rc}")
    return ch[:n_vars]


def tensor_transpose(elems: np.ndarray) -> np.ndarray:
    e = np.ascontiguousarray(elems, dtype=np.uint64).reshape(-1, 2)
    assert e.shape[0] == 128
    out = np.zeros_like(e)
    lib().bo_tensor_transpose(_p64(e), _p64(out))
    return out


def pcs_prove(dom: Domain, packed: np.ndarray, eval_point: np.ndarray, codeword: np.ndarray, tree: np.ndarray,
              log_dim: int, log_inv_rate: int, log_batch: int, arities, n_queries: int,
              transcript: "Transcript", threads: int = 1, scheme=None) -> None:
    """OneBitPCSProver::prove after the caller wrote the codeword root."""
    w = np.array(packed, dtype=np.uint64).reshape(-1, 2)
    pt = np.ascontiguousarray(eval_point, dtype=np.uint64).reshape(-1, 2)
    assert w.shape[0] == 1 << (log_dim + log_batch) and pt.shape[0] == log_dim + log_batch + 7
    cw = np.ascontiguousarray(codeword, dtype=np.uint64).reshape(-1, 2)
    tr = np.ascontiguousarray(tree, dtype=np.uint8).reshape(-1)
    rc = lib().bo_pcs_prove_ex(dom.ptr, _p64(w), _p64(pt), _p64(cw), _p8(tr), log_dim, log_inv_rate, log_batch,
                               _arities(arities), len(arities), n_queries, transcript.ptr, threads, *_scheme(scheme))
    if rc:
        raise ValueError(f"pcs_prove error {rc}")


# ---- Vision-M hash -------------------------------------------------------------
class Vision:
    """Vision-M permutation, sponge digest and (M = 4) compression, built from the
    hash's defining data (B_FWD_COEFFS and the round-key schedule inputs)."""

    def __init__(self, m: int, data: dict):
        def arr(key):
            return to_elems(int(x, 16) for x in data[key])
        self.m = m
        self._b, self._i, self._mat, self._c = (arr("b_fwd_coeffs"), arr("initial_constant"),
                                                arr("constants_matrix"), arr("constants_constant"))
        self.ptr = lib().bo_vision_new(m, _p64(self._b), _p64(self._i), _p64(self._mat), _p64(self._c))
        if not self.ptr:
            raise ValueError("vision setup failed")

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().bo_vision_free(self.ptr)
            self.ptr = None

    def round_constants(self):
        out = np.zeros((17 * self.m, 2), dtype=np.uint64)
        lib().bo_vision_round_constants(self.ptr, _p64(out))
        return to_ints(out)

    def b_maps(self, x: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, 2)
        f, i = np.zeros_like(x), np.zeros_like(x)
        lib().bo_vision_b_maps(self.ptr, _p64(x), _p64(f), _p64(i), x.shape[0])
        return f, i

    def permutation(self, states: np.ndarray) -> np.ndarray:
        st = np.array(states, dtype=np.uint64).reshape(-1, 2)
        assert st.shape[0] % self.m == 0
        lib().bo_vision_permutation(self.ptr, _p64(st), st.shape[0] // self.m)
        return st

    def digest(self, data: bytes) -> bytes:
        buf = np.frombuffer(data, dtype=np.uint8).copy() if data else np.zeros(1, dtype=np.uint8)
        out = np.zeros(32, dtype=np.uint8)
        lib().bo_vision_digest(self.ptr, _p8(buf), len(data), _p8(out))
        return out.tobytes()

    def compress(self, left: bytes, right: bytes) -> bytes:
        assert self.m == 4
        out = np.zeros(32, dtype=np.uint8)
        lib().bo_vision_compress(self.ptr, _p8(np.frombuffer(left, dtype=np.uint8).copy()),
                                 _p8(np.frombuffer(right, dtype=np.uint8).copy()), _p8(out))
        return out.tobytes()


def vision_batch_invert(x: np.ndarray) -> np.ndarray:
    v = np.array(x, dtype=np.uint64).reshape(-1, 2)
    lib().bo_vision_batch_invert(_p64(v), v.shape[0])
    return v


def merkle_build_vision(leaf: Vision, node: Vision, elems: np.ndarray, batch: int, threads: int = 1) -> np.ndarray:
    e = np.ascontiguousarray(elems, dtype=np.uint64).reshape(-1, 2)
    n_leaves = e.shape[0] // batch if batch else 0
    tree = np.zeros((max(2 * n_leaves - 1, 1), 32), dtype=np.uint8)
    rc = lib().bo_merkle_build_vision(leaf.ptr, node.ptr, _p64(e), e.shape[0], batch, _p8(tree), threads)
    if rc:
        raise ValueError(f"merkle error {rc}")
    return tree
