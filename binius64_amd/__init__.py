"""binius64_amd: MI355X implementation of Binius64's B128 polynomial-commitment
hot path (additive NTT / Reed-Solomon encode, SHA-256 Merkle commit, FRI folds).

This Python layer is a thin host mirror over the C ABI of libb64hip.so
(include/b64hip.h).  Names follow the reference's Rust interfaces:
  DomainContext / GenericPreExpanded  verifier/math/src/ntt/{mod.rs,domain_context.rs}
  AdditiveNTT                         verifier/math/src/ntt/mod.rs:58-113
  ReedSolomonCode.encode_batch        verifier/math/src/reed_solomon.rs:128-201
  BinaryMerkleTreeProver              prover/prover/src/merkle_tree/prover.rs
  FRIParams                           verifier/verifier/src/fri/common.rs
  commit_interleaved                  prover/prover/src/fri/commit.rs:25-75
  FRIFoldProver                       prover/prover/src/fri/fold.rs
  ProverTranscript                    verifier/transcript/src/transcript.rs
Field elements on the host are numpy uint64 arrays of shape (n, 2) = (lo, hi).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from ._abi import B64Error, FriParamsC, LIB_PATH, check, lib

__all__ = [
    "B64Error", "DeviceBuffer", "DomainContext", "AdditiveNTT", "ReedSolomonCode",
    "BinaryMerkleTreeProver", "FRIParams", "CommitOutput", "commit_interleaved", "setup_commit_params",
    "ProverTranscript", "FRIFoldProver", "ProveJob", "batch_prove", "gf_mul", "gf_inv", "gf_mul_device", "gf_mul_const_device",
    "pack_witness", "device_count", "config_set", "config_get", "ntt_plan", "LIB_PATH", "lib",
    "CommitJob", "BatchPlan", "Server", "fri_prove", "ring_switch_fold_1b_rows", "ring_switch_fold_elems",
    "Proof", "STD_CHALLENGER_NAME",
]


def _p64(a: np.ndarray):
    assert a.dtype == np.uint64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def _p8(a: np.ndarray):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def _stream(stream) -> Optional[int]:
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return int(getattr(stream, "cuda_stream", stream))


def device_count() -> int:
    return lib().b64_device_count()


def config_set(key: str, value: int) -> None:
    """Set a kernel tuning knob (see b64_config_set in include/b64hip.h)."""
    check(lib().b64_config_set(key.encode(), int(value)), f"config_set {key}")


def ntt_plan(log_len: int, skip_early: int, skip_late: int, inverse: bool = False, encode: bool = False,
             n_cus: int = 0) -> list:
    """The pass schedule the library runs for one transform (b64_ntt_plan): per
    pass a dict of layer0, n_layers, low_bits, extra_bits, kernel (name) and the
    work items of its grid on a device of n_cus compute units (0 = current)."""
    from ._abi import NTT_KERNEL_NAMES, NttPassInfoC
    n = C.c_uint32()
    check(lib().b64_ntt_plan(log_len, skip_early, skip_late, int(inverse), int(encode), n_cus, None, 0,
                             C.byref(n)), "ntt_plan")
    arr = (NttPassInfoC * max(1, n.value))()
    check(lib().b64_ntt_plan(log_len, skip_early, skip_late, int(inverse), int(encode), n_cus, arr, n.value,
                             C.byref(n)), "ntt_plan")
    return [{"layer0": p.layer0, "n_layers": p.n_layers, "low_bits": p.low_bits, "extra_bits": p.extra_bits,
             "kernel": NTT_KERNEL_NAMES[p.kernel], "workitems": p.workitems} for p in arr[:n.value]]


def config_get(key: str) -> int:
    v = C.c_int64()
    check(lib().b64_config_get(key.encode(), C.byref(v)), f"config_get {key}")
    return v.value


def _elems(x) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.uint64)
    return a.reshape(-1, 2)


def _u128(v: int) -> np.ndarray:
    return np.array([v & 0xFFFFFFFFFFFFFFFF, v >> 64], dtype=np.uint64)


def _release(obj, destroy: str) -> None:
    """Frees obj.ptr with the named destroy call.  During interpreter teardown
    the module's globals may already be gone; the process exit then frees the
    library's objects itself."""
    ptr = getattr(obj, "ptr", None)
    obj.ptr = None
    if ptr:
        try:
            getattr(lib(), destroy)(ptr)
        except Exception:
            pass


# ---------------------------------------------------------------------------
class DeviceBuffer:
    """An HBM allocation owned by the library's allocator."""

    def __init__(self, nbytes: int):
        ptr = C.c_void_p()
        check(lib().b64_device_alloc(nbytes, C.byref(ptr)), "device_alloc")
        self.ptr = ptr.value
        self.nbytes = nbytes

    @classmethod
    def from_host(cls, arr: np.ndarray, stream=None) -> "DeviceBuffer":
        a = np.ascontiguousarray(arr)
        buf = cls(max(a.nbytes, 16))
        check(lib().b64_memcpy_htod(buf.ptr, a.ctypes.data, a.nbytes, _stream(stream)), "htod")
        return buf

    def to_host(self, n_elems: Optional[int] = None, stream=None) -> np.ndarray:
        n = n_elems if n_elems is not None else self.nbytes // 16
        out = np.zeros((n, 2), dtype=np.uint64)
        check(lib().b64_memcpy_dtoh(out.ctypes.data, self.ptr, 16 * n, _stream(stream)), "dtoh")
        return out

    def to_bytes(self, nbytes: Optional[int] = None, stream=None) -> bytes:
        n = nbytes if nbytes is not None else self.nbytes
        out = np.zeros(n, dtype=np.uint8)
        check(lib().b64_memcpy_dtoh(out.ctypes.data, self.ptr, n, _stream(stream)), "dtoh")
        return out.tobytes()

    def free(self):
        if getattr(self, "ptr", None):
            lib().b64_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------------
def gf_mul(a: int, b: int) -> int:
    out = np.zeros(2, dtype=np.uint64)
    lib().b64_gf_mul(_p64(_u128(a)), _p64(_u128(b)), _p64(out))
    return int(out[0]) | (int(out[1]) << 64)


def gf_inv(a: int) -> int:
    out = np.zeros(2, dtype=np.uint64)
    lib().b64_gf_inv(_p64(_u128(a)), _p64(out))
    return int(out[0]) | (int(out[1]) << 64)


def gf_mul_device(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a, b = _elems(a), _elems(b)
    da, db = DeviceBuffer.from_host(a), DeviceBuffer.from_host(b)
    do = DeviceBuffer(max(a.nbytes, 16))
    check(lib().b64_gf_mul_device(da.ptr, db.ptr, do.ptr, a.shape[0], None), "gf_mul_device")
    return do.to_host(a.shape[0])


def gf_mul_const_device(a: np.ndarray, c: int) -> np.ndarray:
    a = _elems(a)
    da = DeviceBuffer.from_host(a)
    do = DeviceBuffer(max(a.nbytes, 16))
    check(lib().b64_gf_mul_const_device(da.ptr, _p64(_u128(c)), do.ptr, a.shape[0], None), "gf_mul_const")
    return do.to_host(a.shape[0])


# ---------------------------------------------------------------------------
class DomainContext:
    """GenericPreExpanded domain (ntt/domain_context.rs:99-159)."""

    def __init__(self, basis=None, dim: Optional[int] = None, _ptr=None):
        if _ptr is not None:
            self.ptr = _ptr
        else:
            out = C.c_void_p()
            if basis is None:
                check(lib().b64_domain_create_standard(dim, C.byref(out)), "domain_create_standard")
            else:
                b = _elems(basis)
                check(lib().b64_domain_create(_p64(b), b.shape[0], C.byref(out)), "domain_create")
            self.ptr = out.value

    @classmethod
    def standard(cls, dim: int) -> "DomainContext":
        return cls(dim=dim)

    def __del__(self):
        _release(self, "b64_domain_destroy")

    def log_domain_size(self) -> int:
        return lib().b64_domain_log_size(self.ptr)

    def subspace(self, i: int) -> np.ndarray:
        out = np.zeros((max(i, 1), 2), dtype=np.uint64)
        check(lib().b64_domain_subspace(self.ptr, i, _p64(out)), "subspace")
        return out[:i]

    def twiddle(self, layer: int, block: int) -> int:
        out = np.zeros(2, dtype=np.uint64)
        check(lib().b64_domain_twiddle(self.ptr, layer, block, _p64(out)), "twiddle")
        return int(out[0]) | (int(out[1]) << 64)


class AdditiveNTT:
    """The AdditiveNTT trait (ntt/mod.rs:58-113) on device buffers."""

    def __init__(self, domain_context: DomainContext):
        self.domain_context = domain_context

    def log_domain_size(self) -> int:
        return self.domain_context.log_domain_size()

    def subspace(self, i: int) -> np.ndarray:
        return self.domain_context.subspace(i)

    def twiddle(self, layer: int, block: int) -> int:
        return self.domain_context.twiddle(layer, block)

    def forward_transform(self, data: DeviceBuffer, log_len: int, skip_early: int, skip_late: int, stream=None):
        check(lib().b64_ntt_forward(self.domain_context.ptr, data.ptr, log_len, skip_early, skip_late,
                                    _stream(stream)), "ntt_forward")

    def inverse_transform(self, data: DeviceBuffer, log_len: int, skip_early: int, skip_late: int, stream=None):
        check(lib().b64_ntt_inverse(self.domain_context.ptr, data.ptr, log_len, skip_early, skip_late,
                                    _stream(stream)), "ntt_inverse")


class ReedSolomonCode:
    """ReedSolomonCode::with_domain_context_subspace (reed_solomon.rs:54-65)."""

    def __init__(self, domain_context: DomainContext, log_dim: int, log_inv_rate: int):
        self.domain_context = domain_context
        self.log_dimension = log_dim
        self.log_inv_rate = log_inv_rate

    def log_dim(self) -> int:
        return self.log_dimension

    def log_len(self) -> int:
        return self.log_dimension + self.log_inv_rate

    def encode_batch(self, data: DeviceBuffer, log_batch_size: int, out: Optional[DeviceBuffer] = None,
                     stream=None) -> DeviceBuffer:
        n_out = 1 << (self.log_len() + log_batch_size)
        out = out or DeviceBuffer(16 * n_out)
        check(lib().b64_rs_encode_batch(self.domain_context.ptr, self.log_dimension, self.log_inv_rate, data.ptr,
                                        log_batch_size, out.ptr, _stream(stream)), "encode_batch")
        return out


class BinaryMerkleTreeProver:
    """BinaryMerkleTreeProver<B128, Sha256, Sha256Compression> (merkle_tree/prover.rs)."""

    @staticmethod
    def commit(data: DeviceBuffer, n_elems: int, batch_size: int, tree: Optional[DeviceBuffer] = None,
               stream=None):
        n_leaves = n_elems // batch_size if batch_size else 0
        tree = tree or DeviceBuffer(32 * max(2 * n_leaves - 1, 1))
        root = np.zeros(32, dtype=np.uint8)
        check(lib().b64_merkle_commit(data.ptr, n_elems, batch_size, tree.ptr, _p8(root), _stream(stream)),
              "merkle_commit")
        return root.tobytes(), tree

    @staticmethod
    def layer(tree: DeviceBuffer, log_len: int, depth: int, stream=None) -> List[bytes]:
        out = np.zeros(32 << depth, dtype=np.uint8)
        check(lib().b64_merkle_layer(tree.ptr, log_len, depth, _p8(out), _stream(stream)), "merkle_layer")
        b = out.tobytes()
        return [b[32 * i:32 * i + 32] for i in range(1 << depth)]

    @staticmethod
    def branch(tree: DeviceBuffer, log_len: int, index: int, depth: int, stream=None) -> List[bytes]:
        n = max(log_len - depth, 1)
        out = np.zeros(32 * n, dtype=np.uint8)
        check(lib().b64_merkle_branch(tree.ptr, log_len, index, depth, _p8(out), _stream(stream)), "merkle_branch")
        b = out.tobytes()
        return [b[32 * i:32 * i + 32] for i in range(log_len - depth)]


# ---------------------------------------------------------------------------
@dataclass
class FRIParams:
    log_dim: int
    log_inv_rate: int
    log_batch_size: int
    n_test_queries: int
    fold_arities: List[int] = field(default_factory=list)
    log_terminal_dim: int = 0

    @classmethod
    def from_c(cls, c: FriParamsC) -> "FRIParams":
        return cls(c.log_dim, c.log_inv_rate, c.log_batch_size, c.n_test_queries,
                   [c.fold_arities[i] for i in range(c.n_fold_arities)], c.log_terminal_dim)

    def to_c(self) -> FriParamsC:
        c = FriParamsC()
        arr = (C.c_uint32 * max(len(self.fold_arities), 1))(*self.fold_arities)
        check(lib().b64_fri_params_new(self.log_dim, self.log_inv_rate, self.log_batch_size, arr,
                                       len(self.fold_arities), self.n_test_queries, C.byref(c)), "fri_params_new")
        return c

    @classmethod
    def new(cls, log_dim, log_inv_rate, log_batch_size, fold_arities, n_test_queries) -> "FRIParams":
        p = cls(log_dim, log_inv_rate, log_batch_size, n_test_queries, list(fold_arities))
        return cls.from_c(p.to_c())

    @classmethod
    def choose_with_constant_fold_arity(cls, domain: DomainContext, log_msg_len: int, security_bits: int,
                                        log_inv_rate: int, arity: int) -> "FRIParams":
        c = FriParamsC()
        check(lib().b64_fri_params_choose(domain.ptr, log_msg_len, security_bits, log_inv_rate, arity,
                                          C.byref(c)), "fri_params_choose")
        return cls.from_c(c)

    # accessors named as in fri/common.rs:117-150
    def rs_log_len(self) -> int:
        return self.log_dim + self.log_inv_rate

    def log_len(self) -> int:
        return self.rs_log_len() + self.log_batch_size

    def log_msg_len(self) -> int:
        return self.log_dim + self.log_batch_size

    def n_fold_rounds(self) -> int:
        return self.log_msg_len()

    def n_oracles(self) -> int:
        return 1 + len(self.fold_arities)

    def index_bits(self) -> int:
        return self.rs_log_len()

    def n_final_challenges(self) -> int:
        return self.log_terminal_dim


def calculate_n_test_queries(security_bits: int, log_inv_rate: int) -> int:
    return lib().b64_calculate_n_test_queries(security_bits, log_inv_rate)


def estimate_optimal_arity(log_block_length: int, digest_size: int, field_size: int) -> int:
    return lib().b64_estimate_optimal_arity(log_block_length, digest_size, field_size)


def setup_commit_params(log_witness_elems: int, log_inv_rate: int = 1, security_bits: int = 96):
    """Verifier::setup + Prover::setup chain (verify.rs:61-106, prove.rs:93-115)."""
    c = FriParamsC()
    dom = C.c_void_p()
    check(lib().b64_setup_commit_params(log_witness_elems, log_inv_rate, security_bits, C.byref(c), C.byref(dom)),
          "setup_commit_params")
    return FRIParams.from_c(c), DomainContext(_ptr=dom.value)


@dataclass
class CommitOutput:
    commitment: bytes
    codeword: DeviceBuffer
    tree: DeviceBuffer


# Merkle schemes (the MerkleTreeProver parameter of the reference's provers)
MERKLE_SHA256, MERKLE_VISION4, MERKLE_VISION6 = 0, 4, 6


def commit_interleaved(params: FRIParams, domain: DomainContext, message: DeviceBuffer, stream=None,
                       merkle_hash: int = MERKLE_SHA256) -> CommitOutput:
    cw = DeviceBuffer(16 << params.log_len())
    tree = DeviceBuffer(32 * ((2 << params.rs_log_len()) - 1))
    root = np.zeros(32, dtype=np.uint8)
    pc = params.to_c()
    check(lib().b64_fri_commit_interleaved_ex(C.byref(pc), domain.ptr, merkle_hash, message.ptr, cw.ptr, tree.ptr,
                                              _p8(root), _stream(stream)), "commit_interleaved")
    return CommitOutput(root.tobytes(), cw, tree)


class ProverTranscript:
    """ProverTranscript<HasherChallenger<Sha256>> (transcript.rs:205-420)."""

    def __init__(self):
        out = C.c_void_p()
        check(lib().b64_transcript_create(C.byref(out)), "transcript_create")
        self.ptr = out.value

    def __del__(self):
        _release(self, "b64_transcript_destroy")

    def _buf(self, data: bytes):
        return np.frombuffer(data, dtype=np.uint8).copy() if data else np.zeros(1, dtype=np.uint8)

    def message(self, data: bytes):
        lib().b64_transcript_message(self.ptr, _p8(self._buf(data)), len(data))

    def observe(self, data: bytes):
        lib().b64_transcript_observe(self.ptr, _p8(self._buf(data)), len(data))

    def decommit(self, data: bytes):
        lib().b64_transcript_decommit(self.ptr, _p8(self._buf(data)), len(data))

    def sample_bytes(self, n: int) -> bytes:
        out = np.zeros(max(n, 1), dtype=np.uint8)
        lib().b64_transcript_sample(self.ptr, _p8(out), n)
        return out[:n].tobytes()

    def sample(self) -> int:
        """CanSample<B128>: 16 sampled bytes as a little-endian u128."""
        return int.from_bytes(self.sample_bytes(16), "little")

    def sample_bits(self, bits: int) -> int:
        return lib().b64_transcript_sample_bits(self.ptr, bits)

    def finalize(self) -> bytes:
        n = lib().b64_transcript_tape(self.ptr, None, 0)
        out = np.zeros(max(n, 1), dtype=np.uint8)
        lib().b64_transcript_tape(self.ptr, _p8(out), n)
        return out[:n].tobytes()


# StdChallenger::NAME (verifier/verifier/src/config.rs:19-21): the name the
# reference's prover CLI puts in a Proof.
STD_CHALLENGER_NAME = "HasherChallenger<Sha256>"


class Proof:
    """Proof (verifier/core/src/constraint_system.rs:1014-1100): transcript bytes
    and the name of the challenger that produced them, with the reference's
    wire format (b64_proof_serialize / b64_proof_deserialize).  Serialisation
    errors raise B64Error with the SerializationError counterpart's code."""

    SERIALIZATION_VERSION = 1

    def __init__(self, data: bytes, challenger_type: str):
        self._data = bytes(data)
        self._challenger = str(challenger_type)

    @classmethod
    def owned(cls, data: bytes, challenger_type: str) -> "Proof":
        return cls(data, challenger_type)

    borrowed = owned

    def as_slice(self) -> bytes:
        return self._data

    def challenger_type(self) -> str:
        return self._challenger

    def len(self) -> int:
        return len(self._data)

    __len__ = len

    def is_empty(self) -> bool:
        return not self._data

    def into_owned(self):
        return self._data, self._challenger

    def to_owned(self) -> "Proof":
        return Proof(self._data, self._challenger)

    def __eq__(self, other) -> bool:
        return isinstance(other, Proof) and self.into_owned() == other.into_owned()

    def serialize(self) -> bytes:
        ct = self._challenger.encode("utf-8")
        if b"\0" in ct:
            raise ValueError("challenger type with a NUL byte cannot cross the C ABI")
        data = np.frombuffer(self._data, dtype=np.uint8).copy() if self._data else np.zeros(1, dtype=np.uint8)
        n = C.c_uint64()
        lib().b64_proof_serialize(_p8(data), len(self._data), ct, None, 0, C.byref(n))
        out = np.zeros(max(n.value, 1), dtype=np.uint8)
        check(lib().b64_proof_serialize(_p8(data), len(self._data), ct, _p8(out), n.value, C.byref(n)),
              "proof_serialize")
        return out[:n.value].tobytes()

    @classmethod
    def deserialize(cls, buf: bytes) -> "Proof":
        arr = np.frombuffer(bytes(buf), dtype=np.uint8).copy() if buf else np.zeros(1, dtype=np.uint8)
        co, cl, do, dl = (C.c_uint64() for _ in range(4))
        check(lib().b64_proof_deserialize(_p8(arr), len(buf), C.byref(co), C.byref(cl), C.byref(do), C.byref(dl)),
              "proof_deserialize")
        raw = arr.tobytes()
        return cls(raw[do.value:do.value + dl.value], raw[co.value:co.value + cl.value].decode("utf-8"))


class FRIFoldProver:
    """FRIFoldProver (prover/prover/src/fri/fold.rs:31-245) over device-resident oracles."""

    def __init__(self, params: FRIParams, domain: DomainContext, codeword: DeviceBuffer, tree: DeviceBuffer,
                 stream=None, merkle_hash: int = MERKLE_SHA256):
        self._keep = (domain, codeword, tree)
        self._pc = params.to_c()
        out = C.c_void_p()
        check(lib().b64_fri_prover_create_ex(C.byref(self._pc), domain.ptr, merkle_hash, codeword.ptr, tree.ptr,
                                             _stream(stream), C.byref(out)), "fri_prover_create")
        self.ptr = out.value

    def __del__(self):
        _release(self, "b64_fri_prover_destroy")

    def n_rounds(self) -> int:
        return lib().b64_fri_prover_n_rounds(self.ptr)

    def curr_round(self) -> int:
        return lib().b64_fri_prover_curr_round(self.ptr)

    def receive_challenge(self, challenge: int):
        check(lib().b64_fri_prover_receive_challenge(self.ptr, _p64(_u128(challenge))), "receive_challenge")

    def execute_fold_round(self) -> Optional[bytes]:
        root = np.zeros(32, dtype=np.uint8)
        committed = C.c_int(0)
        check(lib().b64_fri_prover_execute_fold_round(self.ptr, _p8(root), C.byref(committed)), "fold_round")
        return root.tobytes() if committed.value else None

    def finish_proof(self, transcript: ProverTranscript):
        check(lib().b64_fri_prover_finish_proof(self.ptr, transcript.ptr), "finish_proof")

    def oracle(self, k: int) -> np.ndarray:
        ptr, log_len, tree = C.c_void_p(), C.c_uint32(), C.c_void_p()
        check(lib().b64_fri_prover_oracle(self.ptr, k, C.byref(ptr), C.byref(log_len), C.byref(tree)), "oracle")
        n = 1 << log_len.value
        out = np.zeros((n, 2), dtype=np.uint64)
        check(lib().b64_memcpy_dtoh(out.ctypes.data, ptr.value, 16 * n, None), "dtoh")
        return out


def pack_witness(words: DeviceBuffer, n_words: int, log_elems: int, out: Optional[DeviceBuffer] = None,
                 stream=None) -> DeviceBuffer:
    out = out or DeviceBuffer(16 << log_elems)
    check(lib().b64_pack_witness(words.ptr, n_words, log_elems, out.ptr, _stream(stream)), "pack_witness")
    return out


def eq_ind_partial_eval(point, out: Optional[DeviceBuffer] = None, stream=None) -> DeviceBuffer:
    """eq_ind_partial_eval (verifier/math/src/multilinear/eq.rs:123-131) into HBM."""
    pt = np.ascontiguousarray(_elems(point))
    n = pt.shape[0]
    out = out or DeviceBuffer(16 << n)
    check(lib().b64_eq_ind_partial_eval(_p64(pt), n, out.ptr, _stream(stream)), "eq_ind_partial_eval")
    return out


def _ints(arr: np.ndarray) -> List[int]:
    return [int(lo) | (int(hi) << 64) for lo, hi in np.asarray(arr, dtype=np.uint64).reshape(-1, 2)]


class BivariateProductSumcheckProver:
    """BivariateProductSumcheckProver (prover protocols/sumcheck/bivariate_product.rs).

    The prover folds both multilinears in place: like the reference, which takes
    the buffers by value, it consumes a and b."""

    def __init__(self, a: DeviceBuffer, b: DeviceBuffer, n_vars: int, claim: int, stream=None):
        self._keep = (a, b)
        out = C.c_void_p()
        check(lib().b64_sumcheck_create(a.ptr, b.ptr, n_vars, _p64(_u128(claim)), _stream(stream), C.byref(out)),
              "sumcheck_create")
        self.ptr = out.value

    def __del__(self):
        _release(self, "b64_sumcheck_destroy")

    def n_vars(self) -> int:
        return lib().b64_sumcheck_n_vars(self.ptr)

    def execute(self) -> List[int]:
        co = np.zeros(6, dtype=np.uint64)
        check(lib().b64_sumcheck_execute(self.ptr, _p64(co)), "sumcheck_execute")
        return _ints(co)

    def fold(self, challenge: int):
        check(lib().b64_sumcheck_fold(self.ptr, _p64(_u128(challenge))), "sumcheck_fold")

    def finish(self) -> List[int]:
        ev = np.zeros(4, dtype=np.uint64)
        check(lib().b64_sumcheck_finish(self.ptr, _p64(ev)), "sumcheck_finish")
        return _ints(ev)


def prove_single(prover: BivariateProductSumcheckProver, transcript: ProverTranscript) -> dict:
    """prove_single (prover protocols/sumcheck/prove.rs:17-60)."""
    challenges = []
    for _ in range(prover.n_vars()):
        c0, c1, _c2 = prover.execute()
        transcript.message(c0.to_bytes(16, "little") + c1.to_bytes(16, "little"))
        r = transcript.sample()
        challenges.append(r)
        prover.fold(r)
    return {"challenges": challenges, "multilinear_evals": prover.finish()}


def basefold_prove(params: FRIParams, domain: DomainContext, multilinear: DeviceBuffer, transparent: DeviceBuffer,
                   claim: int, codeword: DeviceBuffer, tree: DeviceBuffer, transcript: ProverTranscript,
                   stream=None, merkle_hash: int = MERKLE_SHA256) -> None:
    """BaseFoldProver::new(...).prove(transcript) (prover protocols/basefold.rs:64-151).
    multilinear and transparent are consumed."""
    pc = params.to_c()
    check(lib().b64_basefold_prove_ex(C.byref(pc), domain.ptr, merkle_hash, multilinear.ptr, transparent.ptr,
                                      _p64(_u128(claim)), codeword.ptr, tree.ptr, transcript.ptr, _stream(stream)),
          "basefold_prove")


class OneBitPCSProver:
    """OneBitPCSProver (prover/prover/src/pcs.rs:35-171): commit and prove an
    evaluation of a multilinear over B1 packed 128 bits per B128 element."""

    def __init__(self, domain: DomainContext, params: FRIParams, merkle_hash: int = MERKLE_SHA256):
        self.domain, self.params, self.merkle_hash = domain, params, merkle_hash

    def commit(self, packed: DeviceBuffer, stream=None) -> CommitOutput:
        return commit_interleaved(self.params, self.domain, packed, stream, merkle_hash=self.merkle_hash)

    def prove(self, codeword: DeviceBuffer, tree: DeviceBuffer, packed: DeviceBuffer, evaluation_point,
              transcript: ProverTranscript, stream=None) -> None:
        """packed is consumed (folded in place by the sumcheck)."""
        pt = np.ascontiguousarray(_elems(evaluation_point))
        pc = self.params.to_c()
        check(lib().b64_pcs_prove_ex(C.byref(pc), self.domain.ptr, self.merkle_hash, codeword.ptr, tree.ptr,
                                     packed.ptr, _p64(pt), pt.shape[0], transcript.ptr, _stream(stream)), "pcs_prove")


class VisionHasher:
    """Vision-M (M = 4 or 6) permutation and sponge digest on device buffers
    (verifier/verifier/src/hash/vision_{4,6})."""

    def __init__(self, m: int):
        if m not in (4, 6):
            raise ValueError("Vision width must be 4 or 6")
        self.m = m

    def permute(self, states: DeviceBuffer, n: int, stream=None) -> None:
        check(lib().b64_vision_permute(self.m, states.ptr, n, _stream(stream)), "vision_permute")

    def digest_many(self, msgs: DeviceBuffer, n: int, msg_elems: int, out: Optional[DeviceBuffer] = None,
                    stream=None) -> DeviceBuffer:
        out = out or DeviceBuffer(32 * max(n, 1))
        check(lib().b64_vision_digest(self.m, msgs.ptr, n, msg_elems, out.ptr, _stream(stream)), "vision_digest")
        return out


def merkle_commit_vision(data: DeviceBuffer, n_elems: int, batch_size: int, leaf_width: int = 4,
                         tree: Optional[DeviceBuffer] = None, stream=None):
    """BinaryMerkleTreeProver::commit with Vision leaves and VisionCompression nodes."""
    n_leaves = n_elems // batch_size if batch_size else 0
    tree = tree or DeviceBuffer(32 * max(2 * n_leaves - 1, 1))
    root = np.zeros(32, dtype=np.uint8)
    check(lib().b64_merkle_commit_vision(leaf_width, data.ptr, n_elems, batch_size, tree.ptr, _p8(root),
                                         _stream(stream)), "merkle_commit_vision")
    return root.tobytes(), tree


class CommitJob:
    """One FRI commit phase with pre-seeded challenges (b64_job_desc): encode,
    Merkle commit, every fold round and its commitment."""

    def __init__(self, log_msg_len: int, message, challenges: np.ndarray):
        from ._abi import JobDescC
        self.log_msg_len = log_msg_len
        if isinstance(message, DeviceBuffer):
            self.msg = message
        else:
            self.msg = DeviceBuffer.from_host(_elems(message))
        self.challenges = np.ascontiguousarray(_elems(challenges))
        assert self.challenges.shape[0] >= log_msg_len
        ws = lib().b64_job_workspace_size(log_msg_len)
        if ws == 0:
            raise B64Error(1, "job_workspace_size")
        self.workspace = DeviceBuffer(ws)
        self.n_roots = lib().b64_job_n_roots(log_msg_len)
        self.roots_buf = np.zeros(32 * self.n_roots, dtype=np.uint8)
        self.desc = JobDescC(log_msg_len, self.msg.ptr, _p64(self.challenges), self.workspace.ptr,
                             _p8(self.roots_buf))

    @staticmethod
    def workspace_size(log_msg_len: int) -> int:
        return lib().b64_job_workspace_size(log_msg_len)

    @staticmethod
    def n_roots_for(log_msg_len: int) -> int:
        return lib().b64_job_n_roots(log_msg_len)

    def run(self, stream=None):
        check(lib().b64_job_run(C.byref(self.desc), _stream(stream)), "job_run")

    def roots(self) -> List[bytes]:
        b = self.roots_buf.tobytes()
        return [b[32 * i:32 * i + 32] for i in range(self.n_roots)]

    def oracle(self, k: int) -> np.ndarray:
        ptr, log_len = C.c_void_p(), C.c_uint32()
        check(lib().b64_job_oracle(C.byref(self.desc), k, C.byref(ptr), C.byref(log_len)), "job_oracle")
        n = 1 << log_len.value
        out = np.zeros((n, 2), dtype=np.uint64)
        check(lib().b64_memcpy_dtoh(out.ctypes.data, ptr.value, 16 * n, None), "dtoh")
        return out


class BatchPlan:
    """Many CommitJobs per launch, optionally replayed from a hipGraph."""

    def __init__(self, jobs: List[CommitJob], use_graph: bool = True, stream=None):
        from ._abi import JobDescC
        self.jobs = list(jobs)
        arr = (JobDescC * len(self.jobs))(*[j.desc for j in self.jobs])
        out = C.c_void_p()
        check(lib().b64_batch_plan_create(arr, len(self.jobs), 1 if use_graph else 0, _stream(stream),
                                          C.byref(out)), "batch_plan_create")
        self.ptr = out.value

    def n_launches(self) -> int:
        return lib().b64_batch_plan_n_launches(self.ptr)

    def run(self, stream=None):
        check(lib().b64_batch_plan_run(self.ptr, _stream(stream)), "batch_plan_run")

    def __del__(self):
        _release(self, "b64_batch_plan_destroy")


class ProveJob:
    """One message to commit and prove with FRI under its own transcript
    (b64_prove_job): what commit_interleaved, the root written to the transcript,
    FRIFoldProver's rounds with sampled challenges and finish_proof do for one
    message, with the Verifier::setup parameters of its size."""

    def __init__(self, log_msg_len: int, message, transcript: Optional[ProverTranscript] = None):
        from ._abi import ProveJobC
        self.log_msg_len = log_msg_len
        self.msg = message if isinstance(message, DeviceBuffer) else DeviceBuffer.from_host(_elems(message))
        ws = lib().b64_job_workspace_size(log_msg_len)
        if ws == 0:
            raise B64Error(1, "job_workspace_size")
        self.workspace = DeviceBuffer(ws)
        self.transcript = transcript if transcript is not None else ProverTranscript()
        self.n_roots = lib().b64_job_n_roots(log_msg_len)
        self.roots_buf = np.zeros(32 * self.n_roots, dtype=np.uint8)
        self.desc = ProveJobC(log_msg_len, self.msg.ptr, self.workspace.ptr, self.transcript.ptr,
                              _p8(self.roots_buf))

    def roots(self) -> List[bytes]:
        b = self.roots_buf.tobytes()
        return [b[32 * i:32 * i + 32] for i in range(self.n_roots)]

    def tape(self) -> bytes:
        return self.transcript.finalize()


def batch_prove(jobs: List[ProveJob], stream=None) -> None:
    """b64_batch_prove: every job's commitment and FRI proof, one launch per
    stage across the jobs of each round."""
    from ._abi import ProveJobC
    arr = (ProveJobC * len(jobs))(*[j.desc for j in jobs])
    check(lib().b64_batch_prove(arr, len(jobs), _stream(stream)), "batch_prove")


class Server:
    """Continuous-batching commit server (b64_server_*): jobs are admitted by HBM
    footprint and batched onto up to n_streams streams."""

    def __init__(self, hbm_budget_bytes: int, max_batch: int = 64, n_streams: int = 2):
        from ._abi import ServerConfigC
        cfg = ServerConfigC(hbm_budget_bytes, max_batch, n_streams)
        out = C.c_void_p()
        check(lib().b64_server_create(C.byref(cfg), C.byref(out)), "server_create")
        self.ptr = out.value
        self._held = {}

    def submit(self, log_msg_len: int, message: np.ndarray, challenges: np.ndarray) -> int:
        """Queue a job; the message array is held until poll/wait reports it done
        (the server reads it when the job is admitted)."""
        m = np.ascontiguousarray(_elems(message))
        ch = np.ascontiguousarray(_elems(challenges))
        assert m.shape[0] == 1 << log_msg_len and ch.shape[0] >= log_msg_len
        t = C.c_uint64()
        check(lib().b64_server_submit(self.ptr, log_msg_len, _p64(m), _p64(ch), C.byref(t)), "server_submit")
        self._held[t.value] = m
        return t.value

    def submit_proof(self, log_msg_len: int, message: np.ndarray, transcript: ProverTranscript) -> int:
        """Queue a proof job (b64_server_submit_proof): its commitment and FRI proof
        are written to `transcript`, which is held (with the message) until
        poll/wait reports the job done."""
        m = np.ascontiguousarray(_elems(message))
        assert m.shape[0] == 1 << log_msg_len
        t = C.c_uint64()
        check(lib().b64_server_submit_proof(self.ptr, log_msg_len, _p64(m), transcript.ptr, C.byref(t)),
              "server_submit_proof")
        self._held[t.value] = (m, transcript)
        return t.value

    def poll(self, ticket: int, n_roots: int):
        roots = np.zeros(32 * n_roots, dtype=np.uint8)
        done, lat = C.c_int(0), C.c_double(0)
        check(lib().b64_server_poll(self.ptr, ticket, _p8(roots), C.byref(done), C.byref(lat)), "server_poll")
        if not done.value:
            return None
        self._held.pop(ticket, None)
        b = roots.tobytes()
        return [b[32 * i:32 * i + 32] for i in range(n_roots)], lat.value

    def wait(self, ticket: int, n_roots: int):
        roots = np.zeros(32 * n_roots, dtype=np.uint8)
        lat = C.c_double(0)
        check(lib().b64_server_wait(self.ptr, ticket, _p8(roots), C.byref(lat)), "server_wait")
        self._held.pop(ticket, None)
        b = roots.tobytes()
        return [b[32 * i:32 * i + 32] for i in range(n_roots)], lat.value

    def stats(self):
        d, b, p = C.c_uint64(), C.c_uint64(), C.c_size_t()
        check(lib().b64_server_stats(self.ptr, C.byref(d), C.byref(b), C.byref(p)), "server_stats")
        return {"jobs_done": d.value, "batches": b.value, "peak_pool_bytes": p.value}

    def close(self):
        # destroy drains every queued and running job before it returns
        if getattr(self, "ptr", None):
            lib().b64_server_destroy(self.ptr)
            self.ptr = None
        self._held = {}

    def __del__(self):
        self.close()


def ring_switch_fold_1b_rows(mat: DeviceBuffer, vec: DeviceBuffer, log_len: int,
                             out: Optional[DeviceBuffer] = None, stream=None) -> DeviceBuffer:
    """ring_switch::fold_1b_rows_for_b128 (prover/prover/src/ring_switch.rs:197-265):
    128 elements, out[b] = sum_r vec[r] * bit_b(mat[r])."""
    out = out or DeviceBuffer(16 * 128)
    check(lib().b64_ring_switch_fold_1b_rows(mat.ptr, vec.ptr, log_len, out.ptr, _stream(stream)),
          "ring_switch_fold_1b_rows")
    return out


def ring_switch_fold_elems(elems: DeviceBuffer, log_len: int, vec: np.ndarray, stream=None) -> DeviceBuffer:
    """ring_switch::fold_b128_elems_inplace (ring_switch.rs:130-176), in place:
    x <- sum_b bit_b(x) * vec[b] for 128 host elements vec."""
    v = np.ascontiguousarray(_elems(vec))
    assert v.shape[0] == 128
    check(lib().b64_ring_switch_fold_elems(elems.ptr, log_len, _p64(v), _stream(stream)), "ring_switch_fold_elems")
    return elems


def fri_prove(params: FRIParams, domain: DomainContext, message: np.ndarray, challenges=None,
              merkle_hash: int = MERKLE_SHA256):
    """FRI-only prover flow (prover/prover/src/fri/tests.rs:58-89): commit, fold
    rounds with transcript challenges (or the given ones), finish_proof.
    Returns (proof tape, roots, challenges)."""
    msg = DeviceBuffer.from_host(_elems(message))
    co = commit_interleaved(params, domain, msg, merkle_hash=merkle_hash)
    prover = FRIFoldProver(params, domain, co.codeword, co.tree, merkle_hash=merkle_hash)
    t = ProverTranscript()
    t.message(co.commitment)
    roots = [co.commitment]
    r = prover.execute_fold_round()
    if r is not None:
        t.message(r)
        roots.append(r)
    used = []
    for i in range(params.n_fold_rounds()):
        ch = t.sample() if challenges is None else int(challenges[i])
        used.append(ch)
        prover.receive_challenge(ch)
        r = prover.execute_fold_round()
        if r is not None:
            t.message(r)
            roots.append(r)
    prover.finish_proof(t)
    return t.finalize(), roots, used
