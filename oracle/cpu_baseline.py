"""ctypes view of oracle/libb64cpu.so, the CPU baseline: the reference prover's
x86-64 commit path (4-wide VPCLMULQDQ multiply, SHA extensions, a persistent
pool of all host threads).  Benchmark/test infrastructure: bench.py's
cpu_baseline leg times it, tests/ check it against the scalar oracle."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_lib = None


def lib():
    global _lib
    if _lib is None:
        san = os.environ.get("B64_SANITIZED_LIBS")  # sanitizer builds (tests/sanitize/run.sh)
        path = Path(san) / "libb64cpu.so" if san else HERE / "libb64cpu.so"
        if not path.exists():
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(str(path))
        vp, u64p, u8p = C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint8)
        L.cb_create.restype = vp
        L.cb_create.argtypes = [u64p, C.c_int, C.c_int]
        L.cb_destroy.argtypes = [vp]
        L.cb_features.restype = C.c_int
        L.cb_restrict.argtypes = [C.c_int]
        L.cb_ntt_forward.argtypes = [vp, u64p, C.c_int, C.c_int, C.c_int]
        L.cb_encode_batch.argtypes = [vp, u64p, C.c_int, C.c_int, C.c_int, u64p]
        L.cb_merkle_build.argtypes = [vp, u64p, C.c_size_t, C.c_size_t, u8p]
        L.cb_fold_interleaved16.argtypes = [vp, u64p, C.c_size_t, u64p, u64p]
        L.cb_commit_step.argtypes = [vp, u64p, C.c_int, C.c_int, C.c_int, u64p, u8p, u64p, u64p]
        _lib = L
    return _lib


def _p64(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def features() -> dict:
    f = lib().cb_features()
    return {"vpclmulqdq_4x128": bool(f & 1), "sha_ni": bool(f & 2)}


def restrict(wide: bool = True, sha_ni: bool = True) -> None:
    lib().cb_restrict((1 if wide else 0) | (2 if sha_ni else 0))


class Baseline:
    """A GenericPreExpanded domain over `basis` and a pool of `threads` threads."""

    def __init__(self, basis: np.ndarray, threads: int):
        b = np.ascontiguousarray(basis, dtype=np.uint64).reshape(-1, 2)
        self.dim = b.shape[0]
        self.ptr = lib().cb_create(_p64(b), self.dim, threads)
        if not self.ptr:
            raise RuntimeError("cb_create failed")
        self.threads = threads

    def close(self):
        if getattr(self, "ptr", None):
            lib().cb_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()

    def ntt_forward(self, data: np.ndarray, se: int, sl: int) -> np.ndarray:
        d = np.ascontiguousarray(data, dtype=np.uint64).reshape(-1, 2).copy()
        log_d = d.shape[0].bit_length() - 1
        if lib().cb_ntt_forward(self.ptr, _p64(d), log_d, se, sl):
            raise ValueError("ntt_forward arguments")
        return d

    def encode_batch(self, msg: np.ndarray, log_dim: int, lir: int, lbs: int) -> np.ndarray:
        m = np.ascontiguousarray(msg, dtype=np.uint64).reshape(-1, 2)
        out = np.zeros((1 << (log_dim + lir + lbs), 2), dtype=np.uint64)
        if lib().cb_encode_batch(self.ptr, _p64(m), log_dim, lir, lbs, _p64(out)):
            raise ValueError("encode_batch arguments")
        return out

    def merkle_build(self, elems: np.ndarray, batch: int) -> np.ndarray:
        e = np.ascontiguousarray(elems, dtype=np.uint64).reshape(-1, 2)
        n_leaves = e.shape[0] // batch
        tree = np.zeros((2 * n_leaves - 1, 32), dtype=np.uint8)
        rc = lib().cb_merkle_build(self.ptr, _p64(e), e.shape[0], batch, tree.ctypes.data_as(C.POINTER(C.c_uint8)))
        if rc:
            raise ValueError(f"merkle_build: status {rc}")
        return tree

    def fold_interleaved16(self, cw: np.ndarray, chal: np.ndarray) -> np.ndarray:
        c = np.ascontiguousarray(cw, dtype=np.uint64).reshape(-1, 2)
        ch = np.ascontiguousarray(chal, dtype=np.uint64).reshape(-1, 2)
        out = np.zeros((c.shape[0] // 16, 2), dtype=np.uint64)
        lib().cb_fold_interleaved16(self.ptr, _p64(c), out.shape[0], _p64(ch), _p64(out))
        return out

    def commit_step(self, msg, log_dim, lir, lbs, chal, cw=None, tree=None, folded=None):
        """encode + Merkle tree (16-element leaves) + first interleaved fold; the
        output buffers may be passed in to keep allocation out of a timed loop."""
        m = np.ascontiguousarray(msg, dtype=np.uint64).reshape(-1, 2)
        n = 1 << (log_dim + lir + lbs)
        cw = cw if cw is not None else np.zeros((n, 2), dtype=np.uint64)
        tree = tree if tree is not None else np.zeros((2 * (n // 16) - 1, 32), dtype=np.uint8)
        folded = folded if folded is not None else np.zeros((n // 16, 2), dtype=np.uint64)
        ch = np.ascontiguousarray(chal, dtype=np.uint64).reshape(-1, 2)
        rc = lib().cb_commit_step(self.ptr, _p64(m), log_dim, lir, lbs, _p64(cw),
                                  tree.ctypes.data_as(C.POINTER(C.c_uint8)), _p64(ch), _p64(folded))
        if rc:
            raise ValueError(f"commit_step: status {rc}")
        return cw, tree, folded
