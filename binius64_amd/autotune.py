"""Autotuner for the NTT pass grouping.

The forward NTT of a commit job runs as a few LDS-tiled passes; how many
layers a pass takes (ntt.max_k), the tile size (ntt.tile_log), which layers
multiply through window tables (ntt.variant, ntt.tab_min) and how many tables
stay resident (ntt.tab_group) trade HBM passes against table builds and
occupancy, and the best choice changes with the transform size.  This module
times every candidate on the device, per transform size, and stores the
winners as per-size overrides ("ntt.<knob>:<log_d>", see b64_config_set).

The table is saved as JSON next to the library (ntt_tuning.json) and applied
whenever the library is loaded, so the plain, batched and served paths all use
it.  Every candidate computes the same transform (the parity tests run each
knob combination against the oracle); only the schedule changes.

Usage on the GPU:  python -m binius64_amd.autotune --out binius64_amd/ntt_tuning.json
"""
from __future__ import annotations

import argparse
import json
import time
from pathlib import Path
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

KNOBS = ("ntt.variant", "ntt.max_k", "ntt.tile_log", "ntt.tab_min", "ntt.tab_group", "ntt.waves")
DEFAULT_PATH = Path(__file__).resolve().parent / "ntt_tuning.json"

Candidate = Tuple[int, int, int, int, int, int]


def default_candidates() -> List[Candidate]:
    c: List[Candidate] = [(0, k, 11, 1, 2, 0) for k in (5, 6, 7)]
    for k in (4, 5, 6, 7):
        for tile in (10, 11):
            for tab_min in (32, 64, 128):
                c.append((1, k, tile, tab_min, 2, 0))
            c.append((1, k, tile, 64, 1, 0))
            c.append((1, k, tile, 64, 2, 4))
        c.append((1, k, 11, 64, 4, 0))
    c += [(1, 3, 11, 32, 2, 0), (1, 3, 11, 64, 1, 0), (1, 3, 10, 32, 2, 4)]
    # variant 2: 4-layer forward passes above index bit 10 in registers
    c += [(2, 4, 11, 128, 2, 0), (2, 4, 11, 64, 2, 0), (2, 4, 10, 128, 2, 0), (2, 4, 11, 64, 2, 4)]
    # variants 3 and 4: 2 KiB-table register passes down to index bit 8, and the
    # remaining forward passes on LDS tiles through 2 KiB (tab_group 2) or 1 KiB
    # all-at-once (tab_group 1) tables
    for v in (3, 4):
        for tile in (9, 10, 11):
            for tab_min in (8, 16, 32):
                c.append((v, 4, tile, tab_min, 2, 0))
            c.append((v, 4, tile, 16, 1, 0))
    c += [(3, 5, 10, 16, 1, 0), (3, 3, 9, 8, 1, 0), (3, 5, 11, 16, 2, 0)]
    return c


def apply_table(table: Dict[int, Candidate], lib=None, prefix: str = "ntt") -> None:
    """Install per-size overrides (log_d -> candidate) in the loaded library:
    prefix "ntt" for single-job transforms, "batch" for batch plans."""
    from ._abi import check
    from ._abi import lib as _lib
    L = lib if lib is not None else _lib()
    check(L.b64_config_set(f"{prefix}.clear_sizes".encode(), 0), f"config_set {prefix}.clear_sizes")
    for log_d, cand in table.items():
        for key, v in zip(KNOBS, cand):
            key = prefix + key[3:]
            check(L.b64_config_set(f"{key}:{int(log_d)}".encode(), int(v)), f"config_set {key}:{log_d}")


def load_table(path: Path = DEFAULT_PATH, section: str = "sizes") -> Dict[int, Candidate]:
    """The per-size table of one section ("sizes": single-job, "batch_sizes":
    batch plans) of a tuning file."""
    data = json.loads(Path(path).read_text())
    knobs = [k.split(".", 1)[1] for k in data.get("knobs" if section == "sizes" else "batch_knobs", list(KNOBS))]
    out = {}
    for k, v in data.get(section, {}).items():
        cfg = dict(zip(knobs, v["config"]))
        out[int(k)] = tuple(int(cfg.get(knob.split(".", 1)[1], 0)) for knob in KNOBS)
    return out


class _EncodeBench:
    """A commit job's encode at message size 2^log_n on uniform random data."""

    def __init__(self, b64, log_n: int):
        import numpy as np
        self.b64, self.L = b64, b64.lib()
        self.params, self.dom = b64.setup_commit_params(log_n, 1, 96)
        words = np.random.default_rng(log_n).integers(0, 2 ** 63, size=2 << log_n, dtype=np.uint64)
        self.msg = b64.DeviceBuffer.from_host(words)
        self.cw = b64.DeviceBuffer(32 << log_n)

    def run(self):
        p = self.params
        self.b64.check(self.L.b64_rs_encode_batch(self.dom.ptr, p.log_dim, 1, self.msg.ptr, p.log_batch_size,
                                                  self.cw.ptr, None), "encode")

    def seconds(self, reps: int) -> float:
        self.run()
        self.b64.check(self.L.b64_stream_synchronize(None), "sync")
        t = time.perf_counter()
        for _ in range(reps):
            self.run()
        self.b64.check(self.L.b64_stream_synchronize(None), "sync")
        return (time.perf_counter() - t) / reps


def tune(log_ns: Iterable[int], candidates: Optional[Sequence[Candidate]] = None, budget_s: float = 0.02,
         verbose: bool = False) -> Dict[int, dict]:
    """Time every candidate for commit jobs of each message size log_n (the NTT
    has log_d = log_n + 1) and return {log_d: {"config": best, "ms": t, "all": {...}}}."""
    import binius64_amd as b64
    from ._abi import check, lib

    L = lib()
    cands = list(candidates or default_candidates())
    out: Dict[int, dict] = {}
    for log_n in log_ns:
        log_d = log_n + 1
        bench = _EncodeBench(b64, log_n)
        est = bench.seconds(1)
        reps = max(3, min(200, int(budget_s / max(est, 1e-6))))
        times = {}
        for cand in cands:
            for key, v in zip(KNOBS, cand):
                check(L.b64_config_set(f"{key}:{log_d}".encode(), int(v)), key)
            times[cand] = min(bench.seconds(reps) for _ in range(2))
        best = min(times, key=times.get)
        for key, v in zip(KNOBS, best):
            check(L.b64_config_set(f"{key}:{log_d}".encode(), int(v)), key)
        out[log_d] = {"config": list(best), "ms": round(1e3 * times[best], 5),
                      "all": {",".join(map(str, c)): round(1e3 * t, 5) for c, t in times.items()}}
        if verbose:
            print(json.dumps({"log_n": log_n, "log_d": log_d, "best": best, "ms": out[log_d]["ms"]}), flush=True)
    return out


def batch_candidates() -> List[Candidate]:
    """Batch-plan schedules: the multi-job table kernel (variants 0-1) or the
    multi-job forms of the register and small-table passes (variants 3-4)."""
    c: List[Candidate] = [(0, 7, 11, 1, 2, 0)]
    for k, tile, tab_min in ((4, 11, 128), (4, 10, 128), (5, 11, 64), (7, 11, 128), (4, 11, 64)):
        c.append((1, k, tile, tab_min, 2, 0))
    for v in (3, 4):
        for k, tile, tab_min, group in ((4, 9, 16, 1), (4, 10, 16, 1), (4, 11, 16, 2), (3, 9, 8, 1), (5, 10, 16, 1)):
            c.append((v, k, tile, tab_min, group, 0))
    return c


class _BatchBench:
    """A batch plan (hipGraph) of `n_jobs` commit jobs of message size 2^log_n."""

    def __init__(self, b64, log_n: int, n_jobs: int):
        import numpy as np
        self.b64 = b64
        rng = np.random.default_rng(log_n)
        self.jobs = [b64.CommitJob(log_n, rng.integers(0, 2 ** 63, size=(1 << log_n, 2), dtype=np.uint64),
                                   rng.integers(0, 2 ** 63, size=(log_n, 2), dtype=np.uint64))
                     for _ in range(n_jobs)]

    def seconds(self, reps: int) -> float:
        plan = self.b64.BatchPlan(self.jobs, use_graph=True)  # planned under the current schedule
        plan.run()
        t = time.perf_counter()
        for _ in range(reps):
            plan.run()
        return (time.perf_counter() - t) / reps


def tune_batch(log_ns: Iterable[int], n_jobs: int = 8, candidates: Optional[Sequence[Candidate]] = None,
               budget_s: float = 0.02, verbose: bool = False) -> Dict[int, dict]:
    """Time every batch schedule on a plan of n_jobs jobs of each message size
    log_n (NTT size log_d = log_n + 1); return {log_d: {"config", "ms", "all"}}."""
    import binius64_amd as b64
    from ._abi import check, lib

    L = lib()
    cands = list(candidates or batch_candidates())
    out: Dict[int, dict] = {}
    for log_n in log_ns:
        log_d = log_n + 1
        bench = _BatchBench(b64, log_n, n_jobs)
        est = bench.seconds(1)
        reps = max(3, min(100, int(budget_s / max(est, 1e-6))))
        times = {}
        for cand in cands:
            for key, v in zip(KNOBS, cand):
                check(L.b64_config_set(f"batch{key[3:]}:{log_d}".encode(), int(v)), key)
            times[cand] = min(bench.seconds(reps) for _ in range(2))
        best = min(times, key=times.get)
        for key, v in zip(KNOBS, best):
            check(L.b64_config_set(f"batch{key[3:]}:{log_d}".encode(), int(v)), key)
        out[log_d] = {"config": list(best), "ms": round(1e3 * times[best], 5), "jobs": n_jobs,
                      "all": {",".join(map(str, c)): round(1e3 * t, 5) for c, t in times.items()}}
        if verbose:
            print(json.dumps({"batch_log_n": log_n, "log_d": log_d, "best": best, "ms": out[log_d]["ms"]}), flush=True)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--log-n", type=int, nargs="+", default=list(range(8, 25)))
    ap.add_argument("--batch", action="store_true",
                    help="tune the batch-plan schedules (batch_sizes section) instead of the single-job ones")
    ap.add_argument("--out", default=str(DEFAULT_PATH))
    args = ap.parse_args(argv)
    from ._abi import lib
    if args.batch:
        lib().b64_config_set(b"batch.clear_sizes", 0)
        res = tune_batch(args.log_n, verbose=True)
        out = Path(args.out)
        doc = json.loads(out.read_text()) if out.exists() else {"device": "MI355X (gfx950)", "knobs": list(KNOBS)}
        doc["batch_knobs"] = ["batch" + k[3:] for k in KNOBS]
        doc["batch_what"] = ("per NTT size log_d = log_n + 1: best schedule for a batch plan of 8 jobs of that "
                             "size (hipGraph replay)")
        doc["batch_sizes"] = {str(k): v for k, v in sorted(res.items())}
        out.write_text(json.dumps(doc, indent=1) + "\n")
        print(json.dumps({"written": args.out, "batch_sizes": len(res)}), flush=True)
        return
    lib().b64_config_set(b"ntt.clear_sizes", 0)
    res = tune(args.log_n, verbose=True)
    doc = {"device": "MI355X (gfx950)", "knobs": list(KNOBS),
           "what": "per transform size log_d = log_n + 1 of a commit job's encode: best NTT schedule",
           "sizes": {str(k): v for k, v in sorted(res.items())}}
    out = Path(args.out)
    if out.exists():  # keep the batch-plan table of the file being replaced
        prev = json.loads(out.read_text())
        for key in ("batch_knobs", "batch_sizes"):
            if key in prev:
                doc[key] = prev[key]
    Path(args.out).write_text(json.dumps(doc, indent=1) + "\n")
    print(json.dumps({"written": args.out, "sizes": len(res)}), flush=True)


if __name__ == "__main__":
    main()
