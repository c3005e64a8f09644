"""Writes oracle_fixtures.json: small input/output vectors of the CPU oracle.

They pin the oracle against later changes (tests/test_oracle_fixtures.py) and
give the GPU tests a fixed target besides the live oracle
(tests/test_gpu_fixtures.py).  Inputs come from the oracle's SplitMix64
generator (random_elems / random_words), so only seeds are stored; large outputs
are stored as the SHA-256 of their bytes plus the first elements.

The set follows the fixture list for this path: B128 products, NTT outputs at
log_d = 6 and 10 across skip windows on a standard and a subspace-image domain,
encode_batch codewords for (log_dim, lir, lbs) = (4,2,0), (6,2,1), (8,3,2),
(16,1,0), a SHA-256 leaf, a Sha256Compression node, the Merkle root of 2^10
elements with batch 16, and full FRI-only transcripts for 2^10 and 2^12
message elements.

Run from the repository root:  python tests/golden/gen_oracle_fixtures.py
"""
import hashlib
import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import oracle as orc  # noqa: E402

OUT = Path(__file__).resolve().parent / "oracle_fixtures.json"


def h(arr) -> str:
    return hashlib.sha256(arr.tobytes()).hexdigest()


def hexes(arr, n=None):
    ints = orc.to_ints(arr[:n] if n else arr)
    return [f"{v:032x}" for v in ints]


NTT_WINDOWS = {6: [(0, 0), (1, 0), (0, 2), (2, 1)], 10: [(0, 0), (1, 0), (1, 4), (3, 3)]}
ENCODE = [(4, 2, 0), (6, 2, 1), (8, 3, 2), (16, 1, 0)]
FRI_N = [10, 12]


def domains():
    """name -> (Domain factory for log_d, description)"""
    return {
        "standard": lambda d: orc.Domain(dim=d),
        # the image subspace a commit of 2^10 elements uses: BinarySubspace::with_dim(11).reduce_dim(d)
        "subspace11": lambda d: orc.Domain(orc.Domain(dim=11).subspace(d)),
    }


def generate() -> dict:
    fx = {"generator": "oracle (C restatement), SplitMix64 inputs"}
    a, b = orc.random_elems(11, 64), orc.random_elems(12, 64)
    fx["mul"] = {"seed_a": 11, "seed_b": 12, "n": 64, "ab": hexes(orc.gf_mul(a, b))}

    ntt = []
    for name, mk in domains().items():
        for log_d, wins in NTT_WINDOWS.items():
            dom = mk(log_d)
            data = orc.random_elems(300 + log_d, 1 << log_d)
            for se, sl in wins:
                out = orc.ntt_forward(dom, data, se, sl)
                ntt.append({"domain": name, "log_d": log_d, "skip_early": se, "skip_late": sl,
                            "seed": 300 + log_d, "sha256": h(out), "head": hexes(out, 2)})
    fx["ntt"] = ntt

    enc = []
    for log_dim, lir, lbs in ENCODE:
        dom = orc.Domain(dim=log_dim + lir)
        msg = orc.random_elems(400 + log_dim, 1 << (log_dim + lbs))
        cw = orc.encode_batch(dom, msg, log_dim, lir, lbs, threads=8)
        enc.append({"log_dim": log_dim, "lir": lir, "lbs": lbs, "seed": 400 + log_dim, "sha256": h(cw),
                    "head": hexes(cw, 2)})
    fx["encode"] = enc

    leaf = orc.random_elems(500, 16)
    two = orc.random_elems(501, 32)
    tree2 = orc.merkle_build(two, 16)
    fx["sha256_leaf"] = {"seed": 500, "n_elems": 16, "digest": orc.merkle_build(leaf, 16)[-1].tobytes().hex()}
    fx["compress"] = {"seed": 501, "n_elems": 32, "batch": 16, "left": tree2[0].tobytes().hex(),
                      "right": tree2[1].tobytes().hex(), "node": tree2[2].tobytes().hex()}
    e10 = orc.random_elems(502, 1 << 10)
    fx["merkle_2p10_batch16"] = {"seed": 502, "root": orc.merkle_build(e10, 16)[-1].tobytes().hex()}

    fri = []
    for n in FRI_N:
        p = orc.fri_params_setup(n, 1, 96)
        dom = orc.Domain(orc.Domain(dim=n + 1).subspace(p["rs_log_len"]))
        msg = orc.random_elems(600 + n, 1 << n)
        res = orc.fri_prove(dom, msg, p["log_dim"], 1, p["log_batch_size"], p["fold_arities"],
                            p["n_test_queries"], threads=8)
        fri.append({"n": n, "seed": 600 + n, "log_dim": p["log_dim"], "log_batch_size": p["log_batch_size"],
                    "fold_arities": p["fold_arities"], "n_test_queries": p["n_test_queries"],
                    "tape_len": len(res["tape"]), "tape_sha256": hashlib.sha256(res["tape"]).hexdigest(),
                    "tape_head": res["tape"][:96].hex(), "roots": [r.hex() for r in res["roots"]]})
    fx["fri_transcript"] = fri
    return fx


if __name__ == "__main__":
    OUT.write_text(json.dumps(generate(), indent=1) + "\n")
    print("wrote", OUT)
