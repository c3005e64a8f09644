"""Extracts the data vectors the reference's own sources hold for this path.

Run in the build container only (it reads the reference tree as text):
    python tests/golden/gen_reference_vectors.py /root/reference

Writes tests/golden/reference_vectors.json with
  * aes_to_ghash: the 256-entry AES-field -> GHASH embedding table
    (verifier/field/src/ghash.rs:520-783, LOOKUP_TABLE in From<AESTowerField8b>);
    the reference's proptest (ghash.rs:928-935) states phi(a*b) = phi(a)*phi(b).
  * ghash_mul_kats: the (a, b, a*b) triples of test_ghash_mul (ghash.rs:798-865).
  * mul_x_cases / mul_inv_x_cases: the inputs of test_mul_x / test_mul_inv_x
    (ghash.rs:873-925).
  * multiplicative_generator: BinaryField128bGhash::MULTIPLICATIVE_GENERATOR (ghash.rs:502).
  * fri_kats: calculate_n_test_queries / estimate_optimal_arity expectations
    (verifier/verifier/src/fri/common.rs:225-247).
Only numbers are copied; no source text is kept.
"""
import json
import re
import sys
from pathlib import Path


def main(ref_root: str) -> None:
    src = (Path(ref_root) / "verifier/field/src/ghash.rs").read_text()
    start = src.index("const LOOKUP_TABLE: [BinaryField128bGhash; 256]")
    end = src.index("];", start)
    table = re.findall(r"BinaryField128bGhash\((0x[0-9a-fA-F]+)\)", src[start:end])
    assert len(table) == 256, len(table)
    table = [int(x, 16) for x in table]

    gen = re.search(r"MULTIPLICATIVE_GENERATOR: Self = Self\((0x[0-9a-fA-F]+)\)", src)
    generator = int(gen.group(1), 16)

    # test_ghash_mul: a/b/c assignments in order.
    body = src[src.index("fn test_ghash_mul()"): src.index("fn test_multiplicative_generator")]

    def parse_val(expr: str) -> int:
        expr = expr.strip().replace("u128", "")
        return int(eval(expr, {"__builtins__": {}}, {}))  # literal arithmetic only

    kats = []
    a = b = None
    for line in body.splitlines():
        m = re.search(r"let a = BinaryField128bGhash\((.*)\);", line)
        if m:
            a = parse_val(m.group(1))
        m = re.search(r"let b = BinaryField128bGhash\((.*)\);", line)
        if m:
            b = parse_val(m.group(1))
        m = re.search(r"assert_eq!\(c, BinaryField128bGhash::from\((.*)\)\);", line)
        if m:
            kats.append([hex(a), hex(b), hex(parse_val(m.group(1)))])
    assert len(kats) >= 10, kats

    def cases(fn_name: str):
        seg = src[src.index(f"fn {fn_name}()"):]
        seg = seg[: seg.index("];")]
        vals = re.findall(r"^\s*(0x[0-9a-fA-F]+)(?:u128)?,", seg, re.M)
        return [hex(int(v, 16)) for v in vals]

    out = {
        "aes_to_ghash": [hex(v) for v in table],
        "ghash_mul_kats": kats,
        "mul_x_cases": cases("test_mul_x"),
        "mul_inv_x_cases": cases("test_mul_inv_x"),
        "multiplicative_generator": hex(generator),
        "fri_kats": {
            "n_test_queries": [[96, 1, 232], [96, 2, 142]],
            "optimal_arity_digest256": {"log_block_length": [22, 35], "arity": 4},
            "optimal_arity_digest1024": {"log_block_length": [22, 28], "arity": 6},
            "field_size": 128,
        },
    }
    dst = Path(__file__).with_name("reference_vectors.json")
    dst.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {dst}: {len(table)} table entries, {len(kats)} mul KATs")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
