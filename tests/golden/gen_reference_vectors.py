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
  * vision4 / vision6: the Vision-M hash data of verifier/verifier/src/hash/vision_{4,6}/:
    B_FWD_COEFFS, B_INV_COEFFS and ROUND_CONSTANTS (constants.rs), the round-key
    schedule inputs INITIAL_CONSTANT, CONSTANTS_MATRIX, CONSTANTS_CONSTANT of its
    test module, and the permutation test vectors of test_permutation
    (permutation.rs; the case whose input comes from a seeded RNG is skipped).
Only numbers are copied; no source text is kept.
"""
import json
import re
import sys
from pathlib import Path


GHASH_RE = r"Ghash::new\((0x[0-9a-fA-F]+)\)"


def _static_values(src: str, name: str):
    start = src.index(name)
    start = src.index("= [", start)
    depth, i = 0, start + 2
    while True:  # the matching closing bracket of the initializer
        if src[i] == "[":
            depth += 1
        elif src[i] == "]":
            depth -= 1
            if depth == 0:
                break
        i += 1
    return [hex(int(v, 16)) for v in re.findall(GHASH_RE, src[start:i])]


def vision_data(root: Path, m: int) -> dict:
    base = root / f"verifier/verifier/src/hash/vision_{m}"
    const = (base / "constants.rs").read_text()
    perm = (base / "permutation.rs").read_text()
    data = {
        "m": m,
        "b_fwd_coeffs": _static_values(const, "static B_FWD_COEFFS"),
        "b_inv_coeffs": _static_values(const, "static B_INV_COEFFS"),
        "round_constants": _static_values(const, "ROUND_CONSTANTS: [[Ghash"),
        "initial_constant": _static_values(const, "static INITIAL_CONSTANT"),
        "constants_matrix": _static_values(const, "static CONSTANTS_MATRIX"),
        "constants_constant": _static_values(const, "static CONSTANTS_CONSTANT"),
    }
    assert len(data["b_fwd_coeffs"]) == 4 and len(data["b_inv_coeffs"]) == 129
    assert len(data["round_constants"]) == 17 * m and len(data["constants_matrix"]) == m * m
    # test_permutation cases: (input, expected) pairs in source order
    body = perm[perm.index("fn test_permutation()"):]
    body = body[: body.index("for (input, expected) in cases")]
    kats = []
    for case in re.split(r"\n\t\t\t\(\n", body)[1:]:
        vals = re.findall(GHASH_RE, case)
        if "Ghash::random" in case:
            continue
        if "array::from_fn(|_| Ghash::new(" in case:
            inp = [vals[0]] * m
            exp = vals[1:]
        else:
            inp, exp = vals[:m], vals[m:]
        assert len(inp) == m and len(exp) == m, (m, case)
        kats.append([[hex(int(v, 16)) for v in inp], [hex(int(v, 16)) for v in exp]])
    assert len(kats) == 2, kats
    data["permutation_kats"] = kats
    return data


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

    out_vision = {f"vision{m}": vision_data(Path(ref_root), m) for m in (4, 6)}

    out = {
        **out_vision,
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
