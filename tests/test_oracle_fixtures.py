"""The oracle still produces the committed fixtures (tests/golden/oracle_fixtures.json)."""
import json
import sys
from pathlib import Path

GOLDEN = Path(__file__).parent / "golden"


def test_oracle_reproduces_fixtures():
    sys.path.insert(0, str(GOLDEN))
    try:
        import gen_oracle_fixtures as gen
    finally:
        sys.path.pop(0)
    want = json.loads((GOLDEN / "oracle_fixtures.json").read_text())
    got = gen.generate()
    for key in want:
        assert got[key] == want[key], key
