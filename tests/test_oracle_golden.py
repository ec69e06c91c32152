"""The oracle against the reference's frozen primitive-ERI vectors.

tests/golden/eri_recursive_m8.jsonl is the reference's own fixture
(tests/data/eri_recursive_m8.jsonl: 96 quartets, angular totals 0..4 per
centre, zero and finite field, normalized and unnormalized primitives),
checked at the reference's tolerance (tests/python/test_eri_optimization.py:15-34).
"""

import json
from pathlib import Path

import numpy as np
import pytest

RECORDS = [json.loads(line) for line in (Path(__file__).parent / "golden" / "eri_recursive_m8.jsonl").read_text().splitlines()]


def test_fixture_size():
    assert len(RECORDS) == 96


@pytest.mark.parametrize("index", range(len(RECORDS)))
def test_oracle_primitive_eri_against_frozen_values(port, index):
    r = RECORDS[index]
    for field, expected in zip([(0, 0, 0), r["field"]], r["expected"]):
        f = port.field6(field, r["gauge"])
        actual = port.primitive_eri(r["primitives"], f)
        np.testing.assert_allclose(actual, complex(*expected), atol=5.0e-14, rtol=5.0e-13)


def test_planted_mistake_is_caught(port):
    """A sign error in one Hermite index must fail the frozen-vector check."""
    r = next(x for x in RECORDS if sum(sum(p["angular"]) for p in x["primitives"]) >= 3)
    prims = [dict(p) for p in r["primitives"]]
    a = list(prims[0]["angular"])
    a[0], a[1] = a[1], a[0]  # transposed operand: x <-> y power swap
    prims[0]["angular"] = a
    f = port.field6(r["field"], r["gauge"])
    if a[0] != a[1]:
        assert abs(port.primitive_eri(prims, f) - complex(*r["expected"][1])) > 1e-9
