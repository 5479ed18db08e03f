"""The committed golden fixtures (tests/golden/golden.json, written by
tests/golden/make_golden.py from the FP64 oracle) against the oracle as it
builds today: a regression pin of the restatement, on CPU."""
import json
import sys
from pathlib import Path

import numpy as np
import pytest

from thoth_amd.book import BuiltRequest

GOLDEN_DIR = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN_DIR))
import cases  # noqa: E402

GOLDEN = json.loads((GOLDEN_DIR / "golden.json").read_text())


def close(a, b, rel=1e-13):
    return abs(a - b) <= rel * max(abs(b), 1e-300)


@pytest.mark.parametrize("name", sorted(GOLDEN["books"]))
def test_oracle_reproduces_golden_books(oracle, name):
    spec = cases.cases()[name]
    b = BuiltRequest(spec)
    r = oracle.price(b)
    g = GOLDEN["books"][name]
    for k in ("premium", "premium_trust", "delta", "gamma", "vega", "rho"):
        assert close(getattr(r, k), g[k]), (k, getattr(r, k), g[k])
    n = len(g["contract_premium"])
    for k in ("contract_premium", "contract_trust", "contract_delta", "contract_vega"):
        for c, (x, y) in enumerate(zip(b.contract_values(r, k, n), g[k])):
            assert close(x, y), (k, c, x, y)
    assert r.n_dates == g["n_dates"]


def test_oracle_reproduces_golden_sobol_and_normals(oracle):
    for e in GOLDEN["sobol"]:
        pts = oracle.sobol_int(e["dim"], e["n"], e["skip"]).astype(np.int64)
        assert np.array_equal(pts, np.array(e["points"], np.int64))
    z = oracle.normal_cdf_inv([e["p"] for e in GOLDEN["normal"]])
    assert np.allclose(z, [e["z"] for e in GOLDEN["normal"]], rtol=1e-15, atol=0)
