"""The device library against the committed golden fixtures
(tests/golden/golden.json, FP64 oracle results): FP64 mode to the 1e-10
parity bar, the Sobol integer stage bit for bit."""
import json
import sys
from pathlib import Path

import numpy as np
import pytest

from thoth_amd.book import BuiltRequest

pytestmark = pytest.mark.gpu

GOLDEN_DIR = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN_DIR))
import cases  # noqa: E402

GOLDEN = json.loads((GOLDEN_DIR / "golden.json").read_text())
REL = 1e-10


@pytest.mark.parametrize("name", sorted(GOLDEN["books"]))
def test_device_fp64_matches_golden_books(gpu, name):
    spec = cases.cases()[name]
    b = BuiltRequest(spec)
    r = gpu.price(b)
    g = GOLDEN["books"][name]
    P = abs(g["premium"]) + 1e-12
    assert abs(r.premium - g["premium"]) <= REL * P, (r.premium, g["premium"])
    assert abs(r.premium_trust - g["premium_trust"]) <= max(REL, 1e-9) * g["premium_trust"] + 1e-15
    # Greeks are premium differences over the bump: premium-scaled bounds
    for k, h in (("delta", 1.0), ("gamma", 1.0), ("vega", 1.0), ("rho", 1e-2)):
        assert abs(getattr(r, k) - g[k]) <= REL * P / h * 10, k
    n = len(g["contract_premium"])
    for c, (x, y) in enumerate(zip(b.contract_values(r, "contract_premium", n), g["contract_premium"])):
        assert abs(x - y) <= REL * max(abs(y), 1e-12), (c, x, y)


def test_device_sobol_matches_golden(gpu):
    for e in GOLDEN["sobol"]:
        pts = gpu.sobol_points(e["dim"], e["n"], e["skip"]).astype(np.int64)
        assert np.array_equal(pts, np.array(e["points"], np.int64))
