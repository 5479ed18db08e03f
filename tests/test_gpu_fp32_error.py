"""FP32 production error against the FP64 oracle-matched device mode, on the
same Sobol points, for every premium and Greek of the three pricing
workloads (Black-Scholes call, Heston down-and-out call, 5-asset LSV worst-of
basket).  The table is printed; the bounds are the stated FP32 tolerances."""
import pytest

from thoth_amd import abi, configs
from thoth_amd.book import BuiltRequest

pytestmark = pytest.mark.gpu

GREEKS = abi.GREEK_DELTA | abi.GREEK_GAMMA | abi.GREEK_VEGA | abi.GREEK_RHO

CASES = [
    ("black_scholes_call", lambda p: configs.black_scholes_call(paths=1 << 18, precision=p, greeks=GREEKS), 1e-5),
    ("heston_down_out_call", lambda p: configs.heston_down_out_call(paths=1 << 18, precision=p, greeks=GREEKS), 1e-4),
    ("lsv_worst_of_basket", lambda p: configs.lsv_worst_of_basket(paths=1 << 16, precision=p, greeks=GREEKS), 1e-4),
]


@pytest.mark.parametrize("name,make,tol", CASES, ids=[c[0] for c in CASES])
def test_fp32_error_table(gpu, name, make, tol):
    s64, s32 = make(abi.FP64), make(abi.FP32)
    r64 = gpu.price(BuiltRequest(s64))
    r32 = gpu.price(BuiltRequest(s32))
    rows = []
    for key in ("premium", "delta", "gamma", "vega", "rho"):
        a, b = getattr(r32, key), getattr(r64, key)
        rows.append((key, b, a, abs(a - b), abs(a - b) / max(abs(b), 1e-300)))
    print("\n%s (FP64 trust %.3e)" % (name, r64.premium_trust))
    for key, b, a, ae, re in rows:
        print("  %-8s fp64 %+.10f fp32 %+.10f abs %.2e rel %.2e" % (key, b, a, ae, re))
    # premium: relative tolerance; Greeks: finite differences of premiums, so the
    # premium error divided by the bump size (delta, gamma h = 1 % / 10 % of spot;
    # vega one vol point; rho 1bp reported per 1 %) bounds them
    P = abs(r64.premium)
    assert rows[0][4] <= tol
    assert rows[1][3] <= 4 * tol * P / 1.0 + 1e-6       # delta: (P(1.01 S) - P) / (0.01 S) with S ~ 100
    assert rows[2][3] <= 4 * tol * P / 100.0 + 1e-6     # gamma: central, h^2 = 100
    assert rows[3][3] <= 4 * tol * P + 1e-6             # vega per point
    assert rows[4][3] <= 4 * tol * P * 100 + 1e-6       # rho: 1bp bump reported per 1 %
