"""Device FP64 against the oracle on 32 seeded random books (tests/random_books.py):
mixed models, correlation, every barrier type, discrete and continuous
monitoring, rainbows, Sobol skips, rate curves, dividends, the leverage
surface and the per-date layout, and bump Greeks on half of them.  The bar is
the engine's FP64 parity bar: 1e-10 relative on every premium."""
import pytest

from random_books import random_book
from thoth_amd.book import BuiltRequest

pytestmark = pytest.mark.gpu

REL = 1e-10


def close(a, b, tol=REL, floor=1e-12):
    return abs(a - b) <= tol * max(abs(b), floor)


@pytest.mark.parametrize("seed", range(32))
def test_random_book_fp64_matches_oracle(gpu, oracle, seed):
    spec = random_book(seed)
    print("case: random book seed %d, %d equities, %d contracts, greeks %d"
          % (seed, len(spec.equities), len(spec.contracts), spec.greeks), flush=True)
    b = BuiltRequest(spec)
    g, o = gpu.price(b), oracle.price(b)
    assert close(g.premium, o.premium), (g.premium, o.premium)
    assert close(g.premium_trust, o.premium_trust, 1e-9), (g.premium_trust, o.premium_trust)
    for c in range(len(spec.contracts)):
        assert close(g.contract_premium[c], o.contract_premium[c]), (c, g.contract_premium[c], o.contract_premium[c])
    if spec.greeks:
        # finite differences of premiums: the premium error over the bump
        P = abs(o.premium) + 1e-12
        assert abs(g.delta - o.delta) <= REL * P * 10, (g.delta, o.delta)
        assert abs(g.rho - o.rho) <= REL * P / 1e-2 * 10, (g.rho, o.rho)
