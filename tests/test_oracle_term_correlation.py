"""Term-structured correlation and multi-curve discounting in the oracle
(reference: vol_correl/correlation.cpp:35-130, 528-620, 860-903;
marketdata/currency.cpp:19-27, 129-142; tasks/pricer.cpp:202-217, 344-359),
pinned against what the definitions fix:

* the window average of a correlation entry equals the quadrature of the
  piecewise-linear entry (numpy's interp is exactly "linear between pillars,
  flat outside"); each date's Cholesky factor reproduces the matrix averaged
  over the step into that date (date 0: today's matrix);
* a structure whose pillars all hold one matrix prices as that constant matrix;
* the best-of of two Black-Scholes assets under a term structure reprices the
  exchange-option closed form (Margrabe) at the integrated correlation, as the
  reference's own test does for a basket (test_term_correlation.cpp:164-185);
* with a distinct OIS curve a European flow is discounted on it: price =
  exp(-(r_ois - r) T) x the single-curve price, path by path, and the rho bump
  moves both curves (test_multicurve.cpp:54-125).
"""
import math

import numpy as np
import pytest
from scipy import integrate
from scipy.stats import norm

from thoth_amd import abi
from thoth_amd.book import BookSpec, BuiltRequest, Equity, McConfig, RainbowUdl, Vanilla

MATS = [0.25, 1.0, 1.5]
PILLARS = [
    [[1.0, 0.8, 0.3], [0.8, 1.0, 0.5], [0.3, 0.5, 1.0]],
    [[1.0, 0.1, -0.2], [0.1, 1.0, 0.0], [-0.2, 0.0, 1.0]],
    [[1.0, -0.4, 0.4], [-0.4, 1.0, 0.2], [0.4, 0.2, 1.0]],
]


def three_bs(mats=MATS, pillars=PILLARS, paths=4096, days=500, **kw):
    return BuiltRequest(BookSpec(mcl=McConfig(paths=paths, max_day_step=7),
                                 equities=[Equity(spot=100, volatility=25.0), Equity(spot=80, volatility=30.0),
                                           Equity(spot=120, volatility=20.0)],
                                 contracts=[Vanilla(-1, days, 100.0)], rate_values=[2.0],
                                 rainbows=[RainbowUdl([0, 1, 2], best=False)],
                                 correlation_maturities=mats, correlation_pillars=pillars, **kw))


def interp_avg(mats, pillars, i, j, a, b):
    f = lambda t: np.interp(t, mats, [p[i][j] for p in pillars])
    if b - a < 1e-12:
        return float(f(a))
    return integrate.quad(f, a, b, points=[m for m in mats if a < m < b], epsabs=0, epsrel=1e-13)[0] / (b - a)


@pytest.mark.parametrize("a,b", [(0.0, 0.1), (0.1, 0.7), (0.3, 2.5), (1.2, 1.4), (1.6, 4.0), (0.6, 0.6),
                                 (0.0, 0.0), (5.0, 5.0), (0.25, 1.5)])
def test_window_average_is_the_quadrature(oracle, a, b):
    got = oracle.avg_correlation(three_bs(), a, b)
    for i in range(3):
        for j in range(3):
            want = 1.0 if i == j else interp_avg(MATS, PILLARS, i, j, a, b)
            assert abs(got[i, j] - want) <= 1e-13, (i, j, got[i, j], want)


def test_step_factors_reproduce_the_step_averages(oracle):
    b = three_bs()
    days, t, _ = oracle.grid(b)
    L = oracle.chol_schedule(b)
    assert L.shape == (len(t), 3, 3)
    for d in range(len(t)):
        lo, hi = (t[0], t[0]) if d == 0 else (t[d - 1], t[d])
        want = np.array([[1.0 if i == j else interp_avg(MATS, PILLARS, i, j, lo, hi) for j in range(3)]
                         for i in range(3)])
        assert np.allclose(L[d] @ L[d].T, want, rtol=0, atol=1e-13), d
        assert np.array_equal(L[d], np.tril(L[d])) and (np.diag(L[d]) > 0).all()
    # a constant matrix has no schedule
    assert oracle.chol_schedule(three_bs(mats=[], pillars=[], correlation=PILLARS[0])).shape[0] == 0


def test_identical_pillars_price_as_the_constant_matrix(oracle):
    term = oracle.price(three_bs(mats=[0.5, 2.0], pillars=[PILLARS[0], PILLARS[0]]))
    flat = oracle.price(three_bs(mats=[], pillars=[], correlation=PILLARS[0]))
    assert abs(term.premium - flat.premium) <= 1e-13 * flat.premium
    one = oracle.price(three_bs(mats=[0.7], pillars=[PILLARS[0]]))  # one pillar: a constant matrix
    assert one.premium == flat.premium


def margrabe_best_of_two(S1, S2, v1, v2, rho, T):
    """E[max(100 S1(T)/S1, 100 S2(T)/S2)] discounted, r = q = 0: 100 + exchange option."""
    s = math.sqrt(v1 * v1 + v2 * v2 - 2 * rho * v1 * v2) * math.sqrt(T)
    d1 = 0.5 * s
    return 100.0 + 100.0 * (norm.cdf(d1) - norm.cdf(d1 - s))


def test_best_of_two_reprices_margrabe_at_the_integrated_correlation(oracle):
    mats, c = [0.2, 0.8, 1.2], [0.9, -0.3, 0.2]
    pillars = [[[1.0, x], [x, 1.0]] for x in c]
    T = 1.0
    rho_bar = interp_avg(mats, pillars, 0, 1, 0.0, T)
    base = dict(mcl=McConfig(paths=1 << 15, max_day_step=5),
                equities=[Equity(spot=100, volatility=30.0), Equity(spot=50, volatility=20.0)],
                contracts=[Vanilla(-1, 365, 0.0)], rate_values=[0.0],
                rainbows=[RainbowUdl([0, 1], best=True)])
    term = oracle.price(BuiltRequest(BookSpec(correlation_maturities=mats, correlation_pillars=pillars, **base)))
    flat = oracle.price(BuiltRequest(BookSpec(correlation=[[1, rho_bar], [rho_bar, 1]], **base)))
    ref = margrabe_best_of_two(100, 50, 0.3, 0.2, rho_bar, T)
    print("rho_bar %.6f: term %.5f +- %.5f, flat %.5f, Margrabe %.5f" % (rho_bar, term.premium, term.premium_trust,
                                                                         flat.premium, ref))
    assert abs(term.premium - ref) <= 6 * term.premium_trust + 0.01
    assert abs(term.premium - flat.premium) <= 6 * term.premium_trust
    # and not the price at today's correlation, nor at the middle pillar's
    for x in (c[0], c[1]):
        assert abs(ref - margrabe_best_of_two(100, 50, 0.3, 0.2, x, T)) > 10 * term.premium_trust


def _call(discount=None, greeks=0):
    kw = dict(discount_days=[0, 3650], discount_values=discount) if discount else {}
    return BuiltRequest(BookSpec(mcl=McConfig(paths=1 << 14, max_day_step=30),
                                 equities=[Equity(spot=100, volatility=30.0)], contracts=[Vanilla(0, 365, 100.0)],
                                 rate_days=[0, 3650], rate_values=[8.0, 8.0], greeks=greeks, **kw))


def test_multi_curve_discounts_on_the_ois_curve(oracle):
    single = oracle.price(_call(greeks=abi.GREEK_RHO))
    multi = oracle.price(_call(discount=[5.0, 5.0], greeks=abi.GREEK_RHO))
    f = math.exp(-(0.05 - 0.08) * 1.0)
    assert abs(multi.premium - f * single.premium) <= 1e-13 * multi.premium
    assert abs(multi.premium_trust - f * single.premium_trust) <= 1e-12 * multi.premium_trust
    # rho moves both curves: the ratio of the discount factors is unchanged
    assert abs(multi.rho - f * single.rho) <= 1e-8 * abs(multi.rho)
    # the reference's closed form (test_multicurve.cpp:54-67): BS forward on the
    # projection curve, discount on the OIS curve
    d1 = (math.log(100 / 100) + (0.08 + 0.045) * 1.0) / 0.3
    ref = math.exp(-0.05) * (100 * math.exp(0.08) * norm.cdf(d1) - 100 * norm.cdf(d1 - 0.3))
    assert abs(multi.premium - ref) <= 6 * multi.premium_trust + 1e-2
    # no discount curve: the historic single curve, and a discount curve equal to
    # the rate curve prices the same
    same = oracle.price(_call(discount=[8.0, 8.0]))
    assert same.premium == single.premium


def test_term_structure_refusals(oracle):
    with pytest.raises(RuntimeError):
        oracle.price(three_bs(mats=[1.0, 0.5, 2.0]))  # not increasing
    bad = [[[1.0, 0.99, -0.99], [0.99, 1.0, 0.99], [-0.99, 0.99, 1.0]]] * 3
    with pytest.raises(RuntimeError):
        oracle.price(three_bs(pillars=bad))          # steps not positive-definite
