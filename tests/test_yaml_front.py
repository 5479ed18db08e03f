"""YAML front-end (thoth_amd/yaml_front.py) on the reference's own sample
configurations (pricer/samples, copied as fixtures under
tests/golden/reference_samples): which `!mcl_pricer` tasks fall in the device
engine's subset, the request each translates to, and -- through the CPU oracle
-- the Black-Scholes premium of the 1y ATM call sample."""
import math
from pathlib import Path

import pytest
from scipy.stats import norm

from thoth_amd import abi
from thoth_amd import yaml_front as yf
from thoth_amd.book import Barrier, BuiltRequest, Vanilla

SAMPLES = Path(__file__).resolve().parent / "golden" / "reference_samples"


def _objs(name):
    return yf.load((SAMPLES / (name + ".yaml")).read_text())


SUPPORTED = {
    "vanilla_mono_eu_call_mcl", "barrier_mono_continuous_call_mcl", "barrier_mono_discrete_call_mcl",
    "vanilla_multi_eu_call_best_of_mcl", "vanilla_multi_eu_call_worst_of_mcl", "vanilla_mono_eu_call_heston_mcl",
    "barrier_mono_continuous_in_call_mcl", "barrier_mono_continuous_down_out_call_mcl",
    "barrier_mono_continuous_down_in_call_mcl", "barrier_mono_continuous_up_out_put_mcl",
    "barrier_mono_discrete_in_call_mcl", "vanilla_mono_eu_put_mcl", "vanilla_mono_eu_call_mcl_gpu",
    "vanilla_mono_eu_put_mcl_gpu", "vanilla_mono_eu_call_lsv_mcl", "vanilla_mono_eu_call_ois_mcl",
    "vanilla_mono_eu_call_mcl_pseudo", "vanilla_mono_am_put_mcl", "vanilla_mono_am_put_heston_mcl",
    "vanilla_mono_am_call_mcl",
}

REFUSED = {  # pricer -> a word of the reason
    "vanilla_mono_eu_call_div_mcl": "discrete_dividends",
    "vanilla_mono_eu_call_quanto_mcl": "currency", "vanilla_mono_eu_call_compo_mcl": "composite",
    "vanilla_multi_eu_call_basket_mcl": "basket", "vanilla_mono_eu_call_bates_mcl": "Bates",
    "variance_mono_mcl": "variance", "vanilla_mono_eu_call_sabr_mcl": "sabr_volatility",
    "basket_term_correl_mcl": "basket",
    "vanilla_mono_eu_call_hw_mcl": "rate_model",
    "autocallable_athena_mcl": "autocallable", "asian_mono_eu_call_mcl": "asian",
    "ratchet_mono_mcl": "ratchet", "digital_mono_cash_call_mcl": "digital",
}


def test_matrix_sample_subset():
    objs = _objs("matrix")
    names = yf.mcl_pricers(objs)
    assert len(names) == 44
    ok = set()
    for n in names:
        try:
            yf.translate(objs, n)
            ok.add(n)
        except yf.Unsupported as e:
            if n in REFUSED:
                assert REFUSED[n] in str(e), (n, str(e))
    assert ok == SUPPORTED
    assert set(REFUSED) <= set(names) - ok


def test_simple_call_translation():
    t = yf.translate(_objs("simple_call"), "simple_call_pricing")
    s = t.spec
    assert t.result == "simple_call_result" and t.contracts == ["simple_call"]
    assert t.greeks == ["delta", "gamma", "vega", "rho", "theta"] and t.not_computed == []
    assert s.greeks == abi.GREEK_DELTA | abi.GREEK_GAMMA | abi.GREEK_VEGA | abi.GREEK_RHO
    assert (s.mcl.paths, s.mcl.max_day_step, s.mcl.min_day_step, s.mcl.use_sobol) == (50000, 1, -1, True)
    (e,) = s.equities
    assert (e.spot, e.vol_kind, e.volatility, e.dividend_yield) == (100.0, abi.VOL_BS, 30.0, 0.0)
    (c,) = s.contracts
    assert isinstance(c, Vanilla) and (c.underlying, c.maturity_days, c.strike, c.is_call) == (0, 365, 100.0, True)
    assert list(s.rate_days) == [0, 3653] and list(s.rate_values) == [8.0, 8.0]


def test_matrix_translations():
    objs = _objs("matrix")
    h = yf.translate(objs, "vanilla_mono_eu_call_heston_mcl").spec.equities[0]
    assert (h.vol_kind, h.init_vol, h.long_vol, h.kappa, h.vol_of_vol, h.rho_spot_var) == \
        (abi.VOL_HESTON, 25.0, 25.0, 2.0, 0.5, -0.7)
    w = yf.translate(objs, "vanilla_multi_eu_call_worst_of_mcl").spec
    assert len(w.equities) == 2 and w.rainbows[0].members == [0, 1] and not w.rainbows[0].best
    assert w.correlation[0][1] == 0.5 and w.contracts[0].underlying == -1
    d = yf.translate(objs, "barrier_mono_discrete_call_mcl").spec.contracts[0]
    assert isinstance(d, Barrier) and (d.barrier_type, d.level, d.discrete, d.monitoring_period_days) == \
        (abi.UP_OUT, 130.0, True, 30)
    c = yf.translate(objs, "barrier_mono_continuous_down_out_call_mcl").spec.contracts[0]
    assert (c.barrier_type, c.level, c.discrete) == (abi.DOWN_OUT, 70.0, False)
    lsv = yf.translate(_objs("lsv"), "lsv_mcl").spec.equities[0]
    assert lsv.vol_kind == abi.VOL_LSV and lsv.rho_spot_var == -0.6 and list(lsv.surface.alpha) == [0.22, 0.24]


def test_simple_call_prices_black_scholes(oracle):
    # the sample's own comment: Black-Scholes reference premium ~= 15.71
    t = yf.translate(_objs("simple_call"), "simple_call_pricing", paths=1 << 13)
    t.spec.greeks = 0
    r = oracle.price(BuiltRequest(t.spec))
    S, K, r_, v, T = 100.0, 100.0, 0.08, 0.30, 1.0
    d1 = (math.log(S / K) + (r_ + v * v / 2) * T) / (v * math.sqrt(T))
    bs = S * norm.cdf(d1) - K * math.exp(-r_ * T) * norm.cdf(d1 - v * math.sqrt(T))
    assert abs(bs - 15.71) < 0.01
    assert abs(r.premium - bs) < 3 * r.premium_trust, (r.premium, bs, r.premium_trust)


def test_unsupported_is_a_value_error():
    with pytest.raises(ValueError):
        yf.translate(_objs("matrix"), "vanilla_mono_eu_call_bates_mcl")


def test_american_exercise_translates():
    c = yf.translate(_objs("matrix"), "vanilla_mono_am_put_mcl").spec.contracts[0]
    assert isinstance(c, Vanilla) and c.american and not c.is_call


def test_theta_is_the_one_day_roll(oracle):
    """Theta as the reference computes it (tasks/pricer.cpp:371-384): the
    premium with the valuation date one day later (spot held) minus the
    premium.  For the Black-Scholes call sample the payoff reads the bridge
    endpoint alone (the first Sobol dimension in both runs), so the MC
    difference is the analytic one-day decay up to the integration error."""
    objs = _objs("simple_call")
    t0 = yf.translate(objs, "simple_call_pricing", paths=1 << 13)
    t1 = yf.translate(objs, "simple_call_pricing", paths=1 << 13, roll_days=1)
    assert t1.spec.contracts[0].maturity_days == 364 and list(t1.spec.rate_days) == [-1, 3652]
    t0.spec.greeks = t1.spec.greeks = 0
    r0, r1 = oracle.price(BuiltRequest(t0.spec)), oracle.price(BuiltRequest(t1.spec))
    theta = yf.result_block(t0, r0, r1)["theta"]

    def bs(T):
        d1 = (math.log(100 / 100) + (0.08 + 0.045) * T) / (0.3 * math.sqrt(T))
        return 100 * norm.cdf(d1) - 100 * math.exp(-0.08 * T) * norm.cdf(d1 - 0.3 * math.sqrt(T))
    ref = bs(364 / 365) - bs(1.0)
    assert ref < 0 and abs(theta - ref) < 2e-3, (theta, ref)


def test_multicurve_sample_discounts_on_its_ois_curve(oracle):
    # pricer/samples/multicurve.yaml: projection curve 8 %, OIS curve 5 %; the
    # MCL task prices the 1y ATM call with the BS forward on the projection
    # curve and the OIS discount factor (test_multicurve.cpp:54-67)
    objs = _objs("multicurve")
    assert yf.mcl_pricers(objs) == ["call_mcl"]
    t = yf.translate(objs, "call_mcl", paths=1 << 15)
    assert list(t.spec.rate_values) == [8.0, 8.0] and list(t.spec.discount_values) == [5.0, 5.0]
    r = oracle.price(BuiltRequest(t.spec))
    T = 365 / 365.0
    d1 = (0.08 + 0.045) * T / (0.3 * math.sqrt(T))
    ref = math.exp(-0.05 * T) * (100 * math.exp(0.08 * T) * norm.cdf(d1) - 100 * norm.cdf(d1 - 0.3 * math.sqrt(T)))
    assert abs(r.premium - ref) <= 6 * r.premium_trust + 1e-2
