"""The reference's sample YAML configurations priced on the device through
the YAML front-end: every `!mcl_pricer` of the samples that the engine
prices, FP64 against the oracle on the same request at 1e-10; the simple-call
sample at its own 50,000 paths against Black-Scholes, and the LSV sample at
its own 200,000 paths against Black(SABR) with pricer/tests/test_lsv.cpp:45-58's
tolerance (6 trust + 0.10)."""
import math
from pathlib import Path

import pytest
from scipy.stats import norm

from thoth_amd import yaml_front as yf
from thoth_amd.book import BuiltRequest

pytestmark = pytest.mark.gpu

SAMPLES = Path(__file__).resolve().parent / "golden" / "reference_samples"
REL = 1e-10


def _cases():
    out = []
    for f in ("simple_call", "lsv", "matrix", "multicurve"):
        objs = yf.load((SAMPLES / (f + ".yaml")).read_text())
        for n in yf.mcl_pricers(objs):
            try:
                yf.translate(objs, n)
                out.append((f, n))
            except yf.Unsupported:
                pass
    return out


@pytest.mark.parametrize("sample,pricer", _cases())
def test_sample_pricer_fp64_matches_oracle(gpu, oracle, sample, pricer):
    print("case: %s / %s" % (sample, pricer), flush=True)
    t = yf.translate(yf.load((SAMPLES / (sample + ".yaml")).read_text()), pricer, paths=1 << 12)
    b = BuiltRequest(t.spec)
    g, o = gpu.price(b), oracle.price(b)
    assert abs(g.premium - o.premium) <= REL * abs(o.premium), (g.premium, o.premium)
    block = yf.result_block(t, g)
    for c, name in enumerate(t.contracts):
        assert abs(block[name + "_premium"] - o.contract_premium[c]) <= REL * max(abs(o.contract_premium[c]), 1e-12)
    P = abs(o.premium) + 1e-12
    for gname, h in (("delta", 1.0), ("gamma", 1.0), ("vega", 1.0), ("rho", 1e-2)):
        if gname in t.greeks:
            assert abs(block[gname] - getattr(o, gname)) <= REL * P / h * 10, gname


def test_simple_call_sample_at_its_own_paths(gpu):
    text = (SAMPLES / "simple_call.yaml").read_text()
    res = yf.price(text)["simple_call_result"]
    for key in ("premium", "premium_trust", "delta", "gamma", "vega", "rho", "theta", "simple_call_premium",
                "simple_call_premium_trust", "simple_call_delta", "simple_call_theta"):
        assert key in res, key
    S, K, r, v, T = 100.0, 100.0, 0.08, 0.30, 1.0
    d1 = (math.log(S / K) + (r + v * v / 2) * T) / (v * math.sqrt(T))
    bs = S * norm.cdf(d1) - K * math.exp(-r * T) * norm.cdf(d1 - v * math.sqrt(T))
    print("simple_call: device %.6f +- %.6f, Black-Scholes %.6f" % (res["premium"], res["premium_trust"], bs))
    assert abs(res["premium"] - bs) < 3 * res["premium_trust"]
    assert abs(res["delta"] - norm.cdf(d1)) < 0.01


def test_lsv_sample_reprices_its_sabr_target(gpu, oracle):
    text = (SAMPLES / "lsv.yaml").read_text()
    t = yf.translate(yf.load(text), "lsv_mcl")
    res = yf.price(text, "lsv_mcl")["lsv_mcl_result"]
    b = BuiltRequest(t.spec)
    T, r = 365 / 365, 0.03
    F = 100 * math.exp(r * T)
    vol = oracle.sabr_vol(b.assets[0].surface, 100.0, F, 365)
    d1 = (math.log(F / 100.0) + 0.5 * vol * vol * T) / (vol * math.sqrt(T))
    ref = math.exp(-r * T) * (F * norm.cdf(d1) - 100.0 * norm.cdf(d1 - vol * math.sqrt(T)))
    print("lsv sample: device %.6f +- %.6f, Black(SABR) %.6f" % (res["premium"], res["premium_trust"], ref))
    assert abs(res["premium"] - ref) <= 6 * res["premium_trust"] + 0.10


def test_theta_matches_the_oracle_roll(gpu, oracle):
    """Theta of the vanilla and barrier samples: the device's one-day roll
    against the oracle's, each premium at 1e-10."""
    objs = yf.load((SAMPLES / "matrix.yaml").read_text())
    for name in ("vanilla_mono_eu_put_mcl", "barrier_mono_discrete_call_mcl"):
        t0 = yf.translate(objs, name, paths=1 << 12)
        t1 = yf.translate(objs, name, paths=1 << 12, roll_days=1)
        t0.spec.greeks = t1.spec.greeks = 0
        b0, b1 = BuiltRequest(t0.spec), BuiltRequest(t1.spec)
        dev = yf.result_block(t0, gpu.price(b0), gpu.price(b1))["theta"]
        ora = yf.result_block(t0, oracle.price(b0), oracle.price(b1))["theta"]
        P = abs(oracle.price(b0).premium)
        assert abs(dev - ora) <= 2e-10 * P, (name, dev, ora)
