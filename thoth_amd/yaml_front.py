"""YAML front-end of the device engine for the reference's `!mcl_pricer`
configurations, so that a reference YAML prices unchanged.

The reference reads a YAML object graph (named objects with `!tags`,
core/object_registry.cpp:89-126) and writes one result block per pricer
(tasks/pricer.cpp:536-596, field names tasks/result_schema.hpp).  This module
reads the subset of that graph the device engine prices:

* `!mcl_pricer` (tasks/pricer.cpp:27-41, tasks/pricer_mcl.cpp:67) with
  `!mcl_configuration` (tasks/mcl_configuration.cpp:15-30), Sobol or
  pseudo-random (`use_sobol: false`);
* `!book` of `!vanilla` (european, contracts/vanilla.cpp:18-27) and `!barrier`
  (contracts/barrier.cpp:18-33) contracts paid in the pricer's currency;
* underlyings: `!equity` (marketdata/equity.cpp:20-39) with `!bs_volatility`,
  `!heston_volatility` (no jumps) or `!lsv_volatility` over a `!sabr_volatility`
  surface, an optional flat `continuous_dividends` curve, no repo or discrete
  dividends, calendars of weight 1; and `!rainbow` best/worst-of
  (underlyings/rainbow.cpp:22-27) over such equities;
* the currency's `!yield_curve` (marketdata/curve.cpp:18-21), no discount
  curve or rate model; a `!correlation_matrix` whose `<equity>_var` entries
  carry the spot/variance correlation (tasks/pricer.cpp:676-688).

Anything outside it raises `Unsupported` with the reason: the caller keeps
its CPU engine for that pricer, as the reference's gate does
(tasks/pricer_mcl_gpu_gate.cpp:83-106).  Contract nominals are ignored, as
the reference's vanilla and barrier contracts ignore them.  The indicators
priced are premium, delta, gamma, vega, rho and theta (a second sweep one
calendar day later, spot held).
"""
import datetime
from dataclasses import dataclass, field
from typing import Dict, List

import yaml

from . import abi
from .book import Barrier, BookSpec, BuiltRequest, Equity, McConfig, RainbowUdl, Sabr, Vanilla


class Unsupported(ValueError):
    """The pricer lies outside what the device engine prices."""


class _Loader(yaml.SafeLoader):
    """Safe YAML loader that keeps each `!tag` as the `__tag__` of its mapping."""


def _tagged(loader, suffix, node):
    if isinstance(node, yaml.MappingNode):
        d = loader.construct_mapping(node, deep=True)
        d["__tag__"] = suffix
        return d
    if isinstance(node, yaml.SequenceNode):
        return loader.construct_sequence(node, deep=True)
    return loader.construct_scalar(node)


_Loader.add_multi_constructor("!", _tagged)

_GREEKS = {"delta": abi.GREEK_DELTA, "gamma": abi.GREEK_GAMMA, "vega": abi.GREEK_VEGA, "rho": abi.GREEK_RHO}
_BARRIERS = {"up&out": abi.UP_OUT, "up&in": abi.UP_IN, "down&out": abi.DOWN_OUT, "down&in": abi.DOWN_IN}


def load(text: str) -> Dict[str, dict]:
    """The named objects of a reference YAML document."""
    objs = yaml.load(text, Loader=_Loader)
    if not isinstance(objs, dict):
        raise ValueError("a reference YAML document is a mapping of named objects")
    return objs


def mcl_pricers(objs: Dict[str, dict]) -> List[str]:
    """Names of the `!mcl_pricer` objects, in document order."""
    return [k for k, v in objs.items() if isinstance(v, dict) and v.get("__tag__") == "mcl_pricer"]


@dataclass
class Translation:
    spec: BookSpec
    contracts: List[str]                 # contract names, in request order
    result: str                          # result block name of the pricer
    greeks: List[str]                    # requested Greeks the engine computes
    not_computed: List[str] = field(default_factory=list)


def _obj(objs, name, *tags):
    if name not in objs or not isinstance(objs[name], dict):
        raise ValueError("unknown object '%s'" % name)
    o = objs[name]
    if tags and o["__tag__"] not in tags:
        raise Unsupported("'%s' is a !%s" % (name, o["__tag__"]))
    return o


def _days(today, d):
    if not isinstance(d, datetime.date):
        d = datetime.date.fromisoformat(str(d))
    return (d - today).days


def _flat_matrix(corr, n):
    m = corr["matrix"]
    flat = [x for row in m for x in row] if m and isinstance(m[0], list) else list(m)
    if len(flat) != n * n:
        raise ValueError("correlation matrix is not %d x %d" % (n, n))
    return [flat[i * n:(i + 1) * n] for i in range(n)]


def translate(objs: Dict[str, dict], pricer: str, precision: int = abi.FP64, paths: int = 0,
              roll_days: int = 0) -> Translation:
    """The device request of one `!mcl_pricer` (paths > 0 overrides its path
    count; roll_days moves the valuation date forward, spot held, as the
    reference's theta roll does, tasks/pricer.cpp:371-384)."""
    p = _obj(objs, pricer, "mcl_pricer")
    today = p["today"] if isinstance(p["today"], datetime.date) else datetime.date.fromisoformat(str(p["today"]))
    today += datetime.timedelta(days=roll_days)
    ccy_name = p["currency"]
    ccy = _obj(objs, ccy_name, "currency")
    if "rate_model" in ccy:
        raise Unsupported("currency '%s' has a rate_model" % ccy_name)
    curve = _obj(objs, ccy["rate"], "yield_curve")
    rate_days = [_days(today, d) for d in curve["dates"]]
    rate_values = [float(v) for v in curve["values"]]
    # a distinct discount_rate (OIS) curve discounts the flows (marketdata/currency.cpp:19-27, 129-142)
    disc_days, disc_values = [], []
    if "discount_rate" in ccy and ccy["discount_rate"] != ccy["rate"]:
        dc = _obj(objs, ccy["discount_rate"], "yield_curve")
        disc_days = [_days(today, d) for d in dc["dates"]]
        disc_values = [float(v) for v in dc["values"]]

    mc = _obj(objs, p["mcl_configuration"], "mcl_configuration")
    # use_sobol defaults to true (tasks/mcl_configuration.hpp:5, mcl_configuration.cpp:23)
    cfg = McConfig(paths=int(paths or mc["paths"]), max_day_step=int(mc["max_day_step"]),
                   min_day_step=int(mc.get("min_day_step", 1)), vol_year_step=float(mc.get("vol_year_step", 0.0)),
                   use_sobol=bool(mc.get("use_sobol", True)), seed=int(mc.get("seed", 0)),
                   sobol_skip=int(mc.get("sobol_skip", 0)),
                   allow_gpu=bool(mc.get("allow_gpu", False)))

    corr = _obj(objs, p["correlation"], "correlation_matrix") if "correlation" in p else None
    if corr is not None and "matrix" not in corr:
        raise Unsupported("correlation '%s' is not given as a full matrix" % p["correlation"])
    # rows: the underlyings, then the forexs (vol_correl/correlation.cpp:120-150);
    # with `maturities` the matrix field concatenates one matrix per pillar
    # (correlation.cpp:35-83)
    names = list(corr.get("underlyings", [])) + list(corr.get("forexs", [])) if corr else []
    pillars = [float(x) for x in corr.get("maturities", [])] if corr else []
    cmats = []
    if corr:
        flat = list(corr["matrix"])
        flat = [x for row in flat for x in row] if flat and isinstance(flat[0], list) else flat
        k = max(1, len(pillars))
        nn = len(names) * len(names)
        if len(flat) != k * nn:
            raise ValueError("correlation matrix length is not %d x %d per pillar" % (len(names), len(names)))
        cmats = [_flat_matrix({"matrix": flat[q * nn:(q + 1) * nn]}, len(names)) for q in range(k)]

    def correl(a, b, q=0):
        if a == b:
            return 1.0
        if a in names and b in names:
            return float(cmats[q][names.index(a)][names.index(b)])
        return 0.0

    equities, eq_names, rainbows, rb_names = [], [], [], []

    def equity(name):
        if name in eq_names:
            return eq_names.index(name)
        e = _obj(objs, name, "equity")
        if e.get("currency", ccy_name) != ccy_name:
            raise Unsupported("equity '%s' is quoted in another currency" % name)
        for key in ("discrete_dividends", "repo"):
            if key in e:
                raise Unsupported("equity '%s' has %s" % (name, key))
        div = 0.0
        if "continuous_dividends" in e:
            vals = {float(v) for v in _obj(objs, e["continuous_dividends"])["values"]}
            if len(vals) != 1:
                raise Unsupported("equity '%s' has a term structure of dividends" % name)
            div = vals.pop()
        vol = _obj(objs, e["volatility"])
        cal = vol.get("calendar")
        if cal is not None and float(_obj(objs, cal).get("non_working_days_weight", 1)) != 1.0:
            raise Unsupported("volatility calendar '%s' weights non-working days" % cal)
        out = Equity(spot=float(e["spot"]), dividend_yield=div)
        if vol["__tag__"] == "bs_volatility":
            out.volatility = float(vol["volatility"])
        elif vol["__tag__"] in ("heston_volatility", "lsv_volatility"):
            if float(vol.get("jump_intensity", 0)) != 0.0:
                raise Unsupported("Bates jumps on '%s'" % name)
            out.vol_kind = abi.VOL_HESTON if vol["__tag__"] == "heston_volatility" else abi.VOL_LSV
            out.init_vol, out.long_vol = float(vol["init_vol"]), float(vol["long_vol"])
            out.kappa, out.vol_of_vol = float(vol["kappa"]), float(vol["vol_of_vol"])
            if corr is None:
                raise Unsupported("stochastic-vol equity '%s' needs a correlation matrix" % name)
            out.rho_spot_var = correl(name, name + "_var")
            if any(correl(name, name + "_var", q) != out.rho_spot_var for q in range(len(cmats))):
                # correlation.cpp:270-285: the engines read one scalar spot/variance rho
                raise ValueError("spot/variance entries of '%s' must be identical across maturities" % name)
            if out.vol_kind == abi.VOL_LSV:
                s = _obj(objs, vol["surface"], "sabr_volatility")
                out.surface = Sabr([float(x) for x in s["maturities"]], [float(x) for x in s["alpha"]],
                                   [float(x) for x in s["beta"]], [float(x) for x in s["rho"]],
                                   [float(x) for x in s["nu"]])
        else:
            raise Unsupported("volatility '%s' is a !%s" % (e["volatility"], vol["__tag__"]))
        equities.append(out)
        eq_names.append(name)
        return len(equities) - 1

    def underlying(name):
        o = _obj(objs, name)
        if o["__tag__"] == "equity":
            return equity(name)
        if o["__tag__"] == "rainbow":
            if name not in rb_names:
                kind = o["type"]
                if kind not in ("best_of", "worst_of"):
                    raise ValueError("unknown rainbow type '%s'" % kind)
                rainbows.append(RainbowUdl([equity(u) for u in o["underlyings"]], best=kind == "best_of"))
                rb_names.append(name)
            return -1 - rb_names.index(name)
        raise Unsupported("underlying '%s' is a !%s" % (name, o["__tag__"]))

    contracts, c_names = [], []
    for cname in _obj(objs, p["book"], "book")["contracts"]:
        c = _obj(objs, cname, "vanilla", "barrier")
        if c.get("premium_currency", ccy_name) != ccy_name:
            raise Unsupported("contract '%s' is paid in another currency" % cname)
        udl = underlying(c["underlying"])
        days = _days(today, c["maturity"])
        call = {"call": True, "put": False}[c["type"]]
        absolute = bool(c.get("is_absolute_strike", True))
        if c["__tag__"] == "vanilla":
            # helpers/enums.hpp:64-71: european or american (Longstaff-Schwartz on
            # the recorded paths of the equity, tasks/pricer_mcl_american.cpp)
            exercise = c.get("exercise", "european")
            if exercise not in ("european", "american"):
                raise ValueError("unknown exercise mode '%s'" % exercise)
            american = exercise == "american"
            if american and udl < 0:
                raise Unsupported("american contract '%s' on a rainbow" % cname)
            if american and any(k.american and k.underlying != udl for k in contracts if isinstance(k, Vanilla)):
                raise Unsupported("american contracts on two equities in one book")
            contracts.append(Vanilla(udl, days, float(c["strike"]), is_call=call, is_absolute_strike=absolute,
                                     american=american))
        else:
            bt = _BARRIERS[c["barrier_type"]]
            level = float(c.get("barrier_up_level", 0) if bt in (abi.UP_OUT, abi.UP_IN)
                          else c.get("barrier_down_level", 0))
            mon = c["barrier_monitoring_type"]
            if mon not in ("continuous_monitoring", "discrete_monitoring"):
                raise ValueError("unknown barrier monitoring '%s'" % mon)
            contracts.append(Barrier(udl, days, float(c["strike"]), bt, level, is_call=call,
                                     discrete=mon == "discrete_monitoring",
                                     monitoring_period_days=int(c.get("monitoring_period_days", 0)),
                                     is_absolute_strike=absolute))
        c_names.append(cname)

    indicators = [str(x) for x in p.get("indicators", ["premium"])]
    greeks = [g for g in indicators if g in _GREEKS or g == "theta"]
    not_computed = [g for g in indicators if g != "premium" and g not in greeks]
    n = len(equities)
    sub = lambda q: [[correl(eq_names[i], eq_names[j], q) for j in range(n)] for i in range(n)]
    term = len(pillars) > 1
    spec = BookSpec(mcl=cfg, equities=equities, contracts=contracts, correlation=sub(0),
                    rate_days=rate_days, rate_values=rate_values, rainbows=rainbows,
                    greeks=sum(_GREEKS.get(g, 0) for g in greeks), precision=precision,
                    discount_days=disc_days, discount_values=disc_values,
                    correlation_maturities=pillars if term else [],
                    correlation_pillars=[sub(q) for q in range(len(pillars))] if term else [])
    return Translation(spec, c_names, str(p.get("result", pricer)), greeks, not_computed)


def result_block(t: Translation, r, rolled=None) -> dict:
    """The reference's result fields (tasks/pricer.cpp:536-596) from an engine
    result: premium, premium_trust, the requested Greeks, and per contract
    <name>_premium, <name>_premium_trust and <name>_<greek>.  Theta is the
    premium one calendar day later minus the premium (`rolled`, the result of
    the request rolled by a day, tasks/pricer.cpp:371-384)."""
    out = {"premium": r.premium, "premium_trust": r.premium_trust}
    per = {"delta": r.contract_delta, "gamma": r.contract_gamma, "vega": r.contract_vega, "rho": r.contract_rho}
    greeks = [g for g in t.greeks if g != "theta" or rolled is not None]  # no roll given: no theta
    for g in greeks:
        out[g] = rolled.premium - r.premium if g == "theta" else getattr(r, g)
    for c, name in enumerate(t.contracts):
        out[name + "_premium"] = r.contract_premium[c]
        out[name + "_premium_trust"] = r.contract_trust[c]
        for g in greeks:
            out[name + "_" + g] = (rolled.contract_premium[c] - r.contract_premium[c] if g == "theta"
                                   else per[g][c])
    return out


def price(text: str, pricer: str = None, precision: int = abi.FP64, paths: int = 0) -> Dict[str, dict]:
    """Price the `!mcl_pricer` objects of a reference YAML document on the
    device (all of them, or the one named): {result block name: fields}.
    Raises Unsupported for a pricer outside the engine's subset."""
    from . import mcl
    objs = load(text)
    out = {}
    for name in ([pricer] if pricer else mcl_pricers(objs)):
        t = translate(objs, name, precision, paths)
        rolled = None
        if "theta" in t.greeks:
            t1 = translate(objs, name, precision, paths, roll_days=1)
            t1.spec.greeks = 0
            rolled = mcl.price(BuiltRequest(t1.spec))
        out[t.result] = result_block(t, mcl.price(BuiltRequest(t.spec)), rolled)
    return out
