"""The workloads at their full sizes on the device, checked through
properties that do not need the CPU oracle at that size: FP32 production
against the FP64 oracle-matched mode on the same Sobol points, closed forms,
put-call parity of every vanilla-grid cell, and RQMC coverage."""
import math

import numpy as np
import pytest

from thoth_amd import abi, configs
from thoth_amd.book import BuiltRequest

pytestmark = pytest.mark.gpu


def test_c1_black_scholes_full_size(gpu):
    # 2^18 paths FP64: the analytic premium inside 4 trusts, Greeks near the closed forms
    r = gpu.price(BuiltRequest(configs.black_scholes_call(paths=1 << 18, precision=abi.FP64,
                                                          greeks=abi.GREEK_DELTA | abi.GREEK_VEGA)))
    assert abs(r.premium - 9.4134033839) <= 4 * r.premium_trust
    assert abs(r.vega - 0.3866681168) < 0.01


def test_c2_heston_barrier_full_size(gpu):
    # 2^20 paths, 252 uniform steps: FP32 premium and bump Greeks against FP64
    greeks = abi.GREEK_DELTA | abi.GREEK_GAMMA | abi.GREEK_VEGA
    r = {p: gpu.price(BuiltRequest(configs.heston_down_out_call(paths=1 << 20, precision=p, uniform_steps=252,
                                                                  greeks=greeks)))
         for p in (abi.FP32, abi.FP64)}
    a, b = r[abi.FP32], r[abi.FP64]
    print("C2 fp32 - fp64: premium %.2e delta %.2e gamma %.2e vega %.2e"
          % (a.premium - b.premium, a.delta - b.delta, a.gamma - b.gamma, a.vega - b.vega))
    assert abs(a.premium - b.premium) <= 2e-5 * b.premium
    assert abs(a.delta - b.delta) <= 1e-4 * abs(b.delta)
    assert abs(a.vega - b.vega) <= 1e-3 * abs(b.vega)
    assert abs(a.gamma - b.gamma) <= 1e-3 * abs(b.gamma)


def test_c3_lsv_basket_full_size(gpu):
    # 2^20 paths, 365 steps, 5 LSV equities on the 100 x 50 surface
    r = {p: gpu.price(BuiltRequest(configs.lsv_worst_of_basket(paths=1 << 20, precision=p)))
         for p in (abi.FP32, abi.FP64)}
    a, b = r[abi.FP32], r[abi.FP64]
    print("C3 fp32 %.8f fp64 %.8f (trust %.2e)" % (a.premium, b.premium, b.premium_trust))
    assert abs(a.premium - b.premium) <= 1e-4 * b.premium
    assert abs(a.premium_trust - b.premium_trust) <= 1e-3 * b.premium_trust


def test_c3_rqmc_interval_full_size(gpu):
    # 16 random digital shifts of 2^16 points: the interval is finite, positive
    # and covers the FP64 point estimate of the same book at 2^20 points
    q = gpu.price_rqmc(BuiltRequest(configs.lsv_worst_of_basket(paths=1 << 16, precision=abi.FP32)), 16, 5)
    b = gpu.price(BuiltRequest(configs.lsv_worst_of_basket(paths=1 << 20, precision=abi.FP64)))
    print("C3 RQMC %.6f +- %.6f, FP64 2^20 %.6f +- %.6f" % (q.mean, q.halfwidth, b.premium, 1.96 * b.premium_trust))
    assert 0 < q.halfwidth < 0.05
    assert abs(q.mean - b.premium) <= q.halfwidth + 3 * b.premium_trust


def test_c4_vanilla_grids_full_size(gpu):
    # 5 underlyings x 984 cells x 2^18 paths: put-call parity of every cell
    # (C - P = df (F - K) up to the MC error of S_T) and FP32 against FP64
    books = configs.lsv_vanilla_grids(paths=1 << 18, precision=abi.FP32)
    grids = [gpu.vanilla_grid(**g) for _, g in books]
    plan = gpu.GridPlan([BuiltRequest(s) for s, _ in books], grids, autotune=False)
    r32, _ = plan.run()
    for s, _ in books:
        s.precision = abi.FP64
    r64, _ = gpu.GridPlan([BuiltRequest(s) for s, _ in books], grids, autotune=False).run()
    worst_parity, worst_prec = 0.0, 0.0
    for u, (spec, g) in enumerate(books):
        S0 = spec.equities[0].spot
        for j, d in enumerate(g["maturity_days"]):
            T = d / 365.0
            df = math.exp(-0.03 * T)
            fwd_pv = S0  # no dividends: PV of the forward = spot
            for k, pct in enumerate(g["strikes"]):
                K = pct * S0 / 100.0
                c, p = r32[u]["call"][j, k], r32[u]["put"][j, k]
                tr = math.hypot(r32[u]["call_trust"][j, k], r32[u]["put_trust"][j, k])
                err = abs((c - p) - (fwd_pv - df * K))
                worst_parity = max(worst_parity, err / max(tr, 1e-12))
                assert err <= 6 * tr + 1e-3 * S0, (u, d, pct, c, p, fwd_pv - df * K)
        for key in ("call", "put", "call_delta", "put_delta", "call_vega", "put_vega"):
            diff = np.abs(r32[u][key] - r64[u][key]).max()
            worst_prec = max(worst_prec, diff)
            assert diff <= 2e-4 * max(np.abs(r64[u][key]).max(), 1.0), (u, key, diff)
    print("C4 worst |parity residual| / trust %.2f, worst |fp32 - fp64| %.2e" % (worst_parity, worst_prec))
