"""Device parity: the gfx950 library against the CPU FP64 oracle on the same
Sobol points.  FP64 mode must reproduce the oracle to 1e-10 relative; FP32
mode is held to the tolerance stated on each test."""
import math

import numpy as np
import pytest

from thoth_amd import abi, configs
from thoth_amd.book import Barrier, BookSpec, BuiltRequest, Equity, McConfig, RainbowUdl, Sabr, Vanilla

pytestmark = pytest.mark.gpu

REL = 1e-10  # FP64 parity bar


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def compare(gpu, oracle, spec, greek_scale=None, tol=REL):
    b = BuiltRequest(spec)
    g = gpu.price(b)
    o = oracle.price(b)
    nc = len(spec.contracts)
    assert rel(g.premium, o.premium) < tol, (g.premium, o.premium)
    assert rel(g.premium_trust, o.premium_trust) < max(tol, 1e-9), (g.premium_trust, o.premium_trust)
    for c in range(nc):
        assert rel(g.contract_premium[c], o.contract_premium[c]) < tol
    # Greeks are finite differences of premiums: their error is the premium
    # error divided by the bump, so compare against premium-scaled bounds.
    if spec.greeks:
        P = abs(o.premium) + 1e-12
        for name, h in (("delta", 1.0), ("gamma", 1.0), ("vega", 1.0), ("rho", 1e-2)):
            assert abs(getattr(g, name) - getattr(o, name)) <= tol * P / h * 10, name
    return g, o


def test_device_present(gpu):
    info = gpu.device_info()
    print(info)
    assert "gfx950" in info


@pytest.mark.parametrize("dim,n,skip", [(4097, 300, 0), (10, 20000, 777), (3650, 256, 1 << 20)])
def test_sobol_integer_stage_bit_exact(gpu, oracle, dim, n, skip):
    assert np.array_equal(gpu.sobol_points(dim, n, skip), oracle.sobol_int(dim, n, skip))


def test_config1_black_scholes_fp64(gpu, oracle):
    g, o = compare(gpu, oracle, configs.black_scholes_call(paths=1 << 18, greeks=15))
    assert abs(g.premium - 9.4134033839) < 1e-3


def test_gbm_entry_point(gpu):
    F, K, T, v, r = 100 * math.exp(0.03), 100.0, 1.0, 0.2, 0.03
    p, t = gpu.price_european_gbm(F, K, T, v, math.exp(-r * T), True, 1 << 18, 0)
    assert abs(p - 9.4134033839) < 2e-3 and 0.01 < t < 0.05
    p2, _ = gpu.price_european_gbm(F, K, T, v, math.exp(-r * T), True, 1 << 18, 0)
    assert p2 == p  # deterministic reduction
    assert gpu.price_european_gbm(F, K, T, v, 1.0, True, 0, 0) == (0.0, 0.0)


@pytest.mark.parametrize("seed,paths", [(3, 1 << 16), ((1 << 64) - 1, 1 << 16), ((1 << 31) - 1, 1 << 16),
                                        (1000, 10_000_000)])
def test_gbm_entry_point_prices_every_seed(gpu, seed, paths):
    """Every unsigned long seed prices (mcl_configuration.cpp:24 reads any int,
    cast to unsigned long): a disjoint run of points while it fits below
    2^32, a seeded digital shift past it (a negative seed wraps to 2^64 - k)."""
    F, K, T, v, r = 100 * math.exp(0.03), 100.0, 1.0, 0.2, 0.03
    df = math.exp(-r * T)
    p, t = gpu.price_european_gbm(F, K, T, v, df, True, paths, seed)
    assert abs(p - 9.4134033839) < 5 * t + 1e-3 and t > 0
    assert gpu.price_european_gbm(F, K, T, v, df, True, paths, seed) == (p, t)  # a fixed point set per seed
    if seed == 3:
        # reference mode: points 1 + seed paths .. (seed + 1) paths
        b = gpu.sweep([_c1_book(gpu)], paths, precision=abi.FP64, sobol_skip=seed * paths)[0]
        assert p == pytest.approx(b.book_sum / paths, rel=1e-9, abs=0)  # the same points, discounted in the book
    else:
        assert p != gpu.price_european_gbm(F, K, T, v, df, True, paths, seed - 1)[0]


def test_heston_down_out_fp64(gpu, oracle):
    spec = configs.heston_down_out_call(paths=1 << 12, precision=abi.FP64, greeks=abi.GREEK_DELTA | abi.GREEK_GAMMA | abi.GREEK_VEGA)
    compare(gpu, oracle, spec)


def test_heston_down_out_fp32_error(gpu, oracle):
    spec64 = configs.heston_down_out_call(paths=1 << 14, precision=abi.FP64)
    spec32 = configs.heston_down_out_call(paths=1 << 14, precision=abi.FP32)
    o = oracle.price(BuiltRequest(spec64))
    g = gpu.price(BuiltRequest(spec32))
    err = rel(g.premium, o.premium)
    print("FP32 Heston D&O premium rel error vs FP64 oracle: %.3e" % err)
    assert err < 2e-4


def test_heston_uniform_252_grid_fp64(gpu, oracle):
    compare(gpu, oracle, configs.heston_down_out_call(paths=1 << 12, precision=abi.FP64, uniform_steps=252))


def test_multi_asset_bs_heston_correlated_fp64(gpu, oracle):
    spec = BookSpec(mcl=McConfig(paths=1 << 12, max_day_step=5),
                    equities=[Equity(spot=100, volatility=20.0),
                              Equity(spot=50, vol_kind=abi.VOL_HESTON, init_vol=25, long_vol=22, kappa=2.0,
                                     vol_of_vol=0.6, rho_spot_var=-0.5),
                              Equity(spot=80, volatility=35.0)],
                    correlation=[[1, .5, .2], [.5, 1, .3], [.2, .3, 1]],
                    contracts=[Vanilla(0, 120, 100.0), Vanilla(2, 90, 85.0, is_call=False),
                               Barrier(1, 120, 50.0, abi.UP_OUT, 65.0, discrete=False),
                               Vanilla(-1, 120, 100.0)],
                    rainbows=[RainbowUdl([0, 1, 2], best=True)], rate_values=[2.5],
                    greeks=abi.GREEK_DELTA | abi.GREEK_VEGA | abi.GREEK_RHO)
    compare(gpu, oracle, spec)


def _lsv_basket(n=2, days=40, paths=1 << 12, precision=abi.FP64, corr_stoch=False, surface=None, greeks=0):
    eqs, corr = configs.lsv_assets(n)
    return BookSpec(mcl=McConfig(paths=paths, max_day_step=1), equities=eqs,
                    contracts=[Vanilla(-1, days, 100.0)], correlation=corr.tolist(), rate_values=[3.0],
                    rainbows=[RainbowUdl(list(range(n)), best=False)], precision=precision,
                    correlate_stochastic=corr_stoch, leverage_surface=surface, greeks=greeks)


def test_lsv_worst_of_fp64(gpu, oracle):
    compare(gpu, oracle, _lsv_basket())


@pytest.mark.parametrize("surface", [(100, 12), (64, 50)])
def test_lsv_leverage_surface_fp64(gpu, oracle, surface):
    # the device reads per-date layers folded from the surface on the host; the
    # oracle reads the surface itself with LeverageSurface::GetLeverage
    compare(gpu, oracle, _lsv_basket(n=2, days=40, surface=surface, greeks=abi.GREEK_DELTA | abi.GREEK_VEGA))


def test_lsv_correlated_mode_fp64(gpu, oracle):
    compare(gpu, oracle, _lsv_basket(n=3, days=20, corr_stoch=True))


def test_lsv_leverage_matches_oracle_calibration(gpu, oracle):
    # the sweep fed with the oracle's leverage layers equals the full host pipeline
    spec = _lsv_basket(n=1, days=15)
    b = BuiltRequest(spec)
    days, t, dt = oracle.grid(b)
    lev = oracle.calibrate(b, 0)
    eq = spec.equities[0]
    v0 = (eq.init_vol / 100.0) ** 2
    theta = (eq.long_vol / 100.0) ** 2
    bk = gpu.BookArrays(t, dt, [abi.MODEL_LSV], [eq.spot], np.full((1, len(t)), 0.03), np.full(len(t), 0.03),
                        [dict(kind=0, is_call=1, underlying=-1, flow_index=len(t) - 1, strike=100.0)],
                        heston=dict(v0=[v0], kappa=[eq.kappa], theta=[theta], xi=[eq.vol_of_vol], rho=[eq.rho_spot_var]),
                        lev=dict(ln_step=lev["ln_step"][None], offset=lev["offset"][None], levels=lev["levels"][None]),
                        rainbows=[(False, [0], [100.0 / eq.spot])])
    out = gpu.sweep([bk], 1 << 12)[0]
    o = oracle.sweep(b)
    assert rel(out.book_sum, o["book_sum"]) < REL
    assert rel(out.book_sum2, o["book_sum2"]) < REL


def _c1_book(gpu, vol=0.2):
    t = np.array([0.0, 1.0])
    return gpu.BookArrays(t, t, [abi.MODEL_BS], [100.0], np.full((1, 2), 0.03), np.full(2, 0.03),
                          [dict(kind=0, is_call=1, underlying=0, flow_index=1, strike=100.0)], vol=[vol])


def test_rqmc_replicates_are_unbiased_and_cover(gpu):
    # R random digital shifts of a full 2^k net: each replicate mean is an unbiased
    # estimate, their spread gives the confidence interval
    R, N = 32, 1 << 12
    out = gpu.sweep([_c1_book(gpu)], N, precision=abi.FP64, replicates=R, rqmc_seed=11)
    means = np.array([o.book_sum / o.count for o in out])
    assert len(set(means)) == R  # the shifts differ
    hw = 1.96 * means.std(ddof=1) / math.sqrt(R)
    assert abs(means.mean() - 9.4134033839) < 4 * hw + 1e-9
    assert hw < 0.01  # far tighter than iid sampling at the same budget
    out2 = gpu.sweep([_c1_book(gpu)], N, precision=abi.FP64, replicates=R, rqmc_seed=11)
    assert [o.book_sum for o in out2] == [o.book_sum for o in out]  # reproducible
