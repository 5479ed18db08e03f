"""Term-structured correlation (one Cholesky factor per step, of the matrix
averaged over the step) and multi-curve discounting on the device: the FP64
mode against the CPU oracle at 1e-10 (premiums, contracts and bump Greeks,
rho moving both curves), and the correlated FP32 basket kernel against the
FP64 device mode on the same points."""
import numpy as np
import pytest

from thoth_amd import abi, configs
from thoth_amd.book import Barrier, BookSpec, BuiltRequest, Equity, McConfig, RainbowUdl, Vanilla

pytestmark = pytest.mark.gpu

REL = 1e-10

MATS = [0.1, 0.4, 1.0]
PILLARS = [
    [[1.0, 0.8, 0.3], [0.8, 1.0, 0.5], [0.3, 0.5, 1.0]],
    [[1.0, 0.1, -0.2], [0.1, 1.0, 0.0], [-0.2, 0.0, 1.0]],
    [[1.0, -0.4, 0.4], [-0.4, 1.0, 0.2], [0.4, 0.2, 1.0]],
]


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def compare(gpu, oracle, spec, tol=REL):
    b = BuiltRequest(spec)
    g, o = gpu.price(b), oracle.price(b)
    print("premium gpu %.12f oracle %.12f" % (g.premium, o.premium))
    assert rel(g.premium, o.premium) < tol
    assert rel(g.premium_trust, o.premium_trust) < max(tol, 1e-9)
    for c in range(len(spec.contracts)):
        assert rel(g.contract_premium[c], o.contract_premium[c]) < tol
    P = abs(o.premium) + 1e-12
    for name, h in (("delta", 1.0), ("gamma", 1.0), ("vega", 1.0), ("rho", 1e-2)):
        assert abs(getattr(g, name) - getattr(o, name)) <= tol * P / h * 10, name
    return g, o


def test_term_correlation_multi_curve_mixed_book_fp64(gpu, oracle):
    spec = BookSpec(mcl=McConfig(paths=1 << 12, max_day_step=3),
                    equities=[Equity(spot=100, volatility=25.0),
                              Equity(spot=50, vol_kind=abi.VOL_HESTON, init_vol=25, long_vol=22, kappa=2.0,
                                     vol_of_vol=0.6, rho_spot_var=-0.5),
                              Equity(spot=80, volatility=35.0)],
                    contracts=[Vanilla(0, 200, 100.0), Vanilla(2, 150, 85.0, is_call=False),
                               Barrier(1, 200, 50.0, abi.UP_OUT, 65.0, discrete=False), Vanilla(-1, 200, 100.0)],
                    rainbows=[RainbowUdl([0, 1, 2], best=True)], rate_days=[0, 365], rate_values=[2.5, 3.5],
                    discount_days=[0, 730], discount_values=[1.5, 2.0], correlate_stochastic=True,
                    correlation_maturities=MATS, correlation_pillars=PILLARS,
                    greeks=abi.GREEK_DELTA | abi.GREEK_VEGA | abi.GREEK_RHO, precision=abi.FP64)
    compare(gpu, oracle, spec)


def test_term_correlation_lsv_basket_fp64(gpu, oracle):
    eqs, _ = configs.lsv_assets(3)
    spec = BookSpec(mcl=McConfig(paths=1 << 12, max_day_step=1), equities=eqs,
                    contracts=[Vanilla(-1, 30, 100.0)], rate_values=[3.0],
                    rainbows=[RainbowUdl([0, 1, 2], best=False)], correlate_stochastic=True,
                    correlation_maturities=[0.02, 0.06], correlation_pillars=PILLARS[:2], precision=abi.FP64)
    compare(gpu, oracle, spec)


def test_term_correlation_basket_fp32_against_fp64(gpu):
    # the correlated uniform-book kernel reading one factor per date
    eqs, corr = configs.lsv_assets(5)
    far = np.eye(5) * 0.6 + 0.4
    kw = dict(mcl=McConfig(paths=1 << 14, max_day_step=2), equities=eqs, contracts=[Vanilla(-1, 730, 100.0)],
              rate_values=[3.0], rainbows=[RainbowUdl(list(range(5)), best=False)], correlate_stochastic=True,
              correlation_maturities=[0.5, 2.0], correlation_pillars=[corr.tolist(), far.tolist()],
              leverage_surface=configs.LEVERAGE_SURFACE)
    a = gpu.price(BuiltRequest(BookSpec(precision=abi.FP64, **kw)))
    b = gpu.price(BuiltRequest(BookSpec(precision=abi.FP32, **kw)))
    flat = gpu.price(BuiltRequest(BookSpec(precision=abi.FP64, **dict(kw, correlation_maturities=[],
                                                                        correlation_pillars=[], correlation=corr))))
    print("term fp32 %.9f fp64 %.9f (constant matrix %.9f)" % (b.premium, a.premium, flat.premium))
    assert rel(b.premium, a.premium) < 1e-4
    assert abs(a.premium - flat.premium) > 1e-6 * a.premium
