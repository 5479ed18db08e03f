"""Device pseudo-random draws against the oracle (reference:
nodes/path_generator.cpp:140-157 for the dimensions past the Sobol table,
nodes/noise_node.cpp:20-25 and tasks/pricer_mcl.cpp:387-392 for use_sobol
false, helpers/rng.hpp:36-72 for the stream):

* an 8-equity LSV worst-of on 365 daily steps (16 factors, 5,840 dimensions,
  1,743 from the stream per path): FP64 device = oracle at 1e-10, FP32
  within its Monte-Carlo error of FP64;
* a Heston barrier on 2,100 daily steps (4,200 dimensions) with bump Greeks;
* use_sobol false books (Black-Scholes, Heston barrier with Greeks,
  correlated LSV worst-of, a negative seed): FP64 device = oracle at 1e-10;
  in that mode sobol_skip moves nothing and the seed moves everything;
* the book-level sweep: two calls give the same sums bit for bit, and the
  no-allocation entry point gives the allocating one's.
"""
import numpy as np
import pytest

from thoth_amd import abi, configs, mcl
from thoth_amd.book import BuiltRequest

pytestmark = pytest.mark.gpu

REL = 1e-10


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def lsv8(paths, precision, use_sobol=True, seed=0):
    spec = configs.lsv_worst_of_basket(paths=paths, precision=precision, n_assets=8, maturity_days=365,
                                       max_day_step=1)
    spec.mcl.use_sobol, spec.mcl.seed = use_sobol, seed
    return spec


def test_eight_equity_lsv_365_steps_past_the_sobol_table_fp64(gpu, oracle):
    spec = lsv8(2048, abi.FP64)
    b = BuiltRequest(spec)
    days, t, _ = gpu.grid(b)
    assert len(t) == 366 and 16 * 365 == 5840
    g, o = gpu.price(b), oracle.price(b)
    print("8-equity LSV, 5840 dims: device %.15f oracle %.15f" % (g.premium, o.premium), flush=True)
    assert rel(g.premium, o.premium) < REL
    assert rel(g.premium_trust, o.premium_trust) < 1e-9


def test_eight_equity_lsv_fp32_against_fp64(gpu):
    g64 = gpu.price(BuiltRequest(lsv8(1 << 16, abi.FP64)))
    g32 = gpu.price(BuiltRequest(lsv8(1 << 16, abi.FP32)))
    print("fp64 %.6f +- %.6f, fp32 %.6f" % (g64.premium, g64.premium_trust, g32.premium), flush=True)
    assert abs(g32.premium - g64.premium) < 0.25 * g64.premium_trust + 1e-4


def test_heston_2100_steps_with_greeks_fp64(gpu, oracle):
    spec = configs.heston_down_out_call(paths=3000, precision=abi.FP64, maturity_days=2100,
                                        greeks=abi.GREEK_DELTA | abi.GREEK_GAMMA | abi.GREEK_VEGA)
    spec.mcl.seed = 12345
    b = BuiltRequest(spec)
    g, o = gpu.price(b), oracle.price(b)
    print("heston 2100 steps: device %.12f oracle %.12f" % (g.premium, o.premium), flush=True)
    assert rel(g.premium, o.premium) < REL
    for name in ("delta", "gamma", "vega"):
        assert abs(getattr(g, name) - getattr(o, name)) <= REL * o.premium * 10, name


def _pseudo_books():
    bs = configs.black_scholes_call(paths=5000)
    hb = configs.heston_down_out_call(paths=3000, precision=abi.FP64, maturity_days=120,
                                      greeks=abi.GREEK_DELTA | abi.GREEK_GAMMA | abi.GREEK_VEGA)
    wo = configs.lsv_worst_of_basket(paths=1500, precision=abi.FP64, n_assets=3, maturity_days=180,
                                     correlate_stochastic=True)
    neg = configs.black_scholes_call(paths=4001)
    neg.mcl.seed = -1
    out = {"bs": bs, "heston_barrier": hb, "lsv_worst_of_correlated": wo, "negative_seed": neg}
    for s in out.values():
        s.mcl.use_sobol = False
    return out


@pytest.mark.parametrize("name", sorted(_pseudo_books()))
def test_use_sobol_false_fp64_matches_oracle(gpu, oracle, name):
    spec = _pseudo_books()[name]
    print("case: pseudo %s" % name, flush=True)
    b = BuiltRequest(spec)
    g, o = gpu.price(b), oracle.price(b)
    print("device %.15f oracle %.15f" % (g.premium, o.premium), flush=True)
    assert rel(g.premium, o.premium) < REL
    assert rel(g.premium_trust, o.premium_trust) < 1e-9
    for gname in ("delta", "gamma", "vega"):
        if spec.greeks & getattr(abi, "GREEK_" + gname.upper()):
            assert abs(getattr(g, gname) - getattr(o, gname)) <= REL * o.premium * 10, gname


def test_use_sobol_false_ignores_the_sobol_skip_and_follows_the_seed(gpu):
    spec = configs.black_scholes_call(paths=3000)
    spec.mcl.use_sobol = False
    a = gpu.price(BuiltRequest(spec))
    spec.mcl.sobol_skip = 4096
    b = gpu.price(BuiltRequest(spec))
    assert a.premium == b.premium and a.premium_trust == b.premium_trust
    spec.mcl.seed = 1
    c = gpu.price(BuiltRequest(spec))
    assert c.premium != a.premium


def test_pseudo_and_tail_modes_refuse_randomised_qmc(gpu):
    spec = configs.black_scholes_call(paths=3000)
    spec.mcl.use_sobol = False
    with pytest.raises(mcl.McError) as e:
        gpu.price_rqmc(BuiltRequest(spec), 4, 1)
    assert e.value.code == abi.ERR_UNSUPPORTED


def _long_book(n_dates):
    t = np.arange(n_dates) / 365.0
    dt = np.concatenate([[0.0], np.diff(t)])
    c = np.array([[1.0, 0.0], [0.6, 0.8]])
    return mcl.BookArrays(
        t, dt, [abi.MODEL_BS, abi.MODEL_HESTON], [100.0, 50.0], np.full((2, n_dates), 0.02),
        np.full(n_dates, 0.02),
        [dict(kind=0, is_call=1, underlying=0, flow_index=n_dates - 1, strike=100.0),
         dict(kind=0, is_call=0, underlying=1, flow_index=n_dates - 1, strike=52.0)],
        vol=[0.25, 0.0], heston=dict(v0=[0.0, 0.05], kappa=[0.0, 1.2], theta=[0.0, 0.04], xi=[0.0, 0.6],
                                     rho=[0.0, -0.6]),
        chol=c, correlate_stochastic=True)


def same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert x.count == y.count
        assert x.book_sum == y.book_sum and x.book_sum2 == y.book_sum2
        assert list(x.contract_sum) == list(y.contract_sum)


@pytest.mark.parametrize("pseudo", [False, True])
@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32])
def test_book_sweep_is_reproducible_and_noalloc_equal(gpu, pseudo, precision):
    # 3 factors x 1500 steps = 4500 dimensions: 403 past the Sobol table
    books = [_long_book(1501), _long_book(1401)]
    kw = dict(paths=1000, precision=precision, pseudo_random=pseudo, seed=9)
    first = gpu.sweep(books, **kw)
    same(first, gpu.sweep(books, **kw))
    same(first, gpu.sweep(books, noalloc=True, **kw))
    assert all(np.isfinite(o.book_sum) and o.count == 1000 for o in first)
