"""thoth_mcl_sweep_noalloc against thoth_mcl_sweep: the same books give the
same sums bit for bit, on two calls in a row (the second smaller than the
first, so stale scratch contents would show), and a batch beyond the scratch
is refused before any device work, leaving the device usable."""
import numpy as np
import pytest

from thoth_amd import abi, mcl
from thoth_amd.book import BookSpec, BuiltRequest, Equity, McConfig, Sabr, Vanilla

pytestmark = pytest.mark.gpu


def two_equity_book(n_dates=61):
    t = np.arange(n_dates) / 365.0
    dt = np.concatenate([[0.0], np.diff(t)])
    c = np.array([[1.0, 0.0], [0.6, 0.8]])
    return mcl.BookArrays(
        t, dt, [abi.MODEL_BS, abi.MODEL_HESTON], [100.0, 50.0], np.full((2, n_dates), 0.02),
        np.full(n_dates, 0.02),
        [dict(kind=0, is_call=1, underlying=0, flow_index=n_dates - 1, strike=100.0),
         dict(kind=1, is_call=0, underlying=1, flow_index=n_dates - 1, strike=52.0, is_up=0, is_in=0, level=40.0,
              monitor=list(range(1, n_dates))),
         dict(kind=0, is_call=1, underlying=-1, flow_index=n_dates - 1, strike=100.0)],
        vol=[0.25, 0.0], heston=dict(v0=[0.0, 0.05], kappa=[0.0, 1.2], theta=[0.0, 0.04], xi=[0.0, 0.6],
                                     rho=[0.0, -0.6]),
        chol=c, correlate_stochastic=True, rainbows=[(False, [0, 1], [1.0, 2.0])])


def lsv_book(oracle):
    spec = BookSpec(mcl=McConfig(paths=1 << 12, max_day_step=1),
                    equities=[Equity(spot=100, vol_kind=abi.VOL_LSV, init_vol=18, long_vol=26, kappa=2.0,
                                     vol_of_vol=0.4, rho_spot_var=-0.5,
                                     surface=Sabr([0.5, 2.0], [0.22, 0.24], [1, 1], [-0.5, -0.5], [0.6, 0.5]))],
                    contracts=[Vanilla(0, 20, 100.0)], rate_values=[3.0])
    b = BuiltRequest(spec)
    days, t, dt = oracle.grid(b)
    lev = oracle.calibrate(b, 0)
    eq = spec.equities[0]
    return mcl.BookArrays(t, dt, [abi.MODEL_LSV], [eq.spot], np.full((1, len(t)), 0.03), np.full(len(t), 0.03),
                          [dict(kind=0, is_call=1, underlying=0, flow_index=len(t) - 1, strike=100.0)],
                          heston=dict(v0=[0.0324], kappa=[2.0], theta=[0.0676], xi=[0.4], rho=[-0.5]),
                          lev=dict(ln_step=lev["ln_step"][None], offset=lev["offset"][None],
                                   levels=lev["levels"][None]))


def same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert x.count == y.count
        assert x.book_sum == y.book_sum and x.book_sum2 == y.book_sum2
        assert list(x.contract_sum) == list(y.contract_sum) and list(x.contract_sum2) == list(y.contract_sum2)


def test_noalloc_equals_allocating_sweep_on_two_calls(gpu, oracle):
    # first call: two FP64 books under 8 random digital shifts of 2^16 points
    big = [two_equity_book(), two_equity_book(91)]
    kw = dict(paths=1 << 16, precision=abi.FP64, replicates=8, rqmc_seed=5)
    same(gpu.sweep(big, noalloc=True, **kw), gpu.sweep(big, **kw))
    # second call, smaller: one FP32 LSV book (leverage staged in LDS), reference mode
    small = [lsv_book(oracle)]
    kw = dict(paths=3000, precision=abi.FP32, sobol_skip=77)
    a = gpu.sweep(small, noalloc=True, **kw)
    same(a, gpu.sweep(small, **kw))
    assert a[0].count == 3000 and a[0].book_sum > 0


def test_noalloc_refuses_beyond_its_scratch(gpu):
    book = [two_equity_book()]
    # 4096 digital shifts of 4097 dimensions alone take 67 MB > 64 MiB
    with pytest.raises(gpu.McError) as e:
        gpu.sweep(book, paths=1 << 10, precision=abi.FP32, replicates=4096, rqmc_seed=1, noalloc=True)
    assert e.value.code == abi.ERR_UNSUPPORTED and "no-allocation scratch" in str(e.value)
    same(gpu.sweep(book, paths=1 << 12, precision=abi.FP32, noalloc=True),
         gpu.sweep(book, paths=1 << 12, precision=abi.FP32))
