"""The launch-shape tuner of single-book plans (thoth_mcl_tune,
thoth_mcl_set_shape): every shape -- time steps per LDS barrier G in
{1, 2, 4, 8}, one workgroup per Sobol block or a persistent grid of W
workgroups per CU -- gives the fixed shape's premium and Greeks bit for bit,
on the FP32 basket (uniform LSV kernel), the FP32 Heston barrier with bump
Greeks (uniform Heston kernel) and an FP64 book (generic kernel); the tuner's
choice is one of them, is cached by book shape, and bad shapes are refused."""
import pytest

from thoth_amd import abi, configs, mcl
from thoth_amd.book import BuiltRequest

pytestmark = pytest.mark.gpu

FIELDS = ("premium", "premium_trust", "delta", "gamma", "vega")


def books():
    return {
        "basket_fp32": configs.lsv_worst_of_basket(paths=1 << 15, greeks=abi.GREEK_DELTA | abi.GREEK_VEGA),
        "basket_correlated_fp32": configs.lsv_worst_of_basket(paths=1 << 14, correlate_stochastic=True),
        "heston_barrier_fp32": configs.heston_down_out_call(
            paths=1 << 15, uniform_steps=252, greeks=abi.GREEK_DELTA | abi.GREEK_GAMMA | abi.GREEK_VEGA),
        "heston_barrier_fp64": configs.heston_down_out_call(paths=5000, precision=abi.FP64, maturity_days=90,
                                                           greeks=abi.GREEK_DELTA | abi.GREEK_VEGA),
    }


def values(r):
    return [getattr(r, f) for f in FIELDS]


@pytest.mark.parametrize("name", sorted(books()))
def test_every_shape_gives_the_fixed_shape_result(gpu, name):
    print("case: %s" % name, flush=True)
    plan = mcl.Plan(BuiltRequest(books()[name]))
    g0 = plan.shape().lev_group
    if plan.shape().workgroups_per_cu == -1:
        plan.set_shape(g0, 0)  # the equity-wave form a small basket takes by itself: compare the lane forms
    fixed = values(plan.run())
    for g, w in ((1, 0), (2, 0), (8, 0), (g0, 1), (2, 3), (4, 2)):
        try:
            plan.set_shape(g, w)
        except mcl.McError as e:
            assert e.code == abi.ERR_UNSUPPORTED, (g, w)  # the LDS does not fit this G
            continue
        sh = plan.shape()
        assert (sh.lev_group, sh.workgroups_per_cu) == (g, w)
        assert values(plan.run()) == fixed, (g, w)
    chosen = plan.tune()
    print("%s: G %d, W %d, %.3f ms (default %.3f ms), %d candidates" % (
        name, chosen.lev_group, chosen.workgroups_per_cu, chosen.device_ms, chosen.default_ms, chosen.candidates))
    assert chosen.candidates >= 4 and chosen.lev_group in (1, 2, 4, 8) and 0 <= chosen.workgroups_per_cu <= 4
    assert chosen.device_ms <= chosen.default_ms
    assert values(plan.run()) == fixed
    # a second plan of the same shape takes the cached choice
    other = mcl.Plan(BuiltRequest(books()[name]))
    again = other.tune()
    assert again.candidates == 0
    assert (again.lev_group, again.workgroups_per_cu) == (chosen.lev_group, chosen.workgroups_per_cu)
    assert values(other.run()) == fixed


def test_bad_shapes_are_refused(gpu):
    plan = mcl.Plan(BuiltRequest(configs.black_scholes_call(paths=4096)))
    fixed = values(plan.run())
    # -1 (the equity-wave form) exists for FP32 uniform LSV books only
    for g, w, code in ((0, 0, abi.ERR_INVALID), (3, 0, None), (4, 17, abi.ERR_INVALID), (4, -2, abi.ERR_INVALID),
                       (4, -1, abi.ERR_UNSUPPORTED)):
        try:
            plan.set_shape(g, w)
        except mcl.McError as e:
            assert code is None or e.code == code
    plan.set_shape(4, 0)
    assert values(plan.run()) == fixed
