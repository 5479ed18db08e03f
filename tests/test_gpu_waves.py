"""The equity-wave latency form of the FP32 uniform-LSV sweep (the equities of
64 paths stepped by as many waves of one workgroup): on the five-equity
basket the same path values as the uniform lane kernel on the same Sobol
points (the premiums agree to the FP64 rounding of their sums); on two and
eight equities, whose lane form is the generic kernel (spot kept as S, not
log S), the same to FP32 rounding; it is the form in effect for pricings too
small to fill the device; the two-phase control variate on its samples and
the in-loop one agree to FP32 rounding; FP32 against the FP64 oracle."""
import pytest

from thoth_amd import abi, configs, mcl
from thoth_amd.book import BuiltRequest

pytestmark = pytest.mark.gpu


def basket(paths, precision=abi.FP32, **kw):
    return BuiltRequest(configs.lsv_worst_of_basket(paths=paths, precision=precision, **kw))


@pytest.mark.parametrize("n_assets,days,paths,tol", [(5, 730, 3000, 1e-12), (5, 60, 700, 1e-12),
                                                     (2, 90, 1 << 12, 1e-5), (8, 60, 1000, 1e-4)])
def test_equity_waves_equal_the_lane_form(gpu, n_assets, days, paths, tol):
    print("case: %d equities, %d days, %d paths" % (n_assets, days, paths), flush=True)
    plan = mcl.Plan(basket(paths, n_assets=n_assets, maturity_days=days))
    assert plan.shape().workgroups_per_cu == -1  # a small pricing takes the equity-wave form by itself
    waves = plan.run()
    g = plan.shape().lev_group
    plan.set_shape(g, 0)
    lane = plan.run()
    print("waves %.12f lane %.12f" % (waves.premium, lane.premium))
    assert abs(waves.premium - lane.premium) <= tol * abs(lane.premium)
    assert abs(waves.premium_trust - lane.premium_trust) <= max(tol, 1e-10) * 10 * lane.premium_trust
    plan.set_shape(g, -1)
    assert plan.shape().workgroups_per_cu == -1
    assert plan.run().premium == waves.premium


def test_equity_waves_fp32_against_fp64_oracle(gpu, oracle):
    spec = configs.lsv_worst_of_basket(paths=1 << 12, precision=abi.FP64, maturity_days=120)
    o = oracle.price(BuiltRequest(spec))
    spec.precision = abi.FP32
    plan = mcl.Plan(BuiltRequest(spec))
    assert plan.shape().workgroups_per_cu == -1
    g = plan.run()
    print("fp32 waves %.9f fp64 oracle %.9f" % (g.premium, o.premium))
    assert abs(g.premium - o.premium) <= 1e-4 * abs(o.premium)


def test_rqmc_small_pricing_in_the_equity_wave_form(gpu):
    """16 replicates of 1536 paths (the size of a single pricing to a 1bp
    interval with the control variate) run in the equity-wave form: the
    estimate is reproducible and agrees with 16 replicates of 2^15 paths
    (lane form) within the two intervals."""
    small, smeans = mcl.price_rqmc(basket(1536), 16, 9, per_replicate=True)
    again, ameans = mcl.price_rqmc(basket(1536), 16, 9, per_replicate=True)
    assert (smeans == ameans).all()
    big = mcl.price_rqmc(basket(1 << 15), 16, 9)
    print("16 x 1536: %.6f +- %.6f (%.3f ms); 16 x 2^15: %.6f +- %.6f (%.3f ms)"
          % (small.mean, small.halfwidth, small.sweep_ms, big.mean, big.halfwidth, big.sweep_ms))
    assert abs(small.mean - big.mean) <= small.halfwidth + big.halfwidth


def test_two_phase_control_variate_equals_in_loop(gpu):
    """The control variate evaluated after the sweep on its samples (two-phase,
    whose sweep takes the equity-wave form at this size) and inside the
    stepping loop (lane form): the same network on the same features and
    increments, so the same estimate up to FP32 summation order."""
    b = basket(1536)
    cfg_train = gpu.cv_config(adam_steps=200, train_paths=1 << 16)
    w, _ = gpu.cv_train(b, cfg_train)
    r = {}
    for mode in (abi.CV_IN_LOOP, abi.CV_TWO_PHASE):
        cfg = gpu.cv_config(adam_steps=200, train_paths=1 << 16, eval_mode=mode)
        r[mode] = gpu.cv_price(b, cfg, w, replicates=16, rqmc_seed=4)
    a, c = r[abi.CV_IN_LOOP], r[abi.CV_TWO_PHASE]
    print("in-loop %.9f two-phase %.9f (plain %.9f / %.9f)" % (a.cv_mean, c.cv_mean, a.plain_mean, c.plain_mean))
    assert abs(a.plain_mean - c.plain_mean) <= 1e-6 * abs(a.plain_mean)
    assert abs(a.cv_mean - c.cv_mean) <= 1e-5 * abs(a.cv_mean)
    assert abs(a.cv_halfwidth - c.cv_halfwidth) <= 1e-3 * a.cv_halfwidth
