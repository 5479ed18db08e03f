"""Learned martingale control variate on the device: the training kernels'
loss and gradient against the CPU oracle (oracle/mlp_oracle.py), the zero-mean
property of the hedge increments (unbiasedness for any network), and the
variance a trained network removes on the 5-asset LSV worst-of basket,
measured against plain Monte-Carlo on the same randomised-QMC replicates."""
import sys
from pathlib import Path

import numpy as np
import pytest

from thoth_amd import abi, configs
from thoth_amd.book import BuiltRequest

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu


def random_flat(rng, scale=0.25):
    params = [(rng.normal(0, scale, (i, o)), rng.normal(0, 0.1, o)) for i, o in mo.SIZES]
    return params, mo.pack(params)


def device_samples(rng, B, J):
    x = rng.normal(0, 1, (B, J, 11)).astype(np.float32)
    feat = np.zeros((J, B, 16), np.uint16)
    feat[:, :, :11] = mo.bf16_bits(x).transpose(1, 0, 2)
    dm = rng.normal(0, 1, (B, J, 5)).astype(np.float32)
    dmd = np.zeros((J, B, 8), np.float32)
    dmd[:, :, :5] = dm.transpose(1, 0, 2)
    pay = rng.normal(3, 2, B).astype(np.float32)
    return mo.bf16(x), dm.astype(np.float64), pay.astype(np.float64), feat, dmd, pay


@pytest.mark.parametrize("activation", [abi.CV_SILU, abi.CV_GATE])
@pytest.mark.parametrize("B,J", [(64, 1), (256, 5)])
def test_loss_and_gradient_match_oracle(gpu, B, J, activation):
    rng = np.random.default_rng(100 + J)
    params, flat = random_flat(rng)
    x, dm, pay, feat, dmd, payf = device_samples(rng, B, J)
    w = gpu.cv_weights_from_array(flat)
    loss, grad = gpu.cv_loss_grad(w, feat, dmd, payf, activation)
    # the oracle evaluated at the FP32 weights the device holds
    kind = mo.ACTIVATIONS[activation]
    params32 = mo.unpack(np.float32(flat))
    l_mc, g_mc = mo.loss_and_grads(params32, x, dm, pay, matrix_cores=True, activation=kind)
    l_ex, g_ex = mo.loss_and_grads(params32, x, dm, pay, activation=kind)
    ref_mc, ref_ex = mo.pack(g_mc), mo.pack(g_ex)
    err_mc = np.linalg.norm(grad - ref_mc) / np.linalg.norm(ref_mc)
    err_ex = np.linalg.norm(grad - ref_ex) / np.linalg.norm(ref_ex)
    print("%s loss %.8g oracle(bf16) %.8g oracle(fp64) %.8g  grad rel err bf16 %.2e fp64 %.2e"
          % (kind, loss, l_mc, l_ex, err_mc, err_ex))
    assert abs(loss - l_mc) <= 2e-4 * l_mc
    assert err_mc <= 3e-3        # same rounding points as the matrix cores: FP32 accumulation order only
    assert err_ex <= 5e-2        # BF16 operands vs exact arithmetic
    # padding stays untouched: layer-1 inputs 11..15, output units 5..15
    assert not grad[mo.W_OFF[0]:mo.W_OFF[1]].reshape(64, 16)[:, 11:].any()
    assert not grad[mo.W_OFF[3]:mo.B_OFF[0]].reshape(16, 64)[5:].any()
    assert not grad[mo.B_OFF[3] + 5:].any()


def test_gradient_is_deterministic(gpu):
    rng = np.random.default_rng(5)
    _, flat = random_flat(rng)
    _, _, _, feat, dmd, payf = device_samples(rng, 128, 3)
    w = gpu.cv_weights_from_array(flat)
    a = gpu.cv_loss_grad(w, feat, dmd, payf)
    b = gpu.cv_loss_grad(w, feat, dmd, payf)
    assert a[0] == b[0] and np.array_equal(a[1], b[1])


@pytest.fixture(scope="module")
def basket():
    """5-asset LSV worst-of call, 2y, 365 steps, FP32 (the production basket)."""
    return BuiltRequest(configs.lsv_worst_of_basket(paths=1 << 14, precision=abi.FP32))


def test_rqmc_samples_are_finite(gpu, basket):
    """Regression: under these digital shifts replicate 1 maps the coordinate
    of dimension 662 at point 7765 to exactly 0; the FP32 inverse normal must
    stay finite there (it returned -inf, and the variance step turned the path
    into NaN, hidden from the worst-of payoff by fminf)."""
    feat, dm, pay = gpu.cv_samples(basket, 73, 1 << 14, replicates=4, rqmc_seed=3)
    f = (feat.astype(np.uint32) << 16).view(np.float32)
    assert np.isfinite(f).all() and np.isfinite(dm).all() and np.isfinite(pay).all()
    # the whole-path hedge sums have zero mean: replicate means within 4 standard errors
    tot = dm[..., :5].sum(axis=1)                  # [R, paths, 5]
    rep = tot.mean(axis=1)                          # [R, 5]
    se = rep.std(axis=0, ddof=1) / np.sqrt(rep.shape[0])
    assert (np.abs(rep.mean(axis=0)) <= 4 * se + 1e-3).all(), (rep.mean(axis=0), se)


def test_zero_network_is_plain_monte_carlo(gpu, basket):
    cfg = gpu.cv_config()
    r = gpu.cv_price(basket, cfg, abi.CvWeights(), replicates=4, rqmc_seed=3)
    assert r.cv_mean == r.plain_mean and r.cv_var == r.plain_var


def paired_zero_mean(r, pm, cm, k=2.0):
    """The paired difference plain - (P - CV) per replicate is the replicate's
    mean of CV: zero in expectation for an unbiased estimator.  Its mean must lie
    within k half-widths of its own Student-t interval over the replicates (the
    two estimators share their paths, so this interval is far tighter than
    either estimator's own)."""
    d = pm - cm
    assert abs(d.mean() - r.cv_term_mean) <= 1e-12 * max(1.0, abs(r.cv_term_mean))
    assert np.allclose(pm.mean(), r.plain_mean, rtol=1e-14) and np.allclose(cm.mean(), r.cv_mean, rtol=1e-14)
    print("paired CV term %.3e +- %.3e (plain half-width %.3e)" % (r.cv_term_mean, r.cv_term_halfwidth,
                                                                  r.plain_halfwidth))
    assert abs(r.cv_term_mean) <= k * r.cv_term_halfwidth


def test_hedge_increments_have_zero_mean(gpu, basket):
    """A constant unit hedge on every equity: CV = the whole discounted,
    rebased martingale sum, whose mean is zero under the simulated scheme."""
    flat = np.zeros(abi.CV_WEIGHTS)
    flat[mo.B_OFF[3]:mo.B_OFF[3] + 5] = 1.0
    cfg = gpu.cv_config()
    r, pm, cm = gpu.cv_price(basket, cfg, gpu.cv_weights_from_array(flat), replicates=32, rqmc_seed=11,
                             per_replicate=True)
    print("plain %.6f +- %.6f   const-hedge %.6f +- %.6f" % (r.plain_mean, r.plain_halfwidth, r.cv_mean,
                                                            r.cv_halfwidth))
    assert r.cv_mean != r.plain_mean
    paired_zero_mean(r, pm, cm)


@pytest.mark.parametrize("activation", [abi.CV_SILU, abi.CV_GATE])
def test_random_network_is_unbiased(gpu, basket, activation):
    rng = np.random.default_rng(21)
    _, flat = random_flat(rng, scale=0.3)
    r, pm, cm = gpu.cv_price(basket, gpu.cv_config(activation=activation), gpu.cv_weights_from_array(flat),
                             replicates=32, rqmc_seed=12, per_replicate=True)
    print("plain %.6f +- %.6f   random-net %.6f +- %.6f" % (r.plain_mean, r.plain_halfwidth, r.cv_mean,
                                                           r.cv_halfwidth))
    paired_zero_mean(r, pm, cm)


@pytest.mark.parametrize("activation", [abi.CV_SILU, abi.CV_GATE])
def test_trained_network_reduces_variance(gpu, oracle, basket, activation):
    """What the learned control variate gains: trained on engine paths from a
    disjoint Sobol block, evaluated on fresh randomised-QMC replicates, it must
    cut the per-path variance several-fold while the estimate stays within the
    confidence interval of plain MC and of the FP64 price."""
    cfg = gpu.cv_config(train_paths=1 << 16, adam_steps=1000, batch_paths=2048, activation=activation)
    w, rep = gpu.cv_train(basket, cfg)
    print("train: loss %.5f -> %.5f, data %.2f ms, %d Adam steps %.2f ms"
          % (rep.loss_first, rep.loss_last, rep.data_ms, rep.steps, rep.train_ms))
    assert rep.loss_last < 0.8 * rep.loss_first
    r = gpu.cv_price(basket, cfg, w, replicates=16, rqmc_seed=99)
    vr = r.plain_var / r.cv_var
    print("plain %.6f +- %.6f  cv %.6f +- %.6f  per-path variance %.4f -> %.4f (x%.1f)  sweep %.2f ms eval %.2f ms"
          % (r.plain_mean, r.plain_halfwidth, r.cv_mean, r.cv_halfwidth, r.plain_var, r.cv_var, vr, r.sweep_ms,
             r.eval_ms))
    assert vr >= 3.0
    assert abs(r.cv_mean - r.plain_mean) <= 1.5 * (r.cv_halfwidth + r.plain_halfwidth)
    # the two-phase evaluation (stored samples + separate inference kernel) is
    # the same estimator: same replicates, same plain sums, CV within FP32 rounding
    cfg2 = gpu.cv_config(train_paths=1 << 16, adam_steps=1000, batch_paths=2048, eval_mode=abi.CV_TWO_PHASE,
                         activation=activation)
    r2 = gpu.cv_price(basket, cfg2, w, replicates=16, rqmc_seed=99)
    print("two-phase cv %.6f +- %.6f var %.4f  sweep %.2f ms eval %.2f ms" % (r2.cv_mean, r2.cv_halfwidth, r2.cv_var,
                                                                             r2.sweep_ms, r2.eval_ms))
    assert abs(r2.plain_mean - r.plain_mean) <= 1e-6 * abs(r.plain_mean)  # stored payoffs are FP32
    assert abs(r2.cv_mean - r.cv_mean) <= 1e-3 * r.cv_halfwidth + 1e-6
    assert abs(r2.cv_var / r.cv_var - 1) <= 1e-3
    # against the FP64 device price on 2^18 Sobol points (its own iid trust as scale)
    b64 = BuiltRequest(configs.lsv_worst_of_basket(paths=1 << 18, precision=abi.FP64))
    p64 = gpu.price(b64)
    assert abs(r.cv_mean - p64.premium) <= r.cv_halfwidth + 3 * p64.premium_trust


def test_specified_network_at_full_size(gpu):
    """The learned control variate as specified: 11 -> 3 x 64 SiLU -> 5, BF16
    weights with FP32 accumulation, Adam at learning rate 1e-3 for 2000 steps on
    2^18 engine paths of the 5-asset LSV basket (2y, 365 steps, 100 x 50
    leverage surface, FP32), evaluated on 2^20 fresh paths (256 random digital
    shifts of 4096 Sobol points, disjoint from the training block).  Unbiased
    (paired test against plain MC on the same paths) and variance-reducing."""
    b = BuiltRequest(configs.lsv_worst_of_basket(paths=1 << 12, precision=abi.FP32))
    cfg = gpu.cv_config()
    assert (cfg.activation, cfg.adam_steps, cfg.train_paths) == (abi.CV_SILU, 2000, 1 << 18)
    assert abs(cfg.learning_rate - 1e-3) < 1e-9
    w, rep = gpu.cv_train(b, cfg)
    print("train: loss %.5f -> %.5f in %.0f ms (data %.0f ms)" % (rep.loss_first, rep.loss_last, rep.train_ms,
                                                                  rep.data_ms))
    assert rep.steps == 2000 and rep.loss_last < 0.2 * rep.loss_first
    r, pm, cm = gpu.cv_price(b, cfg, w, replicates=256, rqmc_seed=2026, per_replicate=True)
    vr = r.plain_var / r.cv_var
    print("2^20 paths: plain %.6f +- %.6f, CV %.6f +- %.6f, per-path variance ratio %.2f, %.2f ms"
          % (r.plain_mean, r.plain_halfwidth, r.cv_mean, r.cv_halfwidth, vr, r.sweep_ms))
    assert r.replicates * r.paths == 1 << 20
    assert vr >= 8.0
    paired_zero_mean(r, pm, cm)
