"""Pins of the hedge-network oracle (oracle/mlp_oracle.py) used to check the
device training kernels: gradients against central finite differences, the
zero-network identities, Adam's first step in closed form, the flat weight
layout of the C-ABI and bfloat16 rounding on known values."""
import sys
from pathlib import Path

import numpy as np
import pytest
from scipy.special import expit

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import mlp_oracle as mo  # noqa: E402


def random_params(rng, scale=0.3):
    return [(rng.normal(0, scale, (i, o)), rng.normal(0, 0.1, o)) for i, o in mo.SIZES]


def sample(rng, B=8, J=3):
    return rng.normal(0, 1, (B, J, 11)), rng.normal(0, 1, (B, J, 5)), rng.normal(2, 1, B)


@pytest.mark.parametrize("activation", ["silu", "gate"])
def test_gradients_match_central_differences(activation):
    rng = np.random.default_rng(11)
    params = random_params(rng)
    x, dm, pay = sample(rng)
    _, grads = mo.loss_and_grads(params, x, dm, pay, activation=activation)
    h = 1e-6
    for li in range(4):
        for which in (0, 1):
            p = params[li][which]
            g = grads[li][which]
            for _ in range(6):
                idx = tuple(rng.integers(0, s) for s in p.shape)
                old = p[idx]
                p[idx] = old + h
                lp, _ = mo.loss_and_grads(params, x, dm, pay, activation=activation)
                p[idx] = old - h
                lm, _ = mo.loss_and_grads(params, x, dm, pay, activation=activation)
                p[idx] = old
                fd = (lp - lm) / (2 * h)
                assert abs(fd - g[idx]) <= 1e-6 * max(1.0, abs(fd)), (li, which, idx, fd, g[idx])


def test_zero_output_layer_gives_plain_loss_and_no_hidden_gradient():
    rng = np.random.default_rng(12)
    params = random_params(rng)
    params[3] = (np.zeros((64, 5)), np.zeros(5))
    x, dm, pay = sample(rng)
    loss, grads = mo.loss_and_grads(params, x, dm, pay)
    assert np.isclose(loss, np.mean(pay ** 2), rtol=1e-15)
    for li in range(3):
        assert not grads[li][0].any() and not grads[li][1].any()
    # output-layer bias gradient: d/db_a mean (P - sum_j b.dm_j)^2 = -2 mean P sum_j dm_ja
    want = -2.0 * np.mean(pay[:, None] * dm.sum(axis=1), axis=0)
    assert np.allclose(grads[3][1], want, rtol=1e-12)


def test_adam_first_step_is_signed_learning_rate():
    p = [(np.array([[1.0, -2.0]]), np.array([0.5, 0.0]))]
    g = [(np.array([[3.0, -1e-3]]), np.array([-7.0, 0.0]))]
    out = mo.adam_step(p, g, {}, t=1, lr=1e-3)
    # m_hat = g, v_hat = g^2 at t = 1: step = lr g / (|g| + eps)
    assert np.allclose(out[0][0], [[1.0 - 1e-3 * 3 / (3 + 1e-8), -2.0 + 1e-3 * 1e-3 / (1e-3 + 1e-8)]], rtol=0, atol=1e-15)
    assert np.allclose(out[0][1], [0.5 + 1e-3 * 7 / (7 + 1e-8), 0.0], rtol=0, atol=1e-15)


def test_flat_layout_round_trip_and_positions():
    rng = np.random.default_rng(13)
    params = random_params(rng)
    flat = mo.pack(params)
    back = mo.unpack(flat)
    for (W, b), (W2, b2) in zip(params, back):
        assert np.array_equal(W, W2) and np.array_equal(b, b2)
    # W^T row-major: weight (input i -> output o) of layer 2 sits at W_OFF[1] + o * 64 + i
    assert flat[mo.W_OFF[1] + 7 * 64 + 3] == params[1][0][3, 7]
    # layer-1 inputs 11..15 and output units 5..15 are padding
    assert not flat[mo.W_OFF[0]:mo.W_OFF[1]].reshape(64, 16)[:, 11:].any()
    assert not flat[mo.W_OFF[3]:mo.B_OFF[0]].reshape(16, 64)[5:].any()
    assert flat[mo.B_OFF[3] + 4] == params[3][1][4] and not flat[mo.B_OFF[3] + 5:].any()


def test_bf16_rounding_known_values():
    ulp = 2.0 ** -7
    x = np.array([1.0, 1 + ulp / 2, 1 + 1.5 * ulp, 1 + 0.51 * ulp, -3.0, 0.0, 65280.0])
    want = np.array([1.0, 1.0, 1 + 2 * ulp, 1 + ulp, -3.0, 0.0, 65280.0])
    assert np.array_equal(mo.bf16(x), want)
    assert mo.bf16_bits(np.array([1.0]))[0] == 0x3F80


def test_matrix_core_emulation_is_close_to_exact():
    rng = np.random.default_rng(14)
    params = random_params(rng)
    x, dm, pay = sample(rng, B=64, J=4)
    x = mo.bf16(x)
    l0, g0 = mo.loss_and_grads(params, x, dm, pay)
    l1, g1 = mo.loss_and_grads(params, x, dm, pay, matrix_cores=True)
    assert abs(l1 - l0) <= 2e-2 * l0
    a, b = mo.pack(g0), mo.pack(g1)
    assert np.linalg.norm(a - b) <= 3e-2 * np.linalg.norm(a)


def test_silu_closed_forms():
    # SiLU a(z) = z sigma(z) with scipy's logistic function; a(0) = 0, a'(0) = 1/2,
    # a(z) -> z, a'(z) -> 1 for large z and -> 0 for very negative z (no overflow)
    z = np.concatenate([np.linspace(-40, 40, 8001), [-800.0, 800.0]])
    assert np.allclose(mo.act(z, "silu"), z * expit(z), rtol=1e-15, atol=1e-300)
    assert mo.act(np.array([0.0]), "silu")[0] == 0.0 and mo.dact(np.array([0.0]), "silu")[0] == 0.5
    assert mo.act(np.array([800.0]), "silu")[0] == 800.0 and mo.dact(np.array([800.0]), "silu")[0] == 1.0
    assert abs(mo.act(np.array([-800.0]), "silu")[0]) < 1e-300 and mo.dact(np.array([-800.0]), "silu")[0] == 0.0
    # the derivative against central differences, and a(z) - a(-z) = z (sigma(z) + sigma(-z) = 1)
    zz = np.linspace(-8, 8, 161)
    fd = (mo.act(zz + 1e-6, "silu") - mo.act(zz - 1e-6, "silu")) / 2e-6
    assert np.allclose(fd, mo.dact(zz, "silu"), rtol=0, atol=1e-8)
    assert np.allclose(mo.act(zz, "silu") - mo.act(-zz, "silu"), zz, rtol=0, atol=1e-14)


def test_activation_closed_forms():
    # a(z) = z t^2 (3 - 2t), t = clip(z/6 + 1/2): zero below -3, identity above 3,
    # a(0) = 0 with slope 1/2, a(3/2) = 3/2 * (3/4)^2 * 3/2 = 81/64, continuous
    # value and slope at the joins
    z = np.array([-10.0, -3.0, 0.0, 1.5, 3.0, 10.0])
    assert np.array_equal(mo.act(z), [0.0, 0.0, 0.0, 81.0 / 64.0, 3.0, 10.0])
    assert np.array_equal(mo.dact(z), [0.0, 0.0, 0.5, 0.5625 * 1.5 + 1.5 * 0.75 * 0.25, 1.0, 1.0])
    for j in (-3.0, 3.0):
        e = 1e-7
        assert abs(mo.act(j + e) - mo.act(j - e)) < 3e-7
        assert abs(mo.dact(j + e) - mo.dact(j - e)) < 1e-6
    zz = np.linspace(-5, 5, 101)
    fd = (mo.act(zz + 1e-6) - mo.act(zz - 1e-6)) / 2e-6
    assert np.allclose(fd, mo.dact(zz), atol=1e-5)  # h * the jump of a'' at +-3
