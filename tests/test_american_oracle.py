"""The oracle's Longstaff-Schwartz restatement (reference:
tasks/pricer_mcl_american.cpp:250-447, helpers/maths.cpp:419-470), pinned by:

* the worked example of Longstaff & Schwartz, "Valuing American options by
  simulation" (Rev. Fin. Studies 14, 2001), section 1: 8 paths, K = 1.10,
  r = 6 %, regressions E[Y|S] = -1.070 + 2.983 S - 1.813 S^2 at t = 2 and
  2.038 - 3.335 S + 1.356 S^2 at t = 1, value 0.1144 (each path repeated 10
  times, so that the 50-path minimum of the fit is met; duplicated rows leave
  a least-squares fit unchanged);
* least squares against numpy's on random designs, and the singular design;
* the reference's own American tests (pricer/tests/test_american.cpp) with
  their tolerances, against a Cox-Ross-Rubinstein binomial tree (2,000
  steps) in place of the reference's PDE: the call without dividends is the
  Black-Scholes call, the put exceeds the European put and sits within 5 %
  below the tree, delta and vega track the tree's, gamma > 0, rho < 0, and a
  deep in-the-money put is worth its intrinsic value.
"""
import math

import numpy as np
import pytest

from american_books import american_book, bs, crr_american
from thoth_amd import abi
from thoth_amd.book import BuiltRequest

LS_PATHS = np.array([[1.00, 1.09, 1.08, 1.34], [1.00, 1.16, 1.26, 1.54], [1.00, 1.22, 1.07, 1.03],
                     [1.00, 0.93, 0.97, 0.92], [1.00, 1.11, 1.56, 1.52], [1.00, 0.76, 0.77, 0.90],
                     [1.00, 0.92, 0.84, 1.01], [1.00, 0.88, 1.22, 1.34]])


def test_longstaff_schwartz_worked_example(oracle):
    K = 1.10
    v, trust, b, fit = oracle.lsm(np.repeat(LS_PATHS, 10, axis=0), [0, 1, 2, 3], [0.06] * 4, K, False)
    assert list(fit) == [False, True, True, False]
    # the policy's basis is m = S / K: back to the paper's {1, S, S^2}
    np.testing.assert_allclose([b[2][0], b[2][1] / K, b[2][2] / K ** 2], [-1.070, 2.983, -1.813], atol=6e-4)
    np.testing.assert_allclose([b[1][0], b[1][1] / K, b[1][2] / K ** 2], [2.038, -3.335, 1.356], atol=6e-4)
    assert round(v, 4) == 0.1144


def test_fewer_than_fifty_in_the_money_paths_hold(oracle):
    # 9 copies: 45 in-the-money paths at t = 1 and 2, no fit, no early exercise
    P = np.repeat(LS_PATHS, 9, axis=0)
    v, _, _, fit = oracle.lsm(P, [0, 1, 2, 3], [0.06] * 4, 1.10, False)
    assert not fit.any()
    european = np.mean(np.maximum(1.10 - P[:, 3], 0)) * math.exp(-0.06 * 3)
    # the value is the larger of the held mean and immediate exercise (1.10 - 1.00)
    assert european < 0.1 and v == pytest.approx(0.1, abs=1e-15)
    # below the maturity payoff's mean: the mean with the strike lowered to 0.95
    v2, _, _, _ = oracle.lsm(P, [0, 1, 2, 3], [0.06] * 4, 0.95, False)
    assert abs(v2 - np.mean(np.maximum(0.95 - P[:, 3], 0)) * math.exp(-0.06 * 3)) < 1e-15


def test_least_squares_against_numpy(oracle):
    rng = np.random.default_rng(4)
    for n, p in ((50, 3), (1000, 3), (20, 5)):
        X = rng.normal(size=(n, p))
        y = rng.normal(size=n)
        np.testing.assert_allclose(oracle.least_squares(X, y), np.linalg.lstsq(X, y, rcond=None)[0], rtol=1e-9)
    X = np.ones((60, 3))
    X[:, 2] = 0.0  # a zero column: X^T X has a zero pivot
    assert oracle.least_squares(X, np.ones(60)) is None


def test_american_call_is_the_european_call(oracle):
    res = oracle.price(BuiltRequest(american_book(100.0, True, 200000)))
    assert abs(res.premium - bs(100, 100, 0.08, 0.3, 1.0, True)) <= 6 * res.contract_trust[0] + 5e-2


def test_american_put_against_the_tree(oracle):
    res = oracle.price(BuiltRequest(american_book(100.0, False, 300000)))
    tree = crr_american(100, 100, 0.08, 0.3, 1.0, False)
    euro = bs(100, 100, 0.08, 0.3, 1.0, False)
    lsm, tr = res.premium, res.contract_trust[0]
    print("LSM %.4f +- %.4f, tree %.4f, European %.4f" % (lsm, tr, tree, euro))
    assert tree > euro
    assert lsm > euro
    assert lsm <= tree + 6 * tr + 5e-2
    assert lsm >= 0.95 * tree


def test_american_put_greeks_track_the_tree(oracle):
    g = abi.GREEK_DELTA | abi.GREEK_GAMMA | abi.GREEK_VEGA | abi.GREEK_RHO
    res = oracle.price(BuiltRequest(american_book(100.0, False, 300000, g)))
    h = 0.5
    t_delta = (crr_american(100 + h, 100, 0.08, 0.3, 1, False) - crr_american(100 - h, 100, 0.08, 0.3, 1, False)) / (2 * h)
    t_vega = (crr_american(100, 100, 0.08, 0.31, 1, False) - crr_american(100, 100, 0.08, 0.29, 1, False)) / 2
    print("delta %.4f / %.4f, vega %.4f / %.4f, gamma %.5f, rho %.4f"
          % (res.delta, t_delta, res.vega, t_vega, res.gamma, res.rho))
    assert res.delta < 0 and abs(res.delta - t_delta) <= 0.03
    assert res.vega > 0 and abs(res.vega - t_vega) <= 0.05
    assert res.gamma > 0
    assert res.rho < 0


def test_deep_in_the_money_put_is_near_intrinsic(oracle):
    res = oracle.price(BuiltRequest(american_book(60.0, False, 20000)))
    assert 40.0 - 1e-2 <= res.premium <= 40.0 + 2.0
