"""Vanilla grids whose books are stepped K to a lane (include/thoth_mcl_grid.h,
THOTH_GRID_SEPARATE_BOOKS): single-equity LSV books on the same dates share
the Sobol points, so K of them run as the K equities of one device book on
two shared factors.  Every cell is bit-identical to the cell of the plan that
sweeps each book on its own: the per-equity arithmetic of the grouped kernel
is the one-equity kernel's, operation for operation."""
import numpy as np
import pytest

from thoth_amd import abi, configs
from thoth_amd.book import BuiltRequest

pytestmark = pytest.mark.gpu

KEYS = ("call", "put", "call_trust", "put_trust", "call_delta", "put_delta", "call_vega", "put_vega")


def plans(gpu, n_underlyings, paths, greeks):
    books = configs.lsv_vanilla_grids(paths=paths, precision=abi.FP32, n_underlyings=n_underlyings, greeks=greeks)
    builts = [BuiltRequest(s) for s, _ in books]
    grids = [gpu.vanilla_grid(**g) for _, g in books]
    return gpu.GridPlan(builts, grids, autotune=False, group=True), gpu.GridPlan(builts, grids, autotune=False,
                                                                                  group=False)


@pytest.mark.parametrize("n_underlyings,greeks,k", [(5, abi.GREEK_DELTA | abi.GREEK_VEGA, 5),
                                                    (2, abi.GREEK_DELTA | abi.GREEK_VEGA, 2),
                                                    (3, 0, 1)])
def test_grouped_books_equal_separate_books(gpu, n_underlyings, greeks, k):
    print("case: grouped grid, %d underlyings, greeks %d" % (n_underlyings, greeks), flush=True)
    grouped, separate = plans(gpu, n_underlyings, 1 << 14, greeks)
    assert grouped.tuning().books_per_lane == k
    assert separate.tuning().books_per_lane == 1
    rg, _ = grouped.run()
    rs, _ = separate.run()
    worst = 0.0
    for u in range(n_underlyings):
        for key in KEYS:
            a, b = rg[u][key], rs[u][key]
            assert np.all(np.isfinite(a)), key
            worst = max(worst, np.abs(a - b).max() / max(np.abs(b).max(), 1.0))
            assert np.array_equal(a, b), (u, key)
    print("worst relative difference grouped vs separate: %.2e" % worst)


def test_grouped_plan_replays_bit_identical(gpu):
    grouped, _ = plans(gpu, 5, 1 << 13, abi.GREEK_DELTA | abi.GREEK_VEGA)
    first, _ = grouped.run()
    gpu.price(BuiltRequest(configs.black_scholes_call(paths=1 << 12)))
    again, _ = grouped.run()
    for u in range(5):
        for key in KEYS:
            assert np.array_equal(first[u][key], again[u][key]), key


def test_fp64_books_are_not_grouped(gpu):
    books = configs.lsv_vanilla_grids(paths=1 << 12, precision=abi.FP64, n_underlyings=5)
    plan = gpu.GridPlan([BuiltRequest(s) for s, _ in books], [gpu.vanilla_grid(**g) for _, g in books],
                        autotune=False)
    assert plan.tuning().books_per_lane == 1
