"""Books of uneven instruments in one persistent launch: vanilla grids on three
underlyings with different maturities (so different diffusion grids and step
counts: 91, 365 and 120 daily steps, the longest taken first by the work
queue) and different strike sets, swept together by one graph-replayed
persistent launch, give every book exactly what it gives swept alone (each
work item writes its own partial-sum slot, reduced in a fixed order), in
FP64 and in FP32 (books on different dates are not grouped)."""
import numpy as np
import pytest

from thoth_amd import abi, configs
from thoth_amd.book import BookSpec, BuiltRequest, McConfig

pytestmark = pytest.mark.gpu

GRIDS = [([30, 91], [80.0, 100.0, 120.0]),
         ([45, 182, 365], [70.0, 90.0, 100.0, 110.0, 130.0]),
         ([7, 120], [95.0, 105.0])]


def books(precision, paths=1 << 12):
    eqs, _ = configs.lsv_assets(3)
    out = []
    for eq, (mats, pcts) in zip(eqs, GRIDS):
        spec = BookSpec(mcl=McConfig(paths=paths, max_day_step=1), equities=[eq], contracts=[], rate_days=[0],
                        rate_values=[3.0], greeks=abi.GREEK_DELTA | abi.GREEK_VEGA, precision=precision,
                        leverage_surface=configs.LEVERAGE_SURFACE)
        out.append((BuiltRequest(spec), (mats, pcts)))
    return out


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32])
def test_uneven_books_in_one_launch_equal_each_book_alone(gpu, precision):
    bs = books(precision)
    grids = [gpu.vanilla_grid(0, m, p) for _, (m, p) in bs]
    together, _ = gpu.GridPlan([b for b, _ in bs], grids, autotune=False).run()
    for k, (b, (m, p)) in enumerate(bs):
        alone, _ = gpu.GridPlan([b], [gpu.vanilla_grid(0, m, p)], autotune=False).run()
        for key in together[k]:
            assert together[k][key].shape == (len(m), len(p))
            assert np.array_equal(together[k][key], alone[0][key]), (k, key)
    # the cells are priced: calls fall and puts rise with the strike
    for k, (_, (m, p)) in enumerate(bs):
        assert (np.diff(together[k]["call"], axis=1) < 0).all() and (np.diff(together[k]["put"], axis=1) > 0).all()
