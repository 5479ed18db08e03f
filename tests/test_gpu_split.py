"""The path-split form of the reference's cluster on one device
(thoth_mcl_price_split, thoth_amd/cluster.py): the blocks are the oracle's
blocks (each FP64 = oracle at 1e-10), the pooled premium and bump Greeks equal
the unsplit book's (the union of the blocks is the single-box run: a
path-weighted mean of block means is the mean), the pooled trust is the
quadrature of the blocks' trusts, and separate worker processes sharing the
GPU give the in-process result bit for bit."""
import math

import pytest

from thoth_amd import abi, cluster, configs
from thoth_amd.book import BuiltRequest

pytestmark = pytest.mark.gpu


def book(paths=3001):
    return configs.heston_down_out_call(paths=paths, precision=abi.FP64, maturity_days=60,
                                        greeks=abi.GREEK_DELTA | abi.GREEK_GAMMA | abi.GREEK_VEGA)


def test_split_equals_the_single_run_and_its_blocks(gpu, oracle):
    spec = book()
    b = BuiltRequest(spec)
    whole = gpu.price(b)
    pooled = gpu.price_split(b, 4)
    print("whole %.12f pooled %.12f" % (whole.premium, pooled.premium))
    assert abs(pooled.premium - whole.premium) <= 1e-13 * whole.premium
    for g in ("delta", "gamma", "vega"):
        assert abs(getattr(pooled, g) - getattr(whole, g)) <= 1e-11 * whole.premium, g
    paths, skips = gpu.split_paths(spec.mcl.paths, 4)
    assert list(paths) == [751, 750, 750, 750] and list(skips) == [0, 751, 1501, 2251]
    trust2 = 0.0
    for k in range(4):
        spec.mcl.paths, spec.mcl.sobol_skip, spec.mcl.seed = int(paths[k]), int(skips[k]), k
        o = oracle.price(BuiltRequest(spec))
        g = gpu.price(BuiltRequest(spec))
        assert abs(g.premium - o.premium) <= 1e-10 * o.premium
        trust2 += (paths[k] / 3001 * o.premium_trust) ** 2
    assert abs(pooled.premium_trust - math.sqrt(trust2)) <= 1e-10 * pooled.premium_trust


def test_worker_processes_equal_the_in_process_split(gpu):
    spec = book(paths=1 << 12)
    inproc = gpu.price_split(BuiltRequest(spec), 4)
    pooled, (paths, skips) = cluster.price_processes(spec, 4)
    assert len(paths) == 4
    for f in ("premium", "premium_trust", "delta", "gamma", "vega"):
        assert getattr(pooled, f) == getattr(inproc, f), f
    assert pooled.contract_premium[0] == inproc.contract_premium[0]
