"""The cluster master's path split and pooling (reference:
modes/run_cluster.cpp:308-339, 402-447), host-only entry points of the
library: balanced disjoint blocks that tile the path range, at most one block
per path, and the pooled fields (path-weighted means, trusts in quadrature)."""
import ctypes as C

import numpy as np
import pytest

from thoth_amd import abi, mcl
from thoth_amd.book import BookSpec, BuiltRequest, Equity, McConfig, Vanilla


@pytest.mark.parametrize("total,parts", [(1000, 4), (1003, 4), (7, 16), (1, 3), (2 ** 20 + 5, 7)])
def test_split_is_balanced_disjoint_and_tiles(total, parts):
    paths, skips = mcl.split_paths(total, parts)
    n = min(total, parts)
    assert len(paths) == n and paths.sum() == total
    assert paths.max() - paths.min() <= 1 and list(paths) == sorted(paths, reverse=True)
    assert skips[0] == 0 and all(skips[k] == skips[k - 1] + paths[k - 1] for k in range(1, n))


def test_split_refuses_empty_requests():
    with pytest.raises(mcl.McError):
        mcl.split_paths(0, 4)
    with pytest.raises(mcl.McError):
        mcl.split_paths(100, 0)


def test_pool_fields():
    spec = BookSpec(mcl=McConfig(paths=10), equities=[Equity(spot=100)],
                    contracts=[Vanilla(0, 30, 100.0), Vanilla(0, 60, 90.0)])
    b = BuiltRequest(spec)
    rng = np.random.default_rng(3)
    paths = np.array([400, 300, 300])
    parts, keep = [], []
    for k in range(3):
        r = abi.Result()
        for f in ("premium", "premium_trust", "delta", "gamma", "vega", "rho"):
            setattr(r, f, float(rng.uniform(0.1, 5)))
        for a in ("contract_premium", "contract_trust", "contract_delta", "contract_gamma", "contract_vega",
                  "contract_rho"):
            buf = (C.c_double * 2)(*rng.uniform(0.1, 5, 2))
            keep.append(buf)
            setattr(r, a, C.cast(buf, abi._pd))
        parts.append(r)
    out = mcl.pool_results(b, paths, parts)
    w = paths / paths.sum()
    for f in ("premium", "delta", "gamma", "vega", "rho"):
        assert abs(getattr(out, f) - sum(w[k] * getattr(parts[k], f) for k in range(3))) <= 1e-15
    assert abs(out.premium_trust - np.sqrt(sum((w[k] * parts[k].premium_trust) ** 2 for k in range(3)))) <= 1e-15
    for c in range(2):
        assert abs(out.contract_premium[c] - sum(w[k] * parts[k].contract_premium[c] for k in range(3))) <= 1e-15
        assert abs(out.contract_trust[c] - np.sqrt(sum((w[k] * parts[k].contract_trust[c]) ** 2
                                                       for k in range(3)))) <= 1e-15
