"""Path-split pricing over several processes sharing one GPU: the single-device
part of the reference's cluster mode (pricer/src/modes/run_cluster.cpp:241-447).

The master cuts a book's paths into disjoint Sobol blocks
(thoth_mcl_split_paths, run_cluster.cpp:308-328); each worker process prices
its block through the C-ABI exactly as the reference's slave prices the
request the master sends it -- its path count, seed k and sobol_skip
(run_cluster.cpp:330-339) -- and the master pools the blocks path-weighted,
trusts in quadrature (thoth_mcl_pool_results, run_cluster.cpp:402-447).  The
HTTP transport of the reference's cluster is not reproduced: workers are
child processes of the caller, each with its own HIP context on the device.
"""
import copy
import ctypes as C
import multiprocessing as mp

import numpy as np

from . import abi, mcl
from .book import BuiltRequest

_FIELDS = ("premium", "premium_trust", "delta", "gamma", "vega", "rho", "n_dates", "n_scenarios", "sweep_ms",
           "calib_ms")
_ARRAYS = ("contract_premium", "contract_trust", "contract_delta", "contract_gamma", "contract_vega", "contract_rho")


def _slave(args):
    """One worker: price the block (its paths, seed k and Sobol skip)."""
    spec, k, paths, skip = args
    spec = copy.deepcopy(spec)
    spec.mcl.paths, spec.mcl.seed, spec.mcl.sobol_skip = int(paths), int(k), int(skip)
    b = BuiltRequest(spec)
    r = mcl.price(b)
    nc = max(1, len(spec.contracts))
    out = {f: getattr(r, f) for f in _FIELDS}
    out.update({a: [getattr(r, a)[i] for i in range(nc)] for a in _ARRAYS})
    return out


def price_processes(spec, parts, processes=None):
    """Price `spec` (a book.BookSpec) as `parts` disjoint Sobol blocks in
    separate worker processes (at most `processes` at once, default one per
    block, capped at 16: the processes that may share one GPU), pooled as the
    reference's cluster master pools its slaves.  Returns the pooled
    abi.Result (with per-contract arrays) and the per-block (paths, skips)."""
    paths, skips = mcl.split_paths(spec.mcl.paths, parts)
    n = len(paths)
    workers = max(1, min(16, processes or n))
    with mp.get_context("spawn").Pool(workers) as pool:
        rows = pool.map(_slave, [(spec, k, paths[k], skips[k]) for k in range(n)])
    built = BuiltRequest(spec)
    nc = max(1, len(spec.contracts))
    keep = []
    results = []
    for row in rows:
        r = abi.Result()
        for f in _FIELDS:
            setattr(r, f, row[f])
        for a in _ARRAYS:
            buf = (C.c_double * nc)(*row[a])
            keep.append(buf)
            setattr(r, a, C.cast(buf, abi._pd))
        results.append(r)
    out = mcl.pool_results(built, paths, results)
    out._owner = built  # keeps the per-contract arrays of `out` alive
    return out, (paths, skips)
