"""TEST INFRASTRUCTURE ONLY: ctypes loader of the CPU FP64 oracle
(oracle/lib/libmcl_oracle.so, built by oracle/Makefile).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np

from thoth_amd import abi

ROOT = Path(__file__).resolve().parent
LIB = ROOT / "lib" / "libmcl_oracle.so"
JOE_KUO = ROOT.parent / "thoth_amd" / "data" / "new-joe-kuo-6.21201.4096.txt"

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        raise RuntimeError("oracle not built: run `make -C oracle`")
    L = C.CDLL(str(LIB))
    P = C.POINTER
    L.oracle_last_error.restype = C.c_char_p
    L.oracle_load_joe_kuo.argtypes = [C.c_char_p]
    L.oracle_sobol_directions.argtypes = [C.c_int, P(C.c_uint32)]
    L.oracle_sobol_int.argtypes = [C.c_int, C.c_int64, C.c_int64, P(C.c_uint32)]
    L.oracle_normal_cdf_inv.argtypes = [C.c_int64, P(C.c_double), P(C.c_double)]
    L.oracle_xoshiro.argtypes = [C.c_uint64, C.c_int64, P(C.c_uint64)]
    L.oracle_gaussians.argtypes = [C.c_uint64, C.c_int64, P(C.c_double)]
    L.oracle_bridge_schedule.argtypes = [C.c_int, P(C.c_double)] + [P(C.c_int)] * 3 + [P(C.c_double)] * 3
    L.oracle_path_noise.argtypes = [C.c_int, P(C.c_double), C.c_int, C.c_int64, C.c_int64, P(C.c_double)]
    L.oracle_path_noise_seeded.argtypes = [C.c_int, P(C.c_double), C.c_int, C.c_int64, C.c_int64, C.c_uint64,
                                           C.c_int, P(C.c_double)]
    L.oracle_xoshiro_raw.argtypes = [P(C.c_uint64), C.c_int64, P(C.c_uint64)]
    L.oracle_least_squares.argtypes = [P(C.c_double), C.c_int64, C.c_int, P(C.c_double), P(C.c_double)]
    L.oracle_lsm.argtypes = [P(C.c_double), C.c_int64, C.c_int, P(C.c_double), P(C.c_double), C.c_double, C.c_int,
                             P(C.c_double), P(C.c_int), P(C.c_double)]
    L.oracle_lsm.restype = C.c_double
    L.oracle_sabr_vol.argtypes = [P(abi.SabrSurface), C.c_double, C.c_double, C.c_double]
    L.oracle_sabr_vol.restype = C.c_double
    L.oracle_sabr_hagan.argtypes = [P(abi.SabrSurface), C.c_double, C.c_double, C.c_double]
    L.oracle_sabr_hagan.restype = C.c_double
    L.oracle_local_vol.argtypes = [P(abi.Request), C.c_int, C.c_double, C.c_double]
    L.oracle_local_vol.restype = C.c_double
    L.oracle_cholesky.argtypes = [C.c_int, P(C.c_double)]
    L.oracle_grid.argtypes = [P(abi.Request), C.c_int] + [P(C.c_double)] * 3
    L.oracle_avg_correlation.argtypes = [P(abi.Request), C.c_double, C.c_double, P(C.c_double)]
    L.oracle_chol_schedule.argtypes = [P(abi.Request), C.c_int, P(C.c_double)]
    L.oracle_calibrate.argtypes = [P(abi.Request), C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                   P(C.c_double), P(C.c_double), P(C.c_int64), P(C.c_double)]
    L.oracle_sweep.argtypes = [P(abi.Request), P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_double),
                               P(C.c_int64)]
    L.oracle_price.argtypes = [P(abi.Request), P(abi.Result)]
    L.oracle_sweep_seconds.argtypes = [P(abi.Request), P(C.c_double)]
    L.oracle_sweep_seconds.restype = C.c_double
    if L.oracle_load_joe_kuo(str(JOE_KUO).encode()) != 0:
        raise RuntimeError("cannot load %s" % JOE_KUO)
    _lib = L
    return L


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def sobol_directions(dim):
    out = np.zeros((dim, 32), np.uint32)
    lib().oracle_sobol_directions(dim, _p(out, C.c_uint32))
    return out


def sobol_int(dim, n, skip=0):
    out = np.zeros((n, dim), np.uint32)
    lib().oracle_sobol_int(dim, skip, n, _p(out, C.c_uint32))
    return out


def normal_cdf_inv(p):
    p = np.ascontiguousarray(p, np.float64)
    z = np.empty_like(p)
    lib().oracle_normal_cdf_inv(p.size, _p(p), _p(z))
    return z


def xoshiro(seed, n):
    out = np.zeros(n, np.uint64)
    lib().oracle_xoshiro(seed, n, _p(out, C.c_uint64))
    return out


def gaussians(seed, n):
    out = np.zeros(n, np.float64)
    lib().oracle_gaussians(seed, n, _p(out))
    return out


def bridge_schedule(t):
    t = np.ascontiguousarray(t, np.float64)
    m = t.size - 1
    ints = [np.zeros(m, np.int32) for _ in range(3)]
    dbls = [np.zeros(m, np.float64) for _ in range(3)]
    lib().oracle_bridge_schedule(t.size, _p(t), *[_p(a, C.c_int) for a in ints], *[_p(a) for a in dbls])
    return dict(zip(["bridge_index", "left_index", "right_index", "left_weight", "right_weight", "std_dev"],
                    ints + dbls))


def path_noise(t, factors, paths, skip=0, seed=0, use_sobol=True):
    t = np.ascontiguousarray(t, np.float64)
    out = np.zeros((paths, factors, t.size), np.float64)
    lib().oracle_path_noise_seeded(t.size, _p(t), factors, skip, paths, seed % (1 << 64), int(use_sobol), _p(out))
    return out


def least_squares(X, y):
    X = np.ascontiguousarray(X, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    beta = np.zeros(X.shape[1])
    ok = lib().oracle_least_squares(_p(X), X.shape[0], X.shape[1], _p(y), _p(beta))
    return beta if ok else None


def lsm(paths, tau, rates, strike, is_call):
    """Longstaff-Schwartz on paths [N][M]: (value, trust, betas [M][3], fitted [M])."""
    P = np.ascontiguousarray(paths, np.float64)
    N, M = P.shape
    tau = np.ascontiguousarray(tau, np.float64)
    rates = np.ascontiguousarray(rates, np.float64)
    b = np.zeros((M, 3))
    fit = np.zeros(M, np.int32)
    tr = C.c_double()
    v = lib().oracle_lsm(_p(P), N, M, _p(tau), _p(rates), strike, int(is_call), _p(b), _p(fit, C.c_int),
                         C.byref(tr))
    return v, tr.value, b, fit.astype(bool)


def xoshiro_raw(state, n):
    st = np.ascontiguousarray(state, np.uint64)
    out = np.zeros(n, np.uint64)
    lib().oracle_xoshiro_raw(_p(st, C.c_uint64), n, _p(out, C.c_uint64))
    return out


def sabr_vol(surface, K, F, day):
    return lib().oracle_sabr_vol(C.byref(surface), K, F, day)


def sabr_hagan(surface, K, F, T):
    """The Hagan expansion alone (no wings) at maturity T in years."""
    return lib().oracle_sabr_hagan(C.byref(surface), K, F, T)


def avg_correlation(built, a, b):
    """The request's spot/spot correlation averaged over [a, b] years."""
    n = built.req.n_assets
    out = np.zeros((n, n))
    rc = lib().oracle_avg_correlation(C.byref(built.req), a, b, _p(out))
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return out


def chol_schedule(built, cap=100000):
    """Per-date lower Cholesky factors of a term-structured correlation, [dates, n, n]."""
    n = built.req.n_assets
    out = np.zeros((cap, n, n))
    k = lib().oracle_chol_schedule(C.byref(built.req), cap, _p(out))
    if k < 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return out[:k].copy()


def local_vol(built, asset, K, day):
    return lib().oracle_local_vol(C.byref(built.req), asset, K, day)


def cholesky(a):
    a = np.array(a, np.float64, order="C")
    rc = lib().oracle_cholesky(a.shape[0], _p(a))
    if rc != 0:
        raise ValueError("not positive definite")
    return np.tril(a)


def grid(built, cap=100000):
    days, t, dt = (np.zeros(cap) for _ in range(3))
    n = lib().oracle_grid(C.byref(built.req), cap, _p(days), _p(t), _p(dt))
    if n < 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return days[:n].copy(), t[:n].copy(), dt[:n].copy()


def calibrate(built, asset, spot=None, vol_shift=0.0, rate_shift=0.0):
    """Calibrated leverage layers: one per diffusion date (201 points), or the
    surface's time layers on their shared axis."""
    days, _, _ = grid(built)
    n = days.size
    r = built.req
    points = r.leverage_spot_points if r.leverage_time_layers > 0 else 201
    t, ln_step = np.zeros(n), np.zeros(n)
    offset = np.zeros(n, np.int64)
    levels = np.zeros((n, points))
    if spot is None:
        spot = built.spec.equities[asset].spot
    k = lib().oracle_calibrate(C.byref(r), asset, spot, vol_shift, rate_shift, n, _p(t), _p(ln_step),
                               _p(offset, C.c_int64), _p(levels))
    if k < 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return dict(t=t[:k], ln_step=ln_step[:k], offset=offset[:k], levels=levels[:k])


def sweep(built):
    nc = max(1, len(built.spec.contracts))
    bs, bs2 = C.c_double(), C.c_double()
    s, s2 = np.zeros(nc), np.zeros(nc)
    cnt = C.c_int64()
    rc = lib().oracle_sweep(C.byref(built.req), C.byref(bs), C.byref(bs2), _p(s), _p(s2), C.byref(cnt))
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return dict(book_sum=bs.value, book_sum2=bs2.value, sum=s, sum2=s2, count=cnt.value)


def price(built):
    res = built.result()
    rc = lib().oracle_price(C.byref(built.req), C.byref(res))
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return res


def sweep_seconds(built):
    """Wall seconds of the oracle's path sweep alone (one core)."""
    bs = C.c_double()
    sec = lib().oracle_sweep_seconds(C.byref(built.req), C.byref(bs))
    if sec < 0:
        raise RuntimeError(lib().oracle_last_error().decode())
    return sec, bs.value
