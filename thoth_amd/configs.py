"""The named workloads of this engine, as `book.BookSpec` builders.

All random inputs come from one seeded xoshiro256++ stream (splitmix64 seeding)
with a fixed draw order, so every run rebuilds the same books.

* black_scholes_call     -- single-asset Black-Scholes European call, S=K=100,
                            1y, vol 20 %, rate 3 %, 2^18 Sobol paths, 1 step, FP64.
* heston_down_out_call   -- Heston (v0=theta=0.04, kappa 1.5, xi 0.5, rho -0.7)
                            down-and-out call K=100, H=80 monitored daily, 1y,
                            2^20 paths; calendar grid (365 daily steps) or the
                            uniform 252-step grid.
* lsv_worst_of_basket    -- 5-asset LSV worst-of call (rebased strike 100), 2y in
                            365 two-day steps, seeded Heston parameters, seeded
                            SABR target surfaces, seeded correlation repaired by
                            eigenvalue clipping, leverage on a 100 spot x 50 time
                            surface, 2^20 paths, FP32.
"""
import math

import numpy as np

from . import abi
from .book import Barrier, BookSpec, Equity, McConfig, RainbowUdl, Sabr, Vanilla

SEED = 20261203


class Xoshiro256pp:
    """xoshiro256++ seeded through splitmix64 (input generation only)."""

    M = (1 << 64) - 1

    def __init__(self, seed):
        z = seed
        self.s = []
        for _ in range(4):
            z = (z + 0x9E3779B97F4A7C15) & self.M
            x = z
            x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & self.M
            x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & self.M
            self.s.append(x ^ (x >> 31))

    def next(self):
        s = self.s
        rotl = lambda x, k: ((x << k) | (x >> (64 - k))) & self.M
        r = (rotl((s[0] + s[3]) & self.M, 23) + s[0]) & self.M
        t = (s[1] << 17) & self.M
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = rotl(s[3], 45)
        return r

    def uniform(self, lo=0.0, hi=1.0):
        u = (self.next() >> 11) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self):
        # Box-Muller on two uniforms (input generation only)
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def repair_correlation(m, floor=1e-3):
    """Nearest correlation matrix by eigenvalue clipping + unit-diagonal rescale
    (spec: web/bff/src/common/correlation-math.ts:61-82)."""
    m = np.asarray(m, np.float64)
    d, v = np.linalg.eigh((m + m.T) / 2)
    out = (v * np.maximum(d, floor)) @ v.T
    inv = 1.0 / np.sqrt(np.where(np.diag(out) > 0, np.diag(out), 1.0))
    out = out * inv[:, None] * inv[None, :]
    np.fill_diagonal(out, 1.0)
    return out


def black_scholes_call(paths=1 << 18, precision=abi.FP64, greeks=0):
    return BookSpec(mcl=McConfig(paths=paths, max_day_step=365),
                    equities=[Equity(spot=100.0, vol_kind=abi.VOL_BS, volatility=20.0)],
                    contracts=[Vanilla(0, 365, 100.0, is_call=True)],
                    rate_days=[0], rate_values=[3.0], greeks=greeks, precision=precision)


def heston_down_out_call(paths=1 << 20, precision=abi.FP32, uniform_steps=0, greeks=0, maturity_days=365):
    eq = Equity(spot=100.0, vol_kind=abi.VOL_HESTON, init_vol=20.0, long_vol=20.0, kappa=1.5, vol_of_vol=0.5,
                rho_spot_var=-0.7)
    c = Barrier(0, maturity_days, 100.0, abi.DOWN_OUT, 80.0, is_call=True, discrete=True,
                monitoring_period_days=1)
    return BookSpec(mcl=McConfig(paths=paths, max_day_step=1), equities=[eq], contracts=[c],
                    rate_days=[0], rate_values=[3.0], greeks=greeks, precision=precision,
                    uniform_steps=uniform_steps)


def lsv_assets(n=5, seed=SEED):
    """Seeded LSV equities (Heston part + SABR target) and correlation matrix."""
    rng = Xoshiro256pp(seed)
    eqs = []
    for _ in range(n):
        spot = rng.uniform(50, 150)
        kappa = rng.uniform(0.5, 3.0)
        theta = rng.uniform(0.01, 0.09)
        v0 = rng.uniform(0.01, 0.09)
        xi = rng.uniform(0.2, 0.8)
        rho = rng.uniform(-0.9, -0.3)
        alpha = rng.uniform(0.18, 0.30)
        srho = rng.uniform(-0.6, -0.3)
        nu = rng.uniform(0.4, 0.7)
        eqs.append(Equity(spot=spot, vol_kind=abi.VOL_LSV, init_vol=100 * math.sqrt(v0),
                          long_vol=100 * math.sqrt(theta), kappa=kappa, vol_of_vol=xi, rho_spot_var=rho,
                          surface=Sabr(maturities=[0.5, 2.0], alpha=[alpha, alpha], beta=[1.0, 1.0],
                                       rho=[srho, srho], nu=[nu, nu])))
    a = np.array([[rng.normal() for _ in range(n)] for _ in range(n)])
    c = a @ a.T
    c = c / np.sqrt(np.outer(np.diag(c), np.diag(c)))
    return eqs, repair_correlation(c)


# the LSV leverage function of the basket workloads: 100 log-spot points x 50 time layers
LEVERAGE_SURFACE = (100, 50)


def lsv_worst_of_basket(paths=1 << 20, precision=abi.FP32, n_assets=5, maturity_days=730, max_day_step=2,
                        correlate_stochastic=False, greeks=0, leverage_surface=LEVERAGE_SURFACE):
    eqs, corr = lsv_assets(n_assets)
    return BookSpec(mcl=McConfig(paths=paths, max_day_step=max_day_step), equities=eqs,
                    contracts=[Vanilla(-1, maturity_days, 100.0, is_call=True)],
                    correlation=corr.tolist(), rate_days=[0], rate_values=[3.0],
                    rainbows=[RainbowUdl(members=list(range(n_assets)), best=False)],
                    greeks=greeks, precision=precision, correlate_stochastic=correlate_stochastic,
                    leverage_surface=leverage_surface)


# Vanilla grids on the LSV underlyings: 41 strikes 60..140 % of spot, 12 monthly
# maturities (days round(365 k / 12)), calls and puts, one book per underlying.
GRID_STRIKES = [60.0 + 2.0 * k for k in range(41)]
GRID_MATURITIES = [round(365 * k / 12) for k in range(1, 13)]


def lsv_vanilla_grids(paths=1 << 18, precision=abi.FP32, greeks=abi.GREEK_DELTA | abi.GREEK_VEGA, n_underlyings=5,
                      strikes=None, maturities=None, leverage_surface=LEVERAGE_SURFACE):
    """(BookSpec, grid arguments) per underlying: single-equity LSV books of
    the basket's underlyings, no contracts beyond the grid."""
    eqs, _ = lsv_assets(5)
    out = []
    for a in range(n_underlyings):
        spec = BookSpec(mcl=McConfig(paths=paths, max_day_step=1), equities=[eqs[a]], contracts=[],
                        rate_days=[0], rate_values=[3.0], greeks=greeks, precision=precision,
                        leverage_surface=leverage_surface)
        out.append((spec, dict(underlying=0, maturity_days=maturities or GRID_MATURITIES,
                               strikes=strikes or GRID_STRIKES)))
    return out
