"""American books of the reference's tests (pricer/tests/test_american.cpp,
helpers.hpp:150-199) and their independent values: a Cox-Ross-Rubinstein
binomial tree with early exercise at every node, and Black-Scholes."""
import math

import numpy as np
from scipy.stats import norm

from thoth_amd import abi
from thoth_amd.book import BookSpec, Equity, McConfig, Vanilla


def crr_american(S, K, r, v, T, is_call, n=2000):
    """Cox-Ross-Rubinstein binomial tree with early exercise at every node."""
    dt = T / n
    u = math.exp(v * math.sqrt(dt))
    d = 1 / u
    q = (math.exp(r * dt) - d) / (u - d)
    disc = math.exp(-r * dt)
    ST = S * u ** np.arange(n, -n - 1, -2.0)
    V = np.maximum(ST - K, 0) if is_call else np.maximum(K - ST, 0)
    for k in range(n, 0, -1):
        ST = ST[:-1] / u
        V = disc * (q * V[:-1] + (1 - q) * V[1:])
        V = np.maximum(V, ST - K if is_call else K - ST)
    return V[0]


def american_book(spot, is_call, paths, greeks=0, precision=abi.FP64, use_sobol=False, max_day_step=30):
    """pricer/tests/helpers.hpp:150-199: one year (2000-01-01 to 2000-12-31),
    30-day steps, pseudo-random draws, 8 % rate, 30 % vol, strike 100."""
    return BookSpec(mcl=McConfig(paths=paths, max_day_step=max_day_step, min_day_step=-1, vol_year_step=0.01,
                                 use_sobol=use_sobol),
                    equities=[Equity(spot=spot, vol_kind=abi.VOL_BS, volatility=30.0)],
                    contracts=[Vanilla(0, 365, 100.0, is_call=is_call, american=True)],
                    rate_days=[0], rate_values=[8.0], greeks=greeks, precision=precision)


def bs(S, K, r, v, T, is_call):
    d1 = (math.log(S / K) + (r + v * v / 2) * T) / (v * math.sqrt(T))
    d2 = d1 - v * math.sqrt(T)
    if is_call:
        return S * norm.cdf(d1) - K * math.exp(-r * T) * norm.cdf(d2)
    return K * math.exp(-r * T) * norm.cdf(-d2) - S * norm.cdf(-d1)
