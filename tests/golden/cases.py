"""The small seeded books of the golden fixtures (shared by the script that
writes golden.json and the tests that read it)."""
from thoth_amd import abi, configs
from thoth_amd.book import BookSpec, McConfig, RainbowUdl, Vanilla

GREEKS = abi.GREEK_DELTA | abi.GREEK_GAMMA | abi.GREEK_VEGA | abi.GREEK_RHO


def _lsv_basket(n, days, surface):
    eqs, corr = configs.lsv_assets(n)
    return BookSpec(mcl=McConfig(paths=1 << 10, max_day_step=1), equities=eqs,
                    contracts=[Vanilla(-1, days, 100.0), Vanilla(0, days, 97.0, is_call=False, is_absolute_strike=False)],
                    correlation=corr.tolist(), rate_values=[3.0], rainbows=[RainbowUdl(list(range(n)), best=False)],
                    precision=abi.FP64, leverage_surface=surface, greeks=abi.GREEK_DELTA | abi.GREEK_VEGA)


def cases():
    """name -> BookSpec (FP64)."""
    return {
        "black_scholes_call": configs.black_scholes_call(paths=1 << 12, precision=abi.FP64, greeks=GREEKS),
        "heston_down_out_call_60d": configs.heston_down_out_call(paths=1 << 10, precision=abi.FP64, greeks=GREEKS,
                                                                 maturity_days=60),
        "heston_down_out_call_252_uniform": configs.heston_down_out_call(paths=1 << 9, precision=abi.FP64,
                                                                         uniform_steps=252),
        "lsv_worst_of_2_per_date_layers": _lsv_basket(2, 30, None),
        "lsv_worst_of_3_surface": _lsv_basket(3, 45, (100, 12)),
        "lsv_worst_of_basket_5_surface": configs.lsv_worst_of_basket(paths=1 << 8, precision=abi.FP64,
                                                                     maturity_days=60),
    }


# raw Sobol integers (dimension, first point, count) and inverse-normal probes
SOBOL = [(8, 0, 16), (3650, 1 << 20, 2), (4097, 12345, 1)]
NORMAL_P = [1e-12, 1e-6, 0.02425, 0.1, 0.5, 0.7, 0.97575, 1 - 1e-9]
