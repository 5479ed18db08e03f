"""Python-side builders for the `thoth_mcl_request` of include/thoth_mcl.h.

They mirror the reference's YAML object model (pricer/schema/thoth.schema.json):
`!mcl_configuration`, `!equity` + `!bs_volatility` / `!heston_volatility` /
`!lsv_volatility` (+ `!sabr_volatility` target), `!yield_curve`, `!rainbow`,
`!vanilla`, `!barrier`.  Values use the reference's units (vols and rates in
percent, dates in days from today).  A built request keeps every ctypes array
it points to alive.
"""
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple
import ctypes as C

from . import abi


@dataclass
class McConfig:
    """`!mcl_configuration` (reference: tasks/mcl_configuration.hpp:11-42)."""
    paths: int
    max_day_step: int = 1
    min_day_step: int = 1
    vol_year_step: float = 0.0
    use_sobol: bool = True
    seed: int = 0
    sobol_skip: int = 0
    allow_gpu: bool = True


@dataclass
class Sabr:
    """`!sabr_volatility`: per-maturity (years) Hagan parameters."""
    maturities: Sequence[float]
    alpha: Sequence[float]
    beta: Sequence[float]
    rho: Sequence[float]
    nu: Sequence[float]


@dataclass
class Equity:
    spot: float
    vol_kind: int = abi.VOL_BS
    volatility: float = 20.0          # bs, percent
    init_vol: float = 20.0            # heston/lsv, percent
    long_vol: float = 20.0
    kappa: float = 1.0
    vol_of_vol: float = 0.5
    rho_spot_var: float = 0.0
    dividend_yield: float = 0.0       # percent
    surface: Optional[Sabr] = None


@dataclass
class Vanilla:
    underlying: int                   # equity index, or -(1 + rainbow index)
    maturity_days: int
    strike: float
    is_call: bool = True
    is_absolute_strike: bool = True
    american: bool = False            # exercise: american (Longstaff-Schwartz) or european


@dataclass
class Barrier:
    underlying: int
    maturity_days: int
    strike: float
    barrier_type: int                 # abi.UP_OUT ...
    level: float
    is_call: bool = True
    discrete: bool = True
    monitoring_period_days: int = 1
    is_absolute_strike: bool = True


@dataclass
class RainbowUdl:
    members: Sequence[int]
    best: bool = False


@dataclass
class BookSpec:
    mcl: McConfig
    equities: List[Equity]
    contracts: list
    correlation: Optional[Sequence[Sequence[float]]] = None
    rate_days: Sequence[int] = (0,)
    rate_values: Sequence[float] = (0.0,)   # percent
    rainbows: List[RainbowUdl] = field(default_factory=list)
    greeks: int = 0
    precision: int = abi.FP64
    correlate_stochastic: bool = False
    uniform_steps: int = 0
    # LSV leverage layout: (spot points, time layers) of one shared log-spot axis,
    # or None for one 201-point layer per diffusion date
    leverage_surface: Optional[Tuple[int, int]] = None
    # `!currency discount_rate`: the OIS curve the flows are discounted on
    # (pillar days, percent zero rates); empty: the rate curve discounts
    discount_days: Sequence[int] = ()
    discount_values: Sequence[float] = ()
    # term-structured `!correlation_matrix`: pillar maturities (years) and one
    # n x n matrix per pillar; empty: the constant `correlation`
    correlation_maturities: Sequence[float] = ()
    correlation_pillars: Sequence[Sequence[Sequence[float]]] = ()
    # scenario inputs: a parallel volatility shift of every equity (decimal,
    # the vega scenario's shift) and the rainbow rebasing spots S_i0 (None:
    # the equities' spots)
    vol_shift: float = 0.0
    rebase_spots: Optional[Sequence[float]] = None


class BuiltRequest:
    """A `thoth_mcl_request` plus the buffers it points to."""

    def __init__(self, spec: BookSpec):
        self.spec = spec
        n = len(spec.equities)
        if not 1 <= n <= abi.MAX_ASSETS:
            raise ValueError("1..%d equities supported" % abi.MAX_ASSETS)
        self.assets = (abi.Asset * n)()
        for a, e in enumerate(spec.equities):
            x = self.assets[a]
            x.spot, x.vol_kind, x.volatility = e.spot, e.vol_kind, e.volatility
            x.init_vol, x.long_vol, x.kappa = e.init_vol, e.long_vol, e.kappa
            x.vol_of_vol, x.rho_spot_var, x.dividend_yield = e.vol_of_vol, e.rho_spot_var, e.dividend_yield
            if e.surface is not None:
                s = e.surface
                x.surface.n = len(s.maturities)
                for i in range(len(s.maturities)):
                    x.surface.maturities[i] = s.maturities[i]
                    x.surface.alpha[i] = s.alpha[i]
                    x.surface.beta[i] = s.beta[i]
                    x.surface.rho[i] = s.rho[i]
                    x.surface.nu[i] = s.nu[i]
        corr = spec.correlation
        if corr is None:
            corr = [[1.0 if i == j else 0.0 for j in range(n)] for i in range(n)]
        self.corr = (C.c_double * (n * n))(*[float(corr[i][j]) for i in range(n) for j in range(n)])
        self.rainbows = (abi.Rainbow * max(1, len(spec.rainbows)))()
        for r, rb in enumerate(spec.rainbows):
            self.rainbows[r].best = int(rb.best)
            self.rainbows[r].n = len(rb.members)
            for j, m in enumerate(rb.members):
                self.rainbows[r].members[j] = m
        nc = len(spec.contracts)
        self.contracts = (abi.Contract * max(1, nc))()
        for c, k in enumerate(spec.contracts):
            x = self.contracts[c]
            x.is_call = int(k.is_call)
            x.underlying = k.underlying
            x.maturity_days = k.maturity_days
            x.strike = k.strike
            x.is_absolute_strike = int(k.is_absolute_strike)
            if isinstance(k, Barrier):
                x.kind = abi.BARRIER
                x.barrier_type = k.barrier_type
                x.discrete_monitoring = int(k.discrete)
                x.monitoring_period_days = k.monitoring_period_days
                if k.barrier_type in (abi.UP_OUT, abi.UP_IN):
                    x.barrier_up_level = k.level
                else:
                    x.barrier_down_level = k.level
            else:
                x.kind = abi.VANILLA
                x.american = int(k.american)
        r = abi.Request()
        m = spec.mcl
        r.mcl.paths, r.mcl.max_day_step, r.mcl.min_day_step = m.paths, m.max_day_step, m.min_day_step
        r.mcl.vol_year_step, r.mcl.use_sobol, r.mcl.seed = m.vol_year_step, int(m.use_sobol), m.seed
        r.mcl.sobol_skip, r.mcl.allow_gpu = m.sobol_skip, int(m.allow_gpu)
        r.n_assets = n
        r.assets = self.assets
        r.correlation = self.corr
        r.rate.n = len(spec.rate_days)
        for i, (d, v) in enumerate(zip(spec.rate_days, spec.rate_values)):
            r.rate.days[i] = d
            r.rate.values[i] = v
        r.n_rainbows = len(spec.rainbows)
        r.rainbows = self.rainbows
        r.n_contracts = nc
        r.contracts = self.contracts
        r.request_greeks = spec.greeks
        r.precision = spec.precision
        r.correlate_stochastic = int(spec.correlate_stochastic)
        r.uniform_steps = spec.uniform_steps
        if spec.leverage_surface is not None:
            r.leverage_spot_points, r.leverage_time_layers = spec.leverage_surface
        r.discount.n = len(spec.discount_days)
        for i, (d, v) in enumerate(zip(spec.discount_days, spec.discount_values)):
            r.discount.days[i] = d
            r.discount.values[i] = v
        npl = len(spec.correlation_maturities)
        if npl:
            self.corr_mats = (C.c_double * npl)(*[float(x) for x in spec.correlation_maturities])
            self.corr_pillars = (C.c_double * (npl * n * n))(
                *[float(m[i][j]) for m in spec.correlation_pillars for i in range(n) for j in range(n)])
            r.n_corr_pillars = npl
            r.corr_maturities = self.corr_mats
            r.corr_matrices = self.corr_pillars
        r.vol_shift = spec.vol_shift
        if spec.rebase_spots is not None:
            if len(spec.rebase_spots) != n:
                raise ValueError("one rebasing spot per equity")
            self.rebase = (C.c_double * n)(*[float(x) for x in spec.rebase_spots])
            r.rebase_spots = self.rebase
        self.req = r

    def result(self):
        """An empty `thoth_mcl_result` with per-contract arrays attached."""
        nc = max(1, len(self.spec.contracts))
        res = abi.Result()
        self._arrays = [(C.c_double * nc)() for _ in range(6)]
        (res.contract_premium, res.contract_trust, res.contract_delta, res.contract_gamma,
         res.contract_vega, res.contract_rho) = [C.cast(a, abi._pd) for a in self._arrays]
        return res

    @staticmethod
    def contract_values(res, name, n):
        return [getattr(res, name)[i] for i in range(n)]
