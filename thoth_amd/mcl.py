"""ctypes binding of the gfx950 device library thoth_amd/lib/libthoth_mcl.so
(C-ABI in include/thoth_mcl_gpu.h and include/thoth_mcl.h).

There is no CPU fallback here: if the library is missing, loading fails
loudly; if no gfx950 device is present, the entry points return
THOTH_ERR_UNAVAILABLE and the wrappers raise `Unavailable` -- the caller keeps
its own CPU engine, as the reference's GPU gate does.
"""
import ctypes as C
from pathlib import Path

import numpy as np

from . import abi

LIB = Path(__file__).resolve().parent / "lib" / "libthoth_mcl.so"
_lib = None


class McError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("thoth status %d: %s" % (code, msg))
        self.code = code


class Unavailable(McError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError("device library not built (%s): run `python -m thoth_amd.build`" % LIB)
        L = C.CDLL(str(LIB))
        P = C.POINTER
        L.thoth_last_error.restype = C.c_char_p
        L.thoth_gpu_available.restype = C.c_int
        L.thoth_gpu_device_info.argtypes = [C.c_char_p, C.c_size_t]
        L.thoth_gpu_price_european_gbm.argtypes = [C.c_double] * 5 + [C.c_int, C.c_longlong, C.c_uint64,
                                                                        P(C.c_double), P(C.c_double)]
        L.thoth_gpu_sobol_points.argtypes = [C.c_int32, C.c_int64, C.c_int64, P(C.c_uint32)]
        L.thoth_gpu_special_f32.argtypes = [C.c_int32, C.c_int64, P(C.c_uint32), P(C.c_float)]
        L.thoth_mcl_sweep.argtypes = [C.c_int32, P(abi.Book), P(abi.SweepCfg), P(abi.BookOut)]
        L.thoth_mcl_sweep_noalloc.argtypes = [C.c_int32, P(abi.Book), P(abi.SweepCfg), P(abi.BookOut)]
        L.thoth_prng_state_after.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, P(C.c_uint64)]
        L.thoth_mcl_price.argtypes = [P(abi.Request), P(abi.Result)]
        L.thoth_mcl_prepare.argtypes = [P(abi.Request), P(C.c_void_p)]
        L.thoth_mcl_run.argtypes = [C.c_void_p, P(abi.Result)]
        L.thoth_mcl_tune.argtypes = [C.c_void_p, P(abi.PlanShape)]
        L.thoth_mcl_set_shape.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        L.thoth_mcl_get_shape.argtypes = [C.c_void_p, P(abi.PlanShape)]
        L.thoth_mcl_release.argtypes = [C.c_void_p]
        L.thoth_mcl_release.restype = None
        L.thoth_mcl_grid.argtypes = [P(abi.Request), C.c_int32, P(C.c_double), P(C.c_double), P(C.c_double)]
        L.thoth_mcl_calibrate_leverage.argtypes = [P(abi.Request), C.c_int32, C.c_double, C.c_int32,
                                                   P(C.c_double), P(C.c_double), P(C.c_int64), P(C.c_double)]
        L.thoth_mcl_leverage_tangent.argtypes = [P(abi.Request), C.c_int32, C.c_int32, C.c_int32, P(C.c_double)]
        L.thoth_mcl_price_pathwise.argtypes = [P(abi.Request), P(abi.PathwiseResult)]
        L.thoth_mcl_price_rqmc.argtypes = [P(abi.Request), C.c_int32, C.c_uint64, P(abi.RqmcResult)]
        L.thoth_mcl_split_paths.argtypes = [C.c_int64, C.c_int32, P(C.c_int64), P(C.c_int64)]
        L.thoth_mcl_pool_results.argtypes = [C.c_int32, P(C.c_int64), P(abi.Result), C.c_int32, P(abi.Result)]
        L.thoth_mcl_price_split.argtypes = [P(abi.Request), C.c_int32, P(abi.Result)]
        L.thoth_grid_prepare.argtypes = [C.c_int32, P(abi.Request), P(abi.VanillaGrid), C.c_int32, P(C.c_void_p)]
        L.thoth_grid_prepare_ex.argtypes = [C.c_int32, P(abi.Request), P(abi.VanillaGrid), C.c_int32,
                                            P(C.c_void_p)]
        L.thoth_grid_run.argtypes = [C.c_void_p, P(abi.GridResult), P(C.c_double)]
        L.thoth_grid_get_tuning.argtypes = [C.c_void_p, P(abi.GridTuning)]
        L.thoth_grid_release.argtypes = [C.c_void_p]
        L.thoth_grid_release.restype = None
        L.thoth_cv_train.argtypes = [P(abi.Request), P(abi.CvConfig), P(abi.CvWeights), P(abi.CvTrainReport)]
        L.thoth_cv_price.argtypes = [P(abi.Request), P(abi.CvConfig), P(abi.CvWeights), C.c_int32, C.c_uint64,
                                     P(abi.CvResult)]
        L.thoth_cv_loss_grad.argtypes = [P(abi.CvWeights), P(C.c_uint16), P(C.c_float), P(C.c_float), C.c_int32,
                                         C.c_int32, C.c_int32, P(C.c_double), P(C.c_float)]
        L.thoth_cv_samples.argtypes = [P(abi.Request), C.c_int32, C.c_int64, C.c_int64, C.c_int32, C.c_uint64,
                                       P(C.c_uint16), P(C.c_float), P(C.c_float)]
        _lib = L
    return _lib


def _check(rc):
    if rc != abi.OK:
        msg = lib().thoth_last_error().decode()
        raise (Unavailable if rc == abi.ERR_UNAVAILABLE else McError)(rc, msg)


def available():
    return bool(lib().thoth_gpu_available())


def device_info():
    buf = C.create_string_buffer(256)
    lib().thoth_gpu_device_info(buf, 256)
    return buf.value.decode()


def price_european_gbm(forward, strike, t, vol, df, is_call, paths, seed=0):
    """Reference gpu::PriceEuropeanGbm contract; returns (premium, trust)."""
    p, tr = C.c_double(), C.c_double()
    _check(lib().thoth_gpu_price_european_gbm(forward, strike, t, vol, df, int(is_call), paths, seed,
                                              C.byref(p), C.byref(tr)))
    return p.value, tr.value


def sobol_points(dim, n, skip=0):
    out = np.zeros((n, dim), np.uint32)
    _check(lib().thoth_gpu_sobol_points(dim, skip, n, out.ctypes.data_as(C.POINTER(C.c_uint32))))
    return out


def special_f32(fn, x):
    """Device FP32 special function of the production sweep on uint32 inputs x:
    fn 0 / 1 inverse normal of the Sobol coordinate (generic / split form),
    fn 2 erfc of the float whose bits are x."""
    x = np.ascontiguousarray(x, np.uint32)
    out = np.zeros(x.size, np.float32)
    _check(lib().thoth_gpu_special_f32(fn, x.size, x.ctypes.data_as(C.POINTER(C.c_uint32)),
                                       out.ctypes.data_as(C.POINTER(C.c_float))))
    return out


def price(built):
    """Price a `book.BuiltRequest`; returns the filled `abi.Result`."""
    res = built.result()
    _check(lib().thoth_mcl_price(C.byref(built.req), C.byref(res)))
    return res


def split_paths(total, parts):
    """The cluster master's split: (paths, skips) of the disjoint Sobol blocks."""
    paths = np.zeros(max(1, parts), np.int64)
    skips = np.zeros(max(1, parts), np.int64)
    n = lib().thoth_mcl_split_paths(total, parts, paths.ctypes.data_as(C.POINTER(C.c_int64)),
                                    skips.ctypes.data_as(C.POINTER(C.c_int64)))
    if n < 0:
        _check(-n)
    return paths[:n].copy(), skips[:n].copy()


def pool_results(built, paths, results):
    """Pool per-block `abi.Result`s (each with its per-contract arrays) as the
    cluster master does: path-weighted means, trusts in quadrature."""
    n = len(results)
    arr = (abi.Result * n)(*results)
    p = np.ascontiguousarray(paths, np.int64)
    out = built.result()
    _check(lib().thoth_mcl_pool_results(n, p.ctypes.data_as(C.POINTER(C.c_int64)), arr,
                                        len(built.spec.contracts), C.byref(out)))
    return out


def price_split(built, parts):
    """thoth_mcl_price_split: the book priced as `parts` disjoint Sobol blocks, pooled."""
    res = built.result()
    _check(lib().thoth_mcl_price_split(C.byref(built.req), parts, C.byref(res)))
    return res


def price_rqmc(built, replicates=16, rqmc_seed=1, per_replicate=False):
    """Randomised-QMC premium of the base book with a 95 % confidence interval
    (built.req.mcl.paths paths per replicate); returns abi.RqmcResult, or with
    per_replicate=True (abi.RqmcResult, replicate means)."""
    res = abi.RqmcResult()
    means = np.zeros(replicates)
    if per_replicate:
        res.replicate_means = means.ctypes.data_as(C.POINTER(C.c_double))
    _check(lib().thoth_mcl_price_rqmc(C.byref(built.req), replicates, rqmc_seed, C.byref(res)))
    res.replicate_means = None
    return (res, means) if per_replicate else res


def price_pathwise(built):
    """Pathwise delta / vega (and gamma from pathwise deltas): a single-equity
    Black-Scholes or Heston book with one vanilla or knock-out contract, or a
    book of 1..8 Black-Scholes / Heston / LSV equities with one european vanilla
    on an equity or a rainbow (per-equity deltas in asset_delta[:n_assets])."""
    res = abi.PathwiseResult()
    _check(lib().thoth_mcl_price_pathwise(C.byref(built.req), C.byref(res)))
    return res


class Plan:
    """Device-resident base + bump-scenario books of one request
    (thoth_mcl_prepare / thoth_mcl_run / thoth_mcl_release)."""

    def __init__(self, built):
        self.built = built
        self.handle = C.c_void_p()
        _check(lib().thoth_mcl_prepare(C.byref(built.req), C.byref(self.handle)))

    def run(self):
        res = self.built.result()
        _check(lib().thoth_mcl_run(self.handle, C.byref(res)))
        return res

    def tune(self):
        """Pick the fastest launch shape for this plan's books (cached by book
        shape); returns abi.PlanShape."""
        sh = abi.PlanShape()
        _check(lib().thoth_mcl_tune(self.handle, C.byref(sh)))
        return sh

    def set_shape(self, lev_group, workgroups_per_cu=0):
        _check(lib().thoth_mcl_set_shape(self.handle, lev_group, workgroups_per_cu))

    def shape(self):
        sh = abi.PlanShape()
        _check(lib().thoth_mcl_get_shape(self.handle, C.byref(sh)))
        return sh

    def close(self):
        if self.handle:
            lib().thoth_mcl_release(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def grid(built, cap=100000):
    days, t, dt = (np.zeros(cap) for _ in range(3))
    n = lib().thoth_mcl_grid(C.byref(built.req), cap, *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in (days, t, dt)])
    if n < 0:
        _check(-n)
    return days[:n].copy(), t[:n].copy(), dt[:n].copy()


def calibrate_leverage(built, asset, spot=None):
    """The FP64 calibrated leverage layers of an LSV equity (host only, no device)."""
    days, _, _ = grid(built)
    r = built.req
    surface = r.leverage_time_layers > 0
    n = min(days.size, r.leverage_time_layers) if surface else days.size
    points = r.leverage_spot_points if surface else 201
    t, ln_step, offset, levels = np.zeros(n), np.zeros(n), np.zeros(n, np.int64), np.zeros((n, points))
    if spot is None:
        spot = built.spec.equities[asset].spot
    k = lib().thoth_mcl_calibrate_leverage(C.byref(r), asset, spot, n,
                                           t.ctypes.data_as(C.POINTER(C.c_double)),
                                           ln_step.ctypes.data_as(C.POINTER(C.c_double)),
                                           offset.ctypes.data_as(C.POINTER(C.c_int64)),
                                           levels.ctypes.data_as(C.POINTER(C.c_double)))
    if k < 0:
        _check(-k)
    return dict(t=t[:k], ln_step=ln_step[:k], offset=offset[:k], levels=levels[:k])


def leverage_tangent(built, asset, which):
    """Sensitivity of an LSV equity's per-date calibrated leverage to its spot
    (which 0) or to the parallel vol shift (which 1): the reference's bump
    recalibration read at the base nodes, per unit of bump (host only),
    [dates, points]."""
    days, _, _ = grid(built)
    n = days.size
    points = built.req.leverage_spot_points if built.req.leverage_time_layers > 0 else 201
    dv = np.zeros((n, points))
    k = lib().thoth_mcl_leverage_tangent(C.byref(built.req), asset, which, n,
                                         dv.ctypes.data_as(C.POINTER(C.c_double)))
    if k < 0:
        _check(-k)
    return dv[:k]


class BookArrays:
    """A `thoth_mcl_book` assembled from numpy arrays (kept alive here)."""

    def __init__(self, t, dt, model, spot, drift, disc_rate, contracts, vol=None, heston=None, chol=None,
                 correlate_stochastic=False, lev=None, rainbows=()):
        self._keep = []
        b = abi.Book()
        na = len(model)
        nd = len(t)

        def arr(x, dtype, ctype):
            a = np.ascontiguousarray(x, dtype)
            self._keep.append(a)
            return a.ctypes.data_as(C.POINTER(ctype))

        d = lambda x: arr(x, np.float64, C.c_double)
        i32 = lambda x: arr(x, np.int32, C.c_int32)
        b.n_dates = nd
        b.t, b.dt = d(t), d(dt)
        b.n_assets = na
        b.model = i32(model)
        b.spot = d(spot)
        b.vol = d(vol if vol is not None else np.zeros(na))
        h = heston or {}
        for k in ("v0", "kappa", "theta", "xi", "rho"):
            setattr(b, k, d(h.get(k, np.zeros(na))))
        b.drift = d(np.asarray(drift, np.float64).reshape(na * nd))
        b.chol = d(chol if chol is not None else np.eye(na))
        b.correlate_stochastic = int(correlate_stochastic)
        if lev is not None:
            b.lev_points = lev["levels"].shape[-1]
            b.lev_ln_step = d(np.asarray(lev["ln_step"]).reshape(-1))
            b.lev_offset = arr(np.asarray(lev["offset"]).reshape(-1), np.int64, C.c_int64)
            b.lev_values = d(np.asarray(lev["levels"]).reshape(-1))
        b.disc_rate = d(disc_rate)
        b.n_rainbows = len(rainbows)
        if rainbows:
            mem = np.zeros((len(rainbows), abi.MAX_ASSETS), np.int32)
            w = np.zeros((len(rainbows), abi.MAX_ASSETS))
            for r, (best, members, weights) in enumerate(rainbows):
                mem[r, :len(members)] = members
                w[r, :len(members)] = weights
            b.rb_best = i32([int(r[0]) for r in rainbows])
            b.rb_n = i32([len(r[1]) for r in rainbows])
            b.rb_member = i32(mem.reshape(-1))
            b.rb_weight = d(w.reshape(-1))
        nc = len(contracts)
        b.n_contracts = nc
        get = lambda k, default=0: [c.get(k, default) for c in contracts]
        b.c_kind = i32(get("kind"))
        b.c_is_call = i32(get("is_call", 1))
        b.c_underlying = i32(get("underlying"))
        b.c_flow_index = i32(get("flow_index"))
        b.c_strike = d(get("strike"))
        b.c_is_up = i32(get("is_up"))
        b.c_is_in = i32(get("is_in"))
        b.c_level = d(get("level", 0.0))
        off = [0]
        mon = []
        for c in contracts:
            mon += list(c.get("monitor", []))
            off.append(len(mon))
        b.c_mon_offset = i32(off)
        b.c_mon_index = i32(mon if mon else [0])
        self.book = b


NOALLOC_SCRATCH_BYTES = 64 << 20  # THOTH_NOALLOC_SCRATCH_BYTES


def sweep(books, paths, precision=abi.FP64, sobol_skip=0, replicates=0, rqmc_seed=0, noalloc=False,
          pseudo_random=False, seed=0):
    """Run thoth_mcl_sweep (or, noalloc=True, thoth_mcl_sweep_noalloc, which
    allocates no device memory) over a list of BookArrays; returns the BookOut
    list (len(books) * replicates entries in RQMC mode, book-major).
    pseudo_random=True is use_sobol false; `seed` seeds the xoshiro256++
    stream of those draws and of the dimensions past the Sobol table (an int
    seed is taken modulo 2^64, as the reference's (uint64_t) cast does)."""
    n = len(books)
    arr = (abi.Book * n)(*[b.book for b in books])
    cfg = abi.SweepCfg(paths=paths, sobol_skip=sobol_skip, precision=precision, block=0,
                       replicates=replicates, rqmc_seed=rqmc_seed, pseudo_random=int(bool(pseudo_random)),
                       seed=int(seed) % (1 << 64))
    out = (abi.BookOut * (n * max(1, replicates)))()
    fn = lib().thoth_mcl_sweep_noalloc if noalloc else lib().thoth_mcl_sweep
    _check(fn(n, arr, C.byref(cfg), out))
    return list(out)


def prng_state_after(seed, steps, draws_per_path):
    """xoshiro256++ state after `steps` draws of the stream seeded by `seed`,
    by the device's jump matrices (steps a multiple of draws_per_path)."""
    out = (C.c_uint64 * 4)()
    _check(lib().thoth_prng_state_after(C.c_uint64(int(seed) % (1 << 64)), C.c_uint64(steps),
                                        C.c_uint64(draws_per_path), out))
    return [int(x) for x in out]


# ---- learned martingale control variate (include/thoth_mcl_cv.h) ----------

def cv_config(hedge_dates=97, train_paths=1 << 18, train_sobol_skip=1 << 24, adam_steps=2000, batch_paths=32768,
              learning_rate=1e-3, init_seed=7, eval_mode=abi.CV_IN_LOOP, hedge_spacing=2.0, activation=abi.CV_SILU):
    """Training/evaluation settings; the training paths come from a Sobol block
    (points train_sobol_skip+1 ..) disjoint from the evaluation replicates.
    The defaults are the specified network: SiLU hidden units, Adam at
    learning rate 1e-3, 2000 steps on 2^18 engine paths."""
    return abi.CvConfig(hedge_dates, train_paths, train_sobol_skip, adam_steps, batch_paths, learning_rate, init_seed,
                        eval_mode, hedge_spacing, activation)


def cv_config_gated(**kw):
    """The smoothstep-gated variant (no transcendental in the network) at
    learning rate 2e-3."""
    kw.setdefault("learning_rate", 2e-3)
    return cv_config(activation=abi.CV_GATE, **kw)


def cv_train(built, cfg):
    """Train the hedge network on the base book of `built`; returns
    (abi.CvWeights, abi.CvTrainReport)."""
    w, rep = abi.CvWeights(), abi.CvTrainReport()
    _check(lib().thoth_cv_train(C.byref(built.req), C.byref(cfg), C.byref(w), C.byref(rep)))
    return w, rep


def cv_price(built, cfg, weights, replicates=16, rqmc_seed=1, per_replicate=False):
    """Plain and control-variate RQMC estimates of the base book premium
    (built.req.mcl.paths paths per replicate); returns abi.CvResult, or with
    per_replicate=True (abi.CvResult, plain replicate means, P - CV replicate means)."""
    res = abi.CvResult()
    pm = np.zeros(replicates)
    cm = np.zeros(replicates)
    if per_replicate:
        res.replicate_plain_means = pm.ctypes.data_as(C.POINTER(C.c_double))
        res.replicate_cv_means = cm.ctypes.data_as(C.POINTER(C.c_double))
    _check(lib().thoth_cv_price(C.byref(built.req), C.byref(cfg), C.byref(weights), replicates, rqmc_seed,
                                C.byref(res)))
    res.replicate_plain_means = None
    res.replicate_cv_means = None
    return (res, pm, cm) if per_replicate else res


def cv_samples(built, hedge_dates, paths, sobol_skip=0, replicates=0, rqmc_seed=0):
    """Control-variate samples of the base book (FP32 sweep): feat uint16
    [R', J, paths, 16] (BF16 bits), dm float32 [R', J, paths, 8], pay float32
    [R', paths] with R' = max(1, replicates)."""
    R = max(1, replicates)
    feat = np.zeros((R, hedge_dates, paths, 16), np.uint16)
    dm = np.zeros((R, hedge_dates, paths, 8), np.float32)
    pay = np.zeros((R, paths), np.float32)
    _check(lib().thoth_cv_samples(C.byref(built.req), hedge_dates, paths, sobol_skip, replicates, rqmc_seed,
                                  feat.ctypes.data_as(C.POINTER(C.c_uint16)), dm.ctypes.data_as(C.POINTER(C.c_float)),
                                  pay.ctypes.data_as(C.POINTER(C.c_float))))
    return feat, dm, pay


def cv_loss_grad(weights, feat, dm, pay, activation=abi.CV_SILU):
    """Device loss and gradient for host samples: feat uint16 [J, B, 16] (BF16
    bits), dm float32 [J, B, 8], pay float32 [B]; returns (loss, grad[10448])."""
    feat = np.ascontiguousarray(feat, np.uint16)
    dm = np.ascontiguousarray(dm, np.float32)
    pay = np.ascontiguousarray(pay, np.float32)
    J, B, _ = feat.shape
    assert feat.shape == (J, B, 16) and dm.shape == (J, B, 8) and pay.shape == (B,)
    loss = C.c_double()
    grad = np.zeros(abi.CV_WEIGHTS, np.float32)
    _check(lib().thoth_cv_loss_grad(C.byref(weights), feat.ctypes.data_as(C.POINTER(C.c_uint16)),
                                    dm.ctypes.data_as(C.POINTER(C.c_float)), pay.ctypes.data_as(C.POINTER(C.c_float)),
                                    B, J, activation, C.byref(loss), grad.ctypes.data_as(C.POINTER(C.c_float))))
    return loss.value, grad


def cv_weights_from_array(a):
    w = abi.CvWeights()
    np.ctypeslib.as_array(w.w)[:] = np.asarray(a, np.float32)
    return w


def cv_weights_array(w):
    return np.ctypeslib.as_array(w.w).copy()


# ---- vanilla grids (include/thoth_mcl_grid.h) ------------------------------

def vanilla_grid(underlying, maturity_days, strikes, absolute=False):
    g = abi.VanillaGrid()
    g.underlying = underlying
    g.n_maturities = len(maturity_days)
    for j, d in enumerate(maturity_days):
        g.maturity_days[j] = int(d)
    g.n_strikes = len(strikes)
    for k, x in enumerate(strikes):
        g.strikes[k] = float(x)
    g.is_absolute_strike = 1 if absolute else 0
    return g


class GridPlan:
    """All books (request x bump scenario) of a set of vanilla grids, swept by
    one persistent launch replayed from a hipGraph (thoth_grid_*)."""

    def __init__(self, builts, grids, autotune=True, group=True):
        """group=False keeps one device book per book (no shared normals)."""
        self.builts = list(builts)
        self.grids = list(grids)
        reqs = (abi.Request * len(self.builts))(*[b.req for b in self.builts])
        gr = (abi.VanillaGrid * len(self.grids))(*self.grids)
        self._keep = (reqs, gr)
        self.handle = C.c_void_p()
        flags = (abi.GRID_AUTOTUNE if autotune else 0) | (0 if group else abi.GRID_SEPARATE_BOOKS)
        _check(lib().thoth_grid_prepare_ex(len(self.builts), reqs, gr, flags, C.byref(self.handle)))

    def run(self):
        """Returns (list of dicts of [n_maturities, n_strikes] arrays, device ms)."""
        out, res = [], (abi.GridResult * len(self.grids))()
        for g, r in zip(self.grids, res):
            shape = (g.n_maturities, g.n_strikes)
            d = {k: np.zeros(shape) for k in ("call", "put", "call_trust", "put_trust", "call_delta", "put_delta",
                                               "call_vega", "put_vega")}
            for k, a in d.items():
                setattr(r, k, a.ctypes.data_as(C.POINTER(C.c_double)))
            out.append(d)
        ms = C.c_double()
        _check(lib().thoth_grid_run(self.handle, res, C.byref(ms)))
        return out, ms.value

    def tuning(self):
        t = abi.GridTuning()
        _check(lib().thoth_grid_get_tuning(self.handle, C.byref(t)))
        return t

    def close(self):
        if self.handle:
            lib().thoth_grid_release(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
