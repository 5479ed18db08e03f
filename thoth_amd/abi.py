"""ctypes mirror of the C-ABI structs declared in include/thoth_mcl.h and
include/thoth_mcl_gpu.h (same field order and types).

This is the binding a Python caller (and the test-suite) uses; the reference
engine's own C++ caller includes the headers directly.
"""
import ctypes as C

MAX_ASSETS = 8
MAX_CONTRACTS = 32
MAX_RAINBOWS = 4
MAX_PILLARS = 16

# status codes (thoth_mcl_gpu.h)
OK, ERR_UNAVAILABLE, ERR_INVALID, ERR_DEVICE, ERR_UNSUPPORTED = 0, 1, 2, 3, 4
# thoth_mcl.h enums
VOL_BS, VOL_HESTON, VOL_LSV = 0, 1, 2
VANILLA, BARRIER = 0, 1
UP_OUT, UP_IN, DOWN_OUT, DOWN_IN = 0, 1, 2, 3
GREEK_DELTA, GREEK_GAMMA, GREEK_VEGA, GREEK_RHO = 1, 2, 4, 8
FP64, FP32 = 0, 1
# thoth_mcl_gpu.h enums
MODEL_BS, MODEL_HESTON, MODEL_LSV = 0, 1, 2
PAYOFF_VANILLA, PAYOFF_BARRIER = 0, 1

_d = C.c_double
_i32 = C.c_int32
_i64 = C.c_int64
_pd = C.POINTER(C.c_double)
_pi32 = C.POINTER(C.c_int32)
_pi64 = C.POINTER(C.c_int64)


class McConfiguration(C.Structure):
    _fields_ = [("paths", _i64), ("max_day_step", _i32), ("min_day_step", _i32),
                ("vol_year_step", _d), ("use_sobol", _i32), ("seed", _i32),
                ("sobol_skip", _i64), ("allow_gpu", _i32)]


class SabrSurface(C.Structure):
    _fields_ = [("n", _i32), ("maturities", _d * MAX_PILLARS), ("alpha", _d * MAX_PILLARS),
                ("beta", _d * MAX_PILLARS), ("rho", _d * MAX_PILLARS), ("nu", _d * MAX_PILLARS)]


class Asset(C.Structure):
    _fields_ = [("spot", _d), ("vol_kind", _i32), ("volatility", _d), ("init_vol", _d),
                ("long_vol", _d), ("kappa", _d), ("vol_of_vol", _d), ("rho_spot_var", _d),
                ("dividend_yield", _d), ("surface", SabrSurface)]


class Curve(C.Structure):
    _fields_ = [("n", _i32), ("days", _i32 * MAX_PILLARS), ("values", _d * MAX_PILLARS)]


class Rainbow(C.Structure):
    _fields_ = [("best", _i32), ("n", _i32), ("members", _i32 * MAX_ASSETS)]


class Contract(C.Structure):
    _fields_ = [("kind", _i32), ("is_call", _i32), ("underlying", _i32), ("maturity_days", _i32),
                ("strike", _d), ("is_absolute_strike", _i32), ("barrier_type", _i32),
                ("discrete_monitoring", _i32), ("monitoring_period_days", _i32),
                ("barrier_up_level", _d), ("barrier_down_level", _d), ("american", _i32)]


class Request(C.Structure):
    _fields_ = [("mcl", McConfiguration), ("n_assets", _i32), ("assets", C.POINTER(Asset)),
                ("correlation", _pd), ("rate", Curve), ("n_rainbows", _i32),
                ("rainbows", C.POINTER(Rainbow)), ("n_contracts", _i32),
                ("contracts", C.POINTER(Contract)), ("request_greeks", _i32),
                ("precision", _i32), ("correlate_stochastic", _i32), ("uniform_steps", _i32),
                ("leverage_spot_points", _i32), ("leverage_time_layers", _i32),
                ("discount", Curve), ("n_corr_pillars", _i32), ("corr_maturities", _pd),
                ("corr_matrices", _pd), ("vol_shift", _d), ("rebase_spots", _pd)]


class Result(C.Structure):
    _fields_ = [("premium", _d), ("premium_trust", _d), ("delta", _d), ("gamma", _d),
                ("vega", _d), ("rho", _d), ("contract_premium", _pd), ("contract_trust", _pd),
                ("contract_delta", _pd), ("contract_gamma", _pd), ("contract_vega", _pd),
                ("contract_rho", _pd), ("n_dates", _i32), ("n_scenarios", _i32),
                ("sweep_ms", _d), ("calib_ms", _d)]


class Book(C.Structure):
    _fields_ = [("n_dates", _i32), ("t", _pd), ("dt", _pd),
                ("n_assets", _i32), ("model", _pi32), ("spot", _pd), ("vol", _pd),
                ("v0", _pd), ("kappa", _pd), ("theta", _pd), ("xi", _pd), ("rho", _pd),
                ("drift", _pd), ("chol", _pd), ("correlate_stochastic", _i32),
                ("lev_points", _i32), ("lev_ln_step", _pd), ("lev_offset", _pi64),
                ("lev_values", _pd), ("disc_rate", _pd),
                ("n_rainbows", _i32), ("rb_best", _pi32), ("rb_n", _pi32), ("rb_member", _pi32),
                ("rb_weight", _pd),
                ("n_contracts", _i32), ("c_kind", _pi32), ("c_is_call", _pi32),
                ("c_underlying", _pi32), ("c_flow_index", _pi32), ("c_strike", _pd),
                ("c_is_up", _pi32), ("c_is_in", _pi32), ("c_level", _pd),
                ("c_mon_offset", _pi32), ("c_mon_index", _pi32), ("chol_dates", _pd)]


class PlanShape(C.Structure):
    _fields_ = [("lev_group", _i32), ("workgroups_per_cu", _i32), ("device_ms", _d), ("candidates", _i32),
                ("default_ms", _d)]


class SweepCfg(C.Structure):
    _fields_ = [("paths", _i64), ("sobol_skip", _i64), ("precision", _i32), ("block", _i32),
                ("replicates", _i32), ("rqmc_seed", C.c_uint64), ("pseudo_random", _i32), ("seed", C.c_uint64)]


class BookOut(C.Structure):
    _fields_ = [("book_sum", _d), ("book_sum2", _d), ("contract_sum", _d * MAX_CONTRACTS),
                ("contract_sum2", _d * MAX_CONTRACTS), ("count", _i64), ("device_ms", _d)]


# include/thoth_mcl_cv.h
CV_WEIGHTS = 10448


class CvWeights(C.Structure):
    _fields_ = [("w", C.c_float * CV_WEIGHTS)]


class CvConfig(C.Structure):
    _fields_ = [("hedge_dates", _i32), ("train_paths", _i64), ("train_sobol_skip", _i64),
                ("adam_steps", _i32), ("batch_paths", _i32), ("learning_rate", C.c_float),
                ("init_seed", C.c_uint64), ("eval_mode", _i32), ("hedge_spacing", C.c_float),
                ("activation", _i32)]


CV_IN_LOOP, CV_TWO_PHASE = 0, 1
CV_SILU, CV_GATE = 0, 1


class CvTrainReport(C.Structure):
    _fields_ = [("loss_first", _d), ("loss_last", _d), ("data_ms", _d), ("train_ms", _d), ("steps", _i32)]


class CvResult(C.Structure):
    _fields_ = [("replicates", _i32), ("paths", _i64), ("plain_mean", _d), ("plain_halfwidth", _d),
                ("cv_mean", _d), ("cv_halfwidth", _d), ("plain_var", _d), ("cv_var", _d),
                ("sweep_ms", _d), ("eval_ms", _d), ("cv_term_mean", _d), ("cv_term_halfwidth", _d),
                ("replicate_plain_means", C.POINTER(_d)), ("replicate_cv_means", C.POINTER(_d)),
                ("setup_ms", _d), ("run_wall_ms", _d)]


class RqmcResult(C.Structure):
    _fields_ = [("replicates", _i32), ("paths", _i64), ("mean", _d), ("halfwidth", _d), ("var", _d),
                ("sweep_ms", _d), ("setup_ms", _d), ("run_wall_ms", _d), ("replicate_means", C.POINTER(_d))]


# include/thoth_mcl_grid.h
GRID_MAX_MATURITIES = 16
GRID_MAX_STRIKES = 64


class VanillaGrid(C.Structure):
    _fields_ = [("underlying", _i32), ("n_maturities", _i32), ("maturity_days", _i32 * GRID_MAX_MATURITIES),
                ("n_strikes", _i32), ("strikes", _d * GRID_MAX_STRIKES), ("is_absolute_strike", _i32)]


class GridResult(C.Structure):
    _fields_ = [("call", _pd), ("put", _pd), ("call_trust", _pd), ("put_trust", _pd), ("call_delta", _pd),
                ("put_delta", _pd), ("call_vega", _pd), ("put_vega", _pd)]


GRID_AUTOTUNE, GRID_SEPARATE_BOOKS = 1, 2


class GridTuning(C.Structure):
    _fields_ = [("blocks_per_cu", _i32), ("lev_group", _i32), ("device_ms", _d), ("candidates", _i32),
                ("books_per_lane", _i32)]


class PathwiseResult(C.Structure):
    _fields_ = [("premium", _d), ("premium_trust", _d), ("delta", _d), ("delta_trust", _d), ("gamma", _d),
                ("vega", _d), ("vega_trust", _d), ("sweep_ms", _d), ("n_assets", _i32),
                ("asset_delta", _d * MAX_ASSETS), ("asset_delta_trust", _d * MAX_ASSETS),
                ("asset_gamma", _d * MAX_ASSETS), ("n_books", _i32), ("setup_ms", _d)]
