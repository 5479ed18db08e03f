/*
 * thoth_mcl.h -- C-ABI of the book-level Monte-Carlo engine (host side).
 *
 * A request describes one `!mcl_pricer` book in the reference's own units:
 * the `!mcl_configuration` fields, the equities with their bs / heston / lsv
 * volatility (vols and rates quoted in percent, as the YAML quotes them), the
 * spot/spot correlation matrix, the yield curve, the rainbow underlyings and the
 * vanilla / barrier contracts.  `thoth_mcl_price` runs the whole pricing
 * (diffusion grid, Cholesky, LSV leverage calibration, device path sweep,
 * bump-and-revalue Greeks) and fills caller-owned results.
 *
 * It replaces, for GPU-eligible books, the body of PricerMCL::PriceBook +
 * PricerMCL::ComputeGreeks (reference: pricer/src/tasks/pricer_mcl.cpp:197-240,
 * pricer/src/tasks/pricer_mcl_greeks.cpp:114-299, pricer/src/tasks/pricer.cpp:273-391).
 * Field semantics follow the reference objects cited on each struct.
 *
 * Errors: every entry point returns a THOTH_* status; thoth_last_error() holds
 * the message of the last failure on the calling thread.  No CPU fallback is
 * taken: without a usable device the call returns THOTH_ERR_UNAVAILABLE and the
 * caller keeps its own CPU engine (the reference gate's contract,
 * pricer/src/tasks/pricer_mcl_gpu_gate.cpp:82-105).
 */
#ifndef THOTH_MCL_H
#define THOTH_MCL_H

#include <stddef.h>
#include <stdint.h>

#include "thoth_mcl_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define THOTH_MAX_PILLARS 16

/* volatility kinds (reference: vol_correl/{bs,heston,lsv}_volatility.hpp) */
enum { THOTH_VOL_BS = 0, THOTH_VOL_HESTON = 1, THOTH_VOL_LSV = 2 };
/* contract kinds (reference: contracts/{vanilla,barrier}.hpp) */
enum { THOTH_VANILLA = 0, THOTH_BARRIER = 1 };
/* barrier types (reference: helpers/constants.hpp BARRIER_TYPE_*) */
enum { THOTH_UP_OUT = 0, THOTH_UP_IN = 1, THOTH_DOWN_OUT = 2, THOTH_DOWN_IN = 3 };
/* Greek request bits (reference: Pricer::_request_*) */
enum { THOTH_GREEK_DELTA = 1, THOTH_GREEK_GAMMA = 2, THOTH_GREEK_VEGA = 4, THOTH_GREEK_RHO = 8 };
/* arithmetic of the device path sweep */
enum { THOTH_FP64 = 0, THOTH_FP32 = 1 };

/* !mcl_configuration (reference: tasks/mcl_configuration.hpp:11-42, .cpp:15-47) */
typedef struct thoth_mcl_configuration {
    int64_t paths;         /* > 1 and <= 2e8 */
    int32_t max_day_step;  /* > 0 */
    int32_t min_day_step;
    double vol_year_step;  /* caps the step for stochastic-vol books (pricer_mcl.cpp:71-117) */
    int32_t use_sobol;     /* the device engine supports the Sobol + bridge mode only */
    int32_t seed;
    int64_t sobol_skip;    /* leading Sobol points discarded (disjoint blocks) */
    int32_t allow_gpu;
} thoth_mcl_configuration;

/* !sabr_volatility (reference: vol_correl/sabr_volatility.cpp:27-35) */
typedef struct thoth_sabr_surface {
    int32_t n; /* pillars */
    double maturities[THOTH_MAX_PILLARS];
    double alpha[THOTH_MAX_PILLARS];
    double beta[THOTH_MAX_PILLARS];
    double rho[THOTH_MAX_PILLARS];
    double nu[THOTH_MAX_PILLARS];
} thoth_sabr_surface;

/* !equity + its volatility object (reference: marketdata/equity.cpp, vol_correl/ sources) */
typedef struct thoth_asset {
    double spot;
    int32_t vol_kind;       /* THOTH_VOL_* */
    double volatility;      /* bs: percent (bs_volatility.cpp:24) */
    double init_vol;        /* heston/lsv: percent, v0 = (init_vol/100)^2 (heston_volatility.cpp:38) */
    double long_vol;        /* heston/lsv: percent, theta = (long_vol/100)^2 */
    double kappa;
    double vol_of_vol;
    double rho_spot_var;    /* correlation(<name>, <name>_var) (pricer.cpp:678-687) */
    double dividend_yield;  /* flat continuous dividend curve, percent (0 = none) */
    thoth_sabr_surface surface; /* lsv target implied surface */
} thoth_asset;

/* !yield_curve: pillars in days from today, zero rates in percent (curve.cpp:45-73) */
typedef struct thoth_curve {
    int32_t n;
    int32_t days[THOTH_MAX_PILLARS];
    double values[THOTH_MAX_PILLARS];
} thoth_curve;

/* !rainbow over equities (reference: underlyings/rainbow.cpp:46-61, basket.cpp:31-38);
 * the rebasing reference spots are the members' spots at load. */
typedef struct thoth_rainbow {
    int32_t best; /* 1 best_of (max), 0 worst_of (min) */
    int32_t n;
    int32_t members[THOTH_MCL_MAX_ASSETS];
} thoth_rainbow;

/* !vanilla / !barrier (reference: contracts/vanilla.cpp, contracts/barrier.cpp) */
typedef struct thoth_contract {
    int32_t kind;                 /* THOTH_VANILLA / THOTH_BARRIER */
    int32_t is_call;
    int32_t underlying;           /* equity index >= 0, or -(1 + rainbow index) */
    int32_t maturity_days;        /* days from today */
    double strike;
    int32_t is_absolute_strike;   /* 0: strike is a percent of the underlying spot */
    int32_t barrier_type;         /* THOTH_UP_OUT .. THOTH_DOWN_IN */
    int32_t discrete_monitoring;  /* 1 discrete_monitoring, 0 continuous_monitoring */
    int32_t monitoring_period_days;
    double barrier_up_level;
    double barrier_down_level;
    /* exercise: 0 european, 1 american (vanilla on an equity: Longstaff-Schwartz
     * on the recorded spot paths, pricer_mcl_american.cpp:137-447) */
    int32_t american;
} thoth_contract;

typedef struct thoth_mcl_request {
    thoth_mcl_configuration mcl;
    int32_t n_assets;              /* equities in name order (the reference's factor order) */
    const thoth_asset* assets;
    const double* correlation;     /* n_assets x n_assets spot/spot matrix, row-major */
    thoth_curve rate;
    int32_t n_rainbows;
    const thoth_rainbow* rainbows;
    int32_t n_contracts;           /* in the book's contract order */
    const thoth_contract* contracts;
    int32_t request_greeks;        /* THOTH_GREEK_* bits */
    /* engine extensions (no reference counterpart) */
    int32_t precision;             /* THOTH_FP64 (oracle-matched) or THOTH_FP32 (production) */
    int32_t correlate_stochastic;  /* 1: Cholesky-correlate Heston/LSV spots too (the reference
                                      leaves them uncorrelated, pricer_mcl.cpp:280-284) */
    int32_t uniform_steps;         /* > 0: replace the calendar grid by uniform steps t_i = i*T/n */
    /* LSV leverage layout.  0, 0: one 201-point layer per diffusion date, each on
     * its own +-8 sigma sqrt(t) log-spot axis (single.cpp:175-181, the reference's).
     * spot_points > 1, time_layers > 0: a surface of min(time_layers, dates)
     * layers calibrated on diffusion dates spread evenly in sqrt(t), all on one
     * log-spot axis of spot_points points (+-4 sigma_ATM sqrt(T) at the last
     * date), read between layers linearly in time as LeverageSurface::GetLeverage
     * does (leverage_surface.cpp:43-66). */
    int32_t leverage_spot_points;
    int32_t leverage_time_layers;
    /* !currency discount_rate (reference: marketdata/currency.cpp:19-27, 129-142):
     * the OIS / collateral curve the contract flows are discounted on, while the
     * rate curve above drives the equity drifts; n = 0: the rate curve discounts
     * (single curve).  The rho bump shifts both curves (tasks/pricer.cpp:202-217). */
    thoth_curve discount;
    /* term-structured !correlation_matrix (vol_correl/correlation.cpp:35-130,
     * 528-620, 860-903): n_corr_pillars > 1 increasing pillar maturities (years)
     * with one n_assets x n_assets matrix each (corr_matrices[p * n * n + i * n + j]);
     * each entry is linear in time between pillars and flat outside, and the
     * increments of step i are correlated by the Cholesky factor of the matrix
     * averaged over [t_{i-1}, t_i].  0 or 1 pillar: the constant `correlation`. */
    int32_t n_corr_pillars;
    const double* corr_maturities;
    const double* corr_matrices;
    /* Scenario inputs (engine extensions, zero / NULL: the reference's base
     * scenario).  vol_shift: a parallel volatility shift of every equity, the
     * shift the reference's vega scenario applies (Volatility::ApplyShift of
     * RISK_FACTOR_VOL, tasks/pricer.cpp:325-333): Black-Scholes vol + shift,
     * Heston sqrt(v0) and sqrt(theta) + shift, LSV target implied vols + shift
     * (the leverage recalibrated); the vega bump adds its own 0.01 on top.
     * rebase_spots: n_assets rebasing spots S_i0 of the rainbow members
     * (rainbow.cpp:22-27 captures them at load, so a spot bump does not move
     * them); NULL: the equities' spots. */
    double vol_shift;
    const double* rebase_spots;
} thoth_mcl_request;

typedef struct thoth_mcl_result {
    double premium, premium_trust;
    double delta, gamma, vega, rho;
    /* per-contract arrays of n_contracts entries, caller-owned; any may be NULL */
    double* contract_premium;
    double* contract_trust;
    double* contract_delta;
    double* contract_gamma;
    double* contract_vega;
    double* contract_rho;
    int32_t n_dates;     /* diffusion dates incl. today */
    int32_t n_scenarios; /* books swept (base + bumps) */
    double sweep_ms;     /* device time of the path sweeps */
    double calib_ms;     /* leverage calibration time */
} thoth_mcl_result;

/* Price one book.  Returns THOTH_OK or an error status. */
int thoth_mcl_price(const thoth_mcl_request* req, thoth_mcl_result* res);

/* Split form of thoth_mcl_price for repeated sweeps of one book: prepare
 * resolves the grid, calibrates, packs the base and bump-scenario books and
 * uploads them once (device-resident); run sweeps them and fills res (same
 * fields as thoth_mcl_price); release frees the plan.  A plan is tied to the
 * calling process and its device. */
typedef struct thoth_mcl_plan_s* thoth_mcl_plan;
int thoth_mcl_prepare(const thoth_mcl_request* req, thoth_mcl_plan* plan);
int thoth_mcl_run(thoth_mcl_plan plan, thoth_mcl_result* res);
void thoth_mcl_release(thoth_mcl_plan plan);

/* No reference counterpart (the reference's GPU hook has no launch shape to
 * choose): the launch shape of a plan's sweep.  lev_group: time steps per LDS
 * barrier (leverage layers and Sobol ops staged per group); workgroups_per_cu:
 * 0 launches one workgroup per Sobol block of 256 paths, W > 0 a persistent
 * grid of W workgroups per CU pulling (book, block) items, -1 the equity-wave
 * form (FP32 books whose equities are all LSV with uncorrelated spots and
 * whose contracts are vanillas: the equities of 64 paths stepped by as many
 * waves, for pricings too small to fill the device; chosen by itself where
 * the other forms would leave the SIMDs under two waves each).  The forms
 * 0 and W > 0 give the same per-block partial sums, hence the same result
 * bit for bit, and are thoth_mcl_tune's candidates; the equity-wave form
 * sums the same path values in quarters of a block (the same to FP64
 * rounding of the sums) and is set by itself or by thoth_mcl_set_shape.
 * thoth_mcl_tune times every candidate (lev_group 1, 2, 4, 8 x the launch
 * forms that exist for the plan's books) twice on the plan's own books, keeps
 * the fastest and caches it by book shape for the process (candidates = 0:
 * the cached choice); thoth_mcl_set_shape forces one; thoth_mcl_get_shape
 * reports the current one (device_ms: the tuner's best time, 0 if untuned). */
typedef struct thoth_plan_shape {
    int32_t lev_group;
    int32_t workgroups_per_cu;
    double device_ms;
    int32_t candidates;
    double default_ms;   /* thoth_mcl_tune: time of the shape the plan had before (0 when cached) */
} thoth_plan_shape;
int thoth_mcl_tune(thoth_mcl_plan plan, thoth_plan_shape* shape);
int thoth_mcl_set_shape(thoth_mcl_plan plan, int32_t lev_group, int32_t workgroups_per_cu);
int thoth_mcl_get_shape(thoth_mcl_plan plan, thoth_plan_shape* shape);

/* Randomised QMC price of the base book of req (no Greeks): R = replicates
 * random digital shifts of Sobol points 0 .. req->mcl.paths - 1, each an
 * independent estimate; the 95 % interval is Student-t over the R replicate
 * means (the reference's trust is the iid formula, monte_carlo_node.cpp:35-40,
 * which is not a confidence interval under Sobol points). */
typedef struct thoth_rqmc_result {
    int32_t replicates;
    int64_t paths;       /* per replicate */
    double mean;         /* premium */
    double halfwidth;    /* 95 % confidence half-width of the mean */
    double var;          /* pooled per-path variance */
    double sweep_ms;     /* device time of the sweep */
    double setup_ms;     /* host clock: grid, leverage calibration, packing and upload */
    double run_wall_ms;  /* host clock: launch, sweep, reduction and interval, the book resident */
    double* replicate_means; /* optional caller-owned array of `replicates` entries (NULL: not written) */
} thoth_rqmc_result;
int thoth_mcl_price_rqmc(const thoth_mcl_request* req, int32_t replicates, uint64_t rqmc_seed, thoth_rqmc_result* res);

/* Pathwise Greeks (no reference counterpart; the reference computes bump
 * Greeks only, pricer_mcl_greeks.cpp:43-299) on the same Sobol points, of
 *  - a single-equity Black-Scholes or Heston book holding one vanilla or one
 *    knock-out barrier contract.  Knock-out books are simulated with one-step
 *    survival conditioning (each monitored step draws the spot conditioned on
 *    survival and carries the survival probability as a weight): an unbiased
 *    estimator of the same discretely monitored price whose pathwise
 *    derivative exists;
 *  - any other book of 1..8 Black-Scholes, Heston or LSV equities (correlated
 *    as the sweep correlates them) holding one european vanilla on an equity
 *    or on a best-of / worst-of rainbow.  LSV equities carry the tangent
 *    through the leverage lookup: its slope in log-spot, and the leverage
 *    function's own sensitivity to the spot and to the vol shift (the
 *    reference recalibrates the leverage in every bump scenario), taken as
 *    that recalibration at the reference's bump (+1 % spot, +1 vol point) over
 *    the bump (thoth_mcl_leverage_tangent): pathwise along the paths, the
 *    reference's finite bump for the particle calibration, whose
 *    infinitesimal derivative is dominated by its bin kinks.
 * delta = sum over equities of dP/dS0_a (the reference's book delta sums its
 * per-underlying bumps, pricer.cpp:287-323), each in asset_delta; vega per vol
 * point of the parallel shift of every equity in the reference's bump
 * parametrisation (Black-Scholes vol + 1pt; Heston sqrt(v0) and sqrt(theta)
 * + 0.01; LSV target implied vols + 1pt); gamma (THOTH_GREEK_GAMMA) = sum
 * over equities of (delta_a(S0_a (1 + 10%)) - delta_a(S0_a (1 - 10%))) /
 * (20% S0_a), the reference's gamma bump applied to the pathwise delta. */
typedef struct thoth_pathwise_result {
    double premium, premium_trust;   /* the survival-weighted estimator */
    double delta, delta_trust;       /* book delta: sum of the per-equity deltas */
    double gamma;                    /* book gamma: sum of the per-equity gammas */
    double vega, vega_trust;
    double sweep_ms;
    int32_t n_assets;
    double asset_delta[THOTH_MCL_MAX_ASSETS];        /* dP / dS0_a */
    double asset_delta_trust[THOTH_MCL_MAX_ASSETS];
    double asset_gamma[THOTH_MCL_MAX_ASSETS];        /* the +-10 % bump of equity a applied to its pathwise delta */
    int32_t n_books;                 /* books swept (base, and two per equity for gamma) */
    double setup_ms;                 /* host clock: calibrations and their tangents, packing, upload */
} thoth_pathwise_result;
int thoth_mcl_price_pathwise(const thoth_mcl_request* req, thoth_pathwise_result* res);

/* Diffusion grid of a request (pricer_mcl.cpp:71-117): writes up to cap dates
 * (days from today, year fractions t and step dt) and returns the date count,
 * or a negative THOTH_* status. */
int thoth_mcl_grid(const thoth_mcl_request* req, int32_t cap, double* days, double* t, double* dt);

/* FP64 LSV leverage calibration of equity `asset` on the request's grid, with
 * the equity spot replaced by `spot` (reference: Single::CalibrateLeverage,
 * marketdata/single.cpp:400-473).  Fills up to cap_layers calibrated layers of
 * P points, P = 201 for the per-date layout and leverage_spot_points for the
 * surface layout: layer time t[k] (years), ln_step[k], offset[k] and
 * levels[k*P + i] at log-spot (offset[k] + i) * ln_step[k].  Returns the layer
 * count or a negative THOTH_* status. */
int thoth_mcl_calibrate_leverage(const thoth_mcl_request* req, int32_t asset, double spot, int32_t cap_layers,
                                 double* t, double* ln_step, int64_t* offset, double* levels);

/* No reference counterpart (host only, no device): the sensitivity of an LSV
 * equity's calibrated leverage to its spot (which = 0, per unit of spot) or
 * to the parallel vol shift (which = 1, per unit of shift) that the pathwise
 * basket Greeks carry, on the request's diffusion dates: the leverage
 * recalibrated at spot + 1 % (helpers/constants.hpp:36) or at vol shift + 1
 * vol point (tasks/pricer.hpp:25), as the reference's bump scenarios
 * recalibrate it, read at the nodes of the base per-date layers (points and
 * offsets of thoth_mcl_calibrate_leverage's per-date layers) the way the
 * sweep reads a layer, minus the base values, over the bump:
 * dv[i * points + k] for node k of date i.  Writes at most cap_dates dates and
 * returns the date count, or a negative THOTH_* status. */
int thoth_mcl_leverage_tangent(const thoth_mcl_request* req, int32_t asset, int32_t which, int32_t cap_dates,
                               double* dv);

/* The cluster master's path split (reference: modes/run_cluster.cpp:308-339):
 * `total` paths cut into min(parts, total) disjoint Sobol blocks, the first
 * total % n one path longer, block k skipping the paths of the blocks before
 * it.  Fills paths[k] and skips[k] (room for `parts` entries) and returns the
 * block count n, or a negative THOTH_* status.  Host only. */
int thoth_mcl_split_paths(int64_t total, int32_t parts, int64_t* paths, int64_t* skips);

/* Pools the results of `parts` disjoint blocks as the cluster master does
 * (run_cluster.cpp:402-447): premiums, Greeks and per-contract values as
 * path-weighted means (weight paths[k] / total), trusts in quadrature
 * sqrt(sum_k w_k^2 trust_k^2); out's per-contract arrays (n_contracts entries,
 * any may be NULL) are filled from the parts' arrays.  Host only. */
int thoth_mcl_pool_results(int32_t parts, const int64_t* paths, const thoth_mcl_result* in, int32_t n_contracts,
                           thoth_mcl_result* out);

/* thoth_mcl_price of a book split into `parts` disjoint Sobol blocks, each
 * priced as the cluster's slave would price it (its path count, seed k and
 * sobol_skip), and pooled by thoth_mcl_pool_results: the single-device form
 * of the reference's path-split cluster. */
int thoth_mcl_price_split(const thoth_mcl_request* req, int32_t parts, thoth_mcl_result* res);

#ifdef __cplusplus
}
#endif
#endif
