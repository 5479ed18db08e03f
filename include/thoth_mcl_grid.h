/*
 * thoth_mcl_grid.h -- whole vanilla grids on the device: calls and puts over
 * a strike x maturity grid per underlying, many books (underlyings and their
 * bump scenarios) swept by ONE persistent launch replayed from a hipGraph.
 *
 * Reference semantics: every cell is a `!vanilla` of the book priced by the
 * mcl engine (contracts/vanilla.cpp, vanilla_flow_node.cpp, payoff
 * finance.cpp:9-32), premium = df(T) mean(payoff), trust as
 * monte_carlo_node.cpp:35-40, and the per-contract bump Greeks
 * (pricer_mcl_greeks.cpp:177-209): delta = (P(S (1 + 1%)) - P(S)) / (1% S),
 * vega = P(vol + 1pt) - P per point; LSV bumps recalibrate the leverage
 * (single.cpp:404, lsv_volatility.hpp:45-55).  The reference prices such a
 * grid as 2 x strikes x maturities separate contracts; here a book carries
 * the grid as one payoff-grid observation at each maturity.
 */
#ifndef THOTH_MCL_GRID_H
#define THOTH_MCL_GRID_H

#include <stdint.h>

#include "thoth_mcl.h"

#ifdef __cplusplus
extern "C" {
#endif

#define THOTH_GRID_MAX_MATURITIES 16
#define THOTH_GRID_MAX_STRIKES 64

typedef struct thoth_vanilla_grid {
    int32_t underlying;         /* equity index of the request */
    int32_t n_maturities;
    int32_t maturity_days[THOTH_GRID_MAX_MATURITIES]; /* increasing, days from today */
    int32_t n_strikes;
    double strikes[THOTH_GRID_MAX_STRIKES];
    int32_t is_absolute_strike; /* 0: strikes in percent of the underlying's spot */
} thoth_vanilla_grid;

/* per book, caller-owned arrays [n_maturities][n_strikes]; any may be NULL */
typedef struct thoth_grid_result {
    double* call;
    double* put;
    double* call_trust;
    double* put_trust;
    double* call_delta;   /* requires THOTH_GREEK_DELTA in the request */
    double* put_delta;
    double* call_vega;    /* requires THOTH_GREEK_VEGA */
    double* put_vega;
} thoth_grid_result;

/* Launch shape picked by the autotuner (or forced): persistent workgroups
 * per CU, and leverage layers staged in LDS per barrier (the timestep tile). */
typedef struct thoth_grid_tuning {
    int32_t blocks_per_cu;
    int32_t lev_group;
    double device_ms;     /* best time measured while tuning */
    int32_t candidates;   /* configurations timed */
    int32_t books_per_lane; /* books stepped together by one lane on shared normals (1: none) */
} thoth_grid_tuning;

typedef struct thoth_grid_plan_s* thoth_grid_plan;

/* Prepare n books (request r with grid r; requests share paths, sobol_skip
 * and precision): grids, Cholesky, leverage calibration of every scenario,
 * device packing, autotuning (autotune != 0) and the hipGraph. */
int thoth_grid_prepare(int32_t n, const thoth_mcl_request* reqs, const thoth_vanilla_grid* grids, int32_t autotune,
                       thoth_grid_plan* plan);
/* No reference counterpart: thoth_grid_prepare with flags.  When every book
 * (request x bump scenario) is a single-equity LSV book on the same grid
 * (FP32), the plan steps 5 (else 2) of them per lane as the equities of one
 * device book: they read the same Sobol points, so one set of normals and one
 * Brownian bridge serve all of them and each book's results are those of its
 * own sweep.  THOTH_GRID_SEPARATE_BOOKS keeps one device book per book. */
#define THOTH_GRID_AUTOTUNE 1
#define THOTH_GRID_SEPARATE_BOOKS 2
int thoth_grid_prepare_ex(int32_t n, const thoth_mcl_request* reqs, const thoth_vanilla_grid* grids, int32_t flags,
                          thoth_grid_plan* plan);
/* Replay the graph; results[n]; device_ms: the replay's device time. */
int thoth_grid_run(thoth_grid_plan plan, thoth_grid_result* results, double* device_ms);
int thoth_grid_get_tuning(thoth_grid_plan plan, thoth_grid_tuning* tuning);
void thoth_grid_release(thoth_grid_plan plan);

#ifdef __cplusplus
}
#endif
#endif
