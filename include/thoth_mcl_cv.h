/*
 * thoth_mcl_cv.h -- learned martingale control variate (no reference
 * counterpart: the reference engine, pricer/src/tasks/pricer_mcl*.cpp, has no
 * control variates).
 *
 * A small MLP f(t, log S, v) -> 5 hedge ratios is trained on paths the
 * engine itself generates (FP32 sweep of the book) and then used as the
 * control variate
 *     CV = sum_j f(x_j) . sum_{i in interval j} df(t_i) (100/S0) M_i,
 *     M_i = the first-order Doob increment of S_i (per equity) under the
 *           simulated scheme: S_{i-1} vol dW for a Black-Scholes step,
 *           S_{i-1} (A (v_i - E[v_i | past]) + b z) for the Andersen QE step
 *           (A its variance loading, b z its spot shock, z independent of v_i),
 * whose conditional mean given the path up to t_{i-1} is exactly zero.  CV
 * has zero mean for ANY weights, so P - CV is an unbiased estimator of the
 * premium P.  Networks: 11 -> 64 -> 64
 * -> 64 -> 5 with SiLU hidden units a(z) = z / (1 + e^-z) (THOTH_CV_SILU) or
 * smoothstep-gated linear units a(z) = z t^2 (3 - 2t), t = clamp(z/6 + 1/2, 0, 1)
 * (THOTH_CV_GATE: no transcendental), BF16 operands with FP32 accumulation on
 * the matrix cores, FP32 master weights, Adam.  Plain Monte-Carlo is reported beside every
 * control-variate estimate as the yardstick of bias and variance reduction.
 */
#ifndef THOTH_MCL_CV_H
#define THOTH_MCL_CV_H

#include <stdint.h>

#include "thoth_mcl.h"

#ifdef __cplusplus
extern "C" {
#endif

/* FP32 master weights: per layer W^T row-major, padded [64x16][64x64][64x64][16x64],
 * then the biases [64][64][64][16]. */
#define THOTH_CV_WEIGHTS 10448

typedef struct thoth_cv_weights {
    float w[THOTH_CV_WEIGHTS];
} thoth_cv_weights;

/* hidden activation of the hedge network */
enum { THOTH_CV_SILU = 0, THOTH_CV_GATE = 1 };

typedef struct thoth_cv_config {
    int32_t hedge_dates;       /* J: hedge ratios are re-evaluated at J dates (hedge_spacing) */
    int64_t train_paths;       /* training paths (Sobol points train_sobol_skip+1 ..) */
    int64_t train_sobol_skip;
    int32_t adam_steps;
    int32_t batch_paths;       /* paths per Adam step, multiple of 64 */
    float learning_rate;
    uint64_t init_seed;        /* weight initialisation */
    int32_t eval_mode;         /* THOTH_CV_IN_LOOP or THOTH_CV_TWO_PHASE (pricing only) */
    float hedge_spacing;       /* p in [1, 8]: hedge date j at step m - round(m (1 - j/J)^p), at least one
                                  step apart; 1 = uniform, 2 = uniform in sqrt(T - t) (denser toward
                                  maturity, where the discrete-hedging error concentrates) */
    int32_t activation;        /* THOTH_CV_SILU or THOTH_CV_GATE (training and pricing must agree) */
} thoth_cv_config;

/* pricing with a trained network: the hedge network evaluated on the matrix
 * cores inside the FP32 stepping loop at each hedge date (production), or
 * two phases -- the sweep stores features and increments, a separate
 * inference kernel reads them (same estimator, used as a cross-check) */
enum { THOTH_CV_IN_LOOP = 0, THOTH_CV_TWO_PHASE = 1 };

typedef struct thoth_cv_train_report {
    double loss_first;   /* mean (Y - mean P)^2 of the first minibatch (f = 0 => variance of P) */
    double loss_last;    /* the same for the last minibatch */
    double data_ms;      /* training-path sweep */
    double train_ms;     /* all Adam steps (one hipGraph replayed per step) */
    int32_t steps;
} thoth_cv_train_report;

typedef struct thoth_cv_result {
    int32_t replicates;    /* R random digital shifts */
    int64_t paths;         /* per replicate */
    double plain_mean, plain_halfwidth;   /* 95 % CI from the R replicate means */
    double cv_mean, cv_halfwidth;
    double plain_var, cv_var;             /* pooled per-path variances of P and of P - CV */
    double sweep_ms, eval_ms;             /* device time: FP32 sweep (with in-loop inference) + separate inference */
    /* paired difference plain - (P - CV), i.e. the mean of CV over the replicates, and the
     * 95 % Student-t half-width of that mean: zero mean is the unbiasedness of the estimator */
    double cv_term_mean, cv_term_halfwidth;
    /* optional caller-owned arrays of `replicates` entries (NULL: not written):
     * the replicate means of P and of P - CV */
    double* replicate_plain_means;
    double* replicate_cv_means;
    /* host clock: book setup (grid, leverage calibration, packing, network upload), and
     * the pricing itself with the book resident (launches, sweep, inference, reduction) */
    double setup_ms, run_wall_ms;
} thoth_cv_result;

/* Train the hedge network for the base book of req (FP32 sweep).  w is the
 * output (and is ignored on input). */
int thoth_cv_train(const thoth_mcl_request* req, const thoth_cv_config* cfg, thoth_cv_weights* w,
                   thoth_cv_train_report* rep);

/* Price the base book of req with R = replicates randomised-QMC replicates of
 * req->mcl.paths paths each, plain and with the control variate of weights w. */
int thoth_cv_price(const thoth_mcl_request* req, const thoth_cv_config* cfg, const thoth_cv_weights* w,
                   int32_t replicates, uint64_t rqmc_seed, thoth_cv_result* res);

/* The control-variate samples of the base book of req, FP32 sweep: with
 * replicates = 0, `paths` paths from Sobol points sobol_skip+1 ..; with
 * replicates R > 0, points 0 .. paths-1 under R random digital shifts.
 * feat [R'][J][paths][16] BF16 bits, dm [R'][J][paths][8], pay [R'][paths],
 * R' = max(1, R) (the layouts thoth_cv_loss_grad consumes for R' = 1); hedge
 * dates uniform, round(j m / J). */
int thoth_cv_samples(const thoth_mcl_request* req, int32_t hedge_dates, int64_t paths, int64_t sobol_skip,
                     int32_t replicates, uint64_t rqmc_seed, uint16_t* feat, float* dm, float* pay);

/* Loss mean_p (P_p - sum_j f(x_pj).dm_pj)^2 and its gradient (same layout as
 * thoth_cv_weights) for B paths of host data: feat [J][B][16] BF16, dm [J][B][8],
 * pay [B]; B multiple of 64; activation THOTH_CV_SILU or THOTH_CV_GATE.  The
 * device training step's forward/backward. */
int thoth_cv_loss_grad(const thoth_cv_weights* w, const uint16_t* feat, const float* dm, const float* pay, int32_t B,
                       int32_t J, int32_t activation, double* loss, float* grad);

#ifdef __cplusplus
}
#endif
#endif
