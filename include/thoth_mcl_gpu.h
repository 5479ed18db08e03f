/*
 * thoth_mcl_gpu.h -- C-ABI of the MI355X (gfx950) Monte-Carlo device library.
 *
 * This is the drop-in device boundary of the reference engine's GPU gate.  The
 * reference declares (C++ linkage, pricer/src/tasks/mcl_gpu.hpp:16-46):
 *     bool gpu::Available();
 *     std::string gpu::DeviceInfo();
 *     GbmResult gpu::PriceEuropeanGbm(double Forward, double Strike, double T,
 *                                     double Vol, double Df, bool IsCall,
 *                                     long Paths, unsigned long Seed);
 * and this library exports their C-ABI replacements (thoth_gpu_available,
 * thoth_gpu_device_info, thoth_gpu_price_european_gbm) plus the book-level
 * sweep thoth_mcl_sweep, which prices whole Sobol + Brownian-bridge books
 * (Black-Scholes, Heston-QE and LSV equities, vanilla / barrier / rainbow
 * payoffs) on the device.  All pointers are plain host pointers, all sizes
 * explicit; the library owns and frees every device buffer it uses.
 *
 * Threading: calls are serialised per process (one pricing thread, as the
 * reference server serialises pricing, pricer/src/modes/run_server.cpp:59,96).
 */
#ifndef THOTH_MCL_GPU_H
#define THOTH_MCL_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define THOTH_MCL_MAX_ASSETS 8
#define THOTH_MCL_MAX_CONTRACTS 32
#define THOTH_MCL_MAX_RAINBOWS 4

enum {
    THOTH_OK = 0,
    THOTH_ERR_UNAVAILABLE = 1, /* no usable device: the caller keeps its CPU engine */
    THOTH_ERR_INVALID = 2,     /* malformed input */
    THOTH_ERR_DEVICE = 3,      /* HIP runtime failure */
    THOTH_ERR_UNSUPPORTED = 4  /* book shape outside the device engine */
};

/* per-equity diffusion model */
enum { THOTH_MODEL_BS = 0, THOTH_MODEL_HESTON = 1, THOTH_MODEL_LSV = 2 };
/* payoff kinds */
enum { THOTH_PAYOFF_VANILLA = 0, THOTH_PAYOFF_BARRIER = 1 };

/* Replaces gpu::Available() (mcl_gpu.hpp:20): 1 iff a gfx950 device is usable. */
int thoth_gpu_available(void);

/* Replaces gpu::DeviceInfo() (mcl_gpu.hpp:24): one-line description written
 * into buf (NUL-terminated, truncated to len).  Returns THOTH_OK, or
 * THOTH_ERR_UNAVAILABLE with the reason in buf. */
int thoth_gpu_device_info(char* buf, size_t len);

/* Replaces gpu::PriceEuropeanGbm (mcl_gpu.hpp:43-44), same math contract
 * (mcl_gpu.hpp:26-38, mcl_gpu.cu:100-152):
 *   F_T = Forward * exp(-0.5 Vol^2 T + Vol sqrt(T) Z), payoff max(+-(F_T-K),0),
 *   premium = Df * mean, trust = Df * sqrt(max(E[x^2]-mean^2,0) / Paths).
 * Z comes from the Joe-Kuo Sobol sequence through Acklam's inverse normal
 * (point indices 1 + Seed*Paths .. (Seed+1)*Paths while they stay below
 * 2^32 - 1; for larger seeds, a negative seed wrapped to 2^64 - k included,
 * points 0 .. Paths-1 under one random digital shift seeded by Seed), so a
 * fixed Seed reproduces the same points and bumps share them.  Paths <= 0
 * returns premium = trust = 0 as the reference. */
int thoth_gpu_price_european_gbm(double forward, double strike, double t, double vol,
                                 double df, int is_call, long long paths, uint64_t seed,
                                 double* premium, double* trust);

/* Integer coordinates of Sobol points skip+1 .. skip+n in dimensions 0..dim-1
 * (out[i*dim + d], dim <= 4097), computed on the device by the same integer
 * stage the path sweep uses.  Equal bit for bit to the reference generator's
 * sequential Gray-code recurrence after Skip(skip) (sobol_generator.cpp:80-121).
 * No reference counterpart as an entry point; exposed for QMC users and tests. */
int thoth_gpu_sobol_points(int32_t dim, int64_t skip, int64_t n, uint32_t* out);

/* The FP32 special functions of the production (FP32) sweep evaluated on the
 * device for n inputs (no reference counterpart; exposed so that their error
 * bounds are checked on the hardware that runs them):
 *   fn 0: the inverse normal of the Sobol coordinate in[i] (u = in[i] 2^-32,
 *         u = 0 read as 2^-32), as the generic kernels evaluate it;
 *   fn 1: the same through the uniform-book kernels' split form (central
 *         polynomial for every lane, tail lanes redone);
 *   fn 2: erfc of the float whose bits are in[i] (the QE exponential branch).
 * n in [1, 2^28]; out[i] receives the result. */
int thoth_gpu_special_f32(int32_t fn, int64_t n, const uint32_t* in, float* out);

/* Message of the last failure on the calling thread ("" if none). */
const char* thoth_last_error(void);

/* ------------------------------------------------------------------------ */
/* Book-level sweep.                                                         */
/*                                                                           */
/* One thoth_mcl_book is one fully resolved `!mcl_pricer` tree (base or one  */
/* Greek bump scenario): diffusion grid, equities, correlation, leverage     */
/* layers, payoffs.  It mirrors the node graph the reference builds in       */
/* PricerMCL::Tree_Init (pricer_mcl.cpp:498-581) once fused:                 */
/*  - factors: the spot noises of the equities in order, then the variance   */
/*    noises of the stochastic ones (pricer_mcl.cpp:404-432); Sobol          */
/*    dimension = bridge step * factors + factor (path_generator.cpp:135-157)*/
/*  - BS equities: correlated Brownian + log-Euler step (spot_diffusion_node)*/
/*  - Heston / LSV equities: Andersen QE variance + leveraged Andersen spot   */
/*    (heston_variance_node.cpp:17-57, heston_spot_node.cpp:32-94,           */
/*    local_volatility_node.cpp:19-47)                                        */
/*  - payoffs: vanilla / barrier flows on an equity or a rainbow, discounted */
/*    to today and summed into the book (contract_node.cpp, book_node.cpp).  */
/* ------------------------------------------------------------------------ */
typedef struct thoth_mcl_book {
    /* diffusion grid, date 0 = today */
    int32_t n_dates;
    const double* t;           /* [n_dates] year fraction today -> date i */
    const double* dt;          /* [n_dates] year fraction date i-1 -> date i (dt[0] unused) */
    /* equities */
    int32_t n_assets;
    const int32_t* model;      /* [n_assets] THOTH_MODEL_* */
    const double* spot;        /* [n_assets] diffusion spot at today */
    const double* vol;         /* [n_assets] BS constant vol-node value */
    const double* v0;          /* [n_assets] Heston/LSV parameters as the nodes read them */
    const double* kappa;
    const double* theta;
    const double* xi;
    const double* rho;
    const double* drift;       /* [n_assets * n_dates] drift-node value (r - q) at each date */
    const double* chol;        /* [n_assets * n_assets] lower Cholesky factor (used when n_assets > 1) */
    int32_t correlate_stochastic; /* 1: Heston/LSV spot noises are correlated too */
    /* LSV leverage (reference layout: one log-spot layer per date, single.cpp:400-473) */
    int32_t lev_points;        /* points per layer */
    const double* lev_ln_step; /* [n_assets * n_dates] */
    const int64_t* lev_offset; /* [n_assets * n_dates] */
    const double* lev_values;  /* [n_assets * n_dates * lev_points] */
    /* discounting: premium-curve zero rate at each date (yield_curve_node.cpp:18-38) */
    const double* disc_rate;   /* [n_dates] */
    /* rainbow underlyings: value = best/worst of weight_j * S_member_j */
    int32_t n_rainbows;
    const int32_t* rb_best;    /* [n_rainbows] */
    const int32_t* rb_n;       /* [n_rainbows] */
    const int32_t* rb_member;  /* [n_rainbows * THOTH_MCL_MAX_ASSETS] */
    const double* rb_weight;   /* [n_rainbows * THOTH_MCL_MAX_ASSETS] = 100 / reference spot */
    /* contracts (one flow each, at c_flow_index) */
    int32_t n_contracts;
    const int32_t* c_kind;       /* THOTH_PAYOFF_* */
    const int32_t* c_is_call;
    const int32_t* c_underlying; /* equity index, or -(1 + rainbow index) */
    const int32_t* c_flow_index; /* date index of the flow */
    const double* c_strike;
    const int32_t* c_is_up;      /* barrier flavour */
    const int32_t* c_is_in;
    const double* c_level;       /* barrier level (continuity-corrected when continuous) */
    const int32_t* c_mon_offset; /* [n_contracts + 1] CSR into c_mon_index */
    const int32_t* c_mon_index;  /* monitored date indices */
    /* optional [n_dates * n_assets * n_assets]: the lower Cholesky factor that
     * correlates the increments into date i (term-structured correlation,
     * correlation.cpp:860-903; date 0 unused); NULL: `chol` at every date */
    const double* chol_dates;
} thoth_mcl_book;

typedef struct thoth_mcl_sweep_cfg {
    int64_t paths;        /* paths per book (per replicate in RQMC mode) */
    int64_t sobol_skip;   /* leading Sobol points discarded (reference mode) */
    int32_t precision;    /* 0 = FP64 (oracle-matched), 1 = FP32 stepping, FP64 accumulation */
    int32_t block;        /* threads per block (0 = library default) */
    /* Randomised QMC (no reference counterpart; the reference's trust is the
     * iid formula, monte_carlo_node.cpp:35-40, which is not a valid Sobol CI).
     * replicates = 0: reference mode, points sobol_skip+1 .. sobol_skip+paths.
     * replicates = R > 0: R independent random digital shifts (seeded by
     * rqmc_seed) of points 0 .. paths-1; out then holds n_books * R entries,
     * out[b * R + r]. */
    int32_t replicates;
    uint64_t rqmc_seed;
    /* Pseudo-random draws from one xoshiro256++ stream seeded by `seed`
     * (splitmix64, rng.hpp:36-47; the reference seeds it once per PriceBook,
     * pricer_mcl.cpp:227), path p from position p T (T draws per path):
     * pseudo_random = 0 (use_sobol true): Sobol points with the Brownian
     * bridge, and the dimensions past the 4097 of the Sobol table from the
     * stream in dimension order (path_generator.cpp:140-157);
     * pseudo_random = 1 (use_sobol false): every normal from the stream in
     * (step, factor) order, no bridge (the reference's noise nodes draw in the
     * order of its node schedule, noise_node.cpp:20-25: same law, another
     * assignment of draws to steps).  Not with replicates > 0. */
    int32_t pseudo_random;
    uint64_t seed;
} thoth_mcl_sweep_cfg;

/* Kahan-compensated FP64 sums over paths of the discounted book value and of
 * every contract value (the reference indicators, monte_carlo_node.cpp:73-96). */
typedef struct thoth_mcl_book_out {
    double book_sum, book_sum2;
    double contract_sum[THOTH_MCL_MAX_CONTRACTS];
    double contract_sum2[THOTH_MCL_MAX_CONTRACTS];
    int64_t count;
    double device_ms;     /* kernel time of this book's sweep */
} thoth_mcl_book_out;

/* Sweep n_books books over the same Sobol points (common random numbers).
 * out holds n_books entries (n_books * replicates in RQMC mode). */
int thoth_mcl_sweep(int32_t n_books, const thoth_mcl_book* books,
                    const thoth_mcl_sweep_cfg* cfg, thoth_mcl_book_out* out);

/* No reference counterpart: the xoshiro256++ state after `steps` draws of the
 * stream seeded by `seed`, by the jump matrices of the device (steps a
 * multiple of draws_per_path, the T of a book), so that the jump-ahead can be
 * checked against sequential stepping on the host. */
int thoth_prng_state_after(uint64_t seed, uint64_t steps, uint64_t draws_per_path, uint64_t* state);

/* No reference counterpart: thoth_mcl_sweep for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments and the same results bit for bit;
 * every device buffer the sweep needs (packed books, digital shifts, per-block
 * partial sums) lives in one library-owned static device array of
 * THOTH_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.  A batch
 * that needs more is refused with THOTH_ERR_UNSUPPORTED before any device
 * work.  The device is set up (stream, Sobol tables, the array's address) by
 * the first call of any entry point of the library, thoth_gpu_available()
 * included: make that call outside a capture.  The library serialises its
 * calls; calls that overlapped on one GPU (from several threads bypassing the
 * library's lock, or captured graphs replayed concurrently) would race on the
 * one array, which keeps the last call's contents. */
#define THOTH_NOALLOC_SCRATCH_BYTES (64u << 20)
int thoth_mcl_sweep_noalloc(int32_t n_books, const thoth_mcl_book* books,
                            const thoth_mcl_sweep_cfg* cfg, thoth_mcl_book_out* out);

#ifdef __cplusplus
}
#endif
#endif
