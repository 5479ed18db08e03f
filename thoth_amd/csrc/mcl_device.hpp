// mcl_device.hpp -- device-side book layout shared by the host launcher
// (mcl_gpu.cpp) and the path-sweep kernels (mcl_sweep.hpp).
#pragma once
#include <cstddef>
#include <cstdint>

namespace thoth
{

constexpr int kMaxAssets = 8;
constexpr int kMaxContracts = 32;
constexpr int kMaxRainbows = 4;
constexpr int kWavesPerBlock = 4;          // 256-thread blocks
constexpr int kBlock = 64 * kWavesPerBlock;
// static LDS of the sweep kernels (per-wave sums, persistent work item) with slack
constexpr int kSweepStaticLds = kWavesPerBlock * ( kMaxContracts + 1 ) * 2 * 8 + 64;
constexpr int kSobolDimsDev = 4097;       // implicit first dimension + 4096 Joe-Kuo rows
// the device direction table [kSobolDimsDev][32] is followed by per-dimension
// lane nibble tables [kSobolDimsDev][32] (see the Sobol ring of the sweep)
constexpr size_t kSobolNibbles = (size_t)kSobolDimsDev * 32;
// one GF(2) jump matrix of the xoshiro256++ state: 256 columns of four words
constexpr int kJumpWords = 256 * 4;

// LDS of the in-loop hedge inference: BF16 A-operand image of the network
// (40 tiles x 64 lanes x 8 bytes), FP32 biases (208), and per wave a
// [64][16] BF16 feature block plus a [64][8] FP32 hedge-ratio block
constexpr int kCvImgBytes = 40 * 64 * 8;
constexpr int kCvBiasBytes = 208 * 4;
constexpr int kCvWaveBytes = 64 * 16 * 2 + 64 * 8 * 4;
constexpr int kCvLdsBytes = kCvImgBytes + kCvBiasBytes + kWavesPerBlock * kCvWaveBytes;

// vanilla grids: per book up to 16 maturities x 64 strikes, calls and puts
constexpr int kGridMaxMat = 16;
constexpr int kGridMaxStrikes = 64;
constexpr int kGridMaxCells = 2 * kGridMaxStrikes;  // cell 2k = call k, 2k + 1 = put k
// LDS of the grid observation: per wave a [16 cells][64 lanes] transposition
// block (FP32 or FP64 values), then the [4 waves][128 cells][2] FP64 stage
constexpr int kGridRows = 16;
constexpr int grid_lds_bytes( int value_bytes )
{
    return kWavesPerBlock * kGridRows * 64 * value_bytes + kWavesPerBlock * kGridMaxCells * 2 * 8;
}

// FP32 per-(equity, date) step constants, precomputed on the host in FP64
// from the reference's QE / Andersen formulas (heston_variance_node.cpp:29-33,
// heston_spot_node.cpp:60-75).
enum StepConst
{
    SC_CARRY = 0,  // carry drift D_i t_i - D_{i-1} t_{i-1}
    SC_E,          // exp(-kappa dt)
    SC_M0,         // theta (1 - e)               : m  = SC_M0 + v SC_E
    SC_S1,         // xi^2 e (1 - e) / kappa       : s2 = SC_S0 + v SC_S1
    SC_S0,         // theta xi^2 (1 - e)^2 / (2 kappa)
    SC_K0,         // -rho kappa theta / xi dt     : K0 = L SC_K0
    SC_A1,         // g dt kappa rho / xi          : a  = L SC_A1 - L^2 SC_A2
    SC_A2,         // g dt / 2
    SC_RX,         // rho / xi                     : K1,2 = a -+ L SC_RX
    SC_K3,         // g dt (1 - rho^2)             : K3 = L^2 SC_K3
    SC_DT,         // dt
    SC_SQDT,       // sqrt(dt)
    SC_HALF_V2DT,  // BS: v^2 / 2 dt
    SC_VOL,        // BS: v
    SC_CVA,        // SC_A1 + SC_RX                : variance loading of the hedge increment
    SC_CVB,        // SC_A2 - SC_K3 / 2            :   A = L SC_CVA - L^2 SC_CVB
    SC_COUNT = 16
};

struct DevAsset
{
    int model;       // THOTH_MODEL_*
    int var_factor;  // factor of the variance noise (-1 for BS)
    double spot, vol, v0, kappa, theta, xi, rho;
};

struct DevBook
{
    // grid
    int n_dates, m, n_assets, factors;
    int n_contracts, n_rainbows, corr_stoch, lev_points;
    int lev_group;        // G: steps per LDS barrier (FP32 LSV leverage ring of 2G layers, Sobol ring)
    int sob_slots;        // S: bridge ops held in the Sobol ring of the sweep (a power of two)
    int n_ops;
    double sd0;           // bridge endpoint std dev sqrt(t_m)
    float sd0f;
    const double* t;      // [n_dates]
    const double* dt;     // [n_dates] YearFraction(d_{i-1}, d_i)
    const double* sqrt_dt;
    const double* bsq;    // [n_dates] sqrt(t_i - t_{i-1}) (0 when the step is empty)
    const float* binv;    // [n_dates] 1 / bsq (0 when empty)
    const int* step_ops;  // [m + 1] prefix offsets into the op arrays
    const int* op_b;      // [n_ops] bridge build index of the op (Sobol dim = b * factors + f)
    const int* op_left;   // [n_ops] has a left anchor
    const double* op_w;   // [n_ops * 3] left weight, right weight, std dev
    const float* op_wf;   // FP32 copy, left weight 0 where the op has no left anchor
    // equities
    DevAsset asset[kMaxAssets];
    // lower Cholesky factor that correlates the increments into date i,
    // [n_dates][kMaxAssets][kMaxAssets] (the same factor at every date unless
    // the correlation is term-structured); multi-equity books only
    const double* chol_t;
    const float* cholf_t;
    const double* drift;  // [n_assets * n_dates] drift-node value
    const float* sc;      // date-major [n_dates][n_assets][SC_COUNT]
    const double* lev;    // [n_assets * n_dates * lev_points]
    const float* levf;
    const double* lev_ln_step;  // [n_assets * n_dates]
    const double* lev_offset;   // [n_assets * n_dates] (integral values)
    const float* lev_offset_f;  // FP32 copy (exact: |offset| < 2^24), date-major [n_dates][n_assets]
    const float* lev_inv_step;  // [n_dates][n_assets] 1 / ln_step
    // pathwise basket Greeks: sensitivity of the LSV leverage layers to the
    // equity's own spot (s) and to the parallel vol shift (e) at each node,
    // [n_assets * n_dates * lev_points] (null when the book has none)
    const double* lev_ds;
    const double* lev_de;
    const float* lev_dsf;
    const float* lev_def;
    // payoffs
    int c_kind[kMaxContracts], c_call[kMaxContracts], c_udl[kMaxContracts];
    int c_up[kMaxContracts], c_in[kMaxContracts];
    double c_strike[kMaxContracts], c_level[kMaxContracts], c_df[kMaxContracts];
    const uint32_t* mon_mask;   // [n_dates] contracts monitored at the date
    const uint32_t* flow_mask;  // [n_dates] contracts paying at the date
    int rb_best[kMaxRainbows], rb_n[kMaxRainbows], rb_member[kMaxRainbows][kMaxAssets];
    double rb_weight[kMaxRainbows][kMaxAssets];
    double* partial;            // [replicates][n_blocks][1 + n_contracts][2] per-block sums
    // learned control variate data (FP32 sweeps only; cv_J = 0 disables):
    // hedge dates d_0 = 0 < d_1 < .. < d_J = m split the steps into J intervals
    int cv_J;
    const int* cv_start;        // [n_dates] j when date i = d_j (j < J), else -1
    const int* cv_end;          // [n_dates] j when date i = d_{j+1}, else -1
    const float* cv_scale;      // [n_dates][n_assets] df(t_i) * 100 / S0_a
    float cv_inv_T;             // 1 / t_m
    float cv_lnS0[kMaxAssets];
    float cv_ln_scale[kMaxAssets];   // 1 / (reference vol * sqrt(t_m)): standardised log-moneyness
    float cv_vol_scale[kMaxAssets];  // 1 / reference vol
    int64_t cv_N;               // paths per replicate (sample stride)
    const uint32_t* cv_wimg;    // in-loop inference: BF16 A-operand image of the hedge network
    const float* cv_bias;       //   and its FP32 biases
    uint16_t* cv_feat;          // [R][J][N][16] BF16 features at d_j
    float* cv_dm;               // [R][J][N][8] discounted rebased martingale increments over interval j
    float* cv_pay;              // [R][N] discounted book value
    // vanilla grid (grid_nm = 0: none): calls and puts on one underlying,
    // summed per (replicate, Sobol block, maturity, cell) into grid_part
    int grid_nm, grid_nk, grid_udl;
    int grid_grouped;           // one grid per equity (sweep kind 5), strikes [equity][nk]
    const int* grid_mat_at;     // [n_dates] maturity index j at the date, else -1
    const float* grid_kf;       // [nk] absolute strikes
    const double* grid_kd;
    double* grid_part;          // [R][tiles][nm][2 nk][2] undiscounted payoff sums
    // pseudo-random draws (mcl_prng.hpp): pseudo = 1 draws every normal of the
    // path from the lane's xoshiro256++ stream in (step, factor) order, with no
    // Sobol point and no bridge; otherwise the normals of the dimensions past
    // the Sobol budget come precomputed from `tail` / `tailf`,
    // [tail_dims][tail_stride] with the path (index - first) minor
    int pseudo;
    int tail_dims;
    int jump_levels;            // tabulated jump matrices (paths <= 2^jump_levels)
    int64_t tail_stride;
    const double* tail;
    const float* tailf;
    const uint64_t* jump;       // [jump_levels][256][4]: (M^T)^(2^k), T the draws per path
    const uint64_t* states;     // [4][tail_stride]: the stream state at the start of each path
    uint64_t seed_state[4];     // splitmix64 expansion of the seed (rng.hpp:36-47)
    // American exercise: the spot path of equity rec_udl at dates 0 .. rec_n - 1
    // (rec_n = 0: none), [rec_n][rec_stride] with the path (index - first) minor
    int rec_udl, rec_n;
    int64_t rec_stride;
    double* rec;
    float* recf;
};

}  // namespace thoth
