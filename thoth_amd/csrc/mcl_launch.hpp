// mcl_launch.hpp -- dispatch from (equities, bridge-stack depth) to the
// template instantiation of the path-sweep kernel.
#pragma once
#include <hip/hip_runtime.h>

#include "mcl_device.hpp"

// supported instantiations: equities per book {1, 2, 5, 8}, stack depth {2, 8, 10, 16}
#define THOTH_DISPATCH_DEPTH( FN, NA, depth, ... )     \
    do                                                 \
    {                                                  \
        if ( ( depth ) <= 2 )                          \
            FN<NA, 2>( __VA_ARGS__ );                  \
        else if ( ( depth ) <= 8 )                     \
            FN<NA, 8>( __VA_ARGS__ );                  \
        else if ( ( depth ) <= 10 )                    \
            FN<NA, 10>( __VA_ARGS__ );                 \
        else if ( ( depth ) <= 16 )                    \
            FN<NA, 16>( __VA_ARGS__ );                 \
        else                                           \
            return false;                              \
        return true;                                   \
    } while ( 0 )

#define THOTH_DISPATCH_SWEEP( FN, na, depth, ... )                     \
    do                                                                 \
    {                                                                  \
        if ( ( na ) <= 1 )                                             \
            THOTH_DISPATCH_DEPTH( FN, 1, depth, __VA_ARGS__ );         \
        else if ( ( na ) <= 2 )                                        \
            THOTH_DISPATCH_DEPTH( FN, 2, depth, __VA_ARGS__ );         \
        else if ( ( na ) <= 5 )                                        \
            THOTH_DISPATCH_DEPTH( FN, 5, depth, __VA_ARGS__ );         \
        else if ( ( na ) <= 8 )                                        \
            THOTH_DISPATCH_DEPTH( FN, 8, depth, __VA_ARGS__ );         \
        return false;                                                  \
    } while ( 0 )

namespace thoth
{
constexpr int kMaxStackDepth = 16;
bool launch_sweep_f64( int na, int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books,
                       const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last, const uint32_t* shifts );
// cv_mode: the kernel's control-variate mode (0 none, 1 training samples,
// 2 in-loop inference of a SiLU network, 3 of a smoothstep-gated network)
bool launch_sweep_f32( int na, int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books,
                       const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last, const uint32_t* shifts,
                       int cv_mode = 0 );
// uniform books (every equity LSV: kind 1, or Heston: kind 2, vol-of-vol > 0,
// uncorrelated stochastic spots, exactly na equities; kind 3: every equity LSV
// with Cholesky-correlated spots): false where no specialised instantiation
// exists (the caller then runs the generic kernel)
bool launch_sweep_f32_uniform( int kind, int na, int depth, dim3 grid, hipStream_t s, size_t lds,
                               const DevBook* books, const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last,
                               const uint32_t* shifts, int cv_mode );
// the five-equity LSV instantiations of the above, kinds 1 and 3 (a unit of their own)
bool launch_sweep_f32_uniform5( int kind, int cv_mode, int depth, dim3 grid, hipStream_t s, size_t lds,
                                const DevBook* books, const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last,
                                const uint32_t* shifts );
bool launch_persistent_f32_uniform5( int kind, int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books,
                                     const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last, const int2* items,
                                     int n_items, uint32_t ntiles, int* counter );
bool launch_persistent_f32_uniform( int kind, int na, int depth, dim3 grid, hipStream_t s, size_t lds,
                                    const DevBook* books, const uint32_t* V, uint32_t bq0, uint32_t first,
                                    uint32_t last, const int2* items, int n_items, uint32_t ntiles, int* counter );
void launch_mlp_pack( hipStream_t s, const float* w, uint16_t* fwd, uint16_t* bwd );
// act: the hidden activation of the network (kActSilu 0, kActGate 1, mcl_mlp.hpp)
// the two-phase evaluation: partial [R][cv_eval_blocks(N)][4] (16 paths per block)
uint32_t cv_eval_blocks( int64_t N );
void launch_cv_eval( hipStream_t s, int act, const uint16_t* fwd, const float* bias, const uint16_t* feat,
                     const float* dm, const float* pay, int64_t N, int J, int R, double* partial );
// one Adam step of the hedge network: minibatch of B paths x J dates split
// into `chunks` date chunks of dpc dates; one gradient slice per block
struct CvTrainLayout
{
    int B, J, tiles, dpc, chunks, waves, blocks;
};
CvTrainLayout cv_train_layout( int B, int J );
hipError_t cv_train_init();  // once per process before the first step (LDS limit of the gradient kernel)
// buffers: cvpart [chunks * B], gpart [blocks * THOTH_CV_WEIGHTS], gsum [kCvGradGroups * THOTH_CV_WEIGHTS],
// loss_part [blocks]
constexpr int kCvGradGroups = 32;
bool launch_cv_train_step( hipStream_t s, int act, const CvTrainLayout& l, const uint16_t* fwd, const uint16_t* bwd,
                           float* w, float* m, float* v, const uint16_t* feat, const float* dm, const float* pay,
                           int64_t N, int* step, float* cvpart, float* gpart, float* gsum, float* loss_part, float lr );
bool launch_persistent_f64( int na, int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books,
                           const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last, const int2* items,
                           int n_items, uint32_t ntiles, int* counter );
bool launch_persistent_f32( int na, int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books,
                           const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last, const int2* items,
                           int n_items, uint32_t ntiles, int* counter );
void launch_counter_reset( hipStream_t s, int* counter );
// books whose normals are not all Sobol points (pseudo-random tail dimensions,
// or use_sobol false): sweep_block kind 4, and the fill of the tail normals
// [T][paths] from the jump tables (levels: paths <= 2^levels)
bool launch_sweep_prng_f64( int na, int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books,
                            const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last );
bool launch_sweep_prng_f32( int na, int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books,
                            const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last );
// the stream state at the start of every path, [4][paths] (jump tables of T,
// paths <= 2^levels), then the tail normals [T][paths] drawn from those states
bool launch_prng_start_states( hipStream_t s, const uint64_t* jump, int levels, const uint64_t* seed_state, int T,
                               int64_t paths, uint64_t* states );
bool launch_tail_fill_f64( hipStream_t s, const uint64_t* states, int T, int64_t paths, double* out );
bool launch_tail_fill_f32( hipStream_t s, const uint64_t* states, int T, int64_t paths, float* out );
// Longstaff-Schwartz on recorded paths S [M][N] (precision 0: double, 1: float):
// the fit writes beta[4 t + 0..3] (coefficients, 1 where fitted) for t = M-2 .. 1
// (beta zeroed by the caller; cf [N], part [lsm_blocks(N) * 8] are scratch);
// the apply writes part[(scenario * lsm_blocks(N) + block) * 2 + {0, 1}] sums of
// the discounted value and its square, disc [scenarios][M]
int lsm_blocks( int64_t N );
bool launch_lsm_fit( hipStream_t s, int precision, const void* S, int64_t N, int M, double strike, int is_call,
                     double norm, const double* df, double* beta, double* cf, double* part );
bool launch_lsm_apply( hipStream_t s, int precision, const void* const* paths, int n_scen, int64_t N, int M,
                       double strike, int is_call, double norm, const double* disc, const double* beta, double* part );
bool launch_pathwise( int precision, int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books,
                      const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last );
bool launch_pathwise_f32( int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books, const uint32_t* V,
                          uint32_t bq0, uint32_t first, uint32_t last );
// the equity-wave latency form of the FP32 uniform-LSV sweep (sweep_waves):
// whether an instantiation exists, and its launch, grid (4 x Sobol blocks,
// books, replicates), 64 na threads per workgroup
bool sweep_waves_available( int na, int depth, int cv_mode );
bool launch_sweep_waves( int na, int depth, int cv_mode, dim3 grid, hipStream_t s, size_t lds, const DevBook* books,
                         const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last, const uint32_t* shifts );
// the pathwise basket kernel (3 + n_assets estimators per block); false when
// no instantiation covers the equity count or bridge depth.  lsv: the FP32
// form for books of na LSV equities with uncorrelated stochastic spots and
// the leverage ring in LDS (where pathwise_basket_lsv_available)
bool pathwise_basket_lsv_available( int na, int depth );
bool launch_pathwise_basket( int precision, int na, int depth, bool lsv, dim3 grid, hipStream_t s, size_t lds,
                             const DevBook* books, const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last );
bool launch_pathwise_basket_f32( int na, int depth, bool lsv, dim3 grid, hipStream_t s, size_t lds,
                                 const DevBook* books, const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last );
void launch_slot_sum( hipStream_t s, const double* part, int nslots, int per_slot, double* out );
void launch_grid_sum( hipStream_t s, const double* part, const long long* table, int n_books, int max_per_slot,
                      int nslots, double* out );
bool launch_sobol_points( hipStream_t s, const uint32_t* V, int dim, uint32_t first, uint32_t last, uint32_t* out );
bool launch_special_f32( hipStream_t s, int fn, int64_t n, const uint32_t* in, float* out );
}  // namespace thoth
