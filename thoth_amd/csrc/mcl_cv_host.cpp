// mcl_cv_host.cpp -- host side of the learned martingale control variate
// (include/thoth_mcl_cv.h): training-data sweep, Adam training as one
// hipGraph replayed per step, and randomised-QMC pricing with and without
// the control variate.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "mcl_hostlib.hpp"
#include "mcl_internal.hpp"
#include "mcl_launch.hpp"
#include "thoth_mcl_cv.h"

namespace thoth
{

namespace
{

constexpr int kWSize = 10240, kBSize = 208, kN = kWSize + kBSize;
static_assert( kN == THOTH_CV_WEIGHTS, "weight layout" );
constexpr int kWOff[4] = { 0, 64 * 16, 64 * 16 + 64 * 64, 64 * 16 + 2 * 64 * 64 };
constexpr int kRows[4] = { 64, 64, 64, 16 };
constexpr int kCols[4] = { 16, 64, 64, 64 };
constexpr int kIn[4] = { 11, 64, 64, 64 };   // true fan-in
constexpr int kOut[4] = { 64, 64, 64, 5 };   // true fan-out

struct Dev
{
    std::vector<void*> p;
    ~Dev()
    {
        for ( void* x : p )
            (void)hipFree( x );
    }
    template <class T>
    T* alloc( size_t n )
    {
        void* x = nullptr;
        if ( hipMalloc( &x, n * sizeof( T ) ) != hipSuccess )
            return nullptr;
        p.push_back( x );
        return static_cast<T*>( x );
    }
};

#define HIP_OK( expr )                                                                                      \
    do                                                                                                      \
    {                                                                                                       \
        hipError_t e_ = ( expr );                                                                           \
        if ( e_ != hipSuccess )                                                                             \
            return SetError( THOTH_ERR_DEVICE, std::string( #expr " failed: " ) + hipGetErrorString( e_ ) ); \
    } while ( 0 )

// seeded initialisation: uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero
// biases, zero output layer (the control variate starts at zero)
void InitWeights( uint64_t seed, float* w )
{
    std::mt19937_64 rng( seed );
    std::fill( w, w + kN, 0.0f );
    for ( int L = 0; L < 3; L++ )
    {
        const float a = std::sqrt( 6.0f / (float)kIn[L] );  // variance 2 / fan-in (He)
        std::uniform_real_distribution<float> u( -a, a );
        for ( int o = 0; o < kOut[L]; o++ )
            for ( int i = 0; i < kIn[L]; i++ )
                w[kWOff[L] + o * kCols[L] + i] = u( rng );
    }
}

int CheckConfig( const thoth_cv_config& c )
{
    if ( c.hedge_dates < 1 || !( c.hedge_spacing >= 1.0f && c.hedge_spacing <= 8.0f ) || c.batch_paths < 64 ||
         c.batch_paths % 64 || c.adam_steps < 0 || c.train_paths < c.batch_paths ||
         c.train_sobol_skip < 0 || !( c.learning_rate > 0 ) ||
         ( c.activation != THOTH_CV_SILU && c.activation != THOTH_CV_GATE ) )
        return SetError( THOTH_ERR_INVALID, "control-variate config out of range" );
    return THOTH_OK;
}

// FP32 CV-data sweep of the base book, `paths` paths per replicate: the
// sample buffers and the batch, resident and ready to sweep
int CvPrepare( const thoth_mcl_request& req, int J, float spacing, int64_t paths, int64_t skip, int R,
               uint64_t rqmc_seed, Dev& dev, CvSpec& cv, DeviceBatch& B )
{
    host::BookArrays A;
    host::Grid g;
    int rc = host::ResolveBaseBook( req, A, g );
    if ( rc != THOTH_OK )
        return rc;
    const int Rr = std::max( 1, R );
    cv.J = J;
    cv.spacing = spacing;
    cv.N = paths;
    cv.feat = dev.alloc<uint16_t>( (size_t)Rr * J * paths * 16 );
    cv.dm = dev.alloc<float>( (size_t)Rr * J * paths * 8 );
    cv.pay = dev.alloc<float>( (size_t)Rr * paths );
    if ( !cv.feat || !cv.dm || !cv.pay )
        return SetError( THOTH_ERR_DEVICE, "cannot allocate the control-variate sample buffers" );
    thoth_mcl_sweep_cfg cfg{};
    cfg.paths = paths;
    cfg.sobol_skip = skip;
    cfg.precision = 1;
    cfg.replicates = R;
    cfg.rqmc_seed = rqmc_seed;
    return PrepareBatch( 1, &A.book, &cfg, B, &cv );
}

int CvSweep( const thoth_mcl_request& req, int J, float spacing, int64_t paths, int64_t skip, int R, uint64_t rqmc_seed,
             Dev& dev, CvSpec& cv, std::vector<thoth_mcl_book_out>& out, double& ms )
{
    DeviceBatch B;
    int rc = CvPrepare( req, J, spacing, paths, skip, R, rqmc_seed, dev, cv, B );
    if ( rc != THOTH_OK )
        return rc;
    out.assign( std::max( 1, R ), thoth_mcl_book_out{} );
    rc = RunBatch( B, out.data() );
    ms = out[0].device_ms;
    return rc;
}

}  // namespace

}  // namespace thoth

using namespace thoth;

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_cv_train( const thoth_mcl_request* req,
                                                                        const thoth_cv_config* cfg,
                                                                        thoth_cv_weights* w, thoth_cv_train_report* rep )
{
    if ( !req || !cfg || !w )
        return SetError( THOTH_ERR_INVALID, "null argument" );
    std::lock_guard<std::mutex> lk( DeviceMutex() );
    try
    {
        int rc = EnsureDevice();
        if ( rc != THOTH_OK )
            return rc;
        if ( ( rc = CheckConfig( *cfg ) ) != THOTH_OK )
            return rc;
        Dev dev;
        CvSpec cv;
        std::vector<thoth_mcl_book_out> out;
        double data_ms = 0;
        const int J = cfg->hedge_dates;
        const int64_t N = cfg->train_paths;
        rc = CvSweep( *req, J, cfg->hedge_spacing, N, cfg->train_sobol_skip, 0, 0, dev, cv, out, data_ms );
        if ( rc != THOTH_OK )
            return rc;
        // centre the payoffs on their training mean: E[CV] = 0 for any weights,
        // so the minimiser of mean (P - c - CV)^2 does not depend on c, while
        // the gradient loses the noise of c dCV/dw
        {
            hipStream_t s = Stream();
            std::vector<float> pay( N );
            HIP_OK( hipMemcpyAsync( pay.data(), cv.pay, N * sizeof( float ), hipMemcpyDeviceToHost, s ) );
            HIP_OK( hipStreamSynchronize( s ) );
            double sum = 0;
            for ( float x : pay )
                sum += x;
            const float mean = (float)( sum / (double)N );
            for ( float& x : pay )
                x -= mean;
            HIP_OK( hipMemcpyAsync( cv.pay, pay.data(), N * sizeof( float ), hipMemcpyHostToDevice, s ) );
            HIP_OK( hipStreamSynchronize( s ) );
        }
        const int B = cfg->batch_paths;
        const CvTrainLayout lay = cv_train_layout( B, J );
        float* dw = dev.alloc<float>( kN );
        float* dm_ = dev.alloc<float>( kN );
        float* dv = dev.alloc<float>( kN );
        uint16_t* fwd = dev.alloc<uint16_t>( kWSize );
        uint16_t* bwd = dev.alloc<uint16_t>( kWSize );
        float* cvpart = dev.alloc<float>( (size_t)lay.chunks * B );
        float* gpart = dev.alloc<float>( (size_t)lay.blocks * kN );
        float* gsum = dev.alloc<float>( (size_t)kCvGradGroups * kN );
        float* loss_part = dev.alloc<float>( lay.blocks );
        int* step = dev.alloc<int>( 1 );
        if ( !dw || !dm_ || !dv || !fwd || !bwd || !cvpart || !gpart || !gsum || !loss_part || !step )
            return SetError( THOTH_ERR_DEVICE, "cannot allocate the training buffers" );
        HIP_OK( cv_train_init() );
        std::vector<float> w0( kN );
        InitWeights( cfg->init_seed, w0.data() );
        hipStream_t s = Stream();
        HIP_OK( hipMemcpyAsync( dw, w0.data(), kN * sizeof( float ), hipMemcpyHostToDevice, s ) );
        HIP_OK( hipMemsetAsync( dm_, 0, kN * sizeof( float ), s ) );
        HIP_OK( hipMemsetAsync( dv, 0, kN * sizeof( float ), s ) );
        HIP_OK( hipMemsetAsync( step, 0, sizeof( int ), s ) );
        launch_mlp_pack( s, dw, fwd, bwd );
        // one Adam step = gradient + Adam + BF16 repack + step counter, captured once
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        HIP_OK( hipStreamBeginCapture( s, hipStreamCaptureModeThreadLocal ) );
        const bool launched = launch_cv_train_step( s, cfg->activation, lay, fwd, bwd, dw, dm_, dv, cv.feat, cv.dm,
                                                    cv.pay, N, step,
                                                    cvpart, gpart, gsum, loss_part, cfg->learning_rate );
        HIP_OK( hipStreamEndCapture( s, &graph ) );
        if ( !launched )
        {
            (void)hipGraphDestroy( graph );
            return SetError( THOTH_ERR_INVALID, "batch layout not supported" );
        }
        HIP_OK( hipGraphInstantiate( &exec, graph, nullptr, nullptr, 0 ) );
        std::vector<float> lp( lay.blocks );
        double loss_first = 0, loss_last = 0;
        hipEvent_t e0, e1;
        HIP_OK( hipEventCreate( &e0 ) );
        HIP_OK( hipEventCreate( &e1 ) );
        HIP_OK( hipEventRecord( e0, s ) );
        for ( int k = 0; k < cfg->adam_steps; k++ )
        {
            HIP_OK( hipGraphLaunch( exec, s ) );
            if ( k == 0 || k == cfg->adam_steps - 1 )
            {
                HIP_OK( hipMemcpyAsync( lp.data(), loss_part, lp.size() * sizeof( float ), hipMemcpyDeviceToHost, s ) );
                HIP_OK( hipStreamSynchronize( s ) );
                double l = 0;
                for ( float x : lp )
                    l += x;
                ( k == 0 ? loss_first : loss_last ) = l / B;
            }
        }
        HIP_OK( hipEventRecord( e1, s ) );
        HIP_OK( hipMemcpyAsync( w->w, dw, kN * sizeof( float ), hipMemcpyDeviceToHost, s ) );
        HIP_OK( hipStreamSynchronize( s ) );
        float ms = 0;
        HIP_OK( hipEventElapsedTime( &ms, e0, e1 ) );
        (void)hipEventDestroy( e0 );
        (void)hipEventDestroy( e1 );
        (void)hipGraphExecDestroy( exec );
        (void)hipGraphDestroy( graph );
        if ( rep )
        {
            rep->loss_first = loss_first;
            rep->loss_last = cfg->adam_steps > 1 ? loss_last : loss_first;
            rep->data_ms = data_ms;
            rep->train_ms = ms;
            rep->steps = cfg->adam_steps;
        }
        return THOTH_OK;
    }
    catch ( const host::Invalid& e )
    {
        return SetError( e.code, e.what() );
    }
    catch ( const std::exception& e )
    {
        return SetError( THOTH_ERR_INVALID, e.what() );
    }
}

namespace thoth
{
namespace
{

// 95 % confidence intervals from per-replicate sums (sum Y, sum Y^2, sum P,
// sum P^2) of R replicates of N paths each
void FinishCi( const std::vector<double>& sums, int R, int64_t N, thoth_cv_result* res )
{
    std::vector<double> my( R ), mp( R );
    double sy = 0, sy2 = 0, sp = 0, sp2 = 0;
    for ( int r = 0; r < R; r++ )
    {
        my[r] = sums[4 * r] / N;
        mp[r] = sums[4 * r + 2] / N;
        sy += sums[4 * r];
        sy2 += sums[4 * r + 1];
        sp += sums[4 * r + 2];
        sp2 += sums[4 * r + 3];
    }
    auto ci = [&]( const std::vector<double>& x, double& mean, double& hw )
    {
        mean = 0;
        for ( double a : x )
            mean += a;
        mean /= R;
        double var = 0;
        for ( double a : x )
            var += ( a - mean ) * ( a - mean );
        var /= ( R - 1 );
        hw = StudentT975( R - 1 ) * std::sqrt( var / R );
    };
    ci( my, res->cv_mean, res->cv_halfwidth );
    ci( mp, res->plain_mean, res->plain_halfwidth );
    // paired: plain - (P - CV) per replicate is the replicate's mean of CV
    std::vector<double> md( R );
    for ( int r = 0; r < R; r++ )
        md[r] = mp[r] - my[r];
    ci( md, res->cv_term_mean, res->cv_term_halfwidth );
    for ( int r = 0; r < R; r++ )
    {
        if ( res->replicate_plain_means )
            res->replicate_plain_means[r] = mp[r];
        if ( res->replicate_cv_means )
            res->replicate_cv_means[r] = my[r];
    }
    const double n = (double)N * R;
    res->cv_var = ( sy2 - sy * sy / n ) / ( n - 1 );
    res->plain_var = ( sp2 - sp * sp / n ) / ( n - 1 );
    res->replicates = R;
    res->paths = N;
}

// the hedge network inside the stepping loop: one FP32 sweep of R replicates
int PriceInLoop( const thoth_mcl_request& req, const thoth_cv_config& cfg, const thoth_cv_weights& w, int R,
                 uint64_t rqmc_seed, thoth_cv_result* res )
{
    const auto t0 = std::chrono::steady_clock::now();
    host::BookArrays A;
    host::Grid g;
    int rc = host::ResolveBaseBook( req, A, g );
    if ( rc != THOTH_OK )
        return rc;
    Dev dev;
    float* dw = dev.alloc<float>( kN );
    uint16_t* fwd = dev.alloc<uint16_t>( kWSize );
    uint16_t* bwd = dev.alloc<uint16_t>( kWSize );
    if ( !dw || !fwd || !bwd )
        return SetError( THOTH_ERR_DEVICE, "cannot allocate the network buffers" );
    hipStream_t s = Stream();
    HIP_OK( hipMemcpyAsync( dw, w.w, kN * sizeof( float ), hipMemcpyHostToDevice, s ) );
    launch_mlp_pack( s, dw, fwd, bwd );
    CvSpec cv;
    cv.mode = 2;
    cv.act = cfg.activation;
    cv.J = cfg.hedge_dates;
    cv.spacing = cfg.hedge_spacing;
    cv.N = req.mcl.paths;
    cv.wimg = fwd;
    cv.bias = dw + kWSize;
    thoth_mcl_sweep_cfg sc{};
    sc.paths = req.mcl.paths;
    sc.precision = 1;
    sc.replicates = R;
    sc.rqmc_seed = rqmc_seed;
    DeviceBatch B;
    if ( ( rc = PrepareBatch( 1, &A.book, &sc, B, &cv ) ) != THOTH_OK )
        return rc;
    const auto t1 = std::chrono::steady_clock::now();
    std::vector<thoth_mcl_book_out> out( R );
    std::vector<CvSums> cs( R );
    if ( ( rc = RunBatch( B, out.data(), cs.data() ) ) != THOTH_OK )
        return rc;
    std::vector<double> sums( 4 * (size_t)R );
    for ( int r = 0; r < R; r++ )
    {
        sums[4 * r] = cs[r].sum;
        sums[4 * r + 1] = cs[r].sum2;
        sums[4 * r + 2] = out[r].book_sum;
        sums[4 * r + 3] = out[r].book_sum2;
    }
    FinishCi( sums, R, req.mcl.paths, res );
    const auto t2 = std::chrono::steady_clock::now();
    res->setup_ms = std::chrono::duration<double, std::milli>( t1 - t0 ).count();
    res->run_wall_ms = std::chrono::duration<double, std::milli>( t2 - t1 ).count();
    res->sweep_ms = out[0].device_ms;
    res->eval_ms = 0;
    return THOTH_OK;
}

// two phases: FP32 sweep writing the samples, then the inference kernel
int PriceTwoPhase( const thoth_mcl_request& req, const thoth_cv_config& cfg, const thoth_cv_weights& w, int R,
                   uint64_t rqmc_seed, thoth_cv_result* res )
{
    // set-up: the book resident, the sample and evaluation buffers allocated,
    // the network packed; then the timed run: sweep, evaluation, reduction
    const auto t0 = std::chrono::steady_clock::now();
    Dev dev;
    CvSpec cv;
    const int J = cfg.hedge_dates;
    const int64_t N = req.mcl.paths;
    DeviceBatch B;
    int rc = CvPrepare( req, J, cfg.hedge_spacing, N, 0, R, rqmc_seed, dev, cv, B );
    if ( rc != THOTH_OK )
        return rc;
    float* dw = dev.alloc<float>( kN );
    uint16_t* fwd = dev.alloc<uint16_t>( kWSize );
    uint16_t* bwd = dev.alloc<uint16_t>( kWSize );
    const uint32_t nb = cv_eval_blocks( N );
    double* part = dev.alloc<double>( (size_t)R * nb * 4 );
    if ( !dw || !fwd || !bwd || !part )
        return SetError( THOTH_ERR_DEVICE, "cannot allocate the evaluation buffers" );
    hipStream_t s = Stream();
    HIP_OK( hipMemcpyAsync( dw, w.w, kN * sizeof( float ), hipMemcpyHostToDevice, s ) );
    launch_mlp_pack( s, dw, fwd, bwd );
    hipEvent_t e0, e1;
    HIP_OK( hipEventCreate( &e0 ) );
    HIP_OK( hipEventCreate( &e1 ) );
    // pinned host copy of the evaluation's partial sums
    double* hp = nullptr;
    const size_t nhp = (size_t)R * nb * 4;
    HIP_OK( hipHostMalloc( &hp, nhp * sizeof( double ) ) );
    struct PinnedFree
    {
        double* p;
        ~PinnedFree() { (void)hipHostFree( p ); }
    } hp_free{ hp };
    HIP_OK( hipStreamSynchronize( s ) );
    const auto t1 = std::chrono::steady_clock::now();
    // the sweep, then (same stream, no synchronisation between) the evaluation
    std::vector<thoth_mcl_book_out> out( R );
    rc = RunBatch( B, out.data(), nullptr, [&]( hipStream_t st ) {
        if ( hipEventRecord( e0, st ) != hipSuccess )
            return SetError( THOTH_ERR_DEVICE, "hipEventRecord" );
        launch_cv_eval( st, cfg.activation, fwd, dw + kWSize, cv.feat, cv.dm, cv.pay, N, J, R, part );
        if ( hipGetLastError() != hipSuccess || hipEventRecord( e1, st ) != hipSuccess ||
             hipMemcpyAsync( hp, part, nhp * sizeof( double ), hipMemcpyDeviceToHost, st ) != hipSuccess )
            return SetError( THOTH_ERR_DEVICE, "control-variate evaluation launch failed" );
        return (int)THOTH_OK;
    } );
    if ( rc != THOTH_OK )
        return rc;
    const double sweep_ms = out[0].device_ms;
    float ems = 0;
    HIP_OK( hipEventElapsedTime( &ems, e0, e1 ) );
    (void)hipEventDestroy( e0 );
    (void)hipEventDestroy( e1 );
    std::vector<double> sums( 4 * (size_t)R, 0.0 );
    for ( int r = 0; r < R; r++ )
    {
        double c[4] = { 0, 0, 0, 0 };
        for ( uint32_t b = 0; b < nb; b++ )
            for ( int k = 0; k < 4; k++ )
            {
                // Kahan over blocks in block order
                const double y = hp[( (size_t)r * nb + b ) * 4 + k] - c[k];
                const double t = sums[4 * r + k] + y;
                c[k] = ( t - sums[4 * r + k] ) - y;
                sums[4 * r + k] = t;
            }
    }
    FinishCi( sums, R, N, res );
    const auto t2 = std::chrono::steady_clock::now();
    res->sweep_ms = sweep_ms;
    res->eval_ms = ems;
    res->setup_ms = std::chrono::duration<double, std::milli>( t1 - t0 ).count();
    res->run_wall_ms = std::chrono::duration<double, std::milli>( t2 - t1 ).count();
    return THOTH_OK;
}

}  // namespace
}  // namespace thoth

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_cv_price( const thoth_mcl_request* req,
                                                                        const thoth_cv_config* cfg,
                                                                        const thoth_cv_weights* w, int32_t replicates,
                                                                        uint64_t rqmc_seed, thoth_cv_result* res )
{
    if ( !req || !cfg || !w || !res )
        return SetError( THOTH_ERR_INVALID, "null argument" );
    if ( replicates < 2 )
        return SetError( THOTH_ERR_INVALID, "at least 2 replicates are needed for a confidence interval" );
    std::lock_guard<std::mutex> lk( DeviceMutex() );
    try
    {
        int rc = EnsureDevice();
        if ( rc != THOTH_OK )
            return rc;
        if ( ( rc = CheckConfig( *cfg ) ) != THOTH_OK )
            return rc;
        if ( cfg->eval_mode == THOTH_CV_IN_LOOP )
            return PriceInLoop( *req, *cfg, *w, replicates, rqmc_seed, res );
        if ( cfg->eval_mode == THOTH_CV_TWO_PHASE )
            return PriceTwoPhase( *req, *cfg, *w, replicates, rqmc_seed, res );
        return SetError( THOTH_ERR_INVALID, "unknown evaluation mode" );
    }
    catch ( const host::Invalid& e )
    {
        return SetError( e.code, e.what() );
    }
    catch ( const std::exception& e )
    {
        return SetError( THOTH_ERR_INVALID, e.what() );
    }
}

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_cv_loss_grad( const thoth_cv_weights* w,
                                                                            const uint16_t* feat, const float* dm,
                                                                            const float* pay, int32_t B, int32_t J,
                                                                            int32_t activation, double* loss,
                                                                            float* grad )
{
    if ( !w || !feat || !dm || !pay || B < 64 || B % 64 || J < 1 )
        return SetError( THOTH_ERR_INVALID, "B must be a positive multiple of 64" );
    if ( activation != THOTH_CV_SILU && activation != THOTH_CV_GATE )
        return SetError( THOTH_ERR_INVALID, "unknown activation" );
    std::lock_guard<std::mutex> lk( DeviceMutex() );
    int rc = EnsureDevice();
    if ( rc != THOTH_OK )
        return rc;
    Dev dev;
    const CvTrainLayout lay = cv_train_layout( B, J );
    float* dw = dev.alloc<float>( kN );
    float* dm_ = dev.alloc<float>( kN );
    float* dv = dev.alloc<float>( kN );
    uint16_t* fwd = dev.alloc<uint16_t>( kWSize );
    uint16_t* bwd = dev.alloc<uint16_t>( kWSize );
    uint16_t* dfeat = dev.alloc<uint16_t>( (size_t)J * B * 16 );
    float* ddm = dev.alloc<float>( (size_t)J * B * 8 );
    float* dpay = dev.alloc<float>( B );
    float* cvpart = dev.alloc<float>( (size_t)lay.chunks * B );
    float* gpart = dev.alloc<float>( (size_t)lay.blocks * kN );
    float* gsum = dev.alloc<float>( (size_t)kCvGradGroups * kN );
    float* loss_part = dev.alloc<float>( lay.blocks );
    int* step = dev.alloc<int>( 1 );
    if ( !dw || !dm_ || !dv || !fwd || !bwd || !dfeat || !ddm || !dpay || !cvpart || !gpart || !gsum || !loss_part ||
         !step )
        return SetError( THOTH_ERR_DEVICE, "allocation failed" );
    HIP_OK( cv_train_init() );
    hipStream_t s = Stream();
    HIP_OK( hipMemcpyAsync( dw, w->w, kN * sizeof( float ), hipMemcpyHostToDevice, s ) );
    HIP_OK( hipMemcpyAsync( dfeat, feat, (size_t)J * B * 16 * sizeof( uint16_t ), hipMemcpyHostToDevice, s ) );
    HIP_OK( hipMemcpyAsync( ddm, dm, (size_t)J * B * 8 * sizeof( float ), hipMemcpyHostToDevice, s ) );
    HIP_OK( hipMemcpyAsync( dpay, pay, B * sizeof( float ), hipMemcpyHostToDevice, s ) );
    HIP_OK( hipMemsetAsync( dm_, 0, kN * sizeof( float ), s ) );
    HIP_OK( hipMemsetAsync( dv, 0, kN * sizeof( float ), s ) );
    HIP_OK( hipMemsetAsync( step, 0, sizeof( int ), s ) );
    launch_mlp_pack( s, dw, fwd, bwd );
    // gradient only: the training step with a zero learning rate leaves w unchanged
    if ( !launch_cv_train_step( s, activation, lay, fwd, bwd, dw, dm_, dv, dfeat, ddm, dpay, B, step, cvpart, gpart, gsum,
                                loss_part, 0.0f ) )
        return SetError( THOTH_ERR_INVALID, "batch layout not supported" );
    HIP_OK( hipGetLastError() );
    std::vector<float> gp( (size_t)lay.blocks * kN ), lp( lay.blocks );
    HIP_OK( hipMemcpyAsync( gp.data(), gpart, gp.size() * sizeof( float ), hipMemcpyDeviceToHost, s ) );
    HIP_OK( hipMemcpyAsync( lp.data(), loss_part, lp.size() * sizeof( float ), hipMemcpyDeviceToHost, s ) );
    HIP_OK( hipStreamSynchronize( s ) );
    double l = 0;
    for ( float x : lp )
        l += x;
    if ( loss )
        *loss = l / B;
    if ( grad )
        for ( int e = 0; e < kN; e++ )
        {
            double g = 0;
            for ( int k = 0; k < lay.blocks; k++ )
                g += gp[(size_t)k * kN + e];
            grad[e] = (float)g;
        }
    return THOTH_OK;
}

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_cv_samples( const thoth_mcl_request* req,
                                                                          int32_t hedge_dates, int64_t paths,
                                                                          int64_t sobol_skip, int32_t replicates,
                                                                          uint64_t rqmc_seed, uint16_t* feat,
                                                                          float* dm, float* pay )
{
    if ( !req || !feat || !dm || !pay || paths < 1 || sobol_skip < 0 || replicates < 0 )
        return SetError( THOTH_ERR_INVALID, "null argument or empty sample" );
    std::lock_guard<std::mutex> lk( DeviceMutex() );
    try
    {
        int rc = EnsureDevice();
        if ( rc != THOTH_OK )
            return rc;
        Dev dev;
        CvSpec cv;
        std::vector<thoth_mcl_book_out> out;
        double ms = 0;
        rc = CvSweep( *req, hedge_dates, 1.0f, paths, sobol_skip, replicates, rqmc_seed, dev, cv, out, ms );
        if ( rc != THOTH_OK )
            return rc;
        hipStream_t s = Stream();
        const size_t reps = (size_t)std::max( 1, replicates );
        const size_t rows = reps * hedge_dates * paths;
        HIP_OK( hipMemcpyAsync( feat, cv.feat, rows * 16 * sizeof( uint16_t ), hipMemcpyDeviceToHost, s ) );
        HIP_OK( hipMemcpyAsync( dm, cv.dm, rows * 8 * sizeof( float ), hipMemcpyDeviceToHost, s ) );
        HIP_OK( hipMemcpyAsync( pay, cv.pay, reps * paths * sizeof( float ), hipMemcpyDeviceToHost, s ) );
        HIP_OK( hipStreamSynchronize( s ) );
        return THOTH_OK;
    }
    catch ( const host::Invalid& e )
    {
        return SetError( e.code, e.what() );
    }
    catch ( const std::exception& e )
    {
        return SetError( THOTH_ERR_INVALID, e.what() );
    }
}
