// mcl_grid.cpp -- vanilla grids (include/thoth_mcl_grid.h): every book of
// every request (underlying x bump scenario) swept by one persistent launch
// pulling (book, Sobol block) items from a queue, replayed from a hipGraph,
// with the launch shape chosen by timing candidates on the plan itself.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "mcl_hostlib.hpp"
#include "mcl_internal.hpp"
#include "mcl_launch.hpp"
#include "thoth_mcl_grid.h"

namespace thoth
{

#define HIP_OK( expr )                                                                                      \
    do                                                                                                      \
    {                                                                                                       \
        hipError_t e_ = ( expr );                                                                           \
        if ( e_ != hipSuccess )                                                                             \
            return SetError( THOTH_ERR_DEVICE, std::string( #expr " failed: " ) + hipGetErrorString( e_ ) ); \
    } while ( 0 )

struct GridPlan
{
    DeviceBatch batch;
    int n_req = 0;
    int64_t paths = 0;
    std::vector<int> base_book, delta_book, vega_book;  // per request (-1: not swept)
    std::vector<double> delta_h;                        // spot bump of the grid underlying
    std::vector<int> nm, nk;
    std::vector<std::vector<double>> df;                // per request, per maturity
    int2* ditems = nullptr;
    int* dcounter = nullptr;
    int n_items = 0;
    int cus = 0;
    int blocks_per_cu = 2, lev_group = 4;
    double tune_ms = 0;
    int candidates = 0;
    // grouped books: original book b is equity slot_of[b] of device book
    // group_of[b] (books_per_lane = 1: every book its own device book)
    int books_per_lane = 1;
    std::vector<int> group_of, slot_of;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~GridPlan()
    {
        if ( exec )
            (void)hipGraphExecDestroy( exec );
        if ( graph )
            (void)hipGraphDestroy( graph );
        if ( ditems )
            (void)hipFree( ditems );
        if ( dcounter )
            (void)hipFree( dcounter );
    }
};

namespace
{

// enqueue one sweep of every item + the per-book grid sums on stream s
int Enqueue( const GridPlan& P, hipStream_t s )
{
    const DeviceBatch& B = P.batch;
    launch_counter_reset( s, P.dcounter );
    const dim3 grid( (uint32_t)( P.cus * P.blocks_per_cu ) );
    const bool ok = B.precision == 0
                        ? launch_persistent_f64( B.na_max, B.depth, grid, s, B.lds_bytes, B.dbooks, DeviceDirections(),
                                                 B.q0, B.first, B.last, P.ditems, P.n_items, B.nblocks, P.dcounter )
                        : ( B.kind && launch_persistent_f32_uniform( B.kind, B.na_max, B.depth, grid, s, B.lds_bytes,
                                                                     B.dbooks, DeviceDirections(), B.q0, B.first,
                                                                     B.last, P.ditems, P.n_items, B.nblocks,
                                                                     P.dcounter ) ) ||
                              ( B.kind != 5 &&  // grouped books have no generic form
                                launch_persistent_f32( B.na_max, B.depth, grid, s, B.lds_bytes, B.dbooks,
                                                     DeviceDirections(), B.q0, B.first, B.last, P.ditems, P.n_items,
                                                     B.nblocks, P.dcounter ) );
    if ( !ok )
        return SetError( THOTH_ERR_UNSUPPORTED, "no kernel instantiation for this book shape" );
    HIP_OK( hipGetLastError() );
    return ReduceGrids( B, s );
}

// device time of one direct (uncaptured) launch
int TimeOnce( const GridPlan& P, double& ms )
{
    hipStream_t s = Stream();
    hipEvent_t e0, e1;
    HIP_OK( hipEventCreate( &e0 ) );
    HIP_OK( hipEventCreate( &e1 ) );
    HIP_OK( hipEventRecord( e0, s ) );
    int rc = Enqueue( P, s );
    HIP_OK( hipEventRecord( e1, s ) );
    HIP_OK( hipEventSynchronize( e1 ) );
    float f = 0;
    HIP_OK( hipEventElapsedTime( &f, e0, e1 ) );
    (void)hipEventDestroy( e0 );
    (void)hipEventDestroy( e1 );
    ms = f;
    return rc;
}

// Autotuner: every (workgroups per CU, steps per LDS barrier) candidate
// is launched twice on the plan's own books and the second time counts; the
// fastest is kept.
int Autotune( GridPlan& P )
{
    const int bpc[] = { 1, 2, 4, 6, 8 };
    const int lg[] = { 1, 2, 4, 8 };
    double best = 1e300;
    int best_b = P.blocks_per_cu, best_g = P.lev_group;
    P.candidates = 0;
    for ( int g : lg )
    {
        int rc = SetLevGroup( P.batch, g );
        if ( rc == THOTH_ERR_UNSUPPORTED )
            continue;  // this G does not fit the LDS of a workgroup
        if ( rc != THOTH_OK )
            return rc;
        P.lev_group = g;
        for ( int b : bpc )
        {
            P.blocks_per_cu = b;
            double ms = 0;
            if ( ( rc = TimeOnce( P, ms ) ) != THOTH_OK || ( rc = TimeOnce( P, ms ) ) != THOTH_OK )
                return rc;
            P.candidates++;
            if ( ms < best )
            {
                best = ms;
                best_b = b;
                best_g = g;
            }
        }
    }
    P.blocks_per_cu = best_b;
    P.lev_group = best_g;
    P.tune_ms = best;
    return SetLevGroup( P.batch, best_g );
}

int Capture( GridPlan& P )
{
    hipStream_t s = Stream();
    HIP_OK( hipStreamBeginCapture( s, hipStreamCaptureModeThreadLocal ) );
    const int rc = Enqueue( P, s );
    hipGraph_t g = nullptr;
    HIP_OK( hipStreamEndCapture( s, &g ) );
    if ( rc != THOTH_OK )
    {
        (void)hipGraphDestroy( g );
        return rc;
    }
    P.graph = g;
    HIP_OK( hipGraphInstantiate( &P.exec, P.graph, nullptr, nullptr, 0 ) );
    return THOTH_OK;
}

}  // namespace

namespace host
{

// K single-equity LSV books on the same grid and Sobol points (FP32) can be
// stepped as the K equities of one device book on two shared factors (sweep
// kind 5): the same normals and bridge serve all K.  K = 5, else 2, when it
// divides the book count; 1 otherwise.
static int GroupSize( const std::vector<thoth_mcl_book>& books, const std::vector<GridSpec>& specs, int precision )
{
    const size_t n = books.size();
    if ( precision != THOTH_FP32 || n < 2 )
        return 1;
    const thoth_mcl_book& b0 = books[0];
    for ( size_t b = 0; b < n; b++ )
    {
        const thoth_mcl_book& x = books[b];
        if ( x.n_assets != 1 || x.model[0] != THOTH_MODEL_LSV || !( x.xi[0] > 1e-12 ) || x.n_contracts != 0 ||
             x.n_rainbows != 0 || x.n_dates != b0.n_dates || x.lev_points != b0.lev_points ||
             !std::equal( x.t, x.t + x.n_dates, b0.t ) ||
             !std::equal( x.disc_rate, x.disc_rate + x.n_dates, b0.disc_rate ) || specs[b].nm != specs[0].nm ||
             specs[b].nk != specs[0].nk ||
             !std::equal( specs[b].mat_date, specs[b].mat_date + specs[b].nm, specs[0].mat_date ) )
            return 1;
    }
    return n % 5 == 0 ? 5 : n % 2 == 0 ? 2 : 1;
}

// the device book of K grouped books: their equities in order, an identity
// correlation (the factors are shared, not correlated), no contracts, the
// discount curve they all have
static void MergeBooks( const std::vector<const BookArrays*>& in, const std::vector<thoth_mcl_book>& books, int first,
                        BookArrays& M )
{
    const int K = (int)in.size();
    const thoth_mcl_book& b0 = books[first];
    const int nd = b0.n_dates, P = b0.lev_points;
    for ( int a = 0; a < K; a++ )
    {
        const BookArrays& A = *in[a];
        M.model.push_back( A.model[0] );
        M.spot.push_back( A.spot[0] );
        M.vol.push_back( A.vol[0] );
        M.v0.push_back( A.v0[0] );
        M.kappa.push_back( A.kappa[0] );
        M.theta.push_back( A.theta[0] );
        M.xi.push_back( A.xi[0] );
        M.rho.push_back( A.rho[0] );
        M.drift.insert( M.drift.end(), A.drift.begin(), A.drift.begin() + nd );
        M.lev_ln_step.insert( M.lev_ln_step.end(), A.lev_ln_step.begin(), A.lev_ln_step.begin() + nd );
        M.lev_offset.insert( M.lev_offset.end(), A.lev_offset.begin(), A.lev_offset.begin() + nd );
        M.lev_values.insert( M.lev_values.end(), A.lev_values.begin(), A.lev_values.begin() + (size_t)nd * P );
    }
    M.chol.assign( (size_t)K * K, 0.0 );
    for ( int a = 0; a < K; a++ )
        M.chol[(size_t)a * K + a] = 1.0;
    M.disc = in[0]->disc;
    M.c_mon_off.assign( 1, 0 );
    thoth_mcl_book& b = M.book;
    b = thoth_mcl_book{};
    b.n_dates = nd;
    b.t = b0.t;
    b.dt = b0.dt;
    b.n_assets = K;
    b.model = M.model.data();
    b.spot = M.spot.data();
    b.vol = M.vol.data();
    b.v0 = M.v0.data();
    b.kappa = M.kappa.data();
    b.theta = M.theta.data();
    b.xi = M.xi.data();
    b.rho = M.rho.data();
    b.drift = M.drift.data();
    b.chol = M.chol.data();
    b.lev_points = P;
    b.lev_ln_step = M.lev_ln_step.data();
    b.lev_offset = M.lev_offset.data();
    b.lev_values = M.lev_values.data();
    b.disc_rate = M.disc.data();
    b.c_mon_offset = M.c_mon_off.data();
}

int PrepareGridPlan( int n, const thoth_mcl_request* reqs, const thoth_vanilla_grid* grids, bool autotune,
                     GridPlan& P, bool allow_grouping )
{
    if ( n < 1 || !reqs || !grids )
        throw Invalid( THOTH_ERR_INVALID, "null argument" );
    const thoth_mcl_request& r0 = reqs[0];
    for ( int r = 0; r < n; r++ )
        for ( int c = 0; reqs[r].contracts && c < reqs[r].n_contracts; c++ )
            if ( reqs[r].contracts[c].american )
                throw Invalid( THOTH_ERR_UNSUPPORTED, "vanilla grids price european exercise" );
    P.n_req = n;
    P.paths = r0.mcl.paths;
    std::vector<BookArrays> arrays;
    arrays.reserve( 3 * (size_t)n );  // the books point into these arrays and into the grids
    std::vector<Grid> grids_alive;
    grids_alive.reserve( n );
    std::vector<GridSpec> specs;
    P.base_book.assign( n, -1 );
    P.delta_book.assign( n, -1 );
    P.vega_book.assign( n, -1 );
    P.delta_h.assign( n, 0.0 );
    P.nm.assign( n, 0 );
    P.nk.assign( n, 0 );
    P.df.assign( n, {} );
    for ( int r = 0; r < n; r++ )
    {
        const thoth_mcl_request& q = reqs[r];
        const thoth_vanilla_grid& vg = grids[r];
        if ( q.mcl.paths != r0.mcl.paths || q.mcl.sobol_skip != r0.mcl.sobol_skip || q.precision != r0.precision )
            throw Invalid( THOTH_ERR_INVALID, "the books of one grid plan share paths, sobol_skip and precision" );
        if ( q.mcl.paths <= 1 || q.mcl.paths > 200000000 )
            throw Invalid( THOTH_ERR_INVALID, "paths must be in (1, 2e8]" );
        if ( !q.mcl.use_sobol )
            throw Invalid( THOTH_ERR_UNSUPPORTED, "the device engine runs the Sobol + Brownian-bridge mode only" );
        if ( q.n_assets < 1 || q.n_assets > THOTH_MCL_MAX_ASSETS || !q.assets )
            throw Invalid( THOTH_ERR_INVALID, "1..8 equities" );
        if ( q.n_contracts < 0 || q.n_contracts > THOTH_MCL_MAX_CONTRACTS || ( q.n_contracts > 0 && !q.contracts ) )
            throw Invalid( THOTH_ERR_UNSUPPORTED, "at most 32 contracts per book" );
        if ( vg.n_maturities < 1 || vg.n_maturities > THOTH_GRID_MAX_MATURITIES || vg.n_strikes < 1 ||
             vg.n_strikes > THOTH_GRID_MAX_STRIKES || vg.underlying < 0 || vg.underlying >= q.n_assets )
            throw Invalid( THOTH_ERR_INVALID, "vanilla grid: 1..16 maturities x 1..64 strikes on an equity" );
        std::vector<double> days;
        for ( int j = 0; j < vg.n_maturities; j++ )
        {
            if ( vg.maturity_days[j] <= 0 || ( j > 0 && vg.maturity_days[j] <= vg.maturity_days[j - 1] ) )
                throw Invalid( THOTH_ERR_INVALID, "grid maturities must increase" );
            days.push_back( vg.maturity_days[j] );
        }
        const Market base = BuildMarket( q );
        grids_alive.push_back( BuildGrid( q, base, &days ) );
        const Grid& g = grids_alive.back();
        const std::vector<double> chol = SpotCholesky( q, base );
        const int na = q.n_assets;
        GridSpec gs;
        gs.nm = vg.n_maturities;
        gs.nk = vg.n_strikes;
        gs.udl = vg.underlying;
        const double spot = base.eq[vg.underlying].spot;
        for ( int k = 0; k < vg.n_strikes; k++ )
            gs.strikes[k] = vg.is_absolute_strike ? vg.strikes[k] : vg.strikes[k] * spot / 100.0;  // vanilla.cpp
        for ( int j = 0; j < vg.n_maturities; j++ )
        {
            const auto it = std::find( g.days.begin(), g.days.end(), days[j] );
            gs.mat_date[j] = (int)( it - g.days.begin() );
            const double t = g.t[gs.mat_date[j]];
            P.df[r].push_back( std::exp( -t * base.rate.Value( days[j] ) ) );  // curve.cpp:45-73 zero rate
        }
        P.nm[r] = gs.nm;
        P.nk[r] = gs.nk;
        // scenarios: base, spot + 1 % of the grid underlying, every vol + 1 point
        std::vector<Market> mks{ base };
        std::vector<int> kinds{ 0 };
        if ( q.request_greeks & THOTH_GREEK_DELTA )
        {
            Market m = base;
            P.delta_h[r] = kSpotBump * spot;
            m.eq[vg.underlying].spot = spot + P.delta_h[r];
            mks.push_back( m );
            kinds.push_back( 1 );
        }
        if ( q.request_greeks & THOTH_GREEK_VEGA )
        {
            Market m = base;
            for ( Equity& e : m.eq )
                e.vol_shift += kVolBump;
            mks.push_back( m );
            kinds.push_back( 4 );
        }
        for ( size_t k = 0; k < mks.size(); k++ )
        {
            std::vector<LeverageGrid> lev( na );
            for ( int a = 0; a < na; a++ )
                if ( mks[k].eq[a].kind == THOTH_VOL_LSV )
                {
                    const int rc = CalibrateLeverage( mks[k], a, g, lev[a] );
                    if ( rc != THOTH_OK )
                        return rc;
                }
            arrays.emplace_back();
            Resolve( q, mks[k], g, chol, lev, arrays.back() );
            const int b = (int)arrays.size() - 1;
            ( kinds[k] == 0 ? P.base_book : kinds[k] == 1 ? P.delta_book : P.vega_book )[r] = b;
            specs.push_back( gs );
        }
    }
    std::vector<thoth_mcl_book> books;
    for ( const BookArrays& A : arrays )
        books.push_back( A.book );
    // grouping of the books into device books of K equities on shared factors
    const int K = allow_grouping ? GroupSize( books, specs, r0.precision ) : 1;
    P.books_per_lane = K;
    P.group_of.assign( books.size(), 0 );
    P.slot_of.assign( books.size(), 0 );
    std::vector<BookArrays> merged( K > 1 ? books.size() / K : 0 );
    if ( K > 1 )
    {
        std::vector<thoth_mcl_book> dev;
        std::vector<GridSpec> dev_specs;
        for ( size_t g = 0; g < merged.size(); g++ )
        {
            std::vector<const BookArrays*> in;
            GridSpec gs = specs[g * K];
            gs.grouped = 1;
            gs.udl = 0;
            for ( int a = 0; a < K; a++ )
            {
                const size_t b = g * K + a;
                in.push_back( &arrays[b] );
                std::copy( specs[b].strikes, specs[b].strikes + specs[b].nk, gs.strikes_by_equity[a] );
                P.group_of[b] = (int)g;
                P.slot_of[b] = a;
            }
            MergeBooks( in, books, (int)( g * K ), merged[g] );
            dev.push_back( merged[g].book );
            dev_specs.push_back( gs );
        }
        books.swap( dev );
        specs.swap( dev_specs );
    }
    else
        for ( size_t b = 0; b < books.size(); b++ )
            P.group_of[b] = (int)b;
    thoth_mcl_sweep_cfg cfg{};
    cfg.paths = r0.mcl.paths;
    cfg.sobol_skip = r0.mcl.sobol_skip;
    cfg.precision = r0.precision;
    int rc = PrepareBatch( (int32_t)books.size(), books.data(), &cfg, P.batch, nullptr, specs.data() );
    if ( rc != THOTH_OK )
        return rc;
    // work items, longest books first (steps x equities), Sobol blocks in order
    std::vector<int> order( books.size() );
    for ( size_t b = 0; b < order.size(); b++ )
        order[b] = (int)b;
    std::stable_sort( order.begin(), order.end(), [&]( int a, int b ) {
        return (long)( books[a].n_dates - 1 ) * books[a].n_assets > (long)( books[b].n_dates - 1 ) * books[b].n_assets;
    } );
    std::vector<int2> items;
    for ( int b : order )
        for ( uint32_t t = 0; t < P.batch.nblocks; t++ )
            items.push_back( make_int2( b, (int)t ) );
    P.n_items = (int)items.size();
    HIP_OK( hipMalloc( &P.ditems, items.size() * sizeof( int2 ) ) );
    HIP_OK( hipMalloc( &P.dcounter, sizeof( int ) ) );
    hipStream_t s = Stream();
    HIP_OK( hipMemcpyAsync( P.ditems, items.data(), items.size() * sizeof( int2 ), hipMemcpyHostToDevice, s ) );
    HIP_OK( hipStreamSynchronize( s ) );
    hipDeviceProp_t prop;
    HIP_OK( hipGetDeviceProperties( &prop, 0 ) );
    P.cus = prop.multiProcessorCount;
    if ( autotune )
    {
        if ( ( rc = Autotune( P ) ) != THOTH_OK )
            return rc;
    }
    else if ( ( rc = SetLevGroup( P.batch, P.lev_group ) ) != THOTH_OK )
        return rc;
    return Capture( P );
}

int RunGridPlan( GridPlan& P, thoth_grid_result* res, double* device_ms )
{
    hipStream_t s = Stream();
    hipEvent_t e0, e1;
    HIP_OK( hipEventCreate( &e0 ) );
    HIP_OK( hipEventCreate( &e1 ) );
    HIP_OK( hipEventRecord( e0, s ) );
    HIP_OK( hipGraphLaunch( P.exec, s ) );
    HIP_OK( hipEventRecord( e1, s ) );
    HIP_OK( hipEventSynchronize( e1 ) );
    float f = 0;
    HIP_OK( hipEventElapsedTime( &f, e0, e1 ) );
    (void)hipEventDestroy( e0 );
    (void)hipEventDestroy( e1 );
    if ( device_ms )
        *device_ms = f;
    std::vector<std::vector<double>> dsums;
    int rc = CopyGrids( P.batch, dsums );
    if ( rc != THOTH_OK )
        return rc;
    // per original book: its device book's sums, or its equity's slice of them
    std::vector<std::vector<double>> sums( P.group_of.size() );
    for ( size_t b = 0; b < sums.size(); b++ )
    {
        const std::vector<double>& d = dsums[P.group_of[b]];
        const size_t per = d.size() / (size_t)P.books_per_lane;
        sums[b].assign( d.begin() + P.slot_of[b] * per, d.begin() + ( P.slot_of[b] + 1 ) * per );
    }
    const double n = (double)P.paths;
    for ( int r = 0; r < P.n_req; r++ )
    {
        thoth_grid_result& o = res[r];
        const int nk = P.nk[r];
        auto mean = [&]( int b, int j, int c ) { return sums[b][( j * 2 * nk + c ) * 2] / n * P.df[r][j]; };
        for ( int j = 0; j < P.nm[r]; j++ )
            for ( int k = 0; k < nk; k++ )
                for ( int put = 0; put < 2; put++ )
                {
                    const int c = 2 * k + put, cell = j * nk + k;
                    const int b0 = P.base_book[r];
                    const double s1 = sums[b0][( j * 2 * nk + c ) * 2], s2 = sums[b0][( j * 2 * nk + c ) * 2 + 1];
                    const double df = P.df[r][j];
                    const double p = s1 / n * df;
                    // monte_carlo_node.cpp:35-40 on the discounted flows
                    const double tr = df * std::sqrt( std::max( 0.0, ( s2 - s1 * s1 / n ) / ( n - 1 ) / n ) );
                    double* vp = put ? o.put : o.call;
                    double* vt = put ? o.put_trust : o.call_trust;
                    double* vd = put ? o.put_delta : o.call_delta;
                    double* vv = put ? o.put_vega : o.call_vega;
                    if ( vp )
                        vp[cell] = p;
                    if ( vt )
                        vt[cell] = tr;
                    if ( vd )
                        vd[cell] = P.delta_book[r] >= 0 ? ( mean( P.delta_book[r], j, c ) - p ) / P.delta_h[r] : 0.0;
                    if ( vv )  // (P(vol + 1pt) - P) / 0.01 * 0.01 (pricer.cpp reporting per point)
                        vv[cell] = P.vega_book[r] >= 0 ? ( mean( P.vega_book[r], j, c ) - p ) / kVolBump * 0.01 : 0.0;
                }
    }
    return THOTH_OK;
}

}  // namespace host
}  // namespace thoth

using namespace thoth;

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_grid_prepare_ex( int32_t n,
                                                                               const thoth_mcl_request* reqs,
                                                                               const thoth_vanilla_grid* grids,
                                                                               int32_t flags, thoth_grid_plan* plan )
{
    if ( !plan )
        return SetError( THOTH_ERR_INVALID, "null argument" );
    *plan = nullptr;
    std::lock_guard<std::mutex> lk( DeviceMutex() );
    try
    {
        int rc = EnsureDevice();
        if ( rc != THOTH_OK )
            return rc;
        const bool autotune = ( flags & THOTH_GRID_AUTOTUNE ) != 0;
        auto* P = new GridPlan;
        rc = host::PrepareGridPlan( n, reqs, grids, autotune, *P, ( flags & THOTH_GRID_SEPARATE_BOOKS ) == 0 );
        if ( rc == THOTH_ERR_UNSUPPORTED && P->books_per_lane > 1 )
        {
            // no grouped form for these books (bridge depth, draws past the
            // Sobol table): one device book per book
            delete P;
            P = new GridPlan;
            rc = host::PrepareGridPlan( n, reqs, grids, autotune, *P, false );
        }
        if ( rc != THOTH_OK )
        {
            delete P;
            return rc;
        }
        *plan = reinterpret_cast<thoth_grid_plan>( P );
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

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_grid_prepare( int32_t n,
                                                                            const thoth_mcl_request* reqs,
                                                                            const thoth_vanilla_grid* grids,
                                                                            int32_t autotune, thoth_grid_plan* plan )
{
    return thoth_grid_prepare_ex( n, reqs, grids, autotune ? THOTH_GRID_AUTOTUNE : 0, plan );
}

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_grid_run( thoth_grid_plan plan,
                                                                        thoth_grid_result* results,
                                                                        double* device_ms )
{
    if ( !plan || !results )
        return SetError( THOTH_ERR_INVALID, "null argument" );
    std::lock_guard<std::mutex> lk( DeviceMutex() );
    return host::RunGridPlan( *reinterpret_cast<GridPlan*>( plan ), results, device_ms );
}

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_grid_get_tuning( thoth_grid_plan plan,
                                                                               thoth_grid_tuning* tuning )
{
    if ( !plan || !tuning )
        return SetError( THOTH_ERR_INVALID, "null argument" );
    const GridPlan& P = *reinterpret_cast<GridPlan*>( plan );
    tuning->blocks_per_cu = P.blocks_per_cu;
    tuning->lev_group = P.lev_group;
    tuning->device_ms = P.tune_ms;
    tuning->candidates = P.candidates;
    tuning->books_per_lane = P.books_per_lane;
    return THOTH_OK;
}

extern "C" __attribute__( ( visibility( "default" ) ) ) void thoth_grid_release( thoth_grid_plan plan )
{
    std::lock_guard<std::mutex> lk( DeviceMutex() );
    delete reinterpret_cast<GridPlan*>( plan );
}
