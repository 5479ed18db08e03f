// mcl_host.cpp -- C++23 host engine behind thoth_mcl_price (include/thoth_mcl.h).
//
// It plays the role of PricerMCL::PriceBook / ComputeGreeks for GPU-eligible
// books: builds the diffusion grid (pricer_mcl.cpp:71-117), resolves the market
// into per-date drift / discount arrays, the Cholesky factor (maths.cpp:354-381),
// the LSV leverage layers (FP64, single.cpp:400-473), the payoffs (vanilla.cpp,
// barrier.cpp:122-173, rainbow.cpp:46-61), then sweeps the base book and every
// bump scenario of the reference's Greeks (pricer.cpp:273-391) in ONE device
// launch on common Sobol points, and finite-differences the results exactly as
// the reference does.
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstring>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "mcl_hostlib.hpp"
#include "mcl_launch.hpp"
#include "mcl_internal.hpp"
#include "thoth_mcl.h"

#define HIP_OK( expr )                                                                                      \
    do                                                                                                      \
    {                                                                                                       \
        hipError_t e_ = ( expr );                                                                           \
        if ( e_ != hipSuccess )                                                                             \
            return SetError( THOTH_ERR_DEVICE, std::string( #expr " failed: " ) + hipGetErrorString( e_ ) ); \
    } while ( 0 )

namespace thoth::host
{

namespace
{

// barrier.cpp:56-69
std::set<double> MonitoringDates( const thoth_contract& c )
{
    std::set<double> s;
    if ( c.monitoring_period_days > 0 )
        for ( double d = c.monitoring_period_days; d < c.maturity_days; d += c.monitoring_period_days )
            s.insert( d );
    s.insert( c.maturity_days );
    return s;
}


}  // namespace

// pricer_mcl.cpp:71-117 (+ vanilla/barrier GetFixingDates); uniform-steps extension
Grid BuildGrid( const thoth_mcl_request& req, const Market& mk, const std::vector<double>* extra_days )
{
    Grid g;
    std::set<double> fixing;
    if ( extra_days )
        for ( double d : *extra_days )
        {
            if ( d <= 0 )
                throw Invalid( THOTH_ERR_INVALID, "fixing dates must be after today" );
            fixing.insert( d );
        }
    if ( fixing.empty() && req.n_contracts < 1 )
        throw Invalid( THOTH_ERR_INVALID, "nothing to price" );
    for ( int c = 0; c < req.n_contracts; c++ )
    {
        const thoth_contract& k = req.contracts[c];
        if ( k.maturity_days <= 0 )
            throw Invalid( THOTH_ERR_INVALID, "contract maturity must be after today" );
        fixing.insert( k.maturity_days );
        if ( k.kind == THOTH_BARRIER && k.discrete_monitoring )
        {
            const std::set<double> m = MonitoringDates( k );
            fixing.insert( m.begin(), m.end() );
        }
    }
    const double last = *fixing.rbegin();
    if ( req.uniform_steps > 0 )
    {
        if ( extra_days || req.n_contracts < 1 )
            throw Invalid( THOTH_ERR_UNSUPPORTED, "uniform grid needs every contract at the last date" );
        for ( int c = 0; c < req.n_contracts; c++ )
            if ( req.contracts[c].maturity_days != last )
                throw Invalid( THOTH_ERR_UNSUPPORTED, "uniform grid needs every contract at the last date" );
        g.uniform = true;
        const double T = last / kDaysPerYear;
        for ( int i = 0; i <= req.uniform_steps; i++ )
        {
            const double ti = ( i == req.uniform_steps ) ? T : T * i / req.uniform_steps;
            g.t.push_back( ti );
            g.days.push_back( i == req.uniform_steps ? last : ti * kDaysPerYear );
            g.dt.push_back( i == 0 ? 0.0 : ti - g.t[i - 1] );
        }
        return g;
    }
    if ( req.mcl.max_day_step <= 0 )
        throw Invalid( THOTH_ERR_INVALID, "max_day_step must be > 0" );
    long step = req.mcl.max_day_step;
    if ( req.mcl.vol_year_step > 0 )
    {
        bool stoch = false;
        for ( const Equity& e : mk.eq )
            stoch |= e.kind != THOTH_VOL_BS;
        if ( stoch )
            step = std::min( step, std::max<long>( 1, (long)std::floor( req.mcl.vol_year_step * kDaysPerYear ) ) );
    }
    std::set<double> dates;
    for ( double d = 0; d < last; d += step )
        dates.insert( d );
    dates.insert( fixing.begin(), fixing.end() );
    g.days.assign( dates.begin(), dates.end() );
    for ( size_t i = 0; i < g.days.size(); i++ )
    {
        g.t.push_back( YearFraction( g.days[0], g.days[i] ) );
        g.dt.push_back( i > 0 ? YearFraction( g.days[i - 1], g.days[i] ) : 0.0 );
    }
    return g;
}

void Resolve( const thoth_mcl_request& req, const Market& mk, const Grid& g, const std::vector<double>& chol,
              const std::vector<LeverageGrid>& lev, BookArrays& A )
{
    const int n = req.n_assets;
    const int nd = (int)g.t.size();
    for ( int a = 0; a < n; a++ )
    {
        const Equity& e = mk.eq[a];
        A.model.push_back( e.kind == THOTH_VOL_BS ? THOTH_MODEL_BS
                                                  : e.kind == THOTH_VOL_HESTON ? THOTH_MODEL_HESTON : THOTH_MODEL_LSV );
        A.spot.push_back( e.spot - 0.0 );
        A.vol.push_back( ( e.bs_vol + e.vol_shift ) * 1.0 );  // bs_volatility.cpp:46-47
        const StochParams h = e.kind == THOTH_VOL_BS ? StochParams{} : mk.Stoch( a );
        A.v0.push_back( h.v0 );
        A.kappa.push_back( h.kappa );
        A.theta.push_back( h.theta );
        A.xi.push_back( h.xi );
        A.rho.push_back( h.rho );
        for ( int i = 0; i < nd; i++ )
            A.drift.push_back( mk.Drift( a, g.days[i] ) );
    }
    A.chol = chol;
    if ( n > 1 && mk.TermStructured() )
    {
        // one factor per date, of the matrix averaged over the step into it
        // (date 0: the instantaneous matrix at t_0), correlation.cpp:860-903
        A.chol_dates.assign( (size_t)nd * n * n, 0.0 );
        std::vector<double> step( (size_t)n * n );
        for ( int d = 0; d < nd; d++ )
        {
            const double t1 = g.t[d > 0 ? d - 1 : 0], t2 = g.t[d];
            for ( int i = 0; i < n; i++ )
                for ( int j = 0; j < n; j++ )
                    step[i * n + j] = i == j ? 1.0 : mk.AvgCorrelation( i, j, t1, t2 );
            if ( !CholeskyLower( step, n ) )
                throw Invalid( THOTH_ERR_INVALID, "step-average correlation matrix at date index " +
                                                      std::to_string( d ) + " is not positive-definite" );
            for ( int i = 0; i < n; i++ )
                for ( int j = 0; j <= i; j++ )
                    A.chol_dates[( (size_t)d * n + i ) * n + j] = step[i * n + j];
        }
    }
    // the flows are discounted on the OIS curve of a multi-curve currency
    // (contract.cpp:74 over currency.cpp:129-142), else on the rate curve
    for ( int i = 0; i < nd; i++ )
        A.disc.push_back( mk.multi_curve ? mk.discount.Value( g.days[i] ) : mk.rate.Value( g.days[i] ) );
    bool any_lsv = false;
    for ( int a = 0; a < n; a++ )
        any_lsv |= mk.eq[a].kind == THOTH_VOL_LSV;
    int P = 0;
    if ( any_lsv )
    {
        for ( const LeverageGrid& L : lev )
            P = std::max( P, L.points );
        A.lev_ln_step.assign( (size_t)n * nd, 1.0 );
        A.lev_offset.assign( (size_t)n * nd, 0 );
        A.lev_values.assign( (size_t)n * nd * P, 1.0 );
        for ( int a = 0; a < n; a++ )
        {
            if ( mk.eq[a].kind != THOTH_VOL_LSV )
                continue;
            const LeverageGrid& L = lev[a];
            for ( int i = 0; i < nd; i++ )
            {
                A.lev_ln_step[(size_t)a * nd + i] = L.ln_step[i];
                A.lev_offset[(size_t)a * nd + i] = L.offset[i];
                std::copy( L.values.begin() + (size_t)i * L.points, L.values.begin() + (size_t)( i + 1 ) * L.points,
                           A.lev_values.begin() + ( (size_t)a * nd + i ) * P );
            }
        }
    }
    for ( int r = 0; r < req.n_rainbows; r++ )
    {
        const thoth_rainbow& rb = req.rainbows[r];
        if ( rb.n < 1 || rb.n > THOTH_MCL_MAX_ASSETS )
            throw Invalid( THOTH_ERR_INVALID, "rainbow size" );
        A.rb_best.push_back( rb.best ? 1 : 0 );
        A.rb_n.push_back( rb.n );
        for ( int j = 0; j < THOTH_MCL_MAX_ASSETS; j++ )
        {
            const int m = j < rb.n ? rb.members[j] : 0;
            if ( m < 0 || m >= n )
                throw Invalid( THOTH_ERR_INVALID, "rainbow member index" );
            A.rb_member.push_back( m );
            // rainbow.cpp:56-57: 100 / reference spot captured at load (the unbumped
            // spot, or the caller's rebasing spot)
            const double s0 = req.rebase_spots ? req.rebase_spots[m] : req.assets[m].spot;
            if ( j < rb.n && !( s0 > 0 ) )
                throw Invalid( THOTH_ERR_INVALID, "rainbow rebasing spot must be positive" );
            A.rb_weight.push_back( j < rb.n ? 100.0 / s0 : 0.0 );
        }
    }
    auto index_of = [&]( double day )
    {
        for ( int i = 0; i < nd; i++ )
            if ( g.days[i] == day )
                return i;
        throw Invalid( THOTH_ERR_INVALID, "date not on the diffusion grid" );
    };
    A.c_mon_off.push_back( 0 );
    for ( int c = 0; c < req.n_contracts; c++ )
    {
        const thoth_contract& k = req.contracts[c];
        if ( k.underlying >= n || k.underlying < -req.n_rainbows )
            throw Invalid( THOTH_ERR_INVALID, "contract underlying index" );
        const int fi = index_of( k.maturity_days );
        // vanilla.cpp:29-34 / barrier.cpp:36-41: a relative strike resolves against the base spot
        const double base_spot = k.underlying >= 0 ? req.assets[k.underlying].spot : 100.0;
        const double K = k.is_absolute_strike ? k.strike : k.strike / 100 * base_spot;
        A.c_kind.push_back( k.kind == THOTH_BARRIER ? THOTH_PAYOFF_BARRIER : THOTH_PAYOFF_VANILLA );
        A.c_call.push_back( k.is_call ? 1 : 0 );
        A.c_udl.push_back( k.underlying );
        A.c_flow.push_back( fi );
        A.c_strike.push_back( K );
        double H = 0;
        bool up = false, in = false;
        if ( k.kind == THOTH_BARRIER )
        {
            up = k.barrier_type == THOTH_UP_OUT || k.barrier_type == THOTH_UP_IN;
            in = k.barrier_type == THOTH_UP_IN || k.barrier_type == THOTH_DOWN_IN;
            H = up ? k.barrier_up_level : k.barrier_down_level;
            if ( k.discrete_monitoring && !g.uniform )
            {
                for ( double d : MonitoringDates( k ) )
                    A.c_mon.push_back( index_of( d ) );
            }
            else if ( k.discrete_monitoring )
            {
                for ( int i = 1; i <= fi; i++ )
                    A.c_mon.push_back( i );
            }
            else
            {
                // barrier.cpp:146-161: every diffusion date up to maturity + BGK level shift
                size_t cnt = 0;
                for ( int i = 0; i < nd; i++ )
                    if ( g.days[i] <= k.maturity_days )
                    {
                        A.c_mon.push_back( i );
                        cnt++;
                    }
                if ( k.underlying < 0 )
                    throw Invalid( THOTH_ERR_UNSUPPORTED, "continuous barrier on a rainbow has no implied vol" );
                const double t = YearFraction( 0, k.maturity_days );
                const size_t steps = cnt > 1 ? cnt - 1 : 1;
                const double dt = t / (double)steps;
                const double vol = mk.ImplicitVol( k.underlying, K, k.maturity_days );
                const double beta = 0.5826;
                H *= std::exp( ( up ? -1.0 : 1.0 ) * beta * vol * std::sqrt( dt ) );
            }
        }
        A.c_up.push_back( up ? 1 : 0 );
        A.c_in.push_back( in ? 1 : 0 );
        A.c_level.push_back( H );
        A.c_mon_off.push_back( (int32_t)A.c_mon.size() );
    }
    thoth_mcl_book& b = A.book;
    b.n_dates = nd;
    b.t = g.t.data();
    b.dt = g.dt.data();
    b.n_assets = n;
    b.model = A.model.data();
    b.spot = A.spot.data();
    b.vol = A.vol.data();
    b.v0 = A.v0.data();
    b.kappa = A.kappa.data();
    b.theta = A.theta.data();
    b.xi = A.xi.data();
    b.rho = A.rho.data();
    b.drift = A.drift.data();
    b.chol = A.chol.empty() ? nullptr : A.chol.data();
    b.chol_dates = A.chol_dates.empty() ? nullptr : A.chol_dates.data();
    b.correlate_stochastic = req.correlate_stochastic;
    b.lev_points = P;
    b.lev_ln_step = A.lev_ln_step.empty() ? nullptr : A.lev_ln_step.data();
    b.lev_offset = A.lev_offset.empty() ? nullptr : A.lev_offset.data();
    b.lev_values = A.lev_values.empty() ? nullptr : A.lev_values.data();
    b.disc_rate = A.disc.data();
    b.n_rainbows = req.n_rainbows;
    b.rb_best = A.rb_best.data();
    b.rb_n = A.rb_n.data();
    b.rb_member = A.rb_member.data();
    b.rb_weight = A.rb_weight.data();
    b.n_contracts = req.n_contracts;
    b.c_kind = A.c_kind.data();
    b.c_is_call = A.c_call.data();
    b.c_underlying = A.c_udl.data();
    b.c_flow_index = A.c_flow.data();
    b.c_strike = A.c_strike.data();
    b.c_is_up = A.c_up.data();
    b.c_is_in = A.c_in.data();
    b.c_level = A.c_level.data();
    b.c_mon_offset = A.c_mon_off.data();
    b.c_mon_index = A.c_mon.data();
}


// A request resolved into device-resident books (base + bump scenarios).
struct Plan
{
    DeviceBatch batch;
    std::vector<int> kind;
    std::vector<double> h;
    int n_contracts = 0, n_dates = 0;
    double calib_ms = 0;
    double tuned_ms = 0;  // best sweep time of the launch-shape tuner (0: untuned)
    // American contracts (pricer_mcl_american.cpp:137-248): Longstaff-Schwartz
    // on the equity path every scenario book records; per contract its column
    // count M (dates 0 .. flow index), strike and type
    struct AmericanLeg
    {
        int c, M, is_call;
        double strike;
    };
    std::vector<AmericanLeg> am;
    int rec_udl = -1;
    std::vector<double> t;                   // year fraction of every date
    std::vector<std::vector<double>> rates;  // per scenario: discount zero rate of every date
    std::vector<double> spot0;               // per scenario: the recorded equity's spot today
    // device scratch of the fit and the apply: beta [4 nd], df [nd], disc
    // [scenarios][nd], cf [paths], part, recorded-path pointers [scenarios]
    char* lsm_mem = nullptr;
    double *dbeta = nullptr, *ddf = nullptr, *ddisc = nullptr, *dcf = nullptr, *dpart = nullptr;
    const void** dpaths = nullptr;
    Plan() = default;
    Plan( const Plan& ) = delete;
    Plan& operator=( const Plan& ) = delete;
    ~Plan()
    {
        if ( lsm_mem )
            (void)hipFree( lsm_mem );
    }
};

namespace
{

struct Scenario
{
    Market mk;
    int kind;  // 0 base, 1 delta, 2 gamma up, 3 gamma down, 4 vega, 5 rho
    int asset;
    double h;
};

double Mean( const thoth_mcl_book_out& o, int c ) { return ( c < 0 ? o.book_sum : o.contract_sum[c] ) / (double)o.count; }

double Trust( const thoth_mcl_book_out& o, int c )
{
    const double n = (double)o.count;
    const double s = c < 0 ? o.book_sum : o.contract_sum[c];
    const double s2 = c < 0 ? o.book_sum2 : o.contract_sum2[c];
    return std::sqrt( ( s2 - s * s / n ) / ( n - 1 ) / n );  // monte_carlo_node.cpp:35-40
}

}  // namespace

int PreparePlan( const thoth_mcl_request& req, Plan& plan )
{
    // mcl_configuration.cpp:30-46 guards
    if ( req.mcl.paths <= 1 || req.mcl.paths > 200000000 )
        throw Invalid( THOTH_ERR_INVALID, "paths must be in (1, 2e8]" );
    if ( req.n_assets < 1 || req.n_assets > THOTH_MCL_MAX_ASSETS || !req.assets )
        throw Invalid( THOTH_ERR_INVALID, "1..8 equities" );
    if ( req.n_contracts < 1 || req.n_contracts > THOTH_MCL_MAX_CONTRACTS || !req.contracts )
        throw Invalid( THOTH_ERR_UNSUPPORTED, "1..32 contracts per book" );
    if ( req.precision != THOTH_FP64 && req.precision != THOTH_FP32 )
        throw Invalid( THOTH_ERR_INVALID, "precision" );
    const Market base = BuildMarket( req );
    const Grid g = BuildGrid( req, base );
    // Cholesky of the spot/spot sub-matrix (pricer_mcl.cpp:462-491, correlation.cpp:794-859)
    const std::vector<double> chol = SpotCholesky( req, base );
    const int n = req.n_assets;
    // scenarios of the bump-and-revalue Greeks (pricer.cpp:287-380)
    std::vector<Scenario> sc;
    sc.push_back( { base, 0, -1, 0 } );
    for ( int a = 0; a < n; a++ )
    {
        const double spot = base.eq[a].spot;
        if ( req.request_greeks & THOTH_GREEK_DELTA )
        {
            Scenario s{ base, 1, a, kSpotBump * spot };
            s.mk.eq[a].spot = spot + s.h;
            sc.push_back( s );
        }
        if ( req.request_greeks & THOTH_GREEK_GAMMA )
        {
            Scenario u{ base, 2, a, kGammaBump * spot }, d{ base, 3, a, kGammaBump * spot };
            u.mk.eq[a].spot = spot + u.h;
            d.mk.eq[a].spot = spot - d.h;
            sc.push_back( u );
            sc.push_back( d );
        }
    }
    if ( req.request_greeks & THOTH_GREEK_VEGA )
    {
        Scenario s{ base, 4, -1, kVolBump };
        for ( Equity& e : s.mk.eq )
            e.vol_shift += kVolBump;
        sc.push_back( s );
    }
    if ( req.request_greeks & THOTH_GREEK_RHO )
    {
        Scenario s{ base, 5, -1, kRateBump };
        s.mk.rate.shift = kRateBump;  // the projection AND the discounting curve (pricer.cpp:344-359)
        s.mk.discount.shift = kRateBump;
        sc.push_back( s );
    }
    // leverage calibration per scenario (each scenario rebuilds the leverage node, single.cpp:479-500)
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::vector<LeverageGrid>> levs( sc.size(), std::vector<LeverageGrid>( n ) );
    for ( size_t k = 0; k < sc.size(); k++ )
        for ( int a = 0; a < n; a++ )
            if ( sc[k].mk.eq[a].kind == THOTH_VOL_LSV )
            {
                int rc = CalibrateLeverage( sc[k].mk, a, g, levs[k][a] );
                if ( rc != THOTH_OK )
                    return rc;
            }
    const double calib_ms = std::chrono::duration<double, std::milli>( std::chrono::steady_clock::now() - t0 ).count();
    std::vector<BookArrays> arrays( sc.size() );
    std::vector<thoth_mcl_book> books;
    for ( size_t k = 0; k < sc.size(); k++ )
    {
        Resolve( req, sc[k].mk, g, chol, levs[k], arrays[k] );
        books.push_back( arrays[k].book );
    }
    // American contracts: vanilla on one equity per book, which every scenario
    // book records at dates 0 .. the latest American flow (pricer_mcl_american.cpp:59-94)
    plan.am.clear();
    plan.rec_udl = -1;
    int rec_n = 0;
    for ( int c = 0; c < req.n_contracts; c++ )
    {
        const thoth_contract& k = req.contracts[c];
        if ( !k.american )
            continue;
        if ( k.kind != THOTH_VANILLA || k.underlying < 0 )
            throw Invalid( THOTH_ERR_UNSUPPORTED, "american exercise: vanilla contracts on an equity" );
        if ( plan.rec_udl >= 0 && k.underlying != plan.rec_udl )
            throw Invalid( THOTH_ERR_UNSUPPORTED, "american contracts of one book exercise on one equity" );
        plan.rec_udl = k.underlying;
        const int M = arrays[0].c_flow[c] + 1;
        rec_n = std::max( rec_n, M );
        plan.am.push_back( { c, M, k.is_call ? 1 : 0, arrays[0].c_strike[c] } );
    }
    std::vector<RecordSpec> rec( sc.size() );
    if ( !plan.am.empty() )
    {
        plan.t = g.t;
        plan.rates.assign( sc.size(), {} );
        plan.spot0.assign( sc.size(), 0.0 );
        for ( size_t k = 0; k < sc.size(); k++ )
        {
            rec[k] = { plan.rec_udl, rec_n };
            plan.rates[k] = arrays[k].disc;
            plan.spot0[k] = arrays[k].spot[plan.rec_udl];
        }
    }
    thoth_mcl_sweep_cfg cfg{};
    cfg.paths = req.mcl.paths;
    cfg.sobol_skip = req.mcl.sobol_skip;
    cfg.precision = req.precision;
    // the pseudo-random stream: seeded by the configuration's int seed as
    // Rng::Seed((uint64_t)seed) (pricer_mcl.cpp:227)
    cfg.pseudo_random = req.mcl.use_sobol ? 0 : 1;
    cfg.seed = (uint64_t)(int64_t)req.mcl.seed;
    int rc = PrepareBatch( (int32_t)books.size(), books.data(), &cfg, plan.batch, nullptr, nullptr, false, nullptr, 0,
                           plan.am.empty() ? nullptr : rec.data() );
    if ( rc != THOTH_OK )
        return rc;
    if ( !plan.am.empty() )
    {
        const size_t K = sc.size(), nd = g.t.size(), N = (size_t)req.mcl.paths;
        const size_t nblk = (size_t)lsm_blocks( req.mcl.paths );
        const size_t part = std::max( nblk * 8, K * nblk * 2 );
        const size_t doubles = 4 * nd + nd + K * nd + N + part;
        const size_t bytes = doubles * sizeof( double ) + K * sizeof( void* );
        hipError_t e = hipMalloc( &plan.lsm_mem, bytes );
        if ( e != hipSuccess )
            return SetError( THOTH_ERR_DEVICE, std::string( "hipMalloc: " ) + hipGetErrorString( e ) );
        double* d = reinterpret_cast<double*>( plan.lsm_mem );
        plan.dbeta = d;
        plan.ddf = d + 4 * nd;
        plan.ddisc = plan.ddf + nd;
        plan.dcf = plan.ddisc + K * nd;
        plan.dpart = plan.dcf + N;
        plan.dpaths = reinterpret_cast<const void**>( plan.dpart + part );
        std::vector<const void*> ptrs( K );
        for ( size_t k = 0; k < K; k++ )
            ptrs[k] = req.precision == THOTH_FP64 ? static_cast<const void*>( plan.batch.db[k].rec )
                                                  : static_cast<const void*>( plan.batch.db[k].recf );
        hipStream_t s = Stream();
        if ( hipMemcpyAsync( plan.dpaths, ptrs.data(), K * sizeof( void* ), hipMemcpyHostToDevice, s ) != hipSuccess ||
             hipStreamSynchronize( s ) != hipSuccess )
            return SetError( THOTH_ERR_DEVICE, "upload of the recorded-path table" );
    }
    plan.kind.clear();
    plan.h.clear();
    for ( const Scenario& s : sc )
    {
        plan.kind.push_back( s.kind );
        plan.h.push_back( s.h );
    }
    plan.n_contracts = req.n_contracts;
    plan.n_dates = (int)g.t.size();
    plan.calib_ms = calib_ms;
    return THOTH_OK;
}

// Longstaff-Schwartz of every American contract (pricer_mcl_american.cpp:137-248):
// the policy fitted on the base scenario's recorded paths, applied frozen to
// every scenario's; its value replaces the contract's European value in the
// scenario's contract and book premiums.  Returns the device milliseconds.
int PriceAmerican( const Plan& plan, std::vector<double>& book, std::vector<std::vector<double>>& cprem,
                   std::vector<double>& ctrust, float& ms )
{
    ms = 0;
    hipStream_t s = Stream();
    const DeviceBatch& B = plan.batch;
    const int K = (int)plan.kind.size();
    const int64_t N = B.paths;
    const size_t nd = plan.t.size();
    const int nblk = lsm_blocks( N );
    const void* base_paths = B.precision == 0 ? static_cast<const void*>( B.db[0].rec )
                                              : static_cast<const void*>( B.db[0].recf );
    hipEvent_t e0 = SweepEvent( 0 ), e1 = SweepEvent( 1 );
    HIP_OK( hipEventRecord( e0, s ) );
    std::vector<double> part( (size_t)K * nblk * 2 );
    for ( const Plan::AmericanLeg& leg : plan.am )
    {
        const int M = leg.M;
        const std::vector<double>& t = plan.t;
        const std::vector<double>& r0 = plan.rates[0];
        // discount over [t, t+1] as the ratio of zero-coupon prices
        // (pricer_mcl_american.cpp:307-312), and to today per scenario (:426-439)
        std::vector<double> df( nd, 0.0 ), disc( (size_t)K * nd, 0.0 );
        for ( int i = 0; i + 1 < M; i++ )
            df[i] = std::exp( r0[i] * t[i] - r0[i + 1] * t[i + 1] );
        for ( int k = 0; k < K; k++ )
            for ( int i = 0; i < M; i++ )
                disc[(size_t)k * M + i] = std::exp( -plan.rates[k][i] * ( t[i] - t[0] ) );
        // basis normaliser: the strike (vanilla.hpp:38-41)
        const double norm = leg.strike > 0 ? leg.strike : plan.spot0[0];
        HIP_OK( hipMemcpyAsync( plan.ddf, df.data(), nd * sizeof( double ), hipMemcpyHostToDevice, s ) );
        HIP_OK( hipMemcpyAsync( plan.ddisc, disc.data(), disc.size() * sizeof( double ), hipMemcpyHostToDevice, s ) );
        HIP_OK( hipMemsetAsync( plan.dbeta, 0, 4 * nd * sizeof( double ), s ) );
        if ( !launch_lsm_fit( s, B.precision, base_paths, N, M, leg.strike, leg.is_call, norm, plan.ddf, plan.dbeta,
                              plan.dcf, plan.dpart ) ||
             !launch_lsm_apply( s, B.precision, plan.dpaths, K, N, M, leg.strike, leg.is_call, norm, plan.ddisc,
                                plan.dbeta, plan.dpart ) )
            return SetError( THOTH_ERR_DEVICE, "Longstaff-Schwartz launch failed" );
        HIP_OK( hipMemcpyAsync( part.data(), plan.dpart, part.size() * sizeof( double ), hipMemcpyDeviceToHost, s ) );
        HIP_OK( hipStreamSynchronize( s ) );
        for ( int k = 0; k < K; k++ )
        {
            double sum = 0, sum2 = 0;
            for ( int b = 0; b < nblk; b++ )
            {
                sum += part[( (size_t)k * nblk + b ) * 2];
                sum2 += part[( (size_t)k * nblk + b ) * 2 + 1];
            }
            const double mean = sum / (double)N;
            const double trust = std::sqrt( ( sum2 / (double)N - mean * mean ) / (double)( N - 1 ) );
            const double s0 = plan.spot0[k];
            const double now = std::max( leg.is_call ? s0 - leg.strike : leg.strike - s0, 0.0 );
            const double v = std::max( mean, now );
            book[k] += v - cprem[k][leg.c];
            cprem[k][leg.c] = v;
            if ( k == 0 )
                ctrust[leg.c] = trust;
        }
    }
    HIP_OK( hipEventRecord( e1, s ) );
    HIP_OK( hipEventSynchronize( e1 ) );
    HIP_OK( hipEventElapsedTime( &ms, e0, e1 ) );
    return THOTH_OK;
}

int RunPlan( const Plan& plan, thoth_mcl_result* res )
{
    std::vector<thoth_mcl_book_out> out( plan.kind.size() );
    int rc = RunBatch( plan.batch, out.data() );
    if ( rc != THOTH_OK )
        return rc;
    // Tree_Read (pricer_mcl.cpp:677-709) + ComputeGreeks finite differences
    const int nc = plan.n_contracts;
    const size_t K = plan.kind.size();
    std::vector<double> book( K );
    std::vector<std::vector<double>> cprem( K, std::vector<double>( nc ) );
    std::vector<double> ctrust( nc );
    for ( size_t k = 0; k < K; k++ )
    {
        book[k] = Mean( out[k], -1 );
        for ( int c = 0; c < nc; c++ )
            cprem[k][c] = Mean( out[k], c );
    }
    for ( int c = 0; c < nc; c++ )
        ctrust[c] = Trust( out[0], c );
    float lsm_ms = 0;
    double p0 = book[0];
    if ( !plan.am.empty() )
    {
        if ( ( rc = PriceAmerican( plan, book, cprem, ctrust, lsm_ms ) ) != THOTH_OK )
            return rc;
        // the book premium re-aggregated from its contracts (pricer_mcl_american.cpp:251-257)
        p0 = 0;
        for ( int c = 0; c < nc; c++ )
            p0 += cprem[0][c];
    }
    res->premium = p0;
    res->premium_trust = Trust( out[0], -1 );
    res->delta = res->gamma = res->vega = res->rho = 0;
    std::vector<double> cd( nc, 0 ), cg( nc, 0 ), cv( nc, 0 ), cr( nc, 0 );
    for ( size_t k = 1; k < K; k++ )
    {
        const double h = plan.h[k];
        switch ( plan.kind[k] )
        {
        case 1:
            res->delta += ( book[k] - p0 ) / h;
            for ( int c = 0; c < nc; c++ )
                cd[c] += ( cprem[k][c] - cprem[0][c] ) / h;
            break;
        case 2:
            res->gamma += ( book[k] - 2 * p0 + book[k + 1] ) / ( h * h );
            for ( int c = 0; c < nc; c++ )
                cg[c] += ( cprem[k][c] - 2 * cprem[0][c] + cprem[k + 1][c] ) / ( h * h );
            break;
        case 4:
            res->vega = ( book[k] - p0 ) / kVolBump * 0.01;
            for ( int c = 0; c < nc; c++ )
                cv[c] = ( cprem[k][c] - cprem[0][c] ) / kVolBump * 0.01;
            break;
        case 5:
            res->rho = ( book[k] - p0 ) / kRateBump * 0.01;
            for ( int c = 0; c < nc; c++ )
                cr[c] = ( cprem[k][c] - cprem[0][c] ) / kRateBump * 0.01;
            break;
        default:
            break;
        }
    }
    for ( int c = 0; c < nc; c++ )
    {
        if ( res->contract_premium )
            res->contract_premium[c] = cprem[0][c];
        if ( res->contract_trust )
            res->contract_trust[c] = ctrust[c];
        if ( res->contract_delta )
            res->contract_delta[c] = cd[c];
        if ( res->contract_gamma )
            res->contract_gamma[c] = cg[c];
        if ( res->contract_vega )
            res->contract_vega[c] = cv[c];
        if ( res->contract_rho )
            res->contract_rho[c] = cr[c];
    }
    res->n_dates = plan.n_dates;
    res->n_scenarios = (int32_t)plan.kind.size();
    res->sweep_ms = out[0].device_ms + lsm_ms;
    res->calib_ms = plan.calib_ms;
    return THOTH_OK;
}


// base scenario only (no Greek bumps): grid, Cholesky, calibration, resolution
int ResolveBaseBook( const thoth_mcl_request& req, BookArrays& A, Grid& g )
{
    // its callers (randomised QMC, the control variate) shift Sobol points and
    // price European payoffs
    if ( !req.mcl.use_sobol )
        throw Invalid( THOTH_ERR_UNSUPPORTED, "randomised QMC and the control variate need use_sobol: true" );
    for ( int c = 0; req.contracts && c < req.n_contracts; c++ )
        if ( req.contracts[c].american )
            throw Invalid( THOTH_ERR_UNSUPPORTED, "randomised QMC and the control variate price european exercise" );
    if ( req.n_assets < 1 || req.n_assets > THOTH_MCL_MAX_ASSETS || !req.assets )
        throw Invalid( THOTH_ERR_INVALID, "1..8 equities" );
    if ( req.n_contracts < 1 || req.n_contracts > THOTH_MCL_MAX_CONTRACTS || !req.contracts )
        throw Invalid( THOTH_ERR_UNSUPPORTED, "1..32 contracts per book" );
    const Market mk = BuildMarket( req );
    g = BuildGrid( req, mk );
    const std::vector<double> chol = SpotCholesky( req, mk );
    const int n = req.n_assets;
    std::vector<LeverageGrid> lev( n );
    for ( int a = 0; a < n; a++ )
        if ( mk.eq[a].kind == THOTH_VOL_LSV )
        {
            int rc = CalibrateLeverage( mk, a, g, lev[a] );
            if ( rc != THOTH_OK )
                return rc;
        }
    Resolve( req, mk, g, chol, lev, A );
    return THOTH_OK;
}

int PriceImpl( const thoth_mcl_request& req, thoth_mcl_result* res )
{
    Plan plan;
    int rc = PreparePlan( req, plan );
    if ( rc != THOTH_OK )
        return rc;
    return RunPlan( plan, res );
}

}  // namespace thoth::host

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_mcl_price( const thoth_mcl_request* req,
                                                                         thoth_mcl_result* res )
{
    if ( !req || !res )
        return thoth::SetError( THOTH_ERR_INVALID, "null argument" );
    std::lock_guard<std::mutex> lk( thoth::DeviceMutex() );
    try
    {
        int rc = thoth::EnsureDevice();
        if ( rc != THOTH_OK )
            return rc;
        return thoth::host::PriceImpl( *req, res );
    }
    catch ( const thoth::host::Invalid& e )
    {
        return thoth::SetError( e.code, e.what() );
    }
    catch ( const std::exception& e )
    {
        return thoth::SetError( THOTH_ERR_INVALID, e.what() );
    }
}

namespace thoth
{
double StudentT975( int df )
{
    // two-sided 95 % quantiles of Student's t (df 1..30), normal beyond
    static const double q[] = { 12.706, 4.303, 3.182, 2.776, 2.571, 2.447, 2.365, 2.306, 2.262, 2.228,
                                2.201,  2.179, 2.160, 2.145, 2.131, 2.120, 2.110, 2.101, 2.093, 2.086,
                                2.080,  2.074, 2.069, 2.064, 2.060, 2.056, 2.052, 2.048, 2.045, 2.042 };
    if ( df < 1 )
        return 0.0;
    if ( df <= 30 )
        return q[df - 1];
    return 1.96 + 2.4 / df;  // first Cornish-Fisher correction
}
}  // namespace thoth

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_mcl_price_rqmc( const thoth_mcl_request* req,
                                                                              int32_t replicates, uint64_t rqmc_seed,
                                                                              thoth_rqmc_result* res )
{
    if ( !req || !res )
        return thoth::SetError( THOTH_ERR_INVALID, "null argument" );
    if ( replicates < 2 )
        return thoth::SetError( THOTH_ERR_INVALID, "at least 2 replicates are needed for a confidence interval" );
    std::lock_guard<std::mutex> lk( thoth::DeviceMutex() );
    try
    {
        int rc = thoth::EnsureDevice();
        if ( rc != THOTH_OK )
            return rc;
        const auto t0 = std::chrono::steady_clock::now();
        thoth::host::BookArrays A;
        thoth::host::Grid g;
        if ( ( rc = thoth::host::ResolveBaseBook( *req, A, g ) ) != THOTH_OK )
            return rc;
        thoth_mcl_sweep_cfg sc{};
        sc.paths = req->mcl.paths;
        sc.precision = req->precision;
        sc.replicates = replicates;
        sc.rqmc_seed = rqmc_seed;
        thoth::DeviceBatch B;
        if ( ( rc = thoth::PrepareBatch( 1, &A.book, &sc, B ) ) != THOTH_OK )
            return rc;
        const auto t1 = std::chrono::steady_clock::now();
        std::vector<thoth_mcl_book_out> out( replicates );
        if ( ( rc = thoth::RunBatch( B, out.data() ) ) != THOTH_OK )
            return rc;
        const int64_t N = req->mcl.paths;
        double mean = 0, s1 = 0, s2 = 0;
        for ( const auto& o : out )
        {
            mean += o.book_sum / N;
            s1 += o.book_sum;
            s2 += o.book_sum2;
        }
        mean /= replicates;
        double v = 0;
        for ( const auto& o : out )
            v += ( o.book_sum / N - mean ) * ( o.book_sum / N - mean );
        v /= ( replicates - 1 );
        const double n = (double)N * replicates;
        res->replicates = replicates;
        res->paths = N;
        res->mean = mean;
        res->halfwidth = thoth::StudentT975( replicates - 1 ) * std::sqrt( v / replicates );
        res->var = ( s2 - s1 * s1 / n ) / ( n - 1 );
        res->sweep_ms = out[0].device_ms;
        if ( res->replicate_means )
            for ( int r = 0; r < replicates; r++ )
                res->replicate_means[r] = out[r].book_sum / N;
        const auto t2 = std::chrono::steady_clock::now();
        res->setup_ms = std::chrono::duration<double, std::milli>( t1 - t0 ).count();
        res->run_wall_ms = std::chrono::duration<double, std::milli>( t2 - t1 ).count();
        return THOTH_OK;
    }
    catch ( const thoth::host::Invalid& e )
    {
        return thoth::SetError( e.code, e.what() );
    }
    catch ( const std::exception& e )
    {
        return thoth::SetError( THOTH_ERR_INVALID, e.what() );
    }
}

namespace thoth::host
{

namespace
{

// Run independent host tasks (each may use threads of its own) on a few
// threads; the first exception is rethrown after every task has ended.
void RunConcurrently( std::vector<std::function<void()>>& tasks )
{
    const size_t nt = std::min<size_t>( tasks.size(), std::max( 1u, std::min( 8u, std::thread::hardware_concurrency() ) ) );
    std::atomic<size_t> next{ 0 };
    std::exception_ptr err;
    std::mutex mu;
    auto worker = [&] {
        for ( size_t i; ( i = next.fetch_add( 1 ) ) < tasks.size(); )
            try
            {
                tasks[i]();
            }
            catch ( ... )
            {
                std::lock_guard<std::mutex> lk( mu );
                if ( !err )
                    err = std::current_exception();
            }
    };
    std::vector<std::thread> th;
    for ( size_t t = 1; t < nt; t++ )
        th.emplace_back( worker );
    worker();
    for ( auto& t : th )
        t.join();
    if ( err )
        std::rethrow_exception( err );
}

// Kahan sums over blocks in block order of every estimator of every book
// ([books][ne][sum, sum2]) from the per-block partials of a pathwise batch
std::vector<std::vector<double>> BlockSums( const DeviceBatch& B, const std::vector<double>& part, int nbooks, int ne )
{
    std::vector<std::vector<double>> sums( nbooks, std::vector<double>( 2 * ne, 0.0 ) );
    for ( int b = 0; b < nbooks; b++ )
    {
        std::vector<double> cp( 2 * ne, 0.0 );
        std::vector<double>& sm = sums[b];
        for ( uint32_t k = 0; k < B.nblocks; k++ )
            for ( int j = 0; j < 2 * ne; j++ )
            {
                const double y = part[B.part_off[b] + (size_t)k * 2 * ne + j] - cp[j];
                const double t = sm[j] + y;
                cp[j] = ( t - sm[j] ) - y;
                sm[j] = t;
            }
    }
    return sums;
}

// launch one pathwise batch on the library stream: device time (ms) and the partials
int RunPathwise( const DeviceBatch& B, int kernel, std::vector<double>& part, float& ms )
{
    hipStream_t s = Stream();
    hipEvent_t e0 = SweepEvent( 0 ), e1 = SweepEvent( 1 );
    (void)hipEventRecord( e0, s );
    const dim3 grid( B.nblocks, B.n_books );
    const bool ok = kernel == 1 ? launch_pathwise( B.precision, B.depth, grid, s, B.lds_bytes, B.dbooks,
                                                   DeviceDirections(), B.q0, B.first, B.last )
                                : launch_pathwise_basket( B.precision, B.na_max, B.depth, B.pw_lsv, grid, s,
                                                          B.lds_bytes, B.dbooks, DeviceDirections(), B.q0, B.first,
                                                          B.last );
    if ( !ok )
        return SetError( THOTH_ERR_UNSUPPORTED, "bridge deeper than the pathwise kernels support" );
    if ( hipGetLastError() != hipSuccess )
        return SetError( THOTH_ERR_DEVICE, "pathwise launch failed" );
    (void)hipEventRecord( e1, s );
    part.assign( B.part_total, 0.0 );
    if ( hipMemcpyAsync( part.data(), B.dpart, part.size() * sizeof( double ), hipMemcpyDeviceToHost, s ) !=
             hipSuccess ||
         hipStreamSynchronize( s ) != hipSuccess )
        return SetError( THOTH_ERR_DEVICE, "pathwise copy failed" );
    ms = 0;
    (void)hipEventElapsedTime( &ms, e0, e1 );
    return THOTH_OK;
}

}  // namespace

// single-equity BS / Heston book with one vanilla or knock-out contract:
// base and (gamma) +-10 % spot books swept by the pathwise kernel
int PricePathwiseSingle( const thoth_mcl_request& req, thoth_pathwise_result* res )
{
    const auto t0 = std::chrono::steady_clock::now();
    const Market base = BuildMarket( req );
    const Grid g = BuildGrid( req, base );
    std::vector<Market> mks{ base };
    const double spot = base.eq[0].spot, h = kGammaBump * spot;
    const bool gamma = ( req.request_greeks & THOTH_GREEK_GAMMA ) != 0;
    if ( gamma )
    {
        Market u = base, d = base;
        u.eq[0].spot = spot + h;
        d.eq[0].spot = spot - h;
        mks.push_back( u );
        mks.push_back( d );
    }
    std::vector<BookArrays> arrays( mks.size() );
    std::vector<thoth_mcl_book> books;
    for ( size_t k = 0; k < mks.size(); k++ )
    {
        Resolve( req, mks[k], g, {}, std::vector<LeverageGrid>( 1 ), arrays[k] );
        books.push_back( arrays[k].book );
    }
    thoth_mcl_sweep_cfg cfg{};
    cfg.paths = req.mcl.paths;
    cfg.sobol_skip = req.mcl.sobol_skip;
    cfg.precision = req.precision;
    DeviceBatch B;
    int rc = PrepareBatch( (int32_t)books.size(), books.data(), &cfg, B, nullptr, nullptr, 1 );
    if ( rc != THOTH_OK )
        return rc;
    const auto t1 = std::chrono::steady_clock::now();
    std::vector<double> part;
    float ms = 0;
    if ( ( rc = RunPathwise( B, 1, part, ms ) ) != THOTH_OK )
        return rc;
    // [Y, delta, vega] x [sum, sum2] per book
    const double n = (double)req.mcl.paths;
    const auto sums = BlockSums( B, part, (int)books.size(), 3 );
    auto mean = [&]( int b, int e ) { return sums[b][2 * e] / n; };
    auto trust = [&]( int b, int e ) {
        const double s1 = sums[b][2 * e], s2 = sums[b][2 * e + 1];
        return std::sqrt( std::max( 0.0, ( s2 - s1 * s1 / n ) / ( n - 1 ) / n ) );
    };
    res->premium = mean( 0, 0 );
    res->premium_trust = trust( 0, 0 );
    res->delta = mean( 0, 1 );
    res->delta_trust = trust( 0, 1 );
    res->vega = mean( 0, 2 );
    res->vega_trust = trust( 0, 2 );
    res->gamma = gamma ? ( mean( 1, 1 ) - mean( 2, 1 ) ) / ( 2 * h ) : 0.0;
    res->n_assets = 1;
    res->asset_delta[0] = res->delta;
    res->asset_delta_trust[0] = res->delta_trust;
    res->asset_gamma[0] = res->gamma;
    res->n_books = (int32_t)books.size();
    res->sweep_ms = ms;
    res->setup_ms = std::chrono::duration<double, std::milli>( t1 - t0 ).count();
    return THOTH_OK;
}

// books of 1..8 BS / Heston / LSV equities with one european vanilla on an
// equity or a rainbow: the base book and, for gamma, the +-10 % books of every
// equity (leverage recalibrated at the bumped spot, as the reference's bump
// scenario rebuilds it), swept by the pathwise basket kernel with the LSV
// leverage tangents of each book
int PricePathwiseBasket( const thoth_mcl_request& req, thoth_pathwise_result* res )
{
    const auto t0 = std::chrono::steady_clock::now();
    const int n = req.n_assets;
    const Market base = BuildMarket( req );
    const Grid g = BuildGrid( req, base );
    const std::vector<double> chol = SpotCholesky( req, base );
    const bool gamma = ( req.request_greeks & THOTH_GREEK_GAMMA ) != 0;
    std::vector<Market> mks{ base };
    std::vector<int> bumped{ -1 };
    std::vector<double> hs( n, 0.0 );
    if ( gamma )
        for ( int a = 0; a < n; a++ )
        {
            hs[a] = kGammaBump * base.eq[a].spot;
            Market u = base, d = base;
            u.eq[a].spot += hs[a];
            d.eq[a].spot -= hs[a];
            mks.push_back( u );
            bumped.push_back( a );
            mks.push_back( d );
            bumped.push_back( a );
        }
    const size_t nb = mks.size(), nd = g.t.size();
    // leverage per (book, equity): its own calibration where the book moved
    // the equity (every equity of the base book), else the base book's; the
    // calibrations are independent and run concurrently
    std::vector<std::vector<LeverageGrid>> levs( nb, std::vector<LeverageGrid>( n ) );
    auto own = [&]( size_t k, int a ) { return mks[k].eq[a].kind == THOTH_VOL_LSV && ( k == 0 || a == bumped[k] ); };
    {
        std::vector<std::function<void()>> tasks;
        for ( size_t k = 0; k < nb; k++ )
            for ( int a = 0; a < n; a++ )
                if ( own( k, a ) )
                    tasks.push_back( [&, k, a] {
                        if ( CalibrateLeverage( mks[k], a, g, levs[k][a] ) != THOTH_OK )
                            throw Invalid( THOTH_ERR_INVALID, "leverage calibration failed" );
                    } );
        RunConcurrently( tasks );
        for ( size_t k = 1; k < nb; k++ )
            for ( int a = 0; a < n; a++ )
                if ( mks[k].eq[a].kind == THOTH_VOL_LSV && !own( k, a ) )
                    levs[k][a] = levs[0][a];
    }
    std::vector<BookArrays> arrays( nb );
    std::vector<thoth_mcl_book> books;
    for ( size_t k = 0; k < nb; k++ )
    {
        Resolve( req, mks[k], g, chol, levs[k], arrays[k] );
        books.push_back( arrays[k].book );
    }
    // leverage sensitivities: to the spot (every equity of the base book, the
    // bumped equity of a gamma book) and to the vol shift (base book), zeros
    // elsewhere (those Greeks are not read from those books)
    std::vector<std::vector<double>> ds( nb ), de( nb );
    std::vector<PathwiseLayers> pwl( nb );
    std::vector<std::function<void()>> tasks;
    for ( size_t k = 0; k < nb; k++ )
    {
        const int P = arrays[k].book.lev_points;
        if ( P <= 0 )
            continue;
        ds[k].assign( (size_t)n * nd * P, 0.0 );
        de[k].assign( (size_t)n * nd * P, 0.0 );
        pwl[k] = { ds[k].data(), de[k].data() };
        for ( int a = 0; a < n; a++ )
        {
            if ( !own( k, a ) )
                continue;
            if ( levs[k][a].points != P )
                throw Invalid( THOTH_ERR_UNSUPPORTED, "pathwise Greeks: LSV equities with one leverage layout" );
            for ( int which = 0; which < ( k == 0 ? 2 : 1 ); which++ )
                tasks.push_back( [&, k, a, which, P] {
                    std::vector<double> dv;
                    LeverageTangent( mks[k], a, g, levs[k][a], which, dv );
                    std::copy( dv.begin(), dv.end(), ( which == 0 ? ds[k] : de[k] ).begin() + (size_t)a * nd * P );
                } );
        }
    }
    RunConcurrently( tasks );
    thoth_mcl_sweep_cfg cfg{};
    cfg.paths = req.mcl.paths;
    cfg.sobol_skip = req.mcl.sobol_skip;
    cfg.precision = req.precision;
    DeviceBatch B;
    int rc = PrepareBatch( (int32_t)nb, books.data(), &cfg, B, nullptr, nullptr, 2, nullptr, 0, nullptr, pwl.data() );
    if ( rc != THOTH_OK )
        return rc;
    const auto t1 = std::chrono::steady_clock::now();
    std::vector<double> part;
    float ms = 0;
    if ( ( rc = RunPathwise( B, 2, part, ms ) ) != THOTH_OK )
        return rc;
    // [Y, sum of deltas, vega, delta_0 .. delta_{n-1}] x [sum, sum2] per book
    const int ne = 3 + n;
    const double N = (double)req.mcl.paths;
    const auto sums = BlockSums( B, part, (int)nb, ne );
    auto mean = [&]( size_t b, int e ) { return sums[b][2 * e] / N; };
    auto trust = [&]( size_t b, int e ) {
        const double s1 = sums[b][2 * e], s2 = sums[b][2 * e + 1];
        return std::sqrt( std::max( 0.0, ( s2 - s1 * s1 / N ) / ( N - 1 ) / N ) );
    };
    res->premium = mean( 0, 0 );
    res->premium_trust = trust( 0, 0 );
    res->delta = mean( 0, 1 );
    res->delta_trust = trust( 0, 1 );
    res->vega = mean( 0, 2 );
    res->vega_trust = trust( 0, 2 );
    res->gamma = 0;
    res->n_assets = n;
    for ( int a = 0; a < n; a++ )
    {
        res->asset_delta[a] = mean( 0, 3 + a );
        res->asset_delta_trust[a] = trust( 0, 3 + a );
        // books 1 + 2a (up) and 2 + 2a (down) of equity a
        res->asset_gamma[a] = gamma ? ( mean( 1 + 2 * a, 3 + a ) - mean( 2 + 2 * a, 3 + a ) ) / ( 2 * hs[a] ) : 0.0;
        res->gamma += res->asset_gamma[a];
    }
    res->n_books = (int32_t)nb;
    res->sweep_ms = ms;
    res->setup_ms = std::chrono::duration<double, std::milli>( t1 - t0 ).count();
    return THOTH_OK;
}

int PricePathwise( const thoth_mcl_request& req, thoth_pathwise_result* res )
{
    if ( req.n_assets < 1 || req.n_assets > THOTH_MCL_MAX_ASSETS || !req.assets || req.n_contracts != 1 ||
         !req.contracts )
        throw Invalid( THOTH_ERR_UNSUPPORTED, "pathwise Greeks: 1..8 equities and one contract" );
    if ( req.contracts[0].american )
        throw Invalid( THOTH_ERR_UNSUPPORTED, "pathwise Greeks: european exercise" );
    if ( req.mcl.paths <= 1 || req.mcl.paths > 200000000 || !req.mcl.use_sobol )
        throw Invalid( THOTH_ERR_INVALID, "paths in (1, 2e8], Sobol mode" );
    *res = thoth_pathwise_result{};
    const thoth_contract& k0 = req.contracts[0];
    const bool single = req.n_assets == 1 && req.assets[0].vol_kind != THOTH_VOL_LSV && k0.underlying == 0;
    if ( k0.kind == THOTH_BARRIER )
    {
        if ( !single )
            throw Invalid( THOTH_ERR_UNSUPPORTED,
                           "pathwise Greeks of barrier contracts: single-equity Black-Scholes or Heston books" );
        if ( k0.barrier_type == THOTH_UP_IN || k0.barrier_type == THOTH_DOWN_IN )
            throw Invalid( THOTH_ERR_UNSUPPORTED, "pathwise Greeks: knock-out barriers (knock-in = vanilla - knock-out)" );
    }
    return single ? PricePathwiseSingle( req, res ) : PricePathwiseBasket( req, res );
}

}  // namespace thoth::host

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_mcl_price_pathwise( const thoth_mcl_request* req,
                                                                                  thoth_pathwise_result* res )
{
    if ( !req || !res )
        return thoth::SetError( THOTH_ERR_INVALID, "null argument" );
    std::lock_guard<std::mutex> lk( thoth::DeviceMutex() );
    try
    {
        int rc = thoth::EnsureDevice();
        if ( rc != THOTH_OK )
            return rc;
        return thoth::host::PricePathwise( *req, res );
    }
    catch ( const thoth::host::Invalid& e )
    {
        return thoth::SetError( e.code, e.what() );
    }
    catch ( const std::exception& e )
    {
        return thoth::SetError( THOTH_ERR_INVALID, e.what() );
    }
}

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_mcl_prepare( const thoth_mcl_request* req,
                                                                           thoth_mcl_plan* plan )
{
    if ( !req || !plan )
        return thoth::SetError( THOTH_ERR_INVALID, "null argument" );
    *plan = nullptr;
    std::lock_guard<std::mutex> lk( thoth::DeviceMutex() );
    try
    {
        int rc = thoth::EnsureDevice();
        if ( rc != THOTH_OK )
            return rc;
        auto* p = new thoth::host::Plan();
        rc = thoth::host::PreparePlan( *req, *p );
        if ( rc != THOTH_OK )
        {
            delete p;
            return rc;
        }
        *plan = reinterpret_cast<thoth_mcl_plan>( p );
        return THOTH_OK;
    }
    catch ( const thoth::host::Invalid& e )
    {
        return thoth::SetError( e.code, e.what() );
    }
    catch ( const std::exception& e )
    {
        return thoth::SetError( THOTH_ERR_INVALID, e.what() );
    }
}

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_mcl_run( thoth_mcl_plan plan, thoth_mcl_result* res )
{
    if ( !plan || !res )
        return thoth::SetError( THOTH_ERR_INVALID, "null argument" );
    std::lock_guard<std::mutex> lk( thoth::DeviceMutex() );
    return thoth::host::RunPlan( *reinterpret_cast<thoth::host::Plan*>( plan ), res );
}

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_mcl_tune( thoth_mcl_plan plan,
                                                                        thoth_plan_shape* shape )
{
    if ( !plan || !shape )
        return thoth::SetError( THOTH_ERR_INVALID, "null argument" );
    std::lock_guard<std::mutex> lk( thoth::DeviceMutex() );
    auto* p = reinterpret_cast<thoth::host::Plan*>( plan );
    thoth::ShapeChoice c;
    int rc = thoth::TuneShape( p->batch, c );
    if ( rc != THOTH_OK )
        return rc;
    p->tuned_ms = c.ms;
    shape->lev_group = c.lev_group;
    shape->workgroups_per_cu = c.wpc;
    shape->device_ms = c.ms;
    shape->candidates = c.candidates;
    shape->default_ms = c.default_ms;
    return THOTH_OK;
}

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_mcl_set_shape( thoth_mcl_plan plan,
                                                                             int32_t lev_group,
                                                                             int32_t workgroups_per_cu )
{
    if ( !plan )
        return thoth::SetError( THOTH_ERR_INVALID, "null argument" );
    std::lock_guard<std::mutex> lk( thoth::DeviceMutex() );
    return thoth::SetShape( reinterpret_cast<thoth::host::Plan*>( plan )->batch, lev_group, workgroups_per_cu );
}

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_mcl_get_shape( thoth_mcl_plan plan,
                                                                             thoth_plan_shape* shape )
{
    if ( !plan || !shape )
        return thoth::SetError( THOTH_ERR_INVALID, "null argument" );
    std::lock_guard<std::mutex> lk( thoth::DeviceMutex() );
    const auto* p = reinterpret_cast<const thoth::host::Plan*>( plan );
    shape->lev_group = p->batch.db.empty() ? 0 : p->batch.db[0].lev_group;
    shape->workgroups_per_cu = p->batch.wave_split ? -1 : p->batch.persist_wpc;
    shape->device_ms = p->tuned_ms;
    shape->candidates = 0;
    shape->default_ms = 0;
    return THOTH_OK;
}

extern "C" __attribute__( ( visibility( "default" ) ) ) void thoth_mcl_release( thoth_mcl_plan plan )
{
    std::lock_guard<std::mutex> lk( thoth::DeviceMutex() );
    delete reinterpret_cast<thoth::host::Plan*>( plan );
}

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_mcl_grid( const thoth_mcl_request* req, int32_t cap,
                                                                        double* days, double* t, double* dt )
{
    if ( !req || cap < 0 )
        return -thoth::SetError( THOTH_ERR_INVALID, "null argument" );
    try
    {
        const thoth::host::Market mk = thoth::host::BuildMarket( *req );
        const thoth::host::Grid g = thoth::host::BuildGrid( *req, mk );
        for ( int i = 0; i < (int)g.t.size() && i < cap; i++ )
        {
            days[i] = g.days[i];
            t[i] = g.t[i];
            dt[i] = g.dt[i];
        }
        return (int)g.t.size();
    }
    catch ( const thoth::host::Invalid& e )
    {
        return -thoth::SetError( e.code, e.what() );
    }
    catch ( const std::exception& e )
    {
        return -thoth::SetError( THOTH_ERR_INVALID, e.what() );
    }
}

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_mcl_leverage_tangent(
    const thoth_mcl_request* req, int32_t asset, int32_t which, int32_t cap_dates, double* dv )
{
    if ( !req || asset < 0 || asset >= req->n_assets || ( which != 0 && which != 1 ) || !dv )
        return -thoth::SetError( THOTH_ERR_INVALID, "asset index, tangent kind or null output" );
    try
    {
        const thoth::host::Market mk = thoth::host::BuildMarket( *req );
        if ( mk.eq[asset].kind != THOTH_VOL_LSV )
            return -thoth::SetError( THOTH_ERR_INVALID, "equity is not LSV" );
        const thoth::host::Grid g = thoth::host::BuildGrid( *req, mk );
        thoth::host::LeverageGrid L;
        int rc = thoth::host::CalibrateLeverage( mk, asset, g, L );
        if ( rc != THOTH_OK )
            return -rc;
        std::vector<double> d;
        thoth::host::LeverageTangent( mk, asset, g, L, which, d );
        const int nd = (int)g.t.size();
        for ( int i = 0; i < nd && i < cap_dates; i++ )
        {
            std::copy( d.begin() + (size_t)i * L.points, d.begin() + (size_t)( i + 1 ) * L.points,
                       dv + (size_t)i * L.points );
        }
        return nd;
    }
    catch ( const thoth::host::Invalid& e )
    {
        return -thoth::SetError( e.code, e.what() );
    }
    catch ( const std::exception& e )
    {
        return -thoth::SetError( THOTH_ERR_INVALID, e.what() );
    }
}

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_mcl_calibrate_leverage(
    const thoth_mcl_request* req, int32_t asset, double spot, int32_t cap_layers, double* t, double* ln_step,
    int64_t* offset, double* levels )
{
    if ( !req || asset < 0 || asset >= req->n_assets )
        return -thoth::SetError( THOTH_ERR_INVALID, "asset index" );
    try
    {
        thoth::host::Market mk = thoth::host::BuildMarket( *req );
        if ( mk.eq[asset].kind != THOTH_VOL_LSV )
            return -thoth::SetError( THOTH_ERR_INVALID, "equity is not LSV" );
        mk.eq[asset].spot = spot;
        const thoth::host::Grid g = thoth::host::BuildGrid( *req, mk );
        thoth::host::LeverageGrid L;
        int rc = thoth::host::CalibrateLeverage( mk, asset, g, L );
        if ( rc != THOTH_OK )
            return -rc;
        const bool surf = !L.surface.t.empty();
        const int n = surf ? (int)L.surface.t.size() : (int)L.t.size();
        const std::vector<double>& v = surf ? L.surface.values : L.values;
        for ( int k = 0; k < n && k < cap_layers; k++ )
        {
            t[k] = surf ? L.surface.t[k] : L.t[k];
            ln_step[k] = surf ? L.surface.ln_step : L.ln_step[k];
            offset[k] = surf ? L.surface.offset : L.offset[k];
            std::copy( v.begin() + (size_t)k * L.points, v.begin() + (size_t)( k + 1 ) * L.points,
                       levels + (size_t)k * L.points );
        }
        return n;
    }
    catch ( const thoth::host::Invalid& e )
    {
        return -thoth::SetError( e.code, e.what() );
    }
    catch ( const std::exception& e )
    {
        return -thoth::SetError( THOTH_ERR_INVALID, e.what() );
    }
}
