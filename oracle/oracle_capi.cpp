// oracle_capi.cpp -- TEST INFRASTRUCTURE ONLY: C-ABI of the CPU oracle for the
// Python test-suite (ctypes).  See mcl_oracle.hpp.
#include <chrono>
#include <cstring>
#include <stdexcept>
#include <string>

#include "mcl_oracle.hpp"

using namespace oracle;

namespace
{
thread_local std::string g_err;
int Fail( const std::exception& e )
{
    g_err = e.what();
    return THOTH_ERR_INVALID;
}
}  // namespace

extern "C" {

const char* oracle_last_error( void ) { return g_err.c_str(); }

int oracle_load_joe_kuo( const char* path ) { return LoadJoeKuo( path ) ? 0 : -1; }

int oracle_sobol_max_dimension( void ) { return (int)SobolMaxDimension(); }

// direction numbers V[d][32] of the first dim dimensions
int oracle_sobol_directions( int dim, uint32_t* out )
{
    Sobol s( (unsigned)dim );
    for ( unsigned d = 0; d < s.Dimension(); d++ )
        for ( int k = 0; k < 32; k++ )
            out[d * 32 + k] = s.V()[d][k];
    return (int)s.Dimension();
}

// integer coordinates of points skip+1 .. skip+n (sequential Gray-code recurrence)
int oracle_sobol_int( int dim, int64_t skip, int64_t n, uint32_t* out )
{
    Sobol s( (unsigned)dim );
    s.Skip( (uint64_t)skip );
    std::vector<uint32_t> x;
    for ( int64_t i = 0; i < n; i++ )
    {
        s.NextInt( x );
        std::memcpy( out + i * s.Dimension(), x.data(), sizeof( uint32_t ) * s.Dimension() );
    }
    return (int)s.Dimension();
}

void oracle_normal_cdf_inv( int64_t n, const double* p, double* z )
{
    for ( int64_t i = 0; i < n; i++ )
        z[i] = NormalCdfInv( p[i] );
}

void oracle_xoshiro( uint64_t seed, int64_t n, uint64_t* out )
{
    Rng r( seed );
    for ( int64_t i = 0; i < n; i++ )
        out[i] = r();
}

// raw-state xoshiro256++ (published test vectors start from a chosen state)
void oracle_xoshiro_raw( const uint64_t* state, int64_t n, uint64_t* out )
{
    uint64_t s[4] = { state[0], state[1], state[2], state[3] };
    auto rotl = []( uint64_t x, int k ) { return ( x << k ) | ( x >> ( 64 - k ) ); };
    for ( int64_t i = 0; i < n; i++ )
    {
        // identical update to Rng::operator() (rng.hpp:54-65)
        const uint64_t result = rotl( s[0] + s[3], 23 ) + s[0];
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl( s[3], 45 );
        out[i] = result;
    }
}

// the 256-bit state splitmix64 expands a seed into (rng.hpp:40-51)
void oracle_xoshiro_seed_state( uint64_t seed, uint64_t* state )
{
    uint64_t z = seed;
    for ( int i = 0; i < 4; i++ )
    {
        z += 0x9e3779b97f4a7c15ULL;
        uint64_t x = z;
        x = ( x ^ ( x >> 30 ) ) * 0xbf58476d1ce4e5b9ULL;
        x = ( x ^ ( x >> 27 ) ) * 0x94d049bb133111ebULL;
        state[i] = x ^ ( x >> 31 );
    }
}

void oracle_gaussians( uint64_t seed, int64_t n, double* out )
{
    Rng r( seed );
    for ( int64_t i = 0; i < n; i++ )
        out[i] = r.Gaussian();
}

int oracle_bridge_schedule( int n_times, const double* t, int* bidx, int* lidx, int* ridx, double* lw,
                            double* rw, double* sd )
{
    std::vector<double> tt( t, t + n_times );
    BridgeSchedule s = BuildBridgeSchedule( tt );
    for ( int i = 0; i + 1 < n_times; i++ )
    {
        bidx[i] = s.bridge_index[i];
        lidx[i] = s.left_index[i];
        ridx[i] = s.right_index[i];
        lw[i] = s.left_weight[i];
        rw[i] = s.right_weight[i];
        sd[i] = s.std_dev[i];
    }
    return n_times - 1;
}

// normalized increments of `paths` paths: out[p][f][i], i = 0..n_times-1 (i=0 unused),
// the pseudo-random draws (dimensions past the Sobol table, or every draw when
// use_sobol is 0) from the stream seeded by `seed`
int oracle_path_noise_seeded( int n_times, const double* t, int factors, int64_t skip, int64_t paths, uint64_t seed,
                              int use_sobol, double* out )
{
    std::vector<double> tt( t, t + n_times );
    Rng rng( seed );
    PathGenerator pg( tt, factors, &rng, (uint64_t)skip, use_sobol != 0 );
    for ( int64_t p = 0; p < paths; p++ )
    {
        pg.NextPath();
        for ( int f = 0; f < factors; f++ )
            std::memcpy( out + ( p * factors + f ) * n_times, pg.Buffer( f ).data(), sizeof( double ) * n_times );
    }
    return 0;
}

int oracle_path_noise( int n_times, const double* t, int factors, int64_t skip, int64_t paths, double* out )
{
    return oracle_path_noise_seeded( n_times, t, factors, skip, paths, 0, 1, out );
}

// LeastSquares (maths.cpp:419-470): X [n][p] row-major; returns 1 and beta, 0 if singular
int oracle_least_squares( const double* X, int64_t n, int p, const double* y, double* beta )
{
    std::vector<double> x( X, X + n * p ), yy( y, y + n );
    const std::vector<double> b = LeastSquares( x, (size_t)n, (size_t)p, yy );
    if ( b.empty() )
        return 0;
    std::copy( b.begin(), b.end(), beta );
    return 1;
}

// Longstaff-Schwartz on given paths [N][M]: the fitted policy (b[3 t + k], fit[t])
// and the value of applying it to the same paths, with its trust
double oracle_lsm( const double* paths, int64_t N, int M, const double* tau, const double* rates, double strike,
                   int is_call, double* b, int* fit, double* trust )
{
    std::vector<double> P( paths, paths + N * M ), t( tau, tau + M ), r( rates, rates + M );
    const AmericanPolicy pol = FitAmericanPolicy( strike, is_call != 0, P, (size_t)N, (size_t)M, t, r );
    for ( int k = 0; k < M; k++ )
    {
        b[3 * k] = pol.b0.empty() ? 0 : pol.b0[k];
        b[3 * k + 1] = pol.b1.empty() ? 0 : pol.b1[k];
        b[3 * k + 2] = pol.b2.empty() ? 0 : pol.b2[k];
        fit[k] = pol.has_fit.empty() ? 0 : pol.has_fit[k];
    }
    return ApplyAmericanPolicy( strike, is_call != 0, P, (size_t)N, (size_t)M, t, r, pol, *trust );
}

static Sabr SabrOf( const thoth_sabr_surface* s )
{
    Sabr sb;
    for ( int i = 0; i < s->n; i++ )
    {
        sb.maturities.push_back( s->maturities[i] );
        sb.alpha.push_back( s->alpha[i] );
        sb.beta.push_back( s->beta[i] );
        sb.rho.push_back( s->rho[i] );
        sb.nu.push_back( s->nu[i] );
    }
    return sb;
}

double oracle_sabr_vol( const thoth_sabr_surface* s, double K, double F, double day )
{
    return SabrOf( s ).ImplicitVol( K, F, day );
}

// the Hagan expansion alone (no wings) at a maturity T in years
double oracle_sabr_hagan( const thoth_sabr_surface* s, double K, double F, double T )
{
    return SabrOf( s ).HaganVol( K, F, T );
}

double oracle_local_vol( const thoth_mcl_request* req, int asset, double K, double day )
{
    Market mk = BuildMarket( *req );
    return mk.LocalVol( asset, K, day );
}

int oracle_cholesky( int n, double* a ) { std::vector<double> v( a, a + n * n ); bool ok = CholeskyDecomposeLower( v, n ); std::memcpy( a, v.data(), sizeof( double ) * n * n ); return ok ? 0 : -1; }

// diffusion grid of a request: returns the number of dates (<= cap)
int oracle_grid( const thoth_mcl_request* req, int cap, double* days, double* t, double* dt )
{
    try
    {
        Market mk = BuildMarket( *req );
        Grid g = BuildGrid( *req, mk );
        int n = (int)g.t.size();
        for ( int i = 0; i < n && i < cap; i++ )
        {
            days[i] = g.days[i];
            t[i] = g.t[i];
            dt[i] = g.dt[i];
        }
        return n;
    }
    catch ( const std::exception& e )
    {
        return -Fail( e );
    }
}

// the spot/spot correlation of the request averaged over [a, b] years, n x n
int oracle_avg_correlation( const thoth_mcl_request* req, double a, double b, double* out )
{
    try
    {
        Market mk = BuildMarket( *req );
        const int n = req->n_assets;
        for ( int i = 0; i < n; i++ )
            for ( int j = 0; j < n; j++ )
                out[i * n + j] = i == j ? 1.0 : mk.AvgCorrelation( i, j, a, b );
        return 0;
    }
    catch ( const std::exception& e )
    {
        return -Fail( e );
    }
}

// per-date lower Cholesky factors of a term-structured correlation [dates][n][n];
// returns the date count (0 for a constant matrix)
int oracle_chol_schedule( const thoth_mcl_request* req, int cap, double* out )
{
    try
    {
        Market mk = BuildMarket( *req );
        Grid g = BuildGrid( *req, mk );
        ResolvedBook b = ResolveBook( *req, mk, g );
        const int n = req->n_assets;
        const int nd = (int)b.chol_dates.size();
        for ( int d = 0; d < nd && d < cap; d++ )
            std::memcpy( out + (size_t)d * n * n, b.chol_dates[d].data(), sizeof( double ) * n * n );
        return nd;
    }
    catch ( const std::exception& e )
    {
        return -Fail( e );
    }
}

// leverage layers of an LSV asset on the request's grid (base market + optional spot / vol shift)
int oracle_calibrate( const thoth_mcl_request* req, int asset, double spot, double vol_shift, double rate_shift, int cap,
                      double* t, double* ln_step, int64_t* offset, double* levels )
{
    try
    {
        Market mk = BuildMarket( *req );
        mk.eq[asset].spot = spot;
        for ( Equity& e : mk.eq )
            e.vol_shift = vol_shift;
        mk.rate.shift = rate_shift;
        Grid g = BuildGrid( *req, mk );
        LeverageLayers L = CalibrateLeverage( mk, asset, g.days );
        int n = (int)L.t.size();
        for ( int k = 0; k < n && k < cap; k++ )
        {
            t[k] = L.t[k];
            ln_step[k] = L.ln_step[k];
            offset[k] = L.offset[k];
            std::memcpy( levels + (size_t)k * L.levels[k].size(), L.levels[k].data(), sizeof( double ) * L.levels[k].size() );
        }
        return n;
    }
    catch ( const std::exception& e )
    {
        return -Fail( e );
    }
}

// base-book sums (book + contracts) on the request's own market
int oracle_sweep( const thoth_mcl_request* req, double* book_sum, double* book_sum2, double* sum, double* sum2,
                  int64_t* count )
{
    try
    {
        Market mk = BuildMarket( *req );
        Grid g = BuildGrid( *req, mk );
        BookSums s = SweepBook( ResolveBook( *req, mk, g ), req->mcl.paths, req->mcl.sobol_skip, req->mcl.seed,
                                req->mcl.use_sobol != 0 );
        *book_sum = s.book_sum;
        *book_sum2 = s.book_sum2;
        for ( size_t c = 0; c < s.sum.size(); c++ )
        {
            sum[c] = s.sum[c];
            sum2[c] = s.sum2[c];
        }
        *count = s.count;
        return 0;
    }
    catch ( const std::exception& e )
    {
        return Fail( e );
    }
}

// seconds of the path sweep alone (grid, calibration and book resolution excluded)
double oracle_sweep_seconds( const thoth_mcl_request* req, double* book_sum )
{
    try
    {
        Market mk = BuildMarket( *req );
        Grid g = BuildGrid( *req, mk );
        ResolvedBook rb = ResolveBook( *req, mk, g );
        const auto t0 = std::chrono::steady_clock::now();
        BookSums s = SweepBook( rb, req->mcl.paths, req->mcl.sobol_skip, req->mcl.seed,
                                req->mcl.use_sobol != 0 );
        const double sec = std::chrono::duration<double>( std::chrono::steady_clock::now() - t0 ).count();
        if ( book_sum )
            *book_sum = s.book_sum;
        return sec;
    }
    catch ( const std::exception& e )
    {
        Fail( e );
        return -1.0;
    }
}

int oracle_price( const thoth_mcl_request* req, thoth_mcl_result* res )
{
    try
    {
        return Price( *req, res );
    }
    catch ( const std::exception& e )
    {
        return Fail( e );
    }
}
}
