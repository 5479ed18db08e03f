// mcl_market.cpp -- FP64 market model of the host engine (see mcl_hostlib.hpp):
// curves, SABR + Dupire, the LSV leverage calibration by the binned particle
// method, and the Cholesky factor.  Operation order follows the reference
// sources cited per function so that the calibrated layers reproduce the
// reference's bit for bit on the same platform libm.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "mcl_hostlib.hpp"
#include "thoth_mcl_gpu.h"

namespace thoth::host
{

double Curve::Value( double day ) const
{
    const size_t n = days.size();
    double r;
    if ( n <= 1 || day <= days.front() )
        r = values[0];
    else if ( day >= days.back() )
        r = values[n - 1];
    else
    {
        size_t i = 0;
        while ( i + 1 < n && days[i + 1] <= day )
            i++;
        const double w = YearFraction( days[i], day ) / YearFraction( days[i], days[i + 1] );
        r = values[i] + w * ( values[i + 1] - values[i] );
    }
    return r + shift;
}

// ---- SABR (sabr_volatility.cpp:40-239) ------------------------------------
namespace
{
inline double Ncdf( double x ) { return 0.5 * std::erfc( -x * M_SQRT1_2 ); }  // distributions.cpp:21-26
double BlackCall( double F, double K, double V, double T )
{
    if ( V <= 0 || T <= 0 )
        return std::max( F - K, 0.0 );
    const double s = V * std::sqrt( T );
    const double d1 = std::log( F / K ) / s + 0.5 * s;
    return F * Ncdf( d1 ) - K * Ncdf( d1 - s );
}
double BlackPut( double F, double K, double V, double T )
{
    if ( V <= 0 || T <= 0 )
        return std::max( K - F, 0.0 );
    const double s = V * std::sqrt( T );
    const double d1 = std::log( F / K ) / s + 0.5 * s;
    return K * Ncdf( s - d1 ) - F * Ncdf( -d1 );
}
double TailImplied( double F, double K, double T, double price, bool call )
{
    double lo = 0.0, hi = 20.0;  // IMPLICIT_VOL_MAX, finance.hpp:206
    for ( int i = 0; i < 100 && hi - lo > 1e-13; i++ )
    {
        const double mid = 0.5 * ( lo + hi );
        const double p = call ? BlackCall( F, K, mid, T ) : BlackPut( F, K, mid, T );
        if ( p < price )
            lo = mid;
        else
            hi = mid;
    }
    return 0.5 * ( lo + hi );
}
}  // namespace

double Sabr::Interp( const std::vector<double>& v, double T ) const
{
    if ( T <= mats.front() )
        return v.front();
    if ( T >= mats.back() )
        return v.back();
    for ( size_t i = 1; i < mats.size(); i++ )
        if ( T <= mats[i] )
        {
            const double w = ( T - mats[i - 1] ) / ( mats[i] - mats[i - 1] );
            return v[i - 1] + w * ( v[i] - v[i - 1] );
        }
    return v.back();
}

double Sabr::Hagan( double K, double F, double T ) const
{
    const double al = Interp( alpha, T ), be = Interp( beta, T ), rh = Interp( rho, T ), n = Interp( nu, T );
    const double omb = 1.0 - be;
    const double lfk = std::log( F / K );
    const double fkp = std::pow( F * K, omb / 2.0 );
    const double den = fkp * ( 1.0 + omb * omb / 24.0 * lfk * lfk + omb * omb * omb * omb / 1920.0 * lfk * lfk * lfk * lfk );
    const double z = ( n / al ) * fkp * lfk;
    double zxz = 1.0;
    if ( std::fabs( z ) > 1e-12 )
        zxz = z / std::log( ( std::sqrt( 1.0 - 2.0 * rh * z + z * z ) + z - rh ) / ( 1.0 - rh ) );
    const double corr = 1.0 + ( omb * omb / 24.0 * al * al / std::pow( F * K, omb ) + 0.25 * rh * be * n * al / fkp +
                                ( 2.0 - 3.0 * rh * rh ) / 24.0 * n * n ) *
                                  T;
    return ( al / den ) * zxz * corr + 0.0;
}

double Sabr::Wing( double K, double F, double T, double cut, bool right ) const
{
    const double h = 1e-4 * cut;
    auto price = [&]( double k )
    {
        const double v = Hagan( k, F, T );
        return right ? BlackCall( F, k, v, T ) : BlackPut( F, k, v, T );
    };
    const double p0 = price( cut );
    const double dp = ( price( cut + h ) - price( cut - h ) ) / ( 2 * h );
    if ( p0 <= 0 )
        return Hagan( K, F, T );
    if ( right )
    {
        const double a = -cut * dp / p0;
        if ( a <= 0 )
            return Hagan( K, F, T );
        return TailImplied( F, K, T, p0 * std::pow( cut / K, a ), true );
    }
    const double b = cut * dp / p0;
    if ( b <= 1 )
        return Hagan( K, F, T );
    return TailImplied( F, K, T, p0 * std::pow( K / cut, b ), false );
}

double Sabr::Implied( double strike, double F, double day ) const
{
    double T = YearFraction( 0.0, day );
    if ( T <= 0 )
        T = 1.0 / kDaysPerYear;
    const double K = strike > 0 ? strike : F;
    const double atm = Hagan( F, F, T );
    const double band = 2.5 * atm * std::sqrt( T );
    const double k_lo = F * std::exp( -band );
    const double k_hi = F * std::exp( band );
    if ( K > k_hi )
        return Wing( K, F, T, k_hi, true );
    if ( K < k_lo )
        return Wing( K, F, T, k_lo, false );
    return Hagan( K, F, T );
}

// ---- market ----------------------------------------------------------------
StochParams Market::Stoch( int a ) const
{
    const Equity& e = eq[a];
    if ( e.kind == THOTH_VOL_LSV )  // lsv_volatility.hpp:50-55: no parallel shift on the Heston part
        return { e.v0 + 0.0, e.kappa + 0.0, e.theta + 0.0, e.xi + 0.0, e.rho + 0.0 };
    auto shifted = [&]( double v )  // heston_volatility.hpp:92-96
    {
        const double s = std::sqrt( v ) + e.vol_shift;
        return s * s;
    };
    return { shifted( e.v0 ) + 0.0, e.kappa + 0.0, shifted( e.theta ) + 0.0, e.xi + 0.0, e.rho + 0.0 };
}

double Market::DivRepo( int a, double ) const
{
    double y = eq[a].has_div ? eq[a].div : 0;
    y += 0;
    return y;
}

double Market::Drift( int a, double day ) const
{
    const double r = rate.Value( day );
    const double q = eq[a].has_div ? eq[a].div : 0.0;
    return ( ( r - 0 ) - 0 ) - q;
}

double Market::Forward( int a, double day ) const
{
    const double dt = YearFraction( 0.0, day );
    const double df = std::exp( -rate.Value( day ) * dt );  // yield_curve.cpp:16-21
    return ( eq[a].spot - 0.0 ) * std::exp( -dt * ( DivRepo( a, day ) + 0 ) ) / df;
}

double Market::ImplicitVol( int a, double strike, double day ) const
{
    const Equity& e = eq[a];
    if ( e.kind == THOTH_VOL_BS )
        return ( e.bs_vol + e.vol_shift ) * 1.0;  // bs_volatility.cpp:30-35
    if ( e.kind == THOTH_VOL_HESTON )
        return ( std::sqrt( e.v0 ) + e.vol_shift ) * 1.0;  // heston_volatility.cpp:51-56
    return e.surface.Implied( strike, Forward( a, day ), day ) + e.vol_shift;  // lsv_volatility.cpp:56-61
}

double Market::LocalVol( int a, double K, double day ) const
{
    const Equity& e = eq[a];
    const double r = rate.Value( day );
    double q = 0;
    if ( e.has_div )
        q += e.div;
    const double S = e.spot;
    auto IV = [&]( double k, double f, double d )
    {
        if ( e.kind == THOTH_VOL_BS )
            return ( e.bs_vol + e.vol_shift ) * 1.0;
        if ( e.kind == THOTH_VOL_HESTON )
            return ( std::sqrt( e.v0 ) + e.vol_shift ) * 1.0;
        return e.surface.Implied( k, f, d ) + e.vol_shift;
    };
    const double T = YearFraction( 0.0, day );
    const double dK = .01, dT = 1.0 / kDaysPerYear;
    const double F = S * std::exp( ( r - q ) * T );
    const double F_Td = S * std::exp( ( r - q ) * YearFraction( 0.0, day - 1 ) );
    const double F_Tu = S * std::exp( ( r - q ) * YearFraction( 0.0, day + 1 ) );
    const double V = IV( K, F, day ), V_Kd = IV( K - dK, F, day ), V_Ku = IV( K + dK, F, day );
    const double V_Td = IV( K, F_Td, day - 1 ), V_Tu = IV( K, F_Tu, day + 1 );
    const double sT = std::sqrt( T ), V2 = V * V;
    const double d1 = ( std::log( S / K ) + ( r - q + .5 * V2 ) * T ) / V / sT;
    const double dVdK = ( V_Ku - V_Kd ) / ( 2 * dK );
    const double dVdK2 = ( V_Ku + V_Kd - 2 * V ) / ( dK * dK );
    const double dVdT = ( V_Tu - V_Td ) / ( 2 * dT );
    const double N = 2 * V * T * ( dVdT + ( r - q ) * K * dVdK ) + V2;
    const double D1 = 1 + K * d1 * dVdK * sT;
    const double D = D1 * D1 + K * K * T * V * ( dVdK2 - d1 * dVdK * dVdK * sT );
    const double lv = N / D;
    const double cap = 5.0 * V * ( 5.0 * V );
    if ( lv > cap )
        return 5.0 * V;
    return std::sqrt( lv > 1e-8 ? lv : 1e-8 );
}

// Counts are checked against the fixed arrays of thoth_mcl.h before anything is
// read, and the pillars against what the interpolations divide by.
static void CheckRequestShape( const thoth_mcl_request& req )
{
    if ( !req.assets || req.n_assets < 1 || req.n_assets > THOTH_MCL_MAX_ASSETS )
        throw Invalid( THOTH_ERR_INVALID, "1..8 equities" );
    if ( req.rate.n < 0 || req.rate.n > THOTH_MAX_PILLARS )
        throw Invalid( THOTH_ERR_INVALID, "rate curve: 0..16 pillars" );
    for ( int i = 1; i < req.rate.n; i++ )
        if ( req.rate.days[i] <= req.rate.days[i - 1] )
            throw Invalid( THOTH_ERR_INVALID, "rate curve days must increase" );
    if ( req.n_rainbows < 0 || ( req.n_rainbows > 0 && !req.rainbows ) )
        throw Invalid( THOTH_ERR_INVALID, "rainbow count" );
    for ( int a = 0; a < req.n_assets; a++ )
    {
        const thoth_asset& in = req.assets[a];
        if ( !( in.spot > 0 ) || !std::isfinite( in.spot ) )
            throw Invalid( THOTH_ERR_INVALID, "spot must be positive" );
        if ( in.vol_kind != THOTH_VOL_BS && in.vol_kind != THOTH_VOL_HESTON && in.vol_kind != THOTH_VOL_LSV )
            throw Invalid( THOTH_ERR_INVALID, "unknown volatility model" );
        if ( in.surface.n < 0 || in.surface.n > THOTH_MAX_PILLARS )
            throw Invalid( THOTH_ERR_INVALID, "implied-vol surface: 0..16 pillars" );
        if ( in.vol_kind == THOTH_VOL_LSV && in.surface.n < 1 )
            throw Invalid( THOTH_ERR_INVALID, "an LSV equity needs its implied-vol surface" );
        for ( int i = 1; i < in.surface.n; i++ )
            if ( !( in.surface.maturities[i] > in.surface.maturities[i - 1] ) )
                throw Invalid( THOTH_ERR_INVALID, "surface maturities must increase" );
    }
}

Market BuildMarket( const thoth_mcl_request& req )
{
    CheckRequestShape( req );
    Market mk;
    for ( int a = 0; a < req.n_assets; a++ )
    {
        const thoth_asset& in = req.assets[a];
        Equity e;
        e.spot = in.spot;
        e.kind = in.vol_kind;
        e.bs_vol = in.volatility / 100;                 // bs_volatility.cpp:24
        e.v0 = std::pow( in.init_vol / 100.0, 2 );      // heston_volatility.cpp:38
        e.theta = std::pow( in.long_vol / 100.0, 2 );   // heston_volatility.cpp:39
        e.kappa = in.kappa;
        e.xi = in.vol_of_vol;
        e.rho = in.rho_spot_var;
        e.has_div = in.dividend_yield != 0;
        e.div = in.dividend_yield * 0.01;               // curve.cpp:35-39
        e.vol_shift = req.vol_shift;                    // scenario shift (Volatility::_vol_shift)
        for ( int i = 0; i < in.surface.n; i++ )
        {
            e.surface.mats.push_back( in.surface.maturities[i] );
            e.surface.alpha.push_back( in.surface.alpha[i] );
            e.surface.beta.push_back( in.surface.beta[i] );
            e.surface.rho.push_back( in.surface.rho[i] );
            e.surface.nu.push_back( in.surface.nu[i] );
        }
        mk.eq.push_back( e );
    }
    for ( int i = 0; i < req.rate.n; i++ )
    {
        mk.rate.days.push_back( req.rate.days[i] );
        mk.rate.values.push_back( req.rate.values[i] * 0.01 );
    }
    if ( mk.rate.days.empty() )
    {
        mk.rate.days.push_back( 0 );
        mk.rate.values.push_back( 0 );
    }
    // multi-curve: the flows are discounted on the OIS curve (currency.cpp:129-142)
    if ( req.discount.n < 0 || req.discount.n > THOTH_MAX_PILLARS )
        throw Invalid( THOTH_ERR_INVALID, "discount curve: 0..16 pillars" );
    for ( int i = 0; i < req.discount.n; i++ )
    {
        mk.discount.days.push_back( req.discount.days[i] );
        mk.discount.values.push_back( req.discount.values[i] * 0.01 );
    }
    mk.multi_curve = req.discount.n > 0;
    // spot/spot correlation, constant or term-structured (correlation.cpp:35-130)
    const int n = req.n_assets;
    if ( req.n_corr_pillars < 0 || req.n_corr_pillars > THOTH_MAX_PILLARS )
        throw Invalid( THOTH_ERR_INVALID, "correlation: 0..16 pillars" );
    if ( req.n_corr_pillars > 0 )
    {
        if ( !req.corr_maturities || !req.corr_matrices )
            throw Invalid( THOTH_ERR_INVALID, "correlation pillars need their maturities and matrices" );
        for ( int p = 0; p < req.n_corr_pillars; p++ )
        {
            if ( !( req.corr_maturities[p] > 0 ) || ( p > 0 && !( req.corr_maturities[p] > req.corr_maturities[p - 1] ) ) )
                throw Invalid( THOTH_ERR_INVALID, "correlation maturities must be positive and strictly increasing" );
            const double* m = req.corr_matrices + (size_t)p * n * n;
            for ( int i = 0; i < n; i++ )
                for ( int j = 0; j < n; j++ )
                    if ( !( std::fabs( m[i * n + j] - m[j * n + i] ) <= 1e-12 ) || !( std::fabs( m[i * n + j] ) <= 1.0 ) ||
                         ( i == j && m[i * n + j] != 1.0 ) )
                        throw Invalid( THOTH_ERR_INVALID, "correlation pillar " + std::to_string( p ) +
                                                              " is not a correlation matrix" );
            mk.corr_mats.emplace_back( m, m + n * n );
            mk.corr_pillars.push_back( req.corr_maturities[p] );
        }
        if ( req.n_corr_pillars == 1 )
        {
            mk.correlation = mk.corr_mats[0];
            mk.corr_mats.clear();
            mk.corr_pillars.clear();
        }
        else
        {
            // today's matrix: the instantaneous one at t = 0 (the (0, 0) window)
            mk.correlation.assign( (size_t)n * n, 0.0 );
            for ( int i = 0; i < n; i++ )
                for ( int j = 0; j < n; j++ )
                    mk.correlation[i * n + j] = i == j ? 1.0 : mk.AvgCorrelation( i, j, 0.0, 0.0 );
        }
    }
    else if ( req.correlation )
        mk.correlation.assign( req.correlation, req.correlation + n * n );
    const int sp = req.leverage_spot_points, tl = req.leverage_time_layers;
    if ( sp != 0 || tl != 0 )
    {
        if ( sp < 2 || sp > kMaxLevPoints || tl < 1 )
            throw Invalid( THOTH_ERR_INVALID, "leverage surface needs 2..512 spot points and at least one time layer" );
        mk.lev_spot_points = sp;
        mk.lev_time_layers = tl;
    }
    return mk;
}

// Entry (i, j) of the term structure is e(t): flat at the first pillar's value
// before it, linear between pillars, flat after the last.  Its mean over
// [a, b] is the exact integral of that piecewise-linear function (trapezes on
// the pieces [a, b] cuts) over b - a; a window shorter than 1e-12 years gives
// e(a) (correlation.cpp:528-620).
double Market::AvgCorrelation( int i, int j, double a, double b ) const
{
    const size_t n = eq.size();
    if ( !TermStructured() )
        return correlation[i * n + j];
    const std::vector<double>& m = corr_pillars;
    const size_t np = m.size();
    auto at = [&]( double t )
    {
        if ( t <= m[0] )
            return corr_mats[0][i * n + j];
        if ( t >= m[np - 1] )
            return corr_mats[np - 1][i * n + j];
        size_t p = 1;
        while ( m[p] < t )
            p++;
        const double w = ( t - m[p - 1] ) / ( m[p] - m[p - 1] );
        return corr_mats[p - 1][i * n + j] + w * ( corr_mats[p][i * n + j] - corr_mats[p - 1][i * n + j] );
    };
    a = std::max( a, 0.0 );
    b = std::max( b, a );
    if ( b - a < 1e-12 )
        return at( a );
    // breakpoints inside (a, b), then trapezes between consecutive ones
    double area = 0.0, x0 = a, e0 = at( a );
    for ( size_t p = 0; p <= np; p++ )
    {
        const double x1 = p < np ? std::min( std::max( m[p], a ), b ) : b;
        if ( x1 <= x0 )
            continue;
        const double e1 = at( x1 );
        area += 0.5 * ( e0 + e1 ) * ( x1 - x0 );
        x0 = x1;
        e0 = e1;
    }
    return area / ( b - a );
}

std::vector<double> SpotCholesky( const thoth_mcl_request& req, const Market& mk )
{
    // the spot/spot sub-matrix in name order (pricer_mcl.cpp:462-491, correlation.cpp:794-859)
    const int n = req.n_assets;
    std::vector<double> chol;
    if ( n < 2 )
        return chol;
    if ( mk.correlation.size() != (size_t)n * n )
        throw Invalid( THOTH_ERR_INVALID, "book pricing requires a correlation matrix" );
    chol = mk.correlation;
    if ( !CholeskyLower( chol, n ) )
        throw Invalid( THOTH_ERR_INVALID, "correlation is not SDP" );
    for ( int i = 0; i < n; i++ )
        for ( int j = i + 1; j < n; j++ )
            chol[i * n + j] = 0;
    return chol;
}

bool CholeskyLower( std::vector<double>& A, int n )
{
    for ( int i = 0; i < n; i++ )
        for ( int j = 0; j <= i; j++ )
        {
            double sum = A[i * n + j];
            for ( int k = 0; k < j; k++ )
                sum -= A[i * n + k] * A[j * n + k];
            if ( i == j )
            {
                if ( sum <= 0.0 )
                    return false;
                A[i * n + j] = std::sqrt( sum );
            }
            else
                A[i * n + j] = sum / A[j * n + j];
        }
    return true;
}

// ---- LSV leverage calibration (single.cpp:231-473) -------------------------
namespace
{
constexpr int kGridPoints = 201;          // single.cpp:179
constexpr double kGridSigmas = 8.0;       // single.cpp:177
constexpr double kSurfaceSigmas = 4.0;    // half-width of a surface's shared log-spot axis
constexpr int kParticles = 16384;         // single.cpp:233
constexpr int kBins = 30;
constexpr int kMinBinCount = 20;
constexpr double kMinCondVar = 1e-6;
constexpr double kLevMin = 0.05, kLevMax = 10.0;
constexpr uint64_t kCalibSeed = 20260708;  // single.cpp:246

// xoshiro256++ seeded by splitmix64 (rng.hpp:36-72) and Acklam (distributions.cpp:37-81)
struct Xoshiro
{
    uint64_t s[4];
    explicit Xoshiro( uint64_t seed )
    {
        uint64_t z = seed;
        for ( int i = 0; i < 4; i++ )
        {
            z += 0x9e3779b97f4a7c15ULL;
            uint64_t x = z;
            x = ( x ^ ( x >> 30 ) ) * 0xbf58476d1ce4e5b9ULL;
            x = ( x ^ ( x >> 27 ) ) * 0x94d049bb133111ebULL;
            s[i] = x ^ ( x >> 31 );
        }
    }
    static uint64_t rotl( uint64_t x, int k ) { return ( x << k ) | ( x >> ( 64 - k ) ); }
    uint64_t next()
    {
        const uint64_t r = rotl( s[0] + s[3], 23 ) + s[0];
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl( s[3], 45 );
        return r;
    }
};

double Acklam( double P )
{
    static const double a[] = { -3.969683028665376e+01, 2.209460984245205e+02, -2.759285104469687e+02,
                                1.383577518672690e+02, -3.066479806614716e+01, 2.506628277459239e+00 };
    static const double b[] = { -5.447609879822406e+01, 1.615858368580409e+02, -1.556989798598866e+02,
                                6.680131188771972e+01, -1.328068155288572e+01 };
    static const double c[] = { -7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00,
                                -2.549732539343734e+00, 4.374664141464968e+00, 2.938163982698783e+00 };
    static const double d[] = { 7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00,
                                3.754408661907416e+00 };
    if ( P <= 0.0 )
        return -HUGE_VAL;
    if ( P >= 1.0 )
        return HUGE_VAL;
    if ( P < 0.02425 )
    {
        const double q = std::sqrt( -2.0 * std::log( P ) );
        return ( ( ( ( ( c[0] * q + c[1] ) * q + c[2] ) * q + c[3] ) * q + c[4] ) * q + c[5] ) /
               ( ( ( ( d[0] * q + d[1] ) * q + d[2] ) * q + d[3] ) * q + 1.0 );
    }
    if ( P <= 1.0 - 0.02425 )
    {
        const double q = P - 0.5, r = q * q;
        return ( ( ( ( ( a[0] * r + a[1] ) * r + a[2] ) * r + a[3] ) * r + a[4] ) * r + a[5] ) * q /
               ( ( ( ( ( b[0] * r + b[1] ) * r + b[2] ) * r + b[3] ) * r + b[4] ) * r + 1.0 );
    }
    const double q = std::sqrt( -2.0 * std::log( 1.0 - P ) );
    return -( ( ( ( ( c[0] * q + c[1] ) * q + c[2] ) * q + c[3] ) * q + c[4] ) * q + c[5] ) /
           ( ( ( ( d[0] * q + d[1] ) * q + d[2] ) * q + d[3] ) * q + 1.0 );
}

// The calibration draws the same pseudo-random stream for every equity and
// scenario (a fresh Rng(LSV_CALIB_SEED) per calibration, single.cpp:415): it
// is generated once and cached.  Entry 2*(P*(k-1)+p) is the variance normal of
// particle p into date k, the next entry its spot normal (single.cpp:465-466).
// Calibrations may run concurrently: a longer table is built beside the
// current one, and every table built stays allocated, so a reference handed
// out earlier stays valid.
const std::vector<double>& CalibNormals( size_t steps )
{
    static std::vector<std::unique_ptr<std::vector<double>>> tables;
    static std::mutex mu;
    std::lock_guard<std::mutex> lk( mu );
    const size_t need = steps * 2 * kParticles;
    if ( tables.empty() || tables.back()->size() < need )
    {
        auto z = std::make_unique<std::vector<double>>( need );
        Xoshiro r( kCalibSeed );
        for ( size_t i = 0; i < need; i++ )
        {
            const double u = ( r.next() >> 11 ) * ( 1.0 / 9007199254740992.0 );
            ( *z )[i] = Acklam( u > 0.0 ? u : ( 1.0 / 9007199254740992.0 ) );
        }
        tables.push_back( std::move( z ) );
    }
    return *tables.back();
}

double QeStep( double V, double Dt, double Z, const StochParams& H )  // single.cpp:346-365
{
    const double e = std::exp( -H.kappa * Dt );
    const double m = H.theta + ( V - H.theta ) * e;
    const double s2 = V * H.xi * H.xi * e / H.kappa * ( 1 - e ) + H.theta * H.xi * H.xi / ( 2 * H.kappa ) * ( 1 - e ) * ( 1 - e );
    const double psi = ( m > 0 ) ? s2 / ( m * m ) : 0;
    if ( psi <= 1.5 )
    {
        const double c = 2.0 / psi;
        const double b2 = c - 1.0 + std::sqrt( c ) * std::sqrt( c - 1.0 );
        const double a = m / ( 1.0 + b2 );
        const double bz = std::sqrt( b2 ) + Z;
        return a * bz * bz;
    }
    const double p = ( psi - 1.0 ) / ( psi + 1.0 );
    const double beta = ( 1.0 - p ) / m;
    const double u = 0.5 * std::erfc( -Z / std::sqrt( 2.0 ) );
    return ( u <= p ) ? 0.0 : std::log( ( 1.0 - p ) / ( 1.0 - u ) ) / beta;
}

double LevStep( double Vp, double Vc, double Dt, double Carry, double L, double Z, const StochParams& H )  // single.cpp:369-388
{
    if ( H.xi > 1e-12 )
    {
        const double g = 0.5;
        const double k0 = -L * H.rho * H.kappa * H.theta / H.xi * Dt;
        const double a = g * Dt * ( L * H.kappa * H.rho / H.xi - 0.5 * L * L );
        const double k1 = a - L * H.rho / H.xi;
        const double k2 = a + L * H.rho / H.xi;
        const double k3 = g * Dt * ( 1.0 - H.rho * H.rho ) * L * L;
        double var = k3 * ( Vp + Vc );
        if ( var < 0 )
            var = 0;
        return Carry + k0 + k1 * Vp + k2 * Vc + std::sqrt( var ) * Z;
    }
    return Carry - 0.5 * L * L * Vp * Dt + L * std::sqrt( Vp * Dt ) * Z;
}

template <class Fn>
void ParallelFor( int n, Fn fn )
{
    unsigned nt = std::max( 1u, std::min( 16u, std::thread::hardware_concurrency() ) );
    if ( n < 64 || nt == 1 )
    {
        for ( int i = 0; i < n; i++ )
            fn( i );
        return;
    }
    std::vector<std::thread> th;
    for ( unsigned t = 0; t < nt; t++ )
        th.emplace_back( [&, t] {
            for ( int i = (int)t; i < n; i += (int)nt )
                fn( i );
        } );
    for ( auto& x : th )
        x.join();
}
}  // namespace

int CalibrateLeverage( const Market& mk, int a, const Grid& g, LeverageGrid& lev )
{
    const StochParams h = mk.Stoch( a );
    const std::vector<double>& D = g.days;
    const size_t nd = D.size();
    const double s0 = mk.eq[a].spot - 0.0;
    const double ls0 = std::log( s0 );
    auto cum_carry = [&]( double d ) { return ( mk.rate.Value( d ) - mk.DivRepo( a, d ) ) * YearFraction( 0.0, d ); };
    // calibration dates: every diffusion date (the reference), or for a surface
    // nt of them spread evenly in sqrt(t), at[k] = round(k^2 (nd-1) / (nt-1)^2)
    // and at least one date after the previous one (the leverage moves fastest
    // near t = 0, where linear-in-time reads of evenly spread layers bias LSV
    // vanillas by several trusts at one year)
    const bool surf = mk.lev_time_layers > 0;
    const int P = surf ? mk.lev_spot_points : kGridPoints;
    const size_t nt = surf ? std::min<size_t>( (size_t)mk.lev_time_layers, nd ) : nd;
    std::vector<size_t> at( nt, 0 );
    for ( size_t k = 1; k < nt; k++ )
        at[k] = !surf ? k
                      : std::max( at[k - 1] + 1, ( 2 * k * k * ( nd - 1 ) + ( nt - 1 ) * ( nt - 1 ) ) /
                                                     ( 2 * ( nt - 1 ) * ( nt - 1 ) ) );
    // layer geometry: its own +-8 sigma sqrt(t) axis per date (single.cpp:175-181), or
    // one +-4 sigma sqrt(T) axis at the last date for the whole surface
    std::vector<double> lt( nt ), lstep( nt );
    std::vector<int64_t> loff( nt );
    for ( size_t k = 0; k < nt; k++ )
    {
        const double ax_day = surf ? D[nd - 1] : D[at[k]];
        lt[k] = YearFraction( 0.0, D[at[k]] );
        double T = YearFraction( 0.0, ax_day );
        if ( T <= 0 )
            T = 1.0 / kDaysPerYear;
        const double v_atm = mk.ImplicitVol( a, 0, ax_day );
        const double hw = ( surf ? kSurfaceSigmas : kGridSigmas ) * v_atm * std::sqrt( T );
        double ln_step = 2.0 * hw / ( P - 1 );
        if ( ln_step <= 0 )
            ln_step = 1e-4;
        lstep[k] = ln_step;
        loff[k] = (long)std::floor( ( ls0 - hw ) / ln_step );
    }
    // the Dupire target on every layer is particle-free: evaluated first, in parallel
    std::vector<double> sig( nt * P ), layers( nt * P, 0.0 );
    ParallelFor( (int)( nt * P ), [&]( int j ) {
        const size_t k = j / P;
        const int i = j % P;
        const double x = ( loff[k] + i ) * lstep[k];
        sig[j] = mk.LocalVol( a, std::exp( x ), D[at[k]] );
    } );
    const std::vector<double>& Z = CalibNormals( nd > 1 ? nd - 1 : 1 );
    std::vector<double> spot( kParticles, s0 ), var( kParticles, h.v0 ), ln_spot( kParticles, ls0 );
    std::vector<double> bsum( kBins ), bmean( kBins );
    std::vector<int> bcnt( kBins );
    size_t next = 0;  // next layer to calibrate
    for ( size_t k = 0; k < nd; k++ )
    {
        if ( next < nt && at[next] == k )
        {
            // ConditionalVariance of the current ensemble (single.cpp:263-316)
            double fallback = 0;
            for ( double v : var )
                fallback += v;
            fallback /= (double)var.size();
            const auto [mn, mx] = std::minmax_element( ln_spot.begin(), ln_spot.end() );
            const bool flat = *mx - *mn < 1e-10;
            double x0 = 0, dx = 1;
            if ( !flat )
            {
                const double width = ( *mx - *mn ) / kBins;
                std::fill( bsum.begin(), bsum.end(), 0.0 );
                std::fill( bcnt.begin(), bcnt.end(), 0 );
                for ( int p = 0; p < kParticles; p++ )
                {
                    int b = (int)( ( ln_spot[p] - *mn ) / width );
                    b = std::clamp( b, 0, kBins - 1 );
                    bsum[b] += var[p];
                    bcnt[b]++;
                }
                std::fill( bmean.begin(), bmean.end(), fallback );
                for ( int b = 0; b < kBins; b++ )
                    if ( bcnt[b] >= kMinBinCount )
                        bmean[b] = bsum[b] / bcnt[b];
                for ( int b = 0; b < kBins; b++ )
                {
                    if ( bcnt[b] >= kMinBinCount )
                        continue;
                    for ( int d = 1; d < kBins; d++ )
                    {
                        if ( b - d >= 0 && bcnt[b - d] >= kMinBinCount )
                        {
                            bmean[b] = bsum[b - d] / bcnt[b - d];
                            break;
                        }
                        if ( b + d < kBins && bcnt[b + d] >= kMinBinCount )
                        {
                            bmean[b] = bsum[b + d] / bcnt[b + d];
                            break;
                        }
                    }
                }
                x0 = *mn + 0.5 * width;
                dx = width;
            }
            auto cond = [&]( double X )  // single.cpp:318-337
            {
                if ( flat )
                    return fallback;
                const double f = ( X - x0 ) / dx;
                const double last = (double)( kBins - 1 );
                if ( f <= 0 )
                    return bmean.front();
                if ( f >= last )
                    return bmean.back();
                const size_t i = (size_t)f;
                const double w = f - (double)i;
                return ( 1.0 - w ) * bmean[i] + w * bmean[i + 1];
            };
            double* layer = &layers[next * P];
            for ( int i = 0; i < P; i++ )
            {
                const double x = ( loff[next] + i ) * lstep[next];
                const double ev = std::max( cond( x ), kMinCondVar );
                layer[i] = std::clamp( sig[next * P + i] / std::sqrt( ev ), kLevMin, kLevMax );
            }
            next++;
        }
        if ( k == 0 )
            continue;
        // the particles step into date k on the latest calibrated layer
        const size_t cur = next - 1;
        const double* layer = &layers[cur * P];
        const double dt = YearFraction( D[k - 1], D[k] );
        const double carry = cum_carry( D[k] ) - cum_carry( D[k - 1] );
        const double* z = &Z[( k - 1 ) * 2 * kParticles];
        const double lnstep = lstep[cur], off = (double)loff[cur];
        ParallelFor( kParticles / 256, [&]( int blk ) {
            for ( int p = blk * 256; p < ( blk + 1 ) * 256; p++ )
            {
                // LeverageSurface::LayerAt (leverage_surface.cpp:19-40)
                const double f = std::log( spot[p] ) / lnstep - off;
                double lp;
                if ( !( f > 0.0 ) )
                    lp = layer[0];
                else if ( f >= (double)( P - 1 ) )
                    lp = layer[P - 1];
                else
                {
                    const size_t i = (size_t)f;
                    const double w = f - (double)i;
                    lp = ( 1.0 - w ) * layer[i] + w * layer[i + 1];
                }
                const double vp = var[p];
                const double vc = QeStep( vp, dt, z[2 * p], h );
                const double step = LevStep( vp, vc, dt, carry, lp, z[2 * p + 1], h );
                var[p] = vc;
                spot[p] *= std::exp( step );
                ln_spot[p] += step;
            }
        } );
    }
    lev.points = P;
    if ( !surf )
    {
        lev.t = std::move( lt );
        lev.ln_step = std::move( lstep );
        lev.offset = std::move( loff );
        lev.values = std::move( layers );
        return THOTH_OK;
    }
    // per-date layers of the surface: LeverageSurface::GetLeverage at the date's
    // time (leverage_surface.cpp:43-66), exact on the shared axis since linear
    // interpolation in time and in log-spot commute there
    lev.surface.t = lt;
    lev.surface.ln_step = lstep[0];
    lev.surface.offset = loff[0];
    lev.surface.values = layers;
    lev.t.assign( nd, 0 );
    lev.ln_step.assign( nd, lstep[0] );
    lev.offset.assign( nd, loff[0] );
    lev.values.assign( nd * P, 0 );
    for ( size_t i = 0; i < nd; i++ )
    {
        const double T = YearFraction( 0.0, D[i] );
        lev.t[i] = T;
        double* out = &lev.values[i * P];
        if ( T <= lt.front() || T >= lt.back() )
        {
            const double* src = &layers[( T <= lt.front() ? 0 : nt - 1 ) * P];
            std::copy( src, src + P, out );
            continue;
        }
        size_t hi = 1;
        while ( lt[hi] < T )
            hi++;
        const size_t lo = hi - 1;
        const double w = ( T - lt[lo] ) / ( lt[hi] - lt[lo] );
        for ( int j = 0; j < P; j++ )
            out[j] = ( 1.0 - w ) * layers[lo * P + j] + w * layers[hi * P + j];
    }
    return THOTH_OK;
}

void LeverageTangent( const Market& mk, int a, const Grid& g, const LeverageGrid& L0, int which,
                      std::vector<double>& dv )
{
    // the reference's bumps: +1 % of the spot (helpers/constants.hpp:36), +1
    // vol point (tasks/pricer.hpp:25), each recalibrating the leverage
    const double h = which == 0 ? kSpotBump * mk.eq[a].spot : kVolBump;
    Market up = mk;
    if ( which == 0 )
        up.eq[a].spot += h;
    else
        up.eq[a].vol_shift += h;
    LeverageGrid Lu;
    if ( CalibrateLeverage( up, a, g, Lu ) != THOTH_OK || Lu.points != L0.points )
        throw Invalid( THOTH_ERR_INVALID, "leverage tangent: calibration failed" );
    const int P = L0.points, last = P - 1;
    const size_t nd = g.t.size();
    dv.assign( nd * P, 0.0 );
    for ( size_t i = 0; i < nd; i++ )
    {
        const double* lu = &Lu.values[i * P];
        for ( int k = 0; k < P; k++ )
        {
            // the bumped layer read at the base layer's node k, as the sweep
            // reads a layer (LeverageSurface::LayerAt, leverage_surface.cpp:19-40)
            const double x = (double)( k + L0.offset[i] ) * L0.ln_step[i];
            const double f = x / Lu.ln_step[i] - (double)Lu.offset[i];
            double v;
            if ( f <= 0.0 )
                v = lu[0];
            else if ( f >= (double)last )
                v = lu[last];
            else
            {
                const int j = (int)f;
                const double w = f - (double)j;
                v = ( 1.0 - w ) * lu[j] + w * lu[j + 1];
            }
            dv[i * P + k] = ( v - L0.values[i * P + k] ) / h;
        }
    }
}

}  // namespace thoth::host
