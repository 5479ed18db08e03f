// mcl_oracle.cpp -- TEST INFRASTRUCTURE ONLY: CPU FP64 restatement of the
// reference `mcl` engine (see mcl_oracle.hpp).  Compiled with -ffp-contract=off.
#include "mcl_oracle.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>

namespace oracle
{

// ============================================================================
// Joe-Kuo table + Sobol  (reference: nodes/sobol_generator.cpp:12-121,
// nodes/sobol_direction_numbers.cpp:3-16 for the file-entry -> dimension map)
// ============================================================================
namespace
{
std::vector<unsigned> g_s, g_a;
std::vector<std::vector<unsigned>> g_m;
}  // namespace

bool LoadJoeKuo( const std::string& path )
{
    std::ifstream in( path );
    if ( !in )
        return false;
    std::string line;
    std::getline( in, line );  // header "d s a m_i"
    g_s.clear();
    g_a.clear();
    g_m.clear();
    while ( std::getline( in, line ) )
    {
        std::istringstream ls( line );
        unsigned d, s, a;
        if ( !( ls >> d >> s >> a ) )
            continue;
        std::vector<unsigned> m( s );
        for ( unsigned i = 0; i < s; i++ )
            ls >> m[i];
        g_s.push_back( s );
        g_a.push_back( a );
        g_m.push_back( m );
    }
    return g_s.size() == 4096;
}

// sobol_generator.cpp:12-15: FILE_DIMS tabulated dimensions + the implicit first one
unsigned SobolMaxDimension() { return (unsigned)g_s.size() + 1; }

// sobol_generator.cpp:22-76
Sobol::Sobol( unsigned dim ) : _dim( dim )
{
    const unsigned BITS = 32;
    if ( _dim > SobolMaxDimension() )
        _dim = SobolMaxDimension();
    _v.assign( _dim, std::vector<uint32_t>( BITS, 0 ) );
    _x.assign( _dim, 0 );
    for ( unsigned i = 0; i < BITS; i++ )
        _v[0][i] = 1u << ( BITS - 1 - i );
    for ( unsigned d = 1; d < _dim; d++ )
    {
        unsigned idx = d - 1;
        unsigned s = g_s[idx];
        unsigned a = g_a[idx];
        const std::vector<unsigned>& m = g_m[idx];
        std::vector<uint32_t>& V = _v[d];
        if ( BITS <= s )
        {
            for ( unsigned i = 0; i < BITS; i++ )
                V[i] = (uint32_t)m[i] << ( BITS - 1 - i );
        }
        else
        {
            for ( unsigned i = 0; i < s; i++ )
                V[i] = (uint32_t)m[i] << ( BITS - 1 - i );
            for ( unsigned i = s; i < BITS; i++ )
            {
                V[i] = V[i - s] ^ ( V[i - s] >> s );
                for ( unsigned k = 1; k < s; k++ )
                    V[i] ^= ( ( a >> ( s - 1 - k ) ) & 1u ) * V[i - k];
            }
        }
    }
}

// sobol_generator.cpp:80-100 (integer stage)
void Sobol::NextInt( std::vector<uint32_t>& x )
{
    unsigned c = 0;
    for ( uint64_t value = _count; value & 1; value >>= 1 )
        c++;
    x.resize( _dim );
    for ( unsigned d = 0; d < _dim; d++ )
    {
        _x[d] ^= _v[d][c];
        x[d] = _x[d];
    }
    _count++;
}

// sobol_generator.cpp:80-100
void Sobol::Next( std::vector<double>& point )
{
    const double SCALE = 1.0 / 4294967296.0;
    point.resize( _dim );
    unsigned c = 0;
    for ( uint64_t value = _count; value & 1; value >>= 1 )
        c++;
    for ( unsigned d = 0; d < _dim; d++ )
    {
        _x[d] ^= _v[d][c];
        point[d] = _x[d] * SCALE;
    }
    _count++;
}

// sobol_generator.cpp:105-121
void Sobol::Skip( uint64_t count )
{
    for ( uint64_t n = 0; n < count; n++ )
    {
        unsigned c = 0;
        for ( uint64_t value = _count; value & 1; value >>= 1 )
            c++;
        for ( unsigned d = 0; d < _dim; d++ )
            _x[d] ^= _v[d][c];
        _count++;
    }
}

// ============================================================================
// distributions (reference: helpers/distributions.cpp:21-81)
// ============================================================================
double NormalCdf( double x ) { return 0.5 * std::erfc( -x * M_SQRT1_2 ); }

double NormalCdfInv( double P )
{
    static const double a[] = { -3.969683028665376e+01, 2.209460984245205e+02, -2.759285104469687e+02,
                                1.383577518672690e+02, -3.066479806614716e+01, 2.506628277459239e+00 };
    static const double b[] = { -5.447609879822406e+01, 1.615858368580409e+02, -1.556989798598866e+02,
                                6.680131188771972e+01, -1.328068155288572e+01 };
    static const double c[] = { -7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00,
                                -2.549732539343734e+00, 4.374664141464968e+00, 2.938163982698783e+00 };
    static const double d[] = { 7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00,
                                3.754408661907416e+00 };
    const double p_low = 0.02425;
    const double p_high = 1.0 - p_low;
    if ( P <= 0.0 )
        return -HUGE_VAL;
    if ( P >= 1.0 )
        return HUGE_VAL;
    if ( P < p_low )
    {
        double q = std::sqrt( -2.0 * std::log( P ) );
        return ( ( ( ( ( c[0] * q + c[1] ) * q + c[2] ) * q + c[3] ) * q + c[4] ) * q + c[5] ) /
               ( ( ( ( d[0] * q + d[1] ) * q + d[2] ) * q + d[3] ) * q + 1.0 );
    }
    if ( P <= p_high )
    {
        double q = P - 0.5;
        double r = q * q;
        return ( ( ( ( ( a[0] * r + a[1] ) * r + a[2] ) * r + a[3] ) * r + a[4] ) * r + a[5] ) * q /
               ( ( ( ( ( b[0] * r + b[1] ) * r + b[2] ) * r + b[3] ) * r + b[4] ) * r + 1.0 );
    }
    double q = std::sqrt( -2.0 * std::log( 1.0 - P ) );
    return -( ( ( ( ( c[0] * q + c[1] ) * q + c[2] ) * q + c[3] ) * q + c[4] ) * q + c[5] ) /
           ( ( ( ( d[0] * q + d[1] ) * q + d[2] ) * q + d[3] ) * q + 1.0 );
}

// ============================================================================
// xoshiro256++ / splitmix64 (reference: helpers/rng.hpp:36-72)
// ============================================================================
static inline uint64_t Rotl( uint64_t x, int k ) { return ( x << k ) | ( x >> ( 64 - k ) ); }

void Rng::Seed( uint64_t seed )
{
    uint64_t z = seed;
    for ( int i = 0; i < 4; i++ )
    {
        z += 0x9e3779b97f4a7c15ULL;
        uint64_t x = z;
        x = ( x ^ ( x >> 30 ) ) * 0xbf58476d1ce4e5b9ULL;
        x = ( x ^ ( x >> 27 ) ) * 0x94d049bb133111ebULL;
        _s[i] = x ^ ( x >> 31 );
    }
}

uint64_t Rng::operator()()
{
    const uint64_t result = Rotl( _s[0] + _s[3], 23 ) + _s[0];
    const uint64_t t = _s[1] << 17;
    _s[2] ^= _s[0];
    _s[3] ^= _s[1];
    _s[1] ^= _s[2];
    _s[0] ^= _s[3];
    _s[2] ^= t;
    _s[3] = Rotl( _s[3], 45 );
    return result;
}

double Rng::Uniform()
{
    const double u = ( ( *this )() >> 11 ) * ( 1.0 / 9007199254740992.0 );
    return u > 0.0 ? u : ( 1.0 / 9007199254740992.0 );
}

// ============================================================================
// Brownian bridge (reference: nodes/path_generator.cpp:49-198)
// ============================================================================
BridgeSchedule BuildBridgeSchedule( const std::vector<double>& times )
{
    const int m = (int)times.size() - 1;
    BridgeSchedule s;
    s.bridge_index.assign( m, 0 );
    s.left_index.assign( m, 0 );
    s.right_index.assign( m, 0 );
    s.left_weight.assign( m, 0.0 );
    s.right_weight.assign( m, 0.0 );
    s.std_dev.assign( m, 0.0 );
    if ( m == 0 )
        return s;
    auto bt = [&]( int k ) { return times[k + 1]; };
    std::vector<int> map( m, 0 );
    map[m - 1] = 1;
    s.bridge_index[0] = m - 1;
    s.std_dev[0] = std::sqrt( bt( m - 1 ) );
    s.left_weight[0] = s.right_weight[0] = 0.0;
    int j = 0;
    for ( int i = 1; i < m; i++ )
    {
        while ( map[j] )
            j++;
        int k = j;
        while ( !map[k] )
            k++;
        int l = j + ( ( k - 1 - j ) >> 1 );
        map[l] = i + 1;
        s.bridge_index[i] = l;
        s.left_index[i] = j;
        s.right_index[i] = k;
        if ( j != 0 )
        {
            s.left_weight[i] = ( bt( k ) - bt( l ) ) / ( bt( k ) - bt( j - 1 ) );
            s.right_weight[i] = ( bt( l ) - bt( j - 1 ) ) / ( bt( k ) - bt( j - 1 ) );
            s.std_dev[i] = std::sqrt( ( bt( l ) - bt( j - 1 ) ) * ( bt( k ) - bt( l ) ) / ( bt( k ) - bt( j - 1 ) ) );
        }
        else
        {
            s.left_weight[i] = ( bt( k ) - bt( l ) ) / bt( k );
            s.right_weight[i] = bt( l ) / bt( k );
            s.std_dev[i] = std::sqrt( bt( l ) * ( bt( k ) - bt( l ) ) / bt( k ) );
        }
        j = k + 1;
        if ( j >= m )
            j = 0;
    }
    return s;
}

PathGenerator::PathGenerator( const std::vector<double>& times, int factors, Rng* rng, uint64_t sobol_skip,
                              bool use_sobol )
    : _t( times ), _m( (int)times.size() - 1 ), _factors( factors ), _use_sobol( use_sobol ), _rng( rng )
{
    _b = BuildBridgeSchedule( _t );
    if ( _m > 0 && _use_sobol )
    {
        long total = (long)_m * _factors;
        _sobol_dim = (int)std::min<long>( total, SobolMaxDimension() );
        _sobol_store.emplace_back( (unsigned)_sobol_dim );
        _sobol = &_sobol_store.back();
        if ( sobol_skip > 0 )
            _sobol->Skip( sobol_skip );
    }
    _noise.assign( _factors, std::vector<double>( _m + 1, 0.0 ) );
    _z.assign( _factors, std::vector<double>( _m, 0.0 ) );
    _u.assign( _sobol_dim, 0.0 );
    _w.assign( _m, 0.0 );
}

void PathGenerator::NextPath()
{
    if ( _m == 0 )
        return;
    if ( !_use_sobol )
    {
        // independent draws, no bridge: the normalized increment of step i is
        // the draw itself (noise_node.cpp:20-25)
        for ( int i = 1; i <= _m; i++ )
            for ( int f = 0; f < _factors; f++ )
                _noise[f][i] = _rng->Gaussian();
        return;
    }
    if ( _sobol )
        _sobol->Next( _u );
    for ( int step = 0; step < _m; step++ )
    {
        for ( int f = 0; f < _factors; f++ )
        {
            long dim = (long)step * _factors + f;
            double z;
            if ( dim < _sobol_dim )
            {
                double u = std::min( 1.0 - 1e-12, std::max( 1e-12, _u[dim] ) );
                z = NormalCdfInv( u );
            }
            else
            {
                z = _rng->Gaussian();
            }
            _z[f][step] = z;
        }
    }
    for ( int f = 0; f < _factors; f++ )
    {
        const std::vector<double>& z = _z[f];
        _w[_b.bridge_index[0]] = _b.std_dev[0] * z[0];
        for ( int i = 1; i < _m; i++ )
        {
            int li = _b.left_index[i];
            int ri = _b.right_index[i];
            int bi = _b.bridge_index[i];
            if ( li != 0 )
                _w[bi] = _b.left_weight[i] * _w[li - 1] + _b.right_weight[i] * _w[ri] + _b.std_dev[i] * z[i];
            else
                _w[bi] = _b.right_weight[i] * _w[ri] + _b.std_dev[i] * z[i];
        }
        std::vector<double>& n = _noise[f];
        double prev = 0.0;
        for ( int i = 1; i <= _m; i++ )
        {
            double w = _w[i - 1];
            double dt = _t[i] - _t[i - 1];
            n[i] = ( dt > 0 ) ? ( w - prev ) / std::sqrt( dt ) : 0.0;
            prev = w;
        }
    }
}

// ============================================================================
// Cholesky (reference: helpers/maths.cpp:354-381), row-major n x n in place
// ============================================================================
bool CholeskyDecomposeLower( std::vector<double>& A, int n )
{
    for ( int i = 0; i < n; i++ )
    {
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
            {
                A[i * n + j] = sum / A[j * n + j];
            }
        }
    }
    return true;
}

// ============================================================================
// curves (reference: marketdata/curve.cpp:45-73, yield_curve.cpp:16-21)
// ============================================================================
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
        const double r0 = values[i];
        const double r1 = values[i + 1];
        const double w = ( ( day - days[i] ) / NB_OF_DAYS_A_YEAR ) / ( ( days[i + 1] - days[i] ) / NB_OF_DAYS_A_YEAR );
        r = r0 + w * ( r1 - r0 );
    }
    return r + shift;
}

static inline double YearFraction( double from, double to ) { return ( to - from ) / NB_OF_DAYS_A_YEAR; }

// ============================================================================
// term-structured correlation (reference: vol_correl/correlation.cpp:528-620)
// ============================================================================
// Average of entry (i, j) over [a, b] (years): the entry is linear in time
// between pillars and flat outside, integrated exactly (Correlation::AvgEntry
// over EntryPrimitive); a degenerate window is the instantaneous value at a.
double Market::AvgCorrelation( int i, int j, double a, double b ) const
{
    const size_t n = eq.size();
    if ( !TermStructured() )
        return correlation[i * n + j];
    const std::vector<double>& m = corr_pillars;
    auto entry = [&]( size_t p ) { return corr_mats[p][i * n + j]; };
    auto primitive = [&]( double t )  // EntryPrimitive, correlation.cpp:539-559
    {
        if ( t <= m.front() )
            return entry( 0 ) * t;
        double F = entry( 0 ) * m.front();
        for ( size_t p = 1; p < m.size(); p++ )
        {
            if ( t <= m[p] )
            {
                const double w = ( t - m[p - 1] ) / ( m[p] - m[p - 1] );
                const double e_t = entry( p - 1 ) + w * ( entry( p ) - entry( p - 1 ) );
                return F + ( entry( p - 1 ) + e_t ) / 2 * ( t - m[p - 1] );
            }
            F += ( entry( p - 1 ) + entry( p ) ) / 2 * ( m[p] - m[p - 1] );
        }
        return F + entry( m.size() - 1 ) * ( t - m.back() );
    };
    a = std::max( a, 0.0 );
    b = std::max( b, a );
    if ( b - a < 1e-12 )  // correlation.cpp:571-590
    {
        if ( a <= m.front() )
            return entry( 0 );
        if ( a >= m.back() )
            return entry( m.size() - 1 );
        for ( size_t p = 1; p < m.size(); p++ )
            if ( a <= m[p] )
            {
                const double w = ( a - m[p - 1] ) / ( m[p] - m[p - 1] );
                return entry( p - 1 ) + w * ( entry( p ) - entry( p - 1 ) );
            }
    }
    return ( primitive( b ) - primitive( a ) ) / ( b - a );
}

// ============================================================================
// SABR (reference: vol_correl/sabr_volatility.cpp:40-239)
// ============================================================================
double Sabr::Interp( const std::vector<double>& values, double T ) const
{
    if ( T <= maturities.front() )
        return values.front();
    if ( T >= maturities.back() )
        return values.back();
    for ( size_t i = 1; i < maturities.size(); i++ )
    {
        if ( T <= maturities[i] )
        {
            double w = ( T - maturities[i - 1] ) / ( maturities[i] - maturities[i - 1] );
            return values[i - 1] + w * ( values[i] - values[i - 1] );
        }
    }
    return values.back();
}

namespace
{
double BlackCall( double F, double K, double V, double T )
{
    if ( V <= 0 || T <= 0 )
        return std::max( F - K, 0.0 );
    const double s = V * std::sqrt( T );
    const double d1 = std::log( F / K ) / s + 0.5 * s;
    return F * NormalCdf( d1 ) - K * NormalCdf( d1 - s );
}
double BlackPut( double F, double K, double V, double T )
{
    if ( V <= 0 || T <= 0 )
        return std::max( K - F, 0.0 );
    const double s = V * std::sqrt( T );
    const double d1 = std::log( F / K ) / s + 0.5 * s;
    return K * NormalCdf( s - d1 ) - F * NormalCdf( -d1 );
}
double TailImpliedVol( double F, double K, double T, double Price, bool IsCall )
{
    double lo = 0.0, hi = IMPLICIT_VOL_MAX;
    for ( int i = 0; i < 100 && hi - lo > 1e-13; i++ )
    {
        const double mid = 0.5 * ( lo + hi );
        const double p = IsCall ? BlackCall( F, K, mid, T ) : BlackPut( F, K, mid, T );
        if ( p < Price )
            lo = mid;
        else
            hi = mid;
    }
    return 0.5 * ( lo + hi );
}
}  // namespace

double Sabr::HaganVol( double K, double F, double T ) const
{
    double a_ = Interp( alpha, T );
    double b_ = Interp( beta, T );
    double r_ = Interp( rho, T );
    double n_ = Interp( nu, T );
    double one_mb = 1.0 - b_;
    double log_fk = std::log( F / K );
    double fk_pow = std::pow( F * K, one_mb / 2.0 );
    double den = fk_pow * ( 1.0 + one_mb * one_mb / 24.0 * log_fk * log_fk +
                            one_mb * one_mb * one_mb * one_mb / 1920.0 * log_fk * log_fk * log_fk * log_fk );
    double z = ( n_ / a_ ) * fk_pow * log_fk;
    double zxz = 1.0;
    if ( std::fabs( z ) > 1e-12 )
    {
        double xz = std::log( ( std::sqrt( 1.0 - 2.0 * r_ * z + z * z ) + z - r_ ) / ( 1.0 - r_ ) );
        zxz = z / xz;
    }
    double corr = 1.0 + ( one_mb * one_mb / 24.0 * a_ * a_ / std::pow( F * K, one_mb ) +
                          0.25 * r_ * b_ * n_ * a_ / fk_pow + ( 2.0 - 3.0 * r_ * r_ ) / 24.0 * n_ * n_ ) *
                            T;
    return ( a_ / den ) * zxz * corr + 0.0;
}

double Sabr::WingVol( double K, double F, double T, double Cutoff, bool Right ) const
{
    const double h = 1e-4 * Cutoff;
    auto price = [&]( double k )
    {
        const double v = HaganVol( k, F, T );
        return Right ? BlackCall( F, k, v, T ) : BlackPut( F, k, v, T );
    };
    const double p0 = price( Cutoff );
    const double dp = ( price( Cutoff + h ) - price( Cutoff - h ) ) / ( 2 * h );
    if ( p0 <= 0 )
        return HaganVol( K, F, T );
    if ( Right )
    {
        const double a = -Cutoff * dp / p0;
        if ( a <= 0 )
            return HaganVol( K, F, T );
        const double c = p0 * std::pow( Cutoff / K, a );
        return TailImpliedVol( F, K, T, c, true );
    }
    const double b = Cutoff * dp / p0;
    if ( b <= 1 )
        return HaganVol( K, F, T );
    const double pt = p0 * std::pow( K / Cutoff, b );
    return TailImpliedVol( F, K, T, pt, false );
}

double Sabr::ImplicitVol( double Strike, double Forward, double day ) const
{
    double T = YearFraction( 0.0, day );
    if ( T <= 0 )
        T = 1.0 / NB_OF_DAYS_A_YEAR;
    const double F = Forward;
    const double K = ( Strike > 0 ) ? Strike : F;
    const double atm = HaganVol( F, F, T );
    const double band = 2.5 * atm * std::sqrt( T );
    const double k_lo = F * std::exp( -band );
    const double k_hi = F * std::exp( band );
    if ( K > k_hi )
        return WingVol( K, F, T, k_hi, true );
    if ( K < k_lo )
        return WingVol( K, F, T, k_lo, false );
    return HaganVol( K, F, T );
}

// ============================================================================
// market accessors
// ============================================================================
// heston_volatility.hpp:40-47 / lsv_volatility.hpp:45-55
StochParams Market::Stoch( int a ) const
{
    const Equity& e = eq[a];
    if ( e.vol_kind == THOTH_VOL_LSV )
        return { e.v0 + 0.0, e.kappa + 0.0, e.theta + 0.0, e.xi + 0.0, e.rho + 0.0 };
    auto shifted = [&]( double v )
    {
        double s = std::sqrt( v ) + e.vol_shift;
        return s * s;
    };
    return { shifted( e.v0 ) + 0.0, e.kappa + 0.0, shifted( e.theta ) + 0.0, e.xi + 0.0, e.rho + 0.0 };
}

// equity.cpp:149-154
double Market::DivRepo( int a, double day ) const
{
    const Equity& e = eq[a];
    double y = e.has_div ? e.div : 0;  // flat continuous-dividend curve value
    y += 0;
    (void)day;
    return y;
}

// equity.cpp:160-191 with yield_curve.cpp:16-21
double Market::Forward( int a, double day ) const
{
    double dt = YearFraction( today, day );
    double r = rate.Value( day );
    double df = std::exp( -r * YearFraction( today, day ) );
    double div = DivRepo( a, day );
    double spot = eq[a].spot - 0.0;
    return spot * std::exp( -dt * ( div + 0 ) ) / df;
}

// single.cpp:56-62 -> {bs,heston,lsv}_volatility GetImplicitVol
double Market::ImplicitVol( int a, double strike, double day ) const
{
    const Equity& e = eq[a];
    if ( e.vol_kind == THOTH_VOL_BS )
        return ( e.bs_vol + e.vol_shift ) * 1.0;
    if ( e.vol_kind == THOTH_VOL_HESTON )
        return ( std::sqrt( e.v0 ) + e.vol_shift ) * 1.0;
    return e.surface.ImplicitVol( strike, Forward( a, day ), day ) + e.vol_shift;
}

// equity.cpp:196-211 + volatility.cpp:55-128 (the vol object's GetImplicitVol is virtual)
double Market::LocalVol( int a, double Strike, double day ) const
{
    const Equity& e = eq[a];
    double r = rate.Value( day );
    double q = 0;
    if ( e.has_div )
        q += e.div;
    const double S = e.spot;
    const double K = Strike;
    auto IV = [&]( double k, double f, double d )
    {
        if ( e.vol_kind == THOTH_VOL_BS )
            return ( e.bs_vol + e.vol_shift ) * 1.0;
        if ( e.vol_kind == THOTH_VOL_HESTON )
            return ( std::sqrt( e.v0 ) + e.vol_shift ) * 1.0;
        return e.surface.ImplicitVol( k, f, d ) + e.vol_shift;
    };
    double T = YearFraction( today, day );
    const double dK = .01;
    const double dT = 1.0 / NB_OF_DAYS_A_YEAR;
    double F = S * std::exp( ( r - q ) * T );
    double F_Td = S * std::exp( ( r - q ) * YearFraction( today, day - 1 ) );
    double F_Tu = S * std::exp( ( r - q ) * YearFraction( today, day + 1 ) );
    double V = IV( K, F, day );
    double V_Kd = IV( K - dK, F, day );
    double V_Ku = IV( K + dK, F, day );
    double V_Td = IV( K, F_Td, day - 1 );
    double V_Tu = IV( K, F_Tu, day + 1 );
    double sT = std::sqrt( T );
    double V2 = V * V;
    double d1 = ( std::log( S / K ) + ( r - q + .5 * V2 ) * T ) / V / sT;
    double dVdK = ( V_Ku - V_Kd ) / ( 2 * dK );
    double dVdK2 = ( V_Ku + V_Kd - 2 * V ) / ( dK * dK );
    double dVdT = ( V_Tu - V_Td ) / ( 2 * dT );
    double N = 2 * V * T * ( dVdT + ( r - q ) * K * dVdK ) + V2;
    double D1 = 1 + K * d1 * dVdK * sT;
    double D = D1 * D1 + K * K * T * V * ( dVdK2 - d1 * dVdK * dVdK * sT );
    const double local_var = N / D;
    const double floor = 1e-8;
    const double cap = 5.0 * V * ( 5.0 * V );
    if ( local_var > cap )
        return 5.0 * V;
    return std::sqrt( local_var > floor ? local_var : floor );
}

// ============================================================================
// LSV leverage calibration (reference: marketdata/single.cpp:231-473,
// vol_correl/leverage_surface.cpp:19-40)
// ============================================================================
namespace
{
constexpr int LOCAL_VOL_GRID_POINTS = 201;
constexpr double LOCAL_VOL_GRID_SIGMA_FACTOR = 8.0;
constexpr double SURFACE_SIGMA_FACTOR = 4.0;  // the engine's surface layout: one +-4 sigma sqrt(T) axis
constexpr int LSV_CALIB_PATHS = 16384;
constexpr int LSV_CALIB_BINS = 30;
constexpr int LSV_CALIB_MIN_BIN_COUNT = 20;
constexpr double LSV_MIN_COND_VARIANCE = 1e-6;
constexpr double LSV_LEVERAGE_MIN = 0.05;
constexpr double LSV_LEVERAGE_MAX = 10.0;
constexpr uint64_t LSV_CALIB_SEED = 20260708;

class ConditionalVariance
{
    double _x0 = 0, _dx = 1;
    std::vector<double> _mean;
    double _fallback = 0;

  public:
    ConditionalVariance( const std::vector<double>& LnSpot, const std::vector<double>& Var )
    {
        _fallback = 0;
        for ( double v : Var )
            _fallback += v;
        _fallback /= (double)Var.size();
        const auto [mn, mx] = std::minmax_element( LnSpot.begin(), LnSpot.end() );
        if ( *mx - *mn < 1e-10 )
            return;
        const double width = ( *mx - *mn ) / LSV_CALIB_BINS;
        std::vector<double> sum( LSV_CALIB_BINS, 0.0 );
        std::vector<int> count( LSV_CALIB_BINS, 0 );
        for ( size_t p = 0; p < LnSpot.size(); p++ )
        {
            int b = (int)( ( LnSpot[p] - *mn ) / width );
            b = std::clamp( b, 0, LSV_CALIB_BINS - 1 );
            sum[b] += Var[p];
            count[b]++;
        }
        _mean.assign( LSV_CALIB_BINS, _fallback );
        for ( int b = 0; b < LSV_CALIB_BINS; b++ )
            if ( count[b] >= LSV_CALIB_MIN_BIN_COUNT )
                _mean[b] = sum[b] / count[b];
        for ( int b = 0; b < LSV_CALIB_BINS; b++ )
        {
            if ( count[b] >= LSV_CALIB_MIN_BIN_COUNT )
                continue;
            for ( int d = 1; d < LSV_CALIB_BINS; d++ )
            {
                if ( b - d >= 0 && count[b - d] >= LSV_CALIB_MIN_BIN_COUNT )
                {
                    _mean[b] = sum[b - d] / count[b - d];
                    break;
                }
                if ( b + d < LSV_CALIB_BINS && count[b + d] >= LSV_CALIB_MIN_BIN_COUNT )
                {
                    _mean[b] = sum[b + d] / count[b + d];
                    break;
                }
            }
        }
        _x0 = *mn + 0.5 * width;
        _dx = width;
    }

    double operator()( double X ) const
    {
        if ( _mean.empty() )
            return _fallback;
        const double f = ( X - _x0 ) / _dx;
        const double last = (double)( _mean.size() - 1 );
        if ( f <= 0 )
            return _mean.front();
        if ( f >= last )
            return _mean.back();
        const size_t i = (size_t)f;
        const double w = f - (double)i;
        return ( 1.0 - w ) * _mean[i] + w * _mean[i + 1];
    }
};

// single.cpp:346-365 (== heston_variance_node.cpp:17-57)
double QeVarianceStep( double V, double Dt, double Z, const StochParams& H )
{
    const double e = std::exp( -H.kappa * Dt );
    const double m = H.theta + ( V - H.theta ) * e;
    const double s2 = V * H.xi * H.xi * e / H.kappa * ( 1 - e ) +
                      H.theta * H.xi * H.xi / ( 2 * H.kappa ) * ( 1 - e ) * ( 1 - e );
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

// single.cpp:369-388
double LeveragedLogSpotStep( double Vp, double Vc, double Dt, double Carry, double L, double Z, const StochParams& H )
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
}  // namespace

double LeverageLayers::LayerAt( size_t K, double S ) const
{
    const std::vector<double>& lev = levels[K];
    const double f = std::log( S ) / ln_step[K] - (double)offset[K];
    const size_t last = lev.size() - 1;
    if ( !( f > 0.0 ) )
        return lev[0];
    if ( f >= (double)last )
        return lev[last];
    const size_t i = (size_t)f;
    const double w = f - (double)i;
    return ( 1.0 - w ) * lev[i] + w * lev[i + 1];
}

double LeverageLayers::GetLeverage( double S, double T ) const
{
    if ( t.empty() )
        return 1.0;
    if ( T <= t.front() )
        return LayerAt( 0, S );
    if ( T >= t.back() )
        return LayerAt( t.size() - 1, S );
    size_t hi = 1;
    while ( t[hi] < T )
        hi++;
    const size_t lo = hi - 1;
    const double w = ( T - t[lo] ) / ( t[hi] - t[lo] );
    return ( 1.0 - w ) * LayerAt( lo, S ) + w * LayerAt( hi, S );
}

namespace
{
// The engine's production layout (no reference counterpart): min(time_layers,
// dates) layers calibrated on diffusion dates spread evenly in sqrt(t), all on
// one log-spot axis of spot_points points, +-4 sigma_ATM(T) sqrt(T) around
// log S0 at the last date T.  The particles step through every
// diffusion date exactly as in Single::CalibrateLeverage (single.cpp:400-473),
// each step on the latest calibrated layer; layer k is built from the ensemble
// entering date at[k].
LeverageLayers CalibrateLeverageSurface( const Market& mk, int a, const std::vector<double>& Dates )
{
    const StochParams h = mk.Stoch( a );
    const double today = Dates.front();
    const double s0 = mk.eq[a].spot - 0.0;
    const int P = mk.lev_spot_points;
    const size_t nd = Dates.size();
    const size_t nt = std::min<size_t>( (size_t)mk.lev_time_layers, nd );
    auto cum_carry = [&]( double d )
    {
        const double r = mk.rate.Value( d );
        return ( r - mk.DivRepo( a, d ) ) * YearFraction( today, d );
    };
    double T = YearFraction( today, Dates.back() );
    if ( T <= 0 )
        T = 1.0 / NB_OF_DAYS_A_YEAR;
    const double v_atm = mk.ImplicitVol( a, 0, Dates.back() );
    const double half_width = SURFACE_SIGMA_FACTOR * v_atm * std::sqrt( T );
    double ln_step = 2.0 * half_width / ( P - 1 );
    if ( ln_step <= 0 )
        ln_step = 1e-4;
    const long offset = (long)std::floor( ( std::log( s0 ) - half_width ) / ln_step );
    Rng rng( LSV_CALIB_SEED );
    std::vector<double> spot( LSV_CALIB_PATHS, s0 );
    std::vector<double> var( LSV_CALIB_PATHS, h.v0 );
    std::vector<double> ln_spot( LSV_CALIB_PATHS, std::log( s0 ) );
    // calibration dates evenly spread in sqrt(t): at[k] = round(k^2 (nd-1) / (nt-1)^2),
    // at least one date after the previous one
    std::vector<size_t> at( nt, 0 );
    for ( size_t k = 1; k < nt; k++ )
        at[k] = std::max( at[k - 1] + 1, ( 2 * k * k * ( nd - 1 ) + ( nt - 1 ) * ( nt - 1 ) ) / ( 2 * ( nt - 1 ) * ( nt - 1 ) ) );
    LeverageLayers lev;
    lev.surface = true;
    size_t next = 0;
    for ( size_t k = 0; k < nd; k++ )
    {
        if ( next < nt && at[next] == k )
        {
            const ConditionalVariance cond_var( ln_spot, var );
            std::vector<double> layer( P );
            for ( int i = 0; i < P; i++ )
            {
                const double x = ( offset + i ) * ln_step;
                const double sig = mk.LocalVol( a, std::exp( x ), Dates[k] );
                const double ev = std::max( cond_var( x ), LSV_MIN_COND_VARIANCE );
                layer[i] = std::clamp( sig / std::sqrt( ev ), LSV_LEVERAGE_MIN, LSV_LEVERAGE_MAX );
            }
            lev.t.push_back( YearFraction( today, Dates[k] ) );
            lev.ln_step.push_back( ln_step );
            lev.offset.push_back( offset );
            lev.levels.push_back( std::move( layer ) );
            next++;
        }
        if ( k == 0 )
            continue;
        const size_t cur = next - 1;
        const double dt = YearFraction( Dates[k - 1], Dates[k] );
        const double carry = cum_carry( Dates[k] ) - cum_carry( Dates[k - 1] );
        for ( int p = 0; p < LSV_CALIB_PATHS; p++ )
        {
            const double lp = lev.LayerAt( cur, spot[p] );
            const double vp = var[p];
            const double vc = QeVarianceStep( vp, dt, rng.Gaussian(), h );
            const double step = LeveragedLogSpotStep( vp, vc, dt, carry, lp, rng.Gaussian(), h );
            var[p] = vc;
            spot[p] *= std::exp( step );
            ln_spot[p] += step;
        }
    }
    return lev;
}
}  // namespace

LeverageLayers CalibrateLeverage( const Market& mk, int a, const std::vector<double>& Dates )
{
    if ( mk.lev_time_layers > 0 )
        return CalibrateLeverageSurface( mk, a, Dates );
    const StochParams h = mk.Stoch( a );
    const double today = Dates.front();
    const double s0 = mk.eq[a].spot - 0.0;
    auto cum_carry = [&]( double d )
    {
        const double r = mk.rate.Value( d );
        return ( r - mk.DivRepo( a, d ) ) * YearFraction( today, d );
    };
    Rng rng( LSV_CALIB_SEED );
    std::vector<double> spot( LSV_CALIB_PATHS, s0 );
    std::vector<double> var( LSV_CALIB_PATHS, h.v0 );
    std::vector<double> ln_spot( LSV_CALIB_PATHS, std::log( s0 ) );
    LeverageLayers lev;
    for ( size_t k = 0; k < Dates.size(); k++ )
    {
        const ConditionalVariance cond_var( ln_spot, var );
        double T = YearFraction( today, Dates[k] );
        const double t_layer = T;
        if ( T <= 0 )
            T = 1.0 / NB_OF_DAYS_A_YEAR;
        const double v_atm = mk.ImplicitVol( a, 0, Dates[k] );
        const double half_width = LOCAL_VOL_GRID_SIGMA_FACTOR * v_atm * std::sqrt( T );
        double ln_step = 2.0 * half_width / ( LOCAL_VOL_GRID_POINTS - 1 );
        if ( ln_step <= 0 )
            ln_step = 1e-4;
        const long offset = (long)std::floor( ( std::log( s0 ) - half_width ) / ln_step );
        std::vector<double> layer( LOCAL_VOL_GRID_POINTS );
        for ( int i = 0; i < LOCAL_VOL_GRID_POINTS; i++ )
        {
            const double x = ( offset + i ) * ln_step;
            const double sig = mk.LocalVol( a, std::exp( x ), Dates[k] );
            const double ev = std::max( cond_var( x ), LSV_MIN_COND_VARIANCE );
            layer[i] = std::clamp( sig / std::sqrt( ev ), LSV_LEVERAGE_MIN, LSV_LEVERAGE_MAX );
        }
        lev.t.push_back( t_layer );
        lev.ln_step.push_back( ln_step );
        lev.offset.push_back( offset );
        lev.levels.push_back( std::move( layer ) );
        if ( k == 0 )
            continue;
        const double dt = YearFraction( Dates[k - 1], Dates[k] );
        const double carry = cum_carry( Dates[k] ) - cum_carry( Dates[k - 1] );
        for ( int p = 0; p < LSV_CALIB_PATHS; p++ )
        {
            const double lp = lev.LayerAt( k, spot[p] );
            const double vp = var[p];
            const double vc = QeVarianceStep( vp, dt, rng.Gaussian(), h );
            const double step = LeveragedLogSpotStep( vp, vc, dt, carry, lp, rng.Gaussian(), h );
            var[p] = vc;
            spot[p] *= std::exp( step );
            ln_spot[p] += step;
        }
    }
    return lev;
}

// ============================================================================
// engine: grid, market, book resolution, sweep, Greeks
// ============================================================================
Market BuildMarket( const thoth_mcl_request& req )
{
    Market mk;
    for ( int a = 0; a < req.n_assets; a++ )
    {
        const thoth_asset& in = req.assets[a];
        Equity e{};
        e.spot = in.spot;
        e.vol_kind = in.vol_kind;
        e.bs_vol = in.volatility / 100;                          // bs_volatility.cpp:24
        e.v0 = std::pow( in.init_vol / 100.0, 2 );               // heston_volatility.cpp:38
        e.theta = std::pow( in.long_vol / 100.0, 2 );            // heston_volatility.cpp:39
        e.kappa = in.kappa;
        e.xi = in.vol_of_vol;
        e.rho = in.rho_spot_var;
        e.has_div = in.dividend_yield != 0;
        e.div = in.dividend_yield * 0.01;                        // curve.cpp:35-39
        e.vol_shift = req.vol_shift;                             // scenario shift (Volatility::_vol_shift)
        for ( int i = 0; i < in.surface.n; i++ )
        {
            e.surface.maturities.push_back( in.surface.maturities[i] );
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
    for ( int i = 0; i < req.discount.n; i++ )  // currency.cpp:19-27
    {
        mk.discount.days.push_back( req.discount.days[i] );
        mk.discount.values.push_back( req.discount.values[i] * 0.01 );
    }
    mk.multi_curve = req.discount.n > 0;
    const int n = req.n_assets;
    if ( req.n_corr_pillars > 1 )
    {
        // one matrix per pillar (correlation.cpp:35-130); today's matrix is the
        // instantaneous one at t = 0 (ExtractMatrix over the (0, 0) window)
        for ( int p = 0; p < req.n_corr_pillars; p++ )
        {
            mk.corr_pillars.push_back( req.corr_maturities[p] );
            mk.corr_mats.emplace_back( req.corr_matrices + (size_t)p * n * n, req.corr_matrices + (size_t)( p + 1 ) * n * n );
        }
        mk.correlation.assign( (size_t)n * n, 0.0 );
        for ( int i = 0; i < n; i++ )
            for ( int j = 0; j < n; j++ )
                mk.correlation[i * n + j] = i == j ? 1.0 : mk.AvgCorrelation( i, j, 0.0, 0.0 );
    }
    else if ( req.n_corr_pillars == 1 )
        mk.correlation.assign( req.corr_matrices, req.corr_matrices + n * n );
    else if ( req.correlation )
        mk.correlation.assign( req.correlation, req.correlation + req.n_assets * req.n_assets );
    mk.lev_spot_points = req.leverage_spot_points;
    mk.lev_time_layers = req.leverage_time_layers;
    return mk;
}

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
Grid BuildGrid( const thoth_mcl_request& req, const Market& mk )
{
    Grid g;
    std::set<double> fixing;
    for ( int c = 0; c < req.n_contracts; c++ )
    {
        const thoth_contract& k = req.contracts[c];
        fixing.insert( k.maturity_days );
        if ( k.kind == THOTH_BARRIER && k.discrete_monitoring )
        {
            std::set<double> m = MonitoringDates( k );
            fixing.insert( m.begin(), m.end() );
        }
    }
    const double last_date = *fixing.rbegin();
    if ( req.uniform_steps > 0 )
    {
        g.uniform = true;
        const double T = YearFraction( 0, last_date );
        for ( int i = 0; i <= req.uniform_steps; i++ )
        {
            double ti = ( i == req.uniform_steps ) ? T : T * i / req.uniform_steps;
            g.t.push_back( ti );
            g.days.push_back( i == req.uniform_steps ? last_date : ti * NB_OF_DAYS_A_YEAR );
            g.dt.push_back( i == 0 ? 0.0 : ti - g.t[i - 1] );
        }
        return g;
    }
    long step_days = req.mcl.max_day_step;
    if ( req.mcl.vol_year_step > 0 )
    {
        bool stoch = false;
        for ( const Equity& e : mk.eq )
            stoch |= e.vol_kind != THOTH_VOL_BS;
        if ( stoch )
        {
            const long vol_days = std::max<long>( 1, (long)std::floor( req.mcl.vol_year_step * NB_OF_DAYS_A_YEAR ) );
            step_days = std::min( step_days, vol_days );
        }
    }
    std::set<double> dates;
    for ( double d = 0; d < last_date; d += step_days )
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

ResolvedBook ResolveBook( const thoth_mcl_request& req, const Market& mk, const Grid& g )
{
    ResolvedBook b;
    b.grid = g;
    const int n = req.n_assets;
    const int nd = (int)g.t.size();
    b.n_assets = n;
    int n_stoch = 0;
    for ( int a = 0; a < n; a++ )
    {
        const Equity& e = mk.eq[a];
        b.model.push_back( e.vol_kind );
        b.spot.push_back( e.spot - 0.0 );
        b.vol.push_back( ( e.bs_vol + e.vol_shift ) * 1.0 );      // bs_volatility.cpp:46-47
        StochParams h{ 0, 0, 0, 0, 0 };
        if ( e.vol_kind != THOTH_VOL_BS )
        {
            h = mk.Stoch( a );
            n_stoch++;
        }
        b.v0.push_back( h.v0 );
        b.kappa.push_back( h.kappa );
        b.theta.push_back( h.theta );
        b.xi.push_back( h.xi );
        b.rho.push_back( h.rho );
        std::vector<double> dr( nd );
        for ( int i = 0; i < nd; i++ )
        {
            // drift_node.cpp:20-27 over yield_curve_node.cpp:18-38 (r - 0 - repo - div)
            double r = mk.rate.Value( g.days[i] );
            double q = e.has_div ? e.div : 0.0;
            dr[i] = ( ( r - 0 ) - 0 ) - q;
        }
        b.drift.push_back( dr );
        if ( e.vol_kind == THOTH_VOL_LSV )
            b.lev.push_back( CalibrateLeverage( mk, a, g.days ) );
        else
            b.lev.push_back( LeverageLayers{} );
    }
    b.factors = n + n_stoch;
    b.correlate_stochastic = req.correlate_stochastic != 0;
    if ( n > 1 )
    {
        b.chol = mk.correlation;
        for ( int i = 0; i < n; i++ )
            for ( int j = i + 1; j < n; j++ )
                b.chol[i * n + j] = 0;  // only the lower triangle is read
        std::vector<double> full = mk.correlation;
        if ( !CholeskyDecomposeLower( full, n ) )
            throw std::runtime_error( "correlation is not SDP" );
        for ( int i = 0; i < n; i++ )
            for ( int j = 0; j <= i; j++ )
                b.chol[i * n + j] = full[i * n + j];
        if ( mk.TermStructured() )
        {
            // Correlation::EnsureCholeskySchedule (correlation.cpp:860-903): the
            // factor of date i is that of the matrix averaged over [t_{i-1}, t_i]
            // (date 0: the instantaneous matrix at t_0)
            for ( int d = 0; d < nd; d++ )
            {
                const double t1 = d == 0 ? g.t[0] : g.t[d - 1], t2 = g.t[d];
                std::vector<double> step( (size_t)n * n, 0.0 );
                for ( int i = 0; i < n; i++ )
                    for ( int j = 0; j < n; j++ )
                        step[i * n + j] = i == j ? 1.0 : mk.AvgCorrelation( i, j, t1, t2 );
                if ( !CholeskyDecomposeLower( step, n ) )
                    throw std::runtime_error( "step-average correlation matrix at index " + std::to_string( d ) +
                                              " is not positive-definite" );
                for ( int i = 0; i < n; i++ )
                    for ( int j = i + 1; j < n; j++ )
                        step[i * n + j] = 0;
                b.chol_dates.push_back( std::move( step ) );
            }
        }
    }
    // contract_node.cpp:33-47 over Currency::GetDiscountRateNode (currency.cpp:129-142):
    // the OIS curve when the currency has one, else the rate curve
    for ( int i = 0; i < nd; i++ )
        b.disc_rate.push_back( mk.multi_curve ? mk.discount.Value( g.days[i] ) : mk.rate.Value( g.days[i] ) );
    for ( int r = 0; r < req.n_rainbows; r++ )
    {
        ResolvedBook::Rainbow rb;
        rb.best = req.rainbows[r].best != 0;
        for ( int j = 0; j < req.rainbows[r].n; j++ )
        {
            int m = req.rainbows[r].members[j];
            rb.members.push_back( m );
            // rainbow.cpp:56-57: weight 100 / reference spot captured at load (base
            // spot, or the caller's rebasing spot)
            rb.weight.push_back( 100.0 / ( req.rebase_spots ? req.rebase_spots[m] : req.assets[m].spot ) );
        }
        b.rainbows.push_back( rb );
    }
    auto index_of = [&]( double day )
    {
        for ( int i = 0; i < nd; i++ )
            if ( g.days[i] == day )
                return i;
        throw std::runtime_error( "date not on the diffusion grid" );
    };
    for ( int c = 0; c < req.n_contracts; c++ )
    {
        const thoth_contract& k = req.contracts[c];
        ResolvedBook::Contract rc{};
        rc.kind = k.kind;
        rc.is_call = k.is_call;
        rc.underlying = k.underlying;
        rc.flow_index = index_of( k.maturity_days );
        rc.american = k.american != 0;
        if ( rc.american && ( k.kind != THOTH_VANILLA || k.underlying < 0 ) )
            throw std::runtime_error( "american exercise: vanilla contracts on an equity" );
        // vanilla.cpp:29-34 / barrier.cpp:36-41: relative strike on the BASE spot
        double base_spot = k.underlying >= 0 ? req.assets[k.underlying].spot : 100.0;
        rc.strike = k.is_absolute_strike ? k.strike : k.strike / 100 * base_spot;
        if ( k.kind == THOTH_BARRIER )
        {
            rc.is_up = k.barrier_type == THOTH_UP_OUT || k.barrier_type == THOTH_UP_IN;
            rc.is_in = k.barrier_type == THOTH_UP_IN || k.barrier_type == THOTH_DOWN_IN;
            double H = rc.is_up ? k.barrier_up_level : k.barrier_down_level;
            if ( k.discrete_monitoring && !g.uniform )
            {
                for ( double d : MonitoringDates( k ) )
                    rc.monitor.push_back( index_of( d ) );
            }
            else if ( k.discrete_monitoring && g.uniform )
            {
                for ( int i = 1; i <= rc.flow_index; i++ )
                    rc.monitor.push_back( i );  // uniform grid: every step is a monitoring date
            }
            else
            {
                // barrier.cpp:146-161 (BGK shift on every diffusion date up to maturity)
                for ( int i = 0; i < nd; i++ )
                    if ( g.days[i] <= k.maturity_days )
                        rc.monitor.push_back( i );
                double t = YearFraction( 0, k.maturity_days );
                size_t steps = ( rc.monitor.size() > 1 ) ? rc.monitor.size() - 1 : 1;
                double dt = t / (double)steps;
                if ( k.underlying < 0 )
                    throw std::runtime_error( "continuous barrier on a rainbow has no implied vol (rainbow.cpp)" );
                double vol = mk.ImplicitVol( k.underlying, rc.strike, k.maturity_days );
                const double beta = 0.5826;
                H *= std::exp( ( rc.is_up ? -1.0 : 1.0 ) * beta * vol * std::sqrt( dt ) );
            }
            rc.level = H;
        }
        b.contracts.push_back( rc );
    }
    return b;
}

double BookSums::Mean( int i ) const
{
    return i < 0 ? book_sum / count : sum[i] / count;
}
double BookSums::Trust( int i ) const
{
    const double n = (double)count;
    const double s = i < 0 ? book_sum : sum[i];
    const double s2 = i < 0 ? book_sum2 : sum2[i];
    return std::sqrt( ( s2 - s * s / n ) / ( n - 1 ) / n );
}

namespace
{
inline void Kahan( double x, double& sum, double& c, double& sum2, double& c2 )
{
    // monte_carlo_node.cpp:73-96
    {
        const double y = x - c;
        const double t = sum + y;
        c = ( t - sum ) - y;
        sum = t;
    }
    {
        const double y = x * x - c2;
        const double t = sum2 + y;
        c2 = ( t - sum2 ) - y;
        sum2 = t;
    }
}
}  // namespace

BookSums SweepBook( const ResolvedBook& b, int64_t paths, uint64_t sobol_skip, int seed, bool use_sobol )
{
    const int n = b.n_assets;
    const int nd = (int)b.grid.t.size();
    const int m = nd - 1;
    const std::vector<double>& t = b.grid.t;
    const std::vector<double>& dt = b.grid.dt;
    std::vector<double> sqrt_dt( nd, 0.0 );
    for ( int i = 1; i < nd; i++ )
        sqrt_dt[i] = std::sqrt( dt[i] );  // monte_carlo_node.cpp:121
    Rng rng( (uint64_t)seed );
    PathGenerator pg( t, b.factors, &rng, sobol_skip, use_sobol );
    // factor of each equity's variance noise: stochastic names after every spot noise
    std::vector<int> var_factor( n, -1 );
    {
        int f = n;
        for ( int a = 0; a < n; a++ )
            if ( b.model[a] != THOTH_MODEL_BS )
                var_factor[a] = f++;
    }
    std::vector<std::vector<double>> S( n, std::vector<double>( nd, 0.0 ) );
    std::vector<double> Wp( n ), V( n ), corr( n );
    const int nc = (int)b.contracts.size();
    BookSums out;
    out.sum.assign( nc, 0 );
    out.sum2.assign( nc, 0 );
    out.c.assign( nc, 0 );
    out.c2.assign( nc, 0 );
    std::vector<double> df( nc, 1.0 );
    for ( int c = 0; c < nc; c++ )
    {
        const int fi = b.contracts[c].flow_index;
        if ( fi != 0 )
            df[c] = std::exp( t[0] * b.disc_rate[0] - t[fi] * b.disc_rate[fi] );  // contract_node.cpp:40-44
    }
    auto underlying_value = [&]( int u, int i )
    {
        if ( u >= 0 )
            return S[u][i];
        const ResolvedBook::Rainbow& rb = b.rainbows[-1 - u];  // rainbow_node.cpp:14-27
        double x = rb.best ? -std::numeric_limits<double>::infinity() : std::numeric_limits<double>::infinity();
        for ( size_t j = 0; j < rb.members.size(); j++ )
        {
            double perf = S[rb.members[j]][i] * rb.weight[j];
            x = rb.best ? std::max( x, perf ) : std::min( x, perf );
        }
        return x;
    };
    std::vector<double> cv( nc );
    // path recording of the American contracts' equities (pricer_mcl_american.cpp:59-94)
    out.rec.assign( n, {} );
    for ( const ResolvedBook::Contract& k : b.contracts )
        if ( k.american && out.rec[k.underlying].empty() )
            out.rec[k.underlying].assign( (size_t)paths * nd, 0.0 );
    for ( int64_t p = 0; p < paths; p++ )
    {
        pg.NextPath();
        for ( int a = 0; a < n; a++ )
        {
            S[a][0] = b.spot[a];
            Wp[a] = 0;
            V[a] = b.v0[a];
        }
        for ( int i = 1; i <= m; i++ )
        {
            // correlated_noise_node.cpp:15-27 (lower triangle, name order)
            for ( int a = 0; a < n; a++ )
            {
                const bool stoch = b.model[a] != THOTH_MODEL_BS;
                if ( n > 1 && ( !stoch || b.correlate_stochastic ) )
                {
                    double x = 0;
                    if ( !b.chol_dates.empty() )
                    {
                        // term structure: every lower-triangle coefficient wired,
                        // read at the date (correlation.cpp:906-933)
                        for ( int v = 0; v <= a; v++ )
                            x += pg.Buffer( v )[i] * b.chol_dates[i][a * n + v];
                    }
                    else
                        for ( int v = 0; v <= a; v++ )
                            if ( b.chol[a * n + v] != 0 )  // CholeskyMayBeNonZero, correlation.cpp:923-933
                                x += pg.Buffer( v )[i] * b.chol[a * n + v];
                    corr[a] = x;
                }
                else
                {
                    corr[a] = pg.Buffer( a )[i];
                }
            }
            for ( int a = 0; a < n; a++ )
            {
                const double carry_drift = b.drift[a][i] * t[i] - b.drift[a][i - 1] * t[i - 1];
                if ( b.model[a] == THOTH_MODEL_BS )
                {
                    // brownian_node.cpp:12-24 + spot_diffusion_node.cpp:21-65
                    const double w1 = Wp[a] + sqrt_dt[i] * corr[a];
                    const double dw = w1 - Wp[a];
                    Wp[a] = w1;
                    const double v = b.vol[a];
                    const double expo = carry_drift - v * v / 2 * dt[i] + v * dw;
                    S[a][i] = ( S[a][i - 1] - 0 ) * std::exp( expo ) + 0;
                    continue;
                }
                StochParams h{ b.v0[a], b.kappa[a], b.theta[a], b.xi[a], b.rho[a] };
                const double vp = V[a];
                const double vc = QeVarianceStep( vp, dt[i], pg.Buffer( var_factor[a] )[i], h );
                V[a] = vc;
                double L = 1.0;
                if ( b.model[a] == THOTH_MODEL_LSV )
                {
                    // local_volatility_node.cpp:19-47: layer i at the previous spot
                    // (a surface: LeverageSurface::GetLeverage at the date's time)
                    const LeverageLayers& lv = b.lev[a];
                    if ( lv.surface )
                        L = lv.GetLeverage( S[a][i - 1], b.grid.t[i] );
                    else
                    {
                        const std::vector<double>& vv = lv.levels[i];
                        const double f = std::log( S[a][i - 1] ) / lv.ln_step[i] - (double)lv.offset[i];
                        const size_t last = vv.size() - 1;
                        if ( f <= 0.0 )
                            L = vv[0];
                        else if ( f >= (double)last )
                            L = vv[last];
                        else
                        {
                            const size_t k = (size_t)f;
                            const double w = f - (double)k;
                            L = ( 1.0 - w ) * vv[k] + w * vv[k + 1];
                        }
                    }
                }
                const double z = corr[a];
                double step;
                // heston_spot_node.cpp:32-94
                if ( h.xi > 1e-12 )
                {
                    const double g = 0.5;
                    const double K0 = -L * h.rho * h.kappa * h.theta / h.xi * dt[i];
                    const double aa = g * dt[i] * ( L * h.kappa * h.rho / h.xi - 0.5 * L * L );
                    const double K1 = aa - L * h.rho / h.xi;
                    const double K2 = aa + L * h.rho / h.xi;
                    const double K3 = g * dt[i] * ( 1.0 - h.rho * h.rho ) * L * L;
                    const double K4 = K3;
                    double var = K3 * vp + K4 * vc;
                    if ( var < 0 )
                        var = 0;
                    step = carry_drift + K0 + K1 * vp + K2 * vc + std::sqrt( var ) * z;
                }
                else
                {
                    step = carry_drift - 0.5 * L * L * vp * dt[i] + L * std::sqrt( vp * dt[i] ) * z;
                }
                S[a][i] = S[a][i - 1] * std::exp( step );
            }
        }
        // payoffs: vanilla_flow_node.cpp:18-31, barrier_flow_node.cpp:23-56,
        // contract_node.cpp:16-51, book_node.cpp:16-31
        double book = 0;
        for ( int c = 0; c < nc; c++ )
        {
            const ResolvedBook::Contract& k = b.contracts[c];
            const double s = underlying_value( k.underlying, k.flow_index );
            double flow;
            double vanilla = k.is_call ? ( s - k.strike ) : ( k.strike - s );
            vanilla = std::max( vanilla, 0.0 );
            if ( k.kind == THOTH_BARRIER )
            {
                bool hit = false;
                for ( int idx : k.monitor )
                {
                    double sm = underlying_value( k.underlying, idx );
                    if ( k.is_up ? ( sm >= k.level ) : ( sm <= k.level ) )
                    {
                        hit = true;
                        break;
                    }
                }
                bool alive = k.is_in ? hit : !hit;
                flow = alive ? vanilla : 0;
            }
            else
            {
                flow = vanilla;
            }
            double x = 0;
            x += flow * df[c];
            cv[c] = x;
            book += x * 1.0;
        }
        for ( int c = 0; c < nc; c++ )
            Kahan( cv[c], out.sum[c], out.c[c], out.sum2[c], out.c2[c] );
        Kahan( book, out.book_sum, out.book_c, out.book_sum2, out.book_c2 );
        out.count++;
        for ( int a = 0; a < n; a++ )
            if ( !out.rec[a].empty() )
                std::copy( S[a].begin(), S[a].end(), out.rec[a].begin() + (size_t)p * nd );
    }
    return out;
}

// ---- American Longstaff-Schwartz ---------------------------------------------
std::vector<double> LeastSquares( const std::vector<double>& X, size_t n, size_t p, const std::vector<double>& y )
{
    // normal equations A = X^T X, g = X^T y (maths.cpp:419-441)
    std::vector<double> A( p * p, 0.0 ), g( p, 0.0 );
    for ( size_t a = 0; a < p; a++ )
    {
        for ( size_t b = 0; b <= a; b++ )
        {
            double s = 0.0;
            for ( size_t k = 0; k < n; k++ )
                s += X[k * p + a] * X[k * p + b];
            A[a * p + b] = s;
            A[b * p + a] = s;
        }
        double sg = 0.0;
        for ( size_t k = 0; k < n; k++ )
            sg += X[k * p + a] * y[k];
        g[a] = sg;
    }
    // A = L L^T (maths.cpp:354-381), then L z = g and L^T beta = z (maths.cpp:447-469)
    if ( !CholeskyDecomposeLower( A, (int)p ) )
        return {};
    std::vector<double> z( p ), beta( p );
    for ( size_t i = 0; i < p; i++ )
    {
        double s = g[i];
        for ( size_t k = 0; k < i; k++ )
            s -= A[i * p + k] * z[k];
        z[i] = s / A[i * p + i];
    }
    for ( size_t i = p; i-- > 0; )
    {
        double s = z[i];
        for ( size_t k = i + 1; k < p; k++ )
            s -= A[k * p + i] * beta[k];
        beta[i] = s / A[i * p + i];
    }
    return beta;
}

namespace
{
double Intrinsic( double s, double strike, bool is_call )  // vanilla.cpp:79-82, finance.cpp:9-32
{
    return std::max( is_call ? s - strike : strike - s, 0.0 );
}
}  // namespace

AmericanPolicy FitAmericanPolicy( double strike, bool is_call, const std::vector<double>& paths, size_t N, size_t M,
                                  const std::vector<double>& tau, const std::vector<double>& rates )
{
    AmericanPolicy pol;
    if ( M < 2 )
        return pol;
    const size_t kMinItm = 50;  // pricer_mcl_american.cpp:274
    pol.basis_norm = strike > 0 ? strike : paths[0];  // vanilla.hpp:38-41
    pol.b0.assign( M, 0.0 );
    pol.b1.assign( M, 0.0 );
    pol.b2.assign( M, 0.0 );
    pol.has_fit.assign( M, 0 );
    std::vector<double> cf( N );
    for ( size_t p = 0; p < N; p++ )
        cf[p] = Intrinsic( paths[p * M + M - 1], strike, is_call );
    std::vector<size_t> itm;
    std::vector<double> X, y;
    for ( int t = (int)M - 2; t >= 1; t-- )
    {
        const double df = std::exp( rates[t] * tau[t] - rates[t + 1] * tau[t + 1] );
        for ( size_t p = 0; p < N; p++ )
            cf[p] *= df;
        itm.clear();
        for ( size_t p = 0; p < N; p++ )
            if ( Intrinsic( paths[p * M + t], strike, is_call ) > 0 )
                itm.push_back( p );
        if ( itm.size() < kMinItm )
            continue;
        const size_t ni = itm.size();
        X.assign( ni * 3, 0.0 );
        y.assign( ni, 0.0 );
        for ( size_t k = 0; k < ni; k++ )
        {
            const double m = paths[itm[k] * M + t] / pol.basis_norm;
            X[k * 3] = 1.0;
            X[k * 3 + 1] = m;
            X[k * 3 + 2] = m * m;
            y[k] = cf[itm[k]];
        }
        const std::vector<double> beta = LeastSquares( X, ni, 3, y );
        if ( beta.empty() )
            continue;
        pol.b0[t] = beta[0];
        pol.b1[t] = beta[1];
        pol.b2[t] = beta[2];
        pol.has_fit[t] = 1;
        for ( size_t p : itm )
        {
            const double s = paths[p * M + t];
            const double m = s / pol.basis_norm;
            const double continuation = pol.b0[t] + pol.b1[t] * m + pol.b2[t] * m * m;
            const double intrinsic = Intrinsic( s, strike, is_call );
            if ( intrinsic >= continuation )
                cf[p] = intrinsic;
        }
    }
    return pol;
}

double ApplyAmericanPolicy( double strike, bool is_call, const std::vector<double>& paths, size_t N, size_t M,
                            const std::vector<double>& tau, const std::vector<double>& rates,
                            const AmericanPolicy& pol, double& trust )
{
    trust = 0;
    if ( M < 2 )
        return 0;
    double sum = 0, sum2 = 0;
    for ( size_t p = 0; p < N; p++ )
    {
        double value = 0;
        for ( size_t t = 1; t < M; t++ )
        {
            const double s = paths[p * M + t];
            const double intrinsic = Intrinsic( s, strike, is_call );
            if ( t == M - 1 )
            {
                value = intrinsic * std::exp( -rates[t] * ( tau[t] - tau[0] ) );
                break;
            }
            if ( intrinsic > 0 && pol.has_fit[t] )
            {
                const double m = s / pol.basis_norm;
                const double continuation = pol.b0[t] + pol.b1[t] * m + pol.b2[t] * m * m;
                if ( intrinsic >= continuation )
                {
                    value = intrinsic * std::exp( -rates[t] * ( tau[t] - tau[0] ) );
                    break;
                }
            }
        }
        sum += value;
        sum2 += value * value;
    }
    const double mean = sum / (double)N;
    trust = std::sqrt( ( sum2 / (double)N - mean * mean ) / (double)( N - 1 ) );
    return std::max( mean, Intrinsic( paths[0], strike, is_call ) );
}

// ---- pricing + bump-and-revalue Greeks (pricer.cpp:273-391, pricer_mcl_greeks.cpp:114-218)
// With American contracts (pricer_mcl_american.cpp:137-248): the exercise policy
// of each is fitted once on the base scenario's recorded paths and applied frozen
// to every scenario's; the contract's value replaces its European value in the
// scenario's book premium, and the base book premium is re-aggregated from the
// contracts (the book trust stays the sweep's).
int Price( const thoth_mcl_request& req, thoth_mcl_result* res )
{
    const Market base = BuildMarket( req );
    const Grid g = BuildGrid( req, base );
    const int nc = req.n_contracts;
    struct Scen
    {
        BookSums sums;
        double book;                  // book premium for the finite differences
        std::vector<double> contract;  // contract premiums (American: the policy's value)
        std::vector<double> trust;
    };
    std::vector<int> am;
    std::vector<AmericanPolicy> pol;
    const ResolvedBook rb0 = ResolveBook( req, base, g );
    for ( int c = 0; c < nc; c++ )
        if ( rb0.contracts[c].american )
            am.push_back( c );
    // one contract's exercise data of a scenario: paths [N][M], tau, zero rates
    auto exercise_data = [&]( const ResolvedBook& rb, const BookSums& s, int c, std::vector<double>& P,
                              std::vector<double>& tau, std::vector<double>& rates, size_t& N, size_t& M )
    {
        const ResolvedBook::Contract& k = rb.contracts[c];
        const size_t nd = g.t.size();
        M = (size_t)k.flow_index + 1;
        N = (size_t)s.count;
        P.assign( N * M, 0.0 );
        const std::vector<double>& rec = s.rec[k.underlying];
        for ( size_t p = 0; p < N; p++ )
            std::copy( rec.begin() + p * nd, rec.begin() + p * nd + M, P.begin() + p * M );
        tau.assign( g.t.begin(), g.t.begin() + M );
        rates.assign( rb.disc_rate.begin(), rb.disc_rate.begin() + M );
    };
    auto run = [&]( const Market& mk, bool fit )
    {
        const ResolvedBook rb = ResolveBook( req, mk, g );
        Scen sc;
        sc.sums = SweepBook( rb, req.mcl.paths, req.mcl.sobol_skip, req.mcl.seed, req.mcl.use_sobol != 0 );
        sc.book = sc.sums.Mean( -1 );
        sc.contract.resize( nc );
        sc.trust.resize( nc );
        for ( int c = 0; c < nc; c++ )
        {
            sc.contract[c] = sc.sums.Mean( c );
            sc.trust[c] = sc.sums.Trust( c );
        }
        std::vector<double> P, tau, rates;
        size_t N = 0, M = 0;
        for ( size_t j = 0; j < am.size(); j++ )
        {
            const int c = am[j];
            const ResolvedBook::Contract& k = rb.contracts[c];
            exercise_data( rb, sc.sums, c, P, tau, rates, N, M );
            if ( fit )
                pol.push_back( FitAmericanPolicy( k.strike, k.is_call != 0, P, N, M, tau, rates ) );
            double tr = 0;
            const double v = ApplyAmericanPolicy( k.strike, k.is_call != 0, P, N, M, tau, rates, pol[j], tr );
            sc.book += v - sc.contract[c];
            sc.contract[c] = v;
            sc.trust[c] = tr;
        }
        return sc;
    };
    const Scen b0 = run( base, true );
    double p0 = b0.book;
    if ( !am.empty() )
    {
        p0 = 0;
        for ( int c = 0; c < nc; c++ )
            p0 += b0.contract[c];
    }
    res->premium = p0;
    res->premium_trust = b0.sums.Trust( -1 );
    res->n_dates = (int)g.t.size();
    res->n_scenarios = 1;
    for ( int c = 0; c < nc; c++ )
    {
        if ( res->contract_premium )
            res->contract_premium[c] = b0.contract[c];
        if ( res->contract_trust )
            res->contract_trust[c] = b0.trust[c];
    }
    res->delta = res->gamma = res->vega = res->rho = 0;
    std::vector<double> cd( nc, 0 ), cg( nc, 0 ), cvv( nc, 0 ), cr( nc, 0 );
    for ( int a = 0; a < req.n_assets; a++ )
    {
        const double spot = base.eq[a].spot;
        if ( req.request_greeks & THOTH_GREEK_DELTA )
        {
            const double h = GREEK_SPOT_BUMP * spot;
            Market mk = base;
            mk.eq[a].spot = spot + h;
            const Scen s = run( mk, false );
            res->n_scenarios++;
            res->delta += ( s.book - p0 ) / h;
            for ( int c = 0; c < nc; c++ )
                cd[c] += ( s.contract[c] - b0.contract[c] ) / h;
        }
        if ( req.request_greeks & THOTH_GREEK_GAMMA )
        {
            const double h = GREEK_GAMMA_BUMP * spot;
            Market mu = base, md = base;
            mu.eq[a].spot = spot + h;
            md.eq[a].spot = spot - h;
            const Scen su = run( mu, false ), sd = run( md, false );
            res->n_scenarios += 2;
            res->gamma += ( su.book - 2 * p0 + sd.book ) / ( h * h );
            for ( int c = 0; c < nc; c++ )
                cg[c] += ( su.contract[c] - 2 * b0.contract[c] + sd.contract[c] ) / ( h * h );
        }
    }
    if ( req.request_greeks & THOTH_GREEK_VEGA )
    {
        Market mk = base;
        for ( Equity& e : mk.eq )
            e.vol_shift += GREEK_VOL_BUMP;
        const Scen s = run( mk, false );
        res->n_scenarios++;
        res->vega = ( s.book - p0 ) / GREEK_VOL_BUMP * 0.01;
        for ( int c = 0; c < nc; c++ )
            cvv[c] = ( s.contract[c] - b0.contract[c] ) / GREEK_VOL_BUMP * 0.01;
    }
    if ( req.request_greeks & THOTH_GREEK_RHO )
    {
        Market mk = base;
        mk.rate.shift = GREEK_RATE_BUMP;  // projection AND discounting curves (pricer.cpp:344-359)
        mk.discount.shift = GREEK_RATE_BUMP;
        const Scen s = run( mk, false );
        res->n_scenarios++;
        res->rho = ( s.book - p0 ) / GREEK_RATE_BUMP * 0.01;
        for ( int c = 0; c < nc; c++ )
            cr[c] = ( s.contract[c] - b0.contract[c] ) / GREEK_RATE_BUMP * 0.01;
    }
    for ( int c = 0; c < nc; c++ )
    {
        if ( res->contract_delta )
            res->contract_delta[c] = cd[c];
        if ( res->contract_gamma )
            res->contract_gamma[c] = cg[c];
        if ( res->contract_vega )
            res->contract_vega[c] = cvv[c];
        if ( res->contract_rho )
            res->contract_rho[c] = cr[c];
    }
    return THOTH_OK;
}

}  // namespace oracle
