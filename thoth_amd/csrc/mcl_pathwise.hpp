// mcl_pathwise.hpp -- pathwise (forward-tangent) delta and vega of a
// single-equity Black-Scholes or Heston book holding one vanilla or one
// knock-out barrier contract (no reference counterpart: the reference has
// bump Greeks only, pricer_mcl_greeks.cpp:43-299).
//
// The barrier indicator is not differentiable, so knock-out books are
// simulated with one-step survival conditioning (Glasserman & Staum 2001):
// at every monitored step the spot normal is drawn from its distribution
// conditioned on survival, Z' = Phi^-1(q + (1 - q) U) with q the one-step
// knock-out probability and U = Phi(Z) the path's own normal, and the path
// carries the weight W = prod (1 - q).  Y = df W payoff is an unbiased
// estimator of the same discretely monitored price and is a smooth function
// of the spot and of the variance parameters, so its pathwise derivative is
// propagated step by step:
//   x = log S,  x' = x + mu + sigma Z,  zb = s (h - x - mu) / sigma  (s = +1 down, -1 up),
//   dx' = (dx + dmu)(1 - r) + dsigma (Z - s r zb),  r = (1 - U) phi(zb) / phi(Z'),
//   d log(1 - q) = hz (s (dx + dmu) + zb dsigma) / sigma,  hz = phi(zb) / (1 - q),
// and, without monitoring, dx' = dx + dmu + dsigma Z.  Delta seeds dx0 =
// 1/S0; vega seeds the reference's Heston bump (sqrt(v0) + e)^2, (sqrt(theta)
// + e)^2 (heston_volatility.hpp:45-47) through the QE variance map, or vol + e
// for Black-Scholes.  The QE branch switch and the exponential branch's mass
// point are not differentiated (the pathwise vega ignores those jumps).
#pragma once
#include <hip/hip_runtime.h>

#include "mcl_launch.hpp"
#include "mcl_sweep.hpp"

namespace thoth
{

namespace
{

// Phi^-1 from t and 1 - t (both accurate): FP64 Acklam (distributions.cpp:37-81
// coefficients, upper tail from 1 - t), FP32 the fitted erfinv form.
__device__ __forceinline__ double inv_norm( double t, double omt )
{
    const double a0 = -3.969683028665376e+01, a1 = 2.209460984245205e+02, a2 = -2.759285104469687e+02,
                 a3 = 1.383577518672690e+02, a4 = -3.066479806614716e+01, a5 = 2.506628277459239e+00;
    const double b0 = -5.447609879822406e+01, b1 = 1.615858368580409e+02, b2 = -1.556989798598866e+02,
                 b3 = 6.680131188771972e+01, b4 = -1.328068155288572e+01;
    const double c0 = -7.784894002430293e-03, c1 = -3.223964580411365e-01, c2 = -2.400758277161838e+00,
                 c3 = -2.549732539343734e+00, c4 = 4.374664141464968e+00, c5 = 2.938163982698783e+00;
    const double d0 = 7.784695709041462e-03, d1 = 3.224671290700398e-01, d2 = 2.445134137142996e+00,
                 d3 = 3.754408661907416e+00;
    t = fmax( t, 1e-300 );
    omt = fmax( omt, 1e-300 );
    if ( t < 0.02425 )
    {
        const double q = sqrt( -2.0 * log( t ) );
        return ( ( ( ( ( c0 * q + c1 ) * q + c2 ) * q + c3 ) * q + c4 ) * q + c5 ) /
               ( ( ( ( d0 * q + d1 ) * q + d2 ) * q + d3 ) * q + 1.0 );
    }
    if ( omt < 0.02425 )
    {
        const double q = sqrt( -2.0 * log( omt ) );
        return -( ( ( ( ( c0 * q + c1 ) * q + c2 ) * q + c3 ) * q + c4 ) * q + c5 ) /
               ( ( ( ( d0 * q + d1 ) * q + d2 ) * q + d3 ) * q + 1.0 );
    }
    const double q = t - 0.5, r = q * q;
    return ( ( ( ( ( a0 * r + a1 ) * r + a2 ) * r + a3 ) * r + a4 ) * r + a5 ) * q /
           ( ( ( ( ( b0 * r + b1 ) * r + b2 ) * r + b3 ) * r + b4 ) * r + 1.0 );
}

__device__ __forceinline__ float inv_norm( float t, float omt )
{
    t = fmaxf( t, 1e-12f );
    omt = fmaxf( omt, 1e-12f );
    const float w = -fast_log( 4.0f * t * omt );
    const float xs = t - omt;
    float p;
    if ( w < 5.0f )
    {
        const float u = w - 2.5f;
        p = 3.958320987e-08f;
        p = fmaf( p, u, 4.851791016e-07f );
        p = fmaf( p, u, -4.981194608e-06f );
        p = fmaf( p, u, -6.207880344e-06f );
        p = fmaf( p, u, 3.091150937e-04f );
        p = fmaf( p, u, -1.773041744e-03f );
        p = fmaf( p, u, -5.908129905e-03f );
        p = fmaf( p, u, 3.488026846e-01f );
        p = fmaf( p, u, 2.123313560e+00f );
    }
    else
    {
        const float u = __builtin_amdgcn_sqrtf( w ) - 3.0f;
        p = 5.523418413e-06f;
        p = fmaf( p, u, -2.638466750e-04f );
        p = fmaf( p, u, 1.661917106e-03f );
        p = fmaf( p, u, -4.731942225e-03f );
        p = fmaf( p, u, 8.164933407e-03f );
        p = fmaf( p, u, -1.093819435e-02f );
        p = fmaf( p, u, 1.334743934e-02f );
        p = fmaf( p, u, 1.416595396e+00f );
        p = fmaf( p, u, 4.006434468e+00f );
    }
    return xs * p;
}

// Phi(-y) = erfc(y / sqrt 2) / 2
__device__ __forceinline__ double phic( double y ) { return 0.5 * erfc( y * 0.70710678118654752 ); }
__device__ __forceinline__ float phic( float y ) { return 0.5f * erfc_fast( y * 0.70710678118654752f ); }
template <class R>
__device__ __forceinline__ R expr( R x )
{
    if constexpr ( std::is_same_v<R, double> )
        return exp( x );
    else
        return __expf( x );
}
template <class R>
__device__ __forceinline__ R sqrtr( R x )
{
    if constexpr ( std::is_same_v<R, double> )
        return sqrt( x );
    else
        return __builtin_amdgcn_sqrtf( x );
}

// tangent of the QE variance draw v' = QE(v; theta) (same intermediates as
// qe_f64 / qe_f32): dv' from dv and dtheta.  m = M0 + v e, s2 = S0 + v S1 with
// dM0/dtheta = 1 - e, dS0/dtheta = S0 / theta.
template <class R>
__device__ __forceinline__ R qe_tangent( R v, R z, R e, R M0, R S1, R S0, R theta, R dv, R dth )
{
    const R m = M0 + v * e;
    const R s2 = S0 + v * S1;
    if ( !( m > R( 0 ) ) )
        return R( 0 );
    const R dm = e * dv + ( R( 1 ) - e ) * dth;
    const R ds2 = S1 * dv + ( theta > R( 0 ) ? S0 / theta : R( 0 ) ) * dth;
    const R psi = s2 / ( m * m );
    const R dpsi = ( ds2 - R( 2 ) * psi * m * dm ) / ( m * m );
    if ( psi <= R( 1.5 ) )
    {
        if ( !( psi > R( 0 ) ) )
            return R( 0 );
        const R c = R( 2 ) / psi;
        const R root = sqrtr( c * ( c - R( 1 ) ) );
        const R b2 = c - R( 1 ) + root;
        const R a = m / ( R( 1 ) + b2 );
        const R dc = -c * dpsi / psi;
        const R db2 = dc * ( R( 1 ) + ( R( 2 ) * c - R( 1 ) ) / ( R( 2 ) * root ) );
        const R da = ( dm - a * db2 ) / ( R( 1 ) + b2 );
        const R sb = sqrtr( b2 );
        const R y = sb + z;
        return da * y * y + a * y * db2 / sb;
    }
    const R omp = R( 2 ) / ( psi + R( 1 ) );
    const R omu = phic( z );
    if ( omu >= omp )
        return R( 0 );
    const R lg = std::is_same_v<R, double> ? (R)log( (double)omp / (double)omu ) : (R)fast_log( (float)( omp / omu ) );
    const R binv = m * ( psi + R( 1 ) ) * R( 0.5 );
    const R dlomp = -dpsi / ( psi + R( 1 ) );
    const R dbinv = ( dm * ( psi + R( 1 ) ) + m * dpsi ) * R( 0.5 );
    return dlomp * binv + lg * dbinv;
}

}  // namespace

// one workgroup = 256 consecutive Sobol indices of book blockIdx.y; partial
// slots per block: [Y, delta, vega] x [sum, sum of squares]
template <class R, int DEPTH>
__global__ void __launch_bounds__( kBlock ) pathwise_kernel( const DevBook* __restrict__ books,
                                                            const uint32_t* __restrict__ V, uint32_t bq0,
                                                            uint32_t first_idx, uint32_t last_idx )
{
    constexpr bool F64 = std::is_same_v<R, double>;
    typedef R facv __attribute__( ( ext_vector_type( 2 ) ) );
    const DevBook& B = books[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const int wib = threadIdx.x >> 6;
    const uint32_t bq = bq0 + blockIdx.x;
    const uint32_t idx = bq * 256u + (uint32_t)threadIdx.x;
    const bool valid = idx >= first_idx && idx <= last_idx;
    const int m = B.m, nd = B.n_dates;

    extern __shared__ uint32_t lds[];
    __shared__ double acc[kWavesPerBlock][3][2];
    sobol_block_table( lds, V, m * B.factors, bq, nullptr );
    __syncthreads();
    SobolLane sob;
    sob.V = V;
    sob.tab = lds;
    sob.set_lane( wib, lane );
    BridgeWalk<R, 2, DEPTH> bw;
    bw.start( B, sob );

    const DevAsset& A = B.asset[0];
    const bool heston = A.model == 1;
    const bool barrier = B.c_kind[0] == 1;
    const R sgn = B.c_up[0] ? R( -1 ) : R( 1 );  // survive above a down barrier, below an up barrier
    const R hlog = barrier ? (R)log( B.c_level[0] ) : R( 0 );
    const R K = (R)B.c_strike[0];
    const bool call = B.c_call[0] != 0;
    const R theta = (R)A.theta;
    // vega seeds: d v0 / de = 2 sqrt(v0), d theta / de = 2 sqrt(theta); BS: d vol / de = 1
    const R dth = heston ? R( 2 ) * sqrtr( theta ) : R( 0 );

    R S = (R)A.spot, Wb = R( 0 ), v = (R)A.v0;
    R wt = R( 1 );                       // survival weight
    R gx = R( 1 ) / (R)A.spot, lgd = R( 0 );  // d log S / d S0, d log W / d S0
    R dv = heston ? R( 2 ) * sqrtr( v ) : R( 0 ), gv = R( 0 ), lgv = R( 0 );  // vega tangents
    R y = R( 0 ), dl = R( 0 ), vg = R( 0 );

    facv nz = (facv)( R( 0 ) );
    for ( int p = 0; p < m; p++ )
    {
        const int i = p + 1;
        bw.step( B, sob, p, nz );
        cptr<float> sc = as_const( B.sc ) + (size_t)i * SC_COUNT;
        R eps = nz[0];
        R cd;
        if constexpr ( F64 )
            cd = as_const( B.drift )[i] * as_const( B.t )[i] - as_const( B.drift )[i - 1] * as_const( B.t )[i - 1];
        else
            cd = sc[SC_CARRY];
        // ---- the step's mean and scale of log S, and their vega tangents
        R vc = v, mu, sig, dmu, dsig, dvc = R( 0 );
        R K1 = 0, K2 = 0, K3 = 0;
        if ( !heston )
        {
            if constexpr ( F64 )
            {
                mu = cd - A.vol * A.vol / 2 * as_const( B.dt )[i];
                sig = A.vol * as_const( B.sqrt_dt )[i];
                dmu = -A.vol * as_const( B.dt )[i];
                dsig = as_const( B.sqrt_dt )[i];
            }
            else
            {
                mu = cd - sc[SC_HALF_V2DT];
                sig = sc[SC_VOL] * sc[SC_SQDT];
                dmu = -sc[SC_VOL] * sc[SC_DT];
                dsig = sc[SC_SQDT];
            }
        }
        else
        {
            const R zv = nz[1];
            R e, M0, S1, S0;
            if constexpr ( F64 )
            {
                vc = qe_f64( v, as_const( B.dt )[i], zv, A.kappa, A.theta, A.xi );
                e = exp( -A.kappa * as_const( B.dt )[i] );
                M0 = A.theta * ( 1 - e );
                S1 = A.xi * A.xi * e / A.kappa * ( 1 - e );
                S0 = A.theta * A.xi * A.xi / ( 2 * A.kappa ) * ( 1 - e ) * ( 1 - e );
            }
            else
            {
                vc = qe_f32( v, zv, sc );
                e = sc[SC_E];
                M0 = sc[SC_M0];
                S1 = sc[SC_S1];
                S0 = sc[SC_S0];
            }
            dvc = qe_tangent<R>( v, zv, e, M0, S1, S0, theta, dv, dth );
            R K0, dK0;
            if constexpr ( F64 )
            {
                const double dt = as_const( B.dt )[i];
                if ( A.xi > 1e-12 )
                {
                    const double aa = 0.5 * dt * ( A.kappa * A.rho / A.xi - 0.5 );
                    K0 = -A.rho * A.kappa * A.theta / A.xi * dt;
                    dK0 = -A.rho * A.kappa / A.xi * dt * dth;
                    K1 = aa - A.rho / A.xi;
                    K2 = aa + A.rho / A.xi;
                    K3 = 0.5 * dt * ( 1.0 - A.rho * A.rho );
                }
                else
                {
                    K0 = 0;
                    dK0 = 0;
                    K1 = -0.5 * dt;
                    K2 = 0;
                    K3 = dt;  // sigma^2 = vp dt below
                }
            }
            else
            {
                if ( A.xi > 1e-12 )
                {
                    const float aa = sc[SC_A1] - sc[SC_A2];
                    K0 = sc[SC_K0];
                    dK0 = theta > 0.0f ? sc[SC_K0] / theta * dth : 0.0f;
                    K1 = aa - sc[SC_RX];
                    K2 = aa + sc[SC_RX];
                    K3 = sc[SC_K3];
                }
                else
                {
                    K0 = 0;
                    dK0 = 0;
                    K1 = -0.5f * sc[SC_DT];
                    K2 = 0;
                    K3 = sc[SC_DT];
                }
            }
            const bool degenerate = !( A.xi > 1e-12 );
            R var;
            if constexpr ( F64 )
                var = degenerate ? v * K3 : K3 * v + K3 * vc;  // the sweep's operation order
            else
                var = degenerate ? v * K3 : K3 * ( v + vc );
            mu = cd + K0 + K1 * v + K2 * vc;
            sig = sqrtr( var > R( 0 ) ? var : R( 0 ) );
            dmu = dK0 + K1 * dv + K2 * dvc;
            dsig = sig > R( 0 ) ? ( degenerate ? K3 * dv : K3 * ( dv + dvc ) ) / ( R( 2 ) * sig ) : R( 0 );
        }
        // ---- one-step survival at a monitored date of the knock-out barrier
        const bool monitored = barrier && ( as_const( B.mon_mask )[i] & 1u ) && sig > R( 0 );
        if ( monitored )
        {
            const R x = log( S );
            const R zb = sgn * ( hlog - x - mu ) / sig;
            const R q = phic( -zb );  // knock-out probability Phi(zb)
            const R pk = phic( zb );  // survival 1 - q
            const R u1 = phic( -sgn * eps );  // U' = Phi(s eps)
            const R omu = phic( sgn * eps );  // 1 - U'
            const R tt = q + pk * u1, omt = pk * omu;
            const R zc = inv_norm( tt, omt );  // conditioned s-oriented normal Z'
            const R r = omu * expr<R>( R( 0.5 ) * ( zc - zb ) * ( zc + zb ) );
            const R hz = pk > R( 1e-30 ) ? expr<R>( R( -0.5 ) * zb * zb ) * R( 0.3989422804014327 ) / pk
                                         : zb + R( 1 ) / zb;
            // tangents (delta: dmu = dsig = 0)
            lgd += hz * sgn * gx / sig;
            lgv += hz * ( sgn * ( gv + dmu ) + zb * dsig ) / sig;
            gx = gx * ( R( 1 ) - r );
            gv = ( gv + dmu ) * ( R( 1 ) - r ) + dsig * ( sgn * zc - sgn * r * zb );
            wt *= pk;
            eps = sgn * zc;
        }
        else
            gv = gv + dmu + dsig * eps;
        // ---- the step itself (same arithmetic as the path sweep)
        if ( !heston )
        {
            if constexpr ( F64 )
            {
                const double w1 = Wb + as_const( B.sqrt_dt )[i] * eps;
                const double dw = w1 - Wb;
                Wb = w1;
                const double expo = cd - A.vol * A.vol / 2 * as_const( B.dt )[i] + A.vol * dw;
                S = ( S - 0 ) * exp( expo ) + 0;
            }
            else
            {
                const float w1 = Wb + sc[SC_SQDT] * eps;
                const float dw = w1 - Wb;
                Wb = w1;
                S = S * __expf( cd - sc[SC_HALF_V2DT] + sc[SC_VOL] * dw );
            }
        }
        else
        {
            if constexpr ( F64 )
            {
                if ( A.xi > 1e-12 )
                {
                    double var = K3 * v + K3 * vc;
                    if ( var < 0 )
                        var = 0;
                    S = S * exp( mu + sqrt( var ) * eps );
                }
                else
                    S = S * exp( cd - 0.5 * v * as_const( B.dt )[i] + sqrt( v * as_const( B.dt )[i] ) * eps );
            }
            else
                S = S * __expf( mu + sig * eps );
            v = vc;
            dv = dvc;
        }
        // ---- the contract's flow
        if ( as_const( B.flow_mask )[i] & 1u )
        {
            const R f0 = call ? S - K : K - S;
            const R f = f0 > R( 0 ) ? f0 : R( 0 );
            const R fp = f0 > R( 0 ) ? ( call ? R( 1 ) : R( -1 ) ) : R( 0 );  // d payoff / d S
            const R df = (R)B.c_df[0];
            y = df * wt * f;
            dl = df * wt * ( fp * S * gx + f * lgd );
            vg = df * wt * ( fp * S * gv + f * lgv ) * R( 0.01 );  // per vol point
        }
    }
    // ---- block sums of the three estimators (FP64)
    const double e3[3] = { valid ? (double)y : 0.0, valid ? (double)dl : 0.0, valid ? (double)vg : 0.0 };
#pragma unroll
    for ( int k = 0; k < 3; k++ )
    {
        const double s1 = wave_sum<double>( e3[k] );
        const double s2 = wave_sum<double>( e3[k] * e3[k] );
        if ( lane == 0 )
        {
            acc[wib][k][0] = s1;
            acc[wib][k][1] = s2;
        }
    }
    __syncthreads();
    if ( threadIdx.x < 6 )
    {
        const int k = threadIdx.x >> 1, mom = threadIdx.x & 1;
        double s = 0.0;
        for ( int w = 0; w < kWavesPerBlock; w++ )
            s += acc[w][k][mom];
        B.partial[(size_t)blockIdx.x * 6 + threadIdx.x] = s;
    }
}

}  // namespace thoth
