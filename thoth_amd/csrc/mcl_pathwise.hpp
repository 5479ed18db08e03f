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
    bw.start( B, sob, B.factors );

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
    const int F = B.factors;
    auto point = [&]( int o, int f ) __attribute__( ( always_inline ) ) { return sob( B.op_b[o] * F + f ); };
    for ( int p = 0; p < m; p++ )
    {
        const int i = p + 1;
        bw.step( B, point, p, nz, F );
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

// ---------------------------------------------------------------------------
// Pathwise delta and vega of a book of NA Black-Scholes, Heston or LSV
// equities (correlated as the sweep correlates them) holding one european
// vanilla on an equity or on a best-of / worst-of rainbow.  The path is the
// generic sweep's (the same arithmetic, so the premium estimator is the
// sweep's to rounding); each equity carries the forward tangents of its log
// spot x_a with respect to its own spot (gd) and to the parallel vol shift e
// (ge), and a Heston equity the vega tangent of its variance (dv).  Equities
// depend on each other only through the noise, so d x_a / d S0_b = 0 for
// a != b and the per-equity deltas come from one tangent each.
// LSV: the step x' = x + cd + K0 + K1 vp + K2 vc + sqrt(K3 (vp + vc)) z with
// K0..K3 linear and quadratic in the leverage L(t_i, x), so
//   dx' = dx + G dL,  G = dK0/dL + dK1/dL vp + dK2/dL vc + sqrt(var) / L z,
//   dL = L_x dx + L_theta,
// L_x the slope of the layer in log-spot and L_theta the leverage function's
// own response at fixed x to the spot or the vol shift (the calibration's,
// read linearly between the layer's nodes from DevBook::lev_ds / lev_de).
// Heston vega: as pathwise_kernel, through the QE variance map.
// Black-Scholes: d expo / d vol = -vol dt + dW.
// Payoff: vanilla on s = S_u, or s = the rainbow's extreme w_j S_j (weights
// fixed at the rebasing spots, rainbow.cpp:54-57); d f / d x_j* = f' w S_j*.
// Per block: [Y, sum_a delta_a, vega, delta_0 .. delta_{na-1}] x [sum, sum2].
//
// LSV = true: the FP32 form for books of exactly NA LSV equities with
// uncorrelated stochastic spots (the uniform sweep's kind 1): the model
// branches and the variance-factor lookups are compile-time, and the
// leverage with its two sensitivity layers is staged in LDS as one float4
// {L, L_spot, L_vol, 0} per node, G dates at a time into a ring of 2G date
// slots per equity (the sweep's leverage ring and barrier scheme), instead of
// six gathers from device memory per equity and step.  Same arithmetic on
// the same values, so the same results as the generic form.
// ---------------------------------------------------------------------------
template <class R, int NA, int DEPTH, bool LSV = false>
__global__ void __launch_bounds__( kBlock ) pathwise_basket_kernel( const DevBook* __restrict__ books,
                                                                   const uint32_t* __restrict__ V, uint32_t bq0,
                                                                   uint32_t first_idx, uint32_t last_idx )
{
    constexpr bool F64 = std::is_same_v<R, double>;
    constexpr bool lsv = LSV && !F64;
    static_assert( !LSV || !F64, "the staged LSV form is FP32" );
    constexpr int NF = 2 * NA;
    typedef R facv __attribute__( ( ext_vector_type( NF ) ) );
    const DevBook& B = books[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const int wib = threadIdx.x >> 6;
    const uint32_t bq = bq0 + blockIdx.x;
    const uint32_t idx = bq * 256u + (uint32_t)threadIdx.x;
    const bool valid = idx >= first_idx && idx <= last_idx;
    const int m = B.m, nd = B.n_dates, na = B.n_assets;

    extern __shared__ uint32_t lds[];
    __shared__ double acc[kWavesPerBlock][3 + kMaxAssets][2];
    sobol_block_table( lds, V, m * B.factors, bq, nullptr );
    const int P = B.lev_points;
    const int last = P - 1;
    // LSV form, after the block table (16-byte aligned): the sweep's Sobol
    // ring (op o, factor f: 32 words in slot o mod nslots, the lane's point
    // lo[gray(l) & 15] ^ hi[gray(l) >> 4] with the block part folded into hi),
    // then the leverage ring (layer date i of equity a in slot
    // a 2G + (i - 1) mod 2G, node k at + k)
    const int G = B.lev_group;
    const int F = lsv ? NF : B.factors;
    const int nslots = lsv ? B.sob_slots : 0;
    const int tab_words = ( m * B.factors + 3 ) & ~3;
    [[maybe_unused]] uint32_t* ring = lds + tab_words;
    [[maybe_unused]] float4* lring = reinterpret_cast<float4*>( lds + tab_words + nslots * NF * 32 );
    const uint32_t gln = gray32( (uint32_t)threadIdx.x );
    [[maybe_unused]] const uint32_t* rlo = ring + ( gln & 15u );
    [[maybe_unused]] const uint32_t* rhi = ring + 16 + ( gln >> 4 );
    auto so = [&]( int i ) __attribute__( ( always_inline ) ) { return B.step_ops[min( i, m )]; };
    [[maybe_unused]] auto stage_ops = [&]( int o0, int o1 ) __attribute__( ( always_inline ) )
    {
        // four adjacent lanes per (op, factor), each owning 8 of its 32 words
        typedef uint32_t u32x4 __attribute__( ( ext_vector_type( 4 ) ) );
        const uint32_t* nib = V + kSobolNibbles;
        const int n = ( o1 - o0 ) * NF * 4;
        for ( int k = threadIdx.x; k < n; k += kBlock )
        {
            const int q = k & 3;
            const int pair = k >> 2;
            const int o = o0 + pair / NF;
            const int f = pair - ( pair / NF ) * NF;
            const int d = B.op_b[o] * F + f;
            const u32x4* src = reinterpret_cast<const u32x4*>( nib + (size_t)d * 32 + 8 * q );
            u32x4 x0 = src[0], x1 = src[1];
            if ( q >= 2 )
            {
                x0 ^= lds[d];
                x1 ^= lds[d];
            }
            u32x4* dst = reinterpret_cast<u32x4*>( ring + ( ( o & ( nslots - 1 ) ) * NF + f ) * 32 + 8 * q );
            dst[0] = x0;
            dst[1] = x1;
        }
    };
    [[maybe_unused]] auto stage_group = [&]( int first ) __attribute__( ( always_inline ) )
    {
        // dates first .. first + G - 1 are contiguous in every layer array
        // and in the ring (first - 1 is a multiple of G)
        const int run = ( min( first + G - 1, m ) - first + 1 ) * P;
        const int q0 = ( first - 1 ) % ( 2 * G );
        for ( int e = (int)threadIdx.x; e < run; e += kBlock )
        {
            float v[NA][3];
#pragma unroll
            for ( int a = 0; a < NA; a++ )
            {
                const size_t at = ( (size_t)a * nd + first ) * P + e;
                v[a][0] = B.levf[at];
                v[a][1] = B.lev_dsf[at];
                v[a][2] = B.lev_def[at];
            }
#pragma unroll
            for ( int a = 0; a < NA; a++ )
                lring[( a * 2 * G + q0 ) * P + e] = make_float4( v[a][0], v[a][1], v[a][2], 0.0f );
        }
    };
    // the block table is complete before the ops are staged from it
    __syncthreads();
    if constexpr ( lsv )
    {
        // the ops of bridge step 0 and of the first group; its layers
        stage_ops( 0, so( G + 1 ) );
        if ( m >= 1 )
            stage_group( 1 );
    }
    __syncthreads();
    SobolLane sob;
    sob.V = V;
    sob.tab = lds;
    sob.set_lane( wib, lane );
    BridgeWalk<R, NF, DEPTH> bw;
    bw.start( B, sob, F );
    auto point = [&]( int o, int f ) __attribute__( ( always_inline ) ) -> uint32_t
    {
        if constexpr ( lsv )
        {
            const int base = ( ( o & ( nslots - 1 ) ) * NF + f ) * 32;
            return rlo[base] ^ rhi[base];
        }
        else
            return sob( B.op_b[o] * F + f );
    };

    const int udl = B.c_udl[0];
    const R K = (R)B.c_strike[0];
    const bool call = B.c_call[0] != 0;

    R S[NA], Var[NA], W[NA], gd[NA], ge[NA], dv[NA], dl[NA];
#pragma unroll
    for ( int a = 0; a < NA; a++ )
    {
        S[a] = Var[a] = W[a] = gd[a] = ge[a] = dv[a] = dl[a] = R( 0 );
        if ( a < na )
        {
            S[a] = (R)B.asset[a].spot;
            Var[a] = (R)B.asset[a].v0;
            gd[a] = R( 1 ) / (R)B.asset[a].spot;
            // Heston vega seeds d v0 / de = 2 sqrt(v0) (heston_volatility.hpp:45-47)
            dv[a] = B.asset[a].model == 1 ? R( 2 ) * sqrtr( (R)B.asset[a].v0 ) : R( 0 );
        }
    }
    R y = R( 0 ), vg = R( 0 );

    // bridge step p (the increments of step p + 1) runs at the end of step p
    facv nz = (facv)( R( 0 ) );
    if ( m >= 1 )
        bw.step( B, point, 0, nz, F );
    for ( int i = 1; i <= m; i++ )
    {
        if constexpr ( lsv )
            if ( ( i - 1 ) % G == 0 )
            {
                // this group's layers and ops are visible and the previous
                // group's slots free: stage the next group
                __syncthreads();
                if ( i + G <= m )
                    stage_group( i + G );
                stage_ops( so( i + G ), so( i + 2 * G ) );
            }
#pragma unroll
        for ( int a = 0; a < NA; a++ )
        {
            if ( !lsv && a >= na )
                continue;
            const DevAsset& A = B.asset[a];
            const int model = lsv ? 2 : A.model;
            // correlated noise, as the sweep (correlated_noise_node.cpp:15-27)
            R eps = nz[a];
            const size_t crow = ( (size_t)i * kMaxAssets + a ) * kMaxAssets;
            if ( !lsv && na > 1 && ( model == 0 || B.corr_stoch ) )
            {
                eps = R( 0 );
#pragma unroll
                for ( int v = 0; v <= a; v++ )
                {
                    if constexpr ( F64 )
                    {
                        const double c = as_const( B.chol_t )[crow + v];
                        if ( c != 0.0 )
                            eps += nz[v] * c;
                    }
                    else
                        eps += nz[v] * as_const( B.cholf_t )[crow + v];
                }
            }
            cptr<float> sc = as_const( B.sc ) + ( (size_t)i * na + a ) * SC_COUNT;
            R cd;
            if constexpr ( F64 )
                cd = B.drift[(size_t)a * nd + i] * B.t[i] - B.drift[(size_t)a * nd + i - 1] * B.t[i - 1];
            else
                cd = sc[SC_CARRY];
            if ( model == 0 )
            {
                if constexpr ( F64 )
                {
                    const double w1 = W[a] + B.sqrt_dt[i] * eps;
                    const double dw = w1 - W[a];
                    W[a] = w1;
                    const double v = A.vol;
                    const double expo = cd - v * v / 2 * B.dt[i] + v * dw;
                    S[a] = ( S[a] - 0 ) * exp( expo ) + 0;
                    ge[a] += -v * B.dt[i] + dw;
                }
                else
                {
                    const float w1 = W[a] + sc[SC_SQDT] * eps;
                    const float dw = w1 - W[a];
                    W[a] = w1;
                    const float expo = cd - sc[SC_HALF_V2DT] + sc[SC_VOL] * dw;
                    S[a] = S[a] * __expf( expo );
                    ge[a] += -sc[SC_VOL] * sc[SC_DT] + dw;
                }
                continue;
            }
            // LSV form: every equity stochastic, variance noise NA + a
            const int vf = lsv ? NA + a : __builtin_amdgcn_readfirstlane( A.var_factor );
            const R zv = nz[vf];
            const R vp = Var[a];
            R vc, dvc = R( 0 );
            const R theta = (R)A.theta;
            const R dth = model == 1 ? R( 2 ) * sqrtr( theta ) : R( 0 );  // d theta / de
            if constexpr ( F64 )
                vc = qe_f64( vp, B.dt[i], zv, A.kappa, A.theta, A.xi );
            else
                vc = qe_f32( vp, zv, sc );
            if ( model == 1 )
            {
                R e, M0, S1, S0;
                if constexpr ( F64 )
                {
                    e = exp( -A.kappa * B.dt[i] );
                    M0 = A.theta * ( 1 - e );
                    S1 = A.xi * A.xi * e / A.kappa * ( 1 - e );
                    S0 = A.theta * A.xi * A.xi / ( 2 * A.kappa ) * ( 1 - e ) * ( 1 - e );
                }
                else
                {
                    e = sc[SC_E];
                    M0 = sc[SC_M0];
                    S1 = sc[SC_S1];
                    S0 = sc[SC_S0];
                }
                dvc = qe_tangent<R>( vp, zv, e, M0, S1, S0, theta, dv[a], dth );
            }
            // ---- leverage at the previous spot (local_volatility_node.cpp:19-47)
            // and its tangents: slope in log-spot, derivatives at fixed x
            R L = R( 1 ), Lx = R( 0 ), Ls = R( 0 ), Le = R( 0 );
            if ( model == 2 )
            {
                const size_t li = (size_t)a * nd + i;
                R f;  // layer coordinate
                if constexpr ( F64 )
                    f = log( S[a] ) / B.lev_ln_step[li] - B.lev_offset[li];
                else
                {
                    const size_t lf = (size_t)i * na + a;
                    f = fast_log( S[a] ) * as_const( B.lev_inv_step )[lf] - as_const( B.lev_offset_f )[lf];
                }
                const R* lv;
                const R* ds;
                const R* de;
                R inv_step;
                if constexpr ( F64 )
                {
                    lv = B.lev + li * P;
                    ds = B.lev_ds + li * P;
                    de = B.lev_de + li * P;
                    inv_step = 1.0 / B.lev_ln_step[li];
                }
                else
                {
                    lv = B.levf + li * P;
                    ds = B.lev_dsf + li * P;
                    de = B.lev_def + li * P;
                    inv_step = as_const( B.lev_inv_step )[(size_t)i * na + a];
                }
                if constexpr ( F64 )
                {
                    if ( f <= 0.0 )
                    {
                        L = lv[0];
                        Ls = ds[0];
                        Le = de[0];
                    }
                    else if ( f >= (double)last )
                    {
                        L = lv[last];
                        Ls = ds[last];
                        Le = de[last];
                    }
                    else
                    {
                        const int k = (int)f;
                        const double w = f - (double)k;
                        L = ( 1.0 - w ) * lv[k] + w * lv[k + 1];
                        Lx = ( lv[k + 1] - lv[k] ) * inv_step;
                        Ls = ( 1.0 - w ) * ds[k] + w * ds[k + 1];
                        Le = ( 1.0 - w ) * de[k] + w * de[k + 1];
                    }
                }
                else if constexpr ( lsv )
                {
                    // as below, from the LDS ring: one float4 per node
                    const float lev_top = __uint_as_float( __float_as_uint( (float)last ) - 1u );
                    const float fc = __builtin_amdgcn_fmed3f( f, 0.0f, lev_top );
                    const int k = (int)fc;
                    const float w = fc - (float)k;
                    const float4* lq = lring + ( a * 2 * G + ( i - 1 ) % ( 2 * G ) ) * P;
                    const float4 n0 = lq[k], n1 = lq[k + 1];
                    const float sl = n1.x - n0.x;
                    L = fmaf( w, sl, n0.x );
                    const bool inside = f > 0.0f && f < (float)last;
                    Lx = inside ? sl * inv_step : 0.0f;
                    Ls = fmaf( w, n1.y - n0.y, n0.y );
                    Le = fmaf( w, n1.z - n0.z, n0.z );
                }
                else
                {
                    // the sweep's clamp to [0, last) (k <= last - 1, w < 1)
                    const float lev_top = __uint_as_float( __float_as_uint( (float)last ) - 1u );
                    const float fc = __builtin_amdgcn_fmed3f( f, 0.0f, lev_top );
                    const int k = (int)fc;
                    const float w = fc - (float)k;
                    const float sl = lv[k + 1] - lv[k];
                    L = fmaf( w, sl, lv[k] );
                    const bool inside = f > 0.0f && f < (float)last;
                    Lx = inside ? sl * inv_step : 0.0f;
                    Ls = fmaf( w, ds[k + 1] - ds[k], ds[k] );
                    Le = fmaf( w, de[k + 1] - de[k], de[k] );
                }
            }
            // ---- the step (the sweep's arithmetic) and d step / d L
            const R z = eps;
            R step, G, dmu = R( 0 ), dsig = R( 0 );
            if constexpr ( F64 )
            {
                const double xi = A.xi, rho = A.rho, kappa = A.kappa, dt = B.dt[i];
                if ( xi > 1e-12 )
                {
                    const double gg = 0.5;
                    const double K0 = -L * rho * kappa * theta / xi * dt;
                    const double aa = gg * dt * ( L * kappa * rho / xi - 0.5 * L * L );
                    const double K1 = aa - L * rho / xi;
                    const double K2 = aa + L * rho / xi;
                    const double K3 = gg * dt * ( 1.0 - rho * rho ) * L * L;
                    const double K4 = K3;
                    double var = K3 * vp + K4 * vc;
                    if ( var < 0 )
                        var = 0;
                    const double sq = sqrt( var );
                    step = cd + K0 + K1 * vp + K2 * vc + sq * z;
                    const double dK0 = -rho * kappa * theta / xi * dt;
                    const double daa = gg * dt * ( kappa * rho / xi - L );
                    G = dK0 + ( daa - rho / xi ) * vp + ( daa + rho / xi ) * vc + ( L > 0 ? sq / L : 0.0 ) * z;
                    if ( model == 1 )
                    {
                        dmu = -rho * kappa / xi * dt * dth + K1 * dv[a] + K2 * dvc;
                        dsig = sq > 0 ? K3 * ( dv[a] + dvc ) / ( 2 * sq ) : 0.0;
                    }
                }
                else
                {
                    const double sv = sqrt( vp * dt );
                    step = cd - 0.5 * L * L * vp * dt + L * sv * z;
                    G = -L * vp * dt + sv * z;
                    if ( model == 1 )
                    {
                        dmu = -0.5 * dt * dv[a];
                        dsig = sv > 0 ? dt * dv[a] / ( 2 * sv ) : 0.0;
                    }
                }
                S[a] = S[a] * exp( step );
            }
            else
            {
                float det, bvol;
                if ( lsv || A.xi > 1e-12 )
                {
                    const float sv = vp + vc;
                    const float u = fmaf( sc[SC_RX], vc - vp, fmaf( sc[SC_A1], sv, sc[SC_K0] ) );
                    const float l2s = L * L * sv;
                    det = fmaf( -sc[SC_A2], l2s, fmaf( L, u, cd ) );
                    bvol = __builtin_amdgcn_sqrtf( fmaxf( sc[SC_K3] * l2s, 0.0f ) );
                    G = u - 2.0f * sc[SC_A2] * L * sv + ( L > 0.0f ? bvol / L : 0.0f ) * z;
                    if ( model == 1 )
                    {
                        const float aa = sc[SC_A1] - sc[SC_A2];
                        const float dK0 = theta > 0.0f ? sc[SC_K0] / theta * dth : 0.0f;
                        dmu = dK0 + ( aa - sc[SC_RX] ) * dv[a] + ( aa + sc[SC_RX] ) * dvc;
                        dsig = bvol > 0.0f ? sc[SC_K3] * ( dv[a] + dvc ) / ( 2.0f * bvol ) : 0.0f;
                    }
                }
                else
                {
                    const float sv = __builtin_amdgcn_sqrtf( vp * sc[SC_DT] );
                    det = cd - 0.5f * L * L * vp * sc[SC_DT];
                    bvol = L * sv;
                    G = -L * vp * sc[SC_DT] + sv * z;
                    if ( model == 1 )
                    {
                        dmu = -0.5f * sc[SC_DT] * dv[a];
                        dsig = sv > 0.0f ? sc[SC_DT] * dv[a] / ( 2.0f * sv ) : 0.0f;
                    }
                }
                step = det + bvol * z;
                S[a] = S[a] * __expf( step );
            }
            // ---- tangents of the log-spot (the leverage reads the previous spot)
            if ( model == 2 )
            {
                gd[a] += G * ( Lx * gd[a] + Ls );
                ge[a] += G * ( Lx * ge[a] + Le );
            }
            else
            {
                ge[a] += dmu + dsig * z;
                dv[a] = dvc;
            }
            Var[a] = vc;
        }
        // ---- the contract's flow (finance.cpp:9-32, contract_node.cpp:33-47)
        if ( as_const( B.flow_mask )[i] & 1u )
        {
            R s = R( 0 ), w = R( 1 );
            int js = 0;
            if ( udl >= 0 )
            {
#pragma unroll
                for ( int a = 0; a < NA; a++ )
                    if ( a == udl )
                    {
                        s = S[a];
                        js = a;
                    }
            }
            else
            {
                const int r = -1 - udl;  // rainbow_node.cpp:14-27
                const bool best = B.rb_best[r] != 0;
                s = best ? R( -INFINITY ) : R( INFINITY );
                for ( int j = 0; j < B.rb_n[r]; j++ )
                {
                    const int mj = B.rb_member[r][j];
                    R sj = R( 0 );
#pragma unroll
                    for ( int a = 0; a < NA; a++ )
                        if ( a == mj )
                            sj = S[a];
                    const R perf = sj * R( B.rb_weight[r][j] );
                    if ( best ? ( s < perf ) : ( perf < s ) )
                    {
                        s = perf;
                        js = mj;
                        w = R( B.rb_weight[r][j] );
                    }
                }
            }
            const R f0 = call ? s - K : K - s;
            const R f = f0 > R( 0 ) ? f0 : R( 0 );
            const R fp = f0 > R( 0 ) ? ( call ? R( 1 ) : R( -1 ) ) : R( 0 );
            const R df = (R)B.c_df[0];
            y = df * f;
            R sj = R( 0 ), gdj = R( 0 ), gej = R( 0 );
#pragma unroll
            for ( int a = 0; a < NA; a++ )
                if ( a == js )
                {
                    sj = S[a];
                    gdj = gd[a];
                    gej = ge[a];
                }
            const R chain = df * fp * w * sj;  // d payoff / d x_j*
#pragma unroll
            for ( int a = 0; a < NA; a++ )
                dl[a] = a == js ? chain * gdj : R( 0 );
            vg = chain * gej * R( 0.01 );  // per vol point
        }
        if ( i < m )
            bw.step( B, point, i, nz, F );
    }
    // ---- block sums of the estimators (FP64)
    double e[3 + NA];
    double dsum = 0.0;
#pragma unroll
    for ( int a = 0; a < NA; a++ )
    {
        e[3 + a] = valid && a < na ? (double)dl[a] : 0.0;
        dsum += e[3 + a];
    }
    e[0] = valid ? (double)y : 0.0;
    e[1] = dsum;
    e[2] = valid ? (double)vg : 0.0;
    const int ne = 3 + na;
#pragma unroll
    for ( int k = 0; k < 3 + NA; k++ )
    {
        const double s1 = wave_sum<double>( e[k] );
        const double s2 = wave_sum<double>( e[k] * e[k] );
        if ( lane == 0 && k < ne )
        {
            acc[wib][k][0] = s1;
            acc[wib][k][1] = s2;
        }
    }
    __syncthreads();
    if ( (int)threadIdx.x < 2 * ne )
    {
        const int k = threadIdx.x >> 1, mom = threadIdx.x & 1;
        double s = 0.0;
        for ( int w = 0; w < kWavesPerBlock; w++ )
            s += acc[w][k][mom];
        B.partial[(size_t)blockIdx.x * 2 * ne + threadIdx.x] = s;
    }
}

}  // namespace thoth
