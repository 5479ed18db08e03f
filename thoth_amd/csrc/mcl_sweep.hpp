// mcl_sweep.hpp -- the fused Sobol -> inverse-normal -> Brownian-bridge ->
// QE/Andersen/LSV -> payoff -> reduction path-sweep kernel (gfx950, wave64).
//
// One lane = one path.  A wave owns 64 consecutive Sobol indices aligned on 64,
// so gray(64q + j) = gray(64q) ^ gray(j): the wave-uniform part of every Sobol
// coordinate comes from scalar loads of the direction numbers, the per-lane part
// from the 6 low bits of the lane id.  The Brownian bridge is evaluated in time
// order (an in-order walk of the reference's bridge tree) with a per-factor
// stack of pending right anchors held in VGPRs and indexed by a wave-uniform
// stack pointer.  Included by mcl_sweep_f64.hip (compiled with
// -ffp-contract=off, oracle-matched FP64) and mcl_sweep_f32.hip (FP32 stepping).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "mcl_device.hpp"
#include "mcl_mlp.hpp"
#include "mcl_prng.hpp"

namespace thoth
{

__device__ __forceinline__ uint32_t gray32( uint32_t n ) { return n ^ ( n >> 1 ); }

// Tables of a book that no kernel writes while the sweep runs (bridge program,
// step constants, leverage scales, payoff masks) are read through the
// constant address space: at wave-uniform indices those loads become scalar
// loads into SGPRs (scalar cache) instead of flat vector loads into VGPRs
// that wait on vector memory.
template <class T>
using cptr = const __attribute__( ( address_space( 4 ) ) ) T*;
template <class T>
__device__ __forceinline__ cptr<T> as_const( const T* p )
{
    return (cptr<T>)p;
}

// FP32 hardware transcendentals of the production path (v_rcp_f32, v_log_f32:
// about 1 ulp on the normal range, which is all these arguments take)
__device__ __forceinline__ float fast_rcp( float x ) { return __builtin_amdgcn_rcpf( x ); }
__device__ __forceinline__ float fast_div( float a, float b ) { return a * __builtin_amdgcn_rcpf( b ); }
__device__ __forceinline__ float fast_log( float x ) { return __builtin_amdgcn_logf( x ) * 0.693147180559945309f; }

// erfc(x) by the Chebyshev-fitted form of Numerical Recipes in C (2nd ed.,
// section 6.2, "erfcc"): fractional error below 1.2e-7 for every x in exact
// arithmetic, one reciprocal and one exponential; in FP32 on the MI355X at
// most 2.0e-6 relative for |x| <= 4.5 (the arguments z / sqrt 2 of the QE
// exponential branch), from the FP32 exponential of -x^2 + P(t).
__device__ __forceinline__ float erfc_fast( float x )
{
    const float z = fabsf( x );
    const float t = fast_rcp( fmaf( 0.5f, z, 1.0f ) );
    float p = 0.17087277f;
    p = fmaf( p, t, -0.82215223f );
    p = fmaf( p, t, 1.48851587f );
    p = fmaf( p, t, -1.13520398f );
    p = fmaf( p, t, 0.27886807f );
    p = fmaf( p, t, -0.18628806f );
    p = fmaf( p, t, 0.09678418f );
    p = fmaf( p, t, 0.37409196f );
    p = fmaf( p, t, 1.00002368f );
    p = fmaf( p, t, -1.26551223f );
    const float r = t * __expf( fmaf( -z, z, p ) );
    return x >= 0.0f ? r : 2.0f - r;
}

// Integer Sobol coordinates for the 256 consecutive indices of one block,
// aligned on 256.  The reference's sequential Gray-code recurrence
// (sobol_generator.cpp:80-100) gives x_d(n) = XOR_{k in gray(n)} V[d][k], and for
// n = 256B + 64w + j (w < 4, j < 64) gray(n) = gray(256B) ^ gray(64w) ^ gray(j)
// (the summands occupy disjoint bits), so
//   x_d = T_B[d] ^ XOR_{k in gray(64w)} V[d][k] ^ XOR_{k in gray(j)} V[d][k]
// with T_B[d] (block part, plus the RQMC digital shift) tabulated in LDS once
// per block and the intra-block part gray(64 w + j) (bits 0..7) applied by
// eight per-lane masks.
struct SobolLane
{
    const uint32_t* __restrict__ V;
    const uint32_t* tab;  // LDS, block part per dimension
    uint32_t m[8];        // per-lane masks of the bits of gray(64 w + lane)

    // lane 'lane' of wave w: m[k] = all ones iff bit k of gray(64 w + lane) is set
    __device__ __forceinline__ void set_lane( int w, int lane )
    {
        const uint32_t g = gray32( 64u * (uint32_t)w + (uint32_t)lane );
#pragma unroll
        for ( int k = 0; k < 8; k++ )
            m[k] = ( ( g >> k ) & 1u ) ? 0xffffffffu : 0u;
    }

    __device__ __forceinline__ uint32_t operator()( int d ) const
    {
        // d is wave-uniform: the eight direction numbers come in through the
        // scalar cache (one s_load_dwordx8) instead of per-lane vector loads
        const uint32_t* v = V + (size_t)__builtin_amdgcn_readfirstlane( d ) * 32;
        // x ^= v[k] & m[k] as one v_bitop3_b32 each (truth table 0x6c of
        // f(S0, S1, S2) = S1 ^ (S0 & S2), table bit index 4 S0 + 2 S1 + S2)
        uint32_t x = tab[d];
#pragma unroll
        for ( int k = 0; k < 8; k++ )
            x = __builtin_amdgcn_bitop3_b32( v[k], x, m[k], 0x6c );
        return x;
    }
};

// fill the block table: T_B[d] = shift[d] ^ XOR_{k in gray(256 B)} V[d][k]
__device__ __forceinline__ void sobol_block_table( uint32_t* tab, const uint32_t* __restrict__ V, int dims,
                                                   uint32_t bq, const uint32_t* __restrict__ shift )
{
    const uint32_t G = gray32( bq * 256u );
    for ( int d = threadIdx.x; d < dims; d += kBlock )
    {
        const uint32_t* v = V + (size_t)d * 32;
        uint32_t x = shift ? shift[d] : 0u;
        uint32_t gg = G;
        while ( gg )
        {
            x ^= v[__builtin_ctz( gg )];
            gg &= gg - 1;
        }
        tab[d] = x;
    }
}

// single-point variant (one lane = one index, any alignment) for the raw-point entry
__device__ __forceinline__ uint32_t sobol_x( const uint32_t* __restrict__ V, int d, uint32_t n )
{
    const uint32_t* v = V + (size_t)d * 32;
    uint32_t x = 0;
    uint32_t g = gray32( n );
    while ( g )
    {
        x ^= v[__builtin_ctz( g )];
        g &= g - 1;
    }
    return x;
}

// Acklam inverse normal with the reference clamp (path_generator.cpp:148-150,
// distributions.cpp:37-81), FP64.
__device__ __forceinline__ double normal_f64( uint32_t x )
{
    const double u0 = x * ( 1.0 / 4294967296.0 );
    return acklam_f64( fmin( 1.0 - 1e-12, fmax( 1e-12, u0 ) ) );
}

// FP32 inverse normal from the 32-bit integer coordinate, in the erfinv form
// of M. Giles ("Approximating the erfinv function", GPU Computing Gems 2011)
// with coefficients fitted by thoth_amd/csrc/fit_inverse_normal.py to
// Phi^-1 directly (max relative error 3.0e-7 in the fit's FP32 emulation,
// 3.14e-7 measured on the MI355X, whose v_log_f32 is not correctly rounded):
//   x = 2u - 1, w = -log(4u(1-u)), Phi^-1(u) = x P(w).
// 4u(1-u) is formed from the integers u*2^32 and (1-u)*2^32, so it keeps its
// precision in both tails; one log per normal.  The central polynomial
// (degree 10) reaches w = 8, so the tail polynomial runs for 1.7e-4 of the
// draws.  The FP64 oracle mode keeps Acklam.
//
// Instruction economy: x = 2u - 1 is the exact integer x - 2^31 times 2^-31,
// and the 2^-31 is carried by the polynomial coefficients (scaling by a power
// of two commutes with rounding, so the product is the same to the bit); the
// branch test w < 8 is made on L = log2(4u(1-u)) as L > -8 / ln 2.
constexpr float kInvNormScale = 0x1p-31f;
constexpr float kInvNormLogSplit = -11.5415603271117f;  // -8 / ln 2

__device__ __forceinline__ float inv_normal_central( float t )
{
    constexpr float s = kInvNormScale;
    float p = 5.622668877e-10f * s;
    p = fmaf( p, t, -7.025212874e-10f * s );
    p = fmaf( p, t, -4.983229672e-08f * s );
    p = fmaf( p, t, 2.970335407e-07f * s );
    p = fmaf( p, t, 7.164527608e-07f * s );
    p = fmaf( p, t, -2.224123390e-05f * s );
    p = fmaf( p, t, 1.679629502e-04f * s );
    p = fmaf( p, t, -2.876509369e-04f * s );
    p = fmaf( p, t, -1.021180627e-02f * s );
    p = fmaf( p, t, 3.229416742e-01f * s );
    p = fmaf( p, t, 2.628710527e+00f * s );
    return p;
}

__device__ __forceinline__ float normal_f32( uint32_t x )
{
    // u = 0 (x = 0, reachable under a random digital shift) is read as the
    // smallest grid point 2^-32, keeping the normal finite (-6.23)
    x = x > 1u ? x : 1u;
    const float a = (float)x;
    const float b = (float)( 0u - x );
    const float L = __builtin_amdgcn_logf( a * b * 0x1p-62f );
    const float xi = (float)(int32_t)( x - 0x80000000u );
    float p;
    if ( L > kInvNormLogSplit )
        p = inv_normal_central( fmaf( L, -0.693147180559945309f, -4.0f ) );
    else
    {
        constexpr float s = kInvNormScale;
        const float t = __builtin_amdgcn_sqrtf( L * -0.693147180559945309f ) - 3.0f;
        p = 5.879603428e-05f * s;
        p = fmaf( p, t, -5.247431012e-04f * s );
        p = fmaf( p, t, 2.067292677e-03f * s );
        p = fmaf( p, t, -4.845630704e-03f * s );
        p = fmaf( p, t, 7.938630347e-03f * s );
        p = fmaf( p, t, -1.079775749e-02f * s );
        p = fmaf( p, t, 1.336126330e-02f * s );
        p = fmaf( p, t, 1.416582332e+00f * s );
        p = fmaf( p, t, 4.006434188e+00f * s );
    }
    return xi * p;
}

// The central piece of normal_f32 alone (no branch); `tail` is set where
// normal_f32 takes the tail polynomial instead, and for those lanes the
// caller replaces the value by normal_f32's.  x = 0 gives L = -inf: a tail
// lane, so the clamp of normal_f32 is not needed here.
__device__ __forceinline__ float normal_f32_central( uint32_t x, bool& tail )
{
    const float a = (float)x;
    const float b = (float)( 0u - x );
    const float L = __builtin_amdgcn_logf( a * b * 0x1p-62f );
    const float xi = (float)(int32_t)( x - 0x80000000u );
    tail = !( L > kInvNormLogSplit );
    return xi * inv_normal_central( fmaf( L, -0.693147180559945309f, -4.0f ) );
}

template <class R>
__device__ __forceinline__ R normal_of( uint32_t x )
{
    if constexpr ( std::is_same_v<R, double> )
        return normal_f64( x );
    else
        return normal_f32( x );
}

// one pseudo-random normal: Rng::Gaussian in FP64; the FP32 sweeps read the
// draw's 32 high bits as a grid coordinate, as they read a Sobol point
template <class R>
__device__ __forceinline__ R xoshiro_normal( Xoshiro256& x )
{
    if constexpr ( std::is_same_v<R, double> )
        return xoshiro_gauss_f64( x );
    else
        return normal_f32( (uint32_t)( x.next() >> 32 ) );
}

// Andersen QE variance step, FP64, in the reference's operation order
// (heston_variance_node.cpp:17-57).
__device__ __forceinline__ double qe_f64( double v, double dt, double z, double kappa, double theta, double xi )
{
    const double e = exp( -kappa * dt );
    const double m = theta + ( v - theta ) * e;
    const double s2 = v * xi * xi * e / kappa * ( 1 - e ) + theta * xi * xi / ( 2 * kappa ) * ( 1 - e ) * ( 1 - e );
    const double psi = ( m > 0 ) ? s2 / ( m * m ) : 0;
    if ( psi <= 1.5 )
    {
        const double c = 2.0 / psi;
        const double b2 = c - 1.0 + sqrt( c ) * sqrt( c - 1.0 );
        const double a = m / ( 1.0 + b2 );
        const double bz = sqrt( b2 ) + z;
        return a * bz * bz;
    }
    const double p = ( psi - 1.0 ) / ( psi + 1.0 );
    const double beta = ( 1.0 - p ) / m;
    const double u = 0.5 * erfc( -z / sqrt( 2.0 ) );
    return ( u <= p ) ? 0.0 : log( ( 1.0 - p ) / ( 1.0 - u ) ) / beta;
}

// The QE draw actually taken: quadratic branch v' = a (sqrt(b2) + Z)^2, or
// exponential branch (mass p at 0, else exponential with rate beta).
struct QeDraw
{
    float mean;  // E[v' | v] of the FP32 draw as computed below
};

// Same scheme in FP32 from host-precomputed step constants, in three parts
// so that the uniform-book kernels can run the quadratic branch for every
// lane and the exponential one for the lanes that need it (same arithmetic
// as taking the branch).  1-u is formed directly as Phi(-z) so the
// exponential tail keeps its precision.
// psi = s2 / m^2 is formed only where the exponential branch needs it: the
// branch test is s2 <= 1.5 m^2 and the quadratic branch reads 2 / psi as
// 2 m^2 / s2 (one reciprocal instead of two).
struct QeMoments
{
    float m, m2, s2;
    bool quad;  // psi <= 1.5 (psi = 0 when m <= 0)
};

template <class P>
__device__ __forceinline__ QeMoments qe_moments( float v, P sc )
{
    const float m = fmaf( v, sc[SC_E], sc[SC_M0] );
    const float s2 = fmaf( v, sc[SC_S1], sc[SC_S0] );
    const float m2 = m * m;
    return { m, m2, s2, !( m > 0.0f ) || s2 <= 1.5f * m2 };
}

// psi <= 1.5: v' = a (sqrt(b2) + z)^2 with c = 2 m^2 / s^2 = 2 / psi,
// b2 = c - 1 + sqrt(c (c - 1)) and c (c - 1) = (c - 1)^2 + (c - 1)
__device__ __forceinline__ float qe_quadratic( const QeMoments& q, float z, QeDraw* draw )
{
    const float cm1 = fmaf( q.m2 + q.m2, fast_rcp( q.s2 ), -1.0f );
    const float b2 = cm1 + __builtin_amdgcn_sqrtf( fmaf( cm1, cm1, cm1 ) );
    const float a = fast_div( q.m, 1.0f + b2 );
    const float sb = __builtin_amdgcn_sqrtf( b2 );
    const float bz = sb + z;
    if ( draw )
        draw->mean = a * fmaf( sb, sb, 1.0f );  // E[a (sb + Z)^2]
    return a * bz * bz;
}

// psi > 1.5: p = (psi - 1) / (psi + 1), beta = (1 - p) / m, and
// v' = log((1 - p) / (1 - u)) / beta, 0 where u <= p.  With h = (psi + 1) / 2,
// 1 - p = 1 / h and 1 / beta = m h, so v' = -log(h (1 - u)) m h, 0 where
// h (1 - u) >= 1; 1 - u = erfc(z / sqrt 2) / 2.  E[v'] = (1 - p) / beta = m.
__device__ __forceinline__ float qe_exponential( const QeMoments& q, float z, QeDraw* draw )
{
    const float h4 = fmaf( q.s2, fast_rcp( q.m2 ), 1.0f ) * 0.25f;  // h / 2
    if ( draw )
        draw->mean = q.m;
    const float A = h4 * erfc_fast( z * 0.70710678118654752f );  // h (1 - u)
    return A >= 1.0f ? 0.0f
                     : __builtin_amdgcn_logf( A ) * ( -0.693147180559945309f * 2.0f ) * ( q.m * h4 );
}

template <class P>
__device__ __forceinline__ float qe_f32( float v, float z, P sc, QeDraw* draw = nullptr )
{
    const QeMoments q = qe_moments( v, sc );
    if ( q.quad )
        return qe_quadratic( q, z, draw );
    return qe_exponential( q, z, draw );
}

template <class R>
__device__ __forceinline__ R wave_sum( R x )
{
#pragma unroll
    for ( int off = 32; off >= 1; off >>= 1 )
        x += __shfl_xor( x, off, 64 );
    return x;
}

// Brownian bridge of the reference schedule (path_generator.cpp:49-198)
// replayed as an in-order tree walk: the endpoint first, then per step the
// pending points of the step are built (their right anchors pushed on a
// per-factor stack held in VGPRs, indexed by a wave-uniform position) and
// W(t_i) is popped as the normalised increment.  The walk depends on the
// Sobol points only, never on the path state.
template <class R, int NF, int DEPTH>
struct BridgeWalk
{
    // depth-8 stacks two factors per 16-element vector (indexed by one
    // s_set_gpr_idx move instead of eight selects), as in sweep_block
    static constexpr bool paired = DEPTH == 8 && NF % 2 == 0;
    typedef R stackv __attribute__( ( ext_vector_type( DEPTH ) ) );
    typedef R stack2v __attribute__( ( ext_vector_type( 16 ) ) );
    stackv stk[paired ? 1 : NF];
    stack2v stk2[paired ? NF / 2 : 1];
    R top[NF], prev[NF];
    int sp;

    __device__ __forceinline__ R stk_get( int f, int pos ) const
    {
        if constexpr ( paired )
            return stk2[f >> 1][pos + ( ( f & 1 ) << 3 )];
        else
            return stk[f][pos];
    }
    __device__ __forceinline__ void stk_set( int f, int pos, R x )
    {
        if constexpr ( paired )
            stk2[f >> 1][pos + ( ( f & 1 ) << 3 )] = x;
        else
            stk[f][pos] = x;
    }

    // endpoint first (build index 0, path_generator.cpp:166-167); F the
    // book's factor count (a compile-time NF where the caller knows it)
    __device__ __forceinline__ void start( const DevBook& B, const SobolLane& sob, int F )
    {
        if constexpr ( paired )
        {
#pragma unroll
            for ( int k = 0; k < NF / 2; k++ )
                stk2[k] = (stack2v)( R( 0 ) );
        }
        else
        {
#pragma unroll
            for ( int f = 0; f < NF; f++ )
                stk[f] = (stackv)( R( 0 ) );
        }
#pragma unroll
        for ( int f = 0; f < NF; f++ )
        {
            prev[f] = R( 0 );
            top[f] = R( 0 );
            if ( f < F )
            {
                const R z = normal_of<R>( sob( f ) );
                if constexpr ( std::is_same_v<R, double> )
                    top[f] = B.sd0 * z;
                else
                    top[f] = B.sd0f * z;
            }
        }
        sp = 1;
    }

    // bridge step p: build the pending points, then pop W(t_{p+1}) into nz
    // (normalised increment per factor, path_generator.cpp:175-196);
    // point(o, f) is the Sobol point of op o, factor f (dimension op_b[o] F + f)
    template <class V, class Pt>
    __device__ __forceinline__ void step( const DevBook& B, const Pt& point, int p, V& nz, int F )
    {
        constexpr bool F64 = std::is_same_v<R, double>;
        const int i = p + 1;
        const int o0 = B.step_ops[p], o1 = B.step_ops[p + 1];
        for ( int o = o0; o < o1; o++ )
        {
            const bool left = as_const( B.op_left )[o] != 0;
            const int pos = __builtin_amdgcn_readfirstlane( sp - 1 );
            R lw, rw, sd;
            if constexpr ( F64 )
            {
                lw = B.op_w[3 * o];
                rw = B.op_w[3 * o + 1];
                sd = B.op_w[3 * o + 2];
            }
            else
            {
                lw = as_const( B.op_wf )[3 * o];
                rw = as_const( B.op_wf )[3 * o + 1];
                sd = as_const( B.op_wf )[3 * o + 2];
            }
            // FP32 with four or more factors: the central polynomial for
            // every factor straight-line, the tail lanes (w >= 8) redone in
            // one block (normal_f32's values); otherwise each normal where it
            // is used (with two factors, of which the book may use one, the
            // staging and mask bookkeeping cost more than the branch)
            constexpr bool central = !F64 && NF >= 4;
            auto build = [&]( int f, R z ) __attribute__( ( always_inline ) )
            {
                const R nv = left ? lw * prev[f] + rw * top[f] + sd * z : rw * top[f] + sd * z;
                stk_set( f, pos, top[f] );
                top[f] = nv;
            };
            if constexpr ( central )
            {
                R zf[NF];
                uint32_t tails = 0;
#pragma unroll
                for ( int f = 0; f < NF; f++ )
                {
                    zf[f] = R( 0 );
                    if ( f < F )
                    {
                        bool t;
                        zf[f] = normal_f32_central( point( o, f ), t );
                        tails |= t ? 1u << f : 0u;
                    }
                }
                if ( __builtin_amdgcn_ballot_w64( tails != 0 ) )
                {
#pragma unroll
                    for ( int f = 0; f < NF; f++ )
                        if ( ( tails >> f ) & 1u )
                            zf[f] = normal_f32( point( o, f ) );
                }
#pragma unroll
                for ( int f = 0; f < NF; f++ )
                    if ( f < F )
                        build( f, zf[f] );
            }
            else
            {
#pragma unroll
                for ( int f = 0; f < NF; f++ )
                    if ( f < F )
                        build( f, normal_of<R>( point( o, f ) ) );
            }
            sp++;
        }
        const int pos = __builtin_amdgcn_readfirstlane( sp >= 2 ? sp - 2 : 0 );
#pragma unroll
        for ( int f = 0; f < NF; f++ )
        {
            nz[f] = R( 0 );
            if ( f < F )
            {
                const R w = top[f];
                if constexpr ( F64 )
                    nz[f] = ( B.bsq[i] > 0 ) ? ( w - prev[f] ) / B.bsq[i] : 0.0;
                else
                    nz[f] = ( w - prev[f] ) * as_const( B.binv )[i];
                prev[f] = w;
                top[f] = stk_get( f, pos );  // stale only after the last pop
            }
        }
        sp--;
    }
};

// LDS of the in-loop hedge inference (after the Sobol table and leverage
// layers, 16-byte aligned): the BF16 A-operand image of the network, its
// biases, and per wave a [64 paths][16] BF16 feature block and a [64][8]
// FP32 hedge-ratio block for the lane <-> MFMA-layout exchange.
static_assert( kCvImgBytes == kMlpTiles * 64 * 8 && kCvBiasBytes == kMlpBSize * 4, "hedge-network LDS area" );

// CV: 0 plain sweep; 1 also writes the control-variate training samples;
// 2 (SiLU network) and 3 (smoothstep-gated network) evaluate the hedge network
// at every hedge date inside the stepping loop and accumulate P - CV in an
// extra result slot after the contracts.
// One block of 256 consecutive Sobol indices (Sobol block bq) of book B,
// replicate z (RQMC), partial-sum slot (z, tile) of ntiles per replicate.
// KIND: 0 any book; 1 (2) every equity of every book is LSV (Heston) with
// vol-of-vol > 1e-12 and uncorrelated stochastic spots, and the book has
// exactly NA equities -- the host checks this -- so the per-equity model
// branches fold away and the equities' steps interleave in one basic block;
// 3 as 1 with the spot noises Cholesky-correlated across the equities;
// 4 any book (the arithmetic of 0) whose normals are not all Sobol points:
// the dimensions past the Sobol budget read precomputed pseudo-random
// normals, and pseudo books (use_sobol false) draw every normal from the
// lane's xoshiro256++ stream; 5 NA single-equity LSV books on one grid and
// one set of Sobol points stepped together as the NA equities of one book:
// every equity reads the two shared factors (its book's spot and variance
// noises are the same normals), so the normals and the bridge are computed
// once for the NA books, and every equity carries its own vanilla grid.
template <class R, int NA, int DEPTH, int CV, int KIND = 0>
__device__ __forceinline__ void sweep_block( const DevBook& B, const uint32_t* __restrict__ V, uint32_t bq,
                                             uint32_t first_idx, uint32_t last_idx,
                                             const uint32_t* __restrict__ shifts, int z, uint32_t tile,
                                             uint32_t ntiles )
{
    constexpr bool F64 = std::is_same_v<R, double>;
    constexpr bool prng = KIND == 4;  // pseudo-random draws
    constexpr bool uniform = KIND != 0 && !prng;
    constexpr bool grouped = KIND == 5;                            // NA books sharing their two factors
    constexpr bool lsv_kind = KIND == 1 || KIND == 3 || grouped;  // every equity LSV
    constexpr bool corr_kind = KIND == 3;              // correlated spot noises
    constexpr bool qe_phase = uniform;
    const int nas = uniform ? NA : B.n_assets;  // row length of the date-major per-step arrays
    static_assert( !uniform || !F64, "uniform-book kernels are FP32" );
    static_assert( !prng || CV == 0, "pseudo-random draws run without the control variate" );
    constexpr int NF = grouped ? 2 : 2 * NA;
    constexpr int NFV = NF;
    constexpr int NAV = NA;
    typedef R stackv __attribute__( ( ext_vector_type( DEPTH ) ) );
    typedef R facv __attribute__( ( ext_vector_type( NFV ) ) );
    typedef R assetv __attribute__( ( ext_vector_type( NAV ) ) );

    const int lane = threadIdx.x & 63;
    const int wib = threadIdx.x >> 6;
    const uint32_t idx = bq * 256u + (uint32_t)threadIdx.x;
    const bool valid = idx >= first_idx && idx <= last_idx;
    const uint32_t* sh = shifts ? shifts + (size_t)z * kSobolDimsDev : nullptr;

    // pseudo-random draws: the lane's path (index - first) is its position in
    // the book's one stream of draws, whose start state the batch computed
    // once for every book of the same T; invalid lanes read path 0
    const bool pseudo = prng && B.pseudo != 0;
    [[maybe_unused]] const size_t prow = valid ? (size_t)( idx - first_idx ) : 0;
    [[maybe_unused]] Xoshiro256 prs;
    if constexpr ( prng )
        if ( pseudo )
        {
#pragma unroll
            for ( int q = 0; q < 4; q++ )
                prs.s[q] = B.states[(size_t)q * B.tail_stride + prow];
        }

    // uniform books: every equity stochastic, so 2 NA factors (2 shared
    // ones when grouped) and equity a's variance noise is factor NA + a
    const int F = uniform ? NF : B.factors;
    const int m = B.m;
    const int na = B.n_assets;
    const int nd = B.n_dates;
    const int nc = B.n_contracts;

    // dynamic LDS: [Sobol block table: m F words padded to 4, except in the
    // in-loop inference kernels][Sobol ring: nslots x NF factors x 32 words]
    // [FP32 LSV: 2G x NA x lev_points leverage layers][in-loop hedge inference
    // or grid area, 16-byte aligned]
    constexpr bool block_table = CV < 2;
    extern __shared__ uint32_t lds[];
    const int nslots = B.sob_slots;
    const int tab_words = block_table && !pseudo ? ( min( m * F, kSobolDimsDev ) + 3 ) & ~3 : 0;
    const int ring_words = tab_words + nslots * NF * 32;  // through the end of the ring
    uint32_t* tab = lds;
    uint32_t* ring = lds + tab_words;
    float* levs = reinterpret_cast<float*>( lds + ring_words );
    // vanilla-grid area (never together with the control variate)
    [[maybe_unused]] char* gl = nullptr;
    if constexpr ( CV == 0 )
        if ( B.grid_nm > 0 )
        {
            const int lev_words = ( !F64 && B.lev_points > 0 ) ? 2 * B.lev_group * NA * B.lev_points : 0;
            gl = reinterpret_cast<char*>( lds + ( ( ring_words + lev_words + 3 ) & ~3 ) );
        }
    constexpr bool cv_on = CV != 0 && !F64;
    constexpr bool cv_samples = CV == 1 && !F64;
    constexpr bool cv_infer = CV >= 2 && !F64;
    constexpr int cv_act = CV == 3 ? kActGate : kActSilu;
    [[maybe_unused]] char* cvl = nullptr;
    if constexpr ( cv_infer )
    {
        const int lev_words = B.lev_points > 0 ? 2 * B.lev_group * NA * B.lev_points : 0;
        cvl = reinterpret_cast<char*>( lds + ( ( ring_words + lev_words + 3 ) & ~3 ) );
        const uint32_t* src = B.cv_wimg;
        uint32_t* dst = reinterpret_cast<uint32_t*>( cvl );
        for ( int k = threadIdx.x; k < kCvImgBytes / 4; k += kBlock )
            dst[k] = src[k];
        float* bd = reinterpret_cast<float*>( cvl + kCvImgBytes );
        for ( int k = threadIdx.x; k < kMlpBSize; k += kBlock )
            bd[k] = B.cv_bias[k];
    }
    __shared__ double acc[kWavesPerBlock][kMaxContracts + 1][2];
    for ( int k = threadIdx.x; k < kWavesPerBlock * ( kMaxContracts + 1 ) * 2; k += kBlock )
        ( &acc[0][0][0] )[k] = 0.0;
    // ---- Sobol points.  Point n = 256 bq + l (l = threadIdx.x) of dimension d
    // is shift[d] ^ XOR of V[d][k] over the set bits k of gray(n), and
    // gray(n) = gray(l) ^ gray(256 bq) since gray code is linear in XOR.  The
    // block part (shift and gray(256 bq)) is tabulated per dimension at the
    // start, or, in the in-loop inference kernels whose LDS holds the network,
    // computed where it is needed.  The lane part splits into the nibbles of
    // the 8-bit gray(l), whose 16-entry tables per dimension
    //   lo[c] = XOR_{k < 4, bit k of c} V[d][k],  hi[c] = XOR_{k < 4, bit k of c} V[d][4 + k]
    // follow the direction numbers in device memory.  For every bridge op o
    // (build index b: dimensions b F + f) the block copies, per factor, lo and
    // hi ^ (block part) into an LDS ring (op o in slot o mod nslots), and lane
    // l reads lo[gray(l) & 15] ^ hi[gray(l) >> 4]: two LDS reads and one XOR
    // per normal.  The copies of a group of G steps are made one group ahead
    // and published by the barrier that starts the group (the leverage
    // barrier).
    typedef uint32_t u32x4 __attribute__( ( ext_vector_type( 4 ) ) );
    const uint32_t* nib = V + kSobolNibbles;
    const uint32_t gbq = gray32( bq * 256u );
    const uint32_t gln = gray32( (uint32_t)threadIdx.x );
    const uint32_t* rlo = ring + ( gln & 15u );
    const uint32_t* rhi = ring + 16 + ( gln >> 4 );
    auto stage_ops = [&]( int o0, int o1 ) __attribute__( ( always_inline ) )
    {
        // four adjacent lanes per (op, factor), lane q of the four owning
        // words 8q .. 8q+7 of the direction numbers and of the nibble tables
        // (n and kBlock are multiples of four: the four lanes run together)
        if ( pseudo )
            return;
        const int n = ( o1 - o0 ) * NF * 4;
        for ( int k = threadIdx.x; k < n; k += kBlock )
        {
            const int q = k & 3;
            const int pair = k >> 2;  // (o - o0) NF + f
            const int o = o0 + pair / NF;
            const int f = pair - ( pair / NF ) * NF;
            if ( f >= F )
                continue;
            const int d = B.op_b[o] * F + f;
            if constexpr ( prng )
                if ( d >= kSobolDimsDev )
                    continue;  // past the Sobol budget: read from the tail normals
            uint32_t blk;
            if constexpr ( block_table )
                blk = tab[d];
            else
            {
                // XOR of V[d][k] over the bits k of gray(256 bq), eight words
                // per lane, then across the quad (DPP quad permutations
                // [1 0 3 2] and [2 3 0 1]), and the shift
                const u32x4* vr = reinterpret_cast<const u32x4*>( V + (size_t)d * 32 + 8 * q );
                const u32x4 v0 = vr[0], v1 = vr[1];
                const uint32_t bits = gbq >> ( 8 * q );
                blk = 0;
#pragma unroll
                for ( int j = 0; j < 4; j++ )
                {
                    blk ^= v0[j] & ( 0u - ( ( bits >> j ) & 1u ) );
                    blk ^= v1[j] & ( 0u - ( ( bits >> ( 4 + j ) ) & 1u ) );
                }
                blk ^= (uint32_t)__builtin_amdgcn_mov_dpp( (int)blk, 0xB1, 0xF, 0xF, true );
                blk ^= (uint32_t)__builtin_amdgcn_mov_dpp( (int)blk, 0x4E, 0xF, 0xF, true );
                blk ^= sh ? sh[d] : 0u;
            }
            const u32x4* src = reinterpret_cast<const u32x4*>( nib + (size_t)d * 32 + 8 * q );
            u32x4 x0 = src[0], x1 = src[1];
            if ( q >= 2 )
            {
                x0 ^= blk;
                x1 ^= blk;
            }
            u32x4* dst = reinterpret_cast<u32x4*>( ring + ( ( o & ( nslots - 1 ) ) * NF + f ) * 32 + 8 * q );
            dst[0] = x0;
            dst[1] = x1;
        }
    };
    // the point of op o, factor f (o wave-uniform)
    auto sob_op = [&]( int o, int f ) __attribute__( ( always_inline ) ) -> uint32_t
    {
        const int base = ( ( o & ( nslots - 1 ) ) * NF + f ) * 32;
        return rlo[base] ^ rhi[base];
    };
    // a whole point, for the bridge endpoint (build index 0, dimensions 0 .. F-1)
    auto sob_point = [&]( int d ) __attribute__( ( always_inline ) ) -> uint32_t
    {
        if constexpr ( block_table )
        {
            const uint32_t* t = nib + (size_t)d * 32;
            return tab[d] ^ t[gln & 15u] ^ t[16 + ( gln >> 4 )];
        }
        const uint32_t* v = V + (size_t)d * 32;
        uint32_t x = sh ? sh[d] : 0u;
        for ( uint32_t g = gbq ^ gln; g; g &= g - 1 )
            x ^= v[__builtin_ctz( g )];
        return x;
    };
    auto so = [&]( int i ) __attribute__( ( always_inline ) ) { return B.step_ops[min( i, m )]; };
    // FP32 LSV books stage the leverage layers in LDS, G layers at a time into
    // a ring of 2G slots (layer i in slot (i - 1) mod 2G): one barrier per G
    // steps publishes the current group and frees the previous group's slots
    const bool lev_lds = !F64 && KIND != 2 && B.lev_points > 0;  // uniform Heston books have no leverage
    const int P = B.lev_points;
    const int G = B.lev_group;
    // the largest float below the last leverage point P - 1 (P >= 2)
    [[maybe_unused]] const float lev_top = __uint_as_float( __float_as_uint( (float)( P - 1 ) ) - 1u );
    // The ring holds each equity's 2G layer slots contiguously (equity a's
    // slot q at (a 2G + q) P), so a group of an equity is one contiguous run
    // of layers x P floats in the book's leverage array and in LDS.  Each
    // thread takes elements tid + j kBlock of every equity's run and issues
    // all its loads (up to kStageLoads) before its first store, so a group
    // costs one memory latency rather than one per row.
    constexpr int kStageLoads = 8;
    constexpr int kStageRuns = NA >= kStageLoads ? 1 : kStageLoads / NA;
    auto stage_group = [&]( int first ) __attribute__( ( always_inline ) )
    {
        const int run = ( min( first + G - 1, m ) - first + 1 ) * P;
        // layer first + l goes to slot (first - 1) mod 2G + l (first - 1 is a multiple of G)
        const int q0 = ( first - 1 ) % ( 2 * G );
        for ( int e0 = (int)threadIdx.x; e0 < run; e0 += kBlock * kStageRuns )
        {
            float v[kStageRuns][NA];
#pragma unroll
            for ( int j = 0; j < kStageRuns; j++ )
#pragma unroll
                for ( int a = 0; a < NA; a++ )
                    if ( e0 + j * kBlock < run && ( uniform || a < na ) )
                        v[j][a] = B.levf[( (size_t)a * nd + first ) * P + e0 + j * kBlock];
#pragma unroll
            for ( int j = 0; j < kStageRuns; j++ )
#pragma unroll
                for ( int a = 0; a < NA; a++ )
                    if ( e0 + j * kBlock < run && ( uniform || a < na ) )
                        levs[( a * 2 * G + q0 ) * P + e0 + j * kBlock] = v[j][a];
        }
    };
    if constexpr ( block_table )
        if ( !pseudo )
        {
            sobol_block_table( tab, V, min( m * F, kSobolDimsDev ), bq, sh );
            __syncthreads();
        }
    // the ops of bridge_step(0) and of the first group's steps
    stage_ops( 0, so( G + 1 ) );
    if ( lev_lds && m >= 1 )
        stage_group( 1 );
    __syncthreads();

    // ---- bridge state: endpoint first (build index 0, path_generator.cpp:166-167)
    // Stacks of depth 8 are held two factors per 16-element register
    // vector (factor 2k in elements 0..7, 2k+1 in 8..15): the backend indexes
    // a 16-element vector with s_set_gpr_idx and one move, while it lowers a
    // variable index into an 8-element vector as eight selects.
    constexpr bool paired = DEPTH == 8;
    typedef R stack2v __attribute__( ( ext_vector_type( 16 ) ) );
    [[maybe_unused]] stackv stk[paired ? 1 : NF];
    [[maybe_unused]] stack2v stk2[paired ? NF / 2 : 1];
    auto stk_get = [&]( int f, int pos ) __attribute__( ( always_inline ) ) -> R
    {
        if constexpr ( paired )
            return stk2[f >> 1][pos + ( ( f & 1 ) << 3 )];
        else
            return stk[f][pos];
    };
    auto stk_set = [&]( int f, int pos, R x ) __attribute__( ( always_inline ) )
    {
        if constexpr ( paired )
            stk2[f >> 1][pos + ( ( f & 1 ) << 3 )] = x;
        else
            stk[f][pos] = x;
    };
    if constexpr ( paired )
    {
#pragma unroll
        for ( int k = 0; k < NF / 2; k++ )
            stk2[k] = (stack2v)( R( 0 ) );
    }
    else
    {
#pragma unroll
        for ( int f = 0; f < NF; f++ )
            stk[f] = (stackv)( R( 0 ) );
    }
    R top[NF], prev[NF];
#pragma unroll
    for ( int f = 0; f < NF; f++ )
    {
        prev[f] = R( 0 );
        top[f] = R( 0 );
        if ( f < F && !pseudo )
        {
            const R z = normal_of<R>( sob_point( f ) );
            if constexpr ( F64 )
                top[f] = B.sd0 * z;
            else
                top[f] = B.sd0f * z;
        }
    }
    int sp = 1;

    // ---- equity state.  The uniform LSV kernel keeps log S in S (the leverage
    // is read at log S; the spot is exponentiated only where a payoff or the
    // control variate reads it); every other kernel keeps S.
    constexpr bool log_spot = lsv_kind;
    assetv S, Var, W;
#pragma unroll
    for ( int a = 0; a < NAV; a++ )
    {
        S[a] = R( 0 );
        Var[a] = R( 0 );
        W[a] = R( 0 );
    }
#pragma unroll
    for ( int a = 0; a < NA; a++ )
        if ( a < na )
        {
            S[a] = log_spot ? R( log( B.asset[a].spot ) ) : R( B.asset[a].spot );
            Var[a] = R( B.asset[a].v0 );
        }
    auto spot_of = [&]( int a ) __attribute__( ( always_inline ) ) -> R
    {
        if constexpr ( log_spot )
            return __expf( S[a] );
        else
            return S[a];
    };
    uint32_t hit = 0;
    double book = 0.0;

    // ---- learned control variate (FP32 only)
    const int64_t pidx = (int64_t)idx - (int64_t)first_idx;
    const bool cv_store = cv_samples && valid && pidx >= 0 && pidx < B.cv_N;
    float dmacc[NA];
    float hr[NA];     // hedge ratios of the current interval (in-loop inference)
    float cvacc = 0.0f;
#pragma unroll
    for ( int a = 0; a < NA; a++ )
    {
        dmacc[a] = 0.0f;
        hr[a] = 0.0f;
    }
    // standardised features, BF16 pairs: 5 log-moneyness / (vol_ref sqrt(T)),
    // 5 x 2 (sqrt(v) / vol_ref - 1) (0 for Black-Scholes), 2 t / T - 1
    auto cv_feature_words = [&]( int i, uint32_t* w ) __attribute__( ( always_inline ) )
    {
        float f[16];
#pragma unroll
        for ( int k = 0; k < 16; k++ )
            f[k] = 0.0f;
#pragma unroll
        for ( int a = 0; a < NA; a++ )
            if ( a < na && a < 5 )
            {
                f[a] = ( ( log_spot ? (float)S[a] : fast_log( (float)S[a] ) ) - B.cv_lnS0[a] ) * B.cv_ln_scale[a];
                if ( uniform || B.asset[a].model != 0 )
                    f[5 + a] = 2.0f * fmaf( __builtin_amdgcn_sqrtf( fmaxf( (float)Var[a], 0.0f ) ),
                                            B.cv_vol_scale[a], -1.0f );
            }
        f[10] = fmaf( 2.0f * (float)B.t[i], B.cv_inv_T, -1.0f );
#pragma unroll
        for ( int k = 0; k < 8; k++ )
            w[k] = pack2_bf16( f[2 * k], f[2 * k + 1] );
    };
    auto cv_features = [&]( int i, int j ) __attribute__( ( always_inline ) )
    {
        uint32_t w[8];
        cv_feature_words( i, w );
        uint4* dst = reinterpret_cast<uint4*>( B.cv_feat + ( ( (size_t)z * B.cv_J + j ) * B.cv_N + pidx ) * 16 );
        dst[0] = uint4{ w[0], w[1], w[2], w[3] };
        dst[1] = uint4{ w[4], w[5], w[6], w[7] };
    };
    // in-loop inference: the wave's 64 paths as 4 MFMA sample tiles; features
    // go lane -> LDS -> B operands, hedge ratios D layout -> LDS -> lanes
    auto hedge_eval = [&]( int i ) __attribute__( ( always_inline ) )
    {
        uint32_t w[8];
        cv_feature_words( i, w );
        char* wl = cvl + kCvImgBytes + kCvBiasBytes + wib * kCvWaveBytes;
        uint4* frow = reinterpret_cast<uint4*>( wl ) + 2 * lane;
        frow[0] = uint4{ w[0], w[1], w[2], w[3] };
        frow[1] = uint4{ w[4], w[5], w[6], w[7] };
        float* hbuf = reinterpret_cast<float*>( wl + 64 * 16 * 2 );
        wave_lds_sync();
        const uint2* wimg = reinterpret_cast<const uint2*>( cvl );
        const float* wb = reinterpret_cast<const float*>( cvl + kCvImgBytes );
        // two sample tiles per iteration: their independent MFMA chains and
        // activations interleave (no extra registers at this kernel's peak)
#pragma unroll 2
        for ( int t = 0; t < 4; t++ )
        {
            const uint2 xw = *reinterpret_cast<const uint2*>(
                reinterpret_cast<const uint16_t*>( wl ) + ( 16 * t + ( lane & 15 ) ) * 16 + 4 * ( lane >> 4 ) );
            const f32x4 out = mlp_infer<cv_act>( wimg, wb, bits_bf16x4( xw ), lane );
            const int row = 16 * t + ( lane & 15 );
            if ( lane < 16 )
                *reinterpret_cast<float4*>( hbuf + row * 8 ) = float4{ out[0], out[1], out[2], out[3] };
            else if ( lane < 32 )
                hbuf[row * 8 + 4] = out[0];
        }
        wave_lds_sync();
#pragma unroll
        for ( int a = 0; a < NA; a++ )
            if ( a < 5 )
                hr[a] = hbuf[lane * 8 + a];
        wave_lds_sync();
    };
    auto cv_date = [&]( int i ) __attribute__( ( always_inline ) )
    {
        if constexpr ( cv_infer )
        {
            if ( as_const( B.cv_start )[i] >= 0 )
                hedge_eval( i );
            return;
        }
        const int je = as_const( B.cv_end )[i];
        if ( je >= 0 )
        {
            if ( cv_store )
            {
                float o[8] = { 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f };
#pragma unroll
                for ( int a = 0; a < NA; a++ )
                    if ( a < 8 )
                        o[a] = dmacc[a];
                float4* dst = reinterpret_cast<float4*>( B.cv_dm + ( ( (size_t)z * B.cv_J + je ) * B.cv_N + pidx ) * 8 );
                dst[0] = float4{ o[0], o[1], o[2], o[3] };
                dst[1] = float4{ o[4], o[5], o[6], o[7] };
            }
#pragma unroll
            for ( int a = 0; a < NA; a++ )
                dmacc[a] = 0.0f;
        }
        const int js = as_const( B.cv_start )[i];
        if ( js >= 0 && cv_store )
            cv_features( i, js );
    };

    auto udl_value = [&]( int u ) __attribute__( ( always_inline ) ) -> R
    {
        if ( u >= 0 )
            return spot_of( __builtin_amdgcn_readfirstlane( u ) );
        const int r = -1 - u;  // rainbow_node.cpp:14-27
        const bool best = B.rb_best[r] != 0;
        R x = best ? R( -INFINITY ) : R( INFINITY );
        const int n = B.rb_n[r];
        for ( int j = 0; j < n; j++ )
        {
            const R perf = spot_of( __builtin_amdgcn_readfirstlane( B.rb_member[r][j] ) ) * R( B.rb_weight[r][j] );
            x = best ? ( x < perf ? perf : x ) : ( perf < x ? perf : x );
        }
        return x;
    };

    auto observe = [&]( int i ) __attribute__( ( always_inline ) )
    {
        uint32_t mon = as_const( B.mon_mask )[i];
        while ( mon )
        {
            const int c = __builtin_ctz( mon );
            mon &= mon - 1;
            const R s = udl_value( B.c_udl[c] );
            const R H = R( B.c_level[c] );
            const bool h = B.c_up[c] ? ( s >= H ) : ( s <= H );  // barrier_flow_node.cpp:35-37
            if ( h )
                hit |= 1u << c;
        }
        uint32_t fl = as_const( B.flow_mask )[i];
        while ( fl )
        {
            const int c = __builtin_ctz( fl );
            fl &= fl - 1;
            const R s = udl_value( B.c_udl[c] );
            const R K = R( B.c_strike[c] );
            R vanilla = B.c_call[c] ? ( s - K ) : ( K - s );  // finance.cpp:9-32
            vanilla = vanilla < R( 0 ) ? R( 0 ) : vanilla;
            R flow = vanilla;
            if ( B.c_kind[c] == 1 )
            {
                const bool h = ( hit >> c ) & 1u;
                const bool alive = B.c_in[c] ? h : !h;
                flow = alive ? vanilla : R( 0 );
            }
            double x = (double)flow * B.c_df[c];  // contract_node.cpp:33-47
            if ( !valid )
                x = 0.0;
            book += x;
            const double s1 = wave_sum<double>( x );
            const double s2 = wave_sum<double>( x * x );
            if ( lane == 0 )
            {
                acc[wib][1 + c][0] += s1;
                acc[wib][1 + c][1] += s2;
            }
        }
    };

    // vanilla grid at maturity j: call and put payoffs of every strike, summed
    // over the block's 256 paths (FP64 kernels in FP64 from the first
    // addition, FP32 kernels in FP64 past groups of sixteen), one partial
    // per (replicate, Sobol block, maturity, cell).  Per wave, 16 cells at a
    // time go lane -> LDS [cell][lane]; lane l then adds a quarter row
    // (cell l/4, lanes 16 (l%4) ..) and two shuffles finish the row.
    // one grid (underlying value s, strikes kf / kd) at maturity j into the
    // partial slot `out`
    auto grid_observe_one = [&]( R s, const float* kf, const double* kd, double* out ) __attribute__( ( always_inline ) )
    {
        const int nk = B.grid_nk, ncell = 2 * nk;
        R* tr = reinterpret_cast<R*>( gl ) + wib * kGridRows * 64;
        double* stage = reinterpret_cast<double*>( gl + kWavesPerBlock * kGridRows * 64 * sizeof( R ) );
        // a lane past the book's paths pays nothing: its calls see a zero
        // spot, its puts an infinite one
        const R s_call = valid ? s : R( 0 ), s_put = valid ? s : R( INFINITY );
        for ( int c0 = 0; c0 < ncell; c0 += kGridRows )
        {
            // cell c0 + r: the call (r even; c0 is even) or the put of strike
            // (c0 + r) / 2, read by scalar loads; rows past the grid repeat
            // the last strike and are never staged
#pragma unroll
            for ( int r = 0; r < kGridRows; r++ )
            {
                const int k = min( ( c0 + r ) >> 1, nk - 1 );
                R K;
                if constexpr ( F64 )
                    K = as_const( kd )[k];
                else
                    K = as_const( kf )[k];
                const R x = ( r & 1 ) ? K - s_put : s_call - K;  // finance.cpp:9-32
                tr[r * 64 + lane] = x < R( 0 ) ? R( 0 ) : x;
            }
            wave_lds_sync();
            const int row = lane >> 2, q = lane & 3;
            double s1, s2;
            if constexpr ( F64 )
            {
                s1 = s2 = 0.0;
#pragma unroll
                for ( int t = 0; t < 16; t++ )
                {
                    const double v = tr[row * 64 + q * 16 + t];
                    s1 += v;
                    s2 = fma( v, v, s2 );
                }
            }
            else
            {
                // sixteen payoffs in FP32 (relative error below 1e-6), then FP64
                float f1 = 0.0f, f2 = 0.0f;
#pragma unroll
                for ( int t = 0; t < 16; t++ )
                {
                    const float v = tr[row * 64 + q * 16 + t];
                    f1 += v;
                    f2 = fmaf( v, v, f2 );
                }
                s1 = f1;
                s2 = f2;
            }
            s1 += __shfl_xor( s1, 1, 64 );
            s2 += __shfl_xor( s2, 1, 64 );
            s1 += __shfl_xor( s1, 2, 64 );
            s2 += __shfl_xor( s2, 2, 64 );
            if ( q == 0 && c0 + row < ncell )
            {
                stage[( wib * kGridMaxCells + c0 + row ) * 2] = s1;
                stage[( wib * kGridMaxCells + c0 + row ) * 2 + 1] = s2;
            }
            wave_lds_sync();
        }
        __syncthreads();
        for ( int t = threadIdx.x; t < 2 * ncell; t += kBlock )
        {
            const int c = t >> 1, mom = t & 1;
            double v = stage[c * 2 + mom];
            for ( int w = 1; w < kWavesPerBlock; w++ )
                v += stage[( w * kGridMaxCells + c ) * 2 + mom];
            out[t] = v;
        }
        __syncthreads();
    };
    // the book's grid at maturity j; grouped books: every equity's grid, the
    // slot holding [equity][maturity][cell][2]
    auto grid_observe = [&]( int j ) __attribute__( ( always_inline ) )
    {
        const int ncell = 2 * B.grid_nk;
        if constexpr ( grouped )
        {
#pragma unroll
            for ( int a = 0; a < NA; a++ )
                grid_observe_one( spot_of( a ), B.grid_kf + a * B.grid_nk, B.grid_kd + a * B.grid_nk,
                                  B.grid_part + ( ( ( (size_t)z * ntiles + tile ) * NA + a ) * B.grid_nm + j ) *
                                                    ncell * 2 );
        }
        else
            grid_observe_one( udl_value( B.grid_udl ), B.grid_kf, B.grid_kd,
                              B.grid_part + ( ( (size_t)z * ntiles + tile ) * B.grid_nm + j ) * ncell * 2 );
    };

    // path recording of an American contract's equity (path_recorder; the
    // generic kernels only): its spot at dates 0 .. rec_n - 1, [date][path]
    [[maybe_unused]] const size_t rrow = valid ? (size_t)( idx - first_idx ) : 0;
    auto record = [&]( int i ) __attribute__( ( always_inline ) )
    {
        if constexpr ( !uniform )
            if ( i < B.rec_n && valid )
            {
                R x = R( 0 );
#pragma unroll
                for ( int a = 0; a < NA; a++ )
                    if ( a == B.rec_udl )
                        x = S[a];
                if constexpr ( F64 )
                    B.rec[(size_t)i * B.rec_stride + rrow] = x;
                else
                    B.recf[(size_t)i * B.rec_stride + rrow] = x;
            }
    };
    record( 0 );
    observe( 0 );
    if constexpr ( cv_on )
        cv_date( 0 );

    // ---- bridge step p: build the pending points of step p (in-order tree
    // walk), then pop W(t_{p+1}) as the normalised increment.  It depends on
    // the Sobol points only, so it runs one step ahead of the equities and the
    // scheduler overlaps it with their (path-dependent) arithmetic.
    // the normals of build index bidx for every factor.  Uniform FP32
    // kernels evaluate the central polynomial for all factors straight-line
    // and redo the lanes in the tail (w >= 8) in one block afterwards.
    auto normals = [&]( int o, R* zf ) __attribute__( ( always_inline ) )
    {
        if constexpr ( uniform )
        {
            uint32_t tails = 0;
#pragma unroll
            for ( int f = 0; f < NF; f++ )
            {
                bool t;
                zf[f] = normal_f32_central( sob_op( o, f ), t );
                tails |= t ? 1u << f : 0u;
            }
            if ( __builtin_amdgcn_ballot_w64( tails != 0 ) )
            {
#pragma unroll
                for ( int f = 0; f < NF; f++ )
                    if ( ( tails >> f ) & 1u )
                        zf[f] = normal_f32( sob_op( o, f ) );
            }
        }
        else
        {
            [[maybe_unused]] const int d0 = B.op_b[o] * F;
#pragma unroll
            for ( int f = 0; f < NF; f++ )
                if ( f < F )
                {
                    if constexpr ( prng )
                        if ( d0 + f >= kSobolDimsDev )
                        {
                            // dimension past the Sobol budget (path_generator.cpp:150-154)
                            const size_t at = (size_t)( d0 + f - kSobolDimsDev ) * B.tail_stride + prow;
                            if constexpr ( F64 )
                                zf[f] = B.tail[at];
                            else
                                zf[f] = B.tailf[at];
                            continue;
                        }
                    zf[f] = normal_of<R>( sob_op( o, f ) );
                }
        }
    };
    auto bridge_step = [&]( int p, facv& nz ) __attribute__( ( always_inline ) )
    {
        if constexpr ( prng )
            if ( pseudo )
            {
                // no bridge: the step's normals in factor order, straight
                // from the stream (noise_node.cpp:20-25)
#pragma unroll
                for ( int f = 0; f < NFV; f++ )
                    nz[f] = R( 0 );
#pragma unroll
                for ( int f = 0; f < NF; f++ )
                    if ( f < F )
                        nz[f] = xoshiro_normal<R>( prs );
                return;
            }
        const int i = p + 1;
        const int o0 = B.step_ops[p], o1 = B.step_ops[p + 1];
        if ( o1 - o0 == 1 && sp >= 1 )
        {
            // one build then one pop: the pushed right anchor would be popped
            // straight back, so the new point is consumed without touching the
            // stack (same arithmetic as the general walk below)
            const int o = o0;
            // FP32 left weights are 0 where the op has no left anchor
            const bool left = !F64 || as_const( B.op_left )[o] != 0;
            R lw, rw, sd;
            if constexpr ( F64 )
            {
                lw = B.op_w[3 * o];
                rw = B.op_w[3 * o + 1];
                sd = B.op_w[3 * o + 2];
            }
            else
            {
                lw = as_const( B.op_wf )[3 * o];
                rw = as_const( B.op_wf )[3 * o + 1];
                sd = as_const( B.op_wf )[3 * o + 2];
            }
#pragma unroll
            for ( int f = 0; f < NFV; f++ )
                nz[f] = R( 0 );
            R zf[NF];
            normals( o, zf );
#pragma unroll
            for ( int f = 0; f < NF; f++ )
            {
                if ( f < F )
                {
                    const R z = zf[f];
                    const R w = left ? lw * prev[f] + rw * top[f] + sd * z : rw * top[f] + sd * z;
                    if constexpr ( F64 )
                        nz[f] = ( B.bsq[i] > 0 ) ? ( w - prev[f] ) / B.bsq[i] : 0.0;
                    else
                        nz[f] = ( w - prev[f] ) * as_const( B.binv )[i];
                    prev[f] = w;
                }
            }
            return;
        }
        for ( int o = o0; o < o1; o++ )
        {
            // FP32 left weights are 0 where the op has no left anchor
            const bool left = !F64 || as_const( B.op_left )[o] != 0;
            const int pos = __builtin_amdgcn_readfirstlane( sp - 1 );
            R lw, rw, sd;
            if constexpr ( F64 )
            {
                lw = B.op_w[3 * o];
                rw = B.op_w[3 * o + 1];
                sd = B.op_w[3 * o + 2];
            }
            else
            {
                lw = as_const( B.op_wf )[3 * o];
                rw = as_const( B.op_wf )[3 * o + 1];
                sd = as_const( B.op_wf )[3 * o + 2];
            }
            R zf[NF];
            normals( o, zf );
#pragma unroll
            for ( int f = 0; f < NF; f++ )
            {
                if ( f < F )
                {
                    const R z = zf[f];
                    // path_generator.cpp:175-182
                    const R nv = left ? lw * prev[f] + rw * top[f] + sd * z : rw * top[f] + sd * z;
                    stk_set( f, pos, top[f] );
                    top[f] = nv;
                }
            }
            sp++;
        }
        // ---- pop W at this date -> normalized increment (path_generator.cpp:188-196)
        {
            const int pos = __builtin_amdgcn_readfirstlane( sp >= 2 ? sp - 2 : 0 );
#pragma unroll
            for ( int f = 0; f < NFV; f++ )
                nz[f] = R( 0 );
#pragma unroll
            for ( int f = 0; f < NF; f++ )
            {
                if ( f < F )
                {
                    const R w = top[f];
                    if constexpr ( F64 )
                        nz[f] = ( B.bsq[i] > 0 ) ? ( w - prev[f] ) / B.bsq[i] : 0.0;
                    else
                        nz[f] = ( w - prev[f] ) * as_const( B.binv )[i];
                    prev[f] = w;
                    // sp = 1 only at the last pop (W at the final date, the
                    // bridge's first point, is pending until then), after
                    // which top is never read: no select on sp
                    top[f] = stk_get( f, pos );
                }
            }
            sp--;
        }
    };

    facv nz;
    if ( m >= 1 )
        bridge_step( 0, nz );
    for ( int p = 0; p < m; p++ )
    {
        const int i = p + 1;
        if ( ( i - 1 ) % G == 0 )
        {
            // the group's leverage layers and Sobol tables are visible, the
            // previous group's slots are free: stage the next group
            __syncthreads();
            if ( lev_lds && i + G <= m )
                stage_group( i + G );
            stage_ops( so( i + G ), so( i + 2 * G ) );
        }
        // ---- uniform FP32 kernels: the QE variance draws of every equity,
        // quadratic branch straight-line, the exponential-branch lanes redone
        // in one block (qe_f32's arithmetic either way)
        [[maybe_unused]] float vq[NA];
        [[maybe_unused]] QeDraw dq[NA];
        if constexpr ( qe_phase )
        {
            uint32_t expo = 0;
#pragma unroll
            for ( int a = 0; a < NA; a++ )
            {
                cptr<float> sc = as_const( B.sc ) + ( (size_t)i * nas + a ) * SC_COUNT;
                const int vf = grouped ? 1 : uniform ? NA + a : __builtin_amdgcn_readfirstlane( B.asset[a].var_factor );
                const QeMoments q = qe_moments( Var[a], sc );
                vq[a] = qe_quadratic( q, nz[vf], cv_on ? &dq[a] : nullptr );
                expo |= q.quad ? 0u : 1u << a;
            }
            if ( __builtin_amdgcn_ballot_w64( expo != 0 ) )
            {
#pragma unroll
                for ( int a = 0; a < NA; a++ )
                    if ( ( expo >> a ) & 1u )
                    {
                        cptr<float> sc = as_const( B.sc ) + ( (size_t)i * nas + a ) * SC_COUNT;
                        const int vf = grouped ? 1 : uniform ? NA + a : __builtin_amdgcn_readfirstlane( B.asset[a].var_factor );
                        vq[a] = qe_exponential( qe_moments( Var[a], sc ), nz[vf], cv_on ? &dq[a] : nullptr );
                    }
            }
        }
        // ---- equities
#pragma unroll
        for ( int a = 0; a < NA; a++ )
        {
            if constexpr ( !uniform )
                if ( a >= na )
                    continue;
            const int model = lsv_kind ? 2 : KIND == 2 ? 1 : B.asset[a].model;
            // correlated noise (correlated_noise_node.cpp:15-27, lower triangle, name order)
            R eps = nz[grouped ? 0 : a];
            // the factor of the step into date i (term-structured correlation:
            // one per date, correlation.cpp:860-903), read by scalar loads
            const size_t crow = ( (size_t)i * kMaxAssets + a ) * kMaxAssets;
            if constexpr ( corr_kind )
            {
                // every spot correlated
                eps = R( 0 );
#pragma unroll
                for ( int v = 0; v <= a; v++ )
                    eps += nz[v] * as_const( B.cholf_t )[crow + v];
            }
            else if ( !uniform && na > 1 && ( model == 0 || B.corr_stoch ) )
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
            cptr<float> sc = as_const( B.sc ) + ( (size_t)i * nas + a ) * SC_COUNT;
            R cd;
            if constexpr ( F64 )
                cd = B.drift[(size_t)a * nd + i] * B.t[i] - B.drift[(size_t)a * nd + i - 1] * B.t[i - 1];
            else
                cd = sc[SC_CARRY];
            if ( model == 0 )
            {
                // brownian_node.cpp:12-24 + spot_diffusion_node.cpp:21-65
                if constexpr ( F64 )
                {
                    const double w1 = W[a] + B.sqrt_dt[i] * eps;
                    const double dw = w1 - W[a];
                    W[a] = w1;
                    const double v = B.asset[a].vol;
                    const double expo = cd - v * v / 2 * B.dt[i] + v * dw;
                    S[a] = ( S[a] - 0 ) * exp( expo ) + 0;
                }
                else
                {
                    const float w1 = W[a] + sc[SC_SQDT] * eps;
                    const float dw = w1 - W[a];
                    W[a] = w1;
                    const float expo = cd - sc[SC_HALF_V2DT] + sc[SC_VOL] * dw;
                    const float sold = S[a];
                    S[a] = sold * __expf( expo );
                    if constexpr ( cv_on )
                    {
                        // first-order Doob increment S_{i-1} vol dW: zero conditional mean
                        const float M = sold * sc[SC_VOL] * dw;
                        if constexpr ( cv_infer )
                            cvacc = fmaf( hr[a] * as_const( B.cv_scale )[(size_t)i * nas + a], M, cvacc );
                        else
                            dmacc[a] = fmaf( as_const( B.cv_scale )[(size_t)i * nas + a], M, dmacc[a] );
                    }
                }
                continue;
            }
            // Heston / LSV (heston_variance_node.cpp, heston_spot_node.cpp)
            const int vf = grouped ? 1 : uniform ? NA + a : __builtin_amdgcn_readfirstlane( B.asset[a].var_factor );
            const R zv = nz[vf];
            const R vp = Var[a];
            R vc;
            [[maybe_unused]] QeDraw draw;
            if constexpr ( F64 )
                vc = qe_f64( vp, B.dt[i], zv, B.asset[a].kappa, B.asset[a].theta, B.asset[a].xi );
            else if constexpr ( qe_phase )
            {
                vc = vq[a];
                if constexpr ( cv_on )
                    draw = dq[a];
            }
            else if constexpr ( cv_on )
                vc = qe_f32( vp, zv, sc, &draw );
            else
                vc = qe_f32( vp, zv, sc );
            Var[a] = vc;
            R L = R( 1 );
            if ( model == 2 )
            {
                // local_volatility_node.cpp:19-47: layer i at the previous spot
                const size_t li = (size_t)a * nd + i;
                const int last = B.lev_points - 1;
                R f;
                if constexpr ( F64 )
                    f = log( S[a] ) / B.lev_ln_step[li] - B.lev_offset[li];
                else
                    f = ( log_spot ? S[a] : fast_log( S[a] ) ) * as_const( B.lev_inv_step )[(size_t)i * nas + a] -
                        as_const( B.lev_offset_f )[(size_t)i * nas + a];
                if constexpr ( F64 )
                {
                    const double* lv = B.lev + li * B.lev_points;
                    if ( f <= 0.0 )
                        L = lv[0];
                    else if ( f >= (double)last )
                        L = lv[last];
                    else
                    {
                        const int k = (int)f;
                        const double w = f - (double)k;
                        L = ( 1.0 - w ) * lv[k] + w * lv[k + 1];
                    }
                }
                else
                {
                    const float* lv = levs + ( a * 2 * G + ( i - 1 ) % ( 2 * G ) ) * P;
                    // clamped to [0, last) so that k <= last - 1: past the last
                    // point w = 1 - 2^-24 last rather than 1
                    const float fc = __builtin_amdgcn_fmed3f( f, 0.0f, lev_top );
                    const int k = (int)fc;
                    const float w = fc - (float)k;
                    L = fmaf( w, lv[k + 1] - lv[k], lv[k] );
                }
            }
            const R z = eps;
            R step;
            if constexpr ( F64 )
            {
                const double xi = B.asset[a].xi, rho = B.asset[a].rho, kappa = B.asset[a].kappa,
                             theta = B.asset[a].theta, dt = B.dt[i];
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
                    step = cd + K0 + K1 * vp + K2 * vc + sqrt( var ) * z;
                }
                else
                {
                    step = cd - 0.5 * L * L * vp * dt + L * sqrt( vp * dt ) * z;
                }
                S[a] = S[a] * exp( step );
            }
            else
            {
                float det, bvol;
                if ( uniform || B.asset[a].xi > 1e-12 )
                {
                    // K0 + K1 vp + K2 vc with K1,2 = L A1 - L^2 A2 -+ L RX, grouped as
                    // L (K0 + A1 (vp + vc) + RX (vc - vp)) - L^2 A2 (vp + vc)
                    const float sv = vp + vc;
                    const float u = fmaf( sc[SC_RX], vc - vp, fmaf( sc[SC_A1], sv, sc[SC_K0] ) );
                    const float l2s = L * L * sv;
                    det = fmaf( -sc[SC_A2], l2s, fmaf( L, u, cd ) );
                    bvol = __builtin_amdgcn_sqrtf( fmaxf( sc[SC_K3] * l2s, 0.0f ) );
                }
                else
                {
                    det = cd - 0.5f * L * L * vp * sc[SC_DT];
                    bvol = L * __builtin_amdgcn_sqrtf( vp * sc[SC_DT] );
                }
                step = det + bvol * z;
                float sold;
                if constexpr ( log_spot )
                {
                    sold = cv_on ? __expf( S[a] ) : 0.0f;
                    S[a] += step;
                }
                else
                {
                    sold = S[a];
                    S[a] = sold * __expf( step );
                }
                if constexpr ( cv_on )
                {
                    // first-order Doob increment of the spot.  Given v_i, the step
                    // is lognormal with E[S_i | v_i, past] = S_{i-1} e^{c + A v_i},
                    // A = K2 + L^2 K3 / 2 = L SC_CVA - L^2 SC_CVB; its linear part
                    //   M = S_{i-1} (A (v_i - E[v_i | past]) + b z)
                    // has zero conditional mean exactly: the draw's own mean is
                    // subtracted and z is independent of v_i and of the past.
                    float M = bvol * z;
                    if ( uniform || B.asset[a].xi > 1e-12 )
                    {
                        const float A = L * fmaf( -L, sc[SC_CVB], sc[SC_CVA] );
                        M = fmaf( A, vc - draw.mean, M );
                    }
                    M *= sold;
                    if constexpr ( cv_infer )
                        cvacc = fmaf( hr[a] * as_const( B.cv_scale )[(size_t)i * nas + a], M, cvacc );
                    else
                        dmacc[a] = fmaf( as_const( B.cv_scale )[(size_t)i * nas + a], M, dmacc[a] );
                }
            }
        }
        record( i );
        observe( i );
        if constexpr ( CV == 0 )
            if ( B.grid_nm > 0 && as_const( B.grid_mat_at )[i] >= 0 )
                grid_observe( as_const( B.grid_mat_at )[i] );
        if constexpr ( cv_on )
            cv_date( i );
        if ( i < m )
            bridge_step( i, nz );  // the increments of the next step
    }
    if constexpr ( cv_samples )
        if ( cv_store )
            B.cv_pay[(size_t)z * B.cv_N + pidx] = (float)book;
    if constexpr ( cv_infer )
    {
        const double y = valid ? book - (double)cvacc : 0.0;
        const double s1 = wave_sum<double>( y );
        const double s2 = wave_sum<double>( y * y );
        if ( lane == 0 )
        {
            acc[wib][1 + nc][0] = s1;
            acc[wib][1 + nc][1] = s2;
        }
    }

    // ---- book value reduction
    if ( !valid )
        book = 0.0;
    {
        const double s1 = wave_sum<double>( book );
        const double s2 = wave_sum<double>( book * book );
        if ( lane == 0 )
        {
            acc[wib][0][0] = s1;
            acc[wib][0][1] = s2;
        }
    }
    __syncthreads();
    const int ne = ( 1 + nc + ( cv_infer ? 1 : 0 ) ) * 2;
    if ( (int)threadIdx.x < ne )
    {
        const int e = threadIdx.x >> 1, mom = threadIdx.x & 1;
        double s = 0.0;
        for ( int w = 0; w < kWavesPerBlock; w++ )
            s += acc[w][e][mom];
        B.partial[( (size_t)z * ntiles + tile ) * ne + threadIdx.x] = s;
    }
}

// one Sobol block per workgroup: grid (blocks, books, replicates)
template <class R, int NA, int DEPTH, int CV = 0, int KIND = 0>
__global__ void __launch_bounds__( kBlock ) __attribute__( ( amdgpu_waves_per_eu( CV >= 2 ? 2 : 1 ) ) )
sweep_kernel( const DevBook* __restrict__ books, const uint32_t* __restrict__ V, uint32_t bq0, uint32_t first_idx,
              uint32_t last_idx, const uint32_t* __restrict__ shifts )
{
    sweep_block<R, NA, DEPTH, CV, KIND>( books[blockIdx.y], V, bq0 + blockIdx.x, first_idx, last_idx, shifts, blockIdx.z,
                                   blockIdx.x, gridDim.x );
}

// Persistent form: a fixed grid of workgroups pulls (book, Sobol block) work
// items from a queue (items sorted longest book first by the host; the
// counter is zeroed by a kernel of the same graph).  Every item writes its
// own partial-sum slot, so the result does not depend on which workgroup
// took which item.
// grouped grid books (kind 5) are held to 128 VGPRs: four waves per SIMD ran
// the vanilla grids 8 % faster than three at the unconstrained 145
template <class R, int NA, int DEPTH, int KIND = 0>
__global__ void __launch_bounds__( kBlock )
__attribute__( ( amdgpu_waves_per_eu( KIND == 5 ? 4 : 1 ) ) )
sweep_persistent( const DevBook* __restrict__ books, const uint32_t* __restrict__ V, uint32_t bq0, uint32_t first_idx,
                  uint32_t last_idx, const int2* __restrict__ items, int n_items, uint32_t ntiles,
                  int* __restrict__ counter )
{
    __shared__ int item_s;
    for ( ;; )
    {
        if ( threadIdx.x == 0 )
            item_s = atomicAdd( counter, 1 );
        __syncthreads();
        const int it = item_s;
        __syncthreads();  // every thread has its item before thread 0 takes the next
        if ( it >= n_items )
            break;
        const int2 w = items[it];
        sweep_block<R, NA, DEPTH, 0, KIND>( books[w.x], V, bq0 + (uint32_t)w.y, first_idx, last_idx, nullptr, 0,
                                      (uint32_t)w.y, ntiles );
    }
}

}  // namespace thoth
