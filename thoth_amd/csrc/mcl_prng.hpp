// mcl_prng.hpp -- the reference's pseudo-random stream on the device.
//
// The reference draws every normal that the Sobol budget does not cover from
// one xoshiro256++ stream per pricing (helpers/rng.hpp:36-72, seeded once per
// PriceBook, tasks/pricer_mcl.cpp:227): the dimensions of a path past the
// 4097 Sobol dimensions (nodes/path_generator.cpp:140-157), or, with
// use_sobol false, every normal of the path.  Path p therefore starts at
// position p T of the stream, T the draws per path.  xoshiro's state update
// is linear over GF(2), s_{n+1} = M s_n, so the state at p T is (M^T)^p s_0:
// the host tabulates J_k = (M^T)^(2^k) as 256 x 256 bit matrices and a lane
// applies the J_k of the set bits of p (at most log2(paths) matrix-vector
// products, each 256 conditional XORs of four words).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mcl_device.hpp"

namespace thoth
{

struct Xoshiro256
{
    uint64_t s[4];

    __device__ __forceinline__ static uint64_t rotl( uint64_t x, int k ) { return ( x << k ) | ( x >> ( 64 - k ) ); }

    // next 64 bits (rng.hpp:51-62)
    __device__ __forceinline__ uint64_t next()
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

// s <- A s over GF(2).  Column j of A (the image of state bit j: bit j % 64 of
// word j / 64) is A[4 j .. 4 j + 3].  The column addresses are wave-uniform,
// so the four words come in through the scalar cache; each lane masks them
// by its own state bit.
__device__ __forceinline__ void gf2_apply( const uint64_t* __restrict__ A, uint64_t s[4] )
{
    uint32_t r[8] = { 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u };
    const uint32_t* a = reinterpret_cast<const uint32_t*>( A );
#pragma unroll
    for ( int h = 0; h < 8; h++ )
    {
        const uint32_t word = h & 1 ? (uint32_t)( s[h >> 1] >> 32 ) : (uint32_t)s[h >> 1];
#pragma unroll 8
        for ( int b = 0; b < 32; b++ )
        {
            const uint32_t msk = 0u - ( ( word >> b ) & 1u );
            const uint32_t* c = a + ( h * 32 + b ) * 8;
#pragma unroll
            for ( int k = 0; k < 8; k++ )
                r[k] = __builtin_amdgcn_bitop3_b32( c[k], r[k], msk, 0x6c );  // r ^= c & msk
        }
    }
#pragma unroll
    for ( int q = 0; q < 4; q++ )
        s[q] = (uint64_t)r[2 * q] | ( (uint64_t)r[2 * q + 1] << 32 );
}

// advance the lane's state by n T draws: the product of the tabulated J_k over
// the set bits k of n (levels: tabulated matrices, n < 2^levels).  A level no
// lane of the wave needs is skipped as a whole.
__device__ __forceinline__ void xoshiro_jump( Xoshiro256& x, const uint64_t* __restrict__ tables, int levels,
                                              uint64_t n )
{
    for ( int k = 0; k < levels; k++ )
    {
        const bool bit = ( n >> k ) & 1u;
        if ( __builtin_amdgcn_ballot_w64( bit ) )
        {
            uint64_t t[4] = { x.s[0], x.s[1], x.s[2], x.s[3] };
            gf2_apply( tables + (size_t)k * kJumpWords, t );
            if ( bit )
            {
#pragma unroll
                for ( int q = 0; q < 4; q++ )
                    x.s[q] = t[q];
            }
        }
    }
}

// Acklam's inverse normal without the Sobol clamp (distributions.cpp:37-81),
// FP64; the argument is in (0, 1)
__device__ __forceinline__ double acklam_f64( double P )
{
    const double a0 = -3.969683028665376e+01, a1 = 2.209460984245205e+02, a2 = -2.759285104469687e+02,
                 a3 = 1.383577518672690e+02, a4 = -3.066479806614716e+01, a5 = 2.506628277459239e+00;
    const double b0 = -5.447609879822406e+01, b1 = 1.615858368580409e+02, b2 = -1.556989798598866e+02,
                 b3 = 6.680131188771972e+01, b4 = -1.328068155288572e+01;
    const double c0 = -7.784894002430293e-03, c1 = -3.223964580411365e-01, c2 = -2.400758277161838e+00,
                 c3 = -2.549732539343734e+00, c4 = 4.374664141464968e+00, c5 = 2.938163982698783e+00;
    const double d0 = 7.784695709041462e-03, d1 = 3.224671290700398e-01, d2 = 2.445134137142996e+00,
                 d3 = 3.754408661907416e+00;
    const double p_low = 0.02425;
    const double p_high = 1.0 - p_low;
    if ( P < p_low )
    {
        double q = sqrt( -2.0 * log( P ) );
        return ( ( ( ( ( c0 * q + c1 ) * q + c2 ) * q + c3 ) * q + c4 ) * q + c5 ) /
               ( ( ( ( d0 * q + d1 ) * q + d2 ) * q + d3 ) * q + 1.0 );
    }
    if ( P <= p_high )
    {
        double q = P - 0.5;
        double r = q * q;
        return ( ( ( ( ( a0 * r + a1 ) * r + a2 ) * r + a3 ) * r + a4 ) * r + a5 ) * q /
               ( ( ( ( ( b0 * r + b1 ) * r + b2 ) * r + b3 ) * r + b4 ) * r + 1.0 );
    }
    double q = sqrt( -2.0 * log( 1.0 - P ) );
    return -( ( ( ( ( c0 * q + c1 ) * q + c2 ) * q + c3 ) * q + c4 ) * q + c5 ) /
           ( ( ( ( d0 * q + d1 ) * q + d2 ) * q + d3 ) * q + 1.0 );
}

// Rng::Gaussian (rng.hpp:64-72): the 53 high bits as a uniform in (0, 1)
// (0 read as 2^-53), then the inverse normal
__device__ __forceinline__ double xoshiro_gauss_f64( Xoshiro256& x )
{
    const double u = (double)( x.next() >> 11 ) * ( 1.0 / 9007199254740992.0 );
    return acklam_f64( u > 0.0 ? u : 1.0 / 9007199254740992.0 );
}

}  // namespace thoth
