// mcl_prng.cpp -- host side of the device's pseudo-random stream: the
// splitmix64 seeding of xoshiro256++ (helpers/rng.hpp:36-47) and the GF(2)
// jump matrices that put path p at position p T of the stream (mcl_prng.hpp).
//
// xoshiro256++'s state update (rng.hpp:51-62) is XORs, a shift and a rotation
// of the 256-bit state: a linear map s -> M s over GF(2).  Column j of M is the
// image of the state with only bit j set.  (M^T)^(2^k) follow by repeated
// squaring of 256 x 256 bit matrices (each product: 256 columns, each the XOR
// of the columns of the left factor selected by the bits of one column of the
// right factor), and are cached by T.
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

#include "mcl_internal.hpp"

namespace thoth
{

namespace
{

typedef std::vector<uint64_t> Gf2;  // 256 columns x 4 words

void StepState( uint64_t s[4] )
{
    auto rotl = []( uint64_t x, int k ) { return ( x << k ) | ( x >> ( 64 - k ) ); };
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl( s[3], 45 );
}

void Apply( const Gf2& A, const uint64_t* x, uint64_t* out )
{
    uint64_t r[4] = { 0, 0, 0, 0 };
    for ( int j = 0; j < 256; j++ )
        if ( ( x[j >> 6] >> ( j & 63 ) ) & 1u )
            for ( int q = 0; q < 4; q++ )
                r[q] ^= A[4 * j + q];
    std::memcpy( out, r, sizeof( r ) );
}

Gf2 Mul( const Gf2& A, const Gf2& B )
{
    Gf2 C( kJumpWords );
    for ( int j = 0; j < 256; j++ )
        Apply( A, &B[4 * j], &C[4 * j] );
    return C;
}

Gf2 Transition()
{
    Gf2 M( kJumpWords );
    for ( int j = 0; j < 256; j++ )
    {
        uint64_t s[4] = { 0, 0, 0, 0 };
        s[j >> 6] = uint64_t( 1 ) << ( j & 63 );
        StepState( s );
        std::memcpy( &M[4 * j], s, sizeof( s ) );
    }
    return M;
}

Gf2 Power( uint64_t n )
{
    Gf2 R( kJumpWords, 0 ), base = Transition();
    for ( int j = 0; j < 256; j++ )
        R[4 * j + ( j >> 6 )] = uint64_t( 1 ) << ( j & 63 );
    while ( n )
    {
        if ( n & 1u )
            R = Mul( R, base );
        n >>= 1;
        if ( n )
            base = Mul( base, base );
    }
    return R;
}

}  // namespace

void SplitMixSeed( uint64_t seed, uint64_t st[4] )
{
    uint64_t z = seed;
    for ( int i = 0; i < 4; i++ )
    {
        z += 0x9e3779b97f4a7c15ULL;
        uint64_t x = z;
        x = ( x ^ ( x >> 30 ) ) * 0xbf58476d1ce4e5b9ULL;
        x = ( x ^ ( x >> 27 ) ) * 0x94d049bb133111ebULL;
        st[i] = x ^ ( x >> 31 );
    }
}

const std::vector<uint64_t>& JumpTables( uint64_t T, int levels )
{
    // callers hold the library lock
    static std::map<uint64_t, std::vector<uint64_t>> cache;
    std::vector<uint64_t>& tab = cache[T];
    size_t have = tab.size() / kJumpWords;
    if ( have == 0 )
    {
        const Gf2 J = Power( T );
        tab.assign( J.begin(), J.end() );
        have = 1;
    }
    while ( (int)have < levels )
    {
        const Gf2 last( tab.end() - kJumpWords, tab.end() );
        const Gf2 sq = Mul( last, last );
        tab.insert( tab.end(), sq.begin(), sq.end() );
        have++;
    }
    return tab;
}

int JumpLevels( int64_t paths )
{
    int levels = 1;
    while ( levels < 63 && ( uint64_t( 1 ) << levels ) < (uint64_t)paths )
        levels++;
    return levels;
}

}  // namespace thoth

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_prng_state_after( uint64_t seed, uint64_t steps,
                                                                                   uint64_t draws_per_path,
                                                                                   uint64_t* state )
{
    if ( !state || draws_per_path == 0 || steps % draws_per_path != 0 )
        return thoth::SetError( THOTH_ERR_INVALID, "steps must be a multiple of draws_per_path >= 1" );
    std::lock_guard<std::mutex> lk( thoth::DeviceMutex() );
    const uint64_t p = steps / draws_per_path;
    int levels = 1;
    while ( levels < 63 && ( p >> levels ) )
        levels++;
    const std::vector<uint64_t>& tab = thoth::JumpTables( draws_per_path, levels );
    uint64_t s[4];
    thoth::SplitMixSeed( seed, s );
    for ( int k = 0; k < levels; k++ )
        if ( ( p >> k ) & 1u )
        {
            const std::vector<uint64_t> A( tab.begin() + (size_t)k * thoth::kJumpWords,
                                           tab.begin() + (size_t)( k + 1 ) * thoth::kJumpWords );
            thoth::Apply( A, s, s );
        }
    std::memcpy( state, s, sizeof( s ) );
    return THOTH_OK;
}
