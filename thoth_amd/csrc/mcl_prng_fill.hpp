// mcl_prng_fill.hpp -- the normals of the dimensions past the Sobol budget,
// drawn once per sweep into [T][paths] (path minor) so that the bridge walk,
// which needs them out of stream order, reads them like any other input.
#pragma once
#include <hip/hip_runtime.h>

#include "mcl_device.hpp"
#include "mcl_sweep.hpp"

namespace thoth
{

// The stream state at the start of every path, [4][paths]: thread i takes
// `chunk` consecutive paths, jumps to the first one with the tabulated
// matrices, then reaches each next path either by T sequential steps (T
// small: a matrix application costs about as much as 200 steps) or by one
// application of J_0 = M^T.  One state kernel serves every book of the batch
// that shares T (the sweeps of use_sobol false read their lanes' states, the
// tail fill below draws from them).
static __global__ void __launch_bounds__( 256 ) prng_states_kernel( const uint64_t* __restrict__ jump,
                                                                    int levels,
                                                                    uint64_t s0, uint64_t s1, uint64_t s2,
                                                                    uint64_t s3, int T, int64_t paths, int chunk,
                                                                    int by_matrix, uint64_t* __restrict__ states )
{
    const int64_t c0 = ( (int64_t)blockIdx.x * 256 + threadIdx.x ) * chunk;
    Xoshiro256 x;
    x.s[0] = s0;
    x.s[1] = s1;
    x.s[2] = s2;
    x.s[3] = s3;
    // every lane takes part in the jump (its levels are chosen by wave vote)
    xoshiro_jump( x, jump, levels, c0 < paths ? (uint64_t)c0 : 0 );
    if ( c0 >= paths )
        return;
    for ( int k = 0; k < chunk && c0 + k < paths; k++ )
    {
        const int64_t p = c0 + k;
#pragma unroll
        for ( int q = 0; q < 4; q++ )
            states[(size_t)q * paths + p] = x.s[q];
        if ( by_matrix )
            gf2_apply( jump, x.s );
        else
            for ( int t = 0; t < T; t++ )
                (void)x.next();
    }
}

// the normals of the dimensions past the Sobol table: one lane per path
// draws its T normals in stream order (dimension order,
// path_generator.cpp:140-157) from its start state, [T][paths]
template <class R>
__global__ void __launch_bounds__( 256 ) tail_fill_kernel( const uint64_t* __restrict__ states, int T, int64_t paths,
                                                           R* __restrict__ out )
{
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if ( p >= paths )
        return;
    Xoshiro256 x;
#pragma unroll
    for ( int q = 0; q < 4; q++ )
        x.s[q] = states[(size_t)q * paths + p];
    for ( int t = 0; t < T; t++ )
        out[(size_t)t * paths + p] = xoshiro_normal<R>( x );
}

inline int prng_chunk( int64_t paths )
{
    const int64_t c = paths / 16384;
    return (int)( c < 1 ? 1 : c > 4096 ? 4096 : c );
}

inline bool launch_prng_states( hipStream_t s, const uint64_t* jump, int levels, const uint64_t* seed_state, int T,
                                int64_t paths, uint64_t* states )
{
    if ( T < 1 || paths < 1 || levels < 1 || ( levels < 63 && ( int64_t( 1 ) << levels ) < paths ) )
        return false;
    const int chunk = prng_chunk( paths );
    const int64_t threads = ( paths + chunk - 1 ) / chunk;
    const int64_t blocks = ( threads + 255 ) / 256;
    hipLaunchKernelGGL( prng_states_kernel, dim3( (unsigned)blocks ), dim3( 256 ), 0, s, jump, levels, seed_state[0],
                        seed_state[1], seed_state[2], seed_state[3], T, paths, chunk, T > 200 ? 1 : 0, states );
    return hipGetLastError() == hipSuccess;
}

template <class R>
bool launch_tail_fill( hipStream_t s, const uint64_t* states, int T, int64_t paths, R* out )
{
    if ( T < 1 || paths < 1 )
        return false;
    const int64_t blocks = ( paths + 255 ) / 256;
    hipLaunchKernelGGL( ( tail_fill_kernel<R> ), dim3( (unsigned)blocks ), dim3( 256 ), 0, s, states, T, paths, out );
    return hipGetLastError() == hipSuccess;
}

}  // namespace thoth
