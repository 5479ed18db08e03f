// mcl_prng_fill.hpp -- the normals of the dimensions past the Sobol budget,
// drawn once per sweep into [T][paths] (path minor) so that the bridge walk,
// which needs them out of stream order, reads them like any other input.
#pragma once
#include <hip/hip_runtime.h>

#include "mcl_device.hpp"
#include "mcl_sweep.hpp"

namespace thoth
{

// one lane per path p: jump to position p T of the stream, then the path's T
// draws in stream order (dimension order, path_generator.cpp:140-157)
template <class R>
__global__ void __launch_bounds__( 256 ) tail_fill_kernel( const uint64_t* __restrict__ jump, int levels, uint64_t s0,
                                                           uint64_t s1, uint64_t s2, uint64_t s3, int T, int64_t paths,
                                                           R* __restrict__ out )
{
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    Xoshiro256 x;
    x.s[0] = s0;
    x.s[1] = s1;
    x.s[2] = s2;
    x.s[3] = s3;
    // every lane takes part in the jump (its levels are chosen by wave vote)
    xoshiro_jump( x, jump, levels, p < paths ? (uint64_t)p : 0 );
    if ( p >= paths )
        return;
    for ( int t = 0; t < T; t++ )
        out[(size_t)t * paths + p] = xoshiro_normal<R>( x );
}

template <class R>
bool launch_tail_fill( hipStream_t s, const uint64_t* jump, int levels, const uint64_t* seed_state, int T,
                       int64_t paths, R* out )
{
    if ( T < 1 || paths < 1 || levels < 1 || ( levels < 63 && ( int64_t( 1 ) << levels ) < paths ) )
        return false;
    const int64_t blocks = ( paths + 255 ) / 256;
    hipLaunchKernelGGL( ( tail_fill_kernel<R> ), dim3( (unsigned)blocks ), dim3( 256 ), 0, s, jump, levels,
                        seed_state[0], seed_state[1], seed_state[2], seed_state[3], T, paths, out );
    return hipGetLastError() == hipSuccess;
}

}  // namespace thoth
