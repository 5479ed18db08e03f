// mcl_util.hip -- raw Sobol points through the same wave-decomposed integer
// stage the path sweep uses (exported as thoth_gpu_sobol_points).
#include <hip/hip_runtime.h>

#include "mcl_sweep.hpp"

namespace thoth
{

// out[(idx - first) * dim + d] = integer coordinate d of Sobol point idx,
// for idx in [first, last]; one lane per point, waves aligned on 64 indices.
__global__ void __launch_bounds__( kBlock ) sobol_points_kernel( const uint32_t* __restrict__ V, int dim,
                                                                uint32_t q0, uint32_t first, uint32_t last,
                                                                uint32_t* __restrict__ out )
{
    const int lane = threadIdx.x & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane( blockIdx.x * kWavesPerBlock + ( threadIdx.x >> 6 ) );
    const uint32_t q = q0 + wave;
    const uint32_t idx = q * 64u + (uint32_t)lane;
    const uint32_t G = __builtin_amdgcn_readfirstlane( gray32( q * 64u ) );
    const uint32_t g = gray32( (uint32_t)lane );
    const bool valid = idx >= first && idx <= last;
    for ( int d = 0; d < dim; d++ )
    {
        const uint32_t x = sobol_x( V, d, G, g );
        if ( valid )
            out[(size_t)( idx - first ) * dim + d] = x;
    }
}

bool launch_sobol_points( hipStream_t s, const uint32_t* V, int dim, uint32_t first, uint32_t last, uint32_t* out )
{
    const uint32_t q0 = first / 64;
    const uint32_t nw = last / 64 - q0 + 1;
    const uint32_t nb = ( nw + kWavesPerBlock - 1 ) / kWavesPerBlock;
    hipLaunchKernelGGL( sobol_points_kernel, dim3( nb ), dim3( kBlock ), 0, s, V, dim, q0, first, last, out );
    return hipGetLastError() == hipSuccess;
}

}  // namespace thoth
