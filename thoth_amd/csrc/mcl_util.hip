// mcl_util.hip -- raw Sobol points through exactly the integer stage the path
// sweep uses (block table + wave/lane parts), exported as thoth_gpu_sobol_points.
#include <hip/hip_runtime.h>

#include "mcl_sweep.hpp"

namespace thoth
{

// out[(idx - first) * dim + d] = integer coordinate d of Sobol point idx for
// idx in [first, last]; one lane per index, blocks aligned on 256 indices.
__global__ void __launch_bounds__( kBlock ) sobol_points_kernel( const uint32_t* __restrict__ V, int dim,
                                                                uint32_t bq0, uint32_t first, uint32_t last,
                                                                uint32_t* __restrict__ out )
{
    extern __shared__ uint32_t tab[];
    const uint32_t bq = bq0 + blockIdx.x;
    const uint32_t idx = bq * 256u + (uint32_t)threadIdx.x;
    sobol_block_table( tab, V, dim, bq, nullptr );
    __syncthreads();
    SobolLane sob;
    sob.V = V;
    sob.tab = tab;
    sob.gw = __builtin_amdgcn_readfirstlane( gray32( 64u * ( threadIdx.x >> 6 ) ) );
    const uint32_t g = gray32( threadIdx.x & 63u );
    for ( int k = 0; k < 6; k++ )
        sob.m[k] = ( ( g >> k ) & 1u ) ? 0xffffffffu : 0u;
    const bool valid = idx >= first && idx <= last;
    for ( int d = 0; d < dim; d++ )
    {
        const uint32_t x = sob( d );
        if ( valid )
            out[(size_t)( idx - first ) * dim + d] = x;
    }
}

bool launch_sobol_points( hipStream_t s, const uint32_t* V, int dim, uint32_t first, uint32_t last, uint32_t* out )
{
    const uint32_t bq0 = first / 256;
    const uint32_t nb = last / 256 - bq0 + 1;
    hipLaunchKernelGGL( sobol_points_kernel, dim3( nb ), dim3( kBlock ), dim * sizeof( uint32_t ), s, V, dim, bq0, first,
                        last, out );
    return hipGetLastError() == hipSuccess;
}

}  // namespace thoth
