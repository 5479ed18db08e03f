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
    sob.set_lane( threadIdx.x >> 6, threadIdx.x & 63 );
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

// The FP32 special functions of the production sweep evaluated on given
// inputs, one lane per input, so that their error bounds can be checked on the
// device itself: fn 0 normal_f32 of the 32-bit Sobol coordinate in[i]; fn 1
// the split form of the uniform-book kernels (central polynomial for every
// lane, the tail lanes redone by normal_f32); fn 2 erfc_fast of the float
// whose bits are in[i].
__global__ void __launch_bounds__( 256 ) special_f32_kernel( int fn, int64_t n, const uint32_t* __restrict__ in,
                                                             float* __restrict__ out )
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if ( i >= n )
        return;
    const uint32_t x = in[i];
    float r;
    if ( fn == 0 )
        r = normal_f32( x );
    else if ( fn == 1 )
    {
        bool tail;
        r = normal_f32_central( x, tail );
        if ( tail )
            r = normal_f32( x );
    }
    else
        r = erfc_fast( __uint_as_float( x ) );
    out[i] = r;
}

bool launch_special_f32( hipStream_t s, int fn, int64_t n, const uint32_t* in, float* out )
{
    hipLaunchKernelGGL( special_f32_kernel, dim3( (uint32_t)( ( n + 255 ) / 256 ) ), dim3( 256 ), 0, s, fn, n, in,
                        out );
    return hipGetLastError() == hipSuccess;
}

// zero the persistent sweep's work counter (a kernel, so that the graph that
// replays the sweep holds no memset node)
__global__ void counter_reset_kernel( int* counter ) { *counter = 0; }

void launch_counter_reset( hipStream_t s, int* counter )
{
    hipLaunchKernelGGL( counter_reset_kernel, dim3( 1 ), dim3( 1 ), 0, s, counter );
}

// out[e] = Kahan sum over the slots, in slot order, of part[slot * per_slot + e]
__global__ void slot_sum_kernel( const double* __restrict__ part, int nslots, int per_slot, double* __restrict__ out )
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if ( e >= per_slot )
        return;
    double s = 0.0, c = 0.0;
    for ( int k = 0; k < nslots; k++ )
    {
        const double y = part[(size_t)k * per_slot + e] - c;
        const double t = s + y;
        c = ( t - s ) - y;
        s = t;
    }
    out[e] = s;
}

// Grid sums of every book in one launch: block (x, book) owns 32 consecutive
// elements; its 8 thread groups Kahan-sum the slots g, g + 8, ... and the 8
// group sums are combined in group order (fixed partition: deterministic).
// table[b] = { partial offset, output offset, elements per slot } (doubles).
__global__ void __launch_bounds__( 256 ) grid_sum_kernel( const double* __restrict__ part,
                                                         const long long* __restrict__ table, int nslots,
                                                         double* __restrict__ out )
{
    __shared__ double ps[8][32], pc[8][32];
    const long long* tb = table + 3 * blockIdx.y;
    const int per_slot = (int)tb[2];
    const int e = blockIdx.x * 32 + ( threadIdx.x & 31 );
    const int g = threadIdx.x >> 5;
    double sum = 0.0, c = 0.0;
    if ( e < per_slot )
    {
        const double* p = part + tb[0] + e;
        for ( int k = g; k < nslots; k += 8 )
        {
            const double y = p[(size_t)k * per_slot] - c;
            const double t = sum + y;
            c = ( t - sum ) - y;
            sum = t;
        }
    }
    ps[g][threadIdx.x & 31] = sum;
    pc[g][threadIdx.x & 31] = c;
    __syncthreads();
    if ( g == 0 && e < per_slot )
    {
        for ( int h = 1; h < 8; h++ )
        {
            const double y = ps[h][threadIdx.x] - ( c + pc[h][threadIdx.x] );
            const double t = sum + y;
            c = ( t - sum ) - y;
            sum = t;
        }
        out[tb[1] + e] = sum;
    }
}

void launch_grid_sum( hipStream_t s, const double* part, const long long* table, int n_books, int max_per_slot,
                      int nslots, double* out )
{
    hipLaunchKernelGGL( grid_sum_kernel, dim3( ( max_per_slot + 31 ) / 32, n_books ), dim3( 256 ), 0, s, part, table,
                        nslots, out );
}

void launch_slot_sum( hipStream_t s, const double* part, int nslots, int per_slot, double* out )
{
    hipLaunchKernelGGL( slot_sum_kernel, dim3( ( per_slot + 255 ) / 256 ), dim3( 256 ), 0, s, part, nslots, per_slot,
                        out );
}

}  // namespace thoth
