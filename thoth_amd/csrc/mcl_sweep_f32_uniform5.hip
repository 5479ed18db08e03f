// mcl_sweep_f32_uniform5.hip -- the FP32 uniform-book sweep instantiations
// of five-equity LSV books (the basket workload, uncorrelated as the
// reference diffuses it, or with Cholesky-correlated spots), in a unit of their own so
// that they are scheduled ILP-first (-mllvm -amdgpu-sched-strategy=max-ilp):
// at two waves per SIMD either way the five equities' independent steps
// interleave better (12.07 -> 11.79 ms on the basket), while the one-equity
// kernels lose occupancy under that scheduler and keep the default.
#include "mcl_sweep_uniform_launch.hpp"

namespace thoth
{
bool launch_sweep_f32_uniform5( int kind, int cv_mode, int depth, dim3 grid, hipStream_t s, size_t lds,
                                const DevBook* books, const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last,
                                const uint32_t* shifts )
{
    if ( kind == 3 )  // correlated spots: the plain sweep only
        return uniform_launch::by_cv<5, 3>( cv_mode, depth, grid, s, lds, books, V, bq0, first, last, shifts );
    return uniform_launch::by_cv<5, 1>( cv_mode, depth, grid, s, lds, books, V, bq0, first, last, shifts );
}

bool launch_persistent_f32_uniform5( int kind, int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books,
                                     const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last, const int2* items,
                                     int n_items, uint32_t ntiles, int* counter )
{
    if ( depth > 10 )
        return false;
    if ( kind == 5 )
    {
        if ( depth <= 8 )
            hipLaunchKernelGGL( ( sweep_persistent<float, 5, 8, 5> ), grid, dim3( kBlock ), lds, s, books, V, bq0,
                                first, last, items, n_items, ntiles, counter );
        else
            hipLaunchKernelGGL( ( sweep_persistent<float, 5, 10, 5> ), grid, dim3( kBlock ), lds, s, books, V, bq0,
                                first, last, items, n_items, ntiles, counter );
        return true;
    }
    if ( kind == 3 )
    {
        if ( depth <= 8 )
            hipLaunchKernelGGL( ( sweep_persistent<float, 5, 8, 3> ), grid, dim3( kBlock ), lds, s, books, V, bq0,
                                first, last, items, n_items, ntiles, counter );
        else
            hipLaunchKernelGGL( ( sweep_persistent<float, 5, 10, 3> ), grid, dim3( kBlock ), lds, s, books, V, bq0,
                                first, last, items, n_items, ntiles, counter );
    }
    else if ( depth <= 8 )
        hipLaunchKernelGGL( ( sweep_persistent<float, 5, 8, 1> ), grid, dim3( kBlock ), lds, s, books, V, bq0, first,
                            last, items, n_items, ntiles, counter );
    else
        hipLaunchKernelGGL( ( sweep_persistent<float, 5, 10, 1> ), grid, dim3( kBlock ), lds, s, books, V, bq0, first,
                            last, items, n_items, ntiles, counter );
    return true;
}
}  // namespace thoth
