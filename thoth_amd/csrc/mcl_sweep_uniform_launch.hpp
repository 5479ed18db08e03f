// mcl_sweep_uniform_launch.hpp -- launch helpers of the FP32 uniform-book
// sweep instantiations, shared by the translation units that hold them
// (mcl_sweep_f32_uniform.hip: one-equity books; mcl_sweep_f32_uniform5.hip:
// the five-equity books, built with the ILP-first scheduler).
#pragma once
#include "mcl_launch.hpp"
#include "mcl_sweep.hpp"

namespace thoth
{
namespace uniform_launch
{
template <int NA, int D, int CV, int KIND>
inline void one( dim3 grid, hipStream_t s, size_t lds, const DevBook* books, const uint32_t* V, uint32_t bq0, uint32_t first,
          uint32_t last, const uint32_t* shifts )
{
    hipLaunchKernelGGL( ( sweep_kernel<float, NA, D, CV, KIND> ), grid, dim3( kBlock ), lds, s, books, V, bq0, first,
                        last, shifts );
}

template <int NA, int CV, int KIND>
inline bool by_depth( int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books, const uint32_t* V, uint32_t bq0,
               uint32_t first, uint32_t last, const uint32_t* shifts )
{
    // depth 10 only without the control variate: its 5-equity CV kernels
    // would exceed the register budget of two waves per SIMD
    if ( depth <= 8 )
        one<NA, 8, CV, KIND>( grid, s, lds, books, V, bq0, first, last, shifts );
    else if constexpr ( CV == 0 )
    {
        if ( depth > 10 )
            return false;
        one<NA, 10, CV, KIND>( grid, s, lds, books, V, bq0, first, last, shifts );
    }
    else
        return false;
    return true;
}

template <int NA, int KIND>
inline bool by_cv( int cv_mode, int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books, const uint32_t* V,
            uint32_t bq0, uint32_t first, uint32_t last, const uint32_t* shifts )
{
    if ( cv_mode == 0 )
        return by_depth<NA, 0, KIND>( depth, grid, s, lds, books, V, bq0, first, last, shifts );
    if constexpr ( KIND == 1 )
    {
        if ( cv_mode == 1 )
            return by_depth<NA, 1, KIND>( depth, grid, s, lds, books, V, bq0, first, last, shifts );
        if ( cv_mode == 2 )
            return by_depth<NA, 2, KIND>( depth, grid, s, lds, books, V, bq0, first, last, shifts );
        if ( cv_mode == 3 )
            return by_depth<NA, 3, KIND>( depth, grid, s, lds, books, V, bq0, first, last, shifts );
    }
    return false;
}
}  // namespace uniform_launch
}  // namespace thoth
