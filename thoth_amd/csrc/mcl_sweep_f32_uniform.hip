// mcl_sweep_f32_uniform.hip -- FP32 path-sweep instantiations for uniform
// books: every equity LSV (KIND 1) or Heston (KIND 2) with vol-of-vol > 0 and
// uncorrelated stochastic spots, book size exactly the template's equities.
// The per-equity model branches of the generic kernel fold away (measured
// 21.8 -> 18.0 ms on the 5-equity LSV basket).  Shapes without an
// instantiation here run the generic kernel.
#include "mcl_launch.hpp"
#include "mcl_sweep.hpp"

namespace thoth
{
namespace
{
template <int NA, int D, int CV, int KIND>
void one( dim3 grid, hipStream_t s, size_t lds, const DevBook* books, const uint32_t* V, uint32_t bq0, uint32_t first,
          uint32_t last, const uint32_t* shifts )
{
    hipLaunchKernelGGL( ( sweep_kernel<float, NA, D, CV, KIND> ), grid, dim3( kBlock ), lds, s, books, V, bq0, first,
                        last, shifts );
}

template <int NA, int CV, int KIND>
bool by_depth( int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books, const uint32_t* V, uint32_t bq0,
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
bool by_cv( int cv_mode, int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books, const uint32_t* V,
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
    }
    return false;
}
}  // namespace

bool launch_sweep_f32_uniform( int kind, int na, int depth, dim3 grid, hipStream_t s, size_t lds,
                               const DevBook* books, const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last,
                               const uint32_t* shifts, int cv_mode )
{
    if ( kind == 1 && na == 1 )
        return by_cv<1, 1>( cv_mode, depth, grid, s, lds, books, V, bq0, first, last, shifts );
    if ( kind == 1 && na == 5 )
        return by_cv<5, 1>( cv_mode, depth, grid, s, lds, books, V, bq0, first, last, shifts );
    if ( kind == 2 && na == 1 )
        return by_cv<1, 2>( cv_mode, depth, grid, s, lds, books, V, bq0, first, last, shifts );
    return false;
}

bool launch_persistent_f32_uniform( int kind, int na, int depth, dim3 grid, hipStream_t s, size_t lds,
                                    const DevBook* books, const uint32_t* V, uint32_t bq0, uint32_t first,
                                    uint32_t last, const int2* items, int n_items, uint32_t ntiles, int* counter )
{
    if ( kind != 1 || na != 1 || depth > 10 )
        return false;
    if ( depth <= 8 )
        hipLaunchKernelGGL( ( sweep_persistent<float, 1, 8, 1> ), grid, dim3( kBlock ), lds, s, books, V, bq0, first,
                            last, items, n_items, ntiles, counter );
    else
        hipLaunchKernelGGL( ( sweep_persistent<float, 1, 10, 1> ), grid, dim3( kBlock ), lds, s, books, V, bq0, first,
                            last, items, n_items, ntiles, counter );
    return true;
}
}  // namespace thoth
