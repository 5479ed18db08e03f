// mcl_sweep_waves.hip -- instantiations of the latency form of the FP32
// uniform-LSV sweep (mcl_sweep_waves.hpp): equities per book {2, 3, 4, 5, 8},
// bridge stack depth {8, 10}, plain sweeps and control-variate samples (at
// most five equities).
#include "mcl_sweep_waves.hpp"

namespace thoth
{
namespace
{
template <int NA, int D, int CV>
void launch_w( dim3 grid, hipStream_t s, size_t lds, const DevBook* books, const uint32_t* V, uint32_t bq0,
               uint32_t first, uint32_t last, const uint32_t* shifts )
{
    hipLaunchKernelGGL( ( sweep_waves<NA, D, CV> ), grid, dim3( 64 * NA ), lds, s, books, V, bq0, first, last,
                        shifts );
}

template <int NA, int CV>
bool by_depth( int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books, const uint32_t* V,
               uint32_t bq0, uint32_t first, uint32_t last, const uint32_t* shifts )
{
    if ( depth <= 8 )
        launch_w<NA, 8, CV>( grid, s, lds, books, V, bq0, first, last, shifts );
    else if ( depth <= 10 )
        launch_w<NA, 10, CV>( grid, s, lds, books, V, bq0, first, last, shifts );
    else
        return false;
    return true;
}

template <int NA>
bool by_cv( int cv_mode, int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books, const uint32_t* V,
            uint32_t bq0, uint32_t first, uint32_t last, const uint32_t* shifts )
{
    if ( cv_mode == 0 )
        return by_depth<NA, 0>( depth, grid, s, lds, books, V, bq0, first, last, shifts );
    if constexpr ( NA <= 5 )
        if ( cv_mode == 1 )
            return by_depth<NA, 1>( depth, grid, s, lds, books, V, bq0, first, last, shifts );
    return false;
}
}  // namespace

bool sweep_waves_available( int na, int depth, int cv_mode )
{
    const bool na_ok = na == 2 || na == 3 || na == 4 || na == 5 || na == 8;
    return na_ok && depth <= 10 && ( cv_mode == 0 || ( cv_mode == 1 && na <= 5 ) );
}

bool launch_sweep_waves( int na, int depth, int cv_mode, dim3 grid, hipStream_t s, size_t lds, const DevBook* books,
                         const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last, const uint32_t* shifts )
{
    switch ( na )
    {
    case 2:
        return by_cv<2>( cv_mode, depth, grid, s, lds, books, V, bq0, first, last, shifts );
    case 3:
        return by_cv<3>( cv_mode, depth, grid, s, lds, books, V, bq0, first, last, shifts );
    case 4:
        return by_cv<4>( cv_mode, depth, grid, s, lds, books, V, bq0, first, last, shifts );
    case 5:
        return by_cv<5>( cv_mode, depth, grid, s, lds, books, V, bq0, first, last, shifts );
    case 8:
        return by_cv<8>( cv_mode, depth, grid, s, lds, books, V, bq0, first, last, shifts );
    default:
        return false;
    }
}
}  // namespace thoth
