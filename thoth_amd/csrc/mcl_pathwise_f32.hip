// mcl_pathwise_f32.hip -- FP32 instantiations of the pathwise-Greeks kernel
// (mcl_pathwise.hpp), with contraction on like the FP32 sweep.
#include "mcl_pathwise.hpp"

namespace thoth
{
namespace
{
template <int D>
void launch_pw( dim3 grid, hipStream_t s, size_t lds, const DevBook* books, const uint32_t* V, uint32_t bq0,
                uint32_t first, uint32_t last )
{
    hipLaunchKernelGGL( ( pathwise_kernel<float, D> ), grid, dim3( kBlock ), lds, s, books, V, bq0, first, last );
}
template <int NA, int D, bool LSV>
void launch_pwb( dim3 grid, hipStream_t s, size_t lds, const DevBook* books, const uint32_t* V, uint32_t bq0,
                 uint32_t first, uint32_t last )
{
    hipLaunchKernelGGL( ( pathwise_basket_kernel<float, NA, D, LSV> ), grid, dim3( kBlock ), lds, s, books, V, bq0,
                        first, last );
}
template <int NA, bool LSV = false>
bool launch_pwb_depth( int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books, const uint32_t* V,
                       uint32_t bq0, uint32_t first, uint32_t last )
{
    if ( depth <= 8 )
        launch_pwb<NA, 8, LSV>( grid, s, lds, books, V, bq0, first, last );
    else if ( depth <= 10 )
        launch_pwb<NA, 10, LSV>( grid, s, lds, books, V, bq0, first, last );
    else if ( depth <= 16 && !LSV )
        launch_pwb<NA, 16, LSV>( grid, s, lds, books, V, bq0, first, last );
    else
        return false;
    return true;
}
}  // namespace

bool launch_pathwise_f32( int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books, const uint32_t* V,
                          uint32_t bq0, uint32_t first, uint32_t last )
{
    if ( depth <= 8 )
        launch_pw<8>( grid, s, lds, books, V, bq0, first, last );
    else if ( depth <= 10 )
        launch_pw<10>( grid, s, lds, books, V, bq0, first, last );
    else if ( depth <= 16 )
        launch_pw<16>( grid, s, lds, books, V, bq0, first, last );
    else
        return false;
    return true;
}

bool pathwise_basket_lsv_available( int na, int depth ) { return na == 5 && depth <= 10; }

bool launch_pathwise_basket_f32( int na, int depth, bool lsv, dim3 grid, hipStream_t s, size_t lds,
                                 const DevBook* books, const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last )
{
    if ( lsv )
        return pathwise_basket_lsv_available( na, depth ) &&
               launch_pwb_depth<5, true>( depth, grid, s, lds, books, V, bq0, first, last );
    if ( na <= 1 )
        return launch_pwb_depth<1>( depth, grid, s, lds, books, V, bq0, first, last );
    if ( na <= 2 )
        return launch_pwb_depth<2>( depth, grid, s, lds, books, V, bq0, first, last );
    if ( na <= 5 )
        return launch_pwb_depth<5>( depth, grid, s, lds, books, V, bq0, first, last );
    if ( na <= 8 )
        return launch_pwb_depth<8>( depth, grid, s, lds, books, V, bq0, first, last );
    return false;
}
}  // namespace thoth
