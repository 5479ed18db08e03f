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
}  // namespace thoth
