// mcl_sweep_f32.hip -- FP32-stepping instantiations of the path-sweep
// kernel (contraction on); the cross-path accumulation stays FP64.
#include "mcl_sweep.hpp"
#include "mcl_launch.hpp"

namespace thoth
{
template <int NA, int D>
static void launch_one( dim3 grid, hipStream_t s, size_t lds, const DevBook* books, const uint32_t* V, uint32_t bq0,
                        uint32_t first, uint32_t last, const uint32_t* shifts )
{
    hipLaunchKernelGGL( ( sweep_kernel<float, NA, D> ), grid, dim3( kBlock ), lds, s, books, V, bq0, first, last, shifts );
}

bool launch_sweep_f32( int na, int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books,
                       const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last, const uint32_t* shifts )
{
    THOTH_DISPATCH_SWEEP( launch_one, na, depth, grid, s, lds, books, V, bq0, first, last, shifts );
}
}  // namespace thoth
