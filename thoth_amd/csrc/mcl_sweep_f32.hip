// mcl_sweep_f32.hip -- FP32-stepping instantiations of the path-sweep
// kernel (contraction on); the cross-path accumulation stays FP64.
#include "mcl_sweep.hpp"
#include "mcl_launch.hpp"

namespace thoth
{
template <int NA, int D>
static void launch_one( dim3 grid, hipStream_t s, const DevBook* books, const uint32_t* V, uint32_t q0, uint32_t nw,
                        uint32_t first, uint32_t last, const uint32_t* shifts )
{
    hipLaunchKernelGGL( ( sweep_kernel<float, NA, D> ), grid, dim3( kBlock ), 0, s, books, V, q0, nw, first, last, shifts );
}

bool launch_sweep_f32( int na, int depth, dim3 grid, hipStream_t s, const DevBook* books, const uint32_t* V,
                       uint32_t q0, uint32_t nw, uint32_t first, uint32_t last, const uint32_t* shifts )
{
    THOTH_DISPATCH_SWEEP( launch_one, na, depth, grid, s, books, V, q0, nw, first, last, shifts );
}
}  // namespace thoth
