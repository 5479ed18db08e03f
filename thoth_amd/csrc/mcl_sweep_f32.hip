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

// the same sweep writing the control-variate samples (hedge features and
// increments per hedge date, discounted payoff per path); at most 5 equities
template <int NA, int D>
static void launch_one_cv( dim3 grid, hipStream_t s, size_t lds, const DevBook* books, const uint32_t* V, uint32_t bq0,
                           uint32_t first, uint32_t last, const uint32_t* shifts )
{
    if constexpr ( NA <= 5 )
        hipLaunchKernelGGL( ( sweep_kernel<float, NA, D, true> ), grid, dim3( kBlock ), lds, s, books, V, bq0, first,
                            last, shifts );
}

bool launch_sweep_f32( int na, int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books,
                       const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last, const uint32_t* shifts, bool cv )
{
    if ( cv )
    {
        if ( na > 5 )
            return false;
        THOTH_DISPATCH_SWEEP( launch_one_cv, na, depth, grid, s, lds, books, V, bq0, first, last, shifts );
    }
    THOTH_DISPATCH_SWEEP( launch_one, na, depth, grid, s, lds, books, V, bq0, first, last, shifts );
}
}  // namespace thoth
