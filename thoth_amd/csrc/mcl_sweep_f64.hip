// mcl_sweep_f64.hip -- FP64 (oracle-matched) instantiations of the path-sweep
// kernel.  Built with -ffp-contract=off so that every expression rounds as the
// reference's Release build does (no fused multiply-adds).
#include "mcl_sweep.hpp"
#include "mcl_launch.hpp"

namespace thoth
{
template <int NA, int D>
static void launch_one( dim3 grid, hipStream_t s, size_t lds, const DevBook* books, const uint32_t* V, uint32_t bq0,
                        uint32_t first, uint32_t last, const uint32_t* shifts )
{
    hipLaunchKernelGGL( ( sweep_kernel<double, NA, D> ), grid, dim3( kBlock ), lds, s, books, V, bq0, first, last, shifts );
}

bool launch_sweep_f64( int na, int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books,
                       const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last, const uint32_t* shifts )
{
    THOTH_DISPATCH_SWEEP( launch_one, na, depth, grid, s, lds, books, V, bq0, first, last, shifts );
}
template <int NA, int D>
static void launch_persistent_one( dim3 grid, hipStream_t s, size_t lds, const DevBook* books, const uint32_t* V,
                                   uint32_t bq0, uint32_t first, uint32_t last, const int2* items, int n_items,
                                   uint32_t ntiles, int* counter )
{
    hipLaunchKernelGGL( ( sweep_persistent<double, NA, D> ), grid, dim3( kBlock ), lds, s, books, V, bq0, first, last,
                        items, n_items, ntiles, counter );
}

bool launch_persistent_f64( int na, int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books,
                           const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last, const int2* items,
                           int n_items, uint32_t ntiles, int* counter )
{
    THOTH_DISPATCH_SWEEP( launch_persistent_one, na, depth, grid, s, lds, books, V, bq0, first, last, items, n_items,
                          ntiles, counter );
}
}  // namespace thoth
