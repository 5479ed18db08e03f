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

// the same sweep with the learned control variate (at most 5 equities):
// CV = 1 writes the training samples, CV = 2 (SiLU) and 3 (smoothstep gate)
// run the hedge network in the loop
template <int CV>
struct CvLaunch
{
    template <int NA, int D>
    static void one( dim3 grid, hipStream_t s, size_t lds, const DevBook* books, const uint32_t* V, uint32_t bq0,
                     uint32_t first, uint32_t last, const uint32_t* shifts )
    {
        if constexpr ( NA <= 5 )
            hipLaunchKernelGGL( ( sweep_kernel<float, NA, D, CV> ), grid, dim3( kBlock ), lds, s, books, V, bq0, first,
                                last, shifts );
    }
};

bool launch_sweep_f32( int na, int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books,
                       const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last, const uint32_t* shifts,
                       int cv_mode )
{
    if ( cv_mode != 0 && na > 5 )
        return false;
    if ( cv_mode == 1 )
        THOTH_DISPATCH_SWEEP( CvLaunch<1>::one, na, depth, grid, s, lds, books, V, bq0, first, last, shifts );
    if ( cv_mode == 2 )
        THOTH_DISPATCH_SWEEP( CvLaunch<2>::one, na, depth, grid, s, lds, books, V, bq0, first, last, shifts );
    if ( cv_mode == 3 )
        THOTH_DISPATCH_SWEEP( CvLaunch<3>::one, na, depth, grid, s, lds, books, V, bq0, first, last, shifts );
    THOTH_DISPATCH_SWEEP( launch_one, na, depth, grid, s, lds, books, V, bq0, first, last, shifts );
}
template <int NA, int D>
static void launch_persistent_one( dim3 grid, hipStream_t s, size_t lds, const DevBook* books, const uint32_t* V,
                                   uint32_t bq0, uint32_t first, uint32_t last, const int2* items, int n_items,
                                   uint32_t ntiles, int* counter )
{
    hipLaunchKernelGGL( ( sweep_persistent<float, NA, D> ), grid, dim3( kBlock ), lds, s, books, V, bq0, first, last,
                        items, n_items, ntiles, counter );
}

bool launch_persistent_f32( int na, int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books,
                           const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last, const int2* items,
                           int n_items, uint32_t ntiles, int* counter )
{
    THOTH_DISPATCH_SWEEP( launch_persistent_one, na, depth, grid, s, lds, books, V, bq0, first, last, items, n_items,
                          ntiles, counter );
}
}  // namespace thoth
