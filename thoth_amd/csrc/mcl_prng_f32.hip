// mcl_prng_f32.hip -- FP32 sweeps of books whose normals are not all Sobol
// points (kind 4 of sweep_block) and the FP32 fill of their tail normals.
#include "mcl_launch.hpp"
#include "mcl_prng_fill.hpp"
#include "mcl_sweep.hpp"

namespace thoth
{
template <int NA, int D>
static void launch_one( dim3 grid, hipStream_t s, size_t lds, const DevBook* books, const uint32_t* V, uint32_t bq0,
                        uint32_t first, uint32_t last )
{
    hipLaunchKernelGGL( ( sweep_kernel<float, NA, D, 0, 4> ), grid, dim3( kBlock ), lds, s, books, V, bq0, first,
                        last, nullptr );
}

bool launch_sweep_prng_f32( int na, int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books,
                            const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last )
{
    THOTH_DISPATCH_SWEEP( launch_one, na, depth, grid, s, lds, books, V, bq0, first, last );
}

bool launch_tail_fill_f32( hipStream_t s, const uint64_t* states, int T, int64_t paths, float* out )
{
    return launch_tail_fill<float>( s, states, T, paths, out );
}
}  // namespace thoth
