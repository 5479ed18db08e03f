// mcl_prng_f64.hip -- FP64 sweeps of books whose normals are not all Sobol
// points (kind 4 of sweep_block) and the FP64 fill of their tail normals.
// Built with -ffp-contract=off, as the other oracle-matched FP64 kernels.
#include "mcl_launch.hpp"
#include "mcl_prng_fill.hpp"
#include "mcl_sweep.hpp"

namespace thoth
{
template <int NA, int D>
static void launch_one( dim3 grid, hipStream_t s, size_t lds, const DevBook* books, const uint32_t* V, uint32_t bq0,
                        uint32_t first, uint32_t last )
{
    hipLaunchKernelGGL( ( sweep_kernel<double, NA, D, 0, 4> ), grid, dim3( kBlock ), lds, s, books, V, bq0, first,
                        last, nullptr );
}

bool launch_sweep_prng_f64( int na, int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books,
                            const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last )
{
    THOTH_DISPATCH_SWEEP( launch_one, na, depth, grid, s, lds, books, V, bq0, first, last );
}

bool launch_tail_fill_f64( hipStream_t s, const uint64_t* states, int T, int64_t paths, double* out )
{
    return launch_tail_fill<double>( s, states, T, paths, out );
}

bool launch_prng_start_states( hipStream_t s, const uint64_t* jump, int levels, const uint64_t* seed_state, int T,
                               int64_t paths, uint64_t* states )
{
    return launch_prng_states( s, jump, levels, seed_state, T, paths, states );
}
}  // namespace thoth
