// mcl_sweep_f32_uniform.hip -- FP32 path-sweep instantiations for uniform
// books: every equity LSV (KIND 1) or Heston (KIND 2) with vol-of-vol > 0 and
// uncorrelated stochastic spots, book size exactly the template's equities.
// The per-equity model branches of the generic kernel fold away (measured
// 21.8 -> 18.0 ms on the 5-equity LSV basket).  Shapes without an
// instantiation here run the generic kernel.
#include "mcl_sweep_uniform_launch.hpp"

namespace thoth
{
using uniform_launch::by_cv;

bool launch_sweep_f32_uniform( int kind, int na, int depth, dim3 grid, hipStream_t s, size_t lds,
                               const DevBook* books, const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last,
                               const uint32_t* shifts, int cv_mode )
{
    if ( kind == 1 && na == 1 )
        return by_cv<1, 1>( cv_mode, depth, grid, s, lds, books, V, bq0, first, last, shifts );
    if ( ( kind == 1 || kind == 3 ) && na == 5 )
        return launch_sweep_f32_uniform5( kind, cv_mode, depth, grid, s, lds, books, V, bq0, first, last, shifts );
    if ( kind == 2 && na == 1 )
        return by_cv<1, 2>( cv_mode, depth, grid, s, lds, books, V, bq0, first, last, shifts );
    return false;
}
template <int NA, int KIND>
static bool persistent_by_depth( int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books,
                                 const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last, const int2* items,
                                 int n_items, uint32_t ntiles, int* counter )
{
    if ( depth > 10 )
        return false;
    if ( depth <= 8 )
        hipLaunchKernelGGL( ( sweep_persistent<float, NA, 8, KIND> ), grid, dim3( kBlock ), lds, s, books, V, bq0,
                            first, last, items, n_items, ntiles, counter );
    else
        hipLaunchKernelGGL( ( sweep_persistent<float, NA, 10, KIND> ), grid, dim3( kBlock ), lds, s, books, V, bq0,
                            first, last, items, n_items, ntiles, counter );
    return true;
}

// the persistent (work-queue) form of the same instantiations: one-equity LSV
// (the vanilla grids) and Heston books here, five-equity LSV books in their own
// unit; kind 5 (grouped single-equity LSV grid books) for 2 and 5 books
bool launch_persistent_f32_uniform( int kind, int na, int depth, dim3 grid, hipStream_t s, size_t lds,
                                    const DevBook* books, const uint32_t* V, uint32_t bq0, uint32_t first,
                                    uint32_t last, const int2* items, int n_items, uint32_t ntiles, int* counter )
{
    if ( kind == 1 && na == 1 )
        return persistent_by_depth<1, 1>( depth, grid, s, lds, books, V, bq0, first, last, items, n_items, ntiles,
                                          counter );
    if ( kind == 2 && na == 1 )
        return persistent_by_depth<1, 2>( depth, grid, s, lds, books, V, bq0, first, last, items, n_items, ntiles,
                                          counter );
    if ( ( kind == 1 || kind == 3 || kind == 5 ) && na == 5 )
        return launch_persistent_f32_uniform5( kind, depth, grid, s, lds, books, V, bq0, first, last, items, n_items,
                                               ntiles, counter );
    if ( kind == 5 && na == 2 )  // grouped grid books, two per workgroup lane
        return persistent_by_depth<2, 5>( depth, grid, s, lds, books, V, bq0, first, last, items, n_items, ntiles,
                                          counter );
    return false;
}
}  // namespace thoth
