// mcl_launch.hpp -- dispatch from (equities, bridge-stack depth) to the
// template instantiation of the path-sweep kernel.
#pragma once
#include <hip/hip_runtime.h>

#include "mcl_device.hpp"

// supported instantiations: equities per book {1, 2, 5, 8}, stack depth {2, 8, 10, 16}
#define THOTH_DISPATCH_DEPTH( FN, NA, depth, ... )     \
    do                                                 \
    {                                                  \
        if ( ( depth ) <= 2 )                          \
            FN<NA, 2>( __VA_ARGS__ );                  \
        else if ( ( depth ) <= 8 )                     \
            FN<NA, 8>( __VA_ARGS__ );                  \
        else if ( ( depth ) <= 10 )                    \
            FN<NA, 10>( __VA_ARGS__ );                 \
        else if ( ( depth ) <= 16 )                    \
            FN<NA, 16>( __VA_ARGS__ );                 \
        else                                           \
            return false;                              \
        return true;                                   \
    } while ( 0 )

#define THOTH_DISPATCH_SWEEP( FN, na, depth, ... )                     \
    do                                                                 \
    {                                                                  \
        if ( ( na ) <= 1 )                                             \
            THOTH_DISPATCH_DEPTH( FN, 1, depth, __VA_ARGS__ );         \
        else if ( ( na ) <= 2 )                                        \
            THOTH_DISPATCH_DEPTH( FN, 2, depth, __VA_ARGS__ );         \
        else if ( ( na ) <= 5 )                                        \
            THOTH_DISPATCH_DEPTH( FN, 5, depth, __VA_ARGS__ );         \
        else if ( ( na ) <= 8 )                                        \
            THOTH_DISPATCH_DEPTH( FN, 8, depth, __VA_ARGS__ );         \
        return false;                                                  \
    } while ( 0 )

namespace thoth
{
constexpr int kMaxStackDepth = 16;
bool launch_sweep_f64( int na, int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books,
                       const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last, const uint32_t* shifts );
bool launch_sweep_f32( int na, int depth, dim3 grid, hipStream_t s, size_t lds, const DevBook* books,
                       const uint32_t* V, uint32_t bq0, uint32_t first, uint32_t last, const uint32_t* shifts );
bool launch_sobol_points( hipStream_t s, const uint32_t* V, int dim, uint32_t first, uint32_t last, uint32_t* out );
}  // namespace thoth
