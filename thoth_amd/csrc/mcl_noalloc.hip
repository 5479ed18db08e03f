// mcl_noalloc.hip -- thoth_mcl_sweep_noalloc: the book sweep of
// thoth_mcl_sweep without any device allocation, for callers that cannot
// allocate (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  Everything the library would allocate for a sweep --
// the packed per-book arrays (grid, bridge program, step constants, leverage
// layers, payoff masks), the device books, the random digital shifts and the
// per-block partial sums -- is placed in one file-scope device array of
// kNoAllocScratchBytes, resident for the life of the process on every device
// the library runs on.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "mcl_internal.hpp"
#include "thoth_mcl_gpu.h"

namespace thoth
{
namespace
{
constexpr size_t kNoAllocScratchBytes = size_t( 64 ) << 20;  // 64 MiB
static_assert( kNoAllocScratchBytes == (size_t)THOTH_NOALLOC_SCRATCH_BYTES, "scratch size of the C-ABI" );

// the scratch array; in host code its name is a host-side stand-in, never a
// device address: host code uses the per-device address taken at set-up
static __device__ __align__( 256 ) char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
char* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device
}  // namespace

// At set-up (the library's device set-up, once per device, with that device
// current): the device address of this device's copy of the array.
int NoAllocSetup()
{
    int dev = 0;
    hipError_t e = hipGetDevice( &dev );
    if ( e != hipSuccess || dev < 0 || dev >= kMaxDevices )
        return SetError( THOTH_ERR_DEVICE, "no-allocation set-up: no current device" );
    void* p = nullptr;
    e = hipGetSymbolAddress( &p, HIP_SYMBOL( scratch ) );
    if ( e != hipSuccess || !p )
        return SetError( THOTH_ERR_DEVICE, std::string( "hipGetSymbolAddress(scratch): " ) + hipGetErrorString( e ) );
    g_scratch[dev] = static_cast<char*>( p );
    return THOTH_OK;
}

}  // namespace thoth

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_mcl_sweep_noalloc( int32_t n_books,
                                                                                   const thoth_mcl_book* books,
                                                                                   const thoth_mcl_sweep_cfg* cfg,
                                                                                   thoth_mcl_book_out* out )
{
    using namespace thoth;
    std::lock_guard<std::mutex> lk( DeviceMutex() );
    int rc = EnsureDevice();
    if ( rc != THOTH_OK )
        return rc;
    if ( !out )
        return SetError( THOTH_ERR_INVALID, "null argument" );
    int dev = 0;
    if ( hipGetDevice( &dev ) != hipSuccess || dev < 0 || dev >= kMaxDevices || !g_scratch[dev] )
        return SetError( THOTH_ERR_DEVICE, "no-allocation scratch not set up on the current device" );
    // every input is uploaded again and every partial-sum slot is written by
    // its block, so nothing of the previous call's contents is read
    DeviceBatch B;
    rc = PrepareBatch( n_books, books, cfg, B, nullptr, nullptr, false, g_scratch[dev], kNoAllocScratchBytes );
    if ( rc != THOTH_OK )
        return rc;
    return RunBatch( B, out );
}
