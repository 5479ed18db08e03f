// mcl_gpu.cpp -- host side of the device library: device discovery, the
// Joe-Kuo direction numbers, book preparation (bridge program, FP32 step
// constants, payoff masks) and the C-ABI of include/thoth_mcl_gpu.h.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "mcl_device.hpp"
#include "mcl_internal.hpp"
#include "mcl_launch.hpp"
#include "thoth_mcl_gpu.h"

// The Joe-Kuo "new-joe-kuo-6.21201" table (first 4096 rows), embedded verbatim
// from thoth_amd/data at build time.
asm( ".section .rodata\n"
     ".p2align 3\n"
     ".hidden thoth_jk_begin\n"
     ".hidden thoth_jk_end\n"
     "thoth_jk_begin:\n"
     ".incbin \"" THOTH_JK_PATH "\"\n"
     "thoth_jk_end:\n"
     ".byte 0\n"
     ".previous\n" );
extern "C" const char thoth_jk_begin[];
extern "C" const char thoth_jk_end[];

namespace thoth
{

namespace
{
thread_local std::string g_err;

struct DeviceState
{
    bool probed = false;
    bool ok = false;
    std::string info;
    uint32_t* d_dirs = nullptr;  // [kSobolDims][32] directions, then the lane nibble tables
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = { nullptr, nullptr };  // sweep timing (calls are serialised)
};
DeviceState g_dev;

int Fail( int code, const std::string& msg )
{
    g_err = msg;
    return code;
}

#define HIP_OK( expr )                                                                            \
    do                                                                                            \
    {                                                                                             \
        hipError_t e_ = ( expr );                                                                 \
        if ( e_ != hipSuccess )                                                                   \
            return Fail( THOTH_ERR_DEVICE, std::string( #expr " failed: " ) + hipGetErrorString( e_ ) ); \
    } while ( 0 )

}  // namespace

const std::vector<uint32_t>& SobolDirections()
{
    static std::vector<uint32_t> V;
    static std::once_flag once;
    std::call_once( once, [] {
        // sobol_generator.cpp:22-76: dimension 0 implicit, file entry d-1 -> dimension d
        std::istringstream in( std::string( thoth_jk_begin, thoth_jk_end ) );
        std::string line;
        std::getline( in, line );
        V.assign( (size_t)kSobolDims * 32, 0u );
        for ( int i = 0; i < 32; i++ )
            V[i] = 1u << ( 31 - i );
        int d = 1;
        while ( d < kSobolDims && std::getline( in, line ) )
        {
            std::istringstream ls( line );
            unsigned dim, s, a;
            if ( !( ls >> dim >> s >> a ) )
                continue;
            std::vector<unsigned> m( s );
            for ( unsigned k = 0; k < s; k++ )
                ls >> m[k];
            uint32_t* v = &V[(size_t)d * 32];
            if ( s >= 32 )
            {
                for ( int k = 0; k < 32; k++ )
                    v[k] = (uint32_t)m[k] << ( 31 - k );
            }
            else
            {
                for ( unsigned k = 0; k < s; k++ )
                    v[k] = (uint32_t)m[k] << ( 31 - k );
                for ( unsigned k = s; k < 32; k++ )
                {
                    v[k] = v[k - s] ^ ( v[k - s] >> s );
                    for ( unsigned j = 1; j < s; j++ )
                        v[k] ^= ( ( a >> ( s - 1 - j ) ) & 1u ) * v[k - j];
                }
            }
            d++;
        }
    } );
    return V;
}

int EnsureDevice()
{
    if ( g_dev.probed )
        return g_dev.ok ? THOTH_OK : Fail( THOTH_ERR_UNAVAILABLE, g_dev.info );
    g_dev.probed = true;
    int n = 0;
    if ( hipGetDeviceCount( &n ) != hipSuccess || n == 0 )
    {
        g_dev.info = "no HIP device found";
        return Fail( THOTH_ERR_UNAVAILABLE, g_dev.info );
    }
    hipDeviceProp_t p;
    if ( hipGetDeviceProperties( &p, 0 ) != hipSuccess )
    {
        g_dev.info = "HIP device query failed";
        return Fail( THOTH_ERR_UNAVAILABLE, g_dev.info );
    }
    std::string arch = p.gcnArchName;
    std::ostringstream os;
    os << p.name << " (" << arch << ", " << p.multiProcessorCount << " CUs, "
       << ( p.totalGlobalMem >> 30 ) << " GiB)";
    g_dev.info = os.str();
    if ( arch.rfind( "gfx950", 0 ) != 0 )
    {
        g_dev.info += ": code objects are built for gfx950 only";
        return Fail( THOTH_ERR_UNAVAILABLE, g_dev.info );
    }
    HIP_OK( hipSetDevice( 0 ) );
    HIP_OK( hipStreamCreate( &g_dev.stream ) );
    // [dims][32] direction numbers, then [dims][32] lane nibble tables: words
    // c < 16 XOR V[d][k] over the bits k < 4 of c, words 16 + c XOR V[d][4 + k]
    const std::vector<uint32_t>& V = SobolDirections();
    const size_t dims = V.size() / 32;
    if ( dims != (size_t)kSobolDimsDev )
        return Fail( THOTH_ERR_DEVICE, "Sobol direction table size" );
    std::vector<uint32_t> tables( V );
    tables.resize( 2 * V.size() );
    for ( size_t d = 0; d < dims; d++ )
        for ( int c = 0; c < 16; c++ )
        {
            uint32_t lo = 0, hi = 0;
            for ( int k = 0; k < 4; k++ )
                if ( ( c >> k ) & 1 )
                {
                    lo ^= V[d * 32 + k];
                    hi ^= V[d * 32 + 4 + k];
                }
            tables[kSobolNibbles + d * 32 + c] = lo;
            tables[kSobolNibbles + d * 32 + 16 + c] = hi;
        }
    HIP_OK( hipMalloc( &g_dev.d_dirs, tables.size() * sizeof( uint32_t ) ) );
    HIP_OK( hipMemcpy( g_dev.d_dirs, tables.data(), tables.size() * sizeof( uint32_t ), hipMemcpyHostToDevice ) );
    HIP_OK( hipEventCreate( &g_dev.ev[0] ) );
    HIP_OK( hipEventCreate( &g_dev.ev[1] ) );
    // the device address of the no-allocation entry point's scratch array
    int rc = NoAllocSetup();
    if ( rc != THOTH_OK )
        return rc;
    g_dev.ok = true;
    return THOTH_OK;
}

hipEvent_t SweepEvent( int k ) { return g_dev.ev[k]; }

std::mutex& DeviceMutex()
{
    static std::mutex mu;
    return mu;
}

hipStream_t Stream() { return g_dev.stream; }
const uint32_t* DeviceDirections() { return g_dev.d_dirs; }
int SetError( int code, const std::string& msg ) { return Fail( code, msg ); }

// ---------------------------------------------------------------------------
// Brownian bridge program: the reference's schedule (path_generator.cpp:49-120)
// replayed as an in-order walk, so that W is produced in time order.  Each
// step pops one point; before the pop it builds the pending midpoints between
// the last popped point and the anchor on top of the stack.
// ---------------------------------------------------------------------------
BridgeProgram BuildBridgeProgram( const double* t, int n_dates )
{
    BridgeProgram bp;
    const int m = n_dates - 1;
    std::vector<int> bridge_index( m ), left_index( m ), right_index( m );
    std::vector<double> lw( m ), rw( m ), sd( m );
    auto bt = [&]( int k ) { return t[k + 1]; };
    std::vector<int> map( m, 0 );
    map[m - 1] = 1;
    bridge_index[0] = m - 1;
    sd[0] = std::sqrt( bt( m - 1 ) );
    int j = 0;
    for ( int i = 1; i < m; i++ )
    {
        while ( map[j] )
            j++;
        int k = j;
        while ( !map[k] )
            k++;
        int l = j + ( ( k - 1 - j ) >> 1 );
        map[l] = i + 1;
        bridge_index[i] = l;
        left_index[i] = j;
        right_index[i] = k;
        if ( j != 0 )
        {
            lw[i] = ( bt( k ) - bt( l ) ) / ( bt( k ) - bt( j - 1 ) );
            rw[i] = ( bt( l ) - bt( j - 1 ) ) / ( bt( k ) - bt( j - 1 ) );
            sd[i] = std::sqrt( ( bt( l ) - bt( j - 1 ) ) * ( bt( k ) - bt( l ) ) / ( bt( k ) - bt( j - 1 ) ) );
        }
        else
        {
            lw[i] = ( bt( k ) - bt( l ) ) / bt( k );
            rw[i] = bt( l ) / bt( k );
            sd[i] = std::sqrt( bt( l ) * ( bt( k ) - bt( l ) ) / bt( k ) );
        }
        j = k + 1;
        if ( j >= m )
            j = 0;
    }
    std::vector<int> order( m );
    for ( int i = 0; i < m; i++ )
        order[bridge_index[i]] = i;
    bp.sd0 = sd[0];
    bp.step_ops.push_back( 0 );
    std::vector<int> stack{ m - 1 };
    int depth = 1;
    for ( int p = 0; p < m; p++ )
    {
        while ( stack.back() != p )
        {
            const int k = stack.back();
            const int l = p + ( ( k - 1 - p ) >> 1 );
            const int i = order[l];
            bp.op_b.push_back( i );
            bp.op_left.push_back( p != 0 );
            bp.op_w.push_back( lw[i] );
            bp.op_w.push_back( rw[i] );
            bp.op_w.push_back( sd[i] );
            stack.push_back( l );
            depth = std::max<int>( depth, (int)stack.size() );
        }
        stack.pop_back();
        bp.step_ops.push_back( (int)bp.op_b.size() );
    }
    bp.max_depth = depth;
    return bp;
}

}  // namespace thoth

using namespace thoth;


extern "C" {

__attribute__( ( visibility( "default" ) ) ) const char* thoth_last_error( void ) { return g_err.c_str(); }

__attribute__( ( visibility( "default" ) ) ) int thoth_gpu_available( void )
{
    std::lock_guard<std::mutex> lk( DeviceMutex() );
    return EnsureDevice() == THOTH_OK ? 1 : 0;
}

__attribute__( ( visibility( "default" ) ) ) int thoth_gpu_device_info( char* buf, size_t len )
{
    std::lock_guard<std::mutex> lk( DeviceMutex() );
    int rc = EnsureDevice();
    if ( buf && len )
    {
        std::strncpy( buf, g_dev.info.c_str(), len - 1 );
        buf[len - 1] = 0;
    }
    return rc;
}

__attribute__( ( visibility( "default" ) ) ) int thoth_mcl_sweep( int32_t n_books, const thoth_mcl_book* books, const thoth_mcl_sweep_cfg* cfg,
                     thoth_mcl_book_out* out )
{
    std::lock_guard<std::mutex> lk( DeviceMutex() );
    return SweepBooks( n_books, books, cfg, out );
}

__attribute__( ( visibility( "default" ) ) ) int thoth_gpu_sobol_points( int32_t dim, int64_t skip, int64_t n,
                                                                        uint32_t* out )
{
    std::lock_guard<std::mutex> lk( DeviceMutex() );
    int rc = EnsureDevice();
    if ( rc != THOTH_OK )
        return rc;
    if ( dim < 1 || dim > kSobolDims || skip < 0 || n < 1 || !out )
        return Fail( THOTH_ERR_INVALID, "dim in [1, 4097], skip >= 0, n >= 1" );
    const uint64_t first = (uint64_t)skip + 1, last = (uint64_t)skip + (uint64_t)n;
    if ( last >= 0xffffffffull )
        return Fail( THOTH_ERR_UNSUPPORTED, "Sobol index beyond 2^32 - 2" );
    uint32_t* d_out = nullptr;
    const size_t bytes = (size_t)n * dim * sizeof( uint32_t );
    HIP_OK( hipMalloc( &d_out, bytes ) );
    bool ok = launch_sobol_points( g_dev.stream, g_dev.d_dirs, dim, (uint32_t)first, (uint32_t)last, d_out );
    hipError_t e = hipMemcpyAsync( out, d_out, bytes, hipMemcpyDeviceToHost, g_dev.stream );
    if ( e == hipSuccess )
        e = hipStreamSynchronize( g_dev.stream );
    (void)hipFree( d_out );
    if ( !ok || e != hipSuccess )
        return Fail( THOTH_ERR_DEVICE, "sobol points kernel failed" );
    return THOTH_OK;
}

__attribute__( ( visibility( "default" ) ) ) int thoth_gpu_special_f32( int32_t fn, int64_t n, const uint32_t* in,
                                                                       float* out )
{
    std::lock_guard<std::mutex> lk( DeviceMutex() );
    int rc = EnsureDevice();
    if ( rc != THOTH_OK )
        return rc;
    if ( fn < 0 || fn > 2 || n < 1 || n > ( (int64_t)1 << 28 ) || !in || !out )
        return Fail( THOTH_ERR_INVALID, "fn in [0, 2], n in [1, 2^28], non-null buffers" );
    char* d = nullptr;
    const size_t bytes = (size_t)n * sizeof( uint32_t );
    HIP_OK( hipMalloc( &d, 2 * bytes ) );
    uint32_t* d_in = reinterpret_cast<uint32_t*>( d );
    float* d_out = reinterpret_cast<float*>( d + bytes );
    hipError_t e = hipMemcpyAsync( d_in, in, bytes, hipMemcpyHostToDevice, g_dev.stream );
    const bool ok = e == hipSuccess && launch_special_f32( g_dev.stream, fn, n, d_in, d_out );
    if ( ok )
        e = hipMemcpyAsync( out, d_out, bytes, hipMemcpyDeviceToHost, g_dev.stream );
    if ( ok && e == hipSuccess )
        e = hipStreamSynchronize( g_dev.stream );
    (void)hipFree( d );
    if ( !ok || e != hipSuccess )
        return Fail( THOTH_ERR_DEVICE, "special-function kernel failed" );
    return THOTH_OK;
}

__attribute__( ( visibility( "default" ) ) ) int thoth_gpu_price_european_gbm( double forward, double strike, double t, double vol, double df, int is_call,
                                  long long paths, uint64_t seed, double* premium, double* trust )
{
    // mcl_gpu.cu:100-152 contract; Paths <= 0 -> {0, 0}
    if ( premium )
        *premium = 0;
    if ( trust )
        *trust = 0;
    if ( paths <= 0 )
        return THOTH_OK;
    const double tt[2] = { 0.0, t }, dtt[2] = { 0.0, t };
    const int32_t model = THOTH_MODEL_BS;
    const double spot = forward, zero = 0.0, one = 1.0;
    const double drift[2] = { 0.0, 0.0 }, disc[2] = { 0.0, 0.0 };
    const int32_t kind = THOTH_PAYOFF_VANILLA, call = is_call ? 1 : 0, udl = 0, flow = 1, no = 0;
    const int32_t mon_off[2] = { 0, 0 };
    thoth_mcl_book b{};
    b.n_dates = 2;
    b.t = tt;
    b.dt = dtt;
    b.n_assets = 1;
    b.model = &model;
    b.spot = &spot;
    b.vol = &vol;
    b.v0 = b.kappa = b.theta = b.xi = b.rho = &zero;
    b.drift = drift;
    b.chol = &one;
    b.disc_rate = disc;
    b.n_contracts = 1;
    b.c_kind = &kind;
    b.c_is_call = &call;
    b.c_underlying = &udl;
    b.c_flow_index = &flow;
    b.c_strike = &strike;
    b.c_is_up = b.c_is_in = &no;
    b.c_level = &zero;
    b.c_mon_offset = mon_off;
    b.c_mon_index = nullptr;
    // The reference prices every unsigned long seed (pricer/src/tasks/mcl_configuration.cpp:24).
    // While points 1 + Seed Paths .. (Seed + 1) Paths stay below 2^32 - 1 the
    // seed selects that disjoint run of Sobol points; past it (large seeds,
    // a negative YAML seed wrapped to 2^64 - k) the seed instead seeds one
    // random digital shift of points 0 .. Paths - 1: still a fixed,
    // reproducible point set per seed, shared by every bump.
    thoth_mcl_sweep_cfg c{};
    c.paths = paths;
    c.precision = 0;
    const unsigned __int128 end = ( (unsigned __int128)seed + 1u ) * (unsigned __int128)(uint64_t)paths;
    if ( end < 0xffffffffull )
        c.sobol_skip = (int64_t)( seed * (uint64_t)paths );
    else
    {
        c.replicates = 1;
        c.rqmc_seed = seed;
    }
    thoth_mcl_book_out o{};
    int rc = thoth_mcl_sweep( 1, &b, &c, &o );
    if ( rc != THOTH_OK )
        return rc;
    const double n = (double)paths;
    const double mean = o.book_sum / n;
    double var = o.book_sum2 / n - mean * mean;
    if ( var < 0.0 )
        var = 0.0;
    if ( premium )
        *premium = df * mean;
    if ( trust )
        *trust = df * std::sqrt( var / n );
    return THOTH_OK;
}

}  // extern "C"
