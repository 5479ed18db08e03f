// mcl_cv.hip -- training and evaluation kernels of the learned martingale
// control variate (no reference counterpart).
//
// Data (written by the FP32 path sweep, sample layout [replicate][j][path]):
//   feat [R][J][N][16] BF16 features x_{p,j} at hedge date d_j,
//   dm   [R][J][N][8]  sum over the steps of interval j of df(t_i) 100/S0_a M_{i,a},
//        M_i the first-order Doob increment of S_i (zero conditional mean),
//   pay  [R][N]        discounted book value P_p.
// Control variate CV_p = sum_j f(x_{p,j}) . dm_{p,j} has mean zero for any f,
// so Y_p = P_p - CV_p is an unbiased estimator of the premium; training
// minimises mean_p Y_p^2 over the network weights (Adam, FP32 master weights,
// BF16 MFMA forward/backward).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mcl_launch.hpp"
#include "mcl_mlp.hpp"
#include "thoth_mcl_cv.h"

namespace thoth
{

static_assert( (int)THOTH_CV_SILU == (int)kActSilu && (int)THOTH_CV_GATE == (int)kActGate, "activation codes of the C-ABI" );

namespace
{

// feature B operand of (j, path): lane l holds features 4 (l/16) .. +3
__device__ __forceinline__ bf16x4 load_feat( const uint16_t* feat, size_t row, int lane )
{
    const uint2 w = *reinterpret_cast<const uint2*>( feat + row * 16 + 4 * ( lane >> 4 ) );
    bf16x4 r;
    r[0] = (short)( w.x & 0xffffu );
    r[1] = (short)( w.x >> 16 );
    r[2] = (short)( w.y & 0xffffu );
    r[3] = (short)( w.y >> 16 );
    return r;
}

// increments of outputs 4 (l/16) .. +3 (lane groups 2, 3 hold none)
__device__ __forceinline__ f32x4 load_dm( const float* dm, size_t row, int lane )
{
    if ( lane >= 32 )
        return f32x4{ 0.f, 0.f, 0.f, 0.f };
    const float4 v = *reinterpret_cast<const float4*>( dm + row * 8 + 4 * ( lane >> 4 ) );
    return f32x4{ v.x, v.y, v.z, v.w };
}

}  // namespace

// BF16 images of the master weights (W_L^T row-major [rows][cols]) in the
// forward A-operand order (tiles of W^T) and the backward order (tiles of W).
__global__ void mlp_pack_kernel( const float* __restrict__ w, uint16_t* __restrict__ fwd, uint16_t* __restrict__ bwd )
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if ( e >= kMlpWSize )
        return;
    int L = 3;
    while ( L > 0 && e < kMlpWOff[L] )
        L--;
    const int k = e - kMlpWOff[L];
    const int o = k / kMlpCols[L], i = k % kMlpCols[L];
    const uint16_t h = (uint16_t)to_bf16( w[e] );
    {
        const int tile = kMlpTileOff[L] + ( o / 16 ) * kMlpKTiles[L] + i / 16;
        const int lane = ( o % 16 ) + 16 * ( ( i % 16 ) / 4 );
        fwd[( tile * 64 + lane ) * 4 + ( i % 4 )] = h;
    }
    {
        const int tile = kMlpTileOff[L] + ( i / 16 ) * kMlpOTiles[L] + o / 16;
        const int lane = ( i % 16 ) + 16 * ( ( o % 16 ) / 4 );
        bwd[( tile * 64 + lane ) * 4 + ( o % 4 )] = h;
    }
}

// Forward of every sample: per path Y = P - CV; block partial sums
// (sum Y, sum Y^2, sum P, sum P^2) in FP64 for each replicate.
// Workgroup = 16 paths (one sample column tile) x kEvalChunks waves, wave c
// taking the c-th chunk of the J dates: the dates' inferences are
// independent, so the chunks run side by side (kEvalChunks times the waves,
// each with kEvalChunks times fewer dependent inferences) and meet in LDS,
// summed in chunk order.
constexpr int kEvalChunks = 8;
template <int ACT>
__global__ void __launch_bounds__( 64 * kEvalChunks ) cv_eval_kernel( const uint2* __restrict__ wimg,
                                                                     const float* __restrict__ bias,
                                                                     const uint16_t* __restrict__ feat,
                                                                     const float* __restrict__ dm,
                                                                     const float* __restrict__ pay, int64_t N, int J,
                                                                     double* __restrict__ partial )
{
    const int lane = threadIdx.x & 63;
    const int chunk = threadIdx.x >> 6;
    const int r = blockIdx.y;
    const int64_t p = (int64_t)blockIdx.x * 16 + ( lane & 15 );
    const bool valid = p < N;
    const int64_t pp = valid ? p : 0;
    const int dpc = ( J + kEvalChunks - 1 ) / kEvalChunks;
    const int j0 = chunk * dpc, j1 = min( J, j0 + dpc );
    float cv = 0.0f;
    for ( int j = j0; j < j1; j++ )
    {
        const size_t row = ( (size_t)r * J + j ) * N + pp;
        cv += mlp_dot( mlp_infer<ACT>( wimg, bias, load_feat( feat, row, lane ), lane ), load_dm( dm, row, lane ) );
    }
    __shared__ float part[kEvalChunks][16];
    if ( lane < 16 )
        part[chunk][lane] = cv;
    __syncthreads();
    if ( chunk != 0 )
        return;
    float tot = 0.0f;
    if ( lane < 16 )
        for ( int c = 0; c < kEvalChunks; c++ )
            tot += part[c][lane];
    const float P = pay[(size_t)r * N + pp];
    double y = ( valid && lane < 16 ) ? (double)P - (double)tot : 0.0;
    double pv = ( valid && lane < 16 ) ? (double)P : 0.0;
    double v[4] = { y, y * y, pv, pv * pv };
#pragma unroll
    for ( int k = 0; k < 4; k++ )
    {
        double x = v[k];
        for ( int off = 32; off >= 1; off >>= 1 )
            x += __shfl_xor( x, off, 64 );
        v[k] = x;
    }
    if ( lane < 4 )
        partial[( (size_t)r * gridDim.x + blockIdx.x ) * 4 + lane] =
            lane == 0 ? v[0] : lane == 1 ? v[1] : lane == 2 ? v[2] : v[3];
}

// Training step, part 1: the control variate of the minibatch split by date
// chunks.  Wave (tile, chunk) takes 16 paths of the minibatch
// [ (step B) mod N, + B ) and the dates of its chunk; cvpart[chunk][b] =
// sum over those dates of f(x) . dm.
template <int ACT>
__global__ void __launch_bounds__( 256 ) cv_part_kernel( const uint2* __restrict__ wimg, const float* __restrict__ bias,
                                                        const uint16_t* __restrict__ feat, const float* __restrict__ dm,
                                                        int64_t N, int J, int B, int dpc, const int* __restrict__ step,
                                                        float* __restrict__ cvpart )
{
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * 4 + ( threadIdx.x >> 6 );
    const int tiles = B / 16;
    const int tile = wave % tiles, chunk = wave / tiles;
    const int64_t start = ( (int64_t)( *step ) * B ) % N;
    const int b = tile * 16 + ( lane & 15 );
    const int64_t p = ( start + b ) % N;
    const int j0 = chunk * dpc, j1 = min( J, j0 + dpc );
    float cv = 0.0f;
    for ( int j = j0; j < j1; j++ )
    {
        const size_t row = (size_t)j * N + p;
        cv += mlp_dot( mlp_infer<ACT>( wimg, bias, load_feat( feat, row, lane ), lane ), load_dm( dm, row, lane ) );
    }
    if ( lane < 16 )
        cvpart[(size_t)chunk * B + b] = cv;
}

// Training step, part 2: gradient of mean_b Y_b^2, Y = P - sum_chunks cvpart.
// Each wave back-propagates its (16 paths) x (dates of its chunk) samples with
// the activations in LDS and the weight gradient in registers; the block's 4
// waves are summed in a fixed order through LDS into one slice per block
// (deterministic).  loss_part[block] = sum of Y^2 over the block's paths when
// the block holds chunk 0, else 0.
constexpr int kGradN = kMlpWSize + kMlpBSize;
constexpr int kGradGroups = kCvGradGroups;  // first-level groups of the gradient-slice sum
constexpr size_t kCvGradLds = 4 * sizeof( MlpActLds );
static_assert( kCvGradLds >= kGradN * sizeof( float ), "gradient reduction fits in the activation area" );

template <int ACT>
__global__ void __launch_bounds__( 256 ) cv_grad_kernel( const uint2* __restrict__ wimg, const uint2* __restrict__ wbwd,
                                                        const float* __restrict__ bias, const uint16_t* __restrict__ feat,
                                                        const float* __restrict__ dm, const float* __restrict__ pay,
                                                        int64_t N, int J, int B, int dpc, int chunks,
                                                        const int* __restrict__ step, const float* __restrict__ cvpart,
                                                        float* __restrict__ gpart, float* __restrict__ loss_part )
{
    extern __shared__ __align__( 16 ) char smem[];
    const int lane = threadIdx.x & 63;
    const int wib = threadIdx.x >> 6;
    const int wave = blockIdx.x * 4 + wib;
    const int tiles = B / 16;
    const int tile = wave % tiles, chunk = wave / tiles;
    const int64_t start = ( (int64_t)( *step ) * B ) % N;
    const int b = tile * 16 + ( lane & 15 );
    const int64_t p = ( start + b ) % N;
    float cv = 0.0f;
    for ( int c = 0; c < chunks; c++ )
        cv += cvpart[(size_t)c * B + b];
    const float y = pay[p] - cv;
    const float g = -2.0f * y / (float)B;
    MlpActLds& act = reinterpret_cast<MlpActLds*>( smem )[wib];
    MlpGrad G;
#pragma unroll
    for ( int k = 0; k < kMlpTiles; k++ )
        G.w[k] = f32x4{ 0.f, 0.f, 0.f, 0.f };
#pragma unroll
    for ( int k = 0; k < 13; k++ )
        G.b[k] = f32x4{ 0.f, 0.f, 0.f, 0.f };
    const int j0 = chunk * dpc, j1 = min( J, j0 + dpc );
    // the next date's features and increments are loaded while this date's
    // forward and backward run (one wave per SIMD: nothing else hides them)
    bf16x4 fx = load_feat( feat, (size_t)j0 * N + p, lane );
    f32x4 dx = load_dm( dm, (size_t)j0 * N + p, lane );
    for ( int j = j0; j < j1; j++ )
    {
        const bf16x4 f = fx;
        const f32x4 d = dx;
        if ( j + 1 < j1 )
        {
            fx = load_feat( feat, (size_t)( j + 1 ) * N + p, lane );
            dx = load_dm( dm, (size_t)( j + 1 ) * N + p, lane );
        }
        (void)mlp_forward_lds<ACT>( wimg, bias, f, lane, act );
        mlp_backward_lds<ACT>( wbwd, act, f32x4{ g * d[0], g * d[1], g * d[2], g * d[3] }, lane, G );
    }
    // bias partials: sum the 16 sample columns of each lane group
    constexpr int kBiasTile[4] = { 0, 4, 8, 12 };
#pragma unroll
    for ( int k = 0; k < 13; k++ )
#pragma unroll
        for ( int r = 0; r < 4; r++ )
        {
            float x = G.b[k][r];
            x += __shfl_xor( x, 1, 64 );
            x += __shfl_xor( x, 2, 64 );
            x += __shfl_xor( x, 4, 64 );
            x += __shfl_xor( x, 8, 64 );
            G.b[k][r] = x;
        }
    // ordered block sum: wave 0 stores, waves 1..3 add in turn
    __syncthreads();  // every wave is done with its activation block
    float* red = reinterpret_cast<float*>( smem );
    for ( int w = 0; w < 4; w++ )
    {
        if ( wib == w )
        {
#pragma unroll
            for ( int L = 0; L < 4; L++ )
#pragma unroll
                for ( int ot = 0; ot < kMlpOTiles[L]; ot++ )
                {
#pragma unroll
                    for ( int it = 0; it < kMlpKTiles[L]; it++ )
                    {
                        const f32x4 v = G.w[kMlpTileOff[L] + ot * kMlpKTiles[L] + it];
#pragma unroll
                        for ( int r = 0; r < 4; r++ )
                        {
                            const int o = 16 * ot + 4 * ( lane >> 4 ) + r, i = 16 * it + ( lane & 15 );
                            float& dst = red[kMlpWOff[L] + o * kMlpCols[L] + i];
                            dst = ( w == 0 ) ? v[r] : dst + v[r];
                        }
                    }
                    if ( ( lane & 15 ) == 0 )
                    {
                        const f32x4 v = G.b[kBiasTile[L] + ot];
#pragma unroll
                        for ( int r = 0; r < 4; r++ )
                        {
                            float& dst = red[kMlpWSize + kMlpBOff[L] + 16 * ot + 4 * ( lane >> 4 ) + r];
                            dst = ( w == 0 ) ? v[r] : dst + v[r];
                        }
                    }
                }
        }
        __syncthreads();
    }
    float* gw = gpart + (size_t)blockIdx.x * kGradN;
    for ( int e = threadIdx.x; e < kGradN; e += 256 )
        gw[e] = red[e];
    // loss of the block's paths (chunk-0 waves only)
    float l2 = ( chunk == 0 && lane < 16 ) ? y * y : 0.0f;
    for ( int off = 32; off >= 1; off >>= 1 )
        l2 += __shfl_xor( l2, off, 64 );
    __shared__ float lsum[4];
    if ( lane == 0 )
        lsum[wib] = l2;
    __syncthreads();
    if ( threadIdx.x == 0 )
        loss_part[blockIdx.x] = ( lsum[0] + lsum[1] ) + ( lsum[2] + lsum[3] );
}

// Adam (Kingma & Ba): the gradient of element e is the ordered sum of the
// per-block slices; FP32 master weights and moments.
// first level of the gradient sum: group y of kGradGroups adds its slices in
// order (a fixed grouping, so the sum is the same on every run)
__global__ void cv_gsum_kernel( const float* __restrict__ gpart, int nslices, float* __restrict__ gsum )
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if ( e >= kGradN )
        return;
    const int per = ( nslices + kGradGroups - 1 ) / kGradGroups;
    const int k0 = blockIdx.y * per, k1 = min( nslices, k0 + per );
    float g = 0.0f;
    for ( int k = k0; k < k1; k++ )
        g += gpart[(size_t)k * kGradN + e];
    gsum[(size_t)blockIdx.y * kGradN + e] = g;
}

__global__ void cv_adam_kernel( float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                                const float* __restrict__ gsum, const int* __restrict__ step, float lr, float b1,
                                float b2, float eps )
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if ( e >= kGradN )
        return;
    float g = 0.0f;
    for ( int k = 0; k < kGradGroups; k++ )
        g += gsum[(size_t)k * kGradN + e];
    const int t = *step + 1;
    const float mm = b1 * m[e] + ( 1.0f - b1 ) * g;
    const float vv = b2 * v[e] + ( 1.0f - b2 ) * g * g;
    m[e] = mm;
    v[e] = vv;
    const float mh = mm / ( 1.0f - powf( b1, (float)t ) );
    const float vh = vv / ( 1.0f - powf( b2, (float)t ) );
    w[e] -= lr * mh / ( sqrtf( vh ) + eps );
}

__global__ void cv_step_kernel( int* step ) { *step += 1; }

// ---- launch helpers (host side) -------------------------------------------
void launch_mlp_pack( hipStream_t s, const float* w, uint16_t* fwd, uint16_t* bwd )
{
    hipLaunchKernelGGL( mlp_pack_kernel, dim3( ( kMlpWSize + 255 ) / 256 ), dim3( 256 ), 0, s, w, fwd, bwd );
}

uint32_t cv_eval_blocks( int64_t N ) { return (uint32_t)( ( N + 15 ) / 16 ); }

void launch_cv_eval( hipStream_t s, int act, const uint16_t* fwd, const float* bias, const uint16_t* feat,
                     const float* dm, const float* pay, int64_t N, int J, int R, double* partial )
{
    const uint32_t nb = cv_eval_blocks( N );
    if ( act == kActGate )
        hipLaunchKernelGGL( cv_eval_kernel<kActGate>, dim3( nb, R ), dim3( 64 * kEvalChunks ), 0, s,
                            reinterpret_cast<const uint2*>( fwd ), bias, feat, dm, pay, N, J, partial );
    else
        hipLaunchKernelGGL( cv_eval_kernel<kActSilu>, dim3( nb, R ), dim3( 64 * kEvalChunks ), 0, s,
                            reinterpret_cast<const uint2*>( fwd ), bias, feat, dm, pay, N, J, partial );
}

hipError_t cv_train_init()
{
    hipError_t e = hipFuncSetAttribute( reinterpret_cast<const void*>( cv_grad_kernel<kActSilu> ),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)kCvGradLds );
    if ( e != hipSuccess )
        return e;
    return hipFuncSetAttribute( reinterpret_cast<const void*>( cv_grad_kernel<kActGate> ),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kCvGradLds );
}

CvTrainLayout cv_train_layout( int B, int J )
{
    CvTrainLayout l{};
    l.B = B;
    l.J = J;
    l.tiles = B / 16;
    // about 512 blocks of 4 waves: split the dates into chunks
    const int want = std::max( 1, std::min( J, ( 2048 + l.tiles - 1 ) / l.tiles ) );
    l.dpc = ( J + want - 1 ) / want;
    l.chunks = ( J + l.dpc - 1 ) / l.dpc;
    l.waves = l.tiles * l.chunks;
    l.blocks = l.waves / 4;
    return l;
}

template <int ACT>
static void launch_cv_forward_backward( hipStream_t s, const CvTrainLayout& l, const uint16_t* fwd,
                                        const uint16_t* bwd, const float* w, const uint16_t* feat, const float* dm,
                                        const float* pay, int64_t N, const int* step, float* cvpart, float* gpart,
                                        float* loss_part )
{
    hipLaunchKernelGGL( cv_part_kernel<ACT>, dim3( l.blocks ), dim3( 256 ), 0, s, reinterpret_cast<const uint2*>( fwd ),
                        w + kMlpWSize, feat, dm, N, l.J, l.B, l.dpc, step, cvpart );
    hipLaunchKernelGGL( cv_grad_kernel<ACT>, dim3( l.blocks ), dim3( 256 ), kCvGradLds, s,
                        reinterpret_cast<const uint2*>( fwd ), reinterpret_cast<const uint2*>( bwd ), w + kMlpWSize,
                        feat, dm, pay, N, l.J, l.B, l.dpc, l.chunks, step, cvpart, gpart, loss_part );
}

bool launch_cv_train_step( hipStream_t s, int act, const CvTrainLayout& l, const uint16_t* fwd, const uint16_t* bwd,
                           float* w, float* m, float* v, const uint16_t* feat, const float* dm, const float* pay,
                           int64_t N, int* step, float* cvpart, float* gpart, float* gsum, float* loss_part, float lr )
{
    if ( l.blocks < 1 || l.blocks * 4 != l.waves )
        return false;
    if ( act == kActGate )
        launch_cv_forward_backward<kActGate>( s, l, fwd, bwd, w, feat, dm, pay, N, step, cvpart, gpart, loss_part );
    else
        launch_cv_forward_backward<kActSilu>( s, l, fwd, bwd, w, feat, dm, pay, N, step, cvpart, gpart, loss_part );
    hipLaunchKernelGGL( cv_gsum_kernel, dim3( ( kGradN + 255 ) / 256, kGradGroups ), dim3( 256 ), 0, s, gpart, l.blocks,
                        gsum );
    hipLaunchKernelGGL( cv_adam_kernel, dim3( ( kGradN + 255 ) / 256 ), dim3( 256 ), 0, s, w, m, v, gsum, step, lr,
                        0.9f, 0.999f, 1e-8f );
    hipLaunchKernelGGL( mlp_pack_kernel, dim3( ( kMlpWSize + 255 ) / 256 ), dim3( 256 ), 0, s, w,
                        const_cast<uint16_t*>( fwd ), const_cast<uint16_t*>( bwd ) );
    hipLaunchKernelGGL( cv_step_kernel, dim3( 1 ), dim3( 1 ), 0, s, step );
    return true;
}

}  // namespace thoth
