// mcl_lsm.hip -- Longstaff-Schwartz on the device paths recorded by the sweep
// (reference: tasks/pricer_mcl_american.cpp:250-447, helpers/maths.cpp:419-470).
//
// Fit: backward induction over the exercise dates t = M-2 .. 1.  One kernel
// per date does, per path, the exercise decision of date t+1 (fitted in the
// previous launch), the discounting of the path's cashflow to t and the eight
// sums of the {1, m, m^2} regression over the in-the-money paths of t (FP64,
// per-block partials); a one-block kernel adds the partials in a fixed order,
// factors the 3 x 3 normal matrix by Cholesky and stores the date's
// coefficients.  Apply: the frozen policy walks every path of every scenario
// forward to its first exercise; per-block sums of the discounted value and
// its square.  Built with -ffp-contract=off: every product rounds as the
// reference's does.
#include <hip/hip_runtime.h>

#include "mcl_launch.hpp"

namespace thoth
{

namespace
{

constexpr int kLsmThreads = 256;
constexpr int kLsmSums = 8;  // n, sum m, m^2, m^3, m^4, y, y m, y m^2

__device__ __forceinline__ double intrinsic( double s, double k, int is_call )
{
    return fmax( is_call ? s - k : k - s, 0.0 );
}

template <class R>
__global__ void __launch_bounds__( kLsmThreads )
lsm_step_kernel( const R* __restrict__ S, int64_t N, int M, int t, double strike, int is_call, double norm,
                 const double* __restrict__ df, const double* __restrict__ beta, double* __restrict__ cf,
                 double* __restrict__ part )
{
    // beta[4 t + 0..2]: coefficients of date t, beta[4 t + 3] = 1 where fitted
    const bool roll = t + 1 <= M - 2 && beta[4 * ( t + 1 ) + 3] != 0.0;
    const double b0 = roll ? beta[4 * ( t + 1 )] : 0.0, b1 = roll ? beta[4 * ( t + 1 ) + 1] : 0.0,
                 b2 = roll ? beta[4 * ( t + 1 ) + 2] : 0.0;
    const double d = df[t];
    double acc[kLsmSums];
#pragma unroll
    for ( int k = 0; k < kLsmSums; k++ )
        acc[k] = 0.0;
    for ( int64_t p = (int64_t)blockIdx.x * kLsmThreads + threadIdx.x; p < N; p += (int64_t)gridDim.x * kLsmThreads )
    {
        double c;
        if ( t == M - 2 )
            c = intrinsic( (double)S[(size_t)( M - 1 ) * N + p], strike, is_call );  // maturity payoff
        else
        {
            c = cf[p];
            if ( roll )
            {
                // the exercise decision at t + 1 (pricer_mcl_american.cpp:327-337)
                const double s1 = (double)S[(size_t)( t + 1 ) * N + p];
                const double x1 = intrinsic( s1, strike, is_call );
                if ( x1 > 0 )
                {
                    const double m = s1 / norm;
                    if ( x1 >= b0 + b1 * m + b2 * m * m )
                        c = x1;
                }
            }
        }
        c *= d;
        cf[p] = c;
        const double s = (double)S[(size_t)t * N + p];
        if ( intrinsic( s, strike, is_call ) > 0 )
        {
            const double m = s / norm;
            const double m2 = m * m;
            acc[0] += 1.0;
            acc[1] += m;
            acc[2] += m * m;
            acc[3] += m2 * m;
            acc[4] += m2 * m2;
            acc[5] += c;
            acc[6] += m * c;
            acc[7] += m2 * c;
        }
    }
    // wave sums, then the block's four waves in order
    __shared__ double ws[kLsmThreads / 64][kLsmSums];
#pragma unroll
    for ( int k = 0; k < kLsmSums; k++ )
    {
        double v = acc[k];
        for ( int o = 32; o > 0; o >>= 1 )
            v += __shfl_xor( v, o );
        if ( ( threadIdx.x & 63 ) == 0 )
            ws[threadIdx.x >> 6][k] = v;
    }
    __syncthreads();
    if ( threadIdx.x < kLsmSums )
    {
        double v = 0.0;
        for ( int w = 0; w < kLsmThreads / 64; w++ )
            v += ws[w][threadIdx.x];
        part[(size_t)blockIdx.x * kLsmSums + threadIdx.x] = v;
    }
}

// the date's regression: partials in a fixed order, the normal equations by
// Cholesky (maths.cpp:354-381, 441-469), at least 50 in-the-money paths
__global__ void __launch_bounds__( kLsmThreads ) lsm_solve_kernel( const double* __restrict__ part, int nblk, int t,
                                                                    double* __restrict__ beta )
{
    // thread j adds the blocks b = j / 8 (mod 32) of moment j % 8 (independent
    // loads in flight), then the 32 sub-sums of each moment in order
    constexpr int kRows = kLsmThreads / kLsmSums;
    __shared__ double sub[kRows][kLsmSums];
    __shared__ double sum[kLsmSums];
    {
        const int k = threadIdx.x % kLsmSums, r = threadIdx.x / kLsmSums;
        double v = 0.0;
        for ( int b = r; b < nblk; b += kRows )
            v += part[(size_t)b * kLsmSums + k];
        sub[r][k] = v;
    }
    __syncthreads();
    if ( threadIdx.x < kLsmSums )
    {
        double v = 0.0;
        for ( int r = 0; r < kRows; r++ )
            v += sub[r][threadIdx.x];
        sum[threadIdx.x] = v;
    }
    __syncthreads();
    if ( threadIdx.x != 0 )
        return;
    double* out = beta + 4 * t;
    out[0] = out[1] = out[2] = out[3] = 0.0;
    if ( sum[0] < 50.0 )
        return;  // too few in-the-money paths: hold at this date
    double A[3][3] = { { sum[0], sum[1], sum[2] }, { sum[1], sum[2], sum[3] }, { sum[2], sum[3], sum[4] } };
    const double g[3] = { sum[5], sum[6], sum[7] };
    for ( int i = 0; i < 3; i++ )
        for ( int j = 0; j <= i; j++ )
        {
            double s = A[i][j];
            for ( int k = 0; k < j; k++ )
                s -= A[i][k] * A[j][k];
            if ( i == j )
            {
                if ( s <= 0.0 )
                    return;  // not positive definite: hold
                A[i][j] = sqrt( s );
            }
            else
                A[i][j] = s / A[j][j];
        }
    double z[3], b[3];
    for ( int i = 0; i < 3; i++ )
    {
        double s = g[i];
        for ( int k = 0; k < i; k++ )
            s -= A[i][k] * z[k];
        z[i] = s / A[i][i];
    }
    for ( int i = 2; i >= 0; i-- )
    {
        double s = z[i];
        for ( int k = i + 1; k < 3; k++ )
            s -= A[k][i] * b[k];
        b[i] = s / A[i][i];
    }
    out[0] = b[0];
    out[1] = b[1];
    out[2] = b[2];
    out[3] = 1.0;
}

// the frozen policy on scenario blockIdx.y's paths (pricer_mcl_american.cpp:396-447):
// disc[y][t] = exp(-r_t (tau_t - tau_0)) of the scenario
template <class R>
__global__ void __launch_bounds__( kLsmThreads )
lsm_apply_kernel( const R* const* __restrict__ paths, int64_t N, int M, double strike, int is_call, double norm,
                  const double* __restrict__ disc, const double* __restrict__ beta, double* __restrict__ part )
{
    const R* S = paths[blockIdx.y];
    const double* dsc = disc + (size_t)blockIdx.y * M;
    double sum = 0.0, sum2 = 0.0;
    for ( int64_t p = (int64_t)blockIdx.x * kLsmThreads + threadIdx.x; p < N; p += (int64_t)gridDim.x * kLsmThreads )
    {
        double value = 0.0;
        for ( int t = 1; t < M; t++ )
        {
            const double s = (double)S[(size_t)t * N + p];
            const double x = intrinsic( s, strike, is_call );
            if ( t == M - 1 )
            {
                value = x * dsc[t];
                break;
            }
            if ( x > 0 && beta[4 * t + 3] != 0.0 )
            {
                const double m = s / norm;
                if ( x >= beta[4 * t] + beta[4 * t + 1] * m + beta[4 * t + 2] * m * m )
                {
                    value = x * dsc[t];
                    break;
                }
            }
        }
        sum += value;
        sum2 += value * value;
    }
    __shared__ double ws[kLsmThreads / 64][2];
    for ( int o = 32; o > 0; o >>= 1 )
    {
        sum += __shfl_xor( sum, o );
        sum2 += __shfl_xor( sum2, o );
    }
    if ( ( threadIdx.x & 63 ) == 0 )
    {
        ws[threadIdx.x >> 6][0] = sum;
        ws[threadIdx.x >> 6][1] = sum2;
    }
    __syncthreads();
    if ( threadIdx.x < 2 )
    {
        double v = 0.0;
        for ( int w = 0; w < kLsmThreads / 64; w++ )
            v += ws[w][threadIdx.x];
        part[( (size_t)blockIdx.y * gridDim.x + blockIdx.x ) * 2 + threadIdx.x] = v;
    }
}

}  // namespace

int lsm_blocks( int64_t N )
{
    const int64_t b = ( N + kLsmThreads - 1 ) / kLsmThreads;
    return (int)( b < 256 ? b : 256 );  // each thread then takes a few paths; few partials to add
}

bool launch_lsm_fit( hipStream_t s, int precision, const void* S, int64_t N, int M, double strike, int is_call,
                     double norm, const double* df, double* beta, double* cf, double* part )
{
    if ( M < 3 || N < 1 )
        return true;  // no interior exercise date
    const int nblk = lsm_blocks( N );
    for ( int t = M - 2; t >= 1; t-- )
    {
        if ( precision == 0 )
            hipLaunchKernelGGL( lsm_step_kernel<double>, dim3( nblk ), dim3( kLsmThreads ), 0, s,
                                static_cast<const double*>( S ), N, M, t, strike, is_call, norm, df, beta, cf, part );
        else
            hipLaunchKernelGGL( lsm_step_kernel<float>, dim3( nblk ), dim3( kLsmThreads ), 0, s,
                                static_cast<const float*>( S ), N, M, t, strike, is_call, norm, df, beta, cf, part );
        hipLaunchKernelGGL( lsm_solve_kernel, dim3( 1 ), dim3( kLsmThreads ), 0, s, part, nblk, t, beta );
    }
    return hipGetLastError() == hipSuccess;
}

bool launch_lsm_apply( hipStream_t s, int precision, const void* const* paths, int n_scen, int64_t N, int M,
                       double strike, int is_call, double norm, const double* disc, const double* beta, double* part )
{
    const int nblk = lsm_blocks( N );
    if ( precision == 0 )
        hipLaunchKernelGGL( lsm_apply_kernel<double>, dim3( nblk, n_scen ), dim3( kLsmThreads ), 0, s,
                            reinterpret_cast<const double* const*>( paths ), N, M, strike, is_call, norm, disc, beta,
                            part );
    else
        hipLaunchKernelGGL( lsm_apply_kernel<float>, dim3( nblk, n_scen ), dim3( kLsmThreads ), 0, s,
                            reinterpret_cast<const float* const*>( paths ), N, M, strike, is_call, norm, disc, beta,
                            part );
    return hipGetLastError() == hipSuccess;
}

}  // namespace thoth
