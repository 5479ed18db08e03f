// mcl_sweep_waves.hpp -- latency form of the FP32 sweep of uniform LSV books
// (every equity LSV with vol-of-vol > 0, uncorrelated stochastic spots, no
// barrier monitoring): the NA equities of a path are stepped by the NA waves
// of a workgroup, wave a holding equity a of the same 64 paths.
//
// The lane form (sweep_block) steps all NA equities of a path in one lane, so
// a wave's latency per step is NA equities' worth of dependent arithmetic; a
// pricing of a few tens of thousands of paths puts fewer than one wave on each
// SIMD and runs at that latency.  Here each wave steps one equity (its two
// factors of the Brownian bridge, its QE variance draw and its leverage
// lookup, with the equity's step constants in scalar registers), so the same
// pricing runs NA times as many waves, each NA times shorter.  The equities
// meet only where a payoff reads them (an exchange of the spots through LDS at
// a flow date) and where the control-variate samples are written (each wave
// its own equity's features and martingale increments).  The arithmetic of a
// path is that of the uniform lane kernel.
//
// Workgroup = NA waves = one quarter (64 indices) of a 256-index Sobol block;
// grid (4 x Sobol blocks, books, replicates); per-tile partial sums as the
// lane form writes them, with 4 tiles per Sobol block.
#pragma once
#include <hip/hip_runtime.h>

#include "mcl_sweep.hpp"

namespace thoth
{

template <int NA, int DEPTH, int CV>
__global__ void __launch_bounds__( 64 * NA ) sweep_waves( const DevBook* __restrict__ books,
                                                          const uint32_t* __restrict__ V, uint32_t bq0,
                                                          uint32_t first_idx, uint32_t last_idx,
                                                          const uint32_t* __restrict__ shifts )
{
    static_assert( CV == 0 || ( CV == 1 && NA <= 5 ), "control-variate samples of at most five equities" );
    constexpr int NF = 2 * NA;
    constexpr int kThreads = 64 * NA;
    constexpr bool cv_samples = CV == 1;
    const DevBook& B = books[blockIdx.y];
    const int z = blockIdx.z;
    const int lane = threadIdx.x & 63;
    const int a = __builtin_amdgcn_readfirstlane( (int)( threadIdx.x >> 6 ) );  // this wave's equity
    const uint32_t bq = bq0 + blockIdx.x / 4u;
    const uint32_t l = 64u * ( blockIdx.x & 3u ) + (uint32_t)lane;  // index within the Sobol block
    const uint32_t idx = bq * 256u + l;
    const bool valid = idx >= first_idx && idx <= last_idx;
    const uint32_t* sh = shifts ? shifts + (size_t)z * kSobolDimsDev : nullptr;
    const int F = B.factors;
    const int m = B.m;
    const int nc = B.n_contracts;

    // dynamic LDS: [Sobol block table][Sobol ring: nslots x NF x 32 words]
    // [leverage ring: NA x 2G layers x P][spots at a flow date: NA x 64]
    extern __shared__ uint32_t lds[];
    const int nslots = B.sob_slots;
    const int tab_words = ( min( m * F, kSobolDimsDev ) + 3 ) & ~3;
    uint32_t* tab = lds;
    uint32_t* ring = lds + tab_words;
    const int P = B.lev_points;
    const int G = B.lev_group;
    float* levs = reinterpret_cast<float*>( ring + nslots * NF * 32 );
    float* xs = levs + 2 * G * NA * P;
    __shared__ double acc[kMaxContracts + 1][2];
    for ( int k = threadIdx.x; k < ( kMaxContracts + 1 ) * 2; k += kThreads )
        ( &acc[0][0] )[k] = 0.0;

    // ---- Sobol points, as the lane form: block part tabulated once per
    // workgroup, lane part of gray(l) from the per-op nibble tables in the ring
    const uint32_t* nib = V + kSobolNibbles;
    const uint32_t gln = gray32( l );
    const uint32_t* rlo = ring + ( gln & 15u );
    const uint32_t* rhi = ring + 16 + ( gln >> 4 );
    typedef uint32_t u32x4 __attribute__( ( ext_vector_type( 4 ) ) );
    auto stage_ops = [&]( int o0, int o1 ) __attribute__( ( always_inline ) )
    {
        const int n = ( o1 - o0 ) * NF * 4;
        for ( int k = threadIdx.x; k < n; k += kThreads )
        {
            const int q = k & 3;
            const int pair = k >> 2;
            const int o = o0 + pair / NF;
            const int f = pair - ( pair / NF ) * NF;
            if ( f >= F )
                continue;
            const int d = B.op_b[o] * F + f;
            const uint32_t blk = tab[d];
            const u32x4* src = reinterpret_cast<const u32x4*>( nib + (size_t)d * 32 + 8 * q );
            u32x4 x0 = src[0], x1 = src[1];
            if ( q >= 2 )
            {
                x0 ^= blk;
                x1 ^= blk;
            }
            u32x4* dst = reinterpret_cast<u32x4*>( ring + ( ( o & ( nslots - 1 ) ) * NF + f ) * 32 + 8 * q );
            dst[0] = x0;
            dst[1] = x1;
        }
    };
    auto sob_op = [&]( int o, int f ) __attribute__( ( always_inline ) ) -> uint32_t
    {
        const int base = ( ( o & ( nslots - 1 ) ) * NF + f ) * 32;
        return rlo[base] ^ rhi[base];
    };
    // per-step tables by scalar loads at wave-uniform indices
    auto so = [&]( int i ) __attribute__( ( always_inline ) )
    { return __builtin_amdgcn_readfirstlane( as_const( B.step_ops )[min( i, m )] ); };
    // leverage layers G at a time into a ring of 2G per equity (layer i in
    // slot (i - 1) mod 2G), every equity's group one contiguous run
    const float lev_top = __uint_as_float( __float_as_uint( (float)( P - 1 ) ) - 1u );
    const int nd = B.n_dates;
    auto stage_group = [&]( int first ) __attribute__( ( always_inline ) )
    {
        const int run = ( min( first + G - 1, m ) - first + 1 ) * P;
        const int q0 = ( first - 1 ) % ( 2 * G );
        for ( int e0 = (int)threadIdx.x; e0 < run; e0 += kThreads )
        {
            float v[NA];
#pragma unroll
            for ( int e = 0; e < NA; e++ )
                v[e] = B.levf[( (size_t)e * nd + first ) * P + e0];
#pragma unroll
            for ( int e = 0; e < NA; e++ )
                levs[( e * 2 * G + q0 ) * P + e0] = v[e];
        }
    };
    {
        // T_B[d] = shift[d] ^ XOR of V[d][k] over the bits k of gray(256 bq)
        const uint32_t gb = gray32( bq * 256u );
        const int dims = min( m * F, kSobolDimsDev );
        for ( int d = threadIdx.x; d < dims; d += kThreads )
        {
            const uint32_t* vd = V + (size_t)d * 32;
            uint32_t t = sh ? sh[d] : 0u;
            for ( uint32_t g = gb; g; g &= g - 1 )
                t ^= vd[__builtin_ctz( g )];
            tab[d] = t;
        }
    }
    __syncthreads();
    stage_ops( 0, so( G + 1 ) );
    if ( m >= 1 )
        stage_group( 1 );
    __syncthreads();

    // ---- this wave's two factors of the bridge: the equity's spot noise and
    // its variance noise (in-order walk of the reference's bridge tree)
    const int fv = __builtin_amdgcn_readfirstlane( B.asset[a].var_factor );
    const int fac[2] = { a, fv };
    typedef float stack2v __attribute__( ( ext_vector_type( 2 * DEPTH ) ) );
    stack2v stk = (stack2v)( 0.0f );
    float top[2], prev[2] = { 0.0f, 0.0f };
#pragma unroll
    for ( int k = 0; k < 2; k++ )
    {
        const uint32_t* t = nib + (size_t)fac[k] * 32;
        top[k] = B.sd0f * normal_f32( tab[fac[k]] ^ t[gln & 15u] ^ t[16 + ( gln >> 4 )] );
    }
    int sp = 1;
    auto normals2 = [&]( int o, float* zf ) __attribute__( ( always_inline ) )
    {
        bool t0, t1;
        zf[0] = normal_f32_central( sob_op( o, fac[0] ), t0 );
        zf[1] = normal_f32_central( sob_op( o, fac[1] ), t1 );
        if ( __builtin_amdgcn_ballot_w64( t0 || t1 ) )
        {
            if ( t0 )
                zf[0] = normal_f32( sob_op( o, fac[0] ) );
            if ( t1 )
                zf[1] = normal_f32( sob_op( o, fac[1] ) );
        }
    };
    // bridge step p: build the pending points, pop W(t_{p+1}) as the
    // normalised increments of the two factors (path_generator.cpp:175-196)
    auto bridge_step = [&]( int p, float* nz ) __attribute__( ( always_inline ) )
    {
        const int i = p + 1;
        const int o0 = so( p ), o1 = so( p + 1 );
        if ( o1 - o0 == 1 && sp >= 1 )
        {
            const int o = o0;
            const float lw = as_const( B.op_wf )[3 * o], rw = as_const( B.op_wf )[3 * o + 1],
                        sd = as_const( B.op_wf )[3 * o + 2];
            float zf[2];
            normals2( o, zf );
#pragma unroll
            for ( int k = 0; k < 2; k++ )
            {
                const float w = lw * prev[k] + rw * top[k] + sd * zf[k];
                nz[k] = ( w - prev[k] ) * as_const( B.binv )[i];
                prev[k] = w;
            }
            return;
        }
        for ( int o = o0; o < o1; o++ )
        {
            const int pos = __builtin_amdgcn_readfirstlane( sp - 1 );
            const float lw = as_const( B.op_wf )[3 * o], rw = as_const( B.op_wf )[3 * o + 1],
                        sd = as_const( B.op_wf )[3 * o + 2];
            float zf[2];
            normals2( o, zf );
#pragma unroll
            for ( int k = 0; k < 2; k++ )
            {
                const float nv = lw * prev[k] + rw * top[k] + sd * zf[k];
                stk[pos + k * DEPTH] = top[k];
                top[k] = nv;
            }
            sp++;
        }
        const int pos = __builtin_amdgcn_readfirstlane( sp >= 2 ? sp - 2 : 0 );
#pragma unroll
        for ( int k = 0; k < 2; k++ )
        {
            const float w = top[k];
            nz[k] = ( w - prev[k] ) * as_const( B.binv )[i];
            prev[k] = w;
            if ( sp >= 2 )
                top[k] = stk[pos + k * DEPTH];
        }
        sp--;
    };

    // ---- the equity: log-spot and variance
    float x = (float)log( B.asset[a].spot );
    float v = (float)B.asset[a].v0;
    double book = 0.0;  // wave 0: the path's discounted book value

    // ---- control-variate samples (CV 1): this wave's equity's features and
    // discounted rebased martingale increments, wave 0 the shared entries
    const int64_t pidx = (int64_t)idx - (int64_t)first_idx;
    const bool cv_store = cv_samples && valid && pidx >= 0 && pidx < B.cv_N;
    float dmacc = 0.0f;
    auto cv_date = [&]( int i ) __attribute__( ( always_inline ) )
    {
        const int je = as_const( B.cv_end )[i];
        if ( je >= 0 )
        {
            if ( cv_store )
            {
                float* dst = B.cv_dm + ( ( (size_t)z * B.cv_J + je ) * B.cv_N + pidx ) * 8;
                dst[a] = dmacc;
                if ( a == 0 )
                    for ( int e = NA; e < 8; e++ )
                        dst[e] = 0.0f;
            }
            dmacc = 0.0f;
        }
        const int js = as_const( B.cv_start )[i];
        if ( js >= 0 && cv_store )
        {
            uint16_t* row = B.cv_feat + ( ( (size_t)z * B.cv_J + js ) * B.cv_N + pidx ) * 16;
            row[a] = (uint16_t)to_bf16( ( x - B.cv_lnS0[a] ) * B.cv_ln_scale[a] );
            row[5 + a] = (uint16_t)to_bf16(
                2.0f * fmaf( __builtin_amdgcn_sqrtf( fmaxf( v, 0.0f ) ), B.cv_vol_scale[a], -1.0f ) );
            if ( a == 0 )
            {
                for ( int e = NA; e < 5; e++ )
                    row[e] = row[5 + e] = 0;
                row[10] = (uint16_t)to_bf16( fmaf( 2.0f * (float)as_const( B.t )[i], B.cv_inv_T, -1.0f ) );
                for ( int e = 11; e < 16; e++ )
                    row[e] = 0;
            }
        }
    };
    // ---- the contracts paying at date i: the spots meet in LDS, wave 0
    // values the vanillas (finance.cpp:9-32, rainbow_node.cpp:14-27,
    // contract_node.cpp:33-47)
    auto pay = [&]( int i ) __attribute__( ( always_inline ) )
    {
        uint32_t fl = as_const( B.flow_mask )[i];
        if ( !fl )
            return;
        xs[a * 64 + lane] = __expf( x );
        __syncthreads();
        if ( a == 0 )
            while ( fl )
            {
                const int c = __builtin_ctz( fl );
                fl &= fl - 1;
                const int u = B.c_udl[c];
                float s;
                if ( u >= 0 )
                    s = xs[u * 64 + lane];
                else
                {
                    const int r = -1 - u;
                    const bool best = B.rb_best[r] != 0;
                    s = best ? -INFINITY : INFINITY;
                    for ( int j = 0; j < B.rb_n[r]; j++ )
                    {
                        const float perf = xs[B.rb_member[r][j] * 64 + lane] * (float)B.rb_weight[r][j];
                        s = best ? ( s < perf ? perf : s ) : ( perf < s ? perf : s );
                    }
                }
                const float K = (float)B.c_strike[c];
                float vanilla = B.c_call[c] ? ( s - K ) : ( K - s );
                vanilla = vanilla < 0.0f ? 0.0f : vanilla;
                double y = (double)vanilla * B.c_df[c];
                if ( !valid )
                    y = 0.0;
                book += y;
                const double s1 = wave_sum<double>( y );
                const double s2 = wave_sum<double>( y * y );
                if ( lane == 0 )
                {
                    acc[1 + c][0] += s1;
                    acc[1 + c][1] += s2;
                }
            }
        __syncthreads();
    };
    if constexpr ( cv_samples )
        cv_date( 0 );
    pay( 0 );

    // the equity's step constants of step i, loaded one step ahead: the
    // table (dates x equities x 16 floats) is larger than the scalar cache,
    // so a load at the step that reads it would wait on L2 every step
    struct StepConsts
    {
        float sc[SC_COUNT];
        float inv_step, offset, cv_scale;
    };
    auto step_consts = [&]( int i ) __attribute__( ( always_inline ) )
    {
        StepConsts c;
        const size_t lf = (size_t)min( i, m ) * NA + a;
        cptr<float> sc = as_const( B.sc ) + lf * SC_COUNT;
#pragma unroll
        for ( int k = 0; k < SC_COUNT; k++ )
            c.sc[k] = sc[k];
        c.inv_step = as_const( B.lev_inv_step )[lf];
        c.offset = as_const( B.lev_offset_f )[lf];
        c.cv_scale = cv_samples ? as_const( B.cv_scale )[lf] : 0.0f;
        return c;
    };
    StepConsts cur = step_consts( 1 );

    float nz[2];
    if ( m >= 1 )
        bridge_step( 0, nz );
    for ( int p = 0; p < m; p++ )
    {
        const int i = p + 1;
        if ( ( i - 1 ) % G == 0 )
        {
            __syncthreads();
            if ( i + G <= m )
                stage_group( i + G );
            stage_ops( so( i + G ), so( i + 2 * G ) );
        }
        const StepConsts nxt = step_consts( i + 1 );
        const float* sc = cur.sc;
        // QE variance draw (quadratic branch for every lane, the exponential
        // lanes redone), heston_variance_node.cpp:17-57
        const QeMoments q = qe_moments( v, sc );
        QeDraw dq;
        float vc = qe_quadratic( q, nz[1], cv_samples ? &dq : nullptr );
        if ( __builtin_amdgcn_ballot_w64( !q.quad ) )
            if ( !q.quad )
                vc = qe_exponential( q, nz[1], cv_samples ? &dq : nullptr );
        // leverage at the previous spot (local_volatility_node.cpp:19-47)
        const float* lv = levs + ( a * 2 * G + ( i - 1 ) % ( 2 * G ) ) * P;
        const float f = x * cur.inv_step - cur.offset;
        const float fc = __builtin_amdgcn_fmed3f( f, 0.0f, lev_top );
        const int k = (int)fc;
        const float w = fc - (float)k;
        const float L = fmaf( w, lv[k + 1] - lv[k], lv[k] );
        // Andersen leveraged log-spot step (heston_spot_node.cpp:60-75)
        const float vp = v;
        const float sv = vp + vc;
        const float u = fmaf( sc[SC_RX], vc - vp, fmaf( sc[SC_A1], sv, sc[SC_K0] ) );
        const float l2s = L * L * sv;
        const float det = fmaf( -sc[SC_A2], l2s, fmaf( L, u, sc[SC_CARRY] ) );
        const float bvol = __builtin_amdgcn_sqrtf( fmaxf( sc[SC_K3] * l2s, 0.0f ) );
        const float zs = nz[0];
        if constexpr ( cv_samples )
        {
            // first-order Doob increment of the spot, zero conditional mean
            const float sold = __expf( x );
            const float A = L * fmaf( -L, sc[SC_CVB], sc[SC_CVA] );
            const float M = fmaf( A, vc - dq.mean, bvol * zs ) * sold;
            dmacc = fmaf( cur.cv_scale, M, dmacc );
        }
        x += det + bvol * zs;
        v = vc;
        cur = nxt;
        if constexpr ( cv_samples )
            cv_date( i );
        pay( i );
        if ( i < m )
            bridge_step( i, nz );
    }
    if constexpr ( cv_samples )
        if ( cv_store && a == 0 )
            B.cv_pay[(size_t)z * B.cv_N + pidx] = (float)book;

    // ---- book value reduction (wave 0), per (replicate, tile) slot
    if ( a == 0 )
    {
        if ( !valid )
            book = 0.0;
        const double s1 = wave_sum<double>( book );
        const double s2 = wave_sum<double>( book * book );
        if ( lane == 0 )
        {
            acc[0][0] = s1;
            acc[0][1] = s2;
        }
    }
    __syncthreads();
    const int ne = ( 1 + nc ) * 2;
    if ( (int)threadIdx.x < ne )
        B.partial[( (size_t)z * gridDim.x + blockIdx.x ) * ne + threadIdx.x] = ( &acc[0][0] )[threadIdx.x];
}

}  // namespace thoth
