// mcl_mlp.hpp -- the hedge-ratio MLP on CDNA4 matrix cores (no reference
// counterpart: the reference engine has no control variates).
//
//   f(x) = W4 s(W3 s(W2 s(W1 x + b1) + b2) + b3) + b4,   s = SiLU (or the smoothstep-gated linear unit),
//   x: 11 features (5 log-moneyness, 5 scaled vols, t/T) padded to 16,
//   hidden width 64, 5 outputs (hedge ratios) padded to 16.
//
// Layers are evaluated transposed, Z^T = W^T H^T, with v_mfma_f32_16x16x16_bf16
// (BF16 operands, FP32 accumulation) on tiles of 16 samples.  Operand layouts
// (wave64): A[i][k] -> lane i + 16*(k/4), slot k%4; B[k][j] -> lane j + 16*(k/4),
// slot k%4; D[i][j] -> lane j + 16*(i/4), slot i%4.  The D tile of one layer
// (unit x sample) therefore IS the B operand of the next layer: activations
// stay in registers through the whole chain, only the weights (A operands)
// are read, from a pre-swizzled BF16 image.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace thoth
{

constexpr int kMlpIn = 11, kMlpHidden = 64, kMlpOut = 5;
// A-operand tile counts: L1 4x1, L2 4x4, L3 4x4, L4 1x4 (out-tiles x k-tiles)
constexpr int kMlpTiles = 4 + 16 + 16 + 4;
constexpr int kMlpTileOff[4] = { 0, 4, 20, 36 };
constexpr int kMlpKTiles[4] = { 1, 4, 4, 4 };
constexpr int kMlpOTiles[4] = { 4, 4, 4, 1 };
// FP32 master weights, W^T row-major per layer [out_pad][in_pad]
constexpr int kMlpWOff[4] = { 0, 64 * 16, 64 * 16 + 64 * 64, 64 * 16 + 2 * 64 * 64 };
constexpr int kMlpWSize = 64 * 16 + 2 * 64 * 64 + 16 * 64;  // 10240
constexpr int kMlpBOff[4] = { 0, 64, 128, 192 };
constexpr int kMlpBSize = 64 * 3 + 16;                      // 208
constexpr int kMlpRows[4] = { 64, 64, 64, 16 };
constexpr int kMlpCols[4] = { 16, 64, 64, 64 };

typedef short bf16x4 __attribute__( ( ext_vector_type( 4 ) ) );
typedef float f32x4 __attribute__( ( ext_vector_type( 4 ) ) );

__device__ __forceinline__ f32x4 mfma16( bf16x4 a, bf16x4 b, f32x4 c )
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k( a, b, c, 0, 0, 0 );
}

// FP32 -> BF16, round to nearest even
__device__ __forceinline__ short to_bf16( float x )
{
    uint32_t u = __float_as_uint( x );
    u += 0x7fffu + ( ( u >> 16 ) & 1u );
    return (short)( u >> 16 );
}

__device__ __forceinline__ float from_bf16( short h ) { return __uint_as_float( ( (uint32_t)(uint16_t)h ) << 16 ); }

__device__ __forceinline__ bf16x4 bits_bf16x4( uint2 w )
{
    bf16x4 r;
    r[0] = (short)( w.x & 0xffffu );
    r[1] = (short)( w.x >> 16 );
    r[2] = (short)( w.y & 0xffffu );
    r[3] = (short)( w.y >> 16 );
    return r;
}

typedef float f32x2_t __attribute__( ( ext_vector_type( 2 ) ) );
typedef __bf16 bf16x2_t __attribute__( ( ext_vector_type( 2 ) ) );
typedef uint32_t u32x2_t __attribute__( ( ext_vector_type( 2 ) ) );

// two FP32 -> packed BF16 (a in the low half), round to nearest even: one
// v_cvt_pk_bf16_f32 on gfx950, the same result as to_bf16 for finite values
__device__ __forceinline__ uint32_t pack2_bf16( float a, float b )
{
    return __builtin_bit_cast( uint32_t, __builtin_convertvector( f32x2_t{ a, b }, bf16x2_t ) );
}

__device__ __forceinline__ bf16x4 pack_bf16( f32x4 v )
{
    return __builtin_bit_cast( bf16x4, u32x2_t{ pack2_bf16( v[0], v[1] ), pack2_bf16( v[2], v[3] ) } );
}

// Smoothstep-gated linear unit, the hidden activation:
//   a(z) = z s(t),  t = clamp(z / 6 + 1/2, 0, 1),  s(t) = t^2 (3 - 2t),
// C1, 0 for z <= -3, z for z >= 3, a'(0) = 1/2 as SiLU's; five VALU
// operations and no transcendental (SiLU needs an exp and a reciprocal).
//   a'(z) = s(t) + z t (1 - t)   (s'(t) = 6 t (1 - t), dt/dz = 1/6)
__device__ __forceinline__ float gate_t( float z ) { return __builtin_amdgcn_fmed3f( fmaf( z, 1.0f / 6.0f, 0.5f ), 0.0f, 1.0f ); }

__device__ __forceinline__ float gate_act( float z )
{
    const float t = gate_t( z );
    return z * ( t * t * fmaf( -2.0f, t, 3.0f ) );
}

__device__ __forceinline__ float gate_dact( float z )
{
    const float t = gate_t( z );
    return fmaf( z * t, 1.0f - t, t * t * fmaf( -2.0f, t, 3.0f ) );
}

// SiLU a(z) = z / (1 + e^-z), a'(z) = s (1 + z (1 - s)) with s = 1 / (1 + e^-z):
// one exponential and one reciprocal per unit (e^-z = inf for z < -88 gives
// s = 0, a = -0, a' = 0)
__device__ __forceinline__ float silu_s( float z ) { return __builtin_amdgcn_rcpf( 1.0f + __expf( -z ) ); }
__device__ __forceinline__ float silu_act( float z ) { return z * silu_s( z ); }
__device__ __forceinline__ float silu_dact( float z )
{
    const float s = silu_s( z );
    return fmaf( z * s, 1.0f - s, s );
}

// hidden activation of a network: ACT = kActSilu or kActGate
enum
{
    kActSilu = 0,
    kActGate = 1
};
template <int ACT>
__device__ __forceinline__ float act_f( float z )
{
    if constexpr ( ACT == kActGate )
        return gate_act( z );
    else
        return silu_act( z );
}
template <int ACT>
__device__ __forceinline__ float dact_f( float z )
{
    if constexpr ( ACT == kActGate )
        return gate_dact( z );
    else
        return silu_dact( z );
}
// a(c) of a D-layout tile packed as the next layer's BF16 B operand
template <int ACT>
__device__ __forceinline__ bf16x4 act_pack( f32x4 c )
{
    return pack_bf16( f32x4{ act_f<ACT>( c[0] ), act_f<ACT>( c[1] ), act_f<ACT>( c[2] ), act_f<ACT>( c[3] ) } );
}
template <int ACT>
__device__ __forceinline__ f32x4 dact_mul( f32x4 c, f32x4 z )
{
    return f32x4{ c[0] * dact_f<ACT>( z[0] ), c[1] * dact_f<ACT>( z[1] ), c[2] * dact_f<ACT>( z[2] ),
                  c[3] * dact_f<ACT>( z[3] ) };
}

// A operand of tile (layer L, out-tile ot, k-tile kt) from a swizzled image:
// image[(tile * 64 + lane)] holds the lane's 4 BF16 values.
__device__ __forceinline__ bf16x4 weight_frag( const uint2* img, int L, int ot, int kt, int lane )
{
    const uint2 w = img[( kMlpTileOff[L] + ot * kMlpKTiles[L] + kt ) * 64 + lane];
    bf16x4 r;
    r[0] = (short)( w.x & 0xffffu );
    r[1] = (short)( w.x >> 16 );
    r[2] = (short)( w.y & 0xffffu );
    r[3] = (short)( w.y >> 16 );
    return r;
}

// bias fragment of out-tile ot: units 16 ot + 4 (lane/16) + r
__device__ __forceinline__ f32x4 bias_frag( const float* bias, int L, int ot, int lane )
{
    const float* b = bias + kMlpBOff[L] + ot * 16 + 4 * ( lane >> 4 );
    return f32x4{ b[0], b[1], b[2], b[3] };
}

// Activations of one 16-sample tile kept for the backward pass.
struct MlpTile
{
    f32x4 z[3][4];   // pre-activations of the hidden layers (D layout)
    bf16x4 h[3][4];  // a(z) as BF16 B operands
    f32x4 out;       // outputs 0..15 (D layout; 0..4 meaningful)
};

// forward of one tile; x: B operand of the padded features (lane: features
// 4 (lane/16) .. +3 of sample lane % 16)
template <int ACT>
__device__ __forceinline__ void mlp_forward( const uint2* wimg, const float* bias, bf16x4 x, int lane, MlpTile& t )
{
#pragma unroll
    for ( int ot = 0; ot < 4; ot++ )
    {
        f32x4 c = bias_frag( bias, 0, ot, lane );
        c = mfma16( weight_frag( wimg, 0, ot, 0, lane ), x, c );
        t.z[0][ot] = c;
        t.h[0][ot] = act_pack<ACT>( c );
    }
#pragma unroll
    for ( int L = 1; L < 3; L++ )
    {
#pragma unroll
        for ( int ot = 0; ot < 4; ot++ )
        {
            f32x4 c = bias_frag( bias, L, ot, lane );
#pragma unroll
            for ( int kt = 0; kt < 4; kt++ )
                c = mfma16( weight_frag( wimg, L, ot, kt, lane ), t.h[L - 1][kt], c );
            t.z[L][ot] = c;
            t.h[L][ot] = act_pack<ACT>( c );
        }
    }
    f32x4 c = bias_frag( bias, 3, 0, lane );
#pragma unroll
    for ( int kt = 0; kt < 4; kt++ )
        c = mfma16( weight_frag( wimg, 3, 0, kt, lane ), t.h[2][kt], c );
    t.out = c;
}

// forward of one tile keeping nothing for a backward pass: the output tile
template <int ACT>
__device__ __forceinline__ f32x4 mlp_infer( const uint2* wimg, const float* bias, bf16x4 x, int lane )
{
    bf16x4 h0, h1, h2, h3;
    {
        f32x4 c0 = mfma16( weight_frag( wimg, 0, 0, 0, lane ), x, bias_frag( bias, 0, 0, lane ) );
        f32x4 c1 = mfma16( weight_frag( wimg, 0, 1, 0, lane ), x, bias_frag( bias, 0, 1, lane ) );
        f32x4 c2 = mfma16( weight_frag( wimg, 0, 2, 0, lane ), x, bias_frag( bias, 0, 2, lane ) );
        f32x4 c3 = mfma16( weight_frag( wimg, 0, 3, 0, lane ), x, bias_frag( bias, 0, 3, lane ) );
        h0 = act_pack<ACT>( c0 );
        h1 = act_pack<ACT>( c1 );
        h2 = act_pack<ACT>( c2 );
        h3 = act_pack<ACT>( c3 );
    }
    // hidden layers 2 and 3: one output tile at a time (fewest live registers)
    auto unit_tile = [&]( int L, int ot ) __attribute__( ( always_inline ) ) -> bf16x4
    {
        f32x4 c = bias_frag( bias, L, ot, lane );
        c = mfma16( weight_frag( wimg, L, ot, 0, lane ), h0, c );
        c = mfma16( weight_frag( wimg, L, ot, 1, lane ), h1, c );
        c = mfma16( weight_frag( wimg, L, ot, 2, lane ), h2, c );
        c = mfma16( weight_frag( wimg, L, ot, 3, lane ), h3, c );
        return act_pack<ACT>( c );
    };
#pragma unroll
    for ( int L = 1; L < 3; L++ )
    {
        const bf16x4 n0 = unit_tile( L, 0 ), n1 = unit_tile( L, 1 ), n2 = unit_tile( L, 2 ), n3 = unit_tile( L, 3 );
        h0 = n0;
        h1 = n1;
        h2 = n2;
        h3 = n3;
    }
    f32x4 o = bias_frag( bias, 3, 0, lane );
    o = mfma16( weight_frag( wimg, 3, 0, 0, lane ), h0, o );
    o = mfma16( weight_frag( wimg, 3, 0, 1, lane ), h1, o );
    o = mfma16( weight_frag( wimg, 3, 0, 2, lane ), h2, o );
    o = mfma16( weight_frag( wimg, 3, 0, 3, lane ), h3, o );
    return o;
}

// hedge-ratio . increments of the tile's 16 samples: lane l (l < 16) returns
// sum_a f_a dm_a of sample l; dm4: this lane's 4 increments (units 4 (l/16)+r,
// zero beyond output 4)
__device__ __forceinline__ float mlp_dot( const f32x4& out, const f32x4& dm4 )
{
    float p = out[0] * dm4[0] + out[1] * dm4[1] + out[2] * dm4[2] + out[3] * dm4[3];
    p += __shfl_xor( p, 16, 64 );
    p += __shfl_xor( p, 32, 64 );
    return p;
}

}  // namespace thoth

namespace thoth
{

// A operand of the backward (non-transposed) weight image: tile (in-tile it,
// out-tile ot) of W_L [in][out]: lane i + 16 (k/4) holds W[16 it + i][16 ot + k]
__device__ __forceinline__ bf16x4 weight_frag_bwd( const uint2* img, int L, int it, int ot, int lane )
{
    const uint2 w = img[( kMlpTileOff[L] + it * kMlpOTiles[L] + ot ) * 64 + lane];
    bf16x4 r;
    r[0] = (short)( w.x & 0xffffu );
    r[1] = (short)( w.x >> 16 );
    r[2] = (short)( w.y & 0xffffu );
    r[3] = (short)( w.y >> 16 );
    return r;
}

// 16x16 tile in D layout (row 4 (l/16) + r, column l % 16) -> the same tile in
// A/B-operand layout along its columns (lane l: row l % 16, columns
// 4 (l/16) .. +3), through a per-wave 512-byte LDS buffer.
__device__ __forceinline__ bf16x4 transpose_tile( short* buf, bf16x4 v, int lane )
{
    const int row0 = 4 * ( lane >> 4 ), col = lane & 15;
#pragma unroll
    for ( int r = 0; r < 4; r++ )
        buf[( row0 + r ) * 16 + col] = v[r];
    __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
    const int base = col * 16 + 4 * ( lane >> 4 );
    bf16x4 out;
#pragma unroll
    for ( int r = 0; r < 4; r++ )
        out[r] = buf[base + r];
    __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
    return out;
}

// Gradient accumulators of one wave, D layout of the dW^T tiles.
struct MlpGrad
{
    f32x4 w[kMlpTiles];  // tile (L, ot, it) at kMlpTileOff[L] + ot * kMlpKTiles[L] + it
    f32x4 b[13];         // bias partial sums per out-tile: L0..2 4 tiles each, L3 1
};

// Backward of one tile given dOut = dL/dF^T (D layout, units x samples):
//   dW_L^T += dZ_L^T H_{L-1}   (MFMA over the 16 samples of the tile, both
//                               operands transposed through LDS),
//   db_L   += row sums of dZ_L^T (per-lane partials, reduced by the caller),
//   dH_{L-1}^T = W_L dZ_L^T,  dZ_{L-1} = dH_{L-1} * a'(Z_{L-1}).
template <int ACT>
__device__ __forceinline__ void mlp_backward( const uint2* wbwd, const MlpTile& t, bf16x4 x, f32x4 dout, short* buf,
                                              int lane, MlpGrad& g )
{
    constexpr int kBiasTile[4] = { 0, 4, 8, 12 };
    f32x4 dz[4];
    dz[0] = dout;
#pragma unroll
    for ( int L = 3; L >= 0; L-- )
    {
        const int no = kMlpOTiles[L], ni = kMlpKTiles[L];
        // weight and bias gradients of layer L
#pragma unroll
        for ( int ot = 0; ot < no; ot++ )
        {
            const bf16x4 a = transpose_tile( buf, pack_bf16( dz[ot] ), lane );
#pragma unroll
            for ( int it = 0; it < ni; it++ )
            {
                const bf16x4 hin = ( L == 0 ) ? x : t.h[L - 1][it];
                const int k = kMlpTileOff[L] + ot * ni + it;
                g.w[k] = mfma16( a, transpose_tile( buf, hin, lane ), g.w[k] );
            }
            g.b[kBiasTile[L] + ot] += dz[ot];
        }
        if ( L == 0 )
            break;
        // propagate to the previous layer's pre-activations
        f32x4 nd[4];
#pragma unroll
        for ( int it = 0; it < ni; it++ )
        {
            f32x4 c = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
            for ( int ot = 0; ot < no; ot++ )
                c = mfma16( weight_frag_bwd( wbwd, L, it, ot, lane ), pack_bf16( dz[ot] ), c );
            const f32x4 z = t.z[L - 1][it];
            nd[it] = dact_mul<ACT>( c, z );
        }
#pragma unroll
        for ( int it = 0; it < 4; it++ )
            dz[it] = nd[it];
    }
}

}  // namespace thoth

namespace thoth
{

// Per-wave LDS copy of one tile's activations for the backward pass.  z is
// kept in D layout (lane-contiguous 16 bytes); the layer inputs (features and
// hidden activations, BF16) are kept as row-major [unit][sample] 16x16 tiles,
// which is exactly the B operand of the weight-gradient product
// dW^T += dZ^T H (lane j + 16 (k/4) reads units j, samples 4 (k/4) .. +3 as
// one 8-byte word).
struct MlpActLds
{
    f32x4 z[12][64];       // pre-activations, hidden layers 0..2 x out-tiles 0..3
    short in[13][256];     // inputs of layers: [0] features, [1 + 4 (L-1) + t] hidden tile t of layer L-1
    short tbuf[256];       // transpose staging for the errors
};

// store a D-layout BF16 tile (lane = sample + 16 (unit / 4), slot unit % 4) as
// row-major [unit][sample]
__device__ __forceinline__ void store_rows( short* dst, bf16x4 v, int lane )
{
    const int row0 = 4 * ( lane >> 4 ), col = lane & 15;
#pragma unroll
    for ( int r = 0; r < 4; r++ )
        dst[( row0 + r ) * 16 + col] = v[r];
}

// B operand of dW: lane l gets row l % 16, columns 4 (l/16) .. +3
__device__ __forceinline__ bf16x4 load_cols( const short* src, int lane )
{
    const uint2 w = *reinterpret_cast<const uint2*>( src + ( lane & 15 ) * 16 + 4 * ( lane >> 4 ) );
    bf16x4 r;
    r[0] = (short)( w.x & 0xffffu );
    r[1] = (short)( w.x >> 16 );
    r[2] = (short)( w.y & 0xffffu );
    r[3] = (short)( w.y >> 16 );
    return r;
}

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
}

// forward of one tile keeping what the backward pass needs in the wave's LDS
// block; returns the output tile (D layout)
template <int ACT>
__device__ __forceinline__ f32x4 mlp_forward_lds( const uint2* wimg, const float* bias, bf16x4 x, int lane,
                                                  MlpActLds& act )
{
    // the features arrive in B layout (lane: features 4 (l/16) .. +3 of sample
    // l % 16), which is the D layout of the [feature][sample] tile
    store_rows( act.in[0], x, lane );
    bf16x4 h[4];
#pragma unroll
    for ( int ot = 0; ot < 4; ot++ )
    {
        f32x4 c = bias_frag( bias, 0, ot, lane );
        c = mfma16( weight_frag( wimg, 0, ot, 0, lane ), x, c );
        act.z[ot][lane] = c;
        h[ot] = act_pack<ACT>( c );
        store_rows( act.in[1 + ot], h[ot], lane );
    }
#pragma unroll
    for ( int L = 1; L < 3; L++ )
    {
        bf16x4 hn[4];
#pragma unroll
        for ( int ot = 0; ot < 4; ot++ )
        {
            f32x4 c = bias_frag( bias, L, ot, lane );
#pragma unroll
            for ( int kt = 0; kt < 4; kt++ )
                c = mfma16( weight_frag( wimg, L, ot, kt, lane ), h[kt], c );
            act.z[4 * L + ot][lane] = c;
            hn[ot] = act_pack<ACT>( c );
            store_rows( act.in[1 + 4 * L + ot], hn[ot], lane );
        }
#pragma unroll
        for ( int ot = 0; ot < 4; ot++ )
            h[ot] = hn[ot];
    }
    f32x4 c = bias_frag( bias, 3, 0, lane );
#pragma unroll
    for ( int kt = 0; kt < 4; kt++ )
        c = mfma16( weight_frag( wimg, 3, 0, kt, lane ), h[kt], c );
    wave_lds_sync();
    return c;
}

// mlp_backward with the activations read back from the wave's LDS block
template <int ACT>
__device__ __forceinline__ void mlp_backward_lds( const uint2* wbwd, MlpActLds& act, f32x4 dout, int lane, MlpGrad& g )
{
    constexpr int kBiasTile[4] = { 0, 4, 8, 12 };
    f32x4 dz[4];
    dz[0] = dout;
#pragma unroll
    for ( int L = 3; L >= 0; L-- )
    {
        const int no = kMlpOTiles[L], ni = kMlpKTiles[L];
        bf16x4 hin[4];
#pragma unroll
        for ( int it = 0; it < ni; it++ )
            hin[it] = load_cols( act.in[L == 0 ? 0 : 1 + 4 * ( L - 1 ) + it], lane );
#pragma unroll
        for ( int ot = 0; ot < no; ot++ )
        {
            const bf16x4 a = transpose_tile( act.tbuf, pack_bf16( dz[ot] ), lane );
#pragma unroll
            for ( int it = 0; it < ni; it++ )
            {
                const int k = kMlpTileOff[L] + ot * ni + it;
                g.w[k] = mfma16( a, hin[it], g.w[k] );
            }
            g.b[kBiasTile[L] + ot] += dz[ot];
        }
        if ( L == 0 )
            break;
        f32x4 nd[4];
#pragma unroll
        for ( int it = 0; it < ni; it++ )
        {
            f32x4 c = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
            for ( int ot = 0; ot < no; ot++ )
                c = mfma16( weight_frag_bwd( wbwd, L, it, ot, lane ), pack_bf16( dz[ot] ), c );
            const f32x4 z = act.z[4 * ( L - 1 ) + it][lane];
            nd[it] = dact_mul<ACT>( c, z );
        }
#pragma unroll
        for ( int it = 0; it < 4; it++ )
            dz[it] = nd[it];
    }
    wave_lds_sync();  // the next tile's forward overwrites the block
}

}  // namespace thoth
