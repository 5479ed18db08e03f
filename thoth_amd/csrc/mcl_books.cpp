// mcl_books.cpp -- thoth_mcl_sweep: validates thoth_mcl_book inputs, packs the
// device books (bridge program, FP32 step constants, payoff masks), launches
// the fused sweep kernel once for the whole batch and folds the per-block
// FP64 partial sums with Kahan compensation (monte_carlo_node.cpp:73-96).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "mcl_device.hpp"
#include "mcl_internal.hpp"
#include "mcl_launch.hpp"
#include "thoth_mcl_cv.h"

namespace thoth
{

namespace
{

// host image of the device arena: every array is appended 16-byte aligned and
// addressed by offset until the single upload.
struct Arena
{
    std::vector<char> data;
    template <class T>
    size_t put( const T* p, size_t n )
    {
        size_t off = ( data.size() + 15 ) & ~size_t( 15 );
        data.resize( off + n * sizeof( T ) );
        if ( n )
            std::memcpy( data.data() + off, p, n * sizeof( T ) );
        return off;
    }
    template <class T>
    size_t put( const std::vector<T>& v )
    {
        return put( v.data(), v.size() );
    }
};

struct Offsets
{
    size_t t, dt, sqrt_dt, bsq, binv, step_ops, op_b, op_left, op_w, op_wf, drift, sc, lev, levf, lnstep, offset,
        offsetf, invstep, mon, flow, cv_start, cv_end, cv_scale, grid_mat_at, grid_kf, grid_kd, chol, cholf;
    size_t pw[4];  // leverage sensitivities: ds, de (FP64), then their FP32 copies
    bool has_pw;
};

std::string Why( int b, const std::string& s ) { return "book " + std::to_string( b ) + ": " + s; }

int PackBook( int bi, const thoth_mcl_book& in, DevBook& db, Offsets& of, Arena& ar, int& need_depth,
              std::vector<int>& step_ops )
{
    const int nd = in.n_dates;
    const int na = in.n_assets;
    if ( nd < 2 || !in.t || !in.dt )
        return SetError( THOTH_ERR_INVALID, Why( bi, "need at least 2 dates and t/dt arrays" ) );
    if ( na < 1 || na > kMaxAssets )
        return SetError( THOTH_ERR_INVALID, Why( bi, "1..8 equities supported" ) );
    if ( in.n_contracts < 0 || in.n_contracts > kMaxContracts )
        return SetError( THOTH_ERR_UNSUPPORTED, Why( bi, "at most 32 contracts per book" ) );
    if ( in.n_rainbows < 0 || in.n_rainbows > kMaxRainbows )
        return SetError( THOTH_ERR_UNSUPPORTED, Why( bi, "at most 4 rainbows per book" ) );
    if ( in.t[0] != 0.0 )
        return SetError( THOTH_ERR_INVALID, Why( bi, "t[0] must be 0 (today)" ) );
    for ( int i = 1; i < nd; i++ )
        if ( !( in.t[i] >= in.t[i - 1] ) || !( in.dt[i] >= 0 ) )
            return SetError( THOTH_ERR_INVALID, Why( bi, "dates must be non-decreasing" ) );
    if ( !( in.t[nd - 1] > 0 ) )
        return SetError( THOTH_ERR_INVALID, Why( bi, "last date must be after today" ) );
    std::memset( &db, 0, sizeof( db ) );
    db.n_dates = nd;
    db.m = nd - 1;
    db.n_assets = na;
    db.n_contracts = in.n_contracts;
    db.n_rainbows = in.n_rainbows;
    db.corr_stoch = in.correlate_stochastic ? 1 : 0;
    // factors: spot noises in equity order, then the variance noises of the
    // stochastic equities (pricer_mcl.cpp:404-432)
    int f = na;
    bool any_lsv = false;
    for ( int a = 0; a < na; a++ )
    {
        DevAsset& A = db.asset[a];
        A.model = in.model[a];
        if ( A.model < THOTH_MODEL_BS || A.model > THOTH_MODEL_LSV )
            return SetError( THOTH_ERR_INVALID, Why( bi, "unknown model" ) );
        A.spot = in.spot[a];
        if ( !( A.spot > 0 ) )
            return SetError( THOTH_ERR_INVALID, Why( bi, "spot must be positive" ) );
        A.vol = in.vol ? in.vol[a] : 0.0;
        A.v0 = in.v0 ? in.v0[a] : 0.0;
        A.kappa = in.kappa ? in.kappa[a] : 0.0;
        A.theta = in.theta ? in.theta[a] : 0.0;
        A.xi = in.xi ? in.xi[a] : 0.0;
        A.rho = in.rho ? in.rho[a] : 0.0;
        A.var_factor = -1;
        if ( A.model != THOTH_MODEL_BS )
        {
            if ( !( A.kappa > 0 ) || !( A.v0 >= 0 ) || !( A.theta >= 0 ) || !( A.xi >= 0 ) )
                return SetError( THOTH_ERR_INVALID, Why( bi, "Heston parameters out of range" ) );
            A.var_factor = f++;
        }
        any_lsv |= A.model == THOTH_MODEL_LSV;
    }
    db.factors = f;
    if ( na > 1 && !in.chol && !in.chol_dates )
        return SetError( THOTH_ERR_INVALID, Why( bi, "multi-equity book needs a Cholesky factor" ) );
    of.chol = of.cholf = 0;
    if ( na > 1 )
    {
        // per-date lower factors [date][kMaxAssets][kMaxAssets], FP64 and FP32
        std::vector<double> ct( (size_t)nd * kMaxAssets * kMaxAssets, 0.0 );
        for ( int i = 0; i < nd; i++ )
        {
            const double* L = in.chol_dates ? in.chol_dates + (size_t)i * na * na : in.chol;
            for ( int a = 0; a < na; a++ )
                for ( int v = 0; v <= a; v++ )
                    ct[( (size_t)i * kMaxAssets + a ) * kMaxAssets + v] = L[a * na + v];
        }
        std::vector<float> cf( ct.begin(), ct.end() );
        of.chol = ar.put( ct );
        of.cholf = ar.put( cf );
    }
    // grid derived arrays
    std::vector<double> sqrt_dt( nd, 0.0 ), bsq( nd, 0.0 );
    std::vector<float> binv( nd, 0.0f );
    for ( int i = 1; i < nd; i++ )
    {
        sqrt_dt[i] = std::sqrt( in.dt[i] );
        const double dtb = in.t[i] - in.t[i - 1];
        bsq[i] = dtb > 0 ? std::sqrt( dtb ) : 0.0;
        binv[i] = dtb > 0 ? (float)( 1.0 / bsq[i] ) : 0.0f;
    }
    BridgeProgram bp = BuildBridgeProgram( in.t, nd );
    need_depth = std::max( need_depth, std::max( 1, bp.max_depth - 1 ) );
    if ( bp.max_depth - 1 > kMaxStackDepth )
        return SetError( THOTH_ERR_UNSUPPORTED, Why( bi, "too many steps for the bridge stack" ) );
    db.sd0 = bp.sd0;
    db.sd0f = (float)bp.sd0;
    db.n_ops = (int)bp.op_b.size();
    // FP32 weights: the left weight of an op without a left anchor is 0, so the
    // FP32 kernels form lw W_left + rw W_right + sd z without a select
    // (0 W_left adds an exact zero)
    std::vector<float> op_wf( bp.op_w.begin(), bp.op_w.end() );
    for ( size_t o = 0; o < bp.op_left.size(); o++ )
        if ( !bp.op_left[o] )
            op_wf[3 * o] = 0.0f;
    of.t = ar.put( in.t, nd );
    of.dt = ar.put( in.dt, nd );
    of.sqrt_dt = ar.put( sqrt_dt );
    of.bsq = ar.put( bsq );
    of.binv = ar.put( binv );
    of.step_ops = ar.put( bp.step_ops );
    step_ops = bp.step_ops;
    of.op_b = ar.put( bp.op_b );
    of.op_left = ar.put( bp.op_left );
    of.op_w = ar.put( bp.op_w );
    of.op_wf = ar.put( op_wf );
    // drift and FP32 step constants
    if ( !in.drift )
        return SetError( THOTH_ERR_INVALID, Why( bi, "drift array required" ) );
    of.drift = ar.put( in.drift, (size_t)na * nd );
    std::vector<float> sc( (size_t)na * nd * SC_COUNT, 0.0f );
    for ( int a = 0; a < na; a++ )
    {
        const DevAsset& A = db.asset[a];
        for ( int i = 1; i < nd; i++ )
        {
            float* c = &sc[( (size_t)i * na + a ) * SC_COUNT];  // date-major: one base per step
            const double dt = in.dt[i];
            c[SC_CARRY] = (float)( in.drift[(size_t)a * nd + i] * in.t[i] - in.drift[(size_t)a * nd + i - 1] * in.t[i - 1] );
            c[SC_DT] = (float)dt;
            c[SC_SQDT] = (float)std::sqrt( dt );
            c[SC_VOL] = (float)A.vol;
            c[SC_HALF_V2DT] = (float)( A.vol * A.vol / 2 * dt );
            if ( A.model != THOTH_MODEL_BS )
            {
                const double e = std::exp( -A.kappa * dt );
                c[SC_E] = (float)e;
                c[SC_M0] = (float)( A.theta * ( 1 - e ) );
                c[SC_S1] = (float)( A.xi * A.xi * e / A.kappa * ( 1 - e ) );
                c[SC_S0] = (float)( A.theta * A.xi * A.xi / ( 2 * A.kappa ) * ( 1 - e ) * ( 1 - e ) );
                if ( A.xi > 1e-12 )
                {
                    c[SC_K0] = (float)( -A.rho * A.kappa * A.theta / A.xi * dt );
                    c[SC_A1] = (float)( 0.5 * dt * A.kappa * A.rho / A.xi );
                    c[SC_A2] = (float)( 0.5 * dt * 0.5 );
                    c[SC_RX] = (float)( A.rho / A.xi );
                    c[SC_K3] = (float)( 0.5 * dt * ( 1.0 - A.rho * A.rho ) );
                    c[SC_CVA] = (float)( 0.5 * dt * A.kappa * A.rho / A.xi + A.rho / A.xi );
                    c[SC_CVB] = (float)( 0.5 * dt * 0.5 - 0.5 * ( 0.5 * dt * ( 1.0 - A.rho * A.rho ) ) );
                }
            }
        }
    }
    of.sc = ar.put( sc );
    // leverage layers
    db.lev_points = 0;
    of.lev = of.levf = of.lnstep = of.offset = of.offsetf = of.invstep = 0;
    if ( any_lsv )
    {
        const int P = in.lev_points;
        if ( P < 2 || !in.lev_values || !in.lev_ln_step || !in.lev_offset )
            return SetError( THOTH_ERR_INVALID, Why( bi, "LSV equity needs leverage layers" ) );
        db.lev_points = P;
        const size_t nl = (size_t)na * nd;
        std::vector<float> levf( in.lev_values, in.lev_values + nl * P );
        std::vector<double> off( nl );
        std::vector<float> offf( nl ), inv( nl );  // FP32 copies date-major [i][a]
        for ( size_t k = 0; k < nl; k++ )
        {
            off[k] = (double)in.lev_offset[k];
            const size_t a = k / nd, i = k % nd, kf = i * na + a;
            offf[kf] = (float)in.lev_offset[k];
            inv[kf] = in.lev_ln_step[k] > 0 ? (float)( 1.0 / in.lev_ln_step[k] ) : 0.0f;
        }
        of.lev = ar.put( in.lev_values, nl * P );
        of.levf = ar.put( levf );
        of.lnstep = ar.put( in.lev_ln_step, nl );
        of.offset = ar.put( off );
        of.offsetf = ar.put( offf );
        of.invstep = ar.put( inv );
    }
    // rainbows (rainbow_node.cpp:14-27)
    for ( int r = 0; r < in.n_rainbows; r++ )
    {
        db.rb_best[r] = in.rb_best[r];
        db.rb_n[r] = in.rb_n[r];
        if ( db.rb_n[r] < 1 || db.rb_n[r] > kMaxAssets )
            return SetError( THOTH_ERR_INVALID, Why( bi, "rainbow size" ) );
        for ( int j = 0; j < db.rb_n[r]; j++ )
        {
            db.rb_member[r][j] = in.rb_member[r * THOTH_MCL_MAX_ASSETS + j];
            db.rb_weight[r][j] = in.rb_weight[r * THOTH_MCL_MAX_ASSETS + j];
            if ( db.rb_member[r][j] < 0 || db.rb_member[r][j] >= na )
                return SetError( THOTH_ERR_INVALID, Why( bi, "rainbow member index" ) );
        }
    }
    // contracts: flows and monitoring masks per date
    std::vector<uint32_t> mon( nd, 0u ), flow( nd, 0u );
    for ( int c = 0; c < in.n_contracts; c++ )
    {
        db.c_kind[c] = in.c_kind[c];
        db.c_call[c] = in.c_is_call[c] ? 1 : 0;
        db.c_udl[c] = in.c_underlying[c];
        db.c_strike[c] = in.c_strike[c];
        const int u = db.c_udl[c];
        if ( u >= na || u < -in.n_rainbows )
            return SetError( THOTH_ERR_INVALID, Why( bi, "contract underlying index" ) );
        const int fi = in.c_flow_index[c];
        if ( fi < 0 || fi >= nd )
            return SetError( THOTH_ERR_INVALID, Why( bi, "contract flow index" ) );
        flow[fi] |= 1u << c;
        // contract_node.cpp:33-47: discount the flow date back to today
        const double dfac = fi != 0 ? std::exp( in.t[0] * in.disc_rate[0] - in.t[fi] * in.disc_rate[fi] ) : 1.0;
        db.c_df[c] = dfac;
        if ( db.c_kind[c] == THOTH_PAYOFF_BARRIER )
        {
            db.c_up[c] = in.c_is_up[c] ? 1 : 0;
            db.c_in[c] = in.c_is_in[c] ? 1 : 0;
            db.c_level[c] = in.c_level[c];
            for ( int k = in.c_mon_offset[c]; k < in.c_mon_offset[c + 1]; k++ )
            {
                const int mi = in.c_mon_index[k];
                if ( mi < 0 || mi > fi )
                    return SetError( THOTH_ERR_INVALID, Why( bi, "monitoring date after the flow" ) );
                mon[mi] |= 1u << c;
            }
        }
        else if ( db.c_kind[c] != THOTH_PAYOFF_VANILLA )
            return SetError( THOTH_ERR_INVALID, Why( bi, "unknown payoff kind" ) );
    }
    of.mon = ar.put( mon );
    of.flow = ar.put( flow );
    return THOTH_OK;
}

// hedge dates d_j = m - round(m (1 - j/J)^p), at least one step apart, and the per-date increment scale
// df(t_i) 100 / S0_a of the learned control variate
int PackCv( int bi, const thoth_mcl_book& in, const CvSpec& cv, DevBook& db, Offsets& of, Arena& ar )
{
    const int nd = in.n_dates, m = nd - 1, na = in.n_assets;
    if ( cv.J < 1 || cv.J > m )
        return SetError( THOTH_ERR_INVALID, Why( bi, "hedge dates must be in [1, steps]" ) );
    if ( na > 5 )
        return SetError( THOTH_ERR_INVALID, Why( bi, "the hedge network covers at most 5 equities" ) );
    if ( cv.mode != 1 && cv.mode != 2 )
        return SetError( THOTH_ERR_INVALID, Why( bi, "unknown control-variate mode" ) );
    if ( cv.mode == 2 && ( !cv.wimg || !cv.bias || in.n_contracts >= kMaxContracts ) )
        return SetError( THOTH_ERR_INVALID, Why( bi, "in-loop inference needs weights and a free result slot" ) );
    std::vector<int> start( nd, -1 ), end( nd, -1 );
    if ( !( cv.spacing >= 1.0f && cv.spacing <= 8.0f ) )
        return SetError( THOTH_ERR_INVALID, Why( bi, "hedge spacing must be in [1, 8]" ) );
    const double pw = cv.spacing;
    std::vector<int> dd( cv.J + 1, 0 );
    for ( int j = 1; j <= cv.J; j++ )
    {
        const int g = m - (int)std::lround( m * std::pow( 1.0 - (double)j / cv.J, pw ) );
        dd[j] = std::min( std::max( dd[j - 1] + 1, g ), m - ( cv.J - j ) );
    }
    int prev = 0;
    for ( int j = 0; j < cv.J; j++ )
    {
        const int d0 = dd[j];
        const int d1 = dd[j + 1];
        if ( d1 <= d0 || d0 != prev )
            return SetError( THOTH_ERR_INVALID, Why( bi, "hedge dates must be increasing" ) );
        start[d0] = j;
        end[d1] = j;
        prev = d1;
    }
    std::vector<float> scale( (size_t)na * nd, 0.0f );
    for ( int a = 0; a < na; a++ )
        for ( int i = 1; i < nd; i++ )
            scale[(size_t)i * na + a] =  // date-major
                (float)( std::exp( in.t[0] * in.disc_rate[0] - in.t[i] * in.disc_rate[i] ) * 100.0 / in.spot[a] );
    db.cv_J = cv.J;
    db.cv_N = cv.N;
    db.cv_inv_T = (float)( 1.0 / in.t[m] );
    for ( int a = 0; a < na; a++ )
    {
        // reference vol: the Black-Scholes vol, else sqrt(v0)
        const double ref = in.model[a] == THOTH_MODEL_BS ? in.vol[a] : std::sqrt( std::max( in.v0[a], 1e-8 ) );
        db.cv_lnS0[a] = (float)std::log( in.spot[a] );
        db.cv_ln_scale[a] = (float)( 1.0 / ( std::max( ref, 1e-4 ) * std::sqrt( in.t[m] ) ) );
        db.cv_vol_scale[a] = (float)( 1.0 / std::max( ref, 1e-4 ) );
    }
    db.cv_wimg = reinterpret_cast<const uint32_t*>( cv.wimg );
    db.cv_bias = cv.bias;
    db.cv_feat = cv.feat;
    db.cv_dm = cv.dm;
    db.cv_pay = cv.pay;
    of.cv_start = ar.put( start );
    of.cv_end = ar.put( end );
    of.cv_scale = ar.put( scale );
    return THOTH_OK;
}

int PackGrid( int bi, const thoth_mcl_book& in, const GridSpec& gs, DevBook& db, Offsets& of, Arena& ar )
{
    const int nd = in.n_dates, m = nd - 1;
    if ( gs.nm < 1 || gs.nm > kGridMaxMat || gs.nk < 1 || gs.nk > kGridMaxStrikes )
        return SetError( THOTH_ERR_INVALID, Why( bi, "vanilla grid: 1..16 maturities x 1..64 strikes" ) );
    if ( gs.udl >= in.n_assets || ( gs.udl < 0 && -1 - gs.udl >= in.n_rainbows ) )
        return SetError( THOTH_ERR_INVALID, Why( bi, "vanilla grid underlying" ) );
    std::vector<int> at( nd, -1 );
    for ( int j = 0; j < gs.nm; j++ )
    {
        const int d = gs.mat_date[j];
        if ( d < 1 || d > m || at[d] >= 0 )
            return SetError( THOTH_ERR_INVALID, Why( bi, "vanilla grid maturities must be distinct diffusion dates" ) );
        at[d] = j;
    }
    std::vector<double> kd( gs.strikes, gs.strikes + gs.nk );
    if ( gs.grouped )
    {
        // every equity a single-equity LSV book of its own: two shared factors
        // (spot and variance noise) instead of two per equity
        for ( int a = 0; a < in.n_assets; a++ )
            if ( in.model[a] != THOTH_MODEL_LSV )
                return SetError( THOTH_ERR_INVALID, Why( bi, "grouped grid books are LSV equities" ) );
        if ( in.correlate_stochastic )
            return SetError( THOTH_ERR_INVALID, Why( bi, "grouped grid books are uncorrelated" ) );
        db.factors = 2;
        for ( int a = 0; a < in.n_assets; a++ )
            db.asset[a].var_factor = 1;
        kd.clear();
        for ( int a = 0; a < in.n_assets; a++ )
            kd.insert( kd.end(), gs.strikes_by_equity[a], gs.strikes_by_equity[a] + gs.nk );
    }
    std::vector<float> kf( kd.begin(), kd.end() );
    db.grid_nm = gs.nm;
    db.grid_nk = gs.nk;
    db.grid_udl = gs.udl;
    db.grid_grouped = gs.grouped ? 1 : 0;
    of.grid_mat_at = ar.put( at );
    of.grid_kf = ar.put( kf );
    of.grid_kd = ar.put( kd );
    return THOTH_OK;
}

template <class T>
T* At( char* base, size_t off )
{
    return reinterpret_cast<T*>( base + off );
}

}  // namespace

// Slots of the sweep's Sobol ring for steps per barrier G: the ops of
// bridge_step(0) and of the first group are staged together, then at the
// start of each group those of the next one while the current group's are
// read, so the ring holds any two consecutive groups.  A power of two.
static int RingSlots( const std::vector<int>& so, int G )
{
    const int m = (int)so.size() - 1;
    auto at = [&]( int i ) { return so[std::min( i, m )]; };
    int need = at( G + 1 );
    for ( int i = 1; i <= m; i += G )
        need = std::max( need, at( i + 2 * G ) - at( i ) );
    int S = 1;
    while ( S < need )
        S *= 2;
    return S;
}

static size_t LdsBytes( const DeviceBatch& B, int G )
{
    // the pathwise kernel reads a Sobol block table; the sweep a block table
    // (not in the in-loop inference kernels) and a ring (so does the staged
    // pathwise form).  The leverage ring:
    // 2G layers per equity, of floats in the sweep, of float4 nodes (layer
    // and its two sensitivities) in the staged pathwise form, none in the
    // other pathwise kernels (they read it from device memory)
    size_t ring = 0;
    if ( !B.pathwise || B.pw_lsv )
        for ( const std::vector<int>& so : B.step_ops )
            ring = std::max( ring, (size_t)RingSlots( so, G ) * B.nf_inst * 32 );
    const size_t table = B.pathwise && !B.pw_lsv ? B.lds_dims
                         : B.cv_mode == 2        ? 0
                                                 : ( B.lds_dims + 3 ) & ~size_t( 3 );
    const size_t lev = !B.pathwise ? (size_t)G * B.lds_lev_words : B.pw_lsv ? (size_t)G * 4 * B.lds_lev_words : 0;
    const size_t words = table + ring + lev;
    if ( !B.lds_tail )
        return words * sizeof( uint32_t );
    return ( ( words + 3 ) & ~size_t( 3 ) ) * sizeof( uint32_t ) + B.lds_tail;
}

// dynamic LDS of the equity-wave form at G steps per barrier: Sobol block
// table, ring of the 2 na factors, every equity's 2G leverage layers and the
// spots exchanged at a flow date
static size_t WavesLds( const DeviceBatch& B, int G )
{
    size_t ring = 0;
    int P = 0;
    for ( size_t b = 0; b < B.step_ops.size(); b++ )
    {
        ring = std::max( ring, (size_t)RingSlots( B.step_ops[b], G ) * 2 * B.na_max * 32 );
        P = std::max( P, B.db[b].lev_points );
    }
    const size_t words = ( ( B.lds_dims + 3 ) & ~size_t( 3 ) ) + ring + (size_t)2 * G * B.na_max * P + 64 * B.na_max;
    return words * sizeof( uint32_t );
}

// result slots per block of book b: the book, its contracts, and P - CV when
// the hedge network runs in the loop
static int SlotsOf( const DeviceBatch& B, int b )
{
    if ( B.pathwise == 1 )
        return 3;  // survival-weighted payoff, pathwise delta, pathwise vega
    if ( B.pathwise == 2 )
        return 3 + B.db[b].n_assets;  // payoff, book delta, vega, per-equity deltas
    return 1 + B.db[b].n_contracts + ( B.cv_mode == 2 ? 1 : 0 );
}

DeviceBatch::~DeviceBatch()
{
    if ( dbase && !external )
        (void)hipFree( dbase );
    if ( ditems )
        (void)hipFree( ditems );
    if ( dcounter )
        (void)hipFree( dcounter );
    if ( hpart )
        (void)hipHostFree( hpart );
}

int PrepareBatch( int32_t n_books, const thoth_mcl_book* books, const thoth_mcl_sweep_cfg* cfg, DeviceBatch& B,
                  const CvSpec* cv, const GridSpec* grids, int pathwise, char* ext_base, size_t ext_cap,
                  const RecordSpec* rec, const PathwiseLayers* pwl )
{
    B.pathwise = pathwise;
    int rc = EnsureDevice();
    if ( rc != THOTH_OK )
        return rc;
    if ( n_books < 1 || !books || !cfg )
        return SetError( THOTH_ERR_INVALID, "null argument" );
    if ( cfg->paths < 1 || cfg->sobol_skip < 0 )
        return SetError( THOTH_ERR_INVALID, "paths must be >= 1 and sobol_skip >= 0" );
    if ( cfg->precision != 0 && cfg->precision != 1 )
        return SetError( THOTH_ERR_INVALID, "precision must be 0 (FP64) or 1 (FP32)" );
    if ( cfg->replicates < 0 || cfg->replicates > 4096 )
        return SetError( THOTH_ERR_INVALID, "replicates must be in [0, 4096]" );
    // reference mode: points skip+1 .. skip+paths; RQMC mode: points 0 .. paths-1
    const bool rqmc = cfg->replicates > 0;
    const uint64_t first = rqmc ? 0 : (uint64_t)cfg->sobol_skip + 1;
    const uint64_t last = rqmc ? (uint64_t)cfg->paths - 1 : (uint64_t)cfg->sobol_skip + (uint64_t)cfg->paths;
    if ( last >= 0xffffffffull )
        return SetError( THOTH_ERR_UNSUPPORTED, "Sobol index beyond 2^32 - 2" );
    B.replicates = rqmc ? cfg->replicates : 1;
    B.first = (uint32_t)first;
    B.last = (uint32_t)last;
    // blocks own 256 consecutive Sobol indices aligned on 256
    B.q0 = (uint32_t)( first / kBlock );
    B.nblocks = (uint32_t)( last / kBlock ) - B.q0 + 1;
    B.paths = cfg->paths;
    B.precision = cfg->precision;
    B.n_books = n_books;

    Arena ar;
    B.db.assign( n_books, DevBook{} );
    B.step_ops.assign( n_books, {} );
    std::vector<Offsets> of( n_books );
    B.depth = 1;
    B.na_max = 1;
    for ( int b = 0; b < n_books; b++ )
    {
        rc = PackBook( b, books[b], B.db[b], of[b], ar, B.depth, B.step_ops[b] );
        if ( rc != THOTH_OK )
            return rc;
        of[b].has_pw = false;
        if ( pwl && pathwise == 2 && B.db[b].lev_points > 0 )
        {
            // leverage sensitivities, zeros where the caller gives none
            const size_t n = (size_t)B.db[b].n_assets * B.db[b].n_dates * B.db[b].lev_points;
            const double* src[2] = { pwl[b].ds, pwl[b].de };
            for ( int k = 0; k < 2; k++ )
            {
                std::vector<double> v( n, 0.0 );
                if ( src[k] )
                    std::copy( src[k], src[k] + n, v.begin() );
                of[b].pw[k] = ar.put( v );
                of[b].pw[2 + k] = ar.put( std::vector<float>( v.begin(), v.end() ) );
            }
            of[b].has_pw = true;
        }
        if ( cv && cv->J > 0 )
        {
            if ( cfg->precision != 1 || n_books != 1 )
                return SetError( THOTH_ERR_INVALID, "the control-variate data comes from FP32 sweeps of one book" );
            rc = PackCv( b, books[b], *cv, B.db[b], of[b], ar );
            if ( rc != THOTH_OK )
                return rc;
            B.cv_mode = cv->mode;
            B.cv_act = cv->act;
        }
        if ( grids && grids[b].nm > 0 )
        {
            if ( cv )
                return SetError( THOTH_ERR_INVALID, "vanilla grids and the control variate are separate sweeps" );
            rc = PackGrid( b, books[b], grids[b], B.db[b], of[b], ar );
            if ( rc != THOTH_OK )
                return rc;
        }
        B.na_max = std::max( B.na_max, B.db[b].n_assets );
    }
    // FP32 books whose equities are all LSV (or all Heston) with vol-of-vol
    // > 1e-12, uncorrelated stochastic spots and the same equity count run
    // the uniform-book kernel (no per-equity model branches); all-LSV books
    // with correlated stochastic spots run its correlated form (kind 3)
    B.kind = 0;
    if ( cfg->precision == 1 )
    {
        int kind = -1;
        for ( int b = 0; b < n_books && kind != 0; b++ )
        {
            const DevBook& d = B.db[b];
            for ( int a = 0; a < d.n_assets; a++ )
            {
                const DevAsset& A = d.asset[a];
                int k = A.model == THOTH_MODEL_LSV ? 1 : A.model == THOTH_MODEL_HESTON ? 2 : 0;
                // correlated stochastic spots: the correlated LSV kind (3) only
                if ( d.corr_stoch && d.n_assets > 1 )
                    k = k == 1 ? 3 : 0;
                if ( !( A.xi > 1e-12 ) || d.n_assets != B.na_max )
                    k = 0;
                kind = kind < 0 || kind == k ? k : 0;
            }
        }
        B.kind = std::max( kind, 0 );
    }
    // grouped grid books (every book of the batch, or none): the uniform LSV
    // arithmetic on the two shared factors
    bool grouped = false;
    for ( int b = 0; grids && b < n_books; b++ )
        grouped |= B.db[b].grid_grouped != 0;
    if ( grouped )
    {
        for ( int b = 0; b < n_books; b++ )
            if ( !B.db[b].grid_grouped )
                return SetError( THOTH_ERR_INVALID, "grouped and single grid books in one batch" );
        if ( cfg->precision != 1 || B.kind != 1 )
            return SetError( THOTH_ERR_UNSUPPORTED, "grouped grid books: FP32 LSV books with vol-of-vol > 0" );
        if ( B.depth > 10 )
            return SetError( THOTH_ERR_UNSUPPORTED, "grouped grid books: bridge stack deeper than 10" );
        B.kind = 5;
    }
    // pseudo-random draws: every normal (use_sobol false), or the dimensions
    // of a book past the Sobol table (path_generator.cpp:140-157).  One stream
    // per book, seeded alike: books with the same draws per path T share the
    // jump tables and the tail normals.
    B.pseudo = cfg->pseudo_random != 0;
    B.prng = B.pseudo;
    std::vector<int64_t> draws( n_books, 0 );
    for ( int b = 0; b < n_books; b++ )
    {
        const int64_t dims = (int64_t)B.db[b].m * B.db[b].factors;
        draws[b] = B.pseudo ? dims : std::max<int64_t>( 0, dims - kSobolDims );
        B.prng |= draws[b] > 0;
    }
    if ( B.prng )
    {
        const char* why = rqmc ? "randomised QMC" : cv ? "the control variate" : pathwise ? "pathwise Greeks" : nullptr;
        for ( int b = 0; b < n_books && !why; b++ )
            if ( grids && grids[b].nm > 0 )
                why = "vanilla grids";
        if ( why )
            return SetError( THOTH_ERR_UNSUPPORTED,
                             std::string( why ) + " needs every normal from the Sobol table (use_sobol true, "
                                                  "steps x factors <= 4097)" );
        if ( grouped )
            return SetError( THOTH_ERR_UNSUPPORTED, "grouped grid books need every normal from the Sobol table" );
        B.kind = 0;
        if ( B.pseudo )
            B.depth = 1;  // no bridge
    }
    // path recording (American exercise): the generic kernels, plain sweeps
    B.recording = false;
    for ( int b = 0; rec && b < n_books; b++ )
        if ( rec[b].n > 0 )
        {
            if ( rec[b].udl < 0 || rec[b].udl >= B.db[b].n_assets || rec[b].n > B.db[b].n_dates )
                return SetError( THOTH_ERR_INVALID, Why( b, "recorded equity or date count out of range" ) );
            B.recording = true;
        }
    if ( B.recording )
    {
        if ( rqmc || cv || pathwise )
            return SetError( THOTH_ERR_UNSUPPORTED, "American exercise runs in the reference-mode sweep only" );
        for ( int b = 0; b < n_books; b++ )
            if ( grids && grids[b].nm > 0 )
                return SetError( THOTH_ERR_UNSUPPORTED, "American exercise and vanilla grids are separate sweeps" );
        B.kind = 0;
    }
    // pathwise Greeks of FP32 books of LSV equities with uncorrelated
    // stochastic spots (the uniform kind 1): the staged-leverage form
    B.pw_lsv = pathwise == 2 && pwl && cfg->precision == 1 && B.kind == 1 &&
               pathwise_basket_lsv_available( B.na_max, B.depth );
    // dynamic LDS: the sweep's Sobol ring (the pathwise kernel: the Sobol block
    // table of the widest book), plus (FP32 LSV) a ring of 2G leverage layers
    // per equity of the kernel instantiation (G = 4, halved while two blocks
    // per CU, the occupancy of the 5-asset kernels, would not fit the 160 KiB
    // of LDS with the kernel's static LDS)
    const int na_inst = B.na_max <= 1 ? 1 : B.na_max <= 2 ? 2 : B.na_max <= 5 ? 5 : 8;
    size_t dims = 0, levw = 0;
    bool any_grid = false;
    for ( int b = 0; b < n_books; b++ )
    {
        if ( !B.pseudo )
            dims = std::max( dims, std::min( (size_t)B.db[b].m * B.db[b].factors, (size_t)kSobolDims ) );
        if ( cfg->precision == 1 && B.db[b].lev_points > 0 )
            levw = std::max( levw, (size_t)2 * na_inst * B.db[b].lev_points );
        any_grid |= B.db[b].grid_nm > 0;
    }
    B.lds_dims = dims;
    B.lds_lev_words = levw;
    // tail areas, 16-byte aligned after the leverage ring (sweep_block's layout)
    B.lds_tail = B.cv_mode == 2 ? (size_t)kCvLdsBytes : any_grid ? (size_t)grid_lds_bytes( cfg->precision == 0 ? 8 : 4 ) : 0;
    B.nf_inst = B.kind == 5 ? 2 : 2 * na_inst;  // grouped books: two shared factors
    int lev_group = 4;
    while ( lev_group > 1 && 2 * ( LdsBytes( B, lev_group ) + kSweepStaticLds ) > (size_t)160 * 1024 )
        lev_group /= 2;
    for ( int b = 0; b < n_books; b++ )
    {
        B.db[b].lev_group = lev_group;
        B.db[b].sob_slots = RingSlots( B.step_ops[b], lev_group );
    }
    B.lds_bytes = LdsBytes( B, lev_group );
    if ( B.lds_bytes + kSweepStaticLds > (size_t)160 * 1024 )
        return SetError( THOTH_ERR_UNSUPPORTED, "books need more than the 160 KiB of LDS of a workgroup "
                                                "(fewer steps, equities or leverage points)" );
    // grid partials [R][blocks][nm][2 nk][2] and their sums [nm][2 nk][2], per book
    B.grid_off.assign( n_books, 0 );
    B.grid_out_off.assign( n_books, 0 );
    B.grid_cells.assign( n_books, 0 );
    B.grid_total = B.grid_out_total = 0;
    for ( int b = 0; b < n_books; b++ )
    {
        const DevBook& d = B.db[b];
        B.grid_cells[b] = d.grid_nm * 2 * d.grid_nk * ( d.grid_grouped ? d.n_assets : 1 );
        B.grid_off[b] = B.grid_total;
        B.grid_out_off[b] = B.grid_out_total;
        B.grid_total += (size_t)B.replicates * B.nblocks * B.grid_cells[b] * 2;
        B.grid_out_total += (size_t)B.grid_cells[b] * 2;
    }
    // the equity-wave latency form: FP32 uniform LSV books (one equity count),
    // vanilla contracts only, plain sweeps or control-variate samples
    B.waves_ok = cfg->precision == 1 && B.kind == 1 && !B.prng && !B.recording && !any_grid && !pathwise &&
                 ( B.cv_mode == 0 || B.cv_mode == 1 ) && sweep_waves_available( B.na_max, B.depth, B.cv_mode ) &&
                 WavesLds( B, lev_group ) + kSweepStaticLds <= (size_t)160 * 1024;
    for ( int b = 0; b < n_books && B.waves_ok; b++ )
        for ( int c = 0; c < B.db[b].n_contracts; c++ )
            B.waves_ok = B.waves_ok && B.db[b].c_kind[c] == 0;
    B.wave_split = false;
    if ( B.waves_ok )
    {
        // where the lane form gives fewer than two waves per SIMD (measured on
        // the five-equity basket: 16 replicates of 1536 paths 0.58 ms against
        // 1.30 ms; at 16 x 8192, two lane waves per SIMD, the lane form is
        // the faster, 1.35 against 1.61 ms)
        hipDeviceProp_t prop;
        if ( hipGetDeviceProperties( &prop, 0 ) != hipSuccess )
            return SetError( THOTH_ERR_DEVICE, "hipGetDeviceProperties" );
        const size_t lane_waves = (size_t)B.nblocks * n_books * B.replicates * kWavesPerBlock;
        B.wave_split = lane_waves < (size_t)2 * 4 * prop.multiProcessorCount;
    }
    const int tiles = B.waves_ok ? 4 : 1;  // partial-sum tiles per Sobol block
    B.part_off.assign( n_books, 0 );
    B.part_total = 0;
    for ( int b = 0; b < n_books; b++ )
    {
        B.part_off[b] = B.part_total;
        B.part_total += (size_t)B.replicates * B.nblocks * tiles * SlotsOf( B, b ) * 2;
    }
    // random digital shifts: one 32-bit word per (replicate, dimension) from
    // xoshiro256++ seeded by splitmix64(rqmc_seed) (upper 32 bits of each draw)
    std::vector<uint32_t> shifts;
    if ( rqmc )
    {
        shifts.resize( (size_t)B.replicates * kSobolDims );
        uint64_t z = cfg->rqmc_seed, st[4];
        for ( int i = 0; i < 4; i++ )
        {
            z += 0x9e3779b97f4a7c15ULL;
            uint64_t x = z;
            x = ( x ^ ( x >> 30 ) ) * 0xbf58476d1ce4e5b9ULL;
            x = ( x ^ ( x >> 27 ) ) * 0x94d049bb133111ebULL;
            st[i] = x ^ ( x >> 31 );
        }
        auto rotl = []( uint64_t x, int k ) { return ( x << k ) | ( x >> ( 64 - k ) ); };
        for ( uint32_t& w : shifts )
        {
            const uint64_t r = rotl( st[0] + st[3], 23 ) + st[0];
            const uint64_t t = st[1] << 17;
            st[2] ^= st[0];
            st[3] ^= st[1];
            st[1] ^= st[2];
            st[0] ^= st[3];
            st[2] ^= t;
            st[3] = rotl( st[3], 45 );
            w = (uint32_t)( r >> 32 );
        }
    }
    const size_t shift_off = ar.put( shifts );
    std::vector<long long> gtab;
    for ( int b = 0; b < n_books; b++ )
    {
        gtab.push_back( (long long)B.grid_off[b] );
        gtab.push_back( (long long)B.grid_out_off[b] );
        gtab.push_back( (long long)B.grid_cells[b] * 2 );
    }
    const size_t gtab_off = ar.put( gtab );
    // jump tables per distinct T (arena) and, in Sobol mode, tail normals
    // [T][paths] per distinct T (a region of their own after the sums)
    B.fills.clear();
    std::vector<size_t> jump_off( n_books, 0 ), tail_off( n_books, 0 ), state_off( n_books, 0 );
    size_t tail_total = 0;
    const size_t elem = cfg->precision == 0 ? sizeof( double ) : sizeof( float );
    if ( B.prng )
    {
        B.jump_levels = JumpLevels( cfg->paths );
        SplitMixSeed( cfg->seed, B.seed_state );
        for ( int b = 0; b < n_books; b++ )
        {
            int same = -1;
            for ( int c = 0; c < b && same < 0; c++ )
                if ( draws[c] == draws[b] )
                    same = c;
            if ( same >= 0 )
            {
                jump_off[b] = jump_off[same];
                tail_off[b] = tail_off[same];
                state_off[b] = state_off[same];
                continue;
            }
            if ( draws[b] == 0 )
                continue;
            const std::vector<uint64_t>& tab = JumpTables( (uint64_t)draws[b], B.jump_levels );
            jump_off[b] = ar.put( tab.data(), (size_t)B.jump_levels * kJumpWords );
            // start states [4][paths], then (Sobol mode) the tail normals
            state_off[b] = tail_total;
            tail_total += ( 4 * sizeof( uint64_t ) * (size_t)cfg->paths + 255 ) & ~size_t( 255 );
            tail_off[b] = tail_total;
            if ( !B.pseudo )
                tail_total += ( (size_t)draws[b] * (size_t)cfg->paths * elem + 255 ) & ~size_t( 255 );
            B.fills.push_back( { draws[b], b, nullptr, nullptr, nullptr } );
        }
        if ( tail_total > ( (size_t)128 << 30 ) )
            return SetError( THOTH_ERR_UNSUPPORTED, "the normals past the Sobol table would take " +
                                                        std::to_string( tail_total >> 30 ) +
                                                        " GiB (fewer paths or steps)" );
    }
    const size_t arena_bytes = ( ar.data.size() + 255 ) & ~size_t( 255 );
    const size_t books_bytes = ( n_books * sizeof( DevBook ) + 255 ) & ~size_t( 255 );
    const size_t sums_bytes =
        ( ( B.part_total + B.grid_total + B.grid_out_total ) * sizeof( double ) + 255 ) & ~size_t( 255 );
    // recorded paths [n][paths] per recording book, after the tail normals
    std::vector<size_t> rec_off( n_books, 0 );
    size_t rec_total = 0;
    for ( int b = 0; B.recording && b < n_books; b++ )
        if ( rec[b].n > 0 )
        {
            rec_off[b] = rec_total;
            rec_total += ( (size_t)rec[b].n * (size_t)cfg->paths * elem + 255 ) & ~size_t( 255 );
        }
    const size_t total = arena_bytes + books_bytes + sums_bytes + tail_total + rec_total;
    if ( ext_base )
    {
        if ( total > ext_cap )
            return SetError( THOTH_ERR_UNSUPPORTED,
                             "the books need " + std::to_string( total ) + " bytes of device memory, more than the " +
                                 std::to_string( ext_cap ) + " bytes of the no-allocation scratch (fewer paths, "
                                 "replicates, dates or leverage points, or the allocating entry point)" );
        B.dbase = ext_base;
        B.external = true;
    }
    else
    {
        hipError_t e = hipMalloc( &B.dbase, total );
        if ( e != hipSuccess )
            return SetError( THOTH_ERR_DEVICE, std::string( "hipMalloc: " ) + hipGetErrorString( e ) );
        // pinned host memory for the partial sums (the no-allocation entry
        // point keeps a pageable copy per run)
        if ( hipHostMalloc( &B.hpart, std::max<size_t>( 1, B.part_total ) * sizeof( double ) ) != hipSuccess )
            B.hpart = nullptr;
    }
    char* dbase = B.dbase;
    B.dpart = reinterpret_cast<double*>( dbase + arena_bytes + books_bytes );
    B.dgrid = B.dpart + B.part_total;
    B.dgrid_out = B.dgrid + B.grid_total;
    for ( int b = 0; b < n_books; b++ )
    {
        DevBook& d = B.db[b];
        const Offsets& o = of[b];
        d.t = At<double>( dbase, o.t );
        d.dt = At<double>( dbase, o.dt );
        d.sqrt_dt = At<double>( dbase, o.sqrt_dt );
        d.bsq = At<double>( dbase, o.bsq );
        d.binv = At<float>( dbase, o.binv );
        d.step_ops = At<int>( dbase, o.step_ops );
        d.op_b = At<int>( dbase, o.op_b );
        d.op_left = At<int>( dbase, o.op_left );
        d.op_w = At<double>( dbase, o.op_w );
        d.op_wf = At<float>( dbase, o.op_wf );
        d.drift = At<double>( dbase, o.drift );
        d.sc = At<float>( dbase, o.sc );
        if ( d.n_assets > 1 )
        {
            d.chol_t = At<double>( dbase, o.chol );
            d.cholf_t = At<float>( dbase, o.cholf );
        }
        if ( d.lev_points )
        {
            d.lev = At<double>( dbase, o.lev );
            d.levf = At<float>( dbase, o.levf );
            d.lev_ln_step = At<double>( dbase, o.lnstep );
            d.lev_offset = At<double>( dbase, o.offset );
            d.lev_offset_f = At<float>( dbase, o.offsetf );
            d.lev_inv_step = At<float>( dbase, o.invstep );
        }
        if ( o.has_pw )
        {
            d.lev_ds = At<double>( dbase, o.pw[0] );
            d.lev_de = At<double>( dbase, o.pw[1] );
            d.lev_dsf = At<float>( dbase, o.pw[2] );
            d.lev_def = At<float>( dbase, o.pw[3] );
        }
        d.mon_mask = At<uint32_t>( dbase, o.mon );
        d.flow_mask = At<uint32_t>( dbase, o.flow );
        if ( d.cv_J > 0 )
        {
            d.cv_start = At<int>( dbase, o.cv_start );
            d.cv_end = At<int>( dbase, o.cv_end );
            d.cv_scale = At<float>( dbase, o.cv_scale );
        }
        d.partial = B.dpart + B.part_off[b];
        d.rec_udl = -1;
        if ( B.recording && rec[b].n > 0 )
        {
            char* r = dbase + arena_bytes + books_bytes + sums_bytes + tail_total + rec_off[b];
            d.rec_udl = rec[b].udl;
            d.rec_n = rec[b].n;
            d.rec_stride = cfg->paths;
            if ( cfg->precision == 0 )
                d.rec = reinterpret_cast<double*>( r );
            else
                d.recf = reinterpret_cast<float*>( r );
        }
        if ( B.prng )
        {
            d.pseudo = B.pseudo ? 1 : 0;
            d.jump_levels = B.jump_levels;
            d.jump = draws[b] > 0 ? At<uint64_t>( dbase, jump_off[b] ) : nullptr;
            d.states = draws[b] > 0 ? reinterpret_cast<const uint64_t*>( dbase + arena_bytes + books_bytes +
                                                                         sums_bytes + state_off[b] )
                                    : nullptr;
            d.tail_dims = B.pseudo ? 0 : (int)draws[b];
            d.tail_stride = cfg->paths;
            char* tail = dbase + arena_bytes + books_bytes + sums_bytes + tail_off[b];
            if ( d.tail_dims > 0 )
            {
                if ( cfg->precision == 0 )
                    d.tail = reinterpret_cast<const double*>( tail );
                else
                    d.tailf = reinterpret_cast<const float*>( tail );
            }
            for ( int q = 0; q < 4; q++ )
                d.seed_state[q] = B.seed_state[q];
        }
        if ( d.grid_nm > 0 )
        {
            d.grid_mat_at = At<int>( dbase, o.grid_mat_at );
            d.grid_kf = At<float>( dbase, o.grid_kf );
            d.grid_kd = At<double>( dbase, o.grid_kd );
            d.grid_part = B.dgrid + B.grid_off[b];
        }
    }
    for ( DeviceBatch::TailFill& fl : B.fills )
    {
        fl.jump = B.db[fl.book].jump;
        fl.states = const_cast<uint64_t*>( B.db[fl.book].states );
        fl.dst = B.pseudo ? nullptr : dbase + arena_bytes + books_bytes + sums_bytes + tail_off[fl.book];
    }
    B.dbooks = reinterpret_cast<const DevBook*>( dbase + arena_bytes );
    B.dgrid_table = At<long long>( dbase, gtab_off );
    B.dshifts = rqmc ? At<uint32_t>( dbase, shift_off ) : nullptr;
    hipStream_t s = Stream();
#define HIP_OK( expr )                                                                                      \
    do                                                                                                      \
    {                                                                                                       \
        hipError_t e_ = ( expr );                                                                           \
        if ( e_ != hipSuccess )                                                                             \
            return SetError( THOTH_ERR_DEVICE, std::string( #expr " failed: " ) + hipGetErrorString( e_ ) ); \
    } while ( 0 )
    HIP_OK( hipMemcpyAsync( dbase, ar.data.data(), ar.data.size(), hipMemcpyHostToDevice, s ) );
    HIP_OK( hipMemcpyAsync( dbase + arena_bytes, B.db.data(), n_books * sizeof( DevBook ), hipMemcpyHostToDevice, s ) );
    HIP_OK( hipStreamSynchronize( s ) );
    return THOTH_OK;
}

// the sweep's device work on stream s: the tail normals (once per distinct
// T), then the sweep kernel in the batch's launch shape
static int EnqueueSweep( const DeviceBatch& B, hipStream_t s )
{
    // kernel control-variate mode: in-loop inference of a gated network is mode 3
    const int kcv = B.cv_mode == 2 && B.cv_act == THOTH_CV_GATE ? 3 : B.cv_mode;
    for ( const DeviceBatch::TailFill& fl : B.fills )
    {
        bool ok = launch_prng_start_states( s, fl.jump, B.jump_levels, B.seed_state, (int)fl.T, B.paths, fl.states );
        if ( ok && fl.dst )
            ok = B.precision == 0
                     ? launch_tail_fill_f64( s, fl.states, (int)fl.T, B.paths, reinterpret_cast<double*>( fl.dst ) )
                     : launch_tail_fill_f32( s, fl.states, (int)fl.T, B.paths, reinterpret_cast<float*>( fl.dst ) );
        if ( !ok )
            return SetError( THOTH_ERR_DEVICE, "pseudo-random stream set-up launch failed" );
    }
    bool ok;
    if ( B.wave_split )
    {
        const dim3 grid( 4 * B.nblocks, B.n_books, B.replicates );
        const int G = B.db.empty() ? 1 : B.db[0].lev_group;
        ok = launch_sweep_waves( B.na_max, B.depth, B.cv_mode, grid, s, WavesLds( B, G ), B.dbooks,
                                 DeviceDirections(), B.q0, B.first, B.last, B.dshifts );
    }
    else if ( B.persist_wpc > 0 )
    {
        // work queue of (book, Sobol block) items; the counter is cleared by a kernel
        launch_counter_reset( s, B.dcounter );
        const dim3 pg( (uint32_t)( B.cus * B.persist_wpc ) );
        const DevBook* db = B.dbooks;
        const uint32_t* V = DeviceDirections();
        ok = B.precision == 0 ? launch_persistent_f64( B.na_max, B.depth, pg, s, B.lds_bytes, db, V, B.q0, B.first,
                                                       B.last, B.ditems, B.n_items, B.nblocks, B.dcounter )
             : B.kind         ? launch_persistent_f32_uniform( B.kind, B.na_max, B.depth, pg, s, B.lds_bytes, db, V,
                                                               B.q0, B.first, B.last, B.ditems, B.n_items, B.nblocks,
                                                               B.dcounter )
                              : launch_persistent_f32( B.na_max, B.depth, pg, s, B.lds_bytes, db, V, B.q0, B.first,
                                                       B.last, B.ditems, B.n_items, B.nblocks, B.dcounter );
    }
    else
    {
        const dim3 grid( B.nblocks, B.n_books, B.replicates );
        ok = B.prng ? ( B.precision == 0 ? launch_sweep_prng_f64( B.na_max, B.depth, grid, s, B.lds_bytes, B.dbooks,
                                                                  DeviceDirections(), B.q0, B.first, B.last )
                                         : launch_sweep_prng_f32( B.na_max, B.depth, grid, s, B.lds_bytes, B.dbooks,
                                                                  DeviceDirections(), B.q0, B.first, B.last ) )
             : B.precision == 0
                 ? launch_sweep_f64( B.na_max, B.depth, grid, s, B.lds_bytes, B.dbooks, DeviceDirections(), B.q0,
                                     B.first, B.last, B.dshifts )
                 : ( B.kind && launch_sweep_f32_uniform( B.kind, B.na_max, B.depth, grid, s, B.lds_bytes, B.dbooks,
                                                         DeviceDirections(), B.q0, B.first, B.last, B.dshifts,
                                                         kcv ) ) ||
                       ( B.kind != 5 &&  // grouped grid books have no generic form
                         launch_sweep_f32( B.na_max, B.depth, grid, s, B.lds_bytes, B.dbooks, DeviceDirections(), B.q0,
                                         B.first, B.last, B.dshifts, kcv ) );
    }
    if ( !ok )
        return SetError( THOTH_ERR_UNSUPPORTED, "no kernel instantiation for this book shape" );
    HIP_OK( hipGetLastError() );
    return THOTH_OK;
}

int RunBatch( const DeviceBatch& B, thoth_mcl_book_out* out, CvSums* cv_out,
              const std::function<int( hipStream_t )>& then )
{
    hipStream_t s = Stream();
    hipEvent_t e0 = SweepEvent( 0 ), e1 = SweepEvent( 1 );
    HIP_OK( hipEventRecord( e0, s ) );
    int rc = EnqueueSweep( B, s );
    if ( rc != THOTH_OK )
        return rc;
    HIP_OK( hipEventRecord( e1, s ) );
    if ( then && ( rc = then( s ) ) != THOTH_OK )
        return rc;
    // into the batch's pinned buffer where it has one (a truly asynchronous
    // copy, no staging), else into pageable memory
    std::vector<double> pageable( B.hpart ? 0 : B.part_total );
    double* part = B.hpart ? B.hpart : pageable.data();
    HIP_OK( hipMemcpyAsync( part, B.dpart, B.part_total * sizeof( double ), hipMemcpyDeviceToHost, s ) );
    HIP_OK( hipStreamSynchronize( s ) );
    float ms = 0;
    HIP_OK( hipEventElapsedTime( &ms, e0, e1 ) );
    // partial-sum tiles of the form that ran (4 per Sobol block in the equity-wave form)
    const uint32_t ntiles = B.nblocks * ( B.wave_split ? 4u : 1u );
    for ( int b = 0; b < B.n_books; b++ )
      for ( int r = 0; r < B.replicates; r++ )
    {
        const int nc = B.db[b].n_contracts;
        const int ne = SlotsOf( B, b );
        std::vector<double> sum( ne * 2, 0.0 ), comp( ne * 2, 0.0 );
        const double* pb = part + B.part_off[b] + (size_t)r * ntiles * ne * 2;
        for ( uint32_t k = 0; k < ntiles; k++ )
            for ( int j = 0; j < ne * 2; j++ )
            {
                // Kahan over blocks in block order (deterministic)
                const double y = pb[(size_t)k * ne * 2 + j] - comp[j];
                const double t = sum[j] + y;
                comp[j] = ( t - sum[j] ) - y;
                sum[j] = t;
            }
        thoth_mcl_book_out& o = out[(size_t)b * B.replicates + r];
        std::memset( &o, 0, sizeof( o ) );
        o.book_sum = sum[0];
        o.book_sum2 = sum[1];
        for ( int c = 0; c < nc; c++ )
        {
            o.contract_sum[c] = sum[2 * ( 1 + c )];
            o.contract_sum2[c] = sum[2 * ( 1 + c ) + 1];
        }
        o.count = B.paths;
        o.device_ms = ms;
        if ( cv_out && B.cv_mode == 2 && b == 0 )
        {
            cv_out[r].sum = sum[2 * ( 1 + nc )];
            cv_out[r].sum2 = sum[2 * ( 1 + nc ) + 1];
        }
    }
    return THOTH_OK;
}

int SetLevGroup( DeviceBatch& B, int G )
{
    if ( G < 1 || G > 16 )
        return SetError( THOTH_ERR_INVALID, "leverage group out of range" );
    if ( LdsBytes( B, G ) + kSweepStaticLds > (size_t)160 * 1024 )
        return SetError( THOTH_ERR_UNSUPPORTED, "steps per LDS barrier beyond the LDS of a workgroup" );
    for ( size_t b = 0; b < B.db.size(); b++ )
    {
        B.db[b].lev_group = G;
        B.db[b].sob_slots = RingSlots( B.step_ops[b], G );
    }
    B.lds_bytes = LdsBytes( B, G );
    hipStream_t s = Stream();
    HIP_OK( hipMemcpyAsync( const_cast<DevBook*>( B.dbooks ), B.db.data(), B.db.size() * sizeof( DevBook ),
                            hipMemcpyHostToDevice, s ) );
    HIP_OK( hipStreamSynchronize( s ) );
    return THOTH_OK;
}

// the persistent form exists for the generic kernels and the uniform-book
// instantiations of the workloads (one-equity LSV and Heston, five-equity LSV)
static bool PersistentAvailable( const DeviceBatch& B )
{
    if ( B.prng || B.cv_mode != 0 || B.replicates != 1 || B.pathwise || B.grid_total > 0 )
        return false;
    if ( B.precision == 0 || B.kind == 0 )
        return true;
    if ( B.depth > 10 )
        return false;
    return ( ( B.kind == 1 || B.kind == 2 ) && B.na_max == 1 ) || ( ( B.kind == 1 || B.kind == 3 ) && B.na_max == 5 );
}

int SetShape( DeviceBatch& B, int lev_group, int wpc )
{
    if ( wpc < -1 || wpc > 16 )
        return SetError( THOTH_ERR_INVALID, "workgroups per CU must be in [-1, 16]" );
    if ( wpc > 0 && !PersistentAvailable( B ) )
        return SetError( THOTH_ERR_UNSUPPORTED, "no persistent form of the sweep for these books" );
    if ( wpc == -1 && ( !B.waves_ok || WavesLds( B, lev_group ) + kSweepStaticLds > (size_t)160 * 1024 ) )
        return SetError( THOTH_ERR_UNSUPPORTED, "no equity-wave form of the sweep for these books" );
    int rc = SetLevGroup( B, lev_group );
    if ( rc != THOTH_OK )
        return rc;
    B.wave_split = wpc == -1;
    if ( wpc == -1 )
    {
        B.persist_wpc = 0;
        return THOTH_OK;
    }
    if ( wpc > 0 && !B.ditems )
    {
        // items book-major, Sobol blocks in order
        std::vector<int2> items;
        items.reserve( (size_t)B.n_books * B.nblocks );
        for ( int b = 0; b < B.n_books; b++ )
            for ( uint32_t t = 0; t < B.nblocks; t++ )
                items.push_back( make_int2( b, (int)t ) );
        hipDeviceProp_t prop;
        HIP_OK( hipGetDeviceProperties( &prop, 0 ) );
        B.cus = prop.multiProcessorCount;
        HIP_OK( hipMalloc( &B.ditems, items.size() * sizeof( int2 ) ) );
        HIP_OK( hipMalloc( &B.dcounter, sizeof( int ) ) );
        hipStream_t s = Stream();
        HIP_OK( hipMemcpyAsync( B.ditems, items.data(), items.size() * sizeof( int2 ), hipMemcpyHostToDevice, s ) );
        HIP_OK( hipStreamSynchronize( s ) );
        B.n_items = (int)items.size();
    }
    B.persist_wpc = wpc;
    return THOTH_OK;
}

namespace
{
std::string ShapeKey( const DeviceBatch& B )
{
    std::string k = std::to_string( B.precision ) + "/" + std::to_string( B.kind ) + "/" + std::to_string( B.na_max ) +
                    "/" + std::to_string( B.depth ) + "/" + std::to_string( B.n_books ) + "/" +
                    std::to_string( B.nblocks ) + "/" + std::to_string( B.prng ) + "/" + std::to_string( B.lds_tail );
    for ( const DevBook& d : B.db )
        k += ":" + std::to_string( d.m ) + "," + std::to_string( d.factors ) + "," + std::to_string( d.lev_points ) +
             "," + std::to_string( d.n_contracts ) + "," + std::to_string( d.n_ops );
    return k;
}

int TimeSweep( const DeviceBatch& B, double& ms )
{
    hipStream_t s = Stream();
    hipEvent_t e0 = SweepEvent( 0 ), e1 = SweepEvent( 1 );
    HIP_OK( hipEventRecord( e0, s ) );
    int rc = EnqueueSweep( B, s );
    if ( rc != THOTH_OK )
        return rc;
    HIP_OK( hipEventRecord( e1, s ) );
    HIP_OK( hipEventSynchronize( e1 ) );
    float f = 0;
    HIP_OK( hipEventElapsedTime( &f, e0, e1 ) );
    ms = f;
    return THOTH_OK;
}
}  // namespace

int TuneShape( DeviceBatch& B, ShapeChoice& out )
{
    // callers hold the library lock
    static std::map<std::string, ShapeChoice> cache;
    const std::string key = ShapeKey( B );
    auto it = cache.find( key );
    if ( it != cache.end() )
    {
        int rc = SetShape( B, it->second.lev_group, it->second.wpc );
        if ( rc != THOTH_OK )
            return rc;
        out = it->second;
        out.candidates = 0;
        out.default_ms = 0;
        return THOTH_OK;
    }
    const int groups[] = { 1, 2, 4, 8 };
    // 0: one workgroup per Sobol block; W > 0: persistent, W per CU (the
    // forms with the same partial sums; the equity-wave form is not a
    // candidate: it sums in quarter blocks, and is taken by itself where it pays)
    const int forms[] = { 0, 1, 2, 3, 4 };
    ShapeChoice best;
    best.ms = 1e300;
    const int g0 = B.db.empty() ? 4 : B.db[0].lev_group, w0 = B.persist_wpc;
    int candidates = 0;
    for ( int g : groups )
        for ( int w : forms )
        {
            int rc = SetShape( B, g, w );
            if ( rc == THOTH_ERR_UNSUPPORTED )
                continue;  // LDS does not fit, or no persistent form
            if ( rc != THOTH_OK )
                return rc;
            double ms = 0;
            if ( ( rc = TimeSweep( B, ms ) ) != THOTH_OK || ( rc = TimeSweep( B, ms ) ) != THOTH_OK )
                return rc;
            candidates++;
            if ( g == g0 && w == w0 )
                best.default_ms = ms;
            if ( ms < best.ms )
            {
                const double d = best.default_ms;
                best.ms = ms;
                best.lev_group = g;
                best.wpc = w;
                best.default_ms = d;
            }
        }
    if ( candidates == 0 )
        return SetError( THOTH_ERR_UNSUPPORTED, "no launch shape fits these books" );
    int rc = SetShape( B, best.lev_group, best.wpc );
    if ( rc != THOTH_OK )
        return rc;
    best.candidates = candidates;
    cache[key] = best;
    out = best;
    return THOTH_OK;
}

int ReduceGrids( const DeviceBatch& B, hipStream_t s )
{
    int max_cells = 0;
    for ( int b = 0; b < B.n_books; b++ )
        max_cells = std::max( max_cells, B.grid_cells[b] * 2 );
    if ( max_cells > 0 )
        launch_grid_sum( s, B.dgrid, B.dgrid_table, B.n_books, max_cells, (int)( B.replicates * B.nblocks ),
                         B.dgrid_out );
    HIP_OK( hipGetLastError() );
    return THOTH_OK;
}

int CopyGrids( const DeviceBatch& B, std::vector<std::vector<double>>& out )
{
    hipStream_t s = Stream();
    out.assign( B.n_books, {} );
    std::vector<double> all( B.grid_out_total );
    if ( B.grid_out_total )
        HIP_OK( hipMemcpyAsync( all.data(), B.dgrid_out, all.size() * sizeof( double ), hipMemcpyDeviceToHost, s ) );
    HIP_OK( hipStreamSynchronize( s ) );
    for ( int b = 0; b < B.n_books; b++ )
        out[b].assign( all.begin() + B.grid_out_off[b], all.begin() + B.grid_out_off[b] + B.grid_cells[b] * 2 );
    return THOTH_OK;
}

int SweepBooks( int32_t n_books, const thoth_mcl_book* books, const thoth_mcl_sweep_cfg* cfg,
                thoth_mcl_book_out* out )
{
    if ( !out )
        return SetError( THOTH_ERR_INVALID, "null argument" );
    DeviceBatch B;
    int rc = PrepareBatch( n_books, books, cfg, B );
    if ( rc != THOTH_OK )
        return rc;
    return RunBatch( B, out );
}

}  // namespace thoth
