// mcl_split.cpp -- the cluster master's path split and result pooling
// (reference: pricer/src/modes/run_cluster.cpp:308-339 and 402-447) for one
// device: a book's paths are cut into disjoint Sobol blocks, each block is
// priced as its own request (in this process, or by separate processes
// sharing the GPU), and the results are pooled path-weighted, standard errors
// in quadrature.
#include <cmath>
#include <cstring>
#include <vector>

#include "mcl_internal.hpp"
#include "thoth_mcl.h"

namespace
{
// path-weighted mean of one field over the parts
double Pool( int n, const double* w, const thoth_mcl_result* in, double thoth_mcl_result::*f )
{
    double s = 0.0;
    for ( int k = 0; k < n; k++ )
        s += w[k] * ( in[k].*f );
    return s;
}
}  // namespace

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_mcl_split_paths( int64_t total, int32_t parts,
                                                                               int64_t* paths, int64_t* skips )
{
    if ( total < 1 || parts < 1 || !paths || !skips )
        return -thoth::SetError( THOTH_ERR_INVALID, "split: total >= 1, parts >= 1, non-null outputs" );
    // never a part of 0 paths: at most `total` parts (run_cluster.cpp:311-314)
    const int32_t n = total < (int64_t)parts ? (int32_t)total : parts;
    int64_t run = 0;
    for ( int32_t k = 0; k < n; k++ )
    {
        // balanced: the first total % n parts take one path more; each part
        // skips the paths of the parts before it (run_cluster.cpp:316-328)
        paths[k] = total / n + ( k < total % n ? 1 : 0 );
        skips[k] = run;
        run += paths[k];
    }
    return n;
}

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_mcl_pool_results( int32_t parts,
                                                                                const int64_t* paths,
                                                                                const thoth_mcl_result* in,
                                                                                int32_t n_contracts,
                                                                                thoth_mcl_result* out )
{
    if ( parts < 1 || !paths || !in || !out || n_contracts < 0 || n_contracts > THOTH_MCL_MAX_CONTRACTS )
        return thoth::SetError( THOTH_ERR_INVALID, "pool: parts >= 1, 0..32 contracts, non-null arguments" );
    int64_t total = 0;
    for ( int k = 0; k < parts; k++ )
    {
        if ( paths[k] < 1 )
            return thoth::SetError( THOTH_ERR_INVALID, "pool: every part needs paths" );
        total += paths[k];
    }
    std::vector<double> w( parts );
    for ( int k = 0; k < parts; k++ )
        w[k] = (double)paths[k] / (double)total;  // the part's path share (run_cluster.cpp:426)
    // values: path-weighted means; standard errors: sqrt(sum w^2 se^2) (run_cluster.cpp:427-447)
    auto quad = [&]( auto get )
    {
        double s = 0.0;
        for ( int k = 0; k < parts; k++ )
        {
            const double v = get( in[k] );
            s += w[k] * w[k] * v * v;
        }
        return std::sqrt( s );
    };
    out->premium = Pool( parts, w.data(), in, &thoth_mcl_result::premium );
    out->premium_trust = quad( []( const thoth_mcl_result& r ) { return r.premium_trust; } );
    out->delta = Pool( parts, w.data(), in, &thoth_mcl_result::delta );
    out->gamma = Pool( parts, w.data(), in, &thoth_mcl_result::gamma );
    out->vega = Pool( parts, w.data(), in, &thoth_mcl_result::vega );
    out->rho = Pool( parts, w.data(), in, &thoth_mcl_result::rho );
    double* const* outs[] = { &out->contract_premium, &out->contract_delta, &out->contract_gamma,
                              &out->contract_vega, &out->contract_rho };
    double* thoth_mcl_result::*ins[] = { &thoth_mcl_result::contract_premium, &thoth_mcl_result::contract_delta,
                                         &thoth_mcl_result::contract_gamma, &thoth_mcl_result::contract_vega,
                                         &thoth_mcl_result::contract_rho };
    for ( int f = 0; f < 5; f++ )
    {
        double* dst = *outs[f];
        if ( !dst )
            continue;
        for ( int c = 0; c < n_contracts; c++ )
        {
            double s = 0.0;
            for ( int k = 0; k < parts; k++ )
                s += in[k].*ins[f] ? w[k] * ( in[k].*ins[f] )[c] : 0.0;
            dst[c] = s;
        }
    }
    if ( out->contract_trust )
        for ( int c = 0; c < n_contracts; c++ )
            out->contract_trust[c] =
                quad( [c]( const thoth_mcl_result& r ) { return r.contract_trust ? r.contract_trust[c] : 0.0; } );
    out->n_dates = in[0].n_dates;
    out->n_scenarios = in[0].n_scenarios;
    out->sweep_ms = out->calib_ms = 0.0;
    for ( int k = 0; k < parts; k++ )
    {
        out->sweep_ms += in[k].sweep_ms;
        out->calib_ms += in[k].calib_ms;
    }
    return THOTH_OK;
}

extern "C" __attribute__( ( visibility( "default" ) ) ) int thoth_mcl_price_split( const thoth_mcl_request* req,
                                                                               int32_t parts, thoth_mcl_result* res )
{
    if ( !req || !res )
        return thoth::SetError( THOTH_ERR_INVALID, "null argument" );
    std::vector<int64_t> paths( parts > 0 ? parts : 1 ), skips( paths.size() );
    const int n = thoth_mcl_split_paths( req->mcl.paths, parts, paths.data(), skips.data() );
    if ( n < 1 )
        return -n;
    const int nc = req->n_contracts > 0 && req->n_contracts <= THOTH_MCL_MAX_CONTRACTS ? req->n_contracts : 0;
    std::vector<thoth_mcl_result> part( n );
    std::vector<double> arrays( (size_t)n * 6 * ( nc > 0 ? nc : 1 ), 0.0 );
    for ( int k = 0; k < n; k++ )
    {
        // the slave's request: its path count, seed k and the Sobol skip of the
        // parts before it (run_cluster.cpp:330-339)
        thoth_mcl_request q = *req;
        q.mcl.paths = paths[k];
        q.mcl.seed = k;
        q.mcl.sobol_skip = skips[k];
        thoth_mcl_result& r = part[k];
        std::memset( &r, 0, sizeof( r ) );
        double* a = arrays.data() + (size_t)k * 6 * ( nc > 0 ? nc : 1 );
        const size_t st = nc > 0 ? nc : 1;
        r.contract_premium = a;
        r.contract_trust = a + st;
        r.contract_delta = a + 2 * st;
        r.contract_gamma = a + 3 * st;
        r.contract_vega = a + 4 * st;
        r.contract_rho = a + 5 * st;
        const int rc = thoth_mcl_price( &q, &r );
        if ( rc != THOTH_OK )
            return rc;
    }
    return thoth_mcl_pool_results( n, paths.data(), part.data(), nc, res );
}
