// mcl_hostlib.hpp -- market model of the host engine: curves, equities with
// their bs / heston / lsv volatility, the SABR target surface and its Dupire
// local volatility, the diffusion grid, the Cholesky factor and the LSV
// leverage calibration.  All FP64 (the FP32 production sweep keeps these
// steps in FP64).
#pragma once
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "thoth_mcl.h"
#include "thoth_mcl_gpu.h"

namespace thoth::host
{

constexpr double kDaysPerYear = 365.0;  // helpers/constants.hpp:7 (ACT/365)
constexpr int kMaxLevPoints = 512;      // spot points of a leverage surface (FP32 LDS ring bound)

inline double YearFraction( double from, double to ) { return ( to - from ) / kDaysPerYear; }  // misc.hpp:54-57

// zero curve, linear in rate between pillars, flat outside (curve.cpp:45-73)
struct Curve
{
    std::vector<double> days, values;  // decimal
    double shift = 0;                  // rho bump
    double Value( double day ) const;
};

// Hagan lognormal SABR with BDK power-law wings (sabr_volatility.cpp:40-239)
struct Sabr
{
    std::vector<double> mats, alpha, beta, rho, nu;
    double Interp( const std::vector<double>& v, double t ) const;
    double Hagan( double k, double f, double t ) const;
    double Wing( double k, double f, double t, double cutoff, bool right ) const;
    double Implied( double strike, double forward, double day ) const;
};

struct StochParams
{
    double v0 = 0, kappa = 0, theta = 0, xi = 0, rho = 0;
};

struct Equity
{
    double spot = 0;
    int kind = THOTH_VOL_BS;
    double bs_vol = 0, v0 = 0, theta = 0, kappa = 0, xi = 0, rho = 0;
    bool has_div = false;
    double div = 0;
    Sabr surface;
    double vol_shift = 0;  // vega bump (Volatility::_vol_shift)
};

struct Market
{
    std::vector<Equity> eq;
    Curve rate;
    Curve discount;            // the OIS curve of a multi-curve currency (currency.cpp:19-27)
    bool multi_curve = false;
    // spot/spot correlation: today's matrix, and for a term structure the
    // pillar maturities (years) with one matrix each (correlation.cpp:35-130)
    std::vector<double> correlation;
    std::vector<double> corr_pillars;
    std::vector<std::vector<double>> corr_mats;
    bool TermStructured() const { return corr_pillars.size() > 1; }
    // entry (i, j) averaged over [a, b] years (correlation.cpp:528-620)
    double AvgCorrelation( int i, int j, double a, double b ) const;
    int lev_spot_points = 0, lev_time_layers = 0;  // 0, 0: one 201-point layer per diffusion date
    StochParams Stoch( int a ) const;                       // StochasticParams()
    double DivRepo( int a, double day ) const;              // equity.cpp:149-154
    double Drift( int a, double day ) const;                // drift_node.cpp:20-27
    double Forward( int a, double day ) const;              // equity.cpp:160-191
    double ImplicitVol( int a, double strike, double day ) const;  // single.cpp:56-62
    double LocalVol( int a, double strike, double day ) const;     // equity.cpp:196-211 + volatility.cpp:55-128
};

struct Grid
{
    std::vector<double> days, t, dt;
    bool uniform = false;
};

// leverage layers of one equity on the diffusion grid (leverage_surface.hpp):
// layer i is read by the step into diffusion date i.  A surface calibration
// keeps its calibrated layers in `surface` and fills the per-date layers by
// interpolating them linearly in time (leverage_surface.cpp:43-66).
struct LeverageGrid
{
    int points = 0;
    std::vector<double> t, ln_step;
    std::vector<int64_t> offset;
    std::vector<double> values;  // [dates * points]
    struct Surface
    {
        std::vector<double> t;       // calibration dates, years
        double ln_step = 0;
        int64_t offset = 0;
        std::vector<double> values;  // [layers * points]
    } surface;
};

// the arrays one thoth_mcl_book points to (the book keeps pointers into them)
struct BookArrays
{
    std::vector<int32_t> model;
    std::vector<double> spot, vol, v0, kappa, theta, xi, rho, drift, chol, disc;
    std::vector<double> chol_dates;  // [dates][n][n] per-date factors (term-structured correlation)
    std::vector<double> lev_ln_step, lev_values;
    std::vector<int64_t> lev_offset;
    std::vector<int32_t> rb_best, rb_n, rb_member;
    std::vector<double> rb_weight;
    std::vector<int32_t> c_kind, c_call, c_udl, c_flow, c_up, c_in, c_mon_off, c_mon;
    std::vector<double> c_strike, c_level;
    thoth_mcl_book book{};
};

constexpr double kSpotBump = 0.01;    // helpers/constants.hpp:36
constexpr double kGammaBump = 0.10;   // tasks/pricer.hpp:20
constexpr double kVolBump = 0.01;     // tasks/pricer.hpp:25
constexpr double kRateBump = 0.0001;  // tasks/pricer.hpp:26

Market BuildMarket( const thoth_mcl_request& req );
// the diffusion grid; extra_days adds fixing dates (vanilla-grid maturities)
Grid BuildGrid( const thoth_mcl_request& req, const Market& mk, const std::vector<double>* extra_days = nullptr );
// the device book arrays of one scenario market (pointers into A)
void Resolve( const thoth_mcl_request& req, const Market& mk, const Grid& g, const std::vector<double>& chol,
              const std::vector<LeverageGrid>& lev, BookArrays& A );
bool CholeskyLower( std::vector<double>& a, int n );  // maths.cpp:354-381
// the lower factor of today's spot/spot matrix (empty for one equity); throws
// Invalid when the matrix is missing or not positive-definite
std::vector<double> SpotCholesky( const thoth_mcl_request& req, const Market& mk );
int CalibrateLeverage( const Market& mk, int a, const Grid& g, LeverageGrid& out );
// Sensitivity of equity a's calibrated leverage to its spot (which = 0, per
// unit of spot) or to the parallel vol shift (which = 1, per unit of shift),
// as the reference's bump sees it: the leverage recalibrated at spot + 1 % or
// at vol shift + 1 vol point, read at the base layers' nodes, minus the base,
// over the bump.  dv[i P + k] for node k of date i's layer.
void LeverageTangent( const Market& mk, int a, const Grid& g, const LeverageGrid& L0, int which,
                      std::vector<double>& dv );
// the base scenario's device book (grid, Cholesky, LSV calibration, payoffs);
// throws on invalid input like the other host entry points
int ResolveBaseBook( const thoth_mcl_request& req, BookArrays& A, Grid& g );

struct Invalid : std::runtime_error
{
    int code;
    Invalid( int c, const std::string& s ) : std::runtime_error( s ), code( c ) {}
};

}  // namespace thoth::host
