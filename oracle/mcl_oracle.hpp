// mcl_oracle.hpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU FP64 restatement of the reference `mcl` Monte-Carlo engine (GValas/Thoth,
// pricer/src), used as the checker of the device library.  Only the tests,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
// product never links, loads or calls it.
//
// Every function follows the reference file:line cited on it, in the reference's
// order of floating-point operations, and the library is compiled with
// -ffp-contract=off (the reference Release build forms no FMAs).
//
// What checks each part independently of this restatement (published vectors,
// closed forms, invariants) is exercised by tests/test_oracle.py.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "thoth_mcl.h"

namespace oracle
{

constexpr double NB_OF_DAYS_A_YEAR = 365.0;          // helpers/constants.hpp:7
constexpr double GREEK_SPOT_BUMP = 0.01;             // helpers/constants.hpp:36
constexpr double GREEK_GAMMA_BUMP = 0.10;            // tasks/pricer.hpp:20
constexpr double GREEK_VOL_BUMP = 0.01;              // tasks/pricer.hpp:25
constexpr double GREEK_RATE_BUMP = 0.0001;           // tasks/pricer.hpp:26
constexpr double IMPLICIT_VOL_MAX = 20.0;            // helpers/finance.hpp:206

// ---- leaves ---------------------------------------------------------------
bool LoadJoeKuo( const std::string& path );
unsigned SobolMaxDimension();                        // sobol_generator.cpp:12-15

class Sobol                                          // sobol_generator.cpp:22-121
{
  public:
    explicit Sobol( unsigned dim );
    void Next( std::vector<double>& point );
    void NextInt( std::vector<uint32_t>& x );        // integer stage of Next (no scaling)
    void Skip( uint64_t count );
    unsigned Dimension() const { return _dim; }
    const std::vector<std::vector<uint32_t>>& V() const { return _v; }

  private:
    unsigned _dim;
    uint64_t _count = 0;
    std::vector<std::vector<uint32_t>> _v;
    std::vector<uint32_t> _x;
};

double NormalCdf( double x );                        // distributions.cpp:21-26
double NormalCdfInv( double p );                     // distributions.cpp:37-81

class Rng                                            // helpers/rng.hpp:19-72
{
  public:
    explicit Rng( uint64_t seed = 0 ) { Seed( seed ); }
    void Seed( uint64_t seed );
    uint64_t operator()();
    double Uniform();
    double Gaussian() { return NormalCdfInv( Uniform() ); }

  private:
    uint64_t _s[4];
};

struct BridgeSchedule                                // path_generator.cpp:49-120
{
    std::vector<int> bridge_index, left_index, right_index;
    std::vector<double> left_weight, right_weight, std_dev;
};
BridgeSchedule BuildBridgeSchedule( const std::vector<double>& times );

class PathGenerator                                  // path_generator.cpp:11-198
{
  public:
    // use_sobol false: every normal from rng, step-major and factor-minor,
    // with no bridge (the reference's noise nodes draw in node-schedule
    // order instead, noise_node.cpp:20-25; same law, other assignment)
    PathGenerator( const std::vector<double>& times, int factors, Rng* rng, uint64_t sobol_skip,
                   bool use_sobol = true );
    void NextPath();
    const std::vector<double>& Buffer( int f ) const { return _noise[f]; }
    int SobolDim() const { return _sobol_dim; }

  private:
    std::vector<double> _t;
    int _m, _factors;
    bool _use_sobol;
    Rng* _rng;
    Sobol* _sobol = nullptr;
    std::vector<Sobol> _sobol_store;
    int _sobol_dim = 0;
    BridgeSchedule _b;
    std::vector<std::vector<double>> _noise, _z;
    std::vector<double> _u, _w;
};

bool CholeskyDecomposeLower( std::vector<double>& a, int n );   // helpers/maths.cpp:354-381

// ---- market ----------------------------------------------------------------
struct Curve                                         // marketdata/curve.cpp:45-73
{
    std::vector<double> days, values;                // values already decimal
    double shift = 0;                                // live rho shift
    double Value( double day ) const;                // GetCurveValue
};

struct Sabr                                          // vol_correl/sabr_volatility.cpp
{
    std::vector<double> maturities, alpha, beta, rho, nu;
    double Interp( const std::vector<double>& values, double t ) const;
    double HaganVol( double k, double f, double t ) const;
    double WingVol( double k, double f, double t, double cutoff, bool right ) const;
    double ImplicitVol( double strike, double forward, double day ) const;  // GetImplicitVol
};

struct Equity
{
    double spot;
    int vol_kind;
    double bs_vol;                                   // decimal
    double v0, theta, kappa, xi, rho;                // raw (unshifted) Heston fields
    double div;                                      // decimal continuous dividend yield
    bool has_div;
    Sabr surface;
    double vol_shift = 0;                            // vega bump (Volatility::_vol_shift)
};

struct StochParams { double v0, kappa, theta, xi, rho; };

struct Market
{
    std::vector<Equity> eq;
    Curve rate;
    Curve discount;                                  // currency.cpp:19-27 discount_rate (OIS)
    bool multi_curve = false;                        // HasDistinctDiscountCurve
    std::vector<double> correlation;
    // term-structured correlation (correlation.cpp:35-130): pillar maturities
    // (years) and one n x n matrix per pillar; IsTermStructured = > 1 pillar
    std::vector<double> corr_pillars;
    std::vector<std::vector<double>> corr_mats;
    bool TermStructured() const { return corr_pillars.size() > 1; }
    double AvgCorrelation( int i, int j, double a, double b ) const;  // Correlation::AvgEntry
    int lev_spot_points = 0, lev_time_layers = 0;    // 0, 0: one 201-point layer per diffusion date
    double today = 0;                                // days
    StochParams Stoch( int a ) const;                // StochasticParams()
    double ImplicitVol( int a, double strike, double day ) const;   // Single::GetImplicitVol
    double Forward( int a, double day ) const;       // Equity::GetForward
    double DivRepo( int a, double day ) const;       // Equity::DividendRepoYield
    double LocalVol( int a, double strike, double day ) const;      // Equity/Volatility::GetLocalVolatility
};

struct LeverageLayers                                // leverage_surface.hpp
{
    std::vector<double> t, ln_step;
    std::vector<long> offset;
    std::vector<std::vector<double>> levels;
    bool surface = false;                            // read by time (GetLeverage), not by date index
    double LayerAt( size_t k, double s ) const;      // leverage_surface.cpp:19-40
    double GetLeverage( double s, double T ) const;  // leverage_surface.cpp:43-66
};
LeverageLayers CalibrateLeverage( const Market& mk, int a, const std::vector<double>& date_days );

// ---- engine -----------------------------------------------------------------
struct Grid
{
    std::vector<double> days;                        // diffusion dates (days from today)
    std::vector<double> t, dt;                       // YearFraction(today, d_i), (d_{i-1}, d_i)
    bool uniform = false;
};

struct BookSums
{
    double book_sum = 0, book_sum2 = 0, book_c = 0, book_c2 = 0;
    std::vector<double> sum, sum2, c, c2;
    long count = 0;
    // recorded spot paths of the equities American contracts exercise on
    // (path_recorder: every diffusion date), rec[a] = [paths][n_dates], empty
    // for the other equities
    std::vector<std::vector<double>> rec;
    double Mean( int i ) const;                      // i < 0: book
    double Trust( int i ) const;
};

// The fully resolved inputs of one swept tree (what the device book receives).
struct ResolvedBook
{
    Grid grid;
    int n_assets = 0, factors = 0;
    std::vector<int> model;
    std::vector<double> spot, vol, v0, kappa, theta, xi, rho;
    std::vector<std::vector<double>> drift;          // [asset][date]
    std::vector<double> chol;
    std::vector<std::vector<double>> chol_dates;     // per-date factors (term-structured correlation)
    bool correlate_stochastic = false;
    std::vector<LeverageLayers> lev;                 // per asset (LSV only)
    std::vector<double> disc_rate;
    struct Rainbow { bool best; std::vector<int> members; std::vector<double> weight; };
    std::vector<Rainbow> rainbows;
    struct Contract
    {
        int kind, is_call, underlying, flow_index;
        double strike, level;
        bool is_up, is_in, american;
        std::vector<int> monitor;
    };
    std::vector<Contract> contracts;
};

Grid BuildGrid( const thoth_mcl_request& req, const Market& mk );
Market BuildMarket( const thoth_mcl_request& req );
ResolvedBook ResolveBook( const thoth_mcl_request& req, const Market& mk, const Grid& g );
BookSums SweepBook( const ResolvedBook& b, int64_t paths, uint64_t sobol_skip, int seed, bool use_sobol = true );
int Price( const thoth_mcl_request& req, thoth_mcl_result* res );

// ---- American Longstaff-Schwartz (pricer_mcl_american.cpp:250-447) -----------
// ordinary least squares by the normal equations and Cholesky (maths.cpp:419-470);
// X is [n][p] row-major; empty when X^T X is not positive definite
std::vector<double> LeastSquares( const std::vector<double>& X, size_t n, size_t p, const std::vector<double>& y );
struct AmericanPolicy
{
    double basis_norm = 1;
    std::vector<double> b0, b1, b2;  // continuation = b0 + b1 m + b2 m^2, m = S / basis_norm
    std::vector<char> has_fit;
};
// paths [N][M] (column 0 today, M - 1 maturity), tau and zero rates per column
AmericanPolicy FitAmericanPolicy( double strike, bool is_call, const std::vector<double>& paths, size_t N, size_t M,
                                  const std::vector<double>& tau, const std::vector<double>& rates );
double ApplyAmericanPolicy( double strike, bool is_call, const std::vector<double>& paths, size_t N, size_t M,
                            const std::vector<double>& tau, const std::vector<double>& rates,
                            const AmericanPolicy& pol, double& trust );

}  // namespace oracle
