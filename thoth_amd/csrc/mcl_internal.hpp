// mcl_internal.hpp -- host-side declarations shared by the device library's
// translation units (not part of the public C-ABI).
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "mcl_device.hpp"
#include "thoth_mcl_gpu.h"

namespace thoth
{

constexpr int kSobolDims = 4097;  // implicit first dimension + 4096 Joe-Kuo rows

struct BridgeProgram
{
    double sd0 = 0;
    std::vector<int> step_ops;  // [m + 1]
    std::vector<int> op_b, op_left;
    std::vector<double> op_w;   // [n_ops * 3]
    int max_depth = 1;
};

const std::vector<uint32_t>& SobolDirections();
BridgeProgram BuildBridgeProgram( const double* t, int n_dates );
int EnsureDevice();
std::mutex& DeviceMutex();  // serialises every public entry point
hipStream_t Stream();
hipEvent_t SweepEvent( int k );  // the two events that time a sweep (created at set-up)
const uint32_t* DeviceDirections();
// set-up of the no-allocation entry point (mcl_noalloc.hip): the device
// address of its static scratch array on the current device
int NoAllocSetup();
int SetError( int code, const std::string& msg );
// pseudo-random stream (mcl_prng.cpp): the splitmix64 expansion of a seed,
// the jump matrices (M^T)^(2^k), k < levels (cached by T; callers hold the
// library lock), and the levels that cover path indices 0 .. paths - 1
void SplitMixSeed( uint64_t seed, uint64_t st[4] );
const std::vector<uint64_t>& JumpTables( uint64_t T, int levels );
int JumpLevels( int64_t paths );
double StudentT975( int df );  // two-sided 95 % quantile of Student's t
struct DevBook;
// A packed, uploaded batch of books ready to sweep (device-resident inputs).
struct DeviceBatch
{
    char* dbase = nullptr;
    bool external = false;  // dbase is caller-provided memory (not freed here)
    double* dpart = nullptr;
    double* hpart = nullptr;  // pinned host copy of the partial sums (null: a pageable one per run)
    const DevBook* dbooks = nullptr;
    const uint32_t* dshifts = nullptr;  // [replicates][kSobolDims] in RQMC mode
    int replicates = 1;
    std::vector<DevBook> db;
    std::vector<size_t> part_off;
    size_t part_total = 0;
    uint32_t q0 = 0, nblocks = 0, first = 0, last = 0;  // q0: first 256-index block
    size_t lds_bytes = 0;
    int n_books = 0, depth = 1, na_max = 1, precision = 0;
    int kind = 0;  // uniform-book kernel: 1 all LSV, 2 all Heston (see sweep_block), 0 generic
    int cv_mode = 0;  // 0 none, 1 control-variate samples, 2 in-loop hedge inference
    int cv_act = 0;   // hidden activation of the in-loop network: THOTH_CV_SILU or THOTH_CV_GATE
    // pathwise-Greek sweep: 1 the single-equity kernel (three estimators per
    // block), 2 the basket kernel (3 + n_assets estimators per block)
    int pathwise = 0;
    // vanilla grids: per-(replicate, block) partials and their per-book sums
    // [nm][2 nk][2] (device), offsets per book; grid_cells[b] = nm * 2 nk
    double* dgrid = nullptr;
    double* dgrid_out = nullptr;
    std::vector<size_t> grid_off, grid_out_off;
    std::vector<int> grid_cells;
    const long long* dgrid_table = nullptr;  // [books][3]: partial offset, output offset, doubles per slot
    size_t grid_total = 0, grid_out_total = 0;
    // dynamic-LDS recipe: Sobol table words, leverage words per staged layer,
    // fixed tail bytes (hedge-network or grid area)
    size_t lds_dims = 0, lds_lev_words = 0, lds_tail = 0;
    // host copies of the books' bridge step offsets (Sobol ring sizing) and
    // the factor stride NF = 2 NA of the kernel instantiation
    std::vector<std::vector<int>> step_ops;
    int nf_inst = 2;
    int64_t paths = 0;
    // pseudo-random draws: prng when any book draws from the stream (kind 4
    // of the sweep), pseudo for use_sobol false; per distinct T the start
    // states of the paths (jump: device tables, states: [4][paths]) and, in
    // Sobol mode, the tail normals drawn from them (dst: [T][paths])
    bool prng = false, pseudo = false;
    bool recording = false;  // some book records a path for American exercise
    int jump_levels = 0;
    uint64_t seed_state[4] = { 0, 0, 0, 0 };
    struct TailFill
    {
        int64_t T;
        int book;
        const uint64_t* jump;
        uint64_t* states;
        char* dst;  // nullptr for use_sobol false (no tail)
    };
    std::vector<TailFill> fills;
    // launch shape: one workgroup per Sobol block of every book (persist_wpc
    // = 0), or a persistent grid of persist_wpc workgroups per CU pulling
    // (book, block) items (the same per-block partial sums, so the same
    // result bit for bit); the item list and counter are device-resident
    int persist_wpc = 0, cus = 0, n_items = 0;
    // the equity-wave latency form (FP32 uniform LSV books with vanilla
    // contracts): waves_ok when it applies (the partial sums then have room
    // for its 4 tiles per Sobol block), wave_split when it is the form in
    // effect (chosen where the lane form would leave the SIMDs under two
    // waves each, or by the launch shape)
    bool waves_ok = false, wave_split = false;
    // pathwise basket Greeks of uniform FP32 LSV books: the staged-leverage
    // form (its ring of float4 nodes counted in the LDS bytes)
    bool pw_lsv = false;
    int2* ditems = nullptr;
    int* dcounter = nullptr;
    DeviceBatch() = default;
    DeviceBatch( const DeviceBatch& ) = delete;
    DeviceBatch& operator=( const DeviceBatch& ) = delete;
    ~DeviceBatch();
};
// Learned control variate of an FP32 sweep over J hedge intervals.
// mode 1 writes the training samples into device buffers sized
// [replicates][J][N][16] / [replicates][J][N][8] / [replicates][N];
// mode 2 evaluates the hedge network (BF16 A-operand image + FP32 biases,
// device pointers) inside the stepping loop and accumulates P - CV.
struct CvSpec
{
    int mode = 1;
    int act = 0;  // hidden activation of the network (mode 2): THOTH_CV_SILU or THOTH_CV_GATE
    int J = 0;
    float spacing = 1.0f;  // hedge date j at step m - round(m (1 - j/J)^spacing)
    int64_t N = 0;
    uint16_t* feat = nullptr;
    float* dm = nullptr;
    float* pay = nullptr;
    const uint16_t* wimg = nullptr;
    const float* bias = nullptr;
};
// Vanilla grid of one book: calls and puts of nk absolute strikes at nm
// maturities (date indices of the book's grid) on underlying udl (equity
// index, or -(1 + rainbow index)).
struct GridSpec
{
    int nm = 0, nk = 0, udl = 0;
    int mat_date[kGridMaxMat] = {};
    double strikes[kGridMaxStrikes] = {};
    // grouped: the book is several single-equity LSV books on one grid and one
    // set of Sobol points, one per equity, stepped together on two shared
    // factors (sweep kind 5); every equity has a grid with its own strikes
    int grouped = 0;
    double strikes_by_equity[kMaxAssets][kGridMaxStrikes] = {};
};
// sums of P - CV per replicate (mode 2)
struct CvSums
{
    double sum = 0, sum2 = 0;
};
// Path recording of one book for American exercise: the spot of equity udl at
// dates 0 .. n - 1 (n = 0: none) into [n][paths] device memory of the batch
// (DevBook::rec / recf)
struct RecordSpec
{
    int udl = -1, n = 0;
};
// Leverage sensitivities of one book for the pathwise basket kernel (host
// arrays laid out as DevBook::lev_ds / lev_de; null for zeros)
struct PathwiseLayers
{
    const double* ds = nullptr;
    const double* de = nullptr;
};
// ext_base / ext_cap: device memory of the caller to hold the whole batch
// (the no-allocation entry point's scratch) instead of a hipMalloc; a batch
// that needs more than ext_cap bytes is refused before any device work
int PrepareBatch( int32_t n_books, const thoth_mcl_book* books, const thoth_mcl_sweep_cfg* cfg, DeviceBatch& B,
                  const CvSpec* cv = nullptr, const GridSpec* grids = nullptr, int pathwise = 0,
                  char* ext_base = nullptr, size_t ext_cap = 0, const RecordSpec* rec = nullptr,
                  const PathwiseLayers* pwl = nullptr );
// `then`, when given, enqueues further work on the library stream right after
// the sweep (before the one synchronisation that returns the sums)
int RunBatch( const DeviceBatch& B, thoth_mcl_book_out* out, CvSums* cv_out = nullptr,
              const std::function<int( hipStream_t )>& then = {} );
// Launch shape of a plan's batch (reference-mode sweeps without the control
// variate): steps per LDS barrier G and workgroups per CU of the persistent
// form (0: one workgroup per Sobol block, -1: the equity-wave form, where it
// applies).  SetShape applies one (refusing
// shapes whose LDS does not fit or that have no kernel instantiation);
// TuneShape times every candidate (G in {1, 2, 4, 8} x launch forms) twice
// on the batch itself, keeps the fastest, and caches it by book shape.
struct ShapeChoice
{
    int lev_group = 4, wpc = 0;
    double ms = 0;
    int candidates = 0;   // timed (0: taken from the cache)
    double default_ms = 0;  // the shape in effect before tuning (0 when cached)
};
int SetShape( DeviceBatch& B, int lev_group, int wpc );
int TuneShape( DeviceBatch& B, ShapeChoice& out );
// grid sums of every book, [nm][2 nk][2] each, from the last sweep (device reduction + copy)
int ReduceGrids( const DeviceBatch& B, hipStream_t s );
// re-stage the leverage layers G per barrier (ring of 2G): updates the device
// books and the dynamic-LDS size
int SetLevGroup( DeviceBatch& B, int G );
int CopyGrids( const DeviceBatch& B, std::vector<std::vector<double>>& out );
int SweepBooks( int32_t n_books, const thoth_mcl_book* books, const thoth_mcl_sweep_cfg* cfg,
                thoth_mcl_book_out* out );

}  // namespace thoth
