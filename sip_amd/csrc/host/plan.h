// FFT plan: the host-side state behind an asdFftHandle, and the planner that lays a transform out
// as a sequence of kernel stages.
//
// Reference counterparts: FFTPlan / PlanStep (core/fft/fftplan/include/fftplan/fft_plan.h:43-70),
// the plan builders addFFTSteps / init2DSteps / init3DSteps (core/fft/fft_api.cpp:485-697) and the
// workspace contract (fft_api.cpp:699-733, 989-1045). The decomposition differs on purpose: the
// reference chains whole-array steps (a DFT-matrix core per axis, with AsStrided transposes in
// between); here every axis is one strided Stockham launch that reads and writes in place, long
// axes become four-step launch pairs, and lengths with large prime factors use Bluestein.
#pragma once
#include <cstdint>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../kernels/aux_args.h"
#include "../kernels/stockham_args.h"
#include "common.h"

namespace sip {

constexpr int kTypeC2C = 0x10, kTypeC2R = 0x11, kTypeR2C = 0x12, kTypeC2CSep = 0x30;
constexpr int kDirForward = 0x10, kDirInverse = 0x11;
constexpr int kDimHorizontal = 0x10, kDimVertical = 0x11;
constexpr int64_t kMaxFftSize = int64_t(1) << 27;

// Device-resident coefficient table (twiddles, roots, chirps), shared between plans.
struct DeviceTable {
    void *ptr = nullptr;
    size_t bytes = 0;
    int device = -1;
    ~DeviceTable();
};
using TablePtr = std::shared_ptr<DeviceTable>;

// A table is identified by a key (its content is a pure function of the key) and generated on
// the host in FP64 / long double, rounded once to FP32.
struct TableRef {
    std::string key;
    std::function<std::vector<float2>()> gen;
    TablePtr dev;
    bool used() const { return !key.empty(); }
};

// Returns the cached device table for ref.key, creating it from ref.gen when absent
// (counterpart of FFTensorCache::getCoeff, core/fft/utils/fftensor_cache.cpp:80-92).
TablePtr coeff_get(const TableRef &ref);
size_t coeff_cache_live_tables();
// Map entries (live or expired), live tables and their bytes, over every device.
void coeff_cache_stats(size_t *entries, size_t *live, size_t *live_bytes);

enum BufKind : int { BUF_IN = 0, BUF_OUT = 1, BUF_WS = 2, BUF_ABS = 3 };

struct BufRef {
    int kind = BUF_IN;
    int ws_index = 0;  // workspace region (BUF_WS) or absolute-pointer slot (BUF_ABS)
};

// Per-plan tuning choices (0 / -1 = automatic). The autotuner of the batching layer searches them.
struct Tune {
    int row_F = 0;         // transforms per workgroup for row launches
    int split_lds = -1;    // single-plane LDS row kernel
    int xcd = -1;          // XCD-contiguous tile order
    int64_t split_l1 = 0;  // forced first factor of a four-step split
    int col_F = 0;         // transforms per workgroup for column (strided) launches
    int rc = -1;           // rows-in / columns-out kernel for the second four-step stage
};

enum class StageKind { Stockham, Aux };

struct Stage {
    StageKind kind = StageKind::Stockham;
    BufRef in, out;
    // Stockham
    const StockhamInstance *inst = nullptr;
    int sign = -1;
    StockhamArgs sa{};
    dim3 grid, block;
    size_t lds = 0;
    TableRef tw, rtw, tw4lo, tw4hi, mfw;
    // fused 2D tile stage (both axes of X x Y slices in one launch); tw2 = column twiddles
    const Stockham2dInstance *inst2d = nullptr;
    TableRef tw2;
    // rows-in / columns-out kernel (four-step second stage: contiguous rows, transposed store)
    bool rc = false;
    int in_tab = -1, out_tab = -1;  // row tables (Plan::row_tables) gathered by the in / out side
    // Aux
    AuxArgs xa{};
    TableRef tab0;
    std::string note;
};

struct Plan {
    int fftType = kTypeC2C;
    int direction = kDirForward;
    int dimType = kDimHorizontal;
    std::vector<int64_t> sizes;   // FFT sizes (1..3 dims)
    std::vector<int64_t> strides; // 1 unless vertical
    int64_t batch = 1;
    hipStream_t stream = nullptr;
    void *workspace = nullptr;
    bool initialized = false;
    bool failed = false;
    int device = 0;

    std::vector<Stage> stages;
    std::vector<size_t> ws_offsets;  // byte offset of each workspace region
    size_t ws_bytes = 0;

    std::mutex mu;

    // Gathered-row plans (batching layer): absolute base pointers and row-offset tables.
    std::vector<void *> abs_ptrs;
    std::vector<std::vector<int64_t>> row_tables;
    std::vector<void *> row_tables_dev;
    std::vector<const int64_t *> external_tables;  // caller-owned device row tables (serving layer)
    Tune tune;
    int precision = 0;  // PrecisionMode of the radix-16 stages of C2C launches
    // Single-pass leaves for lengths 16384..32768 from 256 transforms up. A caller whose results
    // must not depend on how many rows share a launch (the serving form's coalescing) turns them
    // off: the four-step split is then taken at every row count.
    bool big_leaf = true;
    // Transforms per launch from which those leaves replace the split; 0 takes the plain-plan
    // default (256). The batching form's groups run concurrently on several chains and set 64.
    int64_t big_leaf_rows = 0;
    // ISTFT plan (asdFftIstftMakePlan): inverse C2C of the (channel, n_fft, frames) spectra into a
    // frame workspace, then the windowed overlap-add into (channel, out_len)
    struct Istft {
        bool active = false;
        int64_t channel = 0, nfft = 0, frames = 0, hop = 0, win_len = 0, out_len = 0;
    } istft;

    int sign() const { return direction == kDirInverse ? 1 : -1; }
    ~Plan();
};

// 1D plan over rows gathered from arbitrary places: row r of the input starts at element
// in_rows[r] of in_base, row r of the output at out_rows[r] of out_base (elements of each buffer's
// own type: complex for C2C and R2C outputs, float for R2C inputs). C2C: any length (auxiliary
// stages such as Bluestein's chirp steps and the length-1 copy read and write through the same row
// tables); R2C: lengths whose (half) transform is one leaf launch (gathered_supported).
void plan_build_gathered(Plan &p, int type, int64_t n, const std::vector<int64_t> &in_rows,
                         const std::vector<int64_t> &out_rows, void *in_base, void *out_base);
void plan_build_gathered_c2c(Plan &p, int64_t n, const std::vector<int64_t> &in_rows,
                             const std::vector<int64_t> &out_rows, void *in_base, void *out_base);
bool gathered_supported(int type, int64_t n);
// plan_materialize with the row tables already on the device (one pointer per Plan::row_tables).
void plan_materialize_external(Plan &p, const std::vector<const int64_t *> &dev_tables);

// Lays the plan out (stages, workspace) from its type / sizes / batch. Host only; throws SipError.
void plan_build(Plan &p);
// Creates the device-side coefficient tables and kernel attributes. Throws SipError.
void plan_materialize(Plan &p);
// Sets each stage's nontemporal-output flag (no later stage reads its output); plan_materialize
// calls it, host-only descriptions call it directly.
void plan_mark_streaming(Plan &p);
// True when every stage's launch shape (threads, static + dynamic LDS) is within what the device
// accepts for its kernel; checked before a tuning candidate is materialized or launched.
bool plan_launchable(const Plan &p);
// Enqueues the stages on p.stream. in/out are device addresses of the user tensors.
void plan_execute(Plan &p, const void *in, void *out);
// Split-plane (C2C_SEP) execution: user buffers are real / imaginary float planes.
void plan_execute_sep(Plan &p, const void *in, const void *in_im, void *out, void *out_im);
// ISTFT execution: `window` (n_fft floats) feeds the overlap-add stage.
void plan_execute_istft(Plan &p, const void *in, const float *window, void *out);
// JSON description of the stages.
std::string plan_describe(const Plan &p);

// Kernel registry (generated instances).
const StockhamInstance *find_instance(int64_t L);
// Split-precision MFMA instances (mode = PrecisionMode) and their DFT-matrix operand table.
const StockhamInstance *find_mfma_instance(int64_t L, int mode);
// Fused 2D tile instance for X x Y slices (X, Y in 32, 64, 128), or null.
const Stockham2dInstance *find_2d_instance(int64_t X, int64_t Y);
TableRef mfma_dft_table(int mode, int sign);
// Operand table of the single-plane K = 32 matrix passes (kernels/mfma_split.h).
TableRef mfma2_dft_table(int mode, int sign);
// Pass twiddles of an arbitrary radix sequence of length L (the leaf kernels' layout; the table of
// a leaf with the same sequence is the same table), and the rows per k of a radix-R pass.
TableRef pass_twiddles_seq(int64_t L, const std::vector<int> &radices, int sign);
int64_t pass_twiddle_rows(int R);
int64_t instance_count();

// Launchers implemented in the .hip files.
hipError_t aux_launch(const AuxArgs &a, hipStream_t stream);

}  // namespace sip
