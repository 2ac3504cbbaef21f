// kamino_runtime.h -- host runtime shared by the single-simulation path and the
// batching / serving engine: shape plans, per-shape device tables, and device
// batches (descriptor table + tile lists + launch sequence).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "kamino.h"
#include "kamino_device.h"

namespace kam {

// Records the message for kam_last_error() and returns `code`.
int set_err(int code, const char* fmt, ...);

#define KAM_HIP(call)                                                               \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess)                                                       \
            return ::kam::set_err(KAM_EHIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// Launch geometry of one simulation shape.
struct Plan {
    int nT = 0, nP = 0, M = 0;
    int nfac = 0;
    int fac[12] = {0};
    int advRows = 0;   // rows per advection tile
    int advCols = 0;   // columns per advection tile
    int fftRows = 0;   // rows per FFT tile
    int triL = 0;      // lanes per mode in the tridiagonal kernel
    int triK = 0;      // template row capacity per lane
    size_t advLds = 0; // bytes of LDS of an advection tile (geometric on)
    size_t fftLds = 0; // bytes of LDS of an FFT tile (forward or inverse)
    int transform = 0; // KAM_TRANSFORM_FFT, _FFT8, or _DFT (FP32-pressure simulations)
};

// Validates a shape and fills its plan; returns KAM_OK or KAM_EINVAL.
int make_plan(int nTheta, int nPhi, Plan& plan);
// KAM_OK when `precision` is known and supported at this shape.
int check_precision(int precision, const Plan& plan);
// Overrides the tile heights (0 keeps the default); used by the autotuner.
int adjust_plan(Plan& plan, int advRows, int fftRows);
// Records tile shapes for n_theta (0 = default); later plans use them.
int tune_set(int nTheta, int advRows, int advCols, int fftRows);
// Records the phi transform for n_theta (KAM_TRANSFORM_*; -1 = default FFT).
int tune_set_transform(int nTheta, int transform);

// Sets or clears SIM_DFT on a simulation's descriptor from its plan and
// precision, acquiring the shape's DFT matrices when needed (cached per
// device and nPhi for the life of the process).
int attach_transform(SimDesc& d, const Plan& p, int precision, int device);

// Device tables shared by every simulation of one (device, shape, dt, R).
struct Tables {
    float* rowf = nullptr;
    double* rowd = nullptr;
    float* twN = nullptr;
    double* twN64 = nullptr;
};
int acquire_tables(int device, int nTheta, double dt, double radius, Tables* out);
void release_tables(int device, int nTheta, double dt, double radius);

// Host values of the per-shape tables (no allocation): rowf[RT_COUNT * nT],
// rowd[RD_COUNT * nT], tw[2 nPhi] (FP32) and tw64[2 nPhi] (FP64).
void compute_tables(int nTheta, double dt, double radius, float* rowf, double* rowd, float* tw,
                    double* tw64);

// Fills the host descriptor of a simulation whose state lives at `mem`
// (7 * nTheta * nPhi floats: two state copies of three fields, then the
// spectrum).
void fill_desc(SimDesc& d, const Plan& p, const Tables& t, const kam_sim_desc& sd, float* mem);
// Sets d.advCols = W and the division constants of its advection tiles.
void set_tile_cols(SimDesc& d, int W);
size_t sim_floats(int nTheta, int nPhi);

// The launch sequence of one step over a set of simulations.
struct TriClass {
    int L, K, off, count;
    bool fp32;  // FP32 elimination (see launch_tridiag)
};

// Tile lists of one step over a set of simulations, as ranges of one array.
struct TileLayout {
    int nsims = 0, total = 0;
    int advOff = 0, nAdv = 0;
    int fftOff = 0, nFft = 0, nFft32 = 0, nFft64 = 0;  // FFT tiles (FP32-pressure ones first)
    int partOff = 0, nPart = 0;
    int dftRowOff = 0, nDftRow = 0, dftOff = 0, nDft = 0;  // SIM_DFT rows and MFMA product tiles
    int ntri = 0;
    TriClass tri[12] = {};
    size_t advLds = 0, fftLds = 0, fftLds64 = 0, divLds = 0;
    bool divW8 = false;  // forward transform by k_div_fft8 (half twiddle table, eight waves)
    int advStride = 0;  // staged row stride of the advection launch
};

// Plans the tiles of n simulations: writes the first min(cap, needed) tiles to
// `out` and returns how many the plan needs (uses no heap).  rowScratch holds
// 3n ints.  sims may be given narrower advection tiles (set_tile_cols) when
// the batch mixes tile widths.
int plan_tiles(SimDesc* sims, const Plan* plans, int n, int* rowScratch, Tile* out, int cap,
               TileLayout& layout);
// Queues the phases of one step over device descriptors and tiles planned as
// `layout`; returns the parity after the step (flipped iff the advection ran).
int launch_step(const SimDesc* dSims, const Tile* dTiles, const TileLayout& layout, int phases,
                int parity, hipStream_t st, int* newParity, hipEvent_t* kernelEvents = nullptr);
// Queues the tridiagonal phase alone (resid: with the per-mode residual report).
int launch_tri(const SimDesc* dSims, const Tile* dTiles, const TileLayout& layout, int parity,
               hipStream_t st, bool resid);

class DeviceBatch {
public:
    ~DeviceBatch();
    // Rebuilds tile lists for these simulations and uploads descriptors and
    // tiles on `st` (synchronously with respect to the host buffers).
    int build(const std::vector<SimDesc>& sims, const std::vector<Plan>& plans, hipStream_t st);
    // Queues the phases of one step.  `parity` selects the current state copy;
    // returns the parity after the step (flipped iff the advection ran).
    int launch(int phases, int parity, hipStream_t st, int* newParity,
               hipEvent_t* kernelEvents = nullptr);
    // Queues the tridiagonal phase alone, with the per-mode residual report
    // (`parity`: the current state copy).
    int launch_tri_residual(hipStream_t st, int parity);
    // Queues the inverse transform of the solved spectrum of a one-simulation
    // batch, writing its pressure to pout (device, float [nTheta][nPhi]); a
    // SIM_DFT simulation writes it over its finished state copy instead.
    int launch_pressure(hipStream_t st, int parity, float* pout);
    int nsims() const { return lay_.nsims; }
    size_t capacityBytes() const { return capSims_ * sizeof(SimDesc) + capTiles_ * sizeof(Tile); }

private:
    int reserve(size_t nsims, size_t ntiles, hipStream_t st);
    SimDesc* dSims_ = nullptr;
    Tile* dTiles_ = nullptr;
    size_t capSims_ = 0, capTiles_ = 0;
    // two pinned staging buffers used in turn; `uploaded_[b]` marks when the
    // copy out of buffer b is done, so a rebuild waits only for the upload
    // two rebuilds back (long done, normally) instead of the last one
    void* hStage_[2] = {nullptr, nullptr};
    size_t stageBytes_[2] = {0, 0};
    hipEvent_t uploaded_[2] = {nullptr, nullptr};
    int cur_ = 0;
    std::vector<Tile> hTiles_;
    TileLayout lay_;
};

// Kernel-time categories reported by profiling.
enum KernelCat { KC_ADVECT = 0, KC_DIVFFT = 1, KC_TRI = 2, KC_IFFT = 3, KC_COUNT = 4 };

int ensure_device(int device);
// Compute units of the last device made ready (for tile-count heuristics).
int device_cus();

}  // namespace kam
