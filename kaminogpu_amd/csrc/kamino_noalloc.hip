// kamino_noalloc.hip -- kam_sim_step_noalloc: full steps of one simulation
// whose fields are caller-owned device buffers, run out of a fixed per-device
// scratch array.  Nothing on its call path allocates: no hipMalloc / hipFree,
// no pinned or heap host memory (host staging is static storage).
//
// The scratch array holds, for the shape of the last call on that device:
//   [descriptor][tile list][row tables FP32][row tables FP64][twiddles FP32]
//   [twiddles FP64][state: 2 copies x 3 fields, spectrum, watchdog word]
// each part 256-byte aligned.  A call copies the caller's fields into state
// copy 0, steps, and copies the result back.  Every call on a device shares
// its array, so calls that overlap on one device (other threads, or other
// streams with work still queued) race; calls queued on one stream are
// ordered by that stream.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kamino_launch.h"
#include "kamino_runtime.h"

namespace kam {
namespace {

constexpr size_t kScratchBytes = (size_t)KAM_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 16;
constexpr int kMaxTiles = 4096;
constexpr int kMaxTheta = 2048;  // the library's n_theta limit (host table staging)

// One copy per device that loads the library.  In host code this name is a
// stand-in, not a device address: host code uses the address that
// hipGetSymbolAddress returns for the current device (DevScratch::base).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Byte offsets of the parts of the scratch array for one shape.
struct ScratchLayout {
    size_t desc, tiles, rowf, rowd, tw, tw64, state, end;
};

ScratchLayout layout_for(int nT, int ntiles) {
    const int nP = 2 * nT;
    ScratchLayout L;
    L.desc = 0;
    L.tiles = align256(sizeof(SimDesc));
    L.rowf = L.tiles + align256((size_t)ntiles * sizeof(Tile));
    L.rowd = L.rowf + align256((size_t)RT_COUNT * nT * sizeof(float));
    L.tw = L.rowd + align256((size_t)RD_COUNT * nT * sizeof(double));
    L.tw64 = L.tw + align256((size_t)2 * nP * sizeof(float));
    L.state = L.tw64 + align256((size_t)2 * nP * sizeof(double));
    L.end = L.state + align256(sim_floats(nT, nP) * sizeof(float));
    return L;
}

// What a device's scratch array holds now.
struct DevScratch {
    bool ready = false;
    unsigned char* base = nullptr;  // this device's address of `scratch`
    bool loaded = false;
    int nT = 0, geometric = 0, precision = 0, advRows = 0, advCols = 0, fftRows = 0;
    double dt = 0.0, radius = 0.0;
    TileLayout tiles;
    SimDesc desc;  // the uploaded descriptor (device pointers into base)
};

DevScratch g_dev[kMaxDevices];
std::mutex g_mu;

// Host staging of one upload; static storage guarded by g_mu, refilled only
// after the previous upload has landed.
struct Stage {
    SimDesc desc;
    int rows[3];
    Tile tiles[kMaxTiles];
    float rowf[RT_COUNT * kMaxTheta];
    double rowd[RD_COUNT * kMaxTheta];
    float tw[4 * kMaxTheta];
    double tw64[4 * kMaxTheta];
};
Stage g_stage;

// Up to four float ranges to clear.
struct ZeroRanges {
    float* p[4];
    int n[4];
};

// Zero-fills the ranges with vector stores (a clear that also works where a
// caller captures the stream into a graph).
__global__ __launch_bounds__(256) void k_zero(ZeroRanges z) {
    const int stride = gridDim.x * blockDim.x;
    for (int r = 0; r < 4; ++r)
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < z.n[r]; i += stride) z.p[r][i] = 0.f;
}

// Loads the descriptor, tiles and tables of this shape into the device's
// scratch array (when it holds another shape), and waits for the upload.
int load_shape(DevScratch& ds, const kam_sim_desc& d, const Plan& plan, hipStream_t st) {
    if (ds.loaded && ds.nT == d.n_theta && ds.dt == d.dt && ds.radius == d.radius &&
        ds.geometric == d.geometric && ds.precision == d.precision && ds.advRows == plan.advRows &&
        ds.advCols == plan.advCols && ds.fftRows == plan.fftRows)
        return KAM_OK;
    const int nT = plan.nT, nP = plan.nP;
    if (nT > kMaxTheta) return set_err(KAM_EINVAL, "n_theta %d exceeds %d", nT, kMaxTheta);
    // the tile count first: the tiles depend on the shape and flags, not on
    // where the tables and state sit
    Tables t;
    unsigned char* base = ds.base;
    fill_desc(g_stage.desc, plan, t, d, reinterpret_cast<float*>(base));
    TileLayout tl;
    const int need = plan_tiles(&g_stage.desc, &plan, 1, g_stage.rows, g_stage.tiles, kMaxTiles, tl);
    if (need > kMaxTiles)
        return set_err(KAM_EINVAL, "n_theta %d needs %d tiles; the no-allocation step holds %d", nT, need,
                       kMaxTiles);
    const ScratchLayout L = layout_for(nT, need);
    if (L.end > kScratchBytes)
        return set_err(KAM_EINVAL,
                       "n_theta %d needs %zu bytes of scratch; the no-allocation step holds %zu (%d MiB)",
                       nT, L.end, kScratchBytes, (int)(kScratchBytes >> 20));
    t.rowf = reinterpret_cast<float*>(base + L.rowf);
    t.rowd = reinterpret_cast<double*>(base + L.rowd);
    t.twN = reinterpret_cast<float*>(base + L.tw);
    t.twN64 = reinterpret_cast<double*>(base + L.tw64);
    fill_desc(g_stage.desc, plan, t, d, reinterpret_cast<float*>(base + L.state));
    plan_tiles(&g_stage.desc, &plan, 1, g_stage.rows, g_stage.tiles, kMaxTiles, tl);
    compute_tables(nT, d.dt, d.radius, g_stage.rowf, g_stage.rowd, g_stage.tw, g_stage.tw64);
    ds.loaded = false;
    KAM_HIP(hipMemcpyAsync(base + L.desc, &g_stage.desc, sizeof(SimDesc), hipMemcpyHostToDevice, st));
    KAM_HIP(hipMemcpyAsync(base + L.tiles, g_stage.tiles, (size_t)need * sizeof(Tile), hipMemcpyHostToDevice,
                           st));
    KAM_HIP(hipMemcpyAsync(t.rowf, g_stage.rowf, (size_t)RT_COUNT * nT * sizeof(float),
                           hipMemcpyHostToDevice, st));
    KAM_HIP(hipMemcpyAsync(t.rowd, g_stage.rowd, (size_t)RD_COUNT * nT * sizeof(double),
                           hipMemcpyHostToDevice, st));
    KAM_HIP(hipMemcpyAsync(t.twN, g_stage.tw, (size_t)2 * nP * sizeof(float), hipMemcpyHostToDevice, st));
    KAM_HIP(hipMemcpyAsync(t.twN64, g_stage.tw64, (size_t)2 * nP * sizeof(double), hipMemcpyHostToDevice,
                           st));
    KAM_HIP(hipStreamSynchronize(st));
    ds.desc = g_stage.desc;
    ds.tiles = tl;
    ds.nT = nT;
    ds.dt = d.dt;
    ds.radius = d.radius;
    ds.geometric = d.geometric;
    ds.precision = d.precision;
    ds.advRows = plan.advRows;
    ds.advCols = plan.advCols;
    ds.fftRows = plan.fftRows;
    ds.loaded = true;
    return KAM_OK;
}

}  // namespace
}  // namespace kam

using namespace kam;

extern "C" int kam_sim_step_noalloc(const kam_sim_desc* d, float* u_phi, float* u_theta, float* rho,
                                    int nsteps, void* stream) {
    if (!d || !u_phi || !u_theta) return set_err(KAM_EINVAL, "null argument");
    if (nsteps < 0) return set_err(KAM_EINVAL, "negative step count");
    if (!(d->dt > 0.0) || !(d->radius > 0.0)) return set_err(KAM_EINVAL, "bad dt / radius");
    if (d->device < 0 || d->device >= kMaxDevices)
        return set_err(KAM_EINVAL, "device %d: the no-allocation step serves devices 0..%d", d->device,
                       kMaxDevices - 1);
    Plan plan;
    int rc = make_plan(d->n_theta, d->n_phi, plan);
    if (rc != KAM_OK) return rc;
    rc = check_precision(d->precision, plan);
    if (rc != KAM_OK) return rc;
    plan.transform = KAM_TRANSFORM_FFT;  // the DFT's matrices would need an allocation
    rc = ensure_device(d->device);  // makes the device current
    if (rc != KAM_OK) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lk(g_mu);
    DevScratch& ds = g_dev[d->device];
    if (!ds.ready) {
        // once per device, with that device current: this device's copy of the array
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_err(KAM_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        ds.base = static_cast<unsigned char*>(p);
        ds.ready = true;
    }
    rc = load_shape(ds, *d, plan, st);
    if (rc != KAM_OK) return rc;
    if (nsteps == 0) return KAM_OK;
    const int nT = plan.nT, nP = plan.nP;
    const size_t N = (size_t)nT * nP, Nt = N - nP;
    const SimDesc& sd = ds.desc;
    // state copy 0 <- the caller's fields; u_theta's padding row (both copies)
    // and, without rho, the density read as zero
    KAM_HIP(hipMemcpyAsync(sd.state[0][F_UPHI], u_phi, N * sizeof(float), hipMemcpyDeviceToDevice, st));
    KAM_HIP(hipMemcpyAsync(sd.state[0][F_UTHETA], u_theta, Nt * sizeof(float), hipMemcpyDeviceToDevice, st));
    if (rho) KAM_HIP(hipMemcpyAsync(sd.state[0][F_RHO], rho, N * sizeof(float), hipMemcpyDeviceToDevice, st));
    ZeroRanges z = {{sd.state[0][F_UTHETA] + Nt, sd.state[1][F_UTHETA] + Nt, reinterpret_cast<float*>(sd.status),
                     sd.state[0][F_RHO]},
                    {nP, nP, 1, rho ? 0 : (int)N}};
    hipLaunchKernelGGL(k_zero, dim3(rho ? 8 : 256), dim3(256), 0, st, z);
    KAM_HIP(hipGetLastError());
    const SimDesc* dSims = reinterpret_cast<const SimDesc*>(ds.base);
    const Tile* dTiles = reinterpret_cast<const Tile*>(ds.base + align256(sizeof(SimDesc)));
    int p = 0;
    for (int k = 0; k < nsteps; ++k) {
        int np = p;
        rc = launch_step(dSims, dTiles, ds.tiles, KAM_PHASE_ALL, p, st, &np);
        if (rc != KAM_OK) return rc;
        p = np;
    }
    KAM_HIP(hipMemcpyAsync(u_phi, sd.state[p][F_UPHI], N * sizeof(float), hipMemcpyDeviceToDevice, st));
    KAM_HIP(hipMemcpyAsync(u_theta, sd.state[p][F_UTHETA], Nt * sizeof(float), hipMemcpyDeviceToDevice, st));
    if (rho) KAM_HIP(hipMemcpyAsync(rho, sd.state[p][F_RHO], N * sizeof(float), hipMemcpyDeviceToDevice, st));
    return KAM_OK;
}
