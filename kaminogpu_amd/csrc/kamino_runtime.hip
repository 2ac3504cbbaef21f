// kamino_runtime.hip -- host runtime: shape plans, per-shape tables, device
// batches and the single-simulation C ABI (kam_sim_*).
#include "kamino_runtime.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

#include "kamino_launch.h"

namespace kam {

static thread_local std::string g_err;

int set_err(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// Reference constants, KaminoHeader.cuh:26-27.
static const double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// Plans
// ---------------------------------------------------------------------------
static int factor_m(int M, int* fac) {
    int n = 0;
    while (M % 4 == 0 && M > 4 && n < 12) {
        fac[n++] = 4;
        M /= 4;
    }
    const int primes[4] = {4, 2, 3, 5};
    for (int p : primes) {
        while (M % p == 0 && M > 1 && n < 12) {
            fac[n++] = p;
            M /= p;
        }
    }
    return M == 1 ? n : -1;
}

// LDS of an advection tile of R rows and W columns at staged row stride SW
// (k_advect): the staged input block, plus with the geometric update the
// advected u_phi and u_theta rows and the antipodal rows (the cubic roots
// reuse the staged block)
static size_t adv_lds(int R, int W, int SW, bool geometric) {
    // three staged planes; with the geometric update the advected u_phi*,
    // u_theta* and antipodal rows go over the third (rho) plane once the rho
    // nodes are done, spilling past it when larger; then three per-row
    // coefficients of rows j0-1 .. j0+R and three maxima per wavefront
    const size_t plane = (size_t)(R + 2 * kAdvHaloRows) * SW;
    const size_t uv = geometric ? (size_t)(R + 1) * (W + 2) + (R + 2) * (W + 1) + 2 * (W + 1) : 0;
    return (2 * plane + std::max(plane, uv) + 3 * (size_t)(R + 2) + 3 * (256 / 64)) * sizeof(float);
}
// Lanes per row in the row FFT (a quarter of M, 8..64, power of two); a
// 256-thread block holds 256 / group lane groups, each with a scratch row.
// Long rows (M >= 1024) take the whole block per row, its passes ordered by
// block barriers: a 1024 x 2048 sphere has only four rows per CU, so with
// one wavefront per row each SIMD would hold a single transform's chain.
static int fft_group_host(int M) {
    if (M >= 1024) return 256;
    int g = 8;
    while (g < 64 && 4 * g < M) g *= 2;
    return g;
}
// LDS of an inverse-transform tile: the nPhi = 2M twiddles (M in FP64), R + 1
// rows of M complex and the lane groups' scratch rows; FP32 complex (8 bytes),
// or FP64 (16 bytes) for the FP64 pressure
static size_t fft_lds(int R, int M, bool p64 = false) {
    return (size_t)((p64 ? 1 : 2) + R + 1 + 256 / fft_group_host(M)) * M * (p64 ? 16 : 8);
}
// LDS of a forward-transform tile: the nPhi = 2M twiddles (k_div_fft) or the
// first M of them rounded up to even (k_div_fft8, `half`), R rows of M complex
// and the scratch rows; with half twiddles 18 KiB at 4 x 256, eight
// workgroups per CU
static size_t div_lds(int R, int M, bool half) {
    const size_t tw = half ? (size_t)((M + 1) & ~1) : (size_t)2 * M;
    return (tw + (size_t)(R + 256 / fft_group_host(M)) * M) * 8;
}
static constexpr size_t kLdsMax = 160 * 1024;

// ---------------------------------------------------------------------------
// Tuning table: tile shapes per n_theta, set by the autotuner (or by hand).
// Tile shapes regroup work only; they never change a computed value.
// ---------------------------------------------------------------------------
struct TuneEntry {
    int advRows = 0, advCols = 0, fftRows = 0;
    int transform = -1;  // KAM_TRANSFORM_*, -1 = default (FFT)
};
static std::mutex g_tune_mu;
static std::map<int, TuneEntry> g_tune;

int tune_set(int nT, int advRows, int advCols, int fftRows) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    TuneEntry& e = g_tune[nT];
    e.advRows = advRows;
    e.advCols = advCols;
    e.fftRows = fftRows;
    return KAM_OK;
}

int tune_set_transform(int nT, int transform) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[nT].transform = transform;
    return KAM_OK;
}

static TuneEntry tune_get(int nT) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(nT);
    return it == g_tune.end() ? TuneEntry{} : it->second;
}

int make_plan(int nT, int nP, Plan& p) {
    if (nT < 4) return set_err(KAM_EINVAL, "n_theta must be >= 4 (got %d)", nT);
    if (nP != 2 * nT) return set_err(KAM_EINVAL, "n_phi must equal 2*n_theta (got %d, %d)", nP, nT);
    if (nT > 2048) return set_err(KAM_EINVAL, "n_theta must be <= 2048 (got %d)", nT);
    p = Plan();
    p.nT = nT;
    p.nP = nP;
    p.M = nP / 2;
    p.nfac = factor_m(p.M, p.fac);
    if (p.nfac < 0)
        return set_err(KAM_EINVAL, "n_theta must factor as 2^a 3^b 5^c (got %d)", nT);
    // advection tile: 16 rows x 64 columns (8 rows below nTheta = 128), about
    // 22 KiB of LDS, the shape the autotuner picks for most sizes of a mixed
    // batch; narrower spheres use their full width.  A batch stages every tile
    // at the widest member's row stride, so one common default width keeps
    // mixed batches compact.
    p.advCols = nP < 64 ? nP : 64;
    p.advRows = nT < 128 ? (nT < 8 ? nT : 8) : 16;
    // FFT tile: at most 8 rows (+1 halo row for the inverse, + scratch rows) in ~48 KiB
    int Rf = (int)(49152 / (8 * (size_t)p.M)) - 3 - 256 / fft_group_host(p.M);
    Rf = Rf < 1 ? 1 : (Rf > 8 ? 8 : Rf);
    // long rows (a block per row): one row per tile.  The inverse transform
    // then repeats its halo row, but twice as many tiles run at once
    // (measured, ms per step at one sphere: 1024 x 2048 0.0811 with two rows
    // -> 0.0798, 2048 x 4096 0.314 -> 0.272)
    if (fft_group_host(p.M) == 256) Rf = 1;
    p.fftRows = Rf > nT ? nT : Rf;
    // tridiagonal: L lanes per mode, at most K rows per lane
    if (nT <= 64) { p.triL = 8; p.triK = 8; }
    else if (nT <= 128) { p.triL = 16; p.triK = 8; }
    else if (nT <= 256) { p.triL = 32; p.triK = 8; }
    else if (nT <= 512) { p.triL = 64; p.triK = 8; }
    else if (nT <= 1024) { p.triL = 64; p.triK = 16; }
    else { p.triL = 64; p.triK = 32; }
    // tuned shapes, then manual overrides for tile sweeps
    const TuneEntry te = tune_get(nT);
    // the MFMA DFT holds nPhi x nPhi matrices: offered up to n_theta = 1024
    if (te.transform == KAM_TRANSFORM_DFT && nT <= 1024) p.transform = KAM_TRANSFORM_DFT;
    if (te.transform == KAM_TRANSFORM_FFT8) p.transform = KAM_TRANSFORM_FFT8;
    if (te.advCols >= 8 && te.advCols <= kAdvMaxCols) p.advCols = te.advCols < nP ? te.advCols : nP;
    if (te.advRows > 0 || te.fftRows > 0) adjust_plan(p, te.advRows, te.fftRows);
    if (const char* v = std::getenv("KAM_ADV_COLS")) {
        const int w = std::atoi(v);
        if (w >= 8 && w <= kAdvMaxCols) p.advCols = w < nP ? w : nP;
    }
    if (const char* v = std::getenv("KAM_ADV_ROWS")) adjust_plan(p, std::atoi(v), 0);
    if (const char* v = std::getenv("KAM_FFT_ROWS")) adjust_plan(p, 0, std::atoi(v));
    p.advLds = adv_lds(p.advRows, p.advCols, adv_stride(p.advCols), true);
    p.fftLds = fft_lds(p.fftRows, p.M);
    return KAM_OK;
}

int adjust_plan(Plan& p, int advRows, int fftRows) {
    if (advRows > 0) {
        const int R = advRows > p.nT ? p.nT : advRows;
        // the kernel's static LDS (row coefficients, reductions) takes under 1 KiB
        if (R > kAdvMaxRows || adv_lds(R, p.advCols, adv_stride(kAdvMaxCols), true) > 159 * 1024)
            return set_err(KAM_EINVAL, "advection tile too tall");
        p.advRows = R;
        p.advLds = adv_lds(p.advRows, p.advCols, adv_stride(p.advCols), true);
    }
    if (fftRows > 0) {
        if (fft_lds(fftRows, p.M) > 160 * 1024) return set_err(KAM_EINVAL, "FFT tile too tall");
        p.fftRows = fftRows > p.nT ? p.nT : fftRows;
        p.fftLds = fft_lds(p.fftRows, p.M);
    }
    return KAM_OK;
}

int check_precision(int precision, const Plan& plan) {
    if (precision != KAM_PREC_FP32 && precision != KAM_PREC_FP64_PRESSURE)
        return set_err(KAM_EINVAL, "unknown precision %d", precision);
    // the FP64 inverse transform's rows must fit in LDS, and the FP64
    // elimination is built for at most 16 rows per lane
    if (precision == KAM_PREC_FP64_PRESSURE && (plan.nT > 1024 || fft_lds(1, plan.M, true) > kLdsMax))
        return set_err(KAM_EINVAL, "the FP64 pressure needs n_theta <= 1024");
    return KAM_OK;
}

// three fields x two copies, the padded spectrum, one watchdog block
size_t sim_floats(int nT, int nP) { return (size_t)6 * nT * nP + (size_t)nP * spec_rows(nT) + 64; }

// ---------------------------------------------------------------------------
// Per-shape tables
// ---------------------------------------------------------------------------
struct TabEntry {
    Tables t;
    int refs = 0;
};
static std::mutex g_tab_mu;
static std::map<std::tuple<int, int, double, double>, TabEntry> g_tabs;

void compute_tables(int nT, double dt, double radius, float* rowf, double* rowd, float* tw, double* tw64) {
    const int nP = 2 * nT;
    const double h = kPi / nT;  // KaminoCore.cu:849
    for (int j = 0; j < nT; ++j) {
        const double tc = (j + 0.5) * h, tt = (j + 1.0) * h;
        const double sc = std::sin(tc), cc = std::cos(tc);
        rowf[RT_COFPHI_C * nT + j] = (float)(dt / (radius * sc * h));
        rowf[RT_COFPHI_T * nT + j] = (j < nT - 1) ? (float)(dt / (radius * std::sin(tt) * h)) : 0.f;
        rowf[RT_G * nT + j] = (float)(dt * cc / (radius * sc));
        rowf[RT_DIVS * nT + j] = (float)(std::sin(tc + 0.5 * h) / (h * sc));
        rowf[RT_DIVN * nT + j] = (float)(std::sin(tc - 0.5 * h) / (h * sc));
        rowf[RT_INVHS * nT + j] = (float)(1.0 / (h * sc));
        rowd[RD_COT2H * nT + j] = cc / (2.0 * h * sc);
        rowd[RD_INVSIN2 * nT + j] = 1.0 / (sc * sc);
        rowd[RD_INVHS * nT + j] = 1.0 / (h * sc);
        // the same FP64 terms as the FP64 elimination forms (1/h^2 is SimDesc::invH2)
        const double h2 = 1.0 / (h * h), ct = rowd[RD_COT2H * nT + j];
        rowf[RT_TRIA * nT + j] = (j == 0) ? 0.f : (float)(h2 - ct);
        rowf[RT_TRIC * nT + j] = (j == nT - 1) ? 0.f : (float)(h2 + ct);
        rowf[RT_TRIS * nT + j] = (float)rowd[RD_INVSIN2 * nT + j];
    }
    // exp(-2 pi i n / nPhi); the second half is the exact negation of the
    // first (w(n + M) = -w(n)), so kernels that stage only the first half
    // read the same values as those that stage all of it
    const int M = nP / 2;
    for (int n = 0; n < nP; ++n) {
        if (n < M) {
            const double a = 2.0 * kPi * n / nP;
            tw64[2 * n] = std::cos(a);
            tw64[2 * n + 1] = -std::sin(a);
        } else {
            tw64[2 * n] = -tw64[2 * (n - M)];
            tw64[2 * n + 1] = -tw64[2 * (n - M) + 1];
        }
        tw[2 * n] = (float)tw64[2 * n];
        tw[2 * n + 1] = (float)tw64[2 * n + 1];
    }
}

int acquire_tables(int device, int nT, double dt, double radius, Tables* out) {
    std::lock_guard<std::mutex> lk(g_tab_mu);
    auto key = std::make_tuple(device, nT, dt, radius);
    auto it = g_tabs.find(key);
    if (it != g_tabs.end()) {
        it->second.refs++;
        *out = it->second.t;
        return KAM_OK;
    }
    const int nP = 2 * nT;
    std::vector<float> rowf((size_t)RT_COUNT * nT);
    std::vector<double> rowd((size_t)RD_COUNT * nT);
    std::vector<float> tw((size_t)2 * nP);
    std::vector<double> tw64((size_t)2 * nP);
    compute_tables(nT, dt, radius, rowf.data(), rowd.data(), tw.data(), tw64.data());
    KAM_HIP(hipSetDevice(device));
    TabEntry e;
    KAM_HIP(hipMalloc(&e.t.rowf, rowf.size() * sizeof(float)));
    KAM_HIP(hipMalloc(&e.t.rowd, rowd.size() * sizeof(double)));
    KAM_HIP(hipMalloc(&e.t.twN, tw.size() * sizeof(float)));
    KAM_HIP(hipMalloc(&e.t.twN64, tw64.size() * sizeof(double)));
    KAM_HIP(hipMemcpy(e.t.rowf, rowf.data(), rowf.size() * sizeof(float), hipMemcpyHostToDevice));
    KAM_HIP(hipMemcpy(e.t.rowd, rowd.data(), rowd.size() * sizeof(double), hipMemcpyHostToDevice));
    KAM_HIP(hipMemcpy(e.t.twN, tw.data(), tw.size() * sizeof(float), hipMemcpyHostToDevice));
    KAM_HIP(hipMemcpy(e.t.twN64, tw64.data(), tw64.size() * sizeof(double), hipMemcpyHostToDevice));
    e.refs = 1;
    g_tabs[key] = e;
    *out = e.t;
    return KAM_OK;
}

void release_tables(int device, int nT, double dt, double radius) {
    std::lock_guard<std::mutex> lk(g_tab_mu);
    auto it = g_tabs.find(std::make_tuple(device, nT, dt, radius));
    if (it == g_tabs.end()) return;
    if (--it->second.refs > 0) return;
    (void)hipSetDevice(device);
    (void)hipFree(it->second.t.rowf);
    (void)hipFree(it->second.t.rowd);
    (void)hipFree(it->second.t.twN);
    (void)hipFree(it->second.t.twN64);
    g_tabs.erase(it);
}

// ---------------------------------------------------------------------------
// DFT matrices of the MFMA phi transform (kamino_dft.hip), per device and
// nPhi, built once and kept for the life of the process: FP16 hi and lo
// planes of 2^13 W, [N][K] with N = nPhi rounded up to 128 output columns and
// K = nPhi rounded up to 32 (padding zero).
//   forward  W[2n][i] = cos(2 pi n i / nPhi), W[2n+1][i] = -sin(...), n < M;
//            W[1][i] = (-1)^i (the Nyquist mode packed into slot 0)
//   inverse  W[i][2n] = 2 cos(2 pi n i / nPhi), W[i][2n+1] = -2 sin(...), 0 < n < M;
//            W[i][0] = 0 (mode 0 dropped), W[i][1] = (-1)^i
// ---------------------------------------------------------------------------
static uint16_t f16_bits(double x) {  // round to nearest even, normal and subnormal range
    const float f = (float)x;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    const uint32_t sign = (u >> 16) & 0x8000u;
    const double a = std::fabs(x);
    if (a == 0.0) return (uint16_t)sign;
    int e;
    (void)std::frexp(a, &e);  // a = m 2^e, m in [0.5, 1)
    const int ex = std::max(e - 1, -14);  // exponent of the binade (subnormals share -14)
    const double q = std::ldexp(a, 10 - ex);  // significand in units of the last place
    double r = std::nearbyint(q);            // ties to even under the default rounding mode
    int exOut = ex;
    if (r >= 2048.0) {  // rounded up into the next binade
        r /= 2.0;
        ++exOut;
    }
    if (exOut > 15) return (uint16_t)(sign | 0x7c00u);
    const uint32_t mant = (uint32_t)r;
    if (mant < 1024) return (uint16_t)(sign | mant);  // subnormal (exOut = -14)
    return (uint16_t)(sign | ((uint32_t)(exOut + 15) << 10) | (mant - 1024));
}
static double f16_value(uint16_t h) {
    const int e = (h >> 10) & 0x1f, m = h & 0x3ff;
    const double v = (e == 0) ? std::ldexp((double)m, -24) : std::ldexp(1024.0 + m, e - 25);
    return (h & 0x8000u) ? -v : v;
}

struct DftTables {
    uint16_t* fwd[2] = {nullptr, nullptr};
    uint16_t* inv[2] = {nullptr, nullptr};
    int K = 0;
};
static std::mutex g_dft_mu;
static std::map<std::pair<int, int>, DftTables> g_dft;

static int acquire_dft(int device, int nP, DftTables* out) {
    std::lock_guard<std::mutex> lk(g_dft_mu);
    auto it = g_dft.find({device, nP});
    if (it != g_dft.end()) {
        *out = it->second;
        return KAM_OK;
    }
    const int M = nP / 2;
    const int N = (nP + kDftTileCols - 1) / kDftTileCols * kDftTileCols;
    const int K = (nP + kDftChunk - 1) / kDftChunk * kDftChunk;
    std::vector<uint16_t> fh((size_t)N * K, 0), fl((size_t)N * K, 0), ih((size_t)N * K, 0), il((size_t)N * K, 0);
    const double scale = 8192.0;
    auto put = [&](std::vector<uint16_t>& hi, std::vector<uint16_t>& lo, int c, int k, double w) {
        const double x = w * scale;
        const uint16_t h = f16_bits(x);
        hi[(size_t)c * K + k] = h;
        lo[(size_t)c * K + k] = f16_bits(x - f16_value(h));
    };
    for (int i = 0; i < nP; ++i) {
        const double alt = (i & 1) ? -1.0 : 1.0;
        for (int n = 0; n < M; ++n) {
            // the angle 2 pi n i / nPhi reduced exactly: n i mod nPhi
            const double a = 2.0 * kPi * (double)(((long long)n * i) % nP) / nP;
            const double cs = std::cos(a), sn = std::sin(a);
            put(fh, fl, 2 * n, i, cs);
            put(fh, fl, 2 * n + 1, i, n == 0 ? alt : -sn);
            put(ih, il, i, 2 * n, n == 0 ? 0.0 : 2.0 * cs);
            put(ih, il, i, 2 * n + 1, n == 0 ? alt : -2.0 * sn);
        }
    }
    KAM_HIP(hipSetDevice(device));
    DftTables t;
    t.K = K;
    const size_t bytes = (size_t)N * K * sizeof(uint16_t);
    std::vector<uint16_t>* src[4] = {&fh, &fl, &ih, &il};
    uint16_t** dst[4] = {&t.fwd[0], &t.fwd[1], &t.inv[0], &t.inv[1]};
    for (int q = 0; q < 4; ++q) {
        if (hipMalloc(dst[q], bytes) != hipSuccess) return set_err(KAM_ENOMEM, "DFT matrix allocation failed");
        KAM_HIP(hipMemcpy(*dst[q], src[q]->data(), bytes, hipMemcpyHostToDevice));
    }
    g_dft[{device, nP}] = t;
    *out = t;
    return KAM_OK;
}

int attach_transform(SimDesc& d, const Plan& p, int precision, int device) {
    d.flags &= ~SIM_DFT;
    d.dftFwd[0] = d.dftFwd[1] = d.dftInv[0] = d.dftInv[1] = nullptr;
    d.dftK = 0;
    if (p.transform != KAM_TRANSFORM_DFT || precision != KAM_PREC_FP32) return KAM_OK;
    DftTables t;
    int rc = acquire_dft(device, p.nP, &t);
    if (rc != KAM_OK) return rc;
    d.flags |= SIM_DFT;
    d.dftFwd[0] = t.fwd[0];
    d.dftFwd[1] = t.fwd[1];
    d.dftInv[0] = t.inv[0];
    d.dftInv[1] = t.inv[1];
    d.dftK = t.K;
    return KAM_OK;
}

void fill_desc(SimDesc& d, const Plan& p, const Tables& t, const kam_sim_desc& sd, float* mem) {
    std::memset(&d, 0, sizeof d);
    const double h = kPi / p.nT;
    d.nTheta = p.nT;
    d.nPhi = p.nP;
    d.M = p.M;
    d.flags = (sd.geometric ? SIM_GEOMETRIC : 0) | (sd.precision == KAM_PREC_FP64_PRESSURE ? SIM_P64 : 0);
    d.cofTheta = (float)(sd.dt / (sd.radius * h));
    d.invH = (float)(1.0 / h);
    d.invH2 = 1.0 / (h * h);
    d.invH64 = 1.0 / h;
    const size_t N = (size_t)p.nT * p.nP;
    for (int b = 0; b < 2; ++b)
        for (int f = 0; f < 3; ++f) d.state[b][f] = mem + (size_t)(3 * b + f) * N;
    d.spec = mem + 6 * N;
    d.status = reinterpret_cast<int*>(mem + 6 * N + (size_t)p.nP * spec_rows(p.nT));  // callers may point it elsewhere
    d.rowf = t.rowf;
    d.rowd = t.rowd;
    d.twN = t.twN;
    d.twN64 = t.twN64;
    d.nfac = p.nfac;
    int Ls = 1;
    for (int k = 0; k < p.nfac; ++k) {
        d.fac[k] = p.fac[k];
        d.twStep[k] = p.nP / (Ls * p.fac[k]);
        d.lsMagic[k] = (Ls & (Ls - 1)) == 0 ? 0u : (unsigned)((((uint64_t)1 << 32) + Ls - 1) / Ls);
        Ls *= p.fac[k];
    }
    d.fftShift = __builtin_ctz((unsigned)fft_group_host(p.M));
    d.advRows = p.advRows;
    d.fftRows = p.fftRows;
    d.magicP = (unsigned)((((uint64_t)1 << 32) + (uint64_t)p.nP - 1) / (uint64_t)p.nP);
    d.magicP2 = (unsigned)((((uint64_t)1 << 32) + (uint64_t)p.M - 1) / (uint64_t)p.M);
    d.invNP = 1.0f / (float)p.nP;
    set_tile_cols(d, p.advCols);
    d.radius = (float)sd.radius;
    d.dt = (float)sd.dt;
}

void set_tile_cols(SimDesc& d, int W) {
    d.advCols = W;
    const int SC = W + 2 * kAdvHaloCols;
    const uint64_t div[5] = {(uint64_t)W, (uint64_t)W + 1, (uint64_t)W + 2, (uint64_t)SC, (uint64_t)SC / 4};
    for (int k = 0; k < 5; ++k) d.magicW[k] = (unsigned)((((uint64_t)1 << 32) + div[k] - 1) / div[k]);
}

// ---------------------------------------------------------------------------
// Device batches
// ---------------------------------------------------------------------------
DeviceBatch::~DeviceBatch() {
    for (int b = 0; b < 2; ++b)
        if (uploaded_[b]) (void)hipEventSynchronize(uploaded_[b]);
    if (dSims_) (void)hipFree(dSims_);
    if (dTiles_) (void)hipFree(dTiles_);
    for (int b = 0; b < 2; ++b) {
        if (hStage_[b]) (void)hipHostFree(hStage_[b]);
        if (uploaded_[b]) (void)hipEventDestroy(uploaded_[b]);
    }
}

// Grows the device tables geometrically; a regrowth frees the old tables
// after the stream has finished with them.
int DeviceBatch::reserve(size_t ns, size_t nt, hipStream_t st) {
    if (ns > capSims_ || nt > capTiles_) KAM_HIP(hipStreamSynchronize(st));
    if (ns > capSims_) {
        if (dSims_) KAM_HIP(hipFree(dSims_));
        dSims_ = nullptr;
        const size_t c = std::max<size_t>(std::max<size_t>(16, ns), 2 * capSims_);
        KAM_HIP(hipMalloc(&dSims_, c * sizeof(SimDesc)));
        capSims_ = c;
    }
    if (nt > capTiles_) {
        if (dTiles_) KAM_HIP(hipFree(dTiles_));
        dTiles_ = nullptr;
        const size_t c = std::max<size_t>(std::max<size_t>(256, nt), 2 * capTiles_);
        KAM_HIP(hipMalloc(&dTiles_, c * sizeof(Tile)));
        capTiles_ = c;
    }
    const size_t need = capSims_ * sizeof(SimDesc) + capTiles_ * sizeof(Tile);
    const int b = cur_;
    if (need > stageBytes_[b]) {
        if (uploaded_[b]) KAM_HIP(hipEventSynchronize(uploaded_[b]));
        if (hStage_[b]) KAM_HIP(hipHostFree(hStage_[b]));
        hStage_[b] = nullptr;
        KAM_HIP(hipHostMalloc(&hStage_[b], need, hipHostMallocDefault));
        stageBytes_[b] = need;
    }
    if (!uploaded_[b]) KAM_HIP(hipEventCreateWithFlags(&uploaded_[b], hipEventDisableTiming));
    return KAM_OK;
}

// Tile heights: each simulation's planned height, halved for every
// simulation while a kernel would get fewer than two workgroups per CU.  The
// height changes only how rows are grouped, never a value.
static int shrink_until_busy(const Plan* plans, int n, int* rows, int minTiles, const int* advCols) {
    for (;;) {
        long tiles = 0;
        bool canShrink = false;
        for (int s = 0; s < n; ++s) {
            const long cols = advCols ? (plans[s].nP + advCols[s] - 1) / advCols[s] : 1;
            tiles += cols * ((plans[s].nT + rows[s] - 1) / rows[s]);
            canShrink |= rows[s] > 1;
        }
        if (tiles >= minTiles || !canShrink) return (int)tiles;
        for (int s = 0; s < n; ++s) rows[s] = rows[s] > 1 ? (rows[s] + 1) / 2 : 1;
    }
}

int plan_tiles(SimDesc* sims, const Plan* plans, int n, int* rowScratch, Tile* out, int cap,
               TileLayout& L) {
    L = TileLayout();
    L.nsims = n;
    int cnt = 0;
    auto put = [&](const Tile& t) {
        if (cnt < cap) out[cnt] = t;
        ++cnt;
    };
    const int minTiles = 2 * device_cus();
    int* advRows = rowScratch;
    int* advCols = rowScratch + n;
    int* fftRows = rowScratch + 2 * n;
    for (int s = 0; s < n; ++s) {
        advRows[s] = plans[s].advRows;
        advCols[s] = plans[s].advCols;
        fftRows[s] = plans[s].fftRows;
    }
    // One advection launch stages every tile at the widest member's row
    // stride and with the largest member's LDS.  When members were tuned to
    // different widths, the wide ones (alone a launch too small to fill the
    // GPU) are restaged as narrow tiles of the same area, so that the narrow
    // majority keeps its occupancy.  Tile shapes never change a value.
    bool narrow = false, wide = false;
    for (int s = 0; s < n; ++s) (advCols[s] > kAdvMaxCols / 2 ? wide : narrow) = true;
    if (narrow && wide)
        for (int s = 0; s < n; ++s)
            if (advCols[s] > kAdvMaxCols / 2) {
                advRows[s] = std::min(2 * advRows[s], kAdvMaxRows);
                advCols[s] = kAdvMaxCols / 2;
                set_tile_cols(sims[s], advCols[s]);
            }
    L.advStride = adv_stride(wide && !narrow ? kAdvMaxCols : kAdvMaxCols / 2);
    shrink_until_busy(plans, n, advRows, minTiles, advCols);
    shrink_until_busy(plans, n, fftRows, minTiles, nullptr);
    // advection tiles, in batch order (tallest simulations first); the last
    // column tile of a row band overlaps its neighbour instead of running short.
    // A tile starting within kAdvHaloRows + 1 rows of a pole (or ending within
    // them) samples with the checked sampler throughout, so the two polar bands
    // are exactly that tall when the tile height allows, and every band between
    // them can take the fast one; the last inner band may be shorter.
    L.advOff = 0;
    constexpr int kPoleRows = kAdvHaloRows + 1;
    for (int s = 0; s < n; ++s) {
        const Plan& p = plans[s];
        const int W = advCols[s], R = advRows[s];
        const bool polar = R >= kPoleRows && p.nT > 2 * kPoleRows;
        for (int j = 0, rows; j < p.nT; j += rows) {
            if (!polar)
                rows = std::min(R, p.nT - j);
            else if (j == 0)
                rows = kPoleRows;
            else if (j >= p.nT - kPoleRows)
                rows = p.nT - j;
            else
                rows = std::min(R, p.nT - kPoleRows - j);
            for (int c = 0; c < p.nP; c += W) put(Tile{s, j, rows, std::min(c, p.nP - W)});
        }
        L.advLds = std::max(L.advLds, adv_lds(advRows[s], W, L.advStride, (sims[s].flags & SIM_GEOMETRIC) != 0));
    }
    L.nAdv = cnt;
    // FFT tiles: simulations with the FP32 pressure first, then those with the
    // FP64 pressure, whose inverse transform is a launch of its own (FP64 rows
    // in LDS: their tiles hold as many rows as fit)
    L.fftOff = cnt;
    // the eight-wave forward kernel when every FFT simulation asks for it
    bool anyFft = false, allW8 = true;
    size_t divFull = 0, divHalf = 0;
    for (int pass = 0; pass < 2; ++pass)
        for (int s = 0; s < n; ++s) {
            const Plan& p = plans[s];
            const bool p64 = (sims[s].flags & SIM_P64) != 0;
            if (p64 != (pass == 1) || (sims[s].flags & SIM_DFT)) continue;
            int R = fftRows[s];
            if (p64)
                while (R > 1 && fft_lds(R, p.M, true) > kLdsMax) --R;
            for (int j = 0; j < p.nT; j += R) {
                put(Tile{s, j, std::min(R, p.nT - j), 0});
                L.nFft64 += p64 ? 1 : 0;
            }
            L.fftLds = std::max(L.fftLds, fft_lds(R, p.M));
            divFull = std::max(divFull, div_lds(R, p.M, false));
            divHalf = std::max(divHalf, div_lds(R, p.M, true));
            anyFft = true;
            allW8 = allW8 && p.transform == KAM_TRANSFORM_FFT8;
            if (p64) L.fftLds64 = std::max(L.fftLds64, fft_lds(R, p.M, true));
        }
    L.divW8 = anyFft && allW8;
    L.divLds = L.divW8 ? divHalf : divFull;
    L.nFft = cnt - L.fftOff;
    L.nFft32 = L.nFft - L.nFft64;
    // SIM_DFT simulations: row runs for the divergence and gradient kernels,
    // then 64-row x 128-column blocks of the DFT products
    L.dftRowOff = cnt;
    for (int s = 0; s < n; ++s)
        if (sims[s].flags & SIM_DFT)
            for (int j = 0; j < plans[s].nT; j += fftRows[s]) put(Tile{s, j, std::min(fftRows[s], plans[s].nT - j), 0});
    L.nDftRow = cnt - L.dftRowOff;
    L.dftOff = cnt;
    for (int s = 0; s < n; ++s)
        if (sims[s].flags & SIM_DFT)
            for (int j = 0; j < plans[s].nT; j += kDftTileRows)
                for (int c = 0; c < plans[s].nP; c += kDftTileCols)
                    put(Tile{s, j, std::min(kDftTileRows, plans[s].nT - j), c});
    L.nDft = cnt - L.dftOff;
    // tridiagonal tiles, grouped by kernel class; when several classes of
    // K = 8 rows per lane are present they share one mixed launch (L = 0),
    // each tile carrying its lanes per mode in col0
    const int classes[6][2] = {{8, 8}, {16, 8}, {32, 8}, {64, 8}, {64, 16}, {64, 32}};
    int k8 = 0;
    for (auto& c : classes)
        if (c[1] == 8)
            for (int s = 0; s < n; ++s)
                if (plans[s].triL == c[0] && plans[s].triK == 8) {
                    ++k8;
                    break;
                }
    const bool mixed = k8 > 1;
    // FP32-pressure simulations run the elimination in FP32 (the low-mode
    // tiles refining once in FP64); FP64-pressure ones in FP64
    for (int fp32 = 0; fp32 < 2; ++fp32)
        for (auto& c : classes) {
            const bool inMixed = mixed && c[1] == 8;
            const int off = cnt;
            for (int s = 0; s < n; ++s) {
                const Plan& p = plans[s];
                if (p.triL != c[0] || p.triK != c[1]) continue;
                const bool p64 = (sims[s].flags & SIM_P64) != 0;
                const int per = 256 / p.triL;
                for (int m = 0; m < p.M; m += per) {
                    if (p64 == (fp32 == 1)) continue;
                    put(Tile{s, m, std::min(per, p.M - m), inMixed ? p.triL : 0});
                }
            }
            const int num = cnt - off;
            if (num == 0) continue;
            if (inMixed && L.ntri > 0 && L.tri[L.ntri - 1].L == 0 && L.tri[L.ntri - 1].fp32 == (fp32 == 1))
                L.tri[L.ntri - 1].count += num;  // classes are contiguous in the tile list
            else
                L.tri[L.ntri++] = TriClass{inMixed ? 0 : c[0], c[1], off, num, fp32 == 1};
        }
    // particle tiles: runs of kPartTile particles
    constexpr int kPartTile = 2048;
    L.partOff = cnt;
    for (int s = 0; s < n; ++s)
        for (int k = 0; k < sims[s].nPart; k += kPartTile) put(Tile{s, k, std::min(kPartTile, sims[s].nPart - k), 0});
    L.nPart = cnt - L.partOff;
    L.total = cnt;
    return cnt;
}

static const int kGridCap = 1 << 20;

int launch_tri(const SimDesc* dSims, const Tile* dTiles, const TileLayout& L, int parity,
               hipStream_t st, bool resid) {
    // one launch per class (all K = 8 classes of a mixed batch form one, L = 0),
    // in stream order (side streams would let small classes overlap, but their
    // fork/join graphs cost far more to capture each time a serving engine's
    // membership changes)
    for (int k = 0; k < L.ntri; ++k) {
        const TriClass& c = L.tri[k];
        KAM_HIP(launch_tridiag(c.L, c.K, dSims, dTiles + c.off, c.count, parity, kGridCap, st, resid, c.fp32));
    }
    return KAM_OK;
}

int launch_step(const SimDesc* dSims, const Tile* dTiles, const TileLayout& L, int phases, int parity,
                hipStream_t st, int* newParity, hipEvent_t* ev) {
    int p = parity;
    if (L.nsims == 0) {
        *newParity = p;
        return KAM_OK;
    }
    if (phases & KAM_PHASE_ADVECT) {
        if (ev) KAM_HIP(hipEventRecord(ev[2 * KC_ADVECT], st));
        KAM_HIP(launch_advect(dSims, dTiles + L.advOff, L.nAdv, p, L.advStride, L.advLds, kGridCap, st));
        KAM_HIP(launch_particles(dSims, dTiles + L.partOff, L.nPart, p, kGridCap, st));
        if (ev) KAM_HIP(hipEventRecord(ev[2 * KC_ADVECT + 1], st));
        p ^= 1;
    }
    if (phases & KAM_PHASE_DIVFFT) {
        if (ev) KAM_HIP(hipEventRecord(ev[2 * KC_DIVFFT], st));
        KAM_HIP(launch_div_fft(dSims, dTiles + L.fftOff, L.nFft, p, L.divLds, kGridCap, st, L.divW8));
        KAM_HIP(launch_dft_forward(dSims, dTiles + L.dftRowOff, L.nDftRow, dTiles + L.dftOff, L.nDft, p, st));
        if (ev) KAM_HIP(hipEventRecord(ev[2 * KC_DIVFFT + 1], st));
    }
    if (phases & KAM_PHASE_TRI) {
        if (ev) KAM_HIP(hipEventRecord(ev[2 * KC_TRI], st));
        int rc = launch_tri(dSims, dTiles, L, p, st, false);
        if (rc != KAM_OK) return rc;
        if (ev) KAM_HIP(hipEventRecord(ev[2 * KC_TRI + 1], st));
    }
    if (phases & KAM_PHASE_IFFT) {
        if (ev) KAM_HIP(hipEventRecord(ev[2 * KC_IFFT], st));
        KAM_HIP(launch_ifft_grad(dSims, dTiles + L.fftOff, L.nFft32, p, L.fftLds, kGridCap, st, false));
        KAM_HIP(launch_ifft_grad(dSims, dTiles + L.fftOff + L.nFft32, L.nFft64, p, L.fftLds64, kGridCap, st,
                                 true));
        KAM_HIP(launch_dft_inverse(dSims, dTiles + L.dftRowOff, L.nDftRow, dTiles + L.dftOff, L.nDft, p, st));
        if (ev) KAM_HIP(hipEventRecord(ev[2 * KC_IFFT + 1], st));
    }
    *newParity = p;
    return KAM_OK;
}

int DeviceBatch::build(const std::vector<SimDesc>& simsIn, const std::vector<Plan>& plans,
                       hipStream_t st) {
    const int n = (int)plans.size();
    std::vector<SimDesc> sims(simsIn);
    std::vector<int> rows(3 * (size_t)n + 1);
    const int need = plan_tiles(sims.data(), plans.data(), n, rows.data(), nullptr, 0, lay_);
    hTiles_.resize((size_t)need);
    plan_tiles(sims.data(), plans.data(), n, rows.data(), hTiles_.data(), need, lay_);
    cur_ ^= 1;
    int rc = reserve(sims.size(), hTiles_.size(), st);
    if (rc != KAM_OK) return rc;
    // this staging buffer is free once its last upload, two rebuilds back,
    // has landed; the copies are queued behind the steps already launched
    KAM_HIP(hipEventSynchronize(uploaded_[cur_]));
    char* stage = static_cast<char*>(hStage_[cur_]);
    const size_t sb = sims.size() * sizeof(SimDesc), tb = hTiles_.size() * sizeof(Tile);
    std::memcpy(stage, sims.data(), sb);
    std::memcpy(stage + capSims_ * sizeof(SimDesc), hTiles_.data(), tb);
    if (sb) KAM_HIP(hipMemcpyAsync(dSims_, stage, sb, hipMemcpyHostToDevice, st));
    if (tb)
        KAM_HIP(hipMemcpyAsync(dTiles_, stage + capSims_ * sizeof(SimDesc), tb,
                               hipMemcpyHostToDevice, st));
    KAM_HIP(hipEventRecord(uploaded_[cur_], st));
    return KAM_OK;
}

int DeviceBatch::launch_tri_residual(hipStream_t st, int parity) {
    return launch_tri(dSims_, dTiles_, lay_, parity, st, true);
}

int DeviceBatch::launch_pressure(hipStream_t st, int parity, float* pout) {
    const TileLayout& L = lay_;
    KAM_HIP(launch_ifft_grad(dSims_, dTiles_ + L.fftOff, L.nFft32, parity, L.fftLds, kGridCap, st, false, pout));
    KAM_HIP(launch_ifft_grad(dSims_, dTiles_ + L.fftOff + L.nFft32, L.nFft64, parity, L.fftLds64, kGridCap, st,
                             true, pout));
    KAM_HIP(launch_dft_pressure(dSims_, dTiles_ + L.dftOff, L.nDft, parity, st));
    return KAM_OK;
}

int DeviceBatch::launch(int phases, int parity, hipStream_t st, int* newParity, hipEvent_t* ev) {
    return launch_step(dSims_, dTiles_, lay_, phases, parity, st, newParity, ev);
}

// ---------------------------------------------------------------------------
// Devices
// ---------------------------------------------------------------------------
static std::mutex g_dev_mu;
static std::map<int, bool> g_dev_ready;
static int g_cus = 256;

int device_cus() { return g_cus; }

int ensure_device(int device) {
    std::lock_guard<std::mutex> lk(g_dev_mu);
    auto it = g_dev_ready.find(device);
    if (it != g_dev_ready.end()) {
        KAM_HIP(hipSetDevice(device));
        return KAM_OK;
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return set_err(KAM_ENODEV, "no HIP device visible");
    if (device < 0 || device >= n) return set_err(KAM_ENODEV, "device %d out of range (%d)", device, n);
    hipDeviceProp_t prop;
    KAM_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_err(KAM_ENODEV, "device %d is %s; this library is built for gfx950", device,
                       prop.gcnArchName);
    KAM_HIP(hipSetDevice(device));
    KAM_HIP(configure_kernels(160 * 1024));
    g_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    g_dev_ready[device] = true;
    return KAM_OK;
}

}  // namespace kam

// ===========================================================================
// C ABI: single simulation
// ===========================================================================
using namespace kam;

struct kam_sim {
    kam_sim_desc sd;
    Plan plan;
    Tables tables;
    SimDesc desc;
    float* mem = nullptr;
    hipStream_t stream = nullptr;
    DeviceBatch batch;
    int parity = 0;
    float* pmem = nullptr;  // particles: 2 copies of nPart float2
    double* resid = nullptr;  // per-mode residual report, on first use
    float* pbuf = nullptr;    // pressure readback, on first use
};

extern "C" {

const char* kam_last_error(void) { return g_err.c_str(); }

const char* kam_version(void) { return "kamino-mi355x 0.1 (gfx950)"; }

int kam_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    int c = 0;
    for (int d = 0; d < n; ++d) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d) == hipSuccess &&
            std::strncmp(prop.gcnArchName, "gfx950", 6) == 0)
            ++c;
    }
    return c;
}

size_t kam_sim_bytes(int n_theta, int n_phi) {
    return sim_floats(n_theta, n_phi) * sizeof(float);
}

static int check_desc(const kam_sim_desc* d, Plan& plan) {
    if (!d) return set_err(KAM_EINVAL, "null descriptor");
    if (!(d->dt > 0.0) || !std::isfinite(d->dt)) return set_err(KAM_EINVAL, "dt must be > 0");
    if (!(d->radius > 0.0) || !std::isfinite(d->radius)) return set_err(KAM_EINVAL, "radius must be > 0");
    int rc = make_plan(d->n_theta, d->n_phi, plan);
    if (rc != KAM_OK) return rc;
    return check_precision(d->precision, plan);
}

int kam_sim_create(const kam_sim_desc* d, kam_sim** out) {
    if (!out) return set_err(KAM_EINVAL, "null output handle");
    *out = nullptr;
    Plan plan;
    int rc = check_desc(d, plan);
    if (rc != KAM_OK) return rc;
    rc = ensure_device(d->device);
    if (rc != KAM_OK) return rc;
    kam_sim* s = new kam_sim();
    s->sd = *d;
    s->plan = plan;
    rc = acquire_tables(d->device, d->n_theta, d->dt, d->radius, &s->tables);
    if (rc != KAM_OK) {
        delete s;
        return rc;
    }
    const size_t bytes = kam_sim_bytes(d->n_theta, d->n_phi);
    if (hipMalloc(&s->mem, bytes) != hipSuccess) {
        release_tables(d->device, d->n_theta, d->dt, d->radius);
        delete s;
        return set_err(KAM_ENOMEM, "hipMalloc of %zu bytes failed", bytes);
    }
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMemsetAsync(s->mem, 0, bytes, s->stream) != hipSuccess) {
        kam_sim_destroy(s);
        return set_err(KAM_EHIP, "stream creation / clear failed");
    }
    fill_desc(s->desc, plan, s->tables, *d, s->mem);
    rc = attach_transform(s->desc, plan, d->precision, d->device);
    if (rc != KAM_OK) {
        kam_sim_destroy(s);
        return rc;
    }
    std::vector<SimDesc> sims{s->desc};
    std::vector<Plan> plans{plan};
    rc = s->batch.build(sims, plans, s->stream);
    if (rc != KAM_OK) {
        kam_sim_destroy(s);
        return rc;
    }
    *out = s;
    return KAM_OK;
}

int kam_sim_destroy(kam_sim* s) {
    if (!s) return KAM_OK;
    (void)hipSetDevice(s->sd.device);
    if (s->stream) {
        (void)hipStreamSynchronize(s->stream);
        (void)hipStreamDestroy(s->stream);
    }
    if (s->mem) (void)hipFree(s->mem);
    if (s->pmem) (void)hipFree(s->pmem);
    if (s->resid) (void)hipFree(s->resid);
    if (s->pbuf) (void)hipFree(s->pbuf);
    if (s->tables.rowf) release_tables(s->sd.device, s->sd.n_theta, s->sd.dt, s->sd.radius);
    delete s;
    return KAM_OK;
}

int kam_sim_set_fields(kam_sim* s, const float* u_phi, const float* u_theta, const float* rho) {
    if (!s || !u_phi || !u_theta) return set_err(KAM_EINVAL, "null argument");
    KAM_HIP(hipSetDevice(s->sd.device));
    const size_t N = (size_t)s->plan.nT * s->plan.nP;
    const size_t Nt = N - s->plan.nP;
    KAM_HIP(hipMemcpyAsync(s->desc.state[s->parity][F_UPHI], u_phi, N * sizeof(float),
                           hipMemcpyHostToDevice, s->stream));
    KAM_HIP(hipMemcpyAsync(s->desc.state[s->parity][F_UTHETA], u_theta, Nt * sizeof(float),
                           hipMemcpyHostToDevice, s->stream));
    if (rho)
        KAM_HIP(hipMemcpyAsync(s->desc.state[s->parity][F_RHO], rho, N * sizeof(float),
                               hipMemcpyHostToDevice, s->stream));
    KAM_HIP(hipStreamSynchronize(s->stream));
    return KAM_OK;
}

int kam_sim_get_fields(kam_sim* s, float* u_phi, float* u_theta, float* rho) {
    if (!s) return set_err(KAM_EINVAL, "null simulation");
    KAM_HIP(hipSetDevice(s->sd.device));
    const size_t N = (size_t)s->plan.nT * s->plan.nP;
    const size_t Nt = N - s->plan.nP;
    if (u_phi)
        KAM_HIP(hipMemcpyAsync(u_phi, s->desc.state[s->parity][F_UPHI], N * sizeof(float),
                               hipMemcpyDeviceToHost, s->stream));
    if (u_theta)
        KAM_HIP(hipMemcpyAsync(u_theta, s->desc.state[s->parity][F_UTHETA], Nt * sizeof(float),
                               hipMemcpyDeviceToHost, s->stream));
    if (rho)
        KAM_HIP(hipMemcpyAsync(rho, s->desc.state[s->parity][F_RHO], N * sizeof(float),
                               hipMemcpyDeviceToHost, s->stream));
    KAM_HIP(hipStreamSynchronize(s->stream));
    return KAM_OK;
}

int kam_sim_set_particles(kam_sim* s, const float* coords, int64_t n) {
    if (!s || n < 0 || (n > 0 && !coords)) return set_err(KAM_EINVAL, "bad particle arguments");
    if (n > (int64_t)1 << 30) return set_err(KAM_EINVAL, "too many particles (%lld)", (long long)n);
    KAM_HIP(hipSetDevice(s->sd.device));
    KAM_HIP(hipStreamSynchronize(s->stream));
    if (n != s->desc.nPart) {
        if (s->pmem) KAM_HIP(hipFree(s->pmem));
        s->pmem = nullptr;
        s->desc.part[0] = s->desc.part[1] = nullptr;
        s->desc.nPart = 0;
        if (n > 0) {
            if (hipMalloc(&s->pmem, (size_t)n * 4 * sizeof(float)) != hipSuccess)
                return set_err(KAM_ENOMEM, "hipMalloc of %lld particles failed", (long long)n);
            s->desc.part[0] = s->pmem;
            s->desc.part[1] = s->pmem + (size_t)n * 2;
            s->desc.nPart = (int)n;
        }
    }
    if (n > 0)
        KAM_HIP(hipMemcpyAsync(s->desc.part[s->parity], coords, (size_t)n * 2 * sizeof(float),
                               hipMemcpyHostToDevice, s->stream));
    std::vector<SimDesc> sims{s->desc};
    std::vector<Plan> plans{s->plan};
    int rc = s->batch.build(sims, plans, s->stream);
    if (rc != KAM_OK) return rc;
    KAM_HIP(hipStreamSynchronize(s->stream));
    return KAM_OK;
}

int kam_sim_get_particles(kam_sim* s, float* coords, int64_t n) {
    if (!s || (n > 0 && !coords)) return set_err(KAM_EINVAL, "bad particle arguments");
    if (n != s->desc.nPart)
        return set_err(KAM_EINVAL, "simulation holds %d particles, not %lld", s->desc.nPart, (long long)n);
    KAM_HIP(hipSetDevice(s->sd.device));
    if (n > 0)
        KAM_HIP(hipMemcpyAsync(coords, s->desc.part[s->parity], (size_t)n * 2 * sizeof(float),
                               hipMemcpyDeviceToHost, s->stream));
    KAM_HIP(hipStreamSynchronize(s->stream));
    return KAM_OK;
}

int64_t kam_sim_particle_count(const kam_sim* s) { return s ? s->desc.nPart : 0; }

int kam_sim_poisson_residual(kam_sim* s, double* per_mode, double* max_rel, double* rms_rel) {
    if (!s) return set_err(KAM_EINVAL, "null simulation");
    KAM_HIP(hipSetDevice(s->sd.device));
    const int M = s->plan.M;
    if (!s->resid) {
        KAM_HIP(hipStreamSynchronize(s->stream));
        if (hipMalloc(&s->resid, (size_t)M * sizeof(double)) != hipSuccess)
            return set_err(KAM_ENOMEM, "residual buffer allocation failed");
        s->desc.resid = s->resid;
        std::vector<SimDesc> sims{s->desc};
        std::vector<Plan> plans{s->plan};
        int rc = s->batch.build(sims, plans, s->stream);
        if (rc != KAM_OK) return rc;
    }
    int rc = s->batch.launch_tri_residual(s->stream, s->parity);
    if (rc != KAM_OK) return rc;
    std::vector<double> h((size_t)M);
    KAM_HIP(hipMemcpyAsync(h.data(), s->resid, (size_t)M * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    KAM_HIP(hipStreamSynchronize(s->stream));
    double mx = 0.0, sq = 0.0;
    for (double v : h) {
        mx = v > mx ? v : mx;
        sq += v * v;
    }
    if (per_mode) std::memcpy(per_mode, h.data(), (size_t)M * sizeof(double));
    if (max_rel) *max_rel = mx;
    if (rms_rel) *rms_rel = std::sqrt(sq / M);
    return KAM_OK;
}

int kam_sim_get_pressure(kam_sim* s, float* p) {
    if (!s || !p) return set_err(KAM_EINVAL, "null argument");
    KAM_HIP(hipSetDevice(s->sd.device));
    const size_t N = (size_t)s->plan.nT * s->plan.nP;
    if (!s->pbuf) {
        KAM_HIP(hipStreamSynchronize(s->stream));
        if (hipMalloc(&s->pbuf, N * sizeof(float)) != hipSuccess)
            return set_err(KAM_ENOMEM, "pressure buffer allocation failed");
    }
    int rc = s->batch.launch_pressure(s->stream, s->parity, s->pbuf);
    if (rc != KAM_OK) return rc;
    // a DFT simulation leaves its pressure rows over the finished state copy
    const float* src = (s->desc.flags & SIM_DFT) ? s->desc.state[s->parity ^ 1][F_UTHETA] : s->pbuf;
    KAM_HIP(hipMemcpyAsync(p, src, N * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    KAM_HIP(hipStreamSynchronize(s->stream));
    return KAM_OK;
}

int kam_sim_get_spectrum(kam_sim* s, float* spec) {
    if (!s || !spec) return set_err(KAM_EINVAL, "null argument");
    KAM_HIP(hipSetDevice(s->sd.device));
    const int M = s->plan.M, nT = s->plan.nT;
    const size_t n = (size_t)M * spec_rows(nT) * 2;
    std::vector<float> blocked(n);
    KAM_HIP(hipMemcpyAsync(blocked.data(), s->desc.spec, n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    KAM_HIP(hipStreamSynchronize(s->stream));
    // the device keeps the spectrum in row blocks (spec_index); the caller
    // gets it mode-major
    for (int m = 0; m < M; ++m)
        for (int j = 0; j < nT; ++j) {
            const size_t src = 2 * spec_index(m, j, M), dst = 2 * ((size_t)m * nT + j);
            spec[dst] = blocked[src];
            spec[dst + 1] = blocked[src + 1];
        }
    return KAM_OK;
}

int kam_sim_step_phases(kam_sim* s, int phases) {
    if (!s) return set_err(KAM_EINVAL, "null simulation");
    if (phases & ~KAM_PHASE_ALL) return set_err(KAM_EINVAL, "unknown phase bits 0x%x", phases);
    KAM_HIP(hipSetDevice(s->sd.device));
    int np = s->parity;
    int rc = s->batch.launch(phases, s->parity, s->stream, &np);
    if (rc != KAM_OK) return rc;
    s->parity = np;
    return KAM_OK;
}

int kam_sim_step(kam_sim* s, int nsteps) {
    if (!s) return set_err(KAM_EINVAL, "null simulation");
    if (nsteps < 0) return set_err(KAM_EINVAL, "negative step count");
    for (int k = 0; k < nsteps; ++k) {
        int rc = kam_sim_step_phases(s, KAM_PHASE_ALL);
        if (rc != KAM_OK) return rc;
    }
    return KAM_OK;
}

int kam_sim_benchmark(kam_sim* s, int nsteps, int per_kernel, double* step_ms, double* kernel_ms,
                      double* total_ms) {
    if (!s || nsteps <= 0) return set_err(KAM_EINVAL, "bad benchmark arguments");
    KAM_HIP(hipSetDevice(s->sd.device));
    std::vector<hipEvent_t> ev((size_t)nsteps + 1);
    hipEvent_t kev[2 * KC_COUNT];
    for (auto& e : ev) KAM_HIP(hipEventCreate(&e));
    for (auto& e : kev) KAM_HIP(hipEventCreate(&e));
    double ksum[KC_COUNT] = {0, 0, 0, 0};
    int rc = KAM_OK;
    KAM_HIP(hipEventRecord(ev[0], s->stream));
    for (int k = 0; k < nsteps && rc == KAM_OK; ++k) {
        int np = s->parity;
        rc = s->batch.launch(KAM_PHASE_ALL, s->parity, s->stream, &np, per_kernel ? kev : nullptr);
        s->parity = np;
        if (rc == KAM_OK && per_kernel) {
            KAM_HIP(hipEventSynchronize(kev[2 * KC_IFFT + 1]));
            for (int c = 0; c < KC_COUNT; ++c) {
                float ms = 0.f;
                KAM_HIP(hipEventElapsedTime(&ms, kev[2 * c], kev[2 * c + 1]));
                ksum[c] += ms;
            }
        }
        KAM_HIP(hipEventRecord(ev[(size_t)k + 1], s->stream));
    }
    KAM_HIP(hipStreamSynchronize(s->stream));
    if (rc == KAM_OK) {
        float tot = 0.f;
        KAM_HIP(hipEventElapsedTime(&tot, ev[0], ev[(size_t)nsteps]));
        if (total_ms) *total_ms = tot;
        if (step_ms)
            for (int k = 0; k < nsteps; ++k) {
                float ms = 0.f;
                KAM_HIP(hipEventElapsedTime(&ms, ev[(size_t)k], ev[(size_t)k + 1]));
                step_ms[k] = ms;
            }
        if (kernel_ms)
            for (int c = 0; c < KC_COUNT; ++c) kernel_ms[c] = per_kernel ? ksum[c] / nsteps : 0.0;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (auto& e : kev) (void)hipEventDestroy(e);
    return rc;
}

int kam_sim_status(kam_sim* s, int* status) {
    if (!s || !status) return set_err(KAM_EINVAL, "null argument");
    KAM_HIP(hipSetDevice(s->sd.device));
    KAM_HIP(hipMemcpyAsync(status, s->desc.status, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    KAM_HIP(hipStreamSynchronize(s->stream));
    return KAM_OK;
}

int kam_sim_sync(kam_sim* s) {
    if (!s) return set_err(KAM_EINVAL, "null simulation");
    KAM_HIP(hipSetDevice(s->sd.device));
    KAM_HIP(hipStreamSynchronize(s->stream));
    return KAM_OK;
}

}  // extern "C"
