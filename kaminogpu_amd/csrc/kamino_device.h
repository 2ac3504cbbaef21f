// kamino_device.h -- device-side data layout shared by the gfx950 kernels and
// the host runtime.  Plain structs, no HIP types in the layout.
//
// HBM layout of one simulation (all FP32, dense theta-major rows, phi
// contiguous; KaminoQuantity.cu:70-73 uses the same [theta][phi] order):
//   state[b][0] u_phi   [nTheta][nPhi]      b = 0, 1 (ping-pong per step)
//   state[b][1] u_theta [nTheta][nPhi]      (row nTheta-1 is padding, zeroed)
//   state[b][2] rho     [nTheta][nPhi]
//   spec        M x nTheta4 float2          half spectrum, M = nPhi/2, in blocks of
//                                           four theta rows (spec_index below;
//                                           nTheta4 = nTheta rounded up to 4);
//                                           slot 0 packs (mode 0, Nyquist mode)
// 28 bytes per cell in total (a little more when nTheta is not a multiple of
// 4; plus one 256-byte block for the watchdog word of
// a stand-alone simulation; engine simulations keep theirs in a slot table).
// With the FP64 pressure (SIM_P64), the solved spectrum p-hat is written as
// double2 [M][nTheta] over the state copy the step has finished with (the
// other parity: 2 of its 3 fields' worth of space), where the inverse
// transform reads it; the spectrum buffer keeps d-hat.
// Particles, when a simulation has them, live in their own allocation:
//   part[b] float2 (phi, theta) [nPart], b = 0, 1 (ping-pong with the fields).  Per-size tables (row metrics, FP64 Poisson
// rows, twiddles) are shared by every simulation of the same shape and dt.
#pragma once

#include <stdint.h>

namespace kam {

enum FieldIdx { F_UPHI = 0, F_UTHETA = 1, F_RHO = 2 };

// Row tables, FP32, each nTheta long, packed one after another.
enum RowTab {
    RT_COFPHI_C = 0,  // dt / (R sin(theta_j) h), theta_j = (j + 1/2) h   (u_phi and rho nodes)
    RT_COFPHI_T = 1,  // dt / (R sin(theta) h),  theta = (j + 1) h        (u_theta nodes)
    RT_G = 2,         // dt cos(theta_j) / (R sin(theta_j))                (geometric)
    RT_DIVS = 3,      // sin(theta_j + h/2) / (h sin theta_j)
    RT_DIVN = 4,      // sin(theta_j - h/2) / (h sin theta_j)
    RT_INVHS = 5,     // 1 / (h sin theta_j)
    // the tridiagonal rows for the FP32 elimination, FP64 values rounded once:
    RT_TRIA = 6,      // sub-diagonal 1/h^2 - cot(theta_j)/(2h); 0 at j = 0 (pole fold)
    RT_TRIC = 7,      // super-diagonal 1/h^2 + cot(theta_j)/(2h); 0 at j = nTheta-1
    RT_TRIS = 8,      // 1 / sin^2(theta_j) (diagonal excess n^2 / sin^2 per mode n)
    RT_COUNT = 9
};

// FP64 tables for the tridiagonal rows, each nTheta long.
enum RowTabD {
    RD_COT2H = 0,    // cot(theta_j) / (2 h)
    RD_INVSIN2 = 1,  // 1 / sin^2(theta_j)
    RD_INVHS = 2,    // 1 / (h sin theta_j), the FP64 pressure gradient's phi factor
    RD_COUNT = 3
};

enum SimFlags {
    SIM_GEOMETRIC = 1,
    SIM_P64 = 2,  // FP64 pressure: p-hat stored, transformed and differentiated in FP64
    SIM_DFT = 4   // phi transforms as dense DFTs on MFMA (FP16 hi/lo matrices) instead of FFTs
};

struct SimDesc {
    int nTheta, nPhi, M;   // M = nPhi / 2
    int flags;             // SimFlags
    float cofTheta;        // dt / (R h): theta cells per unit u_theta
    float invH;            // 1 / h
    double invH2;          // 1 / h^2
    float* state[2][3];    // [parity][field]
    float* spec;           // float2 M x nTheta4, row-blocked (spec_index)
    const float* rowf;     // RT_COUNT * nTheta
    const double* rowd;    // RD_COUNT * nTheta
    const float* twN;      // float2 [nPhi]: exp(-2 pi i n / nPhi)
    const double* twN64;   // double2 [nPhi]: the same in FP64 (the FP64 pressure transform)
    double invH64;         // 1 / h in FP64
    int nfac;              // radix factors of M, innermost first
    int fac[12];
    int twStep[12];        // per pass p: twiddle index stride nPhi / (Ls r), Ls = fac[0] ... fac[p-1]
    unsigned lsMagic[12];  // per pass p: ceil(2^32 / Ls) (0 when Ls is a power of two)
    int fftShift;          // log2 of the lanes per row of the row FFT
    int advRows;           // rows per advection tile
    int fftRows;           // rows per FFT tile
    unsigned magicP;       // ceil(2^32 / nPhi): i / nPhi == umulhi(i, magicP) for i < 2^16
    unsigned magicP2;      // ceil(2^32 / (nPhi / 2))
    float invNP;           // 1 / nPhi
    int advCols;           // columns per advection tile (W <= min(nPhi, 255))
    unsigned magicW[5];    // ceil(2^32 / d) for d = W, W + 1, W + 2, the staged width W + 8 and its quarter
    int* status;           // watchdog word: bit 0 set when a step produced a non-finite value
    float* part[2];        // particles, float2 (phi, theta) radians [nPart], per parity
    int nPart;             // particle count (0: none)
    float radius, dt;      // FP32, as the reference's constants (KaminoCore.cu:868-872)
    double* resid;         // per-mode Poisson residual report (diagnostic launches only)
    // SIM_DFT: FP16 bit patterns of the forward and inverse DFT matrices, hi and
    // lo planes, each [N][dftK] (N = nPhi rounded up to 128 output columns,
    // dftK = nPhi rounded up to 32), scaled by 2^13
    const uint16_t* dftFwd[2];
    const uint16_t* dftInv[2];
    int dftK;
};

// The FP32 half spectrum is stored in blocks of kSpecRows = 4 theta rows,
// mode-major within a block, so mode n of row j is complex entry
//     4b M + 4n + (j - 4b),   b = j / 4;
// rows past nTheta up to the next multiple of 4 are padding (never written
// or read as values).  A transform tile of four rows reads or writes one
// contiguous run of 4 M entries, and the solve reads a tile's modes as runs
// of (modes x 4 rows).
constexpr int kSpecRows = 4;
__host__ __device__ inline int spec_rows(int nT) { return (nT + kSpecRows - 1) & ~(kSpecRows - 1); }
__host__ __device__ inline size_t spec_index(int n, int j, int M) {
    const int j0 = j & ~(kSpecRows - 1);
    return (size_t)j0 * M + (size_t)n * kSpecRows + (j - j0);
}

// One unit of work: rows [row0, row0 + rows) of simulation `sim` (or, for the
// tridiagonal kernel, modes [row0, row0 + rows)); advection tiles also start
// at column col0 and span the simulation's advCols columns.
struct Tile {
    int sim;
    int row0;
    int rows;
    int col0;
};

}  // namespace kam
