// kamino_launch.h -- host launchers of the step kernels (kamino_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "kamino_device.h"

namespace kam {

// An advection tile of R rows and W <= 128 columns stages its input fields
// over R + 2 kAdvHaloRows rows and W + 2 kAdvHaloCols columns of LDS, at a
// row stride of adv_stride(W) floats (W rounded up to 64 or 128, plus halo).
constexpr int kAdvHaloRows = 3, kAdvHaloCols = 4, kAdvMaxCols = 128, kAdvMaxRows = 64;
constexpr int adv_stride(int W) { return (W <= 64 ? 64 : 128) + 2 * kAdvHaloCols; }

hipError_t launch_advect(const SimDesc* sims, const Tile* tiles, int ntiles, int parity,
                         int stride, size_t ldsBytes, int gridCap, hipStream_t st);
hipError_t launch_particles(const SimDesc* sims, const Tile* tiles, int ntiles, int parity,
                            int gridCap, hipStream_t st);
hipError_t launch_div_fft(const SimDesc* sims, const Tile* tiles, int ntiles, int parity,
                          size_t ldsBytes, int gridCap, hipStream_t st, bool half);
// `parity` names the current state copy: a simulation with the FP64 pressure
// keeps its solved spectrum over the other one.
// fp32: run the elimination in FP32 with subtraction-free pivots (the FP32
// pressure); else in FP64 (the FP64 pressure).
hipError_t launch_tridiag(int lanesPerMode, int kmax, const SimDesc* sims, const Tile* tiles,
                          int ntiles, int parity, int gridCap, hipStream_t st, bool resid = false,
                          bool fp32 = false);
// p64: the tiles belong to simulations with the FP64 pressure (SIM_P64).
// pout (one simulation only): write its pressure rows there instead of
// updating the velocities.
hipError_t launch_ifft_grad(const SimDesc* sims, const Tile* tiles, int ntiles, int parity,
                            size_t ldsBytes, int gridCap, hipStream_t st, bool p64 = false,
                            float* pout = nullptr);

// The phi transforms of SIM_DFT simulations (kamino_dft.hip): divergence rows
// + forward DFT product, and inverse DFT product + gradient rows.  rowTiles
// are row runs of the simulations, gemmTiles 64-row x 128-column output
// blocks (row0, rows, col0).
hipError_t launch_dft_forward(const SimDesc* sims, const Tile* rowTiles, int nRow, const Tile* gemmTiles,
                              int nGemm, int parity, hipStream_t st);
hipError_t launch_dft_inverse(const SimDesc* sims, const Tile* rowTiles, int nRow, const Tile* gemmTiles,
                              int nGemm, int parity, hipStream_t st);
// The inverse DFT product alone: pressure rows over the finished state copy.
hipError_t launch_dft_pressure(const SimDesc* sims, const Tile* gemmTiles, int nGemm, int parity,
                               hipStream_t st);
constexpr int kDftTileRows = 64, kDftTileCols = 128, kDftChunk = 32;

// Raises the dynamic-LDS limit of every step kernel to `bytes` (call once).
hipError_t configure_kernels(size_t bytes);

}  // namespace kam
