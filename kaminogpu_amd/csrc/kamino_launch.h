// kamino_launch.h -- host launchers of the step kernels (kamino_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "kamino_device.h"

namespace kam {

hipError_t launch_advect(const SimDesc* sims, const Tile* tiles, int ntiles, int parity,
                         size_t ldsBytes, int gridCap, hipStream_t st);
hipError_t launch_div_fft(const SimDesc* sims, const Tile* tiles, int ntiles, int parity,
                          size_t ldsBytes, int gridCap, hipStream_t st);
hipError_t launch_tridiag(int lanesPerMode, int kmax, const SimDesc* sims, const Tile* tiles,
                          int ntiles, int gridCap, hipStream_t st);
hipError_t launch_ifft_grad(const SimDesc* sims, const Tile* tiles, int ntiles, int parity,
                            size_t ldsBytes, int gridCap, hipStream_t st);

// Raises the dynamic-LDS limit of every step kernel to `bytes` (call once).
hipError_t configure_kernels(size_t bytes);

}  // namespace kam
