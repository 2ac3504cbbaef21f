// kamino_dft.hip -- the phi transform of the projection as a dense DFT on
// MFMA, the alternative to the row FFT that the autotuner can pick per size
// (the reference's cuFFT calls, KaminoCore.cu:763-764 and :812-813).
//
// Forward: the divergence rows d[j][0..nPhi) (fillDivergenceKernel,
// :587-638) times a real nPhi x nPhi matrix give the half spectrum in the
// FFT's layout: spec[n][j] = (Re, Im) of d-hat_n = (1/nPhi) sum_i d_i
// e^{-i n i h}, slot 0 = (d-hat_0, d-hat_Nyquist).  Inverse: the solved
// half spectrum times the synthesis matrix gives the pressure rows, mode 0
// dropped (shiftUKernel, :682-703); a row kernel then subtracts the gradient
// (:705-747).  Both products run as
//     C = A B,   A FP32 [rows][K],   B [K][N] held as FP16 hi + lo planes,
// with A scaled per row by a power of two (max |A| into [2^13, 2^14)) and
// split into FP16 hi + lo on the fly, B scaled by 2^13; each 16-deep block
// takes three v_mfma_f32_32x32x16_f16 (hi.hi + hi.lo + lo.hi) into one FP32
// accumulator.  The dropped lo.lo term is 2^-22 relative, so the products
// carry FP32-level accuracy (a BF16 hi/lo split, 8 + 8 bits, does not).
#include <hip/hip_runtime.h>

#include "kamino_device.h"
#include "kamino_devutil.h"
#include "kamino_launch.h"

namespace kam {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// Workgroup tile: kDftRows rows x kDftCols output columns, K in chunks of
// kDftK; four wavefronts, each owning 32 columns and both 32-row blocks.
constexpr int kDftRows = kDftTileRows, kDftCols = kDftTileCols, kDftK = kDftChunk;
constexpr int kAPad = kDftK + 8;  // LDS row pitch of the A and B chunks (halves)

// Operand layout of v_mfma_f32_32x32x16_f16: lane l holds A[l % 32][8 (l / 32) + t]
// and B[8 (l / 32) + t][l % 32], t = 0..7; accumulator v of lane l is
// C[8 (v / 4) + 4 (l / 32) + v % 4][l % 32].

__device__ __forceinline__ void split_h(float x, _Float16& hi, _Float16& lo) {
    hi = (_Float16)x;
    lo = (_Float16)(x - (float)hi);
}

// Reads A[j][k] of the tile's simulation: the divergence rows (forward) or the
// solved spectrum seen as rows of (Re, Im) pairs (inverse).  k beyond nPhi,
// and rows beyond the field, read as 0.
template <bool INV>
__device__ __forceinline__ float load_a(const SimDesc& s, Rsrc a, int j, int k) {
    if (k >= s.nPhi || j >= s.nTheta) return 0.f;
    if constexpr (INV) return ldf(a, 2 * (int)spec_index(k >> 1, j, s.M) + (k & 1));
    else return ldf(a, j * s.nPhi + k);
}

template <bool INV>
__global__ __launch_bounds__(256) void k_dft_gemm(const SimDesc* __restrict__ sims,
                                                  const Tile* __restrict__ tiles, int ntiles, int parity) {
    __shared__ __attribute__((aligned(16))) _Float16 aHi[kDftRows * kAPad], aLo[kDftRows * kAPad];
    __shared__ __attribute__((aligned(16))) _Float16 bHi[kDftCols * kAPad], bLo[kDftCols * kAPad];
    __shared__ int rowExp[kDftRows];
    __shared__ unsigned rowMax[4][kDftRows];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int t = xcd_tile(ntiles); t < ntiles; t = ntiles) {
        const Tile tile = tiles[t];
        const SimDesc& s = sims[tile.sim];
        const int nT = s.nTheta, nP = s.nPhi, K = s.dftK;
        const int j0 = tile.row0, c0 = tile.col0;
        // A: the divergence rows over the finished state copy (forward) or
        // the spectrum (inverse); B: this direction's FP16 planes, [N][K]
        const Rsrc a = INV ? make_rsrc(s.spec, 2 * s.M * spec_rows(nT)) : make_rsrc(s.state[parity ^ 1][F_UPHI], nT * nP);
        const _Float16* __restrict__ gHi = reinterpret_cast<const _Float16*>(INV ? s.dftInv[0] : s.dftFwd[0]);
        const _Float16* __restrict__ gLo = reinterpret_cast<const _Float16*>(INV ? s.dftInv[1] : s.dftFwd[1]);

        // per-row scale: max |A[j][.]| over the whole row.  Forward: each
        // wavefront walks rows with its lanes along the row (contiguous
        // reads), then reduces across lanes; inverse (the spectrum in row blocks):
        // lanes along the rows, four partial maxima per row.
        if constexpr (INV) {
            const int r = tid & 63, q = tid >> 6;
            unsigned m = 0u;
            for (int k = q; k < nP; k += 4) m = max(m, __float_as_uint(load_a<INV>(s, a, j0 + r, k)) & 0x7fffffffu);
            rowMax[q][r] = m;
        } else {
            for (int r = wave; r < kDftRows; r += 4) {
                unsigned m = 0u;
                for (int k = lane; k < nP; k += 64)
                    m = max(m, __float_as_uint(load_a<INV>(s, a, j0 + r, k)) & 0x7fffffffu);
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o, 64));
                if (lane < 4) rowMax[lane][r] = m;
            }
        }
        __syncthreads();
        if (tid < kDftRows) {
            const unsigned m = max(max(rowMax[0][tid], rowMax[1][tid]), max(rowMax[2][tid], rowMax[3][tid]));
            int e = 0;
            (void)frexpf(__uint_as_float(m), &e);  // m = f 2^e, f in [0.5, 1)
            // finite rows scale into [2^13, 2^14); a non-finite row keeps its values
            rowExp[tid] = (m < 0x7f800000u && m != 0u) ? 14 - e : 0;
        }
        __syncthreads();

        floatx16 acc[2];
        for (int b = 0; b < 2; ++b)
            for (int v = 0; v < 16; ++v) acc[b][v] = 0.f;
        // software pipeline: the next chunk's A values and B halves are loaded
        // into registers while the MFMAs run on the current chunk in LDS.
        // A: 64 rows x 32 k, thread -> row tid / 4, k run (tid % 4) * 8;
        // B: 128 columns x 32 k, thread -> column tid / 2, k run (tid % 2) * 16
        const int ar = tid >> 2, akq = (tid & 3) * 8, bc = tid >> 1, bkh = (tid & 1) * 16;
        float xa[8];
        half8 xb[4];
        auto fetch = [&](int k0) {
#pragma unroll
            for (int u = 0; u < 8; ++u) xa[u] = load_a<INV>(s, a, j0 + ar, k0 + akq + u);
            const size_t g = (size_t)(c0 + bc) * K + k0 + bkh;
            const half8* sh = reinterpret_cast<const half8*>(gHi + g);
            const half8* sl = reinterpret_cast<const half8*>(gLo + g);
            xb[0] = sh[0];
            xb[1] = sh[1];
            xb[2] = sl[0];
            xb[3] = sl[1];
        };
        fetch(0);
        for (int k0 = 0; k0 < K; k0 += kDftK) {
            {
                const int ex = rowExp[ar];
                half8 h, l;
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    _Float16 hh, ll;
                    split_h(ldexpf(xa[u], ex), hh, ll);
                    h[u] = hh;
                    l[u] = ll;
                }
                *reinterpret_cast<half8*>(aHi + ar * kAPad + akq) = h;
                *reinterpret_cast<half8*>(aLo + ar * kAPad + akq) = l;
                *reinterpret_cast<half8*>(bHi + bc * kAPad + bkh) = xb[0];
                *reinterpret_cast<half8*>(bHi + bc * kAPad + bkh + 8) = xb[1];
                *reinterpret_cast<half8*>(bLo + bc * kAPad + bkh) = xb[2];
                *reinterpret_cast<half8*>(bLo + bc * kAPad + bkh + 8) = xb[3];
            }
            __syncthreads();
            if (k0 + kDftK < K) fetch(k0 + kDftK);
#pragma unroll
            for (int ks = 0; ks < kDftK; ks += 16) {
                const int kk = ks + 8 * (lane >> 5), cc = wave * 32 + (lane & 31);
                const half8 bh = *reinterpret_cast<const half8*>(bHi + cc * kAPad + kk);
                const half8 bl = *reinterpret_cast<const half8*>(bLo + cc * kAPad + kk);
#pragma unroll
                for (int rb = 0; rb < 2; ++rb) {
                    const int rr = rb * 32 + (lane & 31);
                    const half8 ah = *reinterpret_cast<const half8*>(aHi + rr * kAPad + kk);
                    const half8 al = *reinterpret_cast<const half8*>(aLo + rr * kAPad + kk);
                    acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, acc[rb], 0, 0, 0);
                    acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, acc[rb], 0, 0, 0);
                    acc[rb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, acc[rb], 0, 0, 0);
                }
            }
            __syncthreads();
        }
        // unscale (B carries 2^13; the forward transform also 1 / nPhi) and store
        const int c = c0 + wave * 32 + (lane & 31);
        if (c < nP) {
            float* __restrict__ specOut = s.spec;
            float* __restrict__ pOut = s.state[parity ^ 1][F_UTHETA];
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int r = rb * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3), j = j0 + r;
                    if (r >= tile.rows) continue;
                    const float val = ldexpf(acc[rb][v], -rowExp[r] - 13);
                    if constexpr (INV)
                        pOut[(size_t)j * nP + c] = val;
                    else
                        specOut[2 * spec_index(c >> 1, j, s.M) + (c & 1)] = val * s.invNP;
                }
        }
        __syncthreads();
    }
}

// Divergence rows of a DFT simulation into the finished state copy, the
// expression of k_div_fft: d = (sin(t+h/2) v_S - sin(t-h/2) v_N) / (h sin t)
// + (u_E - u_W) / (h sin t), v_N = 0 on the first row, v_S = 0 on the last.
__global__ __launch_bounds__(256) void k_div_rows(const SimDesc* __restrict__ sims,
                                                  const Tile* __restrict__ tiles, int ntiles, int parity) {
    for (int t = xcd_tile(ntiles); t < ntiles; t = ntiles) {
        const Tile tile = tiles[t];
        const SimDesc& s = sims[tile.sim];
        const int nT = s.nTheta, nP = s.nPhi;
        const Rsrc upR = make_rsrc(s.state[parity][F_UPHI], nT * nP);
        const Rsrc utR = make_rsrc(s.state[parity][F_UTHETA], (nT - 1) * nP);
        float* __restrict__ d = s.state[parity ^ 1][F_UPHI];
        for (int idx = threadIdx.x; idx < tile.rows * nP; idx += blockDim.x) {
            const int r = idx / nP, i = idx - r * nP, j = tile.row0 + r;
            const float fs = s.rowf[RT_DIVS * nT + j], fn = s.rowf[RT_DIVN * nT + j];
            const float fi = s.rowf[RT_INVHS * nT + j];
            const float uw = ldf(upR, j * nP + i), ue = ldf(upR, j * nP + (i + 1 == nP ? 0 : i + 1));
            const float vn = ldf(utR, j > 0 ? (j - 1) * nP + i : -1), vs = ldf(utR, j * nP + i);
            d[(size_t)j * nP + i] = (fs * vs - fn * vn) + fi * (ue - uw);
        }
    }
}

// Gradient subtraction from the pressure rows of a DFT simulation
// (applyPressureTheta / applyPressurePhi, KaminoCore.cu:705-747), the
// expression of k_ifft_grad.
__global__ __launch_bounds__(256) void k_grad_rows(const SimDesc* __restrict__ sims,
                                                   const Tile* __restrict__ tiles, int ntiles, int parity) {
    for (int t = xcd_tile(ntiles); t < ntiles; t = ntiles) {
        const Tile tile = tiles[t];
        const SimDesc& s = sims[tile.sim];
        const int nT = s.nTheta, nP = s.nPhi;
        const float* __restrict__ p = s.state[parity ^ 1][F_UTHETA];
        float* __restrict__ up = s.state[parity][F_UPHI];
        float* __restrict__ ut = s.state[parity][F_UTHETA];
        const float invH = s.invH;
        for (int idx = threadIdx.x; idx < tile.rows * nP; idx += blockDim.x) {
            const int r = idx / nP, i = idx - r * nP, j = tile.row0 + r;
            const float hs = s.rowf[RT_INVHS * nT + j];
            const float* pr = p + (size_t)j * nP;
            const float p0 = pr[i], pw = pr[i == 0 ? nP - 1 : i - 1];
            const size_t o = (size_t)j * nP + i;
            const float nu = up[o] - (p0 - pw) * hs;
            up[o] = nu;
            bool bad = !isfinite(nu);
            if (j < nT - 1) {
                const float nv = ut[o] - (pr[nP + i] - p0) * invH;
                ut[o] = nv;
                bad |= !isfinite(nv);
            }
            flag_bad(s.status, bad);
        }
    }
}

static inline int dft_grid(int ntiles) { return ((ntiles + kXcds - 1) / kXcds) * kXcds; }

hipError_t launch_dft_forward(const SimDesc* sims, const Tile* rowTiles, int nRow, const Tile* gemmTiles,
                              int nGemm, int parity, hipStream_t st) {
    if (nRow <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_div_rows, dim3(dft_grid(nRow)), dim3(256), 0, st, sims, rowTiles, nRow, parity);
    hipLaunchKernelGGL(k_dft_gemm<false>, dim3(dft_grid(nGemm)), dim3(256), 0, st, sims, gemmTiles, nGemm,
                       parity);
    return hipGetLastError();
}

hipError_t launch_dft_pressure(const SimDesc* sims, const Tile* gemmTiles, int nGemm, int parity,
                               hipStream_t st) {
    if (nGemm <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dft_gemm<true>, dim3(dft_grid(nGemm)), dim3(256), 0, st, sims, gemmTiles, nGemm,
                       parity);
    return hipGetLastError();
}

hipError_t launch_dft_inverse(const SimDesc* sims, const Tile* rowTiles, int nRow, const Tile* gemmTiles,
                              int nGemm, int parity, hipStream_t st) {
    if (nRow <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dft_gemm<true>, dim3(dft_grid(nGemm)), dim3(256), 0, st, sims, gemmTiles, nGemm,
                       parity);
    hipLaunchKernelGGL(k_grad_rows, dim3(dft_grid(nRow)), dim3(256), 0, st, sims, rowTiles, nRow, parity);
    return hipGetLastError();
}

}  // namespace kam
