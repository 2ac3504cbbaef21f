// kamino_devutil.h -- small device helpers shared by the kernel sources:
// buffer-resource loads and stores (32-bit offsets, hardware range check:
// reads past the range return 0), the XCD-aware tile mapping and the
// watchdog flag.
#pragma once

#include <hip/hip_runtime.h>

namespace kam {

typedef __amdgpu_buffer_rsrc_t Rsrc;

__device__ __forceinline__ Rsrc make_rsrc(const float* p, int nfloats) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, nfloats * 4, 0x00020000);
}
__device__ __forceinline__ float ldf(Rsrc r, int elem) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, elem * 4, 0, 0));
}
__device__ __forceinline__ float2 ldf2(Rsrc r, int elem) {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, elem * 4, 0, 0));
}
__device__ __forceinline__ double ldd(Rsrc r, int elem) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, elem * 8, 0, 0));
}
__device__ __forceinline__ float4 ldf4(Rsrc r, int elem) {
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, elem * 4, 0, 0));
}
// A store at byte offset vbytes + sbytes: a row walk passes the lane's column
// in vbytes (loop-invariant) and the row's base in sbytes (a scalar), so the
// address costs no vector instruction.  Stores past the range are dropped.
__device__ __forceinline__ void stf(Rsrc r, int vbytes, int sbytes, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, vbytes, sbytes, 0);
}
__device__ __forceinline__ float lerpf(float a, float b, float t) { return fmaf(t, b - a, a); }

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Block b takes tile (b % 8) * per + b / 8, so every XCD walks a contiguous
// run of the tile list and neighbouring tiles, which share halo rows and
// spectrum lines, meet in the same L2.  Launches use 8 * per blocks.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int ntiles) {
    const int per = (ntiles + kXcds - 1) / kXcds;
    return (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
}

// Watchdog: any non-finite result of a step marks the simulation (a vector
// atomic; taken only on the rare failing path).
__device__ __forceinline__ void flag_bad(int* status, bool bad) {
    if (bad && status) atomicOr(status, 1);
}

}  // namespace kam
