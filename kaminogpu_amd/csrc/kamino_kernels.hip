// kamino_kernels.hip -- gfx950 kernels of the spherical-fluid time step.
//
// One step of a batch is four launches over a device table of simulations
// (SimDesc) and a list of tiles:
//   k_advect      RK2 semi-Lagrangian advection of u_phi, u_theta and rho with
//                 the geometric (metric-term) update fused into its epilogue
//   k_div_fft     divergence + real FFT along phi per latitude row
//   k_tridiag     one tridiagonal solve in theta per Fourier mode
//   k_ifft_grad   inverse real FFT + pressure-gradient subtraction
// All coordinates are handled in cell-index space: an integer node index plus
// a small float offset, so FP32 keeps its precision for the interpolation
// weights however far the node sits from phi = 0.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "kamino_device.h"
#include "kamino_devutil.h"
#include "kamino_launch.h"

namespace kam {

enum Kind { K_VPHI = 0, K_VTHETA = 1, K_CENTERED = 2 };

// ---------------------------------------------------------------------------
// Sampling (KaminoCore.cu:11-184).  The reference folds radian coordinates
// over the poles (validateCoord, :11-29) and picks one of two bilinear
// branches; this is the same map in index space.  `bj + fth` is the
// offset-shifted theta coordinate in cells, `bi + fphi` the phi coordinate.
// Fields are read through buffer resources (32-bit offsets, hardware range
// check), and the branch choice is made with selects, not jumps.
// ---------------------------------------------------------------------------
// Per-shape constants a sampler needs.
struct Grid {
    int nT, nP, half;
    float invNP;
};

// A coordinate in cells as its cell index and the offset within the cell
// (v_cvt_flr_i32_f32 and v_fract_f32).  Every sampling path splits its
// coordinates this way, so they all interpolate with identical weights.
__device__ __forceinline__ int cell_of(float x) { return (int)floorf(x); }
__device__ __forceinline__ float frac_of(float x) { return __builtin_amdgcn_fractf(x); }

template <int KIND>
__device__ __forceinline__ float sample(Rsrc f, const Grid& g, int bi, float fphi, int bj, float fth) {
    const int rows = (KIND == K_VTHETA) ? g.nT - 1 : g.nT;
    const int poleRow = (KIND == K_VTHETA) ? g.nT - 2 : g.nT - 1;
    int ip = bi + cell_of(fphi);
    const float ap = frac_of(fphi);
    int jt = bj + cell_of(fth);
    float at = frac_of(fth);
    float a0, a1, b0, b1, w;
    if (jt >= 1 && jt < poleRow) {
        // ordinary bilinear: no fold, no pole branch, both rows in range
        ip -= __mul24(g.nP, (int)floorf((float)ip * g.invNP));
        ip += (ip < 0) ? g.nP : 0;
        ip -= (ip >= g.nP) ? g.nP : 0;
        int ip1 = ip + 1;
        ip1 = (ip1 == g.nP) ? 0 : ip1;
        const int oA = __mul24(jt, g.nP) + ip, oA1 = __mul24(jt, g.nP) + ip1;
        a0 = ldf(f, oA);
        a1 = ldf(f, oA1);
        b0 = ldf(f, oA + g.nP);
        b1 = ldf(f, oA1 + g.nP);
        w = at;
    } else {
        // fold: t < 0 -> -t and t > nTheta -> 2 nTheta - t, each adding half a turn in phi
        const bool frac = at > 0.f;
        const bool neg = jt < 0;
        const bool high = (jt > g.nT) | ((jt == g.nT) & frac);
        const bool flip = neg | high;
        const int refl = (neg ? -jt : 2 * g.nT - jt) - (int)frac;
        jt = flip ? refl : jt;
        at = (flip & frac) ? 1.f - at : at;
        ip += flip ? g.half : 0;
        // ip mod nPhi (the float quotient is off by at most one)
        ip -= __mul24(g.nP, (int)floorf((float)ip * g.invNP));
        ip += (ip < 0) ? g.nP : 0;
        ip -= (ip >= g.nP) ? g.nP : 0;
        int ip1 = ip + 1;
        ip1 = (ip1 == g.nP) ? 0 : ip1;
        // pole branch: the same row at phi and at phi + pi (alpha halved except for u_phi)
        const bool pole = ((jt == 0) & flip) | (jt == poleRow);
        int ia = ip + g.half;
        ia -= (ia >= g.nP) ? g.nP : 0;
        int ib = ia + 1;
        ib = (ib == g.nP) ? 0 : ib;
        // out-of-range rows (a backtrace past a pole) clamp to the field
        const int rA = min(max(jt, 0), rows - 1);
        const int rB = pole ? rA : min(max(jt + 1, 0), rows - 1);
        const int cB0 = pole ? ia : ip, cB1 = pole ? ib : ip1;
        w = pole ? ((KIND == K_VPHI) ? at : 0.5f * at) : at;
        const int oA = __mul24(rA, g.nP), oB = __mul24(rB, g.nP);
        a0 = ldf(f, oA + ip);
        a1 = ldf(f, oA + ip1);
        b0 = ldf(f, oB + cB0);
        b1 = ldf(f, oB + cB1);
    }
    return lerpf(lerpf(a0, a1, ap), lerpf(b0, b1, ap), w);
}

// Node offsets of each grid in cells (KaminoHeader.cuh:30-41) and the shift
// that maps a raw coordinate onto that grid's sample lattice.
__device__ constexpr float kNodeP[3] = {-0.5f, 0.f, 0.f};
__device__ constexpr float kNodeT[3] = {0.5f, 1.f, 0.5f};
__device__ constexpr float kShiftP[3] = {0.5f, 0.f, 0.f};
__device__ constexpr float kShiftT[3] = {-0.5f, -1.f, -0.5f};

// RK2 backtrace from node (i, j) of grid NODE, sampling field `src` of grid
// SRC at the departure point (KaminoCore.cu:186-319).  hP is half the factor
// that converts u_phi to phi cells per step at this node's latitude, hT half
// the one for u_theta to theta cells: the midpoint lies at x - h u0 and the
// departure point at x - h (u0 + u_mid), one fma per coordinate (every
// advection path uses exactly these roundings).  A field sampled at its own
// node is the node value (both offsets integral, the ordinary branch with
// zero weights), so that read is direct.
template <int NODE, int SRC>
__device__ __forceinline__ float advect_node(Rsrc up, Rsrc ut, Rsrc src, const Grid& g, int i, int j,
                                             float hP, float hT) {
    constexpr float pP = kNodeP[NODE] + kShiftP[K_VPHI], pT = kNodeT[NODE] + kShiftT[K_VPHI];
    constexpr float tP = kNodeP[NODE] + kShiftP[K_VTHETA], tT = kNodeT[NODE] + kShiftT[K_VTHETA];
    constexpr float sP = kNodeP[NODE] + kShiftP[SRC], sT = kNodeT[NODE] + kShiftT[SRC];
    const int o = __mul24(j, g.nP) + i;
    const float u0p = (NODE == K_VPHI) ? ldf(up, o) : sample<K_VPHI>(up, g, i, pP, j, pT);
    const float u0t = (NODE == K_VTHETA) ? ldf(ut, o) : sample<K_VTHETA>(ut, g, i, tP, j, tT);
    const float ump = sample<K_VPHI>(up, g, i, fmaf(-u0p, hP, pP), j, fmaf(-u0t, hT, pT));
    const float umt = sample<K_VTHETA>(ut, g, i, fmaf(-u0p, hP, tP), j, fmaf(-u0t, hT, tT));
    return sample<SRC>(src, g, i, fmaf(-(u0p + ump), hP, sP), j, fmaf(-(u0t + umt), hT, sT));
}

// ---------------------------------------------------------------------------
// Geometric update (KaminoCore.cu:457-514): the real root x of
// G^2 x^3 + (1 + G vbar) x - ubar = 0, or ubar when |G| <= 1e-7.  The
// reference evaluates Cardano's formula in FP32, whose two terms cancel to
// the small root (relative error ~1e-3 near the equator); when 1 + G vbar > 0
// the cubic is monotone and Newton from the smaller of the root's two bounds
// converges to the same unique root to FP32 precision.  Otherwise (three real
// roots possible)
// the reference's own branch choice is evaluated in FP64.
// ---------------------------------------------------------------------------
__device__ double ref_root3_pos(double x) {
    // x > 0.  A non-finite x (from an overflowed velocity) is returned as is,
    // for the watchdog to flag; scaling +inf into [1, 8] would never end.
    if (!(x <= 1.7976931348623157e308)) return x;
    double s = 1.;
    while (x < 1.) {
        x *= 8.;
        s *= 0.5;
    }
    while (x > 8.) {
        x *= 0.125;
        s *= 2.;
    }
    double r = 1.5;
    for (int k = 0; k < 6; ++k) r -= 1. / 3. * (r - x / (r * r));
    return r * s;
}

__device__ __noinline__ float cardano_fp64(double G, double vb, double ub) {
    const double cof = G * G;
    const double B = (G * vb + 1.0) / cof;
    const double Cc = -ub / cof;
    const double q = -3.0 * B / 9.0;
    const double r = 27.0 * Cc / 54.0;
    const double r2 = r * r, q3 = q * q * q;
    if (r2 <= q3 + (double)1e-7f) {
        double t = r / sqrt(q3);
        t = t < -1 ? -1 : (t > 1 ? 1 : t);
        return (float)(-2.0 * sqrt(q) * cos(acos(t) / 3.0));
    }
    const double s = fabs(r) + sqrt(r2 - q3);
    double A = -(s > 0 ? ref_root3_pos(s) : 0.0);
    if (r < 0) A = -A;
    const double Bq = (A == 0) ? 0 : q / A;
    return (float)(A + Bq);
}

__device__ __forceinline__ float geo_root(float G, float vb, float ub) {
    if (!(fabsf(G) > 1e-7f)) return ub;
    const float b = fmaf(G, vb, 1.f);
    if (b > 0.f) {
        const float a = G * G;
        const float rb = __builtin_amdgcn_rcpf(b);
        const float y = ub * rb;
        // x = y phi(e) with e = a y^2 / b >= 0, where e phi^3 + phi - 1 = 0:
        // phi = 1 - e + 3 e^2 - 12 e^3 + 55 e^4 - ...  For e <= 3e-3 the
        // truncation error 55 e^4 |y| stays below 5e-9 |x|, under FP32
        // resolution; this covers every row but the few nearest the poles
        const float e = a * y * y * rb;
        bool done = e <= 3e-3f;
        float x = y * fmaf(e, fmaf(e, fmaf(e, -12.f, 3.f), -1.f), 1.f);
        if (__all(done)) return x;
        // the others: the root is bounded by both |ub| / b and cbrt(|ub| / G^2);
        // Newton on the monotone, convex-on-each-side cubic from the smaller
        // bound converges from above, quadratically within a few steps
        if (!done) {
            x = y;
            if (a * x * x > b) x = copysignf(cbrtf(fabsf(ub) * __builtin_amdgcn_rcpf(a)), ub);
        }
        // at most five steps.  Converging from above, the error after a step
        // of dx is at most 3 (dx / x)^2 |x| (f'' / 2f' <= 3 a |x| / (3 a x^2 + b)
        // between the iterate and the root), so a lane stops once |dx| <= 1e-4 |x|
        // (error below 3e-8 |x|, under FP32 resolution) and then no longer
        // updates; a lane's result depends on its own inputs only, never on
        // its neighbours (the wavefront leaves when every lane has stopped)
        for (int k = 0; k < 5; ++k) {
            const float x2 = x * x;
            const float f = fmaf(a * x2, x, fmaf(b, x, -ub));
            const float fp = fmaf(3.f * a, x2, b);
            const float dx = f * __builtin_amdgcn_rcpf(fp);
            if (!done) {
                x -= dx;
                done = !(fabsf(dx) > 1e-4f * fabsf(x));
            }
            if (__all(done)) break;
        }
        return x;
    }
    return cardano_fp64(G, vb, ub);
}

// ---------------------------------------------------------------------------
// LDS-staged sampling.  A tile stages the three input fields over rows
// [j0 - kER, j0 + R + kER) and columns [c0 - kEC, c0 + W + kEC) into LDS,
// wrapped at the phi seam once, while staging.  A sample whose bilinear
// stencil lies inside the staged block, in rows where the global sampler
// would take its ordinary (no fold, no pole) branch, reads the four values
// from LDS with two ds_read2 at the fixed row stride SW; any other sample
// (departure points beyond the halo near the poles, folds, pole rows) goes
// to the global sampler.  Both read the same values and interpolate with the
// same weights, so the choice never changes a result.
// ---------------------------------------------------------------------------
constexpr int kER = kAdvHaloRows;  // staged rows above and below the tile
constexpr int kEC = kAdvHaloCols;  // staged columns west and east of the tile

struct Stage {
    const float* S;      // staged field k (Kind order: u_phi, u_theta, rho) at S + k * plane
    int plane;           // SR * SW floats
    int rlo[3], rhi[3];  // per Kind: staged rows lt in [rlo, rhi) take the LDS path
    int cmax;            // staged columns lp in [0, cmax) take the LDS path
};

template <int KIND, int SW>
__device__ __forceinline__ float sample_st(const Stage& st, Rsrc f, const Grid& g, int bi, int lc,
                                           float fphi, int bj, int lr, float fth) {
    const int lp = lc + cell_of(fphi), lt = lr + cell_of(fth);
    const bool inside = ((unsigned)(lt - st.rlo[KIND]) < (unsigned)(st.rhi[KIND] - st.rlo[KIND])) &
                        ((unsigned)lp < (unsigned)st.cmax);
    if (inside) {
        const float* q = st.S + KIND * st.plane + __mul24(lt, SW) + lp;
        const float ap = frac_of(fphi), at = frac_of(fth);
        return lerpf(lerpf(q[0], q[1], ap), lerpf(q[SW], q[SW + 1], ap), at);
    }
    return sample<KIND>(f, g, bi, fphi, bj, fth);
}

// advect_node with the samples taken from the staged block where possible;
// (lc, lr) is node (i, j) in staged coordinates.
template <int NODE, int SRC, int SW>
__device__ __forceinline__ float advect_node_st(const Stage& st, Rsrc up, Rsrc ut, Rsrc src,
                                                const Grid& g, int i, int j, int lc, int lr,
                                                float hP, float hT) {
    constexpr float pP = kNodeP[NODE] + kShiftP[K_VPHI], pT = kNodeT[NODE] + kShiftT[K_VPHI];
    constexpr float tP = kNodeP[NODE] + kShiftP[K_VTHETA], tT = kNodeT[NODE] + kShiftT[K_VTHETA];
    constexpr float sP = kNodeP[NODE] + kShiftP[SRC], sT = kNodeT[NODE] + kShiftT[SRC];
    const int o = __mul24(lr, SW) + lc;
    const float u0p = (NODE == K_VPHI) ? st.S[o] : sample_st<K_VPHI, SW>(st, up, g, i, lc, pP, j, lr, pT);
    const float u0t = (NODE == K_VTHETA) ? st.S[st.plane + o]
                                         : sample_st<K_VTHETA, SW>(st, ut, g, i, lc, tP, j, lr, tT);
    const float ump = sample_st<K_VPHI, SW>(st, up, g, i, lc, fmaf(-u0p, hP, pP), j, lr, fmaf(-u0t, hT, pT));
    const float umt =
        sample_st<K_VTHETA, SW>(st, ut, g, i, lc, fmaf(-u0p, hP, tP), j, lr, fmaf(-u0t, hT, tT));
    return sample_st<SRC, SW>(st, src, g, i, lc, fmaf(-(u0p + ump), hP, sP), j, lr,
                              fmaf(-(u0t + umt), hT, sT));
}

// An opaque copy of v, which keeps the compiler from re-associating an
// address: the node's staged offset stays in one register and each sample
// adds its stencil offset to it with one shift-add.
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+v"(v));
    return v;
}
// The same for a wavefront-uniform value, kept in a scalar register: a row
// walk's addresses are then a scalar row base plus the lane's column, not
// loop-carried vector registers.
__device__ __forceinline__ int opaque_s(int v) {
    asm volatile("" : "+s"(v));
    return v;
}

// (int)floorf(x) in one instruction (the compiler selects it only under
// no-NaN math); x is a departure coordinate, |x| < 2^24.
__device__ __forceinline__ int floor_i(float x) {
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// Tile-wide fast path: when the staged velocities bound every RK2
// displacement within the halo (see k_advect) and the tile is clear of the
// pole rows, every sample is an ordinary bilinear read of the staged block,
// with no per-sample check.  Same values, same weights as sample_st.
template <int SW>
__device__ __forceinline__ float sample_fast(const float* __restrict__ nb, float fphi, float fth) {
    // nb: the field's staged value at the node; the stencil starts
    // floor(fth) rows and floor(fphi) columns away, and both rows are read
    // with ds_read2 at immediate offsets from one address
    // one mad for the stencil offset, one shift-add for its address
    const float* q = nb + opaque(__mul24(floor_i(fth), SW) + floor_i(fphi));
    const float ap = frac_of(fphi), at = frac_of(fth);
    return lerpf(lerpf(q[0], q[1], ap), lerpf(q[SW], q[SW + 1], ap), at);
}

// A sample at a constant offset (fphi, fth) from the node: both weights are
// 0 or 1/2, and a zero weight reads one value (a + 0 (b - a) is a for the
// finite velocities of a fast tile).
constexpr int cfloor(float x) { return (x >= 0.f || (float)(int)x == x) ? (int)x : (int)x - 1; }

template <int SW, int IP, int IT, bool HP, bool HT>
__device__ __forceinline__ float sample_half(const float* __restrict__ nb) {
    const float* q = nb + IT * SW + IP;
    auto row = [](const float* r) { return HP ? lerpf(r[0], r[1], 0.5f) : r[0]; };
    return HT ? lerpf(row(q), row(q + SW), 0.5f) : row(q);
}

template <int SW, int NODE, int KIND>
__device__ __forceinline__ float sample_node_const(const float* __restrict__ nb) {
    constexpr float fp = kNodeP[NODE] + kShiftP[KIND], ft = kNodeT[NODE] + kShiftT[KIND];
    constexpr int ip = cfloor(fp), it = cfloor(ft);
    static_assert((fp - ip == 0.f || fp - ip == 0.5f) && (ft - it == 0.f || ft - it == 0.5f),
                  "node offsets are whole or half cells");
    return sample_half<SW, ip, it, (fp - ip) != 0.f, (ft - it) != 0.f>(nb);
}

template <int NODE, int SRC, int SW>
__device__ __forceinline__ float advect_node_fast(const float* __restrict__ S, int plane, int node,
                                                  float hP, float hT) {
    constexpr float pP = kNodeP[NODE] + kShiftP[K_VPHI], pT = kNodeT[NODE] + kShiftT[K_VPHI];
    constexpr float tP = kNodeP[NODE] + kShiftP[K_VTHETA], tT = kNodeT[NODE] + kShiftT[K_VTHETA];
    constexpr float sP = kNodeP[NODE] + kShiftP[SRC], sT = kNodeT[NODE] + kShiftT[SRC];
    const int o = opaque(node);             // the node's staged offset, row * SW + column
    const float* bp = S + o;                // node in the u_phi plane
    const float* bt = S + (o + plane);      // ... u_theta plane
    const float* bs = S + (o + SRC * plane);  // ... source field plane
    const float u0p = (NODE == K_VPHI) ? bp[0] : sample_node_const<SW, NODE, K_VPHI>(bp);
    const float u0t = (NODE == K_VTHETA) ? bt[0] : sample_node_const<SW, NODE, K_VTHETA>(bt);
    const float ump = sample_fast<SW>(bp, fmaf(-u0p, hP, pP), fmaf(-u0t, hT, pT));
    const float umt = sample_fast<SW>(bt, fmaf(-u0p, hP, tP), fmaf(-u0t, hT, tT));
    return sample_fast<SW>(bs, fmaf(-(u0p + ump), hP, sP), fmaf(-(u0t + umt), hT, sT));
}

// Unsigned maximum over the 64 lanes of a wavefront.
__device__ __forceinline__ unsigned wave_max(unsigned v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = max(v, (unsigned)__shfl_xor((int)v, m, 64));
    return v;
}
constexpr int kAdvWaves = 256 / 64;  // wavefronts of an advection workgroup

// Displacement budgets of the fast path, in cells.  Node columns sit at
// staged columns kEC-1 .. W+kEC and sample offsets lie within half a cell,
// so phi departures below 2.5 cells keep the stencil in [0, W + 2 kEC); node
// rows sit at staged rows kER-1 .. R+kER, so theta departures below 1.5 cells
// keep it in [0, R + 2 kER).  The margins absorb FP32 rounding.
constexpr float kFastDPhi = 2.4f, kFastDTheta = 1.4f;

// ---------------------------------------------------------------------------
// k_advect: one tile = rows [j0, j0 + R) x columns [c0, c0 + W) of one
// simulation, W <= SW - 2 kEC.  The tile first stages its input block (see
// Stage).  With the geometric update on, it advects the rho nodes first
// (straight to HBM) and then, over the dead staged rho plane, into LDS the u_phi
// values of rows j0..j0+R and columns c0-1..c0+W, and the u_theta values of
// rows j0-1..j0+R and columns c0-1..c0+W-1 (the update reads one row either
// side and one column to the west; the restaggered u_phi one more column to
// the east), plus, in a polar tile, u_theta of the pole-adjacent row at the
// antipodal columns.  It then solves the cubic per cell and restaggers
// (assignPhiKernel :517-534, assignThetaKernel :537-549) before anything
// leaves the CU; the cubic roots reuse the staged block's LDS.
// Flat item loops split an index into (row, column) with a magic multiply.
// ---------------------------------------------------------------------------
template <int SW>
__global__ __launch_bounds__(256) void k_advect(const SimDesc* __restrict__ sims,
                                                const Tile* __restrict__ tiles, int ntiles,
                                                int parity) {
    extern __shared__ __attribute__((aligned(16))) float lds[];
    constexpr int nthr = kAdvWaves * 64;  // launch_advect's block size
    const int tid = threadIdx.x;
    for (int t = xcd_tile(ntiles); t < ntiles; t = ntiles) {
        const Tile tile = tiles[t];
        const SimDesc& s = sims[tile.sim];
        const Grid g{s.nTheta, s.nPhi, s.nPhi >> 1, s.invNP};
        const int nT = g.nT, nP = g.nP;
        const int N = nT * nP;
        const Rsrc up = make_rsrc(s.state[parity][F_UPHI], N);
        const Rsrc ut = make_rsrc(s.state[parity][F_UTHETA], N - nP);
        const Rsrc rh = make_rsrc(s.state[parity][F_RHO], N);
        float* __restrict__ oup = s.state[parity ^ 1][F_UPHI];
        float* __restrict__ out = s.state[parity ^ 1][F_UTHETA];
        float* __restrict__ orh = s.state[parity ^ 1][F_RHO];
        const Rsrc oupR = make_rsrc(oup, N), outR = make_rsrc(out, N - nP), orhR = make_rsrc(orh, N);
        int* const status = s.status;
        const float* __restrict__ cofPC = s.rowf + RT_COFPHI_C * nT;
        const float* __restrict__ cofPT = s.rowf + RT_COFPHI_T * nT;
        const float hT = 0.5f * s.cofTheta;  // half of theta cells per unit u_theta
        const int j0 = tile.row0, R = tile.rows, c0 = tile.col0;
        const int W = s.advCols;
        const unsigned mW = s.magicW[0], mW1 = s.magicW[1], mW2 = s.magicW[2];
        // tile column c (from -1) -> grid column, wrapping at the phi seam
        auto col = [&](int c) {
            int i = c0 + c;
            i += (i < 0) ? nP : 0;
            return i - ((i >= nP) ? nP : 0);
        };

        // ---- stage rows j0-kER .. j0+R+kER-1, columns c0-kEC .. c0+W+kEC-1
        const int SR = R + 2 * kER, SC = W + 2 * kEC, plane = SR * SW;
        const int js = j0 - kER;
        Stage st;
        st.S = lds;
        st.plane = plane;
        st.cmax = SC - 1;
        // stencil rows that read as the ordinary branch: jt >= 0 and jt < pole
        // row (nT-1; nT-2 for u_theta).  A top row jt = 0 reached without a
        // fold (the raw row index is 0, not -1) takes rows 0 and 1 at the
        // same columns and weights in the global sampler's fold branch too.
        // With nT >= 4 and SR >= 7 every range is non-empty (rhi > rlo).
        st.rlo[K_VPHI] = st.rlo[K_VTHETA] = st.rlo[K_CENTERED] = max(0, -js);
        st.rhi[K_VPHI] = st.rhi[K_CENTERED] = min(SR - 1, nT - 1 - js);
        st.rhi[K_VTHETA] = min(SR - 1, nT - 2 - js);
        // half phi coefficients of node rows j0-1 .. j0+R (u_phi / rho nodes,
        // u_theta nodes) and the geometric coefficient G of those rows, after
        // the staged planes and the advected u block
        const int W1 = W + 1, W2 = W + 2;
        const int uvN = (s.flags & SIM_GEOMETRIC) ? (R + 1) * W2 + (R + 2) * W1 + 2 * W1 : 0;
        float* cofRow[3];
        cofRow[0] = lds + 2 * plane + max(plane, uvN);
        cofRow[1] = cofRow[0] + (R + 2);
        cofRow[2] = cofRow[1] + (R + 2);
        // per wavefront: max |u_phi|, |u_theta| of its staged items and max
        // half phi coefficient of its node rows (float bits), after the
        // coefficients.  The kernel has no static LDS, so the dynamic block
        // starts at address 0 and LDS offsets fold into the instructions.
        unsigned(*wmax)[kAdvWaves] = reinterpret_cast<unsigned(*)[kAdvWaves]>(cofRow[2] + (R + 2));
        unsigned mc = 0u;
        if (tid < R + 2) {  // node rows j0-1 .. j0+R, one per thread
            const int j = j0 - 1 + tid;
            const bool in = j >= 0 && j < nT;
            const float cc = in ? 0.5f * cofPC[j] : 0.f, ct = in ? 0.5f * cofPT[j] : 0.f;
            cofRow[0][tid] = cc;
            cofRow[1][tid] = ct;
            cofRow[2][tid] = in ? s.rowf[RT_G * nT + j] : 0.f;
            const float cm = fmaxf(cc, ct);
            mc = __float_as_uint(cm != cm ? INFINITY : cm);
        }
        // running max of |u| as bits: non-negative floats (and NaN / Inf,
        // above every finite value) order as unsigned, so a non-finite value
        // forces the checked path (fmaxf would drop a NaN)
        unsigned mp = 0u, mt = 0u;
        auto absbits = [](float v) { return __float_as_uint(v) & 0x7fffffffu; };
        if (((nP | W) & 3) == 0) {
            // four columns per item: c0 - kEC, SC, nP and SW are multiples of
            // 4, so a float4 never straddles the phi seam and lands 16-byte
            // aligned in LDS
            // (two items per thread in flight: their loads all go out before
            // the first LDS write)
            const int SC4 = SC >> 2, n4 = SR * SC4;
            const unsigned mSC4 = s.magicW[4];  // e / SC4, e < 2^16
            for (int e0 = tid; e0 < n4; e0 += 2 * nthr) {
                float4 a[2], b[2], d[2];
                int l[2];
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const int e = e0 + q * nthr;
                    const int r = (int)__umulhi((unsigned)e, mSC4), c = (e - __mul24(r, SC4)) << 2;
                    int gi = c0 - kEC + c;
                    gi += (gi < 0) ? nP : 0;
                    gi -= (gi >= nP) ? nP : 0;
                    // rows outside the field (and items past the block) read as 0
                    // through the buffer range check; the former are never sampled
                    const int o = (js + r) * nP + gi;
                    l[q] = r * SW + c;
                    a[q] = ldf4(up, o);
                    b[q] = ldf4(ut, o);
                    d[q] = ldf4(rh, o);
                }
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    if (e0 + q * nthr >= n4) break;
                    *reinterpret_cast<float4*>(lds + l[q]) = a[q];
                    *reinterpret_cast<float4*>(lds + plane + l[q]) = b[q];
                    *reinterpret_cast<float4*>(lds + 2 * plane + l[q]) = d[q];
                    mp = max(mp, max(max(absbits(a[q].x), absbits(a[q].y)), max(absbits(a[q].z), absbits(a[q].w))));
                    mt = max(mt, max(max(absbits(b[q].x), absbits(b[q].y)), max(absbits(b[q].z), absbits(b[q].w))));
                }
            }
        } else {
            const unsigned mSC = s.magicW[3];  // e / SC, e < 2^16
            for (int e = tid; e < SR * SC; e += nthr) {
                const int r = (int)__umulhi((unsigned)e, mSC), c = e - __mul24(r, SC);
                int gi = c0 - kEC + c;
                gi += (gi < 0) ? nP : 0;
                gi -= (gi >= nP) ? nP : 0;
                const int o = (js + r) * nP + gi, l = r * SW + c;
                const float a = ldf(up, o), b = ldf(ut, o);
                lds[l] = a;
                lds[plane + l] = b;
                lds[2 * plane + l] = ldf(rh, o);
                mp = max(mp, absbits(a));
                mt = max(mt, absbits(b));
            }
        }
        mp = wave_max(mp);
        mt = wave_max(mt);
        mc = wave_max(mc);
        if ((tid & 63) == 0) {
            wmax[0][tid >> 6] = mp;
            wmax[1][tid >> 6] = mt;
            wmax[2][tid >> 6] = mc;
        }
        __syncthreads();
        // fast tile: every sample an ordinary bilinear read of the staged block
        bool fast;
        {
            unsigned m[3] = {0u, 0u, 0u};
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int w = 0; w < kAdvWaves; ++w) m[k] = max(m[k], wmax[k][w]);
            const float Up = __uint_as_float(m[0]), Ut = __uint_as_float(m[1]);
            const float hmax = __uint_as_float(m[2]);
            // a departure moves at most 2 h max|u| cells
            fast = (j0 >= kER + 1) && (j0 + R <= nT - kER - 1) && (Up * hmax <= 0.5f * kFastDPhi) &&
                   (Ut * hT <= 0.5f * kFastDTheta);
        }

        // one backtrace at node (c, j) of the tile (c from -1, j global): the
        // branch-free staged sampler on fast tiles, the checked one otherwise;
        // ofs = the node's staged offset (j - js) SW + c + kEC, which row walks
        // carry from one row to the next
        auto nodeAt = [&](auto fastTag, auto nodeTag, int c, int j, float hP, int ofs) -> float {
            constexpr int NK = decltype(nodeTag)::value;
            if constexpr (decltype(fastTag)::value) {
                return advect_node_fast<NK, NK, SW>(lds, plane, ofs, hP, hT);
            } else {
                const Rsrc src = (NK == K_VPHI) ? up : ((NK == K_VTHETA) ? ut : rh);
                return advect_node_st<NK, NK, SW>(st, up, ut, src, g, col(c), j, c + kEC, j - js, hP, hT);
            }
        };
        auto node = [&](auto fastTag, auto nodeTag, int c, int j, float hP) -> float {
            return nodeAt(fastTag, nodeTag, c, j, hP, __mul24(j - js, SW) + c + kEC);
        };
        using KPhi = std::integral_constant<int, K_VPHI>;
        using KTheta = std::integral_constant<int, K_VTHETA>;
        using KRho = std::integral_constant<int, K_CENTERED>;

        if (!(s.flags & SIM_GEOMETRIC)) {
            auto plain = [&](auto F) {
                bool bad = false;
                for (int idx = tid; idx < R * W; idx += nthr) {
                    const int r = (int)__umulhi((unsigned)idx, mW), c = idx - r * W;
                    const int j = j0 + r, o = j * nP + c0 + c;
                    oup[o] = node(F, KPhi{}, c, j, 0.5f * cofPC[j]);
                    if (j < nT - 1) out[o] = node(F, KTheta{}, c, j, 0.5f * cofPT[j]);
                    const float rr = node(F, KRho{}, c, j, 0.5f * cofPC[j]);
                    orh[o] = rr;
                    bad |= !isfinite(rr);
                }
                flag_bad(status, bad);
            };
            if (fast)
                plain(std::true_type{});
            else
                plain(std::false_type{});
            __syncthreads();
            continue;
        }

        const int uLo = j0, uHi = min(j0 + R, nT - 1);
        const int vLo = max(j0 - 1, 0), vHi = min(j0 + R, nT - 2);
        // the rho nodes go first; their staged plane then holds U, V, A0, AL
        float* U = lds + 2 * plane;      // (R+1) x W2: u_phi*, columns c0-1 .. c0+W
        float* V = U + (R + 1) * W2;     // (R+2) x W1: u_theta*, columns c0-1 .. c0+W-1
        float* A0 = V + (R + 2) * W1;    // W1: u_theta* of row 0 at the antipodal columns
        float* AL = A0 + W1;             // W1: u_theta* of row nT-2 at the antipodal columns
        float* X = lds;                  // R x W1: cubic roots, over the staged block once it is dead
        const bool north = (j0 == 0), south = (uHi == nT - 1);
        const int half = g.half;
        const int nU = (uHi - uLo + 1) * W2, nV = (vHi - vLo + 1) * W1, nR = R * W;
        // One backtrace per item.  Tiles 64 or 128 columns wide map each thread
        // to one tile column (cm) and a run of rows (rm0, rm0 + rmS, ...), with
        // no index split per item; the u blocks' edge columns (u_phi at -1 and
        // W, u_theta at -1) then take one short flat pass, from the last
        // wavefront down.  Other widths split a flat item index; those loops end
        // at wavefront granularity, with no duplicated tail items (occupancy
        // hides the LDS latency).  The mapping never changes a value.
        // A stride-SW launch holds widths up to SW - 2 kEC, and the column map
        // serves exactly that width: the row step rmS is then a constant, and
        // rm0 (a whole wavefront's tile row, kColW being a multiple of 64) is
        // uniform over the wavefront, so the walks' trip counts are scalar.
        constexpr int kColW = SW - 2 * kEC, rmS = nthr / kColW;
        static_assert(kColW % 64 == 0 && nthr % kColW == 0, "a column map row is whole wavefronts");
        const bool colMap = (W == kColW);
        const int cm = tid & (kColW - 1), rm0 = __builtin_amdgcn_readfirstlane(tid / kColW);
        const int colB = (c0 + cm) * 4;  // the lane's column in a row of the grid, in bytes
        const int nUr = uHi - uLo + 1, nVr = vHi - vLo + 1;
        auto advect_all = [&](auto F) {
            if (colMap) {
                const int dOfs = __mul24(rmS, SW);
                int ofs = __mul24(j0 + rm0 - js, SW) + cm + kEC;
                bool bad = false;
                for (int r = rm0; r < R; r += rmS, ofs += dOfs) {
                    const float a = nodeAt(F, KRho{}, cm, j0 + r, cofRow[0][r + 1], ofs);
                    stf(orhR, colB, (j0 + r) * nP * 4, a);
                    bad |= !isfinite(a);
                }
                flag_bad(status, bad);
                __syncthreads();  // the staged rho plane is dead from here on
                ofs = __mul24(uLo + rm0 - js, SW) + cm + kEC;
                float* u = U + __mul24(rm0, W2) + cm + 1;
                for (int r = rm0; r < nUr; r += rmS, ofs += dOfs, u += __mul24(rmS, W2))
                    *u = nodeAt(F, KPhi{}, cm, uLo + r, cofRow[0][uLo + r - j0 + 1], ofs);
                ofs = __mul24(vLo + rm0 - js, SW) + cm + kEC;
                float* v = V + __mul24(rm0, W1) + cm + 1;
                for (int r = rm0; r < nVr; r += rmS, ofs += dOfs, v += __mul24(rmS, W1))
                    *v = nodeAt(F, KTheta{}, cm, vLo + r, cofRow[1][vLo + r - j0 + 1], ofs);
                for (int idx = nthr - 1 - tid; idx < 2 * nUr + nVr; idx += nthr) {
                    if (idx < 2 * nUr) {
                        const int r = idx >> 1, c = (idx & 1) ? W : -1;
                        U[__mul24(r, W2) + c + 1] = node(F, KPhi{}, c, uLo + r, cofRow[0][uLo + r - j0 + 1]);
                    } else {
                        const int r = idx - 2 * nUr;
                        V[__mul24(r, W1)] = node(F, KTheta{}, -1, vLo + r, cofRow[1][vLo + r - j0 + 1]);
                    }
                }
                return;
            }
            bool bad = false;
            for (int idx = tid; idx < nR; idx += nthr) {
                const int r = (int)__umulhi((unsigned)idx, mW), c = idx - __mul24(r, W), j = j0 + r;
                const float a = node(F, KRho{}, c, j, cofRow[0][r + 1]);
                orh[j * nP + c0 + c] = a;
                bad |= !isfinite(a);
            }
            flag_bad(status, bad);
            __syncthreads();  // the staged rho plane is dead from here on
            for (int idx = tid; idx < nU; idx += nthr) {
                const int r = (int)__umulhi((unsigned)idx, mW2), c = idx - __mul24(r, W2) - 1, j = uLo + r;
                U[idx] = node(F, KPhi{}, c, j, cofRow[0][j - j0 + 1]);
            }
            for (int idx = tid; idx < nV; idx += nthr) {
                const int r = (int)__umulhi((unsigned)idx, mW1), c = idx - __mul24(r, W1) - 1, j = vLo + r;
                V[idx] = node(F, KTheta{}, c, j, cofRow[1][j - j0 + 1]);
            }
        };
        if (fast)
            advect_all(std::true_type{});
        else
            advect_all(std::false_type{});
        if (north || south) {
            for (int idx = tid; idx < 2 * W1; idx += nthr) {
                const bool s0 = idx < W1;
                if (s0 ? !north : !south) continue;
                const int c = s0 ? idx : idx - W1;
                int io = col(c - 1) + half;
                io -= (io >= nP) ? nP : 0;
                const int j = s0 ? 0 : nT - 2;
                const float a = advect_node<K_VTHETA, K_VTHETA>(up, ut, ut, g, io, j, 0.5f * cofPT[j], hT);
                (s0 ? A0 : AL)[c] = a;
            }
        }
        __syncthreads();

        // geometricFillKernel, one cell per item: the root x and v' = vbar + G x^2
        // of rows j0..uHi (the row below the tile only feeds the u_theta
        // restagger), both into the dead staged block
        const int nG = (uHi - j0 + 1) * W1;
        float* VP = X + (R + 1) * W1;  // (R+1) x W1: v'
        auto geo = [&](int r, int c) {
            const int idx = r * W1 + c, j = j0 + r;
            const float* Ur = U + (j - uLo) * W2;
            const float ub = 0.5f * (Ur[c] + Ur[c + 1]);
            float vb;
            if (j == 0)
                vb = 0.75f * V[(0 - vLo) * W1 + c] + 0.25f * A0[c];
            else if (j == nT - 1)
                vb = 0.75f * V[(nT - 2 - vLo) * W1 + c] + 0.25f * AL[c];
            else
                vb = 0.5f * (V[(j - 1 - vLo) * W1 + c] + V[(j - vLo) * W1 + c]);
            const float G = cofRow[2][r + 1];
            const float x = geo_root(G, vb, ub);
            X[idx] = x;
            VP[idx] = fmaf(G * x, x, vb);
        };
        if (colMap) {
            const int nGr = uHi - j0 + 1;
            for (int r = rm0; r < nGr; r += rmS) geo(opaque_s(r), cm);
            for (int r = nthr - 1 - tid; r < nGr; r += nthr) geo(r, W);  // column W
        } else {
            for (int idx = tid; idx < nG; idx += nthr) {
                const int r = (int)__umulhi((unsigned)idx, mW1);
                geo(r, idx - r * W1);
            }
        }
        __syncthreads();
        // assignPhiKernel (rows j0..j0+R-1) and assignThetaKernel (rows
        // j0..uHi-1), columns c0..c0+W-1
        auto assign = [&](int r, int c) {
            const int o = (j0 + r) * nP + c0 + c, x = r * W1 + c;
            oup[o] = 0.5f * (X[x] + X[x + 1]);
            if (j0 + r < uHi) out[o] = 0.5f * (VP[x + 1] + VP[x + W1 + 1]);
        };
        if (colMap) {
            for (int r = rm0; r < R; r += rmS) {
                const int rr = opaque_s(r), x = rr * W1 + cm, rowB = (j0 + rr) * nP * 4;
                stf(oupR, colB, rowB, 0.5f * (X[x] + X[x + 1]));
                if (j0 + rr < uHi) stf(outR, colB, rowB, 0.5f * (VP[x + 1] + VP[x + W1 + 1]));
            }
        } else {
            for (int idx = tid; idx < nR; idx += nthr) {
                const int r = (int)__umulhi((unsigned)idx, mW);
                assign(r, idx - r * W);
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Particles (advectionParticles, KaminoCore.cu:321-342): one forward-Euler
// step of each particle with the pre-step velocities, sampled at the particle
// by the u_phi and u_theta samplers; no phi motion where R sin(theta) <= 1e-7;
// then validateCoord.  The fold keeps the reference's arithmetic: FP32
// positions, the 2 pi and pi constants and the remainder in FP64.  A tile is
// a run of particles of one simulation; each thread moves one at a time.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void validate_coord(float& phi, float& theta) {
    constexpr double k2Pi = 6.28318530717958647692, kPiD = 3.14159265358979323846;
    theta = (float)((double)theta - (double)(int)floorf((float)((double)theta / k2Pi)) * k2Pi);
    if ((double)theta > kPiD) {
        theta = (float)(k2Pi - (double)theta);
        phi = (float)((double)phi + kPiD);
    }
    if (theta < 0.f) {
        theta = -theta;
        phi = (float)((double)phi + kPiD);
    }
    phi = (float)((double)phi - (double)(int)floorf((float)((double)phi / k2Pi)) * k2Pi);
}

__global__ __launch_bounds__(256) void k_particles(const SimDesc* __restrict__ sims,
                                                   const Tile* __restrict__ tiles, int ntiles,
                                                   int parity) {
    for (int t = xcd_tile(ntiles); t < ntiles; t = ntiles) {
        const Tile tile = tiles[t];
        const SimDesc& s = sims[tile.sim];
        const Grid g{s.nTheta, s.nPhi, s.nPhi >> 1, s.invNP};
        const int N = g.nT * g.nP;
        const Rsrc up = make_rsrc(s.state[parity][F_UPHI], N);
        const Rsrc ut = make_rsrc(s.state[parity][F_UTHETA], N - g.nP);
        const float2* __restrict__ in = reinterpret_cast<const float2*>(s.part[parity]);
        float2* __restrict__ outp = reinterpret_cast<float2*>(s.part[parity ^ 1]);
        const float invH = s.invH, R = s.radius, dt = s.dt;
        const float cofTheta = dt / R;
        const int end = tile.row0 + tile.rows;
        for (int k = tile.row0 + (int)threadIdx.x; k < end; k += (int)blockDim.x) {
            const float2 p = in[k];
            // offset-shifted coordinates in cells: u_phi nodes sit at
            // ((i - 1/2) h, (j + 1/2) h), u_theta nodes at (i h, (j + 1) h)
            const float x = p.x * invH, y = p.y * invH;
            const float uP = sample<K_VPHI>(up, g, 0, x + 0.5f, 0, y - 0.5f);
            const float uT = sample<K_VTHETA>(ut, g, 0, x, 0, y - 1.f);
            const float latRadius = R * sinf(p.y);
            float th = p.y + uT * cofTheta;
            float ph = p.x;
            if (latRadius > 1e-7f) ph = p.x + uP * (dt / latRadius);
            validate_coord(ph, th);
            outp[k] = make_float2(ph, th);
            flag_bad(s.status, !isfinite(ph) || !isfinite(th));
        }
    }
}

// ---------------------------------------------------------------------------
// Row FFT, one wavefront per row: Stockham autosort passes (radix 2/3/4/5)
// ping-pong between the row and the wavefront's own scratch row in LDS.  LDS
// operations of one wavefront complete in issue order, so consecutive passes
// are separated by a wavefront-scope fence only, never a workgroup barrier.
// DIR = -1 for exp(-i ...), +1 for exp(+i ...); unnormalised.
// ---------------------------------------------------------------------------
// Complex arithmetic on float2 (FP32 transforms) and double2 (the FP64
// pressure transform); R is the component type.
template <typename C> struct CplxT;
template <> struct CplxT<float2> { typedef float R; };
template <> struct CplxT<double2> { typedef double R; };
__device__ __forceinline__ float2 mkc(float x, float y) { return make_float2(x, y); }
__device__ __forceinline__ double2 mkc(double x, double y) { return make_double2(x, y); }
template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
    return mkc(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C>
__device__ __forceinline__ C cadd(C a, C b) { return mkc(a.x + b.x, a.y + b.y); }
template <typename C>
__device__ __forceinline__ C csub(C a, C b) { return mkc(a.x - b.x, a.y - b.y); }
template <typename C>
__device__ __forceinline__ C cscale(C a, typename CplxT<C>::R s) { return mkc(a.x * s, a.y * s); }
// DIR * i * a
template <int DIR, typename C>
__device__ __forceinline__ C cmuli(C a) {
    return DIR > 0 ? mkc(-a.y, a.x) : mkc(a.y, -a.x);
}

template <int RADIX, int DIR, typename C>
__device__ __forceinline__ void butterfly(C* x) {
    typedef typename CplxT<C>::R R;
    if constexpr (RADIX == 2) {
        const C a = x[0], b = x[1];
        x[0] = cadd(a, b);
        x[1] = csub(a, b);
    } else if constexpr (RADIX == 3) {
        const C t = cadd(x[1], x[2]), u = csub(x[1], x[2]);
        const C m = csub(x[0], cscale(t, R(0.5)));
        const C n = cmuli<DIR>(cscale(u, R(0.86602540378443864676)));
        x[0] = cadd(x[0], t);
        x[1] = cadd(m, n);
        x[2] = csub(m, n);
    } else if constexpr (RADIX == 4) {
        const C a = cadd(x[0], x[2]), b = csub(x[0], x[2]);
        const C c = cadd(x[1], x[3]), d = cmuli<DIR>(csub(x[1], x[3]));
        x[0] = cadd(a, c);
        x[2] = csub(a, c);
        x[1] = cadd(b, d);
        x[3] = csub(b, d);
    } else {  // 5
        const R c1 = R(0.30901699437494742410), c2 = R(-0.80901699437494742410);
        const R s1 = R(0.95105651629515357212), s2 = R(0.58778525229247312917);
        const C t1 = cadd(x[1], x[4]), t2 = cadd(x[2], x[3]);
        const C t3 = csub(x[1], x[4]), t4 = csub(x[2], x[3]);
        const C a1 = cadd(x[0], cadd(cscale(t1, c1), cscale(t2, c2)));
        const C a2 = cadd(x[0], cadd(cscale(t1, c2), cscale(t2, c1)));
        const C b1 = cmuli<DIR>(cadd(cscale(t3, s1), cscale(t4, s2)));
        const C b2 = cmuli<DIR>(csub(cscale(t3, s2), cscale(t4, s1)));
        x[0] = cadd(x[0], cadd(t1, t2));
        x[1] = cadd(a1, b1);
        x[4] = csub(a1, b1);
        x[2] = cadd(a2, b2);
        x[3] = csub(a2, b2);
    }
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Orders the LDS traffic of one lane group of the row FFT (SimDesc::fftShift:
// a power of two from 8 up to the whole 256-thread block): a wavefront fence
// for groups within one wavefront, a block barrier for a block-wide group
// (long rows).  gs is uniform over the block.
__device__ __forceinline__ void group_sync(int gs) {
    if (gs > 64)
        __syncthreads();
    else
        wave_sync();
}

template <int RADIX, int DIR, bool HALF, typename C>
__device__ __forceinline__ void wave_pass(const C* __restrict__ src, C* __restrict__ dst,
                                          int M, int Ls, int lsShift, unsigned lsMagic,
                                          const C* __restrict__ tw, int twScale, int lane,
                                          int gs) {
    const int nb = M / RADIX;
    for (int b = lane; b < nb; b += gs) {
        int k, j;
        if (lsShift >= 0) {
            k = b >> lsShift;
            j = b & (Ls - 1);
        } else {
            k = (int)__umulhi((unsigned)b, lsMagic);  // b / Ls, exact for b < 2^16
            j = b - k * Ls;
        }
        const C* in = src + j + k * Ls;
        C x[RADIX];
#pragma unroll
        for (int q = 0; q < RADIX; ++q) {
            C v = in[q * nb];
            if (q > 0) {
                // a half table holds n < M only: w(n + M) = -w(n)
                const int t = q * j * twScale;
                C w;
                if constexpr (HALF)
                    w = t < M ? tw[t] : mkc(-tw[t - M].x, -tw[t - M].y);
                else
                    w = tw[t];
                if (DIR > 0) w.y = -w.y;
                v = cmul(v, w);
            }
            x[q] = v;
        }
        butterfly<RADIX, DIR>(x);
        C* out = dst + j + k * Ls * RADIX;
#pragma unroll
        for (int m = 0; m < RADIX; ++m) out[m * Ls] = x[m];
    }
}

// Transforms the M complex values at `row` (LDS) in place; `scratch` is the
// calling lane group's own M-entry LDS row, `tw` the nPhi twiddles
// exp(-2 pi i n / nPhi) staged in LDS (with HALF only the first M of them,
// the others being w(n + M) = -w(n)), `lane` the lane within the group of gs
// lanes.  Every lane of the wavefront (of the block, for a block-wide group)
// calls it (rowless groups pass active = false) so that the fences and
// barriers stay uniform.
template <int DIR, bool HALF, typename C>
__device__ void wave_fft(C* row, C* scratch, const C* tw, const SimDesc& s, int lane, int gs,
                         bool active) {
    const int M = s.M;
    C* src = row;
    C* dst = scratch;
    int Ls = 1;
    for (int p = 0; p < s.nfac; ++p) {
        const int r = s.fac[p];
        const int twScale = s.twStep[p];  // nPhi / (Ls r)
        const unsigned lsMagic = s.lsMagic[p];
        const int lsShift = ((Ls & (Ls - 1)) == 0) ? __builtin_ctz(Ls) : -1;
        if (active) {
            switch (r) {
                case 2: wave_pass<2, DIR, HALF>(src, dst, M, Ls, lsShift, lsMagic, tw, twScale, lane, gs); break;
                case 3: wave_pass<3, DIR, HALF>(src, dst, M, Ls, lsShift, lsMagic, tw, twScale, lane, gs); break;
                case 4: wave_pass<4, DIR, HALF>(src, dst, M, Ls, lsShift, lsMagic, tw, twScale, lane, gs); break;
                default: wave_pass<5, DIR, HALF>(src, dst, M, Ls, lsShift, lsMagic, tw, twScale, lane, gs); break;
            }
        }
        group_sync(gs);
        C* tmp = src;
        src = dst;
        dst = tmp;
        Ls *= r;
    }
    if (src != row) {
        if (active)
            for (int i = lane; i < M; i += gs) row[i] = src[i];
        group_sync(gs);
    }
}

// Memory phases of the FFT kernels issue a batch of loads per thread before the
// first use, so that a tile's HBM latency is paid once per batch rather than
// once per item.
constexpr int kDivBatch = 4;   // row points per lane in k_div_fft
constexpr int kSpecBatch = 8;  // spectrum entries per thread in k_ifft_grad
constexpr int kGradBatch = 4;  // column pairs per thread in k_ifft_grad

// Workgroup size of the FFT kernels (their launches use exactly this).
constexpr int kFftBlockShift = 8, kFftBlock = 1 << kFftBlockShift;

// Stages the nPhi twiddles of s into LDS, two per float4 (nPhi is even), or
// with `half` the first M = nPhi / 2 of them (M rounded up to even: the table
// holds nPhi entries).
__device__ __forceinline__ int tw_slots(int M) { return (M + 1) & ~1; }
__device__ __forceinline__ void stage_twiddles(float2* twl, const SimDesc& s, int tid, int nthr,
                                               bool half) {
    const int n4 = (half ? tw_slots(s.M) : s.nPhi) >> 1;
    const Rsrc tr = make_rsrc(reinterpret_cast<const float*>(s.twN), 2 * s.nPhi);
    float4* dst = reinterpret_cast<float4*>(twl);
    for (int i0 = tid; i0 < n4; i0 += 4 * nthr) {
        float4 v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = ldf4(tr, 4 * (i0 + q * nthr));
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (i0 + q * nthr < n4) dst[i0 + q * nthr] = v[q];
    }
}

// ---------------------------------------------------------------------------
// k_div_fft: divergence (fillDivergenceKernel, KaminoCore.cu:587-638) of R
// rows, packed two reals per complex (z_k = d_2k + i d_2k+1), FFT of length
// M, split into the half spectrum X_n = E_n + exp(-i pi n / M) O_n and scaled
// by 1/nPhi.  Mode n of row j goes to spec[spec_index(n, j)] (row blocks of
// four, mode-major within a block); slot 0 holds (d-hat_0, d-hat_Nyquist).
// This is the reference's cuFFT CUFFT_INVERSE call plus shiftFKernel
// (:640-656): d-hat_n = (1/nPhi) sum_i d_i e^{-i n i h}.
// HALF: only the first M twiddles are staged (w(n + M) = -w(n)), and the
// launch (k_div_fft8) asks for eight waves per SIMD; else all nPhi twiddles
// (k_div_fft).  The two give the same values.
// ---------------------------------------------------------------------------
template <bool HALF>
__device__ __forceinline__ void div_fft_tiles(const SimDesc* __restrict__ sims, const Tile* __restrict__ tiles,
                                              int ntiles, int parity) {
    extern __shared__ __attribute__((aligned(16))) float2 zlds[];
    const int tid = threadIdx.x, nthr = blockDim.x;
    for (int t = xcd_tile(ntiles); t < ntiles; t = ntiles) {
        const Tile tile = tiles[t];
        const SimDesc& s = sims[tile.sim];
        const int nT = s.nTheta, nP = s.nPhi, M = s.M;
        const int j0 = tile.row0, R = tile.rows;
        const int gsh = s.fftShift, gs = 1 << gsh;  // lanes per row (a power of two)
        // kFftBlock threads: kFftBlock / gs lane groups (a power of two)
        const int ngsh = kFftBlockShift - gsh, grp = tid >> gsh, ngrp = 1 << ngsh, lane = tid & (gs - 1);
        const float* __restrict__ up = s.state[parity][F_UPHI];
        const float* __restrict__ ut = s.state[parity][F_UTHETA];
        const float* __restrict__ divS = s.rowf + RT_DIVS * nT;
        const float* __restrict__ divN = s.rowf + RT_DIVN * nT;
        const float* __restrict__ invHs = s.rowf + RT_INVHS * nT;
        float2* twl = zlds;                                        // M or nPhi twiddles
        float2* zbuf = zlds + (HALF ? tw_slots(M) : nP);           // R rows of M
        float2* scratch = zbuf + (size_t)(R + grp) * M;            // this lane group's row
        const Rsrc upR = make_rsrc(up, nT * nP), utR = make_rsrc(ut, (nT - 1) * nP);
        const int nPass = (R + ngrp - 1) >> ngsh;                  // uniform trip count
        for (int it = 0; it < nPass; ++it) {
            const int r = it * ngrp + grp;
            const bool active = r < R;
            const int j = j0 + (active ? r : 0);
            const float fs = divS[j], fn = divN[j], fi = invHs[j];
            float2* row = zbuf + (size_t)(active ? r : 0) * M;
            // kDivBatch points per lane at a time, all their loads issued
            // before the first use; rows beyond the field (j - 1 < 0, j = nT - 1
            // for u_theta) read as 0 through the buffer range check
            const int on = (j - 1) * nP, oc = j * nP;
            for (int kb = 0; kb < M; kb += kDivBatch * gs) {
                float2 uw[kDivBatch], vn[kDivBatch], vs[kDivBatch];
                float ue1[kDivBatch];
#pragma unroll
                for (int q = 0; q < kDivBatch; ++q) {
                    const int i0 = 2 * (kb + q * gs + lane), i2 = (i0 + 2 == nP) ? 0 : i0 + 2;
                    uw[q] = ldf2(upR, oc + i0);
                    ue1[q] = ldf(upR, oc + i2);
                    vn[q] = ldf2(utR, j > 0 ? on + i0 : -1);
                    vs[q] = ldf2(utR, oc + i0);
                }
                // the twiddles are staged while the first rows are in flight
                if (it == 0 && kb == 0) stage_twiddles(twl, s, tid, nthr, HALF);
#pragma unroll
                for (int q = 0; q < kDivBatch; ++q) {
                    const int k = kb + q * gs + lane;
                    const float d0 = (fs * vs[q].x - fn * vn[q].x) + fi * (uw[q].y - uw[q].x);
                    const float d1 = (fs * vs[q].y - fn * vn[q].y) + fi * (ue1[q] - uw[q].y);
                    if (active && k < M) row[k] = make_float2(d0, d1);
                }
            }
            if (it == 0) __syncthreads();  // twiddles visible to every lane group
            group_sync(gs);
            wave_fft<-1, HALF>(row, scratch, twl, s, lane, gs, active);
        }
        __syncthreads();
        float2* __restrict__ spec = reinterpret_cast<float2*>(s.spec);
        const float2* tw = twl;
        const float scale = 1.0f / (float)nP;
        // ceil(2^32 / R) in its 32-bit form (R = 1 splits nothing)
        const unsigned magicR = (R == 1) ? 0u : 0xffffffffu / (unsigned)R + 1u;
        for (int idx = tid; idx < R * M; idx += nthr) {
            const int n = (R == 1) ? idx : (int)__umulhi((unsigned)idx, magicR), r = idx - n * R, j = j0 + r;
            const float2 Zn = zbuf[(size_t)r * M + n];
            float2 out;
            if (n == 0) {
                out = make_float2((Zn.x + Zn.y) * scale, (Zn.x - Zn.y) * scale);
            } else {
                const float2 Zm = zbuf[(size_t)r * M + (M - n)];
                const float2 E = make_float2(0.5f * (Zn.x + Zm.x), 0.5f * (Zn.y - Zm.y));
                const float2 O = make_float2(0.5f * (Zn.y + Zm.y), -0.5f * (Zn.x - Zm.x));
                const float2 X = cadd(E, cmul(tw[n], O));
                out = cscale(X, scale);
            }
            spec[spec_index(n, j, M)] = out;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kFftBlock) void k_div_fft(const SimDesc* __restrict__ sims,
                                                 const Tile* __restrict__ tiles, int ntiles, int parity) {
    div_fft_tiles<false>(sims, tiles, ntiles, parity);
}
__global__ __launch_bounds__(kFftBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_div_fft8(
    const SimDesc* __restrict__ sims, const Tile* __restrict__ tiles, int ntiles, int parity) {
    div_fft_tiles<true>(sims, tiles, ntiles, parity);
}

// ---------------------------------------------------------------------------
// k_tridiag: T_n p-hat = d-hat for every Fourier mode n (crKernel, tdm.cu:3-96,
// with the rows of precomputeABCKernel, KaminoSolver.cu:117-163: sub-diagonal
// a = 1/h^2 - cot/(2h), diagonal b = -2/h^2 - n^2/sin^2, super-diagonal
// c = 1/h^2 + cot/(2h), pole rows folded b += a / b += c), in FP32 with
// subtraction-free pivots or in FP64 (the FP64 pressure).
//
// L lanes share one mode; lane l owns K consecutive rows and runs Thomas on
// them in registers, keeping each row in the form
//     alpha_i x_{s-1} + x_i + gamma_i x_{i+1} = delta_i
// (s = the lane's first row).  With P/Q/S such that x_s = P + Q x_{s-1} + S x_e,
// the lanes' last unknowns X_l = x_e satisfy one tridiagonal system across the
// L lanes, solved by parallel cyclic reduction with cross-lane shuffles; every
// lane then back-substitutes its own rows.  Mode 0 is the identity in the
// reference and its result is discarded (cacheZeroComponents :672-680 and
// shiftUKernel :682-703 subtract it); its slot carries the Nyquist mode.
// ---------------------------------------------------------------------------
// 1/x to FP64 precision: the hardware estimate refined by one Newton step
// (which squares its relative error), shorter than an IEEE division sequence;
// the FP64 residual of the solve measures ~1e-16.
__device__ __forceinline__ double rcp64(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);  // one Newton step doubles the hardware estimate's bits
}

// 1/x to FP32 precision (the hardware estimate is within 1 ulp).
__device__ __forceinline__ float rcpT(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcpT(double x) { return rcp64(x); }
__device__ __forceinline__ float absT(float x) { return fabsf(x); }
__device__ __forceinline__ double absT(double x) { return fabs(x); }

template <int L, typename T>
__device__ __forceinline__ T shfl_grp(T v, int srcLaneInGroup, int laneBase) {
    return __shfl(v, laneBase + srcLaneInGroup, 64);
}

// The FP32 elimination forms every pivot without subtraction.  Each row has
// a, c >= 0 and the diagonal excess m = -b - a - c = n^2 / sin^2 > 0 (pole
// folds included), so the low modes' systems are nearly singular
// (m / (a + c) ~ 5e-6 at mode 1, n_theta = 1024): the textbook pivot
// b_k - a_k ga_{k-1} cancels, and FP32 Thomas loses up to 1e-4 relative
// there.  Instead, with D_k = -(pivot) > 0,
//     D_k = c_k + sig_k,  sig_k = m_k + a_k tau_{k-1},  tau = sig / D = 1 - |ga|,
// and the excess of each normalised row, ep_k = 1 - |al_k| - |ga_k|
// = (m_k + a_k ep_{k-1}) / D_k, are sums of non-negative terms.  The lane's
// reduced row keeps its excess E = B - |A| - |C| the same way (from ep at
// its last row and, for the next lane, R = 1 - Q - S = sum of ep_k times the
// product of |ga| before k), and parallel cyclic reduction updates it as
// E' = E + |k1| E_minus + |k2| E_plus, with B' = |A'| + |C'| + E'.  Every
// pivot then carries FP32 relative accuracy (the usual accurate LDU of a
// diagonally dominant M-matrix); CPU emulation of this exact sequence gives
// <= 8.4e-7 relative error at mode 1 up to n_theta = 2048 (6e-8, FP32
// rounding, at the high modes).  The FP64 elimination keeps the textbook
// pivots (exact to ~1e-13 there).

// One tile of the tridiagonal solve: G = 256 / L modes of one simulation,
// L lanes per mode; T is the arithmetic of the elimination (the coefficients
// are formed in FP64 and rounded to T).  The block's G mode columns and the FP64 row terms pass
// through LDS (zs: G x KMAX x (L+1) float2; cs, ss: KMAX x (L+1) doubles):
// global reads and writes stay contiguous, and lane l finds its row l K + k
// at [k][l] (a padded row per k keeps the staging writes spread over the
// banks).  RESID additionally writes, per mode, ||T x - d|| / ||d|| of the
// solution x in T, formed in FP64 (to SimDesc::resid).  A
// simulation with the FP64 pressure (SIM_P64) keeps x in FP64: each lane
// stores its rows straight to the double2 spectrum over the finished state
// copy (parity ^ 1), K consecutive entries per lane, and skips the FP32
// write-back.
constexpr int kTriBatch = 4;  // float4 spectrum loads per thread in flight while staging

template <int L, int KMAX, bool RESID, typename T>
__device__ __forceinline__ void tri_tile(const SimDesc* __restrict__ sims, const Tile& tile,
                                         int parity, float2* zsp, double* csp, double* ssp,
                                         float* fsp) {
    constexpr int G = 256 / L;  // modes per block
    (void)G;
    // SF: the subtraction-free pivots of the FP32 elimination, from FP32 row
    // tables; the FP64 row terms are staged for the FP64 elimination and for
    // the residual report
    constexpr bool SF = std::is_same<T, float>::value;
    constexpr bool kStage64 = !SF || RESID;
    constexpr int NR = KMAX * (L + 1);
    auto zs = [&](int gg, int k, int ll) -> float2& { return zsp[(gg * KMAX + k) * (L + 1) + ll]; };
    auto cs = [&](int k, int ll) -> double& { return csp[k * (L + 1) + ll]; };
    auto ss = [&](int k, int ll) -> double& { return ssp[k * (L + 1) + ll]; };
    // FP32 rows: sub-diagonal, super-diagonal (pole folds included), 1/sin^2
    auto fa = [&](int k, int ll) -> float& { return fsp[k * (L + 1) + ll]; };
    auto fc = [&](int k, int ll) -> float& { return fsp[NR + k * (L + 1) + ll]; };
    auto fs = [&](int k, int ll) -> float& { return fsp[2 * NR + k * (L + 1) + ll]; };
    const int tid = threadIdx.x;
    const int l = tid % L;               // lane within the mode group
    const int g = tid / L;               // mode within the block
    const int wl = threadIdx.x & 63;     // lane within the wavefront
    const int base = wl - l;             // first wavefront lane of this group
    {
        const SimDesc& s = sims[tile.sim];
        const int nT = s.nTheta;
        const bool p64 = (s.flags & SIM_P64) != 0;
        double2* __restrict__ spec64 = reinterpret_cast<double2*>(s.state[parity ^ 1][0]);
        const bool active = g < tile.rows;
        const int n = tile.row0 + (active ? g : 0);
        const int K = (nT + L - 1) / L;
        // i / K and e / nT as multiply-high (exact for dividends below 2^16):
        // ceil(2^32 / K) in its 32-bit form (K = 1, whose multiplier 2^32 does
        // not fit, is taken apart), and e / nT = 2e / nPhi
        const unsigned magicK = (K == 1) ? 0u : 0xffffffffu / (unsigned)K + 1u;
        auto split = [&](int i, int& li, int& ki) {
            li = (K == 1) ? i : (int)__umulhi((unsigned)i, magicK);
            ki = i - li * K;
        };
        const int r0 = l * K;
        int cnt = nT - r0;
        cnt = cnt < 0 ? 0 : (cnt > K ? K : cnt);
        if (!active) cnt = 0;
        // staging: a thread's loads of a batch all go out before its first
        // LDS write; reads past the tables or the spectrum return 0 and are
        // not kept
        if constexpr (kStage64) {
            const Rsrc rowR = make_rsrc(reinterpret_cast<const float*>(s.rowd), 2 * RD_COUNT * nT);
            for (int i0 = tid; i0 < nT; i0 += 4 * 256) {
                double c[4], v[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    c[q] = ldd(rowR, RD_COT2H * nT + i0 + q * 256);
                    v[q] = ldd(rowR, RD_INVSIN2 * nT + i0 + q * 256);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int i = i0 + q * 256;
                    if (i >= nT) break;
                    int li, ki;
                    split(i, li, ki);
                    cs(ki, li) = c[q];
                    ss(ki, li) = v[q];
                }
            }
        }
        if constexpr (SF) {
            const Rsrc rowF = make_rsrc(s.rowf, RT_COUNT * nT);
            for (int i0 = tid; i0 < nT; i0 += 4 * 256) {
                float a[4], c[4], v[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    a[q] = ldf(rowF, RT_TRIA * nT + i0 + q * 256);
                    c[q] = ldf(rowF, RT_TRIC * nT + i0 + q * 256);
                    v[q] = ldf(rowF, RT_TRIS * nT + i0 + q * 256);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int i = i0 + q * 256;
                    if (i >= nT) break;
                    int li, ki;
                    split(i, li, ki);
                    fa(ki, li) = a[q];
                    fc(ki, li) = c[q];
                    fs(ki, li) = v[q];
                }
            }
        }
        // The tile's entries in spectrum order (spec_index): block by block of
        // kSpecRows rows, in each the tile's modes as runs of kSpecRows rows.
        // Pair p of the tile (two entries, one float4) is mode row0 + g, rows
        // i and i + 1 of block b:  b = p / (kHalf rows), w = p - kHalf rows b,
        // g = w / kHalf, i = kSpecRows b + 2 (w % kHalf), at float2 offset
        // kSpecRows (b M + row0) + 2 w (even, so float4-aligned).  Rows past
        // nTheta are padding and are skipped.
        const Rsrc specR = make_rsrc(s.spec, 2 * s.M * spec_rows(nT));
        constexpr int kHalf = kSpecRows / 2;  // pairs per mode in a block
        const int pairN = kHalf * tile.rows, nPairs = tile.rows * (spec_rows(nT) / 2), Ms = s.M;
        const unsigned magicB = 0xffffffffu / (unsigned)pairN + 1u;  // p / pairN, p < 2^16, pairN >= 2
        auto pair = [&](int p, int& g, int& i) -> int {
            const int b = (int)__umulhi((unsigned)p, magicB);
            const int w = p - b * pairN;
            g = w / kHalf;
            i = kSpecRows * b + 2 * (w % kHalf);
            return kSpecRows * b * Ms + kSpecRows * tile.row0 + 2 * w;
        };
        // (li, ki) of row i + 1 from those of row i
        auto step_row = [&](int& li, int& ki) {
            if (++ki == K) {
                ki = 0;
                ++li;
            }
        };
        for (int p0 = tid; p0 < nPairs; p0 += 256 * kTriBatch) {
            float4 v[kTriBatch];
            int gq[kTriBatch], iq[kTriBatch];
#pragma unroll
            for (int q = 0; q < kTriBatch; ++q) {
                const int p = p0 + 256 * q;
                const int a = pair(p, gq[q], iq[q]);
                v[q] = ldf4(specR, p < nPairs ? 2 * a : -1);  // pairs past the tile read as 0, unused
            }
#pragma unroll
            for (int q = 0; q < kTriBatch; ++q) {
                if (p0 + 256 * q < nPairs && iq[q] < nT) {
                    int li, ki;
                    split(iq[q], li, ki);
                    zs(gq[q], ki, li) = make_float2(v[q].x, v[q].y);
                    if (iq[q] + 1 < nT) {
                        step_row(li, ki);
                        zs(gq[q], ki, li) = make_float2(v[q].z, v[q].w);
                    }
                }
            }
        }
        __syncthreads();
        const double h2 = s.invH2;
        const double nn = (n == 0) ? (double)s.M * (double)s.M : (double)n * (double)n;
        const float nnf = (float)nn;  // exact: n^2 <= 2^22
        // row i = r0 + k of this lane, a x_{i-1} + b x_i + c x_{i+1} = f, in FP64
        // (the pole rows folded, KaminoSolver.cu:117-163)
        auto coef = [&](int k, double& a, double& b, double& c) {
            const int i = r0 + k;
            const double ct = cs(k, l);
            a = h2 - ct;
            c = h2 + ct;
            b = -2.0 * h2 - nn * ss(k, l);
            if (i == 0) {
                b += a;
                a = 0.0;
            }
            if (i == nT - 1) {
                b += c;
                c = 0.0;
            }
        };
        // Solves the group's system for the staged right-hand side (mode 0
        // dropped: its slot's real part) into (xr, xi), in T: Thomas within
        // the lane, then parallel cyclic reduction of the lanes' last
        // unknowns.
        // The solve, from the lanes' Thomas passes to the solution's staging
        // for the write-back, for a row count per lane `cntArg`; a block whose
        // every lane holds KMAX rows (all its modes active, nTheta = L KMAX)
        // takes it with the count as a constant, so every row guard folds away.
        T xr[KMAX], xi[KMAX];
        auto solve = [&](auto cntArg) {
        const int cnt = cntArg;
        {
            T al[KMAX], ga[KMAX], dr[KMAX], di[KMAX];
#pragma unroll
            for (int k = 0; k < KMAX; ++k) {
                al[k] = T(0);
                ga[k] = T(0);
                dr[k] = T(0);
                di[k] = T(0);
            }
            // SF running terms: tau_{k-1}, ep_{k-1}, the lane's R = sum over
            // k < cnt - 1 of ep_k |ga_0 ... ga_{k-1}| with its running product,
            // and ep at the lane's last row
            T tau = T(1), epPrev = T(0), R = T(0), prodGa = T(1), eE = T(0);
            // forward elimination within the lane's rows
#pragma unroll
            for (int k = 0; k < KMAX; ++k) {
                if (k < cnt) {
                    const float2 f = zs(g, k, l);
                    const T fr = (n == 0) ? T(0) : (T)f.x, fi = (T)f.y;
                    if constexpr (SF) {
                        const T a = fa(k, l), c = fc(k, l), m = nnf * fs(k, l);
                        const T sig = m + a * tau;
                        const T inv = rcpT(c + sig);
                        const T ep = (m + a * epPrev) * inv;
                        tau = sig * inv;
                        epPrev = ep;
                        al[k] = a * (k == 0 ? T(-1) : al[k > 0 ? k - 1 : 0]) * inv;
                        ga[k] = -c * inv;
                        dr[k] = (a * (k == 0 ? T(0) : dr[k > 0 ? k - 1 : 0]) - fr) * inv;
                        di[k] = (a * (k == 0 ? T(0) : di[k > 0 ? k - 1 : 0]) - fi) * inv;
                        if (k < cnt - 1) {
                            R += ep * prodGa;
                            prodGa *= c * inv;
                        } else {
                            eE = ep;
                        }
                    } else {
                        double a64, b64, c64;
                        coef(k, a64, b64, c64);
                        const T a = (T)a64, b = (T)b64, c = (T)c64;
                        if (k == 0) {
                            const T inv = rcpT(b);
                            al[k] = a * inv;
                            ga[k] = c * inv;
                            dr[k] = fr * inv;
                            di[k] = fi * inv;
                        } else {
                            const T inv = rcpT(b - a * ga[k - 1]);
                            al[k] = -a * al[k - 1] * inv;
                            ga[k] = c * inv;
                            dr[k] = (fr - a * dr[k - 1]) * inv;
                            di[k] = (fi - a * di[k - 1]) * inv;
                        }
                    }
                }
            }
            // x_s = P + Q x_{s-1} + S x_e, by backward recurrence from the last row
            T Pr = T(0), Pi = T(0), Q = T(0), S = T(1);
#pragma unroll
            for (int k = KMAX - 2; k >= 0; --k) {
                if (k < cnt - 1) {
                    Pr = dr[k] - ga[k] * Pr;
                    Pi = di[k] - ga[k] * Pi;
                    if constexpr (!SF) Q = -al[k] - ga[k] * Q;
                    S = -ga[k] * S;
                }
            }
            // reduced equation of this lane: A X_{l-1} + B X_l + C X_{l+1} = D,
            // with SF's excess E = B - |A| - |C|
            const int src = (l + 1 < L) ? l + 1 : l;
            const T nPr = shfl_grp<L>(Pr, src, base), nPi = shfl_grp<L>(Pi, src, base);
            const T nQ = SF ? T(0) : shfl_grp<L>(Q, src, base), nS = shfl_grp<L>(S, src, base);
            const T nR = SF ? shfl_grp<L>(R, src, base) : T(0);
            T A = T(0), B = T(1), C = T(0), Dr = T(0), Di = T(0), E = T(1);
            if (cnt > 0) {
                T gE = T(0), aE = T(0), drE = T(0), diE = T(0);
#pragma unroll
                for (int k = 0; k < KMAX; ++k)
                    if (k == cnt - 1) {
                        gE = ga[k];
                        aE = al[k];
                        drE = dr[k];
                        diE = di[k];
                    }
                const bool hasNext = (l + 1 < L);
                A = aE;
                C = hasNext ? gE * nS : T(0);
                Dr = drE - (hasNext ? gE * nPr : T(0));
                Di = diE - (hasNext ? gE * nPi : T(0));
                if constexpr (SF) {
                    E = eE + absT(gE) * (hasNext ? nR : T(1));
                    B = absT(A) + absT(C) + E;
                } else {
                    B = T(1) + (hasNext ? gE * nQ : T(0));
                }
            }
            // parallel cyclic reduction across the L lanes of the group
#pragma unroll
            for (int st = 1; st < L; st <<= 1) {
                const int lm = l - st, lp = l + st;
                const int sm = lm >= 0 ? lm : l, sp = lp < L ? lp : l;
                T Am = shfl_grp<L>(A, sm, base), Bm = shfl_grp<L>(B, sm, base);
                T Dmr = shfl_grp<L>(Dr, sm, base), Dmi = shfl_grp<L>(Di, sm, base);
                T Bp = shfl_grp<L>(B, sp, base), Cp = shfl_grp<L>(C, sp, base);
                T Dpr = shfl_grp<L>(Dr, sp, base), Dpi = shfl_grp<L>(Di, sp, base);
                // SF: the neighbours' excesses; else their other couplings
                T Xm = SF ? shfl_grp<L>(E, sm, base) : shfl_grp<L>(C, sm, base);
                T Xp = SF ? shfl_grp<L>(E, sp, base) : shfl_grp<L>(A, sp, base);
                if (lm < 0) {
                    Am = T(0);
                    Bm = T(1);
                    Dmr = T(0);
                    Dmi = T(0);
                    Xm = SF ? T(1) : T(0);
                }
                if (lp >= L) {
                    Bp = T(1);
                    Cp = T(0);
                    Dpr = T(0);
                    Dpi = T(0);
                    Xp = SF ? T(1) : T(0);
                }
                const T k1 = A * rcpT(Bm), k2 = C * rcpT(Bp);
                const T nA = -Am * k1, nC = -Cp * k2;
                T nB;
                if constexpr (SF) {
                    E = E + absT(k1) * Xm + absT(k2) * Xp;
                    nB = absT(nA) + absT(nC) + E;
                } else {
                    nB = B - Xm * k1 - Xp * k2;
                }
                const T nDr = Dr - Dmr * k1 - Dpr * k2;
                const T nDi = Di - Dmi * k1 - Dpi * k2;
                A = nA;
                B = nB;
                C = nC;
                Dr = nDr;
                Di = nDi;
            }
            const T iB = rcpT(B);
            const T Xr = Dr * iB, Xi = Di * iB;
            const int sprev = (l > 0) ? l - 1 : l;
            T xmr = shfl_grp<L>(Xr, sprev, base), xmi = shfl_grp<L>(Xi, sprev, base);
            if (l == 0) {
                xmr = T(0);
                xmi = T(0);
            }
            // back substitution: x_i = delta_i - alpha_i x_{s-1} - gamma_i x_{i+1}
            T vr = Xr, vi = Xi;
#pragma unroll
            for (int k = KMAX - 1; k >= 0; --k) {
                xr[k] = T(0);
                xi[k] = T(0);
                if (k < cnt) {
                    if (k < cnt - 1) {
                        vr = dr[k] - al[k] * xmr - ga[k] * vr;
                        vi = di[k] - al[k] * xmi - ga[k] * vi;
                    }
                    xr[k] = vr;
                    xi[k] = vi;
                }
            }
        }
        if constexpr (RESID) {
            // ||f - T x|| / ||f|| in FP64 over the mode, f the staged
            // right-hand side (FP32 values, exact); the neighbouring unknowns
            // across lane boundaries come by shuffles
            const int sprev = (l > 0) ? l - 1 : l, snext = (l + 1 < L) ? l + 1 : l;
            T lastR = T(0), lastI = T(0);
#pragma unroll
            for (int k = 0; k < KMAX; ++k)
                if (k == cnt - 1) {
                    lastR = xr[k];
                    lastI = xi[k];
                }
            double pmr = (double)shfl_grp<L>(lastR, sprev, base), pmi = (double)shfl_grp<L>(lastI, sprev, base);
            const double x0r = (double)shfl_grp<L>(xr[0], snext, base);
            const double x0i = (double)shfl_grp<L>(xi[0], snext, base);
            if (l == 0) {
                pmr = 0.0;
                pmi = 0.0;
            }
            double rsum = 0.0, dsum = 0.0;
#pragma unroll
            for (int k = 0; k < KMAX; ++k) {
                if (k < cnt) {
                    double a, b, c;
                    coef(k, a, b, c);
                    const double pr = (k == 0) ? pmr : (double)xr[k > 0 ? k - 1 : 0];
                    const double pi = (k == 0) ? pmi : (double)xi[k > 0 ? k - 1 : 0];
                    const double qr = (k == cnt - 1) ? x0r : (double)xr[k + 1 < KMAX ? k + 1 : k];
                    const double qi = (k == cnt - 1) ? x0i : (double)xi[k + 1 < KMAX ? k + 1 : k];
                    const float2 d = zs(g, k, l);
                    const double fre = (n == 0) ? 0.0 : (double)d.x, fim = (double)d.y;
                    const double er = fre - (a * pr + b * (double)xr[k] + c * qr);
                    const double ei = fim - (a * pi + b * (double)xi[k] + c * qi);
                    rsum += er * er + ei * ei;
                    dsum += fre * fre + fim * fim;
                }
            }
            for (int m = 1; m < L; m <<= 1) {
                rsum += __shfl_xor(rsum, m, 64);
                dsum += __shfl_xor(dsum, m, 64);
            }
            if (active && l == 0) s.resid[n] = dsum > 0.0 ? sqrt(rsum / dsum) : 0.0;
        }
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            if (k < cnt) {
                if (p64)
                    spec64[(size_t)n * nT + r0 + k] = make_double2((n == 0) ? 0.0 : (double)xr[k], (double)xi[k]);
                else
                    zs(g, k, l) = make_float2((n == 0) ? 0.f : (float)xr[k], (float)xi[k]);
            }
        }
        };
        if (tile.rows == 256 / L && nT == L * KMAX)
            solve(std::integral_constant<int, KMAX>{});
        else
            solve(cnt);
        __syncthreads();
        if (p64) return;  // the solution is already stored (uniform over the block)
        // write-back in spectrum order, pairs as float4 where both rows are real
        float* __restrict__ spw = s.spec;
        for (int p = tid; p < nPairs; p += 256) {
            int g, i;
            const int a = pair(p, g, i);
            if (i >= nT) continue;
            int li, ki;
            split(i, li, ki);
            const float2 x = zs(g, ki, li);
            if (i + 1 < nT) {
                step_row(li, ki);
                const float2 y = zs(g, ki, li);
                *reinterpret_cast<float4*>(spw + 2 * (size_t)a) = make_float4(x.x, x.y, y.x, y.y);
            } else {
                *reinterpret_cast<float2*>(spw + 2 * (size_t)a) = x;
            }
        }
    }
}


template <int L, int KMAX, bool RESID, typename T>
__global__ __launch_bounds__(256) void k_tridiag(const SimDesc* __restrict__ sims,
                                                 const Tile* __restrict__ tiles, int ntiles, int parity) {
    __shared__ float2 zs[(256 / L) * KMAX * (L + 1)];
    constexpr bool SF = std::is_same<T, float>::value;
    constexpr int NR = KMAX * (L + 1);
    __shared__ double cs[(!SF || RESID) ? NR : 1], ss[(!SF || RESID) ? NR : 1];
    __shared__ float fr[SF ? 3 * NR : 1];
    for (int t = xcd_tile(ntiles); t < ntiles; t = ntiles)
        tri_tile<L, KMAX, RESID, T>(sims, tiles[t], parity, zs, cs, ss, fr);
}

// Every class with K = 8 rows per lane in one launch: a tile carries its lanes
// per mode in col0, and the block takes that variant (mixed-size batches
// would otherwise queue one small launch per class).
template <bool RESID, typename T>
__global__ __launch_bounds__(256) void k_tridiag_mixed(const SimDesc* __restrict__ sims,
                                                       const Tile* __restrict__ tiles, int ntiles,
                                                       int parity) {
    constexpr int K = 8, ZS = 32 * K * 9, CS = K * 65;  // largest over L = 8 ... 64
    __shared__ float2 zs[ZS];
    constexpr bool SF = std::is_same<T, float>::value;
    __shared__ double cs[(!SF || RESID) ? CS : 1], ss[(!SF || RESID) ? CS : 1];
    __shared__ float fr[SF ? 3 * CS : 1];
    for (int t = xcd_tile(ntiles); t < ntiles; t = ntiles) {
        const Tile tile = tiles[t];
        switch (tile.col0) {
            case 8: tri_tile<8, K, RESID, T>(sims, tile, parity, zs, cs, ss, fr); break;
            case 16: tri_tile<16, K, RESID, T>(sims, tile, parity, zs, cs, ss, fr); break;
            case 32: tri_tile<32, K, RESID, T>(sims, tile, parity, zs, cs, ss, fr); break;
            default: tri_tile<64, K, RESID, T>(sims, tile, parity, zs, cs, ss, fr); break;
        }
    }
}

// ---------------------------------------------------------------------------
// k_ifft_grad: pressure rows j0..j0+R (one halo row) from the solved half
// spectrum by a length-M inverse FFT of
//   Y_n = (P_n + conj P_{M-n}) + i e^{i pi n / M} (P_n - conj P_{M-n}),
// whose output y_k = p_2k + i p_2k+1 is the real pressure row in natural order
// (cuFFT CUFFT_FORWARD plus shiftUKernel, KaminoCore.cu:682-703, with the
// n = 0 mode dropped), then the gradient subtraction of applyPressureTheta
// (:705-722) and applyPressurePhi (:724-747), in place.
//
// C = float2 is the FP32 form (p-hat from the spectrum buffer, FP32 twiddles
// and gradient factors); C = double2 the FP64 pressure (SIM_P64): p-hat from
// the double2 spectrum over the finished state copy, FP64 twiddles, FFT and
// differences, one rounding of each updated velocity to FP32.  PONLY writes
// the tile's pressure rows to pout (float [nTheta][nPhi], one simulation)
// instead of updating the velocities: the pressure readback.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_spec(Rsrc r, int idx, float2& v) { v = ldf2(r, 2 * idx); }
__device__ __forceinline__ void load_spec(Rsrc r, int idx, double2& v) {
    v = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, idx * 16, 0, 0));
}

// Stages the first M = nPhi / 2 FP64 twiddles of s into LDS, one double2
// (16 bytes) each; the transform takes the others as w(n + M) = -w(n).
__device__ __forceinline__ void stage_twiddles(double2* twl, const SimDesc& s, int tid, int nthr) {
    const Rsrc tr = make_rsrc(reinterpret_cast<const float*>(s.twN64), 4 * s.nPhi);
    float4* dst = reinterpret_cast<float4*>(twl);
    for (int i0 = tid; i0 < s.M; i0 += 4 * nthr) {
        float4 v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = ldf4(tr, 4 * (i0 + q * nthr));
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (i0 + q * nthr < s.M) dst[i0 + q * nthr] = v[q];
    }
}

template <typename C, bool PONLY>
__global__ __launch_bounds__(kFftBlock) void k_ifft_grad(const SimDesc* __restrict__ sims,
                                                   const Tile* __restrict__ tiles, int ntiles,
                                                   int parity, float* __restrict__ pout) {
    typedef typename CplxT<C>::R R;
    constexpr bool kP64 = std::is_same<C, double2>::value;
    extern __shared__ __attribute__((aligned(16))) float4 zlds4[];
    C* zlds = reinterpret_cast<C*>(zlds4);
    const int tid = threadIdx.x, nthr = blockDim.x;
    for (int t = xcd_tile(ntiles); t < ntiles; t = ntiles) {
        const Tile tile = tiles[t];
        const SimDesc& s = sims[tile.sim];
        const int nT = s.nTheta, nP = s.nPhi, M = s.M;
        const int j0 = tile.row0, R_ = tile.rows;
        const int Rp = min(R_ + 1, nT - j0);  // rows of p needed
        const int gsh = s.fftShift, gs = 1 << gsh;  // lanes per row (a power of two)
        // kFftBlock threads: kFftBlock / gs lane groups (a power of two)
        const int ngsh = kFftBlockShift - gsh, grp = tid >> gsh, ngrp = 1 << ngsh, lane = tid & (gs - 1);
        C* tw = zlds;                                               // nPhi twiddles (M in FP64)
        C* zbuf = zlds + (kP64 ? M : nP);                           // Rp rows of M
        C* scratch = zbuf + (size_t)(Rp + grp) * M;                 // this lane group's row
        const Rsrc specR = kP64 ? make_rsrc(s.state[parity ^ 1][0], 4 * M * nT)
                                : make_rsrc(s.spec, 2 * M * spec_rows(nT));
        if constexpr (kP64)
            stage_twiddles(tw, s, tid, nthr);
        else
            stage_twiddles(tw, s, tid, nthr, false);
        __syncthreads();
        // ceil(2^32 / Rp) (Rp = 1 splits nothing); idx / Rp is exact for idx < 2^16
        const unsigned magicR = (Rp == 1) ? 0u : 0xffffffffu / (unsigned)Rp + 1u;
        for (int idx0 = tid; idx0 < Rp * M; idx0 += kSpecBatch * nthr) {
            C Pn[kSpecBatch], Pm[kSpecBatch];
#pragma unroll
            for (int q = 0; q < kSpecBatch; ++q) {
                const int idx = idx0 + q * nthr;
                const int n = (Rp == 1) ? idx : (int)__umulhi((unsigned)idx, magicR), j = j0 + idx - n * Rp;
                // past the spectrum (idx beyond the tile, or M - n = M) reads as 0
                // or is not used; the FP64 p-hat is mode-major, the FP32 one
                // in row blocks
                if constexpr (kP64) {
                    load_spec(specR, n * nT + j, Pn[q]);
                    load_spec(specR, (M - n) * nT + j, Pm[q]);
                } else {
                    load_spec(specR, (int)spec_index(n, j, M), Pn[q]);
                    load_spec(specR, (int)spec_index(M - n, j, M), Pm[q]);
                }
            }
#pragma unroll
            for (int q = 0; q < kSpecBatch; ++q) {
                const int idx = idx0 + q * nthr;
                if (idx >= Rp * M) break;
                const int n = (Rp == 1) ? idx : (int)__umulhi((unsigned)idx, magicR), r = idx - n * Rp;
                C Y;
                if (n == 0) {
                    Y = mkc(Pn[q].y, -Pn[q].y);  // P_0 = 0 (mode dropped), P_M = Nyquist
                } else {
                    const C ev = mkc(Pn[q].x + Pm[q].x, Pn[q].y - Pm[q].y);
                    const C od = mkc(Pn[q].x - Pm[q].x, Pn[q].y + Pm[q].y);
                    C w = tw[n];
                    w.y = -w.y;  // exp(+i pi n / M)
                    const C o = cmul(w, od);
                    Y = mkc(ev.x - o.y, ev.y + o.x);
                }
                zbuf[(size_t)r * M + n] = Y;
            }
        }
        __syncthreads();
        const int nPass = (Rp + ngrp - 1) >> ngsh;
        for (int it = 0; it < nPass; ++it) {
            const int r = it * ngrp + grp;
            wave_fft<+1, kP64>(zbuf + (size_t)(r < Rp ? r : 0) * M, scratch, (const C*)tw, s, lane, gs, r < Rp);
        }
        __syncthreads();
        const R* p = reinterpret_cast<const R*>(zbuf);  // p[r][i], row stride nP
        if constexpr (PONLY) {
            for (int idx = tid; idx < R_ * nP; idx += nthr) pout[(size_t)j0 * nP + idx] = (float)p[idx];
            __syncthreads();
            continue;
        }
        float* __restrict__ up = s.state[parity][F_UPHI];
        float* __restrict__ ut = s.state[parity][F_UTHETA];
        const Rsrc upR = make_rsrc(up, nT * nP), utR = make_rsrc(ut, (nT - 1) * nP);
        const Rsrc hsR = kP64 ? make_rsrc(reinterpret_cast<const float*>(s.rowd + RD_INVHS * nT), 2 * nT)
                              : make_rsrc(s.rowf + RT_INVHS * nT, nT);
        R invH;
        if constexpr (kP64) invH = s.invH64;
        else invH = s.invH;
        // column pairs (i, i + 1) (nPhi is even); a batch's u loads all go out
        // before its first store, and u_theta of the last row reads as 0
        const int half = nP >> 1, nPair = R_ * half;
        const unsigned magicH = s.magicP2;  // idx / (nPhi / 2)
        for (int idx0 = tid; idx0 < nPair; idx0 += kGradBatch * nthr) {
            float2 u[kGradBatch], v[kGradBatch];
            R hs[kGradBatch];
#pragma unroll
            for (int q = 0; q < kGradBatch; ++q) {
                const int idx = idx0 + q * nthr;
                const int r = (int)__umulhi((unsigned)idx, magicH), i = 2 * (idx - r * half);
                const int o = (j0 + r) * nP + i;
                u[q] = ldf2(upR, o);
                v[q] = ldf2(utR, o);
                if constexpr (kP64) hs[q] = ldd(hsR, j0 + r);
                else hs[q] = ldf(hsR, j0 + r);
            }
#pragma unroll
            for (int q = 0; q < kGradBatch; ++q) {
                const int idx = idx0 + q * nthr;
                if (idx >= nPair) break;
                const int r = (int)__umulhi((unsigned)idx, magicH), i = 2 * (idx - r * half), j = j0 + r;
                const int il = (i == 0) ? nP - 1 : i - 1;
                const R* pr = p + (size_t)r * nP;
                const R p0 = pr[i], p1 = pr[i + 1];
                const float2 nu = make_float2((float)((R)u[q].x - (p0 - pr[il]) * hs[q]),
                                              (float)((R)u[q].y - (p1 - p0) * hs[q]));
                const size_t o = (size_t)j * nP + i;
                *reinterpret_cast<float2*>(up + o) = nu;
                bool bad = !isfinite(nu.x) || !isfinite(nu.y);
                if (j < nT - 1) {
                    const float2 nv = make_float2((float)((R)v[q].x - (pr[nP + i] - p0) * invH),
                                                  (float)((R)v[q].y - (pr[nP + i + 1] - p1) * invH));
                    *reinterpret_cast<float2*>(ut + o) = nv;
                    bad |= !isfinite(nv.x) || !isfinite(nv.y);
                }
                flag_bad(s.status, bad);
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
// one block per tile, rounded up to a multiple of the XCD count (see xcd_tile)
static inline int grid_for(int ntiles, int cap) {
    (void)cap;
    return ((ntiles + kXcds - 1) / kXcds) * kXcds;
}

hipError_t launch_advect(const SimDesc* sims, const Tile* tiles, int ntiles, int parity,
                         int stride, size_t ldsBytes, int gridCap, hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    if (stride == adv_stride(64))
        hipLaunchKernelGGL(k_advect<adv_stride(64)>, dim3(grid_for(ntiles, gridCap)), dim3(256),
                           ldsBytes, st, sims, tiles, ntiles, parity);
    else if (stride == adv_stride(128))
        hipLaunchKernelGGL(k_advect<adv_stride(128)>, dim3(grid_for(ntiles, gridCap)), dim3(256),
                           ldsBytes, st, sims, tiles, ntiles, parity);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_particles(const SimDesc* sims, const Tile* tiles, int ntiles, int parity,
                            int gridCap, hipStream_t st) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_particles, dim3(grid_for(ntiles, gridCap)), dim3(256), 0, st, sims, tiles,
                       ntiles, parity);
    return hipGetLastError();
}

hipError_t launch_div_fft(const SimDesc* sims, const Tile* tiles, int ntiles, int parity,
                          size_t ldsBytes, int gridCap, hipStream_t st, bool half) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(half ? k_div_fft8 : k_div_fft, dim3(grid_for(ntiles, gridCap)), dim3(kFftBlock),
                       ldsBytes, st, sims, tiles, ntiles, parity);
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_tridiag_t(int lanesPerMode, int kmax, const SimDesc* sims, const Tile* tiles,
                                   int ntiles, int parity, dim3 grid, hipStream_t st, bool resid) {
    const dim3 block(256);
    if (lanesPerMode == 0 && kmax == 8) {
        if (resid)
            hipLaunchKernelGGL((k_tridiag_mixed<true, T>), grid, block, 0, st, sims, tiles, ntiles, parity);
        else
            hipLaunchKernelGGL((k_tridiag_mixed<false, T>), grid, block, 0, st, sims, tiles, ntiles, parity);
        return hipGetLastError();
    }
#define KAM_TRI(LL, KK)                                                                        \
    if (lanesPerMode == LL && kmax == KK) {                                                    \
        if (resid)                                                                             \
            hipLaunchKernelGGL((k_tridiag<LL, KK, true, T>), grid, block, 0, st, sims, tiles, ntiles, \
                               parity);                                                        \
        else                                                                                   \
            hipLaunchKernelGGL((k_tridiag<LL, KK, false, T>), grid, block, 0, st, sims, tiles, ntiles, \
                               parity);                                                        \
        return hipGetLastError();                                                              \
    }
    KAM_TRI(8, 8)
    KAM_TRI(16, 8)
    KAM_TRI(32, 8)
    KAM_TRI(64, 8)
    KAM_TRI(64, 16)
    KAM_TRI(64, 32)
#undef KAM_TRI
    return hipErrorInvalidValue;
}

hipError_t launch_tridiag(int lanesPerMode, int kmax, const SimDesc* sims, const Tile* tiles, int ntiles,
                          int parity, int gridCap, hipStream_t st, bool resid, bool fp32) {
    if (ntiles <= 0) return hipSuccess;
    const dim3 grid(grid_for(ntiles, gridCap));
    return fp32 ? launch_tridiag_t<float>(lanesPerMode, kmax, sims, tiles, ntiles, parity, grid, st, resid)
                : launch_tridiag_t<double>(lanesPerMode, kmax, sims, tiles, ntiles, parity, grid, st, resid);
}

hipError_t launch_ifft_grad(const SimDesc* sims, const Tile* tiles, int ntiles, int parity,
                            size_t ldsBytes, int gridCap, hipStream_t st, bool p64, float* pout) {
    if (ntiles <= 0) return hipSuccess;
    const dim3 grid(grid_for(ntiles, gridCap)), block(kFftBlock);
    if (p64 && pout)
        hipLaunchKernelGGL((k_ifft_grad<double2, true>), grid, block, ldsBytes, st, sims, tiles, ntiles, parity,
                           pout);
    else if (p64)
        hipLaunchKernelGGL((k_ifft_grad<double2, false>), grid, block, ldsBytes, st, sims, tiles, ntiles, parity,
                           pout);
    else if (pout)
        hipLaunchKernelGGL((k_ifft_grad<float2, true>), grid, block, ldsBytes, st, sims, tiles, ntiles, parity,
                           pout);
    else
        hipLaunchKernelGGL((k_ifft_grad<float2, false>), grid, block, ldsBytes, st, sims, tiles, ntiles, parity,
                           pout);
    return hipGetLastError();
}

}  // namespace kam

namespace kam {

// Raises each kernel's dynamic-LDS limit to `bytes` less its static LDS.
static hipError_t allow_lds(const void* fn, size_t bytes) {
    hipFuncAttributes a;
    hipError_t e = hipFuncGetAttributes(&a, fn);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)(bytes - a.sharedSizeBytes));
}

hipError_t configure_kernels(size_t bytes) {
    const void* fns[] = {(const void*)k_advect<adv_stride(64)>, (const void*)k_advect<adv_stride(128)>,
                         (const void*)k_div_fft, (const void*)k_ifft_grad<float2, false>,
                         (const void*)k_ifft_grad<double2, false>, (const void*)k_ifft_grad<float2, true>,
                         (const void*)k_ifft_grad<double2, true>};
    for (const void* fn : fns) {
        const hipError_t e = allow_lds(fn, bytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace kam
