/*
 * kamino_oracle.c -- FP64 host restatement of the KaminoGPU time step.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and CPU baseline for the HIP
 * path.  The product library never links or calls this file.
 *
 * Each function restates one reference kernel body in double precision, in
 * the reference's radian coordinates and in its branch order.  Constants are
 * the reference's (KaminoHeader.cuh:26-41); the FP32 epsilon 1e-7f of
 * KaminoCore.cu:419 is kept at its float value.  One output per loop
 * iteration, parallel over cells with OpenMP, just as the reference launches
 * one GPU thread per cell.
 */
#include "kamino_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* KaminoHeader.cuh:26-28 */
#define REF_PI 3.14159265358979323846
#define REF_2PI 6.28318530717958647692
/* KaminoCore.cu:419 -- a float literal promoted where it is compared */
#define REF_EPS ((double)1e-7f)

/* Field staggering offsets in cells (KaminoHeader.cuh:30-41). */
static const double kPhiOff[3] = {-0.5, 0.0, 0.0};
static const double kThetaOff[3] = {0.5, 1.0, 0.5};

int ko_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* validateCoord -- KaminoCore.cu:11-29                                      */
/* ------------------------------------------------------------------------ */
double ko_validate_coord(double *phi, double *theta) {
    double ret = 1.0;
    *theta = *theta - (double)(int)floor(*theta / REF_2PI) * REF_2PI;
    if (*theta > REF_PI) {
        *theta = REF_2PI - *theta;
        *phi += REF_PI;
        ret = -ret;
    }
    if (*theta < 0) {
        *theta = -*theta;
        *phi += REF_PI;
        ret = -ret;
    }
    *phi = *phi - (double)(int)floor(*phi / REF_2PI) * REF_2PI;
    return ret;
}

/* kaminoLerp -- KaminoCore.cu:31-34 */
static inline double lerp(double from, double to, double alpha) {
    return (1.0 - alpha) * from + alpha * to;
}

static inline const double *row_ptr(const double *f, int rows, int nPhi, long row,
                                    int64_t *esc) {
    if (row < 0) {
        if (esc) (*esc)++;
        row = 0;
    } else if (row >= rows) {
        if (esc) (*esc)++;
        row = rows - 1;
    }
    return f + (size_t)row * nPhi;
}

/* sampleVPhi / sampleVTheta / sampleCentered -- KaminoCore.cu:36-184.
 * The three reference samplers differ only in their offsets, in the row that
 * triggers the southern pole branch (nTheta-1, or nTheta-2 for the u_theta
 * grid) and in whether the pole branch halves alphaTheta (not for u_phi). */
double ko_sample(int kind, const double *field, int nTheta, int nPhi, double h,
                 double phiRaw, double thetaRaw, int64_t *escapes) {
    const int rows = (kind == KO_VTHETA) ? nTheta - 1 : nTheta;
    const int poleRow = (kind == KO_VTHETA) ? nTheta - 2 : nTheta - 1;
    double phi = phiRaw - h * kPhiOff[kind];
    double theta = thetaRaw - h * kThetaOff[kind];
    const double invGridLen = 1.0 / h;
    const double isFlippedPole = ko_validate_coord(&phi, &theta);
    const double normedPhi = phi * invGridLen;
    const double normedTheta = theta * invGridLen;
    const long phiIndex = (long)floor(normedPhi);
    const long thetaIndex = (long)floor(normedTheta);
    const double alphaPhi = normedPhi - (double)phiIndex;
    double alphaTheta = normedTheta - (double)thetaIndex;

    if ((thetaIndex == 0 && isFlippedPole == -1.0) || thetaIndex == poleRow) {
        /* pole branch: same row at phi and at phi + pi (KaminoCore.cu:52-67,
         * 102-117, 152-167) */
        const double *r = row_ptr(field, rows, nPhi, thetaIndex, escapes);
        size_t phiLower = (size_t)phiIndex % (size_t)nPhi;
        size_t phiHigher = (phiLower + 1) % (size_t)nPhi;
        const double lowerBelt = lerp(r[phiLower], r[phiHigher], alphaPhi);
        phiLower = (phiLower + (size_t)nPhi / 2) % (size_t)nPhi;
        phiHigher = (phiLower + 1) % (size_t)nPhi;
        const double higherBelt = lerp(r[phiLower], r[phiHigher], alphaPhi);
        if (kind != KO_VPHI) alphaTheta = 0.5 * alphaTheta;
        return lerp(lowerBelt, higherBelt, alphaTheta);
    } else {
        const double *r0 = row_ptr(field, rows, nPhi, thetaIndex, escapes);
        const double *r1 = row_ptr(field, rows, nPhi, thetaIndex + 1, escapes);
        const size_t phiLower = (size_t)phiIndex % (size_t)nPhi;
        const size_t phiHigher = (phiLower + 1) % (size_t)nPhi;
        const double lowerBelt = lerp(r0[phiLower], r0[phiHigher], alphaPhi);
        const double higherBelt = lerp(r1[phiLower], r1[phiHigher], alphaPhi);
        return lerp(lowerBelt, higherBelt, alphaTheta);
    }
}

/* advectionParticles (KaminoCore.cu:321-342): one forward-Euler step of each
 * particle (phi, theta) with the pre-step velocities sampled at the particle
 * (sampleVPhi, sampleVTheta), no phi motion where R sin(theta) <= 1e-7, then
 * validateCoord.  amb[k] = 1 when either sample lay within tol cells of a
 * sampler branch break (see near_break), exact hits included: a particle
 * clamped to theta = 0 (KaminoParticles.cu:52-55) samples u_theta exactly at
 * the flipped idx 0 -> 1 change, and there the reference's own FP32 fold
 * (h - 2 pi + 2 pi) picks the branch by rounding. */
static int near_break(int kind, int nTheta, double h, double thetaRaw, double tol);
static int on_break(int kind, int nTheta, double h, double thetaRaw, double tol);

void ko_advect_particles(int nTheta, int nPhi, double radius, double dt, const double *u_phi,
                         const double *u_theta, int64_t n, const double *pos, double *out,
                         uint8_t *amb, double tol) {
    const double h = REF_PI / nTheta; /* KaminoCore.cu:849 */
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) {
        const double phi = pos[2 * k], theta = pos[2 * k + 1];
        const double uPhi = ko_sample(KO_VPHI, u_phi, nTheta, nPhi, h, phi, theta, NULL);
        const double uTheta = ko_sample(KO_VTHETA, u_theta, nTheta, nPhi, h, phi, theta, NULL);
        const double latRadius = radius * sin(theta);
        double th = theta + uTheta * (dt / radius);
        double ph = phi;
        if (latRadius > 1e-7) ph = phi + uPhi * (dt / latRadius);
        ko_validate_coord(&ph, &th);
        out[2 * k] = ph;
        out[2 * k + 1] = th;
        if (amb)
            amb[k] = (uint8_t)(tol > 0.0 && (on_break(KO_VPHI, nTheta, h, theta, tol) ||
                                             on_break(KO_VTHETA, nTheta, h, theta, tol)));
    }
}

/* Unfolded, offset-shifted theta coordinate in cells, and whether it lies
 * within tol of a point where the sampler's branch selection jumps.  The
 * break points follow from the branch structure above: the north fold at 0,
 * the flipped idx 0 -> 1 change at -1, and in the south the fold at pi for
 * u_phi / centred grids, or the first out-of-range row for the u_theta grid.
 * Exact hits are not ambiguous: both implementations compute them exactly. */
static int near_break(int kind, int nTheta, double h, double thetaRaw, double tol) {
    const double t = (thetaRaw - h * kThetaOff[kind]) / h;
    const double south = (kind == KO_VTHETA) ? (double)(nTheta - 1) : (double)nTheta;
    const double br[3] = {-1.0, 0.0, south};
    for (int k = 0; k < 3; ++k) {
        const double d = fabs(t - br[k]);
        if (d > 0.0 && d < tol) return 1;
    }
    return 0;
}

static int on_break(int kind, int nTheta, double h, double thetaRaw, double tol) {
    const double t = (thetaRaw - h * kThetaOff[kind]) / h;
    const double south = (kind == KO_VTHETA) ? (double)(nTheta - 1) : (double)nTheta;
    const double br[3] = {-1.0, 0.0, south};
    for (int k = 0; k < 3; ++k)
        if (fabs(t - br[k]) < tol) return 1;
    return 0;
}

/* One RK2 semi-Lagrangian backtrace (advectionVPhiKernel KaminoCore.cu:186-229,
 * advectionVThetaKernel :231-274, advectionCentered :276-319).  `node_kind`
 * gives the node's staggering, `src_kind`/`src` the sampled field. */
static double advect_node(int node_kind, int src_kind, const double *src, int i, int j,
                          int nTheta, int nPhi, double radius, double dt, double h,
                          const double *u_phi, const double *u_theta, double tol,
                          int *amb, int64_t *esc) {
    const double gPhi = ((double)i + kPhiOff[node_kind]) * h;
    const double gTheta = ((double)j + kThetaOff[node_kind]) * h;

    const double guPhi = ko_sample(KO_VPHI, u_phi, nTheta, nPhi, h, gPhi, gTheta, esc);
    const double guTheta = ko_sample(KO_VTHETA, u_theta, nTheta, nPhi, h, gPhi, gTheta, esc);

    const double latRadius = radius * sin(gTheta);
    const double cofPhi = dt / latRadius;
    const double cofTheta = dt / radius;

    double deltaPhi = guPhi * cofPhi;
    double deltaTheta = guTheta * cofTheta;

    const double midPhi = gPhi - 0.5 * deltaPhi;
    const double midTheta = gTheta - 0.5 * deltaTheta;
    const double muPhi = ko_sample(KO_VPHI, u_phi, nTheta, nPhi, h, midPhi, midTheta, esc);
    const double muTheta = ko_sample(KO_VTHETA, u_theta, nTheta, nPhi, h, midPhi, midTheta, esc);

    const double averuPhi = 0.5 * (muPhi + guPhi);
    const double averuTheta = 0.5 * (muTheta + guTheta);
    deltaPhi = averuPhi * cofPhi;
    deltaTheta = averuTheta * cofTheta;

    const double pPhi = gPhi - deltaPhi;
    const double pTheta = gTheta - deltaTheta;

    if (amb) {
        const double scale = 1.0 + fabs(deltaTheta) / h;
        const double t = tol * scale;
        *amb = near_break(KO_VPHI, nTheta, h, midTheta, t) ||
               near_break(KO_VTHETA, nTheta, h, midTheta, t) ||
               near_break(src_kind, nTheta, h, pTheta, t);
    }
    return ko_sample(src_kind, src, nTheta, nPhi, h, pPhi, pTheta, esc);
}

void ko_advect(int nTheta, int nPhi, double radius, double dt,
               const double *u_phi, const double *u_theta, const double *rho,
               double *u_phi_out, double *u_theta_out, double *rho_out,
               uint8_t *amb_phi, uint8_t *amb_theta, uint8_t *amb_rho,
               double tol, ko_stats *stats) {
    const double h = REF_PI / nTheta; /* KaminoCore.cu:849 */
    int64_t esc = 0, na_p = 0, na_t = 0, na_r = 0;

#pragma omp parallel for collapse(2) reduction(+ : esc, na_p) schedule(static)
    for (int j = 0; j < nTheta; ++j)
        for (int i = 0; i < nPhi; ++i) {
            int amb = 0;
            u_phi_out[(size_t)j * nPhi + i] =
                advect_node(KO_VPHI, KO_VPHI, u_phi, i, j, nTheta, nPhi, radius, dt, h,
                            u_phi, u_theta, tol, amb_phi ? &amb : NULL, &esc);
            if (amb_phi) amb_phi[(size_t)j * nPhi + i] = (uint8_t)amb;
            na_p += amb;
        }

#pragma omp parallel for collapse(2) reduction(+ : esc, na_t) schedule(static)
    for (int j = 0; j < nTheta - 1; ++j)
        for (int i = 0; i < nPhi; ++i) {
            int amb = 0;
            u_theta_out[(size_t)j * nPhi + i] =
                advect_node(KO_VTHETA, KO_VTHETA, u_theta, i, j, nTheta, nPhi, radius, dt,
                            h, u_phi, u_theta, tol, amb_theta ? &amb : NULL, &esc);
            if (amb_theta) amb_theta[(size_t)j * nPhi + i] = (uint8_t)amb;
            na_t += amb;
        }

    if (rho && rho_out) {
#pragma omp parallel for collapse(2) reduction(+ : esc, na_r) schedule(static)
        for (int j = 0; j < nTheta; ++j)
            for (int i = 0; i < nPhi; ++i) {
                int amb = 0;
                rho_out[(size_t)j * nPhi + i] =
                    advect_node(KO_CENTERED, KO_CENTERED, rho, i, j, nTheta, nPhi, radius,
                                dt, h, u_phi, u_theta, tol, amb_rho ? &amb : NULL, &esc);
                if (amb_rho) amb_rho[(size_t)j * nPhi + i] = (uint8_t)amb;
                na_r += amb;
            }
    }
    if (stats) {
        stats->pole_escapes += esc;
        stats->ambiguous_uphi += na_p;
        stats->ambiguous_utheta += na_t;
        stats->ambiguous_rho += na_r;
    }
}

/* ------------------------------------------------------------------------ */
/* Cubic -- KaminoCore.cu:386-454                                            */
/* ------------------------------------------------------------------------ */
static double ref_root3_pos(double x) { /* _root3, KaminoCore.cu:386-407 */
    /* x > 0; +inf / NaN return as is (the reference's scaling loop never ends on +inf) */
    if (!(x <= DBL_MAX)) return x;
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

static double ref_root3(double x) { /* root3, KaminoCore.cu:409-417 */
    if (x > 0) return ref_root3_pos(x);
    if (x < 0) return -ref_root3_pos(-x);
    return 0.0;
}

double ko_solve_cubic(double a, double b, double c) {
    const double a2 = a * a;
    double q = (a2 - 3 * b) / 9.0;
    const double r = (a * (2.0 * a2 - 9.0 * b) + 27.0 * c) / 54.0;
    const double r2 = r * r;
    const double q3 = q * q * q;
    if (r2 <= (q3 + REF_EPS)) {
        double t = r / sqrt(q3);
        if (t < -1) t = -1;
        if (t > 1) t = 1;
        t = acos(t);
        a /= 3.0;
        q = -2.0 * sqrt(q);
        return q * cos(t / 3.0) - a;
    } else {
        double A = -ref_root3(fabs(r) + sqrt(r2 - q3));
        if (r < 0) A = -A;
        const double B = (A == 0) ? 0 : q / A;
        a /= 3.0;
        return (A + B) - a;
    }
}

/* ------------------------------------------------------------------------ */
/* Geometric -- geometricFillKernel KaminoCore.cu:457-514, assignPhiKernel   */
/* :517-534, assignThetaKernel :537-549, orchestration :551-583              */
/* ------------------------------------------------------------------------ */
void ko_geometric(int nTheta, int nPhi, double radius, double dt, double *u_phi,
                  double *u_theta) {
    const double h = REF_PI / nTheta;
    const size_t N = (size_t)nTheta * nPhi;
    double *xs = (double *)malloc(N * sizeof(double));
    double *vs = (double *)malloc(N * sizeof(double));

#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 0; j < nTheta; ++j)
        for (int i = 0; i < nPhi; ++i) {
            const double gTheta = ((double)j + 0.5) * h;
            const size_t phiLeft = (size_t)i;
            const size_t phiRight = (size_t)(i + 1) % (size_t)nPhi;
            const double uPrev =
                0.5 * (u_phi[phiLeft + (size_t)nPhi * j] + u_phi[phiRight + (size_t)nPhi * j]);
            double vPrev;
            if (j == 0) {
                const size_t opp = (size_t)(i + nPhi / 2) % (size_t)nPhi;
                vPrev = 0.75 * u_theta[i] + 0.25 * u_theta[opp];
            } else if (j == nTheta - 1) {
                const size_t opp = (size_t)(i + nPhi / 2) % (size_t)nPhi;
                vPrev = 0.75 * u_theta[i + (size_t)nPhi * (j - 1)] +
                        0.25 * u_theta[opp + (size_t)nPhi * (j - 1)];
            } else {
                vPrev = 0.5 * (u_theta[i + (size_t)nPhi * (j - 1)] +
                               u_theta[i + (size_t)nPhi * j]);
            }
            const double G = dt * cos(gTheta) / (radius * sin(gTheta));
            double uNext;
            if (fabs(G) > REF_EPS) {
                const double cof = G * G;
                const double A = 0.0;
                const double B = (G * vPrev + 1.0) / cof;
                const double C = -uPrev / cof;
                uNext = ko_solve_cubic(A, B, C);
            } else {
                uNext = uPrev;
            }
            const double vNext = vPrev + G * uNext * uNext;
            xs[(size_t)i + (size_t)nPhi * j] = uNext;
            vs[(size_t)i + (size_t)nPhi * j] = vNext;
        }

#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 0; j < nTheta; ++j)
        for (int i = 0; i < nPhi; ++i) {
            const size_t left = (i == 0) ? (size_t)nPhi - 1 : (size_t)i - 1;
            u_phi[(size_t)i + (size_t)nPhi * j] =
                0.5 * (xs[left + (size_t)nPhi * j] + xs[(size_t)i + (size_t)nPhi * j]);
        }

#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 0; j < nTheta - 1; ++j)
        for (int i = 0; i < nPhi; ++i)
            u_theta[(size_t)i + (size_t)nPhi * j] =
                0.5 * (vs[(size_t)i + (size_t)nPhi * j] + vs[(size_t)i + (size_t)nPhi * (j + 1)]);

    free(xs);
    free(vs);
}

/* The geometric outputs read a 3-wide phi stencil of u_phi and the u_theta
 * rows either side of each cell row (plus the antipodal column in the pole
 * rows); an output is ambiguous if any advected input it reads is. */
void ko_geometric_mask(int nTheta, int nPhi, const uint8_t *amb_phi,
                       const uint8_t *amb_theta, uint8_t *out_phi, uint8_t *out_theta) {
    const size_t N = (size_t)nTheta * nPhi;
    uint8_t *cell = (uint8_t *)calloc(N, 1);
    for (int j = 0; j < nTheta; ++j)
        for (int i = 0; i < nPhi; ++i) {
            const int ip = (i + 1) % nPhi, io = (i + nPhi / 2) % nPhi;
            uint8_t m = amb_phi[(size_t)j * nPhi + i] | amb_phi[(size_t)j * nPhi + ip];
            if (j == 0)
                m |= amb_theta[i] | amb_theta[io];
            else if (j == nTheta - 1)
                m |= amb_theta[(size_t)(j - 1) * nPhi + i] | amb_theta[(size_t)(j - 1) * nPhi + io];
            else
                m |= amb_theta[(size_t)(j - 1) * nPhi + i] | amb_theta[(size_t)j * nPhi + i];
            cell[(size_t)j * nPhi + i] = m;
        }
    for (int j = 0; j < nTheta; ++j)
        for (int i = 0; i < nPhi; ++i) {
            const int il = (i + nPhi - 1) % nPhi;
            out_phi[(size_t)j * nPhi + i] = cell[(size_t)j * nPhi + il] | cell[(size_t)j * nPhi + i];
        }
    for (int j = 0; j < nTheta - 1; ++j)
        for (int i = 0; i < nPhi; ++i)
            out_theta[(size_t)j * nPhi + i] =
                cell[(size_t)j * nPhi + i] | cell[(size_t)(j + 1) * nPhi + i];
    free(cell);
}

/* ------------------------------------------------------------------------ */
/* Projection pieces                                                         */
/* ------------------------------------------------------------------------ */

/* fillDivergenceKernel -- KaminoCore.cu:587-638 (no 1/R factor). */
void ko_divergence(int nTheta, int nPhi, const double *u_phi, const double *u_theta,
                   double *div) {
    const double h = REF_PI / nTheta;
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 0; j < nTheta; ++j)
        for (int i = 0; i < nPhi; ++i) {
            const double gridThetaCoord = ((double)j + 0.5) * h;
            double vNorth = 0.0, vSouth = 0.0;
            const double halfStep = 0.5 * h;
            const double thetaSouth = gridThetaCoord + halfStep;
            const double thetaNorth = gridThetaCoord - halfStep;
            const int phiIdWest = i;
            const int phiIdEast = (phiIdWest + 1) % nPhi;
            const double uWest = u_phi[(size_t)j * nPhi + phiIdWest];
            const double uEast = u_phi[(size_t)j * nPhi + phiIdEast];
            if (j != 0) vNorth = u_theta[(size_t)(j - 1) * nPhi + i];
            if (j != nTheta - 1) vSouth = u_theta[(size_t)j * nPhi + i];
            const double invGridSine = 1.0 / sin(gridThetaCoord);
            const double sinNorth = sin(thetaNorth);
            const double sinSouth = sin(thetaSouth);
            const double factor = invGridSine / h;
            const double termTheta = factor * (vSouth * sinSouth - vNorth * sinNorth);
            const double termPhi = factor * (uEast - uWest);
            div[(size_t)j * nPhi + i] = termTheta + termPhi;
        }
}

/* precomputeABCKernel -- KaminoSolver.cu:117-163.  The reference's gridLen
 * there is 2*pi/nPhi (KaminoSolver.cu:14), identical to pi/nTheta. */
void ko_poisson_rows(int nTheta, int nPhi, int n, double *a, double *b, double *c) {
    const double h = REF_2PI / nPhi;
    for (int i = 0; i < nTheta; ++i) {
        const double thetaI = ((double)i + 0.5) * h;
        const double cosThetaI = cos(thetaI);
        const double sinThetaI = sin(thetaI);
        double valB = -2.0 / (h * h) - (double)n * (double)n / (sinThetaI * sinThetaI);
        double valA = 1.0 / (h * h) - cosThetaI / 2.0 / h / sinThetaI;
        double valC = 1.0 / (h * h) + cosThetaI / 2.0 / h / sinThetaI;
        if (n != 0) {
            if (i == 0) {
                valB += valA;
                valA = 0.0;
            }
            if (i == nTheta - 1) {
                valB += valC;
                valC = 0.0;
            }
        } else {
            valA = 0.0;
            valB = 1.0;
            valC = 0.0;
        }
        a[i] = valA;
        b[i] = valB;
        c[i] = valC;
    }
}

void ko_thomas(int n, const double *a, const double *b, const double *c, const double *d,
               double *x) {
    double *cp = (double *)malloc((size_t)n * sizeof(double));
    double *dp = (double *)malloc((size_t)n * sizeof(double));
    cp[0] = c[0] / b[0];
    dp[0] = d[0] / b[0];
    for (int i = 1; i < n; ++i) {
        const double m = b[i] - a[i] * cp[i - 1];
        cp[i] = c[i] / m;
        dp[i] = (d[i] - a[i] * dp[i - 1]) / m;
    }
    x[n - 1] = dp[n - 1];
    for (int i = n - 2; i >= 0; --i) x[i] = dp[i] - cp[i] * x[i + 1];
    free(cp);
    free(dp);
}

/* crKernel -- tdm.cu:3-96, with the thread loop written out sequentially.
 * The reference launches nTheta/2 threads on a system of nTheta rows. */
int ko_cyclic_reduction(int n, const double *a_in, const double *b_in, const double *c_in,
                        const double *d_in, double *x_out) {
    if (n < 2 || (n & (n - 1)) != 0) return -1;
    double *a = (double *)malloc(5 * (size_t)n * sizeof(double));
    double *b = a + n, *c = b + n, *d = c + n, *x = d + n;
    memcpy(a, a_in, (size_t)n * sizeof(double));
    memcpy(b, b_in, (size_t)n * sizeof(double));
    memcpy(c, c_in, (size_t)n * sizeof(double));
    memcpy(d, d_in, (size_t)n * sizeof(double));
    memset(x, 0, (size_t)n * sizeof(double));
    const int systemSize = n;
    int numThreads = n / 2;
    int stride = 1;
    int iteration = 0;
    for (int s = systemSize / 2; s > 1; s >>= 1) ++iteration; /* log2(systemSize/2) */

    for (int j = 0; j < iteration; j++) { /* forward elimination, tdm.cu:43-63 */
        stride *= 2;
        const int delta = stride / 2;
        for (int t = 0; t < numThreads; ++t) {
            const int i = stride * t + stride - 1;
            const int iLeft = i - delta;
            int iRight = i + delta;
            if (iRight >= systemSize) iRight = systemSize - 1;
            const double tmp1 = a[i] / b[iLeft];
            const double tmp2 = c[i] / b[iRight];
            b[i] = b[i] - c[iLeft] * tmp1 - a[iRight] * tmp2;
            d[i] = d[i] - d[iLeft] * tmp1 - d[iRight] * tmp2;
            a[i] = -a[iLeft] * tmp1;
            c[i] = -c[iRight] * tmp2;
        }
        numThreads /= 2;
    }
    { /* 2x2 solve, tdm.cu:65-72 */
        const int addr1 = stride - 1;
        const int addr2 = 2 * stride - 1;
        const double tmp3 = b[addr2] * b[addr1] - c[addr1] * a[addr2];
        x[addr1] = (b[addr2] * d[addr1] - c[addr1] * d[addr2]) / tmp3;
        x[addr2] = (d[addr2] * b[addr1] - d[addr1] * a[addr2]) / tmp3;
    }
    numThreads = 2; /* backward substitution, tdm.cu:74-90 */
    for (int j = 0; j < iteration; j++) {
        const int delta = stride / 2;
        for (int t = 0; t < numThreads; ++t) {
            const int i = stride * t + stride / 2 - 1;
            if (i == delta - 1)
                x[i] = (d[i] - c[i] * x[i + delta]) / b[i];
            else
                x[i] = (d[i] - a[i] * x[i - delta] - c[i] * x[i + delta]) / b[i];
        }
        stride /= 2;
        numThreads *= 2;
    }
    memcpy(x_out, x, (size_t)n * sizeof(double));
    free(a);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FFT: recursive mixed-radix decimation in time, any length.  Stands in for */
/* cuFFT's C2C transform (plan KaminoSolver.cu:61-66, calls                  */
/* KaminoCore.cu:763-764, 812-813): unnormalised, sign +1 = CUFFT_INVERSE.   */
/* ------------------------------------------------------------------------ */
typedef struct {
    double re, im;
} cplx;

static int smallest_factor(int n) {
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    for (int p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

/* X[k] = sum_j x[j*istride] W^(jk), W = exp(sign*2*pi*i/n); tw holds W_N^e for
 * the top-level length N, and n divides N. */
static void fft_rec(int n, const cplx *in, int istride, cplx *out, const cplx *tw, int N,
                    cplx *scratch) {
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    const int p = smallest_factor(n);
    const int m = n / p;
    for (int q = 0; q < p; ++q) fft_rec(m, in + (size_t)q * istride, istride * p, out + (size_t)q * m, tw, N, scratch);
    const int step = N / n; /* W_n^e = W_N^(e*step) */
    cplx *tmp = scratch;    /* p entries */
    for (int k = 0; k < m; ++k) {
        for (int q = 0; q < p; ++q) {
            const cplx y = out[(size_t)q * m + k];
            const cplx w = tw[((size_t)q * k * step) % (size_t)N];
            tmp[q].re = y.re * w.re - y.im * w.im;
            tmp[q].im = y.re * w.im + y.im * w.re;
        }
        for (int s = 0; s < p; ++s) {
            double sr = 0.0, si = 0.0;
            for (int q = 0; q < p; ++q) {
                const cplx w = tw[((size_t)q * s % (size_t)p) * (size_t)m * step];
                sr += tmp[q].re * w.re - tmp[q].im * w.im;
                si += tmp[q].re * w.im + tmp[q].im * w.re;
            }
            out[(size_t)s * m + k].re = sr;
            out[(size_t)s * m + k].im = si;
        }
    }
}

static cplx *make_twiddles(int N, int sign) {
    cplx *tw = (cplx *)malloc((size_t)N * sizeof(cplx));
    for (int e = 0; e < N; ++e) {
        const double ang = REF_2PI * (double)e / (double)N;
        tw[e].re = cos(ang);
        tw[e].im = (double)sign * sin(ang);
    }
    return tw;
}

static void fft_rows(int rows, int n, cplx *data, int sign) {
    cplx *tw = make_twiddles(n, sign);
#pragma omp parallel
    {
        cplx *buf = (cplx *)malloc((size_t)n * sizeof(cplx));
        cplx *scr = (cplx *)malloc((size_t)n * sizeof(cplx));
#pragma omp for schedule(static)
        for (int r = 0; r < rows; ++r) {
            cplx *row = data + (size_t)r * n;
            fft_rec(n, row, 1, buf, tw, n, scr);
            memcpy(row, buf, (size_t)n * sizeof(cplx));
        }
        free(buf);
        free(scr);
    }
    free(tw);
}

int ko_fft(int n, double *re, double *im, int sign) {
    if (n < 1 || (sign != 1 && sign != -1)) return -1;
    cplx *x = (cplx *)malloc((size_t)n * sizeof(cplx));
    cplx *y = (cplx *)malloc((size_t)n * sizeof(cplx));
    cplx *s = (cplx *)malloc((size_t)n * sizeof(cplx));
    for (int i = 0; i < n; ++i) {
        x[i].re = re[i];
        x[i].im = im[i];
    }
    cplx *tw = make_twiddles(n, sign);
    fft_rec(n, x, 1, y, tw, n, s);
    for (int i = 0; i < n; ++i) {
        re[i] = y[i].re;
        im[i] = y[i].im;
    }
    free(tw);
    free(x);
    free(y);
    free(s);
    return 0;
}

/* projection() -- KaminoCore.cu:749-842, stage by stage. */
int ko_project(int nTheta, int nPhi, double *u_phi, double *u_theta, double *spec_re,
               double *spec_im, double *p_out, int solver) {
    if (solver == 1 && (nTheta & (nTheta - 1)) != 0) return -1;
    const double h = REF_PI / nTheta;
    const size_t N = (size_t)nTheta * nPhi;
    double *div = (double *)malloc(N * sizeof(double));
    cplx *F = (cplx *)malloc(N * sizeof(cplx));
    double *FRe = (double *)malloc(N * sizeof(double));
    double *FIm = (double *)malloc(N * sizeof(double));
    double *URe = (double *)malloc(N * sizeof(double));
    double *UIm = (double *)malloc(N * sizeof(double));
    double *zero = (double *)malloc((size_t)nTheta * sizeof(double));
    double *p = (double *)malloc(N * sizeof(double));

    ko_divergence(nTheta, nPhi, u_phi, u_theta, div); /* :753-758 */
    for (size_t k = 0; k < N; ++k) {
        F[k].re = div[k];
        F[k].im = 0.0;
    }
    fft_rows(nTheta, nPhi, F, +1); /* CUFFT_INVERSE, :763-764 */

    /* shiftFKernel :640-656 -- mode-major, divided by nPhi */
#pragma omp parallel for collapse(2) schedule(static)
    for (int t = 0; t < nTheta; ++t)
        for (int nIdx = 0; nIdx < nPhi; ++nIdx) {
            int fftIndex = nPhi / 2 - nIdx;
            if (fftIndex < 0) fftIndex += nPhi;
            FRe[(size_t)nIdx * nTheta + t] = F[(size_t)t * nPhi + fftIndex].re / (double)nPhi;
            FIm[(size_t)nIdx * nTheta + t] = F[(size_t)t * nPhi + fftIndex].im / (double)nPhi;
        }
    if (spec_re) memcpy(spec_re, FRe, N * sizeof(double));
    if (spec_im) memcpy(spec_im, FIm, N * sizeof(double));

    /* crKernel x2 :779-792 -- one system per mode, real and imaginary parts */
    int bad = 0;
#pragma omp parallel reduction(| : bad)
    {
        double *a = (double *)malloc((size_t)nTheta * 3 * sizeof(double));
        double *b = a + nTheta, *c = b + nTheta;
#pragma omp for schedule(static)
        for (int nIdx = 0; nIdx < nPhi; ++nIdx) {
            ko_poisson_rows(nTheta, nPhi, nIdx - nPhi / 2, a, b, c);
            double *fr = FRe + (size_t)nIdx * nTheta, *fi = FIm + (size_t)nIdx * nTheta;
            double *ur = URe + (size_t)nIdx * nTheta, *ui = UIm + (size_t)nIdx * nTheta;
            if (solver == 1) {
                bad |= ko_cyclic_reduction(nTheta, a, b, c, fr, ur) != 0;
                bad |= ko_cyclic_reduction(nTheta, a, b, c, fi, ui) != 0;
            } else {
                ko_thomas(nTheta, a, b, c, fr, ur);
                ko_thomas(nTheta, a, b, c, fi, ui);
            }
        }
        free(a);
    }

    /* copy2UFourier :658-670 and cacheZeroComponents :672-680 */
#pragma omp parallel for collapse(2) schedule(static)
    for (int t = 0; t < nTheta; ++t)
        for (int nIdx = 0; nIdx < nPhi; ++nIdx) {
            F[(size_t)t * nPhi + nIdx].re = URe[(size_t)nIdx * nTheta + t];
            F[(size_t)t * nPhi + nIdx].im = UIm[(size_t)nIdx * nTheta + t];
        }
    for (int t = 0; t < nTheta; ++t) zero[t] = F[(size_t)t * nPhi + nPhi / 2].re;

    fft_rows(nTheta, nPhi, F, -1); /* CUFFT_FORWARD, :812-813 */

    /* shiftUKernel :682-703 */
#pragma omp parallel for collapse(2) schedule(static)
    for (int t = 0; t < nTheta; ++t)
        for (int i = 0; i < nPhi; ++i) {
            const int fftIndex = (i != 0) ? nPhi - i : 0;
            const double g = F[(size_t)t * nPhi + fftIndex].re;
            p[(size_t)t * nPhi + i] = ((i % 2 == 0) ? g : -g) - zero[t];
        }
    if (p_out) memcpy(p_out, p, N * sizeof(double));

    /* applyPressureTheta :705-722 */
#pragma omp parallel for collapse(2) schedule(static)
    for (int t = 0; t < nTheta - 1; ++t)
        for (int i = 0; i < nPhi; ++i) {
            const double pN = p[(size_t)t * nPhi + i];
            const double pS = p[(size_t)(t + 1) * nPhi + i];
            u_theta[(size_t)t * nPhi + i] += (pS - pN) / (-h);
        }
    /* applyPressurePhi :724-747 */
#pragma omp parallel for collapse(2) schedule(static)
    for (int t = 0; t < nTheta; ++t)
        for (int i = 0; i < nPhi; ++i) {
            const int w = (i == 0) ? nPhi - 1 : i - 1;
            const double pW = p[(size_t)t * nPhi + w];
            const double pE = p[(size_t)t * nPhi + i];
            const double thetaBelt = ((double)t + 0.5) * h;
            u_phi[(size_t)t * nPhi + i] += (pE - pW) / (-h * sin(thetaBelt));
        }

    free(div);
    free(F);
    free(FRe);
    free(FIm);
    free(URe);
    free(UIm);
    free(zero);
    free(p);
    return bad ? -1 : 0;
}

/* stepForward -- KaminoSolver.cu:197-221: advection (all three fields read
 * the pre-step velocities, KaminoCore.cu:344-384), geometric, projection. */
int ko_step(int nTheta, int nPhi, double radius, double dt, int phases, double *u_phi,
            double *u_theta, double *rho, ko_stats *stats) {
    const size_t N = (size_t)nTheta * nPhi;
    if (phases & KO_PHASE_ADVECT) {
        double *a = (double *)malloc(3 * N * sizeof(double));
        ko_advect(nTheta, nPhi, radius, dt, u_phi, u_theta, rho, a, a + N, rho ? a + 2 * N : NULL,
                  NULL, NULL, NULL, 0.0, stats);
        memcpy(u_phi, a, N * sizeof(double));
        memcpy(u_theta, a + N, (N - (size_t)nPhi) * sizeof(double));
        if (rho) memcpy(rho, a + 2 * N, N * sizeof(double));
        free(a);
    }
    if (phases & KO_PHASE_GEOMETRIC) ko_geometric(nTheta, nPhi, radius, dt, u_phi, u_theta);
    if (phases & KO_PHASE_PROJECT)
        return ko_project(nTheta, nPhi, u_phi, u_theta, NULL, NULL, NULL, 0);
    return 0;
}

double ko_run(int nTheta, int nPhi, double radius, double dt, int phases, int steps,
              int threads, double *u_phi, double *u_theta, double *rho) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int s = 0; s < steps; ++s) ko_step(nTheta, nPhi, radius, dt, phases, u_phi, u_theta, rho, NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* ------------------------------------------------------------------------ */
/* Synthetic initial conditions                                              */
/* ------------------------------------------------------------------------ */
uint64_t ko_splitmix64(uint64_t *state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

double ko_uniform(uint64_t *state) { return (double)(ko_splitmix64(state) >> 11) * 0x1.0p-53; }

double ko_normal(uint64_t *state) {
    const double u1 = 1.0 - ko_uniform(state); /* (0, 1] */
    const double u2 = ko_uniform(state);
    return sqrt(-2.0 * log(u1)) * cos(REF_2PI * u2);
}

/* Normalised associated Legendre functions Pbar_l^m(x) = N_lm P_l^m(x),
 * N_lm = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!), without the Condon-Shortley
 * phase; out[l*(L+1)+m] for 0 <= m <= l <= L. */
static void legendre_table(int L, double x, double *out) {
    const double s = sqrt(fmax(0.0, 1.0 - x * x));
    for (int k = 0; k < (L + 1) * (L + 1); ++k) out[k] = 0.0;
    double pmm = sqrt(1.0 / (4.0 * REF_PI));
    for (int m = 0; m <= L; ++m) {
        if (m > 0) pmm *= sqrt((2.0 * m + 1.0) / (2.0 * m)) * s;
        out[m * (L + 1) + m] = pmm;
        if (m + 1 <= L) out[(m + 1) * (L + 1) + m] = sqrt(2.0 * m + 3.0) * x * pmm;
        for (int l = m + 2; l <= L; ++l) {
            const double alm = sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
            const double blm = sqrt((((double)(l - 1) * (l - 1)) - (double)m * m) /
                                    (4.0 * (double)(l - 1) * (l - 1) - 1.0));
            out[l * (L + 1) + m] = alm * (x * out[(l - 1) * (L + 1) + m] - blm * out[(l - 2) * (L + 1) + m]);
        }
    }
}

double ko_real_sh(int l, int m, double theta, double phi) {
    const int am = m < 0 ? -m : m;
    if (am > l) return 0.0;
    double *tab = (double *)malloc((size_t)(l + 1) * (l + 1) * sizeof(double));
    legendre_table(l, cos(theta), tab);
    const double P = tab[l * (l + 1) + am];
    free(tab);
    if (m == 0) return P;
    if (m > 0) return sqrt(2.0) * P * cos(m * phi);
    return sqrt(2.0) * P * sin(am * phi);
}

void ko_init_fields(int nTheta, int nPhi, double radius, int L, uint64_t seed,
                    int density_kind, double *u_phi, double *u_theta, double *rho) {
    const double h = REF_PI / nTheta;
    uint64_t st = seed;
    /* c[l][m+l] ~ N(0, 0.1/l), drawn l = 1..L, m = -l..l */
    double *coef = (double *)calloc((size_t)(L + 1) * (2 * L + 1), sizeof(double));
    for (int l = 1; l <= L; ++l)
        for (int m = -l; m <= l; ++m) coef[(size_t)l * (2 * L + 1) + (m + L)] = (0.1 / l) * ko_normal(&st);

    /* psi at cell corners (theta = j h, phi = (i - 1/2) h), j = 0..nTheta,
     * via per-row Fourier coefficients A_m(theta), B_m(theta). */
    const int nc = nPhi + 1;
    double *psi = (double *)malloc((size_t)(nTheta + 1) * nc * sizeof(double));
#pragma omp parallel
    {
        double *tab = (double *)malloc((size_t)(L + 1) * (L + 1) * sizeof(double));
        double *Am = (double *)malloc((size_t)(L + 1) * sizeof(double));
        double *Bm = (double *)malloc((size_t)(L + 1) * sizeof(double));
#pragma omp for schedule(static)
        for (int j = 0; j <= nTheta; ++j) {
            legendre_table(L, cos((double)j * h), tab);
            for (int m = 0; m <= L; ++m) {
                double a = 0.0, b = 0.0;
                for (int l = (m > 1 ? m : 1); l <= L; ++l) {
                    const double P = tab[l * (L + 1) + m];
                    a += coef[(size_t)l * (2 * L + 1) + (m + L)] * P;
                    if (m > 0) b += coef[(size_t)l * (2 * L + 1) + (-m + L)] * P;
                }
                Am[m] = a;
                Bm[m] = b;
            }
            for (int i = 0; i < nc; ++i) {
                const double phi = ((double)i - 0.5) * h;
                double v = Am[0];
                for (int m = 1; m <= L; ++m) v += sqrt(2.0) * (Am[m] * cos(m * phi) + Bm[m] * sin(m * phi));
                psi[(size_t)j * nc + i] = v;
            }
        }
        free(tab);
        free(Am);
        free(Bm);
    }
    /* psi is 2*pi-periodic and its pole rows are phi-independent in exact
     * arithmetic; make both exact, which keeps the discrete divergence of the
     * curl exactly zero */
    for (int j = 0; j <= nTheta; ++j) psi[(size_t)j * nc + nPhi] = psi[(size_t)j * nc];
    for (int jj = 0; jj <= nTheta; jj += nTheta) {
        double mean = 0.0;
        for (int i = 0; i < nPhi; ++i) mean += psi[(size_t)jj * nc + i];
        mean /= nPhi;
        for (int i = 0; i < nc; ++i) psi[(size_t)jj * nc + i] = mean;
    }
    if (u_phi)
        for (int j = 0; j < nTheta; ++j)
            for (int i = 0; i < nPhi; ++i)
                u_phi[(size_t)j * nPhi + i] =
                    (psi[(size_t)(j + 1) * nc + i] - psi[(size_t)j * nc + i]) / (radius * h);
    if (u_theta)
        for (int j = 0; j < nTheta - 1; ++j)
            for (int i = 0; i < nPhi; ++i)
                u_theta[(size_t)j * nPhi + i] =
                    -(psi[(size_t)(j + 1) * nc + i + 1] - psi[(size_t)(j + 1) * nc + i]) /
                    (radius * h * sin((double)(j + 1) * h));
    if (rho) {
        for (int j = 0; j < nTheta; ++j)
            for (int i = 0; i < nPhi; ++i) {
                double v;
                if (density_kind == 0) {
                    v = 0.5 + 0.1 * ko_uniform(&st);
                } else {
                    const double th = ((double)j + 0.5) * h;
                    v = 0.5 + (0.3 - fabs(th - 0.5 * REF_PI)) / (2.0 * h);
                    v = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
                }
                rho[(size_t)j * nPhi + i] = v;
            }
    }
    free(psi);
    free(coef);
}

/* ---- the reference's default initial velocity ---------------------------
 * KaminoSolver::initialize_velocity and its FBM helpers
 * (KaminoInitializer.cu:3-133).  The noise hashes FP32 lattice coordinates,
 * so this part is restated in the reference's own types, not in FP64: fReal
 * (float) variables, double where a double literal or the vec2 class
 * (vectorUtil.cuh:12, double components) promotes, float on assignment. */
static float ref_rand(float ax, float ay) { /* KaminoInitializer.cu:124-133 */
    const double vecB0 = 12.9898, vecB1 = 4.1414;
    const double a0 = (double)ax, a1 = (double)ay;
    const float dotProd = (float)(a0 * vecB0 + a1 * vecB1);
    const float val = (float)sin((double)dotProd * 43758.5453);
    return val - floorf(val);
}

static float ref_lerp_host(float from, float to, float alpha) { /* :103-106 */
    const double lhs = (1.0 - (double)alpha) * (double)from;
    const float rhs = alpha * to;
    return (float)(lhs + (double)rhs);
}

static float ref_interp_noise(float x, float y) { /* :108-122 */
    const float intX = floorf(x);
    const float fractX = x - intX;
    const float intY = floorf(y);
    const float fractY = y - intY;
    const float v1 = ref_rand(intX, intY);
    const float v2 = ref_rand(intX + 1, intY);
    const float v3 = ref_rand(intX, intY + 1);
    const float v4 = ref_rand(intX + 1, intY + 1);
    const float i1 = ref_lerp_host(v1, v2, fractX);
    const float i2 = ref_lerp_host(v3, v4, fractX);
    return ref_lerp_host(i1, i2, fractY);
}

float ko_fbm(float x, float y) { /* :86-101 */
    float total = 0.0f;
    const float resolutionX = (float)0.15, resolutionY = (float)0.5, persistance = (float)0.5;
    for (int i = 0; i < 4; i++) {
        const float freq = (float)pow(2.0, (double)i);
        const float amp = (float)pow((double)persistance, (double)i);
        total += amp * ref_interp_noise(x * freq / resolutionX, y * freq / resolutionY);
    }
    const float a = 1 - persistance;
    return a * total / 2.0f;
}

float ko_fbm_hash(float x, float y) { return ref_rand(x, y); }

void ko_init_fbm(int nTheta, int nPhi, double radius_d, float *u_phi, float *u_theta) {
    const float gridLen = (float)(REF_2PI / (double)nPhi); /* KaminoSolver.cu:14 */
    const float radius = (float)radius_d;
    const float gain = (float)(4096.0 / (double)nPhi);
    /* u_phi, i >= 1 (:11-34) and the seam i = 0 (:36-55) */
    for (size_t j = 0; j < (size_t)nTheta; ++j)
        for (size_t i = 0; i < (size_t)nPhi; ++i) {
            float ur_x, ul_x;
            if (i == 0) {
                ur_x = gridLen / 2;
                ul_x = (float)(2 * REF_PI - (double)(gridLen / 2));
            } else {
                ur_x = (float)i * gridLen + gridLen / 2;
                ul_x = (float)i * gridLen - gridLen / 2;
            }
            const float lr_x = ur_x, ll_x = ul_x;
            const float ur_y = (float)(j + 1) * gridLen, lr_y = (float)j * gridLen;
            const float ul_y = (float)(j + 1) * gridLen, ll_y = (float)j * gridLen;
            const float noise_ur = ko_fbm(ur_x, ur_y);
            const float noise_lr = ko_fbm(lr_x, lr_y);
            const float noise_ul = ko_fbm(ul_x, ul_y);
            const float noise_ll = ko_fbm(ll_x, ll_y);
            const float noiseDy_l = (noise_ur - noise_lr) / (radius * gridLen);
            const float noiseDy_r = (noise_ul - noise_ll) / (radius * gridLen);
            const float avgNoise = (float)((double)(noiseDy_l + noiseDy_r) / 2.0);
            u_phi[j * (size_t)nPhi + i] = avgNoise * gain;
        }
    /* u_theta (:57-84), including ll_y = j h + h/2 (:69) */
    for (size_t j = 1; j < (size_t)nTheta; ++j)
        for (size_t i = 0; i < (size_t)nPhi; ++i) {
            const float ur_x = (float)(i + 1) * gridLen;
            const float ur_y = (float)j * gridLen + gridLen / 2;
            const float lr_x = (float)(i + 1) * gridLen;
            const float lr_y = (float)j * gridLen - gridLen / 2;
            const float ul_x = (float)i * gridLen;
            const float ul_y = (float)j * gridLen + gridLen / 2;
            const float ll_x = (float)i * gridLen;
            const float ll_y = (float)j * gridLen + gridLen / 2;
            const float noise_ur = ko_fbm(ur_x, ur_y);
            const float noise_lr = ko_fbm(lr_x, lr_y);
            const float noise_ul = ko_fbm(ul_x, ul_y);
            const float noise_ll = ko_fbm(ll_x, ll_y);
            const float noiseDy_u = -1 * (noise_ur - noise_ul) / (radius * gridLen);
            const float noiseDy_d = -1 * (noise_lr - noise_ll) / (radius * gridLen);
            const float avgNoise = (float)((double)(noiseDy_u + noiseDy_d) / 2.0);
            u_theta[(j - 1) * (size_t)nPhi + i] = avgNoise * gain;
        }
}
