/*
 * kamino_oracle.h -- FP64 host restatement of the KaminoGPU time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker and the CPU
 * baseline for the HIP path in kaminogpu_amd/.  Only tests/, the smoke entry
 * point and bench.py's CPU-baseline leg may load it; the product path never
 * calls it and never falls back to it.
 *
 * Every function restates a reference kernel body in double precision, in the
 * reference's own radian coordinates and branch order (file:line citations are
 * relative to the KaminoGPU checkout).  Arrays are dense theta-major with phi
 * contiguous: u_phi[nTheta][nPhi], u_theta[nTheta-1][nPhi], rho[nTheta][nPhi]
 * (KaminoQuantity.cu:70-73, KaminoSolver.cu:50-57).
 */
#ifndef KAMINO_ORACLE_H
#define KAMINO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field kinds, in the order of the reference samplers. */
enum { KO_VPHI = 0, KO_VTHETA = 1, KO_CENTERED = 2 };

/* Phase bits for ko_step. */
enum { KO_PHASE_ADVECT = 1, KO_PHASE_GEOMETRIC = 2, KO_PHASE_PROJECT = 4, KO_PHASE_ALL = 7 };

typedef struct ko_stats {
    int64_t pole_escapes;     /* reads that fell outside a field and were clamped */
    int64_t ambiguous_uphi;   /* advected outputs flagged branch-ambiguous */
    int64_t ambiguous_utheta;
    int64_t ambiguous_rho;
} ko_stats;

/* ---- elementary pieces, exposed for tests ------------------------------ */

/* validateCoord (KaminoCore.cu:11-29): folds (phi, theta) in place, returns +-1. */
double ko_validate_coord(double *phi, double *theta);

/* Bilinear samplers (KaminoCore.cu:36-184).  kind selects sampleVPhi /
 * sampleVTheta / sampleCentered.  rows = rows of `field`. */
double ko_sample(int kind, const double *field, int nTheta, int nPhi, double h,
                 double phiRaw, double thetaRaw, int64_t *escapes);

/* solveCubic(0, b, c) with the reference's Cardano branches and Newton cube
 * root (KaminoCore.cu:386-454). */
double ko_solve_cubic(double a, double b, double c);

/* Poisson tridiagonal rows of Fourier mode n (KaminoSolver.cu:117-163):
 * a = sub-diagonal, b = diagonal, c = super-diagonal, length nTheta. */
void ko_poisson_rows(int nTheta, int nPhi, int n, double *a, double *b, double *c);

/* Thomas algorithm (exact FP64 tridiagonal solve). */
void ko_thomas(int n, const double *a, const double *b, const double *c,
               const double *d, double *x);

/* Cyclic reduction as in crKernel (tdm.cu:3-96), FP64; n must be a power of 2. */
int ko_cyclic_reduction(int n, const double *a, const double *b, const double *c,
                        const double *d, double *x);

/* In-place complex DFT of any length n >= 1, unnormalised,
 * X[k] = sum_j x[j] exp(sign * 2*pi*i*j*k/n), sign = -1 or +1. */
int ko_fft(int n, double *re, double *im, int sign);

/* ---- initial conditions (synthetic inputs) ----------------------------- */

uint64_t ko_splitmix64(uint64_t *state);
double ko_uniform(uint64_t *state);      /* [0,1), 53 bits */
double ko_normal(uint64_t *state);       /* Box-Muller, cosine branch only */

/* Orthonormal real spherical harmonic, no Condon-Shortley phase. */
double ko_real_sh(int l, int m, double theta, double phi);

/* Seeded curl-of-streamfunction velocity and density.
 * density_kind: 0 = 0.5 + 0.1*U[0,1) noise (drawn after the coefficients),
 *               1 = band rho=1 for |theta-pi/2|<0.3, 2-cell linear ramp.
 * Any output pointer may be NULL. */
void ko_init_fields(int nTheta, int nPhi, double radius, int L, uint64_t seed,
                    int density_kind, double *u_phi, double *u_theta, double *rho);

/* ---- step phases ------------------------------------------------------- */

/* Advection of the three fields from the pre-step state (KaminoCore.cu:186-319,
 * 344-384).  Outputs are separate arrays.  amb_* (optional, uint8 per output)
 * flags outputs whose non-node samples fell within `tol` cells of a sampler
 * branch discontinuity. */
void ko_advect(int nTheta, int nPhi, double radius, double dt,
               const double *u_phi, const double *u_theta, const double *rho,
               double *u_phi_out, double *u_theta_out, double *rho_out,
               uint8_t *amb_phi, uint8_t *amb_theta, uint8_t *amb_rho,
               double tol, ko_stats *stats);

/* One particle step (advectionParticles, KaminoCore.cu:321-342): pos and out
 * hold n (phi, theta) pairs in radians; amb (optional, one byte per particle)
 * flags samples within tol cells of a sampler branch discontinuity. */
void ko_advect_particles(int nTheta, int nPhi, double radius, double dt, const double *u_phi,
                         const double *u_theta, int64_t n, const double *pos, double *out,
                         uint8_t *amb, double tol);

/* Geometric (metric-term) update, in place (KaminoCore.cu:457-583). */
void ko_geometric(int nTheta, int nPhi, double radius, double dt,
                  double *u_phi, double *u_theta);

/* Dilate advected-field ambiguity masks through the geometric stencil. */
void ko_geometric_mask(int nTheta, int nPhi, const uint8_t *amb_phi,
                       const uint8_t *amb_theta, uint8_t *out_phi,
                       uint8_t *out_theta);

/* fillDivergenceKernel (KaminoCore.cu:587-638): div[nTheta][nPhi]. */
void ko_divergence(int nTheta, int nPhi, const double *u_phi,
                   const double *u_theta, double *div);

/* Pressure projection, in place on u (KaminoCore.cu:749-842).
 * Optional outputs: spec_re/im[nPhi][nTheta] = d-hat in the reference's
 * mode-major shifted order (nIdx = n + nPhi/2), p[nTheta][nPhi] = pressure.
 * solver: 0 = Thomas, 1 = cyclic reduction (power-of-two nTheta only).
 * Returns 0 on success. */
int ko_project(int nTheta, int nPhi, double *u_phi, double *u_theta,
               double *spec_re, double *spec_im, double *p, int solver);

/* One full step in place: phases is a KO_PHASE_* mask. */
int ko_step(int nTheta, int nPhi, double radius, double dt, int phases,
            double *u_phi, double *u_theta, double *rho, ko_stats *stats);

/* Runs `steps` full steps with `threads` OpenMP threads (0 = default) and
 * returns the wall seconds spent in the stepping loop. */
double ko_run(int nTheta, int nPhi, double radius, double dt, int phases,
              int steps, int threads, double *u_phi, double *u_theta,
              double *rho);

int ko_max_threads(void);

/* The reference's default initial velocity (KaminoInitializer.cu:3-133):
 * FBM value-noise curl, restated in the reference's FP32/FP64 mix because
 * the noise hashes FP32 coordinates.  ko_fbm is FBM(x, y), ko_fbm_hash the
 * lattice hash rand(vec2(x, y)). */
float ko_fbm(float x, float y);
float ko_fbm_hash(float x, float y);
void ko_init_fbm(int nTheta, int nPhi, double radius, float *u_phi, float *u_theta);

#ifdef __cplusplus
}
#endif
#endif
