/*
 * kamino.h -- C ABI of the MI355X-native KaminoGPU time step (libkamino.so).
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes
 * (KAM_OK = 0, negative on error).  Nothing here exits the process; the text
 * of the last error on the calling thread is returned by kam_last_error().
 * Handles are thread-compatible, not thread-safe.  Caller buffers are host
 * memory owned by the caller, dense theta-major with phi contiguous:
 *     u_phi[nTheta][nPhi], u_theta[nTheta-1][nPhi], rho[nTheta][nPhi]
 * (the reference's KaminoQuantity layout, KaminoQuantity.cu:70-73, without
 * the pitch).  Reference citations are relative to the KaminoGPU checkout.
 */
#ifndef KAMINO_H
#define KAMINO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
enum {
    KAM_OK = 0,
    KAM_EINVAL = -1,   /* bad argument or unsupported shape */
    KAM_ENOMEM = -2,   /* device or host allocation failed */
    KAM_EHIP = -3,     /* a HIP runtime call failed */
    KAM_EAGAIN = -4,   /* engine has no slot / HBM budget for the request now */
    KAM_ENOTFOUND = -5,/* unknown request id, or an input file that cannot be read */
    KAM_ESTATE = -6,   /* call not valid in the object's current state */
    KAM_ENODEV = -7    /* no usable gfx950 device */
};

/* ---- step phases (kam_sim_step_phases) ----------------------------------- */
enum {
    KAM_PHASE_ADVECT = 1,    /* advection (+ the geometric update if enabled) */
    KAM_PHASE_DIVFFT = 8,    /* divergence + row FFT into the spectrum buffer */
    KAM_PHASE_TRI = 16,      /* per-mode tridiagonal solves, in place */
    KAM_PHASE_IFFT = 32,     /* inverse row FFT + gradient subtraction */
    KAM_PHASE_PROJECT = 8 | 16 | 32,
    KAM_PHASE_ALL = 1 | 8 | 16 | 32
};

/* ---- phi transforms (kam_tune_set_transform) ----------------------------- */
enum {
    KAM_TRANSFORM_FFT = 0,  /* the row FFT in LDS (default) */
    KAM_TRANSFORM_DFT = 1,  /* dense DFT products on MFMA: FP16 hi/lo operands, FP32
                               accumulate (FP32-pressure simulations, n_theta <= 1024) */
    KAM_TRANSFORM_FFT8 = 2  /* the row FFT with a forward kernel that stages half the
                               twiddle table and runs eight waves per SIMD; a batch takes
                               it when all its FFT simulations do.  Same values as FFT. */
};

/* ---- precision modes ----------------------------------------------------- */
enum {
    KAM_PREC_FP32 = 0,          /* FP32 fields, advection and FFTs; FP64 tridiagonal solves */
    KAM_PREC_FP64_PRESSURE = 1  /* as FP32, but the solved pressure spectrum is kept in FP64
                                   and the inverse FFT and gradient subtraction run in FP64,
                                   rounding each updated velocity once: the post-projection
                                   divergence then sits at the FP32 storage floor */
};

/* Shape and parameters of one simulation.  Replaces the KaminoSolver ctor
 * arguments (KaminoSolver.cuh:99-100) and the __constant__ grid uploads the
 * caller had to make (KaminoCore.cu:868-872): nPhi = 2 nTheta, h = pi/nTheta
 * (KaminoCore.cu:849). */
typedef struct kam_sim_desc {
    int n_theta;     /* latitude rows, 4 ... 2048, of the form 2^a 3^b 5^c (1024 at most
                        with KAM_PREC_FP64_PRESSURE) */
    int n_phi;       /* must equal 2 * n_theta */
    double radius;   /* sphere radius R */
    double dt;       /* time step (constant, as timeStepGlobal, KaminoCore.cu:871) */
    int geometric;   /* 1: the reference step (advection, geometric, projection,
                        KaminoSolver.cu:197-221); 0: advection + projection only */
    int precision;   /* KAM_PREC_* */
    int device;      /* HIP device ordinal */
} kam_sim_desc;

typedef struct kam_sim kam_sim;

/* Text of the last error on this thread ("" if none). */
const char *kam_last_error(void);

/* Library version string. */
const char *kam_version(void);

/* Number of gfx950 devices visible (0 if none). */
int kam_device_count(void);

/* Creates one simulation with its own stream and device buffers (28 bytes per
 * cell, plus 256 bytes).  Replaces KaminoSolver::KaminoSolver (KaminoSolver.cu:12-67). */
int kam_sim_create(const kam_sim_desc *desc, kam_sim **out);

/* Frees the simulation.  Replaces KaminoSolver::~KaminoSolver
 * (KaminoSolver.cu:69-104), without its cudaDeviceReset. */
int kam_sim_destroy(kam_sim *sim);

/* Uploads the state (host, FP32).  Replaces KaminoQuantity::copyToGPU
 * (KaminoQuantity.cu:3-12) for u_phi, u_theta and rho.  rho may be NULL. */
int kam_sim_set_fields(kam_sim *sim, const float *u_phi, const float *u_theta, const float *rho);

/* Downloads the state after all queued steps.  Replaces
 * KaminoQuantity::copyBackToCPU (KaminoQuantity.cu:14-18).  Any pointer may
 * be NULL. */
int kam_sim_get_fields(kam_sim *sim, float *u_phi, float *u_theta, float *rho);

/* Queues nsteps full steps on the simulation's stream.  Replaces
 * KaminoSolver::stepForward (KaminoSolver.cu:197-221); like the reference's
 * kernels, every step uses the creation-time dt. */
int kam_sim_step(kam_sim *sim, int nsteps);

/* Queues one step restricted to the KAM_PHASE_* bits in `phases`, applied in
 * the order advect, div+FFT, tridiagonal, inverse FFT + gradient.  Projection
 * phases act on the current state in place. */
int kam_sim_step_phases(kam_sim *sim, int phases);

/* Copies the half spectrum buffer (float2 [M = nPhi/2][nTheta],
 * mode-major; slot 0 holds (mode 0, Nyquist)).  After KAM_PHASE_DIVFFT it
 * holds d-hat, after KAM_PHASE_TRI p-hat (with KAM_PREC_FP64_PRESSURE the
 * solve stores p-hat in FP64 elsewhere and this buffer keeps d-hat). */
int kam_sim_get_spectrum(kam_sim *sim, float *spec);

/* Pressure of the last projection (host, float p[nTheta][nPhi], the
 * reference's pressure quantity, KaminoSolver.cu:54, as shiftUKernel leaves
 * it, KaminoCore.cu:682-703: the n = 0 mode removed): the inverse transform
 * of the solved spectrum, recomputed on demand.  Valid after a full step or
 * KAM_PHASE_TRI, until the next step's advection (FP64 pressure) or
 * KAM_PHASE_DIVFFT (FP32 pressure).  The reference keeps p on the device
 * and never writes it out; the first call allocates a device buffer of 4
 * bytes per cell. */
int kam_sim_get_pressure(kam_sim *sim, float *p);

/* Runs nsteps full steps and times them with HIP events on the simulation's
 * stream.  step_ms[nsteps] (optional) receives each step's duration,
 * kernel_ms[4] (optional, needs per_kernel = 1) the mean per-step duration of
 * the advection, div+FFT, tridiagonal and inverse-FFT+gradient kernels, and
 * total_ms the whole interval.  per_kernel = 1 records events between the
 * kernels and waits for each step; use 0 for throughput. */
int kam_sim_benchmark(kam_sim *sim, int nsteps, int per_kernel, double *step_ms,
                      double *kernel_ms, double *total_ms);

/* Solves the tridiagonal systems of the current spectrum in place, exactly as
 * KAM_PHASE_TRI does (call it after KAM_PHASE_DIVFFT), and reports the
 * relative Poisson residual ||T_n p_n - d_n|| / ||d_n|| of each mode's
 * solution (the FP32 elimination of the FP32 pressure, the FP64 one of the
 * FP64 pressure), formed in FP64 before the store: per_mode[M = n_phi/2]
 * (optional; slot 0 is the Nyquist mode), their maximum and root mean square.
 * A diagnostic of the solve (no counterpart in the reference, whose crKernel
 * solves in FP32, tdm.cu:3-96). */
int kam_sim_poisson_residual(kam_sim *sim, double *per_mode, double *max_rel, double *rms_rel);

/* Waits for all queued work of the simulation. */
int kam_sim_sync(kam_sim *sim);

/* Watchdog word after all queued steps: 0 = healthy, bit 0 set once a step
 * produced a non-finite value (no counterpart in the reference). */
int kam_sim_status(kam_sim *sim, int *status);

/* The reference's default initial velocity (KaminoSolver::initialize_velocity,
 * KaminoInitializer.cu:3-133, called from the solver's constructor at
 * KaminoSolver.cu:59): the curl of four-octave FBM value noise, gain
 * 4096 / n_phi, with its quirks kept (the separate phi = 0 seam loop :36-55;
 * ll_y = ul_y in the u_theta loop, :69).  Host, FP32, computed in the
 * reference's FP32/FP64 mix (the noise hashes FP32 coordinates).  Either
 * output may be NULL. */
int kam_init_fbm(int n_theta, int n_phi, double radius, float *u_phi, float *u_theta);

/* Particles (KaminoParticles.cu, advectionParticles KaminoCore.cu:321-342).
 * Sets n particles as (phi, theta) pairs in radians (host, FP32, 2n floats,
 * the layout of KaminoParticles::coordCPUBuffer); every later step moves them
 * one forward-Euler step with the pre-step velocities, inside the advection
 * phase.  n = 0 removes them.  Needs 16 bytes of device memory per particle.
 * Replaces KaminoParticles::KaminoParticles's upload (KaminoParticles.cu:3-83,
 * copy2GPU :94-101). */
int kam_sim_set_particles(kam_sim *sim, const float *coords, int64_t n);

/* Downloads the n = current particle count (phi, theta) pairs after all
 * queued steps.  Replaces KaminoParticles::copyBack2CPU
 * (KaminoParticles.cu:84-91). */
int kam_sim_get_particles(kam_sim *sim, float *coords, int64_t n);

/* Particle count of the simulation (0 if none). */
int64_t kam_sim_particle_count(const kam_sim *sim);

/* Seeded particle layout of KaminoParticles.cu:20-80: sqrt(density) * n_theta
 * particles per meridian and twice that per latitude, particle (i, j) at
 * (i delta, j delta) + a jitter of up to delta / 2 in each coordinate with a
 * random sign, negative coordinates clamped to 0, stored at i * n_j + j.
 * The reference draws with C rand(), whose sequence is platform-defined; this
 * draws the same four numbers per particle from SplitMix64(seed).  coords may
 * be NULL to query *count. */
int kam_particles_seed(int n_theta, double density, uint64_t seed, float *coords, int64_t *count);

/* Image inputs.  The reference reads them with OpenCV 2.4 (imread, flip around
 * the vertical axis, resize to n_phi x n_theta, INTER_LINEAR); here images are
 * PPM / PGM files (P6, P3, P5, P2) and the flip and bilinear resize restate
 * OpenCV's 8-bit fixed-point algorithm.  KAM_ENOTFOUND when the file cannot
 * be read.
 * kam_image_density: rho[n_theta][n_phi] = (B + G + R) / 3 of the resized
 * image over 255, as KaminoSolver::initDensityfromPic (KaminoSolver.cu:243-277).
 * kam_image_colors: colors[3n] = (G, R, B) / 255 of the pixel under each
 * particle's (phi, theta), the reference's colorBGR (KaminoParticles.cu:62-76);
 * positions outside the image clamp to its edge. */
int kam_image_density(const char *path, int n_theta, int n_phi, float *rho);
int kam_image_colors(const char *path, int n_theta, int n_phi, const float *coords, int64_t n,
                     float *colors);

/* Bytes of device memory one simulation of this shape needs (28 per cell + 256). */
size_t kam_sim_bytes(int n_theta, int n_phi);

/* Seeded synthetic initial state (host, FP32): the discrete curl of a real
 * spherical-harmonic streamfunction with coefficients c_lm ~ N(0, 0.1/l),
 * l = 1..L, drawn by SplitMix64 + Box-Muller, and a density of
 * 0.5 + 0.1 U[0,1) (density_kind 0) or a band around the equator (1).  No
 * counterpart in the reference, whose initialisers are FBM noise and images
 * (KaminoInitializer.cu:3-134, KaminoSolver.cu:243-277).  Any output may be
 * NULL. */
int kam_init_fields(int n_theta, int n_phi, double radius, int L, uint64_t seed,
                    int density_kind, float *u_phi, float *u_theta, float *rho);

/* ==== Stepping without allocation ========================================
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Runs nsteps full steps (as kam_sim_step, the same
 * kernels) of one simulation whose fields are DEVICE buffers owned by the
 * caller, dense theta-major as above (u_theta has nTheta - 1 rows), updated
 * in place; rho may be NULL (zero density, not written back).  Work is
 * queued on `stream` (a hipStream_t; NULL = the null stream).  Intermediate
 * state, the spectrum, tables and tile lists live in one fixed scratch array
 * per device of KAM_NOALLOC_SCRATCH_BYTES, which each device that runs the
 * library holds until the process ends; shapes that do not fit (n_theta
 * above about 512) return KAM_EINVAL.  The call allocates nothing: no
 * hipMalloc / hipFree and no host heap.  The first call for a shape on a
 * device uploads its tables and waits for that upload on `stream`; later
 * calls with the same shape queue only device work.  Every call on a device
 * shares its scratch array: calls that overlap on one device (other threads,
 * or other streams with work still queued) race. */
#define KAM_NOALLOC_SCRATCH_BYTES (16u << 20)
int kam_sim_step_noalloc(const kam_sim_desc *desc, float *u_phi, float *u_theta, float *rho,
                         int nsteps, void *stream);

/* ==== Batching and serving engine ========================================
 * No counterpart in the reference, which runs one simulation per process
 * (KaminoSolver.cu:20 hard-codes device 0 and the dtor resets the device).
 * An engine owns one stream and an HBM arena; simulations of any supported
 * shape are admitted into slots while the arena and slot count allow, and
 * every kam_engine_step advances all active simulations together with one
 * batched launch per kernel (captured in a hipGraph per step parity).  A
 * request retires after its own step count; its final fields are copied to
 * pinned host memory on the engine stream and handed out by
 * kam_engine_fetch. */

typedef struct kam_engine kam_engine;

typedef struct kam_engine_desc {
    int device;
    int max_slots;            /* concurrently active simulations (0 = 4096) */
    size_t hbm_budget_bytes;  /* arena for simulation state (0 = half of free HBM) */
    int use_graph;            /* 1: replay steps from hipGraphs */
    int max_inflight;         /* launches (steps or graphs) queued ahead of the GPU at most (0 = 2) */
    int graph_steps;          /* with use_graph: G steps per graph (0 = 1, at most 64).  G = 1
                                 replays every step from a one-step graph.  G > 1: a run of G steps
                                 in which no simulation finishes is one graph launch, the other
                                 steps are direct launches.  Results equal direct launches. */
    int groups;               /* concurrent groups (0 = 1, at most 4): the active simulations are split
                                 into groups of about equal cell count, each stepped on a stream of
                                 its own so that their kernels overlap; results are unchanged */
} kam_engine_desc;

/* KAM_REQ_FAILED: quarantined because a step produced a non-finite value in
 * that simulation (checked every 8 steps); the others are unaffected. */
enum { KAM_REQ_ACTIVE = 1, KAM_REQ_DONE = 2, KAM_REQ_CANCELLED = 3, KAM_REQ_FAILED = 4 };

typedef struct kam_req_status {
    int state;              /* KAM_REQ_* */
    int reserved;
    int64_t steps_done;
    int64_t steps_total;
    double admit_s;         /* host clock (s since engine creation) at admission */
    double done_s;          /* host clock when its last step was queued; 0 if not yet */
} kam_req_status;

typedef struct kam_stats {
    int active;              /* simulations in the batch now */
    int reserved;
    int64_t steps;           /* engine steps since the last reset */
    int64_t cell_steps;      /* cells advanced by those steps */
    double gpu_ms;           /* GPU time of those steps (HIP events) */
    double step_p50_ms;      /* per-step GPU latency percentiles over those steps */
    double step_p99_ms;
    double step_max_ms;
    int64_t admitted, completed, rejected, cancelled;
    size_t bytes_in_use, bytes_budget;
    int64_t failed;          /* requests quarantined by the non-finite watchdog */
} kam_stats;

int kam_engine_create(const kam_engine_desc *desc, kam_engine **out);
int kam_engine_destroy(kam_engine *eng);

/* Admits one simulation with `steps` steps to run.  Fields are host FP32 in
 * the layout above (rho may be NULL: zero density).  Returns KAM_EAGAIN when
 * no slot or arena space is free now; the caller retries after steps. */
int kam_engine_admit(kam_engine *eng, const kam_sim_desc *desc, const float *u_phi,
                     const float *u_theta, const float *rho, int64_t steps, uint64_t *id);

/* Runs nsteps shared steps of every active simulation (queued; returns once
 * launched, except when membership changed and tables were re-uploaded). */
int kam_engine_step(kam_engine *eng, int nsteps);

int kam_engine_poll(kam_engine *eng, uint64_t id, kam_req_status *status);

/* Ids of requests whose final fields reached the host since the last call
 * (up to max; *count receives how many were written). */
int kam_engine_completed(kam_engine *eng, uint64_t *ids, int max, int *count);

/* Copies a finished request's fields (waits for its copy) and forgets it. */
int kam_engine_fetch(kam_engine *eng, uint64_t id, float *u_phi, float *u_theta, float *rho);

/* Removes an active request and frees its slot. */
int kam_engine_cancel(kam_engine *eng, uint64_t id);

/* Waits for all queued engine work. */
int kam_engine_sync(kam_engine *eng);

/* Statistics since the last kam_engine_reset_stats (synchronises). */
int kam_engine_stats(kam_engine *eng, kam_stats *stats);
int kam_engine_reset_stats(kam_engine *eng);

/* Runs nsteps steps without graphs, with HIP events around each kernel on the
 * engine stream; kernel_ms[4] receives the mean per-step time of the
 * advection, div+FFT, tridiagonal and inverse-FFT+gradient launches. */
int kam_engine_profile(kam_engine *eng, int nsteps, double *kernel_ms);

/* Host clock of the engine (s since creation), the time base of the request
 * timestamps. */
double kam_engine_now(kam_engine *eng);

/* ==== Autotuning (no counterpart in the reference, whose launch shape is
 * fixed by determineLayout, KaminoSolver.cu:165-180) =======================
 * Tile shapes per n_theta: advection rows x columns and FFT rows.  They
 * regroup work only and never change a computed value. */

/* Records tile shapes for n_theta (0 keeps a default); later plans use them. */
int kam_tune_set(int n_theta, int adv_rows, int adv_cols, int fft_rows);

/* The shapes a simulation of n_theta would get now. */
int kam_tune_get(int n_theta, int *adv_rows, int *adv_cols, int *fft_rows);

/* Records the phi transform (KAM_TRANSFORM_*) for simulations of n_theta
 * created or admitted from now on; kam_autotune picks it by timing them.
 * The reference's transform is cuFFT (KaminoCore.cu:763-764, 812-813;
 * plan KaminoSolver.cu:61-66).  The transform never changes results beyond
 * rounding (both are FP32-accurate). */
int kam_tune_set_transform(int n_theta, int transform);
int kam_tune_get_transform(int n_theta, int *transform);

/* Times candidate shapes on `batch` synthetic simulations of n_theta (steps
 * timed steps each): advection tiles, FFT tile heights, then the two row FFT
 * forms against each other and against the MFMA DFT (n_theta <= 1024).  Records the fastest with
 * kam_tune_set and kam_tune_set_transform, appends a line
 * "gfx950 n_theta batch adv_rows adv_cols fft_rows adv_ms fft_ms transform"
 * to cache_path (may be NULL), and returns the best per-kernel times in
 * kernel_ms[4] (advection, div+transform, tridiagonal, inverse
 * transform+gradient). */
int kam_autotune(int device, int n_theta, int batch, int steps, const char *cache_path,
                 double *kernel_ms);

/* Loads gfx950 entries of a cache written by kam_autotune. */
int kam_tune_load(const char *cache_path);

#ifdef __cplusplus
}
#endif
#endif /* KAMINO_H */
