/* Exercises the FP64 oracle under AddressSanitizer and UBSan (built by
 * tests/test_sanitizers.py with gcc -fsanitize=address,undefined): seeded
 * steps at small and odd sizes, the pole-escape path (departure points past
 * a pole), particles and the FBM initializer.  Exit status 0 = clean. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "kamino_oracle.h"

static int run_case(int nTheta, double dt, int L) {
    const int nPhi = 2 * nTheta;
    const size_t N = (size_t)nTheta * nPhi;
    double *up = calloc(N, sizeof(double)), *ut = calloc(N, sizeof(double)), *rho = calloc(N, sizeof(double));
    if (!up || !ut || !rho) return 1;
    ko_init_fields(nTheta, nPhi, 1.0, L, 7, 1, up, ut, rho);
    ko_stats st;
    memset(&st, 0, sizeof st);
    for (int k = 0; k < 3; ++k)
        if (ko_step(nTheta, nPhi, 1.0, dt, KO_PHASE_ALL, up, ut, rho, &st) != 0) return 2;
    /* particles across the whole sphere, poles included */
    const int64_t np = 64;
    double pos[2 * 64], out[2 * 64];
    uint8_t amb[64];
    for (int64_t k = 0; k < np; ++k) {
        pos[2 * k] = 6.283185307179586 * (double)k / (double)np;
        pos[2 * k + 1] = 3.141592653589793 * (double)(k % 17) / 16.0;
    }
    ko_advect_particles(nTheta, nPhi, 1.0, dt, up, ut, np, pos, out, amb, 1e-5);
    double s = 0.0;
    for (size_t i = 0; i < N; ++i) s += up[i] + rho[i];
    for (int64_t k = 0; k < 2 * np; ++k) s += out[k];
    free(up);
    free(ut);
    free(rho);
    return isfinite(s) ? 0 : 3;
}

int main(void) {
    /* dt large enough that departure points cross the poles (escapes) */
    const int sizes[] = {4, 6, 8, 12, 30, 32};
    for (size_t k = 0; k < sizeof sizes / sizeof sizes[0]; ++k) {
        int rc = run_case(sizes[k], 0.5 / sizes[k], 3);
        if (!rc) rc = run_case(sizes[k], 4.0 / sizes[k], 3);
        if (rc) {
            fprintf(stderr, "case nTheta=%d failed: %d\n", sizes[k], rc);
            return rc;
        }
    }
    float fu[8 * 16], fv[7 * 16];
    ko_init_fbm(8, 16, 1.0, fu, fv);
    printf("oracle sanitizer run clean\n");
    return 0;
}
