// kamino_init.cpp -- seeded synthetic initial conditions (host, FP64 -> FP32).
//
// The workloads define their own initial state, since the reference's image-
// and FBM-driven initialisers are not part of the step:
//   * SplitMix64 seeded with `seed`; normals by Box-Muller (cosine branch, one
//     normal per pair of uniforms); the coefficients are drawn first
//     (l = 1..L, m = -l..l), then any density noise, row by row;
//   * streamfunction psi = sum c_lm Y_lm with orthonormal real spherical
//     harmonics (no Condon-Shortley phase) and c_lm ~ N(0, 0.1 / l);
//   * velocity = discrete curl of psi sampled at the cell corners
//     (theta = j h, phi = (i - 1/2) h):
//       u_phi[j][i]   =  (psi[j+1][i]   - psi[j][i])   / (R h)
//       u_theta[j][i] = -(psi[j+1][i+1] - psi[j+1][i]) / (R h sin((j+1) h)),
//     which the reference's divergence (KaminoCore.cu:587-638) maps to zero;
//   * density: 0.5 + 0.1 U[0,1) per cell, or the band rho = 1 for
//     |theta - pi/2| < 0.3 with a linear ramp over two cells.
#include <cmath>
#include <cstdint>
#include <vector>

#include "kamino.h"

namespace {

const double kPi = 3.14159265358979323846;

struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {
        const double a = 1.0 - uniform();  // (0, 1]
        const double b = uniform();
        return std::sqrt(-2.0 * std::log(a)) * std::cos(2.0 * kPi * b);
    }
};

// Pbar[m][l] = N_lm P_l^m(x) for 0 <= m <= l <= L (no Condon-Shortley phase),
// by the standard three-term recurrence in l at fixed m.
void assoc_legendre(int L, double x, std::vector<double>& P) {
    P.assign((size_t)(L + 1) * (L + 1), 0.0);
    const double s = std::sqrt(std::max(0.0, 1.0 - x * x));
    double diag = 1.0 / std::sqrt(4.0 * kPi);
    for (int m = 0; m <= L; ++m) {
        if (m > 0) diag *= s * std::sqrt((2.0 * m + 1.0) / (2.0 * m));
        double* row = &P[(size_t)m * (L + 1)];
        row[m] = diag;
        if (m < L) row[m + 1] = x * std::sqrt(2.0 * m + 3.0) * diag;
        for (int l = m + 2; l <= L; ++l) {
            const double ll = (double)l * l, mm = (double)m * m, l1 = (double)(l - 1) * (l - 1);
            const double a = std::sqrt((4.0 * ll - 1.0) / (ll - mm));
            const double b = std::sqrt((l1 - mm) / (4.0 * l1 - 1.0));
            row[l] = a * (x * row[l - 1] - b * row[l - 2]);
        }
    }
}

}  // namespace

extern "C" int kam_init_fields(int n_theta, int n_phi, double radius, int L, uint64_t seed,
                               int density_kind, float* u_phi, float* u_theta, float* rho) {
    if (n_theta < 2 || n_phi < 2 || L < 1 || !(radius > 0.0)) return KAM_EINVAL;
    const double h = kPi / n_theta;
    SplitMix rng{seed};
    // coefficient c(l, m) at [l][m + L]
    std::vector<double> c((size_t)(L + 1) * (2 * L + 1), 0.0);
    for (int l = 1; l <= L; ++l)
        for (int m = -l; m <= l; ++m) c[(size_t)l * (2 * L + 1) + m + L] = (0.1 / l) * rng.normal();

    const int nc = n_phi + 1;  // corner columns, the last one equal to the first
    std::vector<double> psi((size_t)(n_theta + 1) * nc);
    std::vector<double> P, A(L + 1), B(L + 1);
    for (int j = 0; j <= n_theta; ++j) {
        assoc_legendre(L, std::cos(j * h), P);
        for (int m = 0; m <= L; ++m) {
            double a = 0.0, b = 0.0;
            for (int l = std::max(m, 1); l <= L; ++l) {
                const double p = P[(size_t)m * (L + 1) + l];
                a += c[(size_t)l * (2 * L + 1) + m + L] * p;
                if (m > 0) b += c[(size_t)l * (2 * L + 1) - m + L] * p;
            }
            A[m] = a;
            B[m] = b;
        }
        for (int i = 0; i < n_phi; ++i) {
            const double phi = (i - 0.5) * h;
            double v = A[0];
            for (int m = 1; m <= L; ++m) v += std::sqrt(2.0) * (A[m] * std::cos(m * phi) + B[m] * std::sin(m * phi));
            psi[(size_t)j * nc + i] = v;
        }
        psi[(size_t)j * nc + n_phi] = psi[(size_t)j * nc];
    }
    // the pole rows are phi-independent exactly
    for (int j : {0, n_theta}) {
        double mean = 0.0;
        for (int i = 0; i < n_phi; ++i) mean += psi[(size_t)j * nc + i];
        mean /= n_phi;
        for (int i = 0; i < nc; ++i) psi[(size_t)j * nc + i] = mean;
    }
    if (u_phi)
        for (int j = 0; j < n_theta; ++j)
            for (int i = 0; i < n_phi; ++i)
                u_phi[(size_t)j * n_phi + i] =
                    (float)((psi[(size_t)(j + 1) * nc + i] - psi[(size_t)j * nc + i]) / (radius * h));
    if (u_theta)
        for (int j = 0; j + 1 < n_theta; ++j) {
            const double sn = std::sin((j + 1) * h);
            for (int i = 0; i < n_phi; ++i)
                u_theta[(size_t)j * n_phi + i] =
                    (float)(-(psi[(size_t)(j + 1) * nc + i + 1] - psi[(size_t)(j + 1) * nc + i]) /
                            (radius * h * sn));
        }
    if (rho)
        for (int j = 0; j < n_theta; ++j)
            for (int i = 0; i < n_phi; ++i) {
                double v;
                if (density_kind == 0) {
                    v = 0.5 + 0.1 * rng.uniform();
                } else {
                    const double th = (j + 0.5) * h;
                    v = 0.5 + (0.3 - std::fabs(th - 0.5 * kPi)) / (2.0 * h);
                    v = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
                }
                rho[(size_t)j * n_phi + i] = (float)v;
            }
    return KAM_OK;
}
