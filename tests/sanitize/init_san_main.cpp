// Exercises the library's host-side C++ (kamino_init.cpp: seeded fields, the
// FBM initializer, particle seeding, the PPM / PGM reader and resampler)
// under AddressSanitizer and UBSan (built by tests/test_sanitizers.py with
// g++; no device code).  Arguments: image files to read, including malformed
// ones, which must be rejected without a memory error.  Exit status 0 = clean.
#include <cstdio>
#include <vector>

#include "kamino.h"

int main(int argc, char** argv) {
    for (int nt : {4, 6, 16, 48}) {
        const int np = 2 * nt;
        std::vector<float> up((size_t)nt * np), ut((size_t)(nt - 1) * np), rho((size_t)nt * np);
        if (kam_init_fields(nt, np, 1.0, 4, 3, 0, up.data(), ut.data(), rho.data()) != KAM_OK) return 1;
        if (kam_init_fbm(nt, np, 1.0, up.data(), ut.data()) != KAM_OK) return 2;
        int64_t n = 0;
        if (kam_particles_seed(nt, 3.0, 5, nullptr, &n) != KAM_OK) return 3;
        std::vector<float> coords((size_t)n * 2), colors((size_t)n * 3);
        if (kam_particles_seed(nt, 3.0, 5, coords.data(), &n) != KAM_OK) return 4;
        for (int a = 1; a < argc; ++a) {
            const int rc = kam_image_density(argv[a], nt, np, rho.data());
            const int rc2 = kam_image_colors(argv[a], nt, np, coords.data(), n, colors.data());
            std::printf("%s nTheta=%d density %d colors %d\n", argv[a], nt, rc, rc2);
        }
    }
    std::printf("host sanitizer run clean\n");
    return 0;
}
