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
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

// The FBM noise below is a hash of exact FP32 / FP64 operation sequences; no
// multiply-add may be fused.
#pragma STDC FP_CONTRACT OFF

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

// Particle layout of KaminoParticles.cu:20-80 with SplitMix64 in place of C
// rand(): per particle the phi sign, the theta sign, the phi jitter and the
// theta jitter, in that order.  The reference keeps these in fReal (FP32).
extern "C" int kam_particles_seed(int n_theta, double density, uint64_t seed, float* coords,
                                  int64_t* count) {
    if (n_theta < 2 || !(density >= 0.0) || !count) return KAM_EINVAL;
    const float linearDensity = std::sqrt((float)density);
    const float delta = (float)(kPi / n_theta / linearDensity);
    const float halfDelta = (float)(delta / 2.0);
    const unsigned nTh = (unsigned)(linearDensity * (float)n_theta);
    const unsigned nPh = 2 * nTh;
    *count = (int64_t)nTh * nPh;
    if (!coords) return KAM_OK;
    SplitMix rng{seed};
    for (unsigned i = 0; i < nPh; ++i)
        for (unsigned j = 0; j < nTh; ++j) {
            const float signPhi = (float)rng.uniform() >= 0.5f ? 1.f : -1.f;
            const float signTheta = (float)rng.uniform() >= 0.5f ? 1.f : -1.f;
            const float randPhi = signPhi * halfDelta * (float)rng.uniform();
            const float randTheta = signTheta * halfDelta * (float)rng.uniform();
            float phi = (float)i * delta + randPhi;
            float theta = (float)j * delta + randTheta;
            if (phi < 0.f) phi = 0.f;
            if (theta < 0.f) theta = 0.f;
            const size_t idx = (size_t)i * nTh + j;
            coords[2 * idx] = phi;
            coords[2 * idx + 1] = theta;
        }
    return KAM_OK;
}

// ---------------------------------------------------------------------------
// The reference's default initial velocity (KaminoSolver::initialize_velocity,
// KaminoInitializer.cu:3-88): curl of FBM value noise.  The noise is a hash
// of FP32 lattice coordinates, so every operation keeps the reference's types:
// fReal (FP32) variables, double where a double literal or the vec2 class
// (vectorUtil.cuh: double components) promotes, and FP32 again on assignment.
// ---------------------------------------------------------------------------
namespace {

// KaminoSolver::rand (KaminoInitializer.cu:124-133): frac(sin(43758.5453 a.b))
float fbm_hash(float x, float y) {
    const double dot = (double)x * 12.9898 + (double)y * 4.1414;
    const float d = (float)dot;
    const float val = (float)std::sin((double)d * 43758.5453);
    return val - std::floor(val);
}

// kaminoLerpHost (KaminoInitializer.cu:103-106): (1.0 - a) * from + a * to,
// the first product in double, the second in FP32
float fbm_lerp(float from, float to, float alpha) {
    const float at = alpha * to;
    return (float)((1.0 - (double)alpha) * (double)from + (double)at);
}

// interpNoise2D (KaminoInitializer.cu:108-122)
float fbm_noise(float x, float y) {
    const float intX = std::floor(x), fractX = x - intX;
    const float intY = std::floor(y), fractY = y - intY;
    const float v1 = fbm_hash(intX, intY);
    const float v2 = fbm_hash(intX + 1.f, intY);
    const float v3 = fbm_hash(intX, intY + 1.f);
    const float v4 = fbm_hash(intX + 1.f, intY + 1.f);
    const float i1 = fbm_lerp(v1, v2, fractX);
    const float i2 = fbm_lerp(v3, v4, fractX);
    return fbm_lerp(i1, i2, fractY);
}

// FBM (KaminoInitializer.cu:86-101): four octaves, persistence 0.5,
// resolutions 0.15 in x and 0.5 in y, scaled by (1 - 0.5) / 2
float fbm(float x, float y) {
    const float resX = 0.15f, resY = 0.5f, persistence = 0.5f;
    float total = 0.0f;
    for (int i = 0; i < 4; ++i) {
        const float freq = (float)std::pow(2.0, i);
        const float amp = (float)std::pow((double)persistence, i);
        total += amp * fbm_noise(x * freq / resX, y * freq / resY);
    }
    const float a = 1.f - persistence;
    return a * total / 2.0f;
}

}  // namespace

extern "C" int kam_init_fbm(int n_theta, int n_phi, double radius, float* u_phi, float* u_theta) {
    if (n_theta < 2 || n_phi < 2 || !(radius > 0.0)) return KAM_EINVAL;
    const double kPi2 = 6.28318530717958647692;
    const float gridLen = (float)(kPi2 / n_phi);  // KaminoSolver.cu:14
    const float R = (float)radius;
    const float gain = (float)(4096.0 / n_phi);
    const float rh = R * gridLen;
    const float g2 = gridLen / 2;
    auto uphi_at = [&](float ur_x, float ul_x, size_t j) {
        const float ur_y = (float)(j + 1) * gridLen, lr_y = (float)j * gridLen;
        const float n_ur = fbm(ur_x, ur_y), n_lr = fbm(ur_x, lr_y);
        const float n_ul = fbm(ul_x, ur_y), n_ll = fbm(ul_x, lr_y);
        const float dyl = (n_ur - n_lr) / rh;
        const float dyr = (n_ul - n_ll) / rh;
        const float avg = (float)((dyl + dyr) / 2.0);
        return avg * gain;
    };
    if (u_phi)
        for (size_t j = 0; j < (size_t)n_theta; ++j) {
            for (size_t i = 1; i < (size_t)n_phi; ++i)
                u_phi[j * n_phi + i] = uphi_at((float)i * gridLen + g2, (float)i * gridLen - g2, j);
            // the phi = 0 seam (KaminoInitializer.cu:36-55): the west corners at 2 pi - h/2
            u_phi[j * n_phi] = uphi_at(g2, (float)(2 * 3.14159265358979323846 - g2), j);
        }
    if (u_theta)
        for (size_t j = 1; j < (size_t)n_theta; ++j)
            for (size_t i = 0; i < (size_t)n_phi; ++i) {
                const float ur_x = (float)(i + 1) * gridLen, ul_x = (float)i * gridLen;
                const float yu = (float)j * gridLen + g2, yd = (float)j * gridLen - g2;
                const float n_ur = fbm(ur_x, yu), n_lr = fbm(ur_x, yd);
                const float n_ul = fbm(ul_x, yu);
                // ll_y = j h + h/2, the same point as ul (KaminoInitializer.cu:69)
                const float n_ll = fbm(ul_x, yu);
                const float du = -1 * (n_ur - n_ul) / rh;
                const float dd = -1 * (n_lr - n_ll) / rh;
                const float avg = (float)((du + dd) / 2.0);
                u_theta[(j - 1) * n_phi + i] = avg * gain;
            }
    return KAM_OK;
}

// ---------------------------------------------------------------------------
// Image inputs (KaminoSolver::initDensityfromPic, KaminoSolver.cu:243-277, and
// the particle colours of KaminoParticles.cu:8-18, 62-76).  The reference
// reads images with OpenCV 2.4.13 (imread IMREAD_COLOR, flip around the
// vertical axis, resize to nPhi x nTheta with the default INTER_LINEAR).
// OpenCV is not part of this build: images are binary or ASCII PPM / PGM
// (P6, P3, P5, P2; grey replicated to B = G = R), and flip and resize restate
// OpenCV's scalar algorithm for 8-bit images: per axis the source coordinate
// (d + 0.5) * scale - 0.5 in FP32, weights rounded to 11-bit fixed point
// (columns clamped to the edge with weights (1, 0), rows clamped with their
// fractional weights kept), a horizontal then a vertical pass, and the result
// rounded as (sum + 2^21) >> 22.  OpenCV's SSE2 vertical pass pre-shifts by 4
// bits and may differ from this in the last bit of a pixel.
// ---------------------------------------------------------------------------
namespace {

struct Image {
    int w = 0, h = 0;
    std::vector<unsigned char> bgr;  // row-major, 3 bytes per pixel (B, G, R)
};

bool read_token(FILE* f, std::string& t) {
    t.clear();
    int c;
    while ((c = std::fgetc(f)) != EOF) {
        if (c == '#') {
            while ((c = std::fgetc(f)) != EOF && c != '\n') {}
            continue;
        }
        if (!std::isspace(c)) break;
    }
    if (c == EOF) return false;
    do t.push_back((char)c);
    while ((c = std::fgetc(f)) != EOF && !std::isspace(c));
    return true;
}

bool load_pnm(const char* path, Image& im) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    std::string magic, tw, th, tm;
    bool ok = read_token(f, magic) && read_token(f, tw) && read_token(f, th) && read_token(f, tm);
    const bool color = magic == "P6" || magic == "P3", binary = magic == "P6" || magic == "P5";
    ok = ok && (color || magic == "P5" || magic == "P2");
    if (ok) {
        im.w = std::atoi(tw.c_str());
        im.h = std::atoi(th.c_str());
        const int maxv = std::atoi(tm.c_str());
        ok = im.w > 0 && im.h > 0 && im.w <= 1 << 15 && im.h <= 1 << 15 && maxv > 0 && maxv < 256;
        if (ok) {
            const size_t n = (size_t)im.w * im.h, ch = color ? 3 : 1;
            std::vector<unsigned char> raw(n * ch);
            if (binary) {
                ok = std::fread(raw.data(), 1, raw.size(), f) == raw.size();
            } else {
                std::string t;
                for (size_t k = 0; k < raw.size() && ok; ++k) {
                    ok = read_token(f, t);
                    if (ok) raw[k] = (unsigned char)std::atoi(t.c_str());
                }
            }
            im.bgr.resize(n * 3);
            for (size_t p = 0; ok && p < n; ++p) {
                const unsigned char r = raw[p * ch], g = raw[p * ch + (ch - 1) / 2], b = raw[p * ch + ch - 1];
                // maxv < 255 is rescaled to 8 bits as an 8-bit decoder would
                auto s8 = [&](unsigned char v) { return (unsigned char)((v * 255 + maxv / 2) / maxv); };
                im.bgr[p * 3] = s8(b);
                im.bgr[p * 3 + 1] = s8(g);
                im.bgr[p * 3 + 2] = s8(r);
            }
        }
    }
    std::fclose(f);
    return ok;
}

// cv::flip(src, dst, 1): mirror around the vertical axis
void flip_columns(Image& im) {
    for (int y = 0; y < im.h; ++y)
        for (int x = 0; x < im.w / 2; ++x)
            for (int c = 0; c < 3; ++c)
                std::swap(im.bgr[((size_t)y * im.w + x) * 3 + c],
                          im.bgr[((size_t)y * im.w + (im.w - 1 - x)) * 3 + c]);
}

// The horizontal axis of INTER_LINEAR: source column and the two 11-bit
// weights per destination column, the coordinate clamped to the edge columns
// with weight (1, 0) there.
void linear_x(int ssize, int dsize, std::vector<int>& ofs, std::vector<short>& w0,
              std::vector<short>& w1) {
    const double scale = (double)ssize / dsize;
    ofs.resize(dsize);
    w0.resize(dsize);
    w1.resize(dsize);
    for (int d = 0; d < dsize; ++d) {
        float fx = (float)((d + 0.5) * scale - 0.5);
        int sx = (int)std::floor(fx);
        fx -= (float)sx;
        if (sx < 0) {
            fx = 0.f;
            sx = 0;
        }
        if (sx >= ssize - 1) {
            fx = 0.f;
            sx = ssize - 1;
        }
        ofs[d] = sx;
        w0[d] = (short)std::lrint((1.f - fx) * 2048.f);
        w1[d] = (short)std::lrint(fx * 2048.f);
    }
}

// cv::resize(src, dst, Size(dw, dh)) with INTER_LINEAR, 8-bit, 3 channels.
// Rows: the two source rows floor(fy) and floor(fy) + 1, each clamped to the
// image, with the unclamped fraction's weights.
Image resize_linear(const Image& s, int dw, int dh) {
    Image d;
    d.w = dw;
    d.h = dh;
    if (s.w == dw && s.h == dh) {
        d.bgr = s.bgr;
        return d;
    }
    d.bgr.resize((size_t)dw * dh * 3);
    std::vector<int> xo;
    std::vector<short> xa, xb;
    linear_x(s.w, dw, xo, xa, xb);
    const double scale_y = (double)s.h / dh;
    std::vector<int> row0((size_t)dw * 3), row1((size_t)dw * 3);
    auto hpass = [&](int sy, std::vector<int>& out) {
        const unsigned char* S = &s.bgr[(size_t)sy * s.w * 3];
        for (int x = 0; x < dw; ++x) {
            const int sx = xo[x], sx1 = sx + 1 < s.w ? sx + 1 : sx;
            for (int c = 0; c < 3; ++c)
                out[(size_t)x * 3 + c] = S[sx * 3 + c] * xa[x] + S[sx1 * 3 + c] * xb[x];
        }
    };
    auto clip = [&](int y) { return y < 0 ? 0 : (y >= s.h ? s.h - 1 : y); };
    for (int y = 0; y < dh; ++y) {
        float fy = (float)((y + 0.5) * scale_y - 0.5);
        const int sy = (int)std::floor(fy);
        fy -= (float)sy;
        const int b0 = (short)std::lrint((1.f - fy) * 2048.f), b1 = (short)std::lrint(fy * 2048.f);
        hpass(clip(sy), row0);
        hpass(clip(sy + 1), row1);
        unsigned char* D = &d.bgr[(size_t)y * dw * 3];
        for (int k = 0; k < dw * 3; ++k) {
            const int v = (row0[k] * b0 + row1[k] * b1 + (1 << 21)) >> 22;
            D[k] = (unsigned char)(v < 0 ? 0 : (v > 255 ? 255 : v));
        }
    }
    return d;
}

bool load_resized(const char* path, int n_theta, int n_phi, Image& out) {
    Image im;
    if (!path || !load_pnm(path, im)) return false;
    flip_columns(im);
    out = resize_linear(im, n_phi, n_theta);
    return true;
}

}  // namespace

extern "C" int kam_image_density(const char* path, int n_theta, int n_phi, float* rho) {
    if (n_theta < 2 || n_phi < 2 || !rho) return KAM_EINVAL;
    Image im;
    if (!load_resized(path, n_theta, n_phi, im)) return KAM_ENOTFOUND;
    // KaminoSolver.cu:263-273: (B + G + R) / 3 of the channels divided by 255
    for (int j = 0; j < n_theta; ++j)
        for (int i = 0; i < n_phi; ++i) {
            const unsigned char* p = &im.bgr[((size_t)j * n_phi + i) * 3];
            const float B = (float)(p[0] / 255.0), G = (float)(p[1] / 255.0), R = (float)(p[2] / 255.0);
            rho[(size_t)j * n_phi + i] = (float)((double)(B + G + R) / 3.0);
        }
    return KAM_OK;
}

extern "C" int kam_image_colors(const char* path, int n_theta, int n_phi, const float* coords, int64_t n,
                                float* colors) {
    if (n_theta < 2 || n_phi < 2 || n < 0 || (n > 0 && (!coords || !colors))) return KAM_EINVAL;
    Image im;
    if (!load_resized(path, n_theta, n_phi, im)) return KAM_ENOTFOUND;
    const float gridLen = (float)(6.28318530717958647692 / n_phi);
    // KaminoParticles.cu:62-76: the pixel under the particle's seeded
    // position, stored as (G, R, B) / 255 (the reference's colorBGR order)
    for (int64_t k = 0; k < n; ++k) {
        long x = (long)std::floor(coords[2 * k] / gridLen), y = (long)std::floor(coords[2 * k + 1] / gridLen);
        x = x < 0 ? 0 : (x >= n_phi ? n_phi - 1 : x);
        y = y < 0 ? 0 : (y >= n_theta ? n_theta - 1 : y);
        const unsigned char* p = &im.bgr[((size_t)y * n_phi + x) * 3];
        colors[3 * k] = (float)(p[1] / 255.0);
        colors[3 * k + 1] = (float)(p[2] / 255.0);
        colors[3 * k + 2] = (float)(p[0] / 255.0);
    }
    return KAM_OK;
}
