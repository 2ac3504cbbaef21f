// kamino -- configKamino.txt driver over the C ABI (include/kamino.h).
//
//   kamino configKamino.txt [key=value ...]
//
// The config holds the reference's 16 whitespace-separated tokens in its
// order (main.cu:17-45):
//   radius nTheta particleDensity dt DT frames A B C D E
//   gridPath particlePath densityImage solidImage colorImage
// followed, here or on the command line, by optional key=value settings:
//   seed=<u64> L=<int> density=band|noise geometric=0|1 format=bgeo|raw
//   device=<int> particles=0|1 init=fbm|sh resume=<raw snapshot>
//   resume_particles=<raw particles>
// A run resumed from the raw snapshot of frame k (format=raw writes one per
// frame) writes frames k+1 .. frames, equal to the uninterrupted run's.
// The frame schedule is Kamino::run's (KaminoCore.cu:886-904): steps of dt
// while T < i*DT, then one remainder step (which, as in the reference's
// kernels, also advances by dt), then frame i is written to <gridPath><i>.bgeo
// and, with particles (on by default, as the reference builds with
// WRITE_PARTICLES), the particles to <particlePath><i>.bgeo.  The particle
// count follows particleDensity truncated to an integer, as
// initParticlesfromPic receives it (KaminoSolver.cu:279).  The initial
// velocity is the reference's FBM curl noise (kam_init_fbm, init=fbm, the
// default) or the seeded spherical-harmonic flow of kam_init_fields (init=sh,
// with seed and L); the density is the band or the seeded noise of
// kam_init_fields.  A..E and the image inputs are not used, and particle
// colours are 0 (the reference's colours when it has no colour image).
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "kamino.h"

namespace {

struct Config {
    double radius = 5.0;
    int nTheta = 128;
    double particleDensity = 200.0;
    float dt = 0.005f, DT = 1.0f / 24.0f;
    int frames = 1000;
    double A = 0.0;
    int B = 1, C = 1, D = 1, E = 1;
    std::string gridPath = "output/frame", particlePath = "particles/frame";
    std::string densityImage, solidImage, colorImage;
    std::map<std::string, std::string> kv;
};

bool parse_config(const std::string& path, Config& c, std::string& err) {
    std::ifstream f(path);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    std::vector<std::string> tok;
    for (std::string t; f >> t;) tok.push_back(t);
    if (tok.size() < 16) {
        err = "config needs 16 tokens (radius nTheta particleDensity dt DT frames A B C D E "
              "gridPath particlePath densityImage solidImage colorImage), found " +
              std::to_string(tok.size());
        return false;
    }
    try {
        c.radius = std::stod(tok[0]);
        c.nTheta = std::stoi(tok[1]);
        c.particleDensity = std::stod(tok[2]);
        c.dt = std::stof(tok[3]);
        c.DT = std::stof(tok[4]);
        c.frames = std::stoi(tok[5]);
        c.A = std::stod(tok[6]);
        c.B = std::stoi(tok[7]);
        c.C = std::stoi(tok[8]);
        c.D = std::stoi(tok[9]);
        c.E = std::stoi(tok[10]);
    } catch (const std::exception& e) {
        err = std::string("bad numeric token: ") + e.what();
        return false;
    }
    c.gridPath = tok[11];
    c.particlePath = tok[12];
    c.densityImage = tok[13] == "null" ? "" : tok[13];
    c.solidImage = tok[14] == "null" ? "" : tok[14];
    c.colorImage = tok[15] == "null" ? "" : tok[15];
    for (size_t k = 16; k < tok.size(); ++k) {
        const auto eq = tok[k].find('=');
        if (eq == std::string::npos) {
            err = "trailing token is not key=value: " + tok[k];
            return false;
        }
        c.kv[tok[k].substr(0, eq)] = tok[k].substr(eq + 1);
    }
    if (c.nTheta < 4 || !(c.dt > 0) || !(c.DT > 0) || c.frames < 0 || !(c.radius > 0)) {
        err = "need nTheta >= 4, dt > 0, DT > 0, frames >= 0, radius > 0";
        return false;
    }
    return true;
}

// ---- Houdini binary geometry (.bgeo, version 5), big-endian -------------
void put_be32(std::ofstream& o, uint32_t v) {
    const unsigned char b[4] = {(unsigned char)(v >> 24), (unsigned char)(v >> 16),
                                (unsigned char)(v >> 8), (unsigned char)v};
    o.write((const char*)b, 4);
}
void put_be16(std::ofstream& o, uint16_t v) {
    const unsigned char b[2] = {(unsigned char)(v >> 8), (unsigned char)v};
    o.write((const char*)b, 2);
}
void put_f(std::ofstream& o, float f) {
    uint32_t v;
    std::memcpy(&v, &f, 4);
    put_be32(o, v);
}
void put_str(std::ofstream& o, const char* s) {
    put_be16(o, (uint16_t)std::strlen(s));
    o.write(s, std::strlen(s));
}

// A point attribute of a .bgeo file: `size` floats per point; vectors are
// Houdini type 5, plain floats type 0.
struct Attr {
    const char* name;
    int size;
    const float* data;
};

// Houdini geometry V5 (the layout Partio::write produces for particles):
// points with P (and w = 1) plus the attributes, then one particle-system
// primitive that lists every point.  write_data_bgeo (KaminoSolver.cu:284-357)
// writes "v" and "density"; write_particles_bgeo (:359-401) "v" and "color".
bool write_bgeo(const std::string& file, int n, const float* pos, const std::vector<Attr>& attrs) {
    std::ofstream o(file, std::ios::binary);
    if (!o) return false;
    o.write("Bgeo", 4);
    o.put('V');
    put_be32(o, 5);
    put_be32(o, (uint32_t)n);             // points
    put_be32(o, 1);                       // primitives
    put_be32(o, 0);                       // point groups
    put_be32(o, 0);                       // primitive groups
    put_be32(o, (uint32_t)attrs.size());  // point attributes (besides P)
    put_be32(o, 0);                       // vertex attributes
    put_be32(o, 0);                       // primitive attributes
    put_be32(o, 0);                       // detail attributes
    for (const Attr& a : attrs) {
        put_str(o, a.name);
        put_be16(o, (uint16_t)a.size);
        put_be32(o, a.size == 3 ? 5 : 0);  // vector or float
        for (int k = 0; k < a.size; ++k) put_f(o, 0.f);
    }
    for (int p = 0; p < n; ++p) {
        for (int k = 0; k < 3; ++k) put_f(o, pos[3 * p + k]);
        put_f(o, 1.f);  // homogeneous w
        for (const Attr& a : attrs)
            for (int k = 0; k < a.size; ++k) put_f(o, a.data[(size_t)a.size * p + k]);
    }
    put_be32(o, 0x8000);  // particle-system primitive
    put_be32(o, (uint32_t)n);
    for (int p = 0; p < n; ++p) {
        if (n <= 65536)
            put_be16(o, (uint16_t)p);
        else
            put_be32(o, (uint32_t)p);
    }
    o.put((char)0x00);
    o.put((char)0xff);
    return (bool)o;
}

// Particle positions on the sphere (mapPToSphere, KaminoSolver.cu:403-409).
void particle_points(double R, const std::vector<float>& coords, std::vector<float>& pos) {
    const size_t n = coords.size() / 2;
    pos.resize(3 * n);
    for (size_t k = 0; k < n; ++k) {
        const float phi = coords[2 * k], theta = coords[2 * k + 1];
        pos[3 * k] = (float)(R * std::sin(theta) * std::cos(phi));
        pos[3 * k + 1] = (float)(R * std::cos(theta));
        pos[3 * k + 2] = (float)(R * std::sin(theta) * std::sin(phi));
    }
}

// Cell-centred velocity on the sphere (KaminoSolver.cu:305-353, with
// mapPToSphere / mapVToSphere :403-423): u_phi averaged west/east, u_theta
// north/south (the pole rows use their single neighbouring u_theta row).
void frame_points(int nT, int nP, double R, const std::vector<float>& up,
                  const std::vector<float>& ut, std::vector<float>& pos, std::vector<float>& vel) {
    const double h = 2.0 * 3.14159265358979323846 / nP;
    pos.resize((size_t)3 * nT * nP);
    vel.resize((size_t)3 * nT * nP);
    for (int j = 0; j < nT; ++j)
        for (int i = 0; i < nP; ++i) {
            const int ie = (i == nP - 1) ? 0 : i + 1;
            const int jn = (j == 0) ? 0 : (j == nT - 1 ? nT - 2 : j - 1);
            const int js = (j == 0) ? 0 : (j == nT - 1 ? nT - 2 : j);
            const double vphi = 0.5 * (up[(size_t)j * nP + i] + up[(size_t)j * nP + ie]);
            const double vth = 0.5 * (ut[(size_t)jn * nP + i] + ut[(size_t)js * nP + i]);
            const double phi = i * h, th = (j + 0.5) * h;
            const size_t o = 3 * ((size_t)j * nP + i);
            pos[o] = (float)(R * std::sin(th) * std::cos(phi));
            pos[o + 1] = (float)(R * std::cos(th));
            pos[o + 2] = (float)(R * std::sin(th) * std::sin(phi));
            vel[o] = (float)(std::cos(th) * std::cos(phi) * vth - std::sin(phi) * vphi);
            vel[o + 1] = (float)(-std::sin(th) * vth);
            vel[o + 2] = (float)(std::cos(th) * std::sin(phi) * vth + std::cos(phi) * vphi);
        }
}

// Raw snapshot: "KAMR", int32 nTheta, int32 nPhi, int64 step, float64 time,
// then u_phi, u_theta, rho as little-endian float32 (a resumable checkpoint).
bool write_raw(const std::string& file, int nT, int nP, int64_t step, double t,
               const std::vector<float>& up, const std::vector<float>& ut, const std::vector<float>& rho) {
    std::ofstream o(file, std::ios::binary);
    if (!o) return false;
    o.write("KAMR", 4);
    o.write((const char*)&nT, 4);
    o.write((const char*)&nP, 4);
    o.write((const char*)&step, 8);
    o.write((const char*)&t, 8);
    o.write((const char*)up.data(), up.size() * 4);
    o.write((const char*)ut.data(), ut.size() * 4);
    o.write((const char*)rho.data(), rho.size() * 4);
    return (bool)o;
}

// Raw particles: "KAMP", int64 count, then (phi, theta) float32 pairs.
bool write_raw_particles(const std::string& file, const std::vector<float>& coords) {
    std::ofstream o(file, std::ios::binary);
    if (!o) return false;
    const int64_t n = (int64_t)(coords.size() / 2);
    o.write("KAMP", 4);
    o.write((const char*)&n, 8);
    o.write((const char*)coords.data(), coords.size() * 4);
    return (bool)o;
}

bool read_raw_particles(const std::string& file, std::vector<float>& coords) {
    std::ifstream f(file, std::ios::binary);
    char magic[4];
    int64_t n = 0;
    if (!f.read(magic, 4) || std::memcmp(magic, "KAMP", 4) != 0 || !f.read((char*)&n, 8)) return false;
    if (n < 0 || n > ((int64_t)1 << 30)) return false;
    coords.resize((size_t)n * 2);
    f.read((char*)coords.data(), coords.size() * 4);
    return (bool)f;
}

bool read_raw(const std::string& file, int nT, int nP, int64_t& step, double& t, std::vector<float>& up,
              std::vector<float>& ut, std::vector<float>& rho) {
    std::ifstream f(file, std::ios::binary);
    char magic[4];
    int a = 0, b = 0;
    if (!f.read(magic, 4) || std::memcmp(magic, "KAMR", 4) != 0) return false;
    f.read((char*)&a, 4);
    f.read((char*)&b, 4);
    f.read((char*)&step, 8);
    f.read((char*)&t, 8);
    if (a != nT || b != nP) return false;
    f.read((char*)up.data(), up.size() * 4);
    f.read((char*)ut.data(), ut.size() * 4);
    f.read((char*)rho.data(), rho.size() * 4);
    return (bool)f;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cout << "Please provide the path to configKamino.txt as an argument.\n"
                  << "Usage example: ./kamino ./configKamino.txt [key=value ...]\n";
        return 1;
    }
    Config c;
    std::string err;
    if (!parse_config(argv[1], c, err)) {
        std::cerr << "kamino: " << err << "\n";
        return 1;
    }
    for (int k = 2; k < argc; ++k) {
        const std::string a = argv[k];
        const auto eq = a.find('=');
        if (eq == std::string::npos) {
            std::cerr << "kamino: argument is not key=value: " << a << "\n";
            return 1;
        }
        c.kv[a.substr(0, eq)] = a.substr(eq + 1);
    }
    auto get = [&](const char* k, const std::string& d) { return c.kv.count(k) ? c.kv[k] : d; };
    const uint64_t seed = std::strtoull(get("seed", "1").c_str(), nullptr, 10);
    const int L = std::atoi(get("L", "8").c_str());
    const int densityKind = get("density", "band") == "noise" ? 0 : 1;
    const int geometric = std::atoi(get("geometric", "1").c_str());
    const std::string format = get("format", "bgeo");
    const int device = std::atoi(get("device", "0").c_str());
    const bool particles = std::atoi(get("particles", "1").c_str()) != 0;
    const std::string init = get("init", "fbm");
    if (init != "fbm" && init != "sh") {
        std::cerr << "kamino: init must be fbm or sh\n";
        return 1;
    }
    if (!c.solidImage.empty())
        std::cerr << "kamino: the solid image is not used (nor is it by the reference's step)\n";

    const int nT = c.nTheta, nP = 2 * nT;
    kam_sim_desc d{};
    d.n_theta = nT;
    d.n_phi = nP;
    d.radius = c.radius;
    d.dt = c.dt;
    d.geometric = geometric;
    d.precision = KAM_PREC_FP32;
    d.device = device;
    kam_sim* sim = nullptr;
    if (kam_sim_create(&d, &sim) != KAM_OK) {
        std::cerr << "kamino: " << kam_last_error() << "\n";
        return 2;
    }
    std::vector<float> up((size_t)nT * nP), ut((size_t)(nT - 1) * nP), rho((size_t)nT * nP);
    int64_t step = 0;
    double Tstart = 0.0;
    if (c.kv.count("resume")) {
        if (!read_raw(c.kv["resume"], nT, nP, step, Tstart, up, ut, rho)) {
            std::cerr << "kamino: cannot resume from " << c.kv["resume"] << "\n";
            return 2;
        }
    } else if (kam_init_fields(nT, nP, c.radius, L, seed, densityKind, up.data(), ut.data(),
                               rho.data()) != KAM_OK ||
               (init == "fbm" && kam_init_fbm(nT, nP, c.radius, up.data(), ut.data()) != KAM_OK)) {
        std::cerr << "kamino: bad initial-state parameters\n";
        return 2;
    }
    // initDensityfromPic (KaminoSolver.cu:243-277) replaces the density
    if (!c.kv.count("resume") && !c.densityImage.empty() &&
        kam_image_density(c.densityImage.c_str(), nT, nP, rho.data()) != KAM_OK) {
        std::cerr << "kamino: cannot read density image " << c.densityImage << " (PPM / PGM)\n";
        return 2;
    }
    if (kam_sim_set_fields(sim, up.data(), ut.data(), rho.data()) != KAM_OK) {
        std::cerr << "kamino: " << kam_last_error() << "\n";
        return 2;
    }
    std::vector<float> coords;
    if (particles) {
        if (c.kv.count("resume_particles")) {
            if (!read_raw_particles(c.kv["resume_particles"], coords)) {
                std::cerr << "kamino: cannot resume particles from " << c.kv["resume_particles"] << "\n";
                return 2;
            }
        } else {
            const double perCell = (double)(size_t)(c.particleDensity > 0 ? c.particleDensity : 0);
            int64_t n = 0;
            if (kam_particles_seed(nT, perCell, seed, nullptr, &n) != KAM_OK) {
                std::cerr << "kamino: bad particle density\n";
                return 2;
            }
            coords.resize((size_t)n * 2);
            kam_particles_seed(nT, perCell, seed, coords.data(), &n);
        }
        if (kam_sim_set_particles(sim, coords.data(), (int64_t)(coords.size() / 2)) != KAM_OK) {
            std::cerr << "kamino: " << kam_last_error() << "\n";
            return 2;
        }
    }
    // particle colours from the colour image (KaminoParticles.cu:62-76), else 0;
    // write_particles_bgeo writes colorBGR[3i+1], [3i+2], [3i+3] for particle i
    // (KaminoSolver.cu:382-384), reading one entry past the particle's own
    std::vector<float> colorBGR;
    if (particles) {
        const size_t n = coords.size() / 2;
        colorBGR.assign(3 * n + 1, 0.f);
        if (!c.colorImage.empty() &&
            kam_image_colors(c.colorImage.c_str(), nT, nP, coords.data(), (int64_t)n, colorBGR.data()) != KAM_OK) {
            std::cerr << "kamino: cannot read colour image " << c.colorImage << " (PPM / PGM)\n";
            return 2;
        }
    }
    std::vector<float> pos, vel, color;
    auto write_frame = [&](int frame) {
        if (kam_sim_get_fields(sim, up.data(), ut.data(), rho.data()) != KAM_OK) return false;
        const std::string file = c.gridPath + std::to_string(frame) + (format == "raw" ? ".kamr" : ".bgeo");
        std::cout << "Writing to: " << file << std::endl;
        if (format == "raw") {
            if (!write_raw(file, nT, nP, step, Tstart, up, ut, rho)) return false;
        } else {
            frame_points(nT, nP, c.radius, up, ut, pos, vel);
            if (!write_bgeo(file, nT * nP, pos.data(), {{"v", 3, vel.data()}, {"density", 1, rho.data()}}))
                return false;
        }
        if (!particles) return true;
        if (kam_sim_get_particles(sim, coords.data(), (int64_t)(coords.size() / 2)) != KAM_OK) return false;
        const std::string pfile = c.particlePath + std::to_string(frame) + (format == "raw" ? ".kamp" : ".bgeo");
        std::cout << "Writing to: " << pfile << std::endl;
        if (format == "raw") return write_raw_particles(pfile, coords);
        const size_t n = coords.size() / 2;
        particle_points(c.radius, coords, pos);
        vel.assign(3 * n, 0.f);  // write_particles_bgeo writes v = 0
        color.resize(3 * n);
        for (size_t k = 0; k < 3 * n; ++k) color[k] = colorBGR[k + 1];
        return write_bgeo(pfile, (int)n, pos.data(), {{"v", 3, vel.data()}, {"color", 3, color.data()}});
    };
    // A resumed run continues the frame schedule of the run that wrote the
    // snapshot: the snapshot's time is k DT after frame k, so the run goes on
    // with frame k + 1 from T = k DT (the same float the uninterrupted run holds
    // there) and does not rewrite frame k.
    int first = 1;
    float T = 0.0f;  // simulation time, float as in KaminoCore.cu:886
    if (c.kv.count("resume")) {
        const double kf = Tstart / (double)c.DT;
        const long k = std::lround(kf);
        if (k < 0 || k > c.frames || std::fabs(kf - (double)k) > 1e-3) {
            std::cerr << "kamino: snapshot time " << Tstart << " is not a frame time of this config (frame "
                      << kf << " of " << c.frames << ")\n";
            return 2;
        }
        first = (int)k + 1;
        T = (int)k * c.DT;
    } else if (!write_frame(0)) {
        std::cerr << "kamino: cannot write frame 0 (" << kam_last_error() << ")\n";
        return 3;
    }
    const auto t0 = std::chrono::steady_clock::now();
    double stepMs = 0.0;
    for (int i = first; i <= c.frames; ++i) {
        int n = 0;
        while (T < i * c.DT) {
            ++n;
            T += c.dt;
        }
        ++n;  // the remainder step stepForward(dt + i*DT - T); its kernels use dt
        T = i * c.DT;
        const auto s0 = std::chrono::steady_clock::now();
        if (kam_sim_step(sim, n) != KAM_OK || kam_sim_sync(sim) != KAM_OK) {
            std::cerr << "kamino: " << kam_last_error() << "\n";
            return 2;
        }
        stepMs += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - s0).count();
        step += n;
        Tstart = T;
        std::cout << "Frame " << i << " is ready" << std::endl;
        if (!write_frame(i)) {
            std::cerr << "kamino: cannot write frame " << i << "\n";
            return 3;
        }
    }
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "Time spent: " << ms << "ms" << std::endl;
    std::cout << "Performance: " << (ms > 0 ? 1000.0 * (c.frames - first + 1) / ms : 0.0) << " frames per second"
              << std::endl;
    std::cout << "Steps: " << step << ", stepping " << stepMs << " ms, "
              << (stepMs > 0 ? (double)step * nT * nP / (stepMs * 1e-3) : 0.0) << " cell-steps/s" << std::endl;
    kam_sim_destroy(sim);
    return 0;
}
